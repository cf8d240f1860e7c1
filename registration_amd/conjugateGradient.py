"""Nonlinear conjugate gradient, drop-in for py-lddmm's conjugateGradient.cg
(py-lddmm/conjugateGradient.py:33-217).

The optimiser is written once, as a generator ``cg_steps`` that *yields* every request that
needs the device (objective, gradient, RKHS products, line-search trials) and receives the
result.  ``cg`` drives it synchronously against one optimisable object, exactly as the
reference does; the serving layer drives many generators in lockstep and answers their
requests with batched launches.  Every decision (step acceptance, PR+ beta, stopping tests)
is taken on host FP64 scalars with the reference's constants: epsMin = 1e-10, small-variation
test 1e-6, step x100 after an accepted iteration, x0.5 / x1.25 line-search moves, restart
every ``restartRate`` (100) iterations.

Optimisable protocol (as in the reference): getVariable, objectiveFun, updateTry(dir, eps,
objRef), acceptVarTry, getGradient(coeff), and optionally dotProduct, addProd, copyDir,
randomDir, startOptim, startOfIteration, endOfIteration, endOptim; attributes gradEps,
cgBurnIn, gradCoeff, restartRate.
"""
from __future__ import annotations

import logging

import numpy as np

OBJECTIVE = "objectiveFun"
GRADIENT = "getGradient"
DOT = "dotProduct"
TRY = "updateTry"
ACCEPT = "acceptVarTry"
END_ITER = "endOfIteration"


def cg_steps(opt, verb=True, maxIter=1000, TestGradient=False, epsInit=10.):
    """Generator form of cg.  Yields (name, *args) requests; send back their results."""
    if not all(hasattr(opt, f) for f in ('getVariable', 'objectiveFun', 'updateTry',
                                          'acceptVarTry', 'getGradient')):
        logging.error('Error: required functions are not provided')
        return None

    if hasattr(opt, 'startOptim'):
        opt.startOptim()
    gradEps = getattr(opt, 'gradEps', 1.0)
    cgBurnIn = getattr(opt, 'cgBurnIn', -1)
    gradCoeff = getattr(opt, 'gradCoeff', 1.0)
    restartRate = getattr(opt, 'restartRate', 100)
    hasDot = hasattr(opt, 'dotProduct')

    eps = epsInit
    epsMin = 1e-10
    opt.converged = False
    opt.stop_reason = 'maxIter'

    obj = yield (OBJECTIVE,)
    logging.info('iteration 0: obj = {0: .5f}'.format(obj))
    opt.history = [obj]

    skipCG = 0
    for it in range(maxIter):
        if it % restartRate == 0:
            skipCG = 1
        if hasattr(opt, 'startOfIteration'):
            opt.startOfIteration()
        grd = yield (GRADIENT, gradCoeff)

        if TestGradient:
            if hasattr(opt, 'randomDir'):
                dirfoo = opt.randomDir()
            else:
                dirfoo = np.random.normal(size=grd.shape)
            epsfoo = 1e-7
            objfoo = yield (TRY, dirfoo, epsfoo, obj - 1e10)
            if hasDot:
                [grdfoo] = yield (DOT, grd, [dirfoo])
            else:
                grdfoo = np.multiply(grd, dirfoo).sum()
            logging.info('Test Gradient: %.4f %.4f' % ((objfoo - obj) / epsfoo, -grdfoo * gradCoeff))

        if it == 0 or it == cgBurnIn:
            if hasDot:
                [grdOld2] = yield (DOT, grd, [grd])
            else:
                grdOld2 = np.multiply(grd, grd).sum()
            grd2 = grdOld2
            grdTry = np.sqrt(max(1e-20, grdOld2))
            if hasattr(opt, 'copyDir'):
                oldDir = opt.copyDir(grd)
                grdOld = opt.copyDir(grd)
                dir0 = opt.copyDir(grd)
            else:
                oldDir = np.copy(grd)
                grdOld = np.copy(grd)
                dir0 = np.copy(grd)
            beta = 0
        else:
            if hasDot:
                [grd2, grd12] = yield (DOT, grd, [grd, grdOld])
            else:
                grd2 = np.multiply(grd, grd).sum()
                grd12 = np.multiply(grd, grdOld).sum()
            if skipCG == 0:
                beta = max(0, (grd2 - grd12) / grdOld2)
            else:
                beta = 0
            grdOld2 = grd2
            grdTry = np.sqrt(max(1e-20, grd2 + beta * grd12))
            if hasattr(opt, 'addProd'):
                dir0 = opt.addProd(grd, oldDir, beta)
            else:
                dir0 = grd + beta * oldDir
            if hasattr(opt, 'copyDir'):
                oldDir = opt.copyDir(dir0)
                grdOld = opt.copyDir(grd)
            else:
                oldDir = np.copy(dir0)
                grdOld = np.copy(grd)

        objTry = yield (TRY, dir0, eps, obj)

        noUpdate = 0
        if objTry > obj:
            epsSmall = 0.000001 / grdTry
            objTry0 = yield (TRY, dir0, epsSmall, obj)
            if objTry0 > obj:
                if (skipCG == 1) | (beta < 1e-10):
                    logging.info('iteration {0:d}: obj = {1:.5f}, eps = {2:.5f}, beta = {3:.5f}, '
                                 'gradient: {4:.5f}'.format(it + 1, obj, eps, beta, np.sqrt(grd2)))
                    logging.info('Stopping Gradient Descent: bad direction')
                    opt.stop_reason = 'bad direction'
                    break
                else:
                    if verb:
                        logging.info('Disabling CG: bad direction')
                        skipCG = 1
                        noUpdate = 1
            else:
                while (objTry > obj) and (eps > epsMin):
                    eps = eps / 2
                    objTry = yield (TRY, dir0, eps, obj)

        if noUpdate == 0:
            contt = 1
            while contt == 1:
                objTry2 = yield (TRY, dir0, .5 * eps, objTry)
                if objTry > objTry2:
                    eps = eps / 2
                    objTry = objTry2
                else:
                    contt = 0
            contt = 1
            while contt == 1:
                objTry2 = yield (TRY, dir0, 1.25 * eps, objTry)
                if objTry > objTry2:
                    eps *= 1.25
                    objTry = objTry2
                else:
                    contt = 0

            if np.fabs(obj - objTry) < .000001:
                if (skipCG == 1) | (beta < 1e-10):
                    logging.info('iteration {0:d}: obj = {1:.5f}, eps = {2:.5f}, beta = {3:.5f}, '
                                 'gradient: {4:.5f}'.format(it + 1, obj, eps, beta, np.sqrt(grd2)))
                    logging.info('Stopping Gradient Descent: small variation')
                    opt.converged = True
                    opt.stop_reason = 'small variation'
                    break
                else:
                    if verb:
                        logging.info('Disabling CG: small variation')
                    skipCG = 1
                    eps = 1.0
            else:
                skipCG = 0

            yield (ACCEPT,)
            obj = objTry
            opt.history.append(obj)
            if verb | (it == maxIter):
                logging.info('iteration {0:d}: obj = {1:.5f}, eps = {2:.5f}, beta = {3:.5f}, '
                             'gradient: {4:.5f}'.format(it + 1, obj, eps, beta, np.sqrt(grd2)))

            if np.sqrt(grd2) < gradEps:
                logging.info('Stopping Gradient Descent: small gradient')
                opt.converged = True
                opt.stop_reason = 'small gradient'
                break
            eps = 100 * eps

            if hasattr(opt, 'endOfIteration'):
                yield (END_ITER,)

    if hasattr(opt, 'endOptim'):
        opt.endOptim()
    return opt.getVariable()


def drive(opt, gen):
    """Run a cg_steps generator to completion against ``opt`` (one request at a time)."""
    try:
        req = next(gen)
        while True:
            res = getattr(opt, req[0])(*req[1:])
            req = gen.send(res)
    except StopIteration as stop:
        return stop.value


def cg(opt, verb=True, maxIter=1000, TestGradient=False, epsInit=10.):
    """Same signature and behaviour as the reference's cg; returns opt.getVariable()."""
    return drive(opt, cg_steps(opt, verb=verb, maxIter=maxIter, TestGradient=TestGradient,
                               epsInit=epsInit))
