"""Second-order (spline) regression of a template surface through a time series of targets:
the drop-in for py-lddmm/secondOrderMatching.py (``SurfaceMatching``, :43-509).

Model.  The state (x, a) of the template's points follows x' = K(x) a (+ A x + b),
a' = -K'(x)(a, a) + rho (- a A) for T = nTarg T1 Euler steps (pointEvolution.secondOrder*,
GPU).  The variables are the initial momentum a0 (N, d), the controls rhot (T, N, d) and the
affine coefficients Afft (T, k) on an AffineBasis.  The objective is

    1/2 <a0, K(x0) a0> + dt cw sum_t |rho_t|^2 / 2 + dt sum_t <w, Afft_t^2> / 2
        + sum_k |S(x_{(k+1) T1}) - target_k|^2 / sigmaError^2     (+ the targets' norms)

with the currents, measure or varifold data term.  The gradient comes from the GPU adjoint
(pointEvolution.secondOrderGradient): d a0 = K(x0) a0 - pa_0, d rho_t = cw rho_t - pa_{t+1},
and dA, db for the affine part.

Design.  A registration here is device-resident: the variables, their trial values and every
direction are device tensors; the target geometry (centers, surfels and measure weights) is
uploaded once; each request (objective, line-search trial, gradient, dot products) gathers its
scalars into one device buffer and reads it back once.  The optimiser is
conjugateGradient.cg with the reference's protocol.

The optimisation path is the reference's, including what decides it: the regression type is
'affine' (the a0 and rho gradients set to zero) for the first affBurnIn = 50 iterations, then
the requested one ('spline2': both; 'spline': rho only; 'geodesic': a0 only); the affine
gradient is taken only during the burn-in, where cg also re-initialises its direction
(cgBurnIn); dot products are Euclidean, with the rho part divided by T and the affine part
weighted by coeffAff = 1; the adjoint re-shoots the trajectory without the affine part, as the
reference's secondOrderCovector does (pointEvolution.py:986).
"""
from __future__ import annotations

import logging
import os

import numpy as np
import torch

from . import conjugateGradient as cg
from . import engine
from . import pointEvolution as evol
from . import runtime as rt
from . import surfaces
from .affineBasis import AffineBasis
from .surfaceMatching import SurfaceMatchingParam


class Direction:
    """A direction in the regression's variables: a0 (N, d), rhot (T, N, d) and aff (T, k),
    device tensors (secondOrderMatching.py:21-25 holds the same three parts)."""

    __slots__ = ("a0", "rhot", "aff")

    def __init__(self, a0=None, rhot=None, aff=None):
        self.a0, self.rhot, self.aff = a0, rhot, aff

    def axpy(self, beta, other):
        """self + beta * other."""
        return Direction(self.a0 + beta * other.a0, self.rhot + beta * other.rhot,
                         self.aff + beta * other.aff)

    def clone(self):
        return Direction(self.a0.clone(), self.rhot.clone(), self.aff.clone())


class _TargetSeries:
    """The data terms of the template's deformations against the time series of targets: the
    targets' device geometry, one device shape per target for the deformed template, and the
    FP64 energies / gradients of all of them."""

    def __init__(self, faces, x0, targets, kind, Kdist, sigmaError):
        self.kind = kind
        self.K = Kdist
        self.s2 = sigmaError ** 2
        weights = kind != 'current'
        self.c = [rt.as_dev(t.centers) for t in targets]
        self.cr = [rt.as_dev(t.surfel) for t in targets]
        self.a = [rt.as_dev(engine.measure_weights(t.surfel)) for t in targets] if weights else None
        self.shapes = [surfaces.DeviceSurface(faces, x0, with_weights=weights) for _ in targets]
        self.norms = None

    def _energy_job(self, k, c1, cr1, a1, total, with_target=True):
        Ke = self.K.energy()
        c2 = self.c[k] if with_target else None
        if self.kind == 'current':
            if not with_target:
                return engine.KNormJob(c1, cr1, Ke, total)
            return engine.CurrentsJob(c1, cr1, c2, self.cr[k], Ke, total)
        if self.kind == 'measure':
            return engine.MeasureJob(c1, a1, c2, self.a[k] if with_target else None, Ke, total)
        return engine.VarifoldJob(c1, a1, cr1, c2, self.a[k] if with_target else None,
                                  self.cr[k] if with_target else None, Ke, total)

    def _run(self, jobs):
        if self.kind == 'measure':
            engine.measure(jobs)
        elif self.kind == 'varifold':
            engine.varifold(jobs)
        else:
            engine.currents([j for j in jobs if isinstance(j, engine.CurrentsJob)])
            engine.knorm([j for j in jobs if isinstance(j, engine.KNormJob)])

    def target_norms(self, out):
        """|target_k|^2 into out[k] (the reference's fun_obj0, with the measure offset)."""
        jobs = [self._energy_job(k, self.c[k], self.cr[k], self.a[k] if self.a else None,
                                 out[k:k + 1], with_target=False) for k in range(len(self.c))]
        self._run(jobs)

    def energies(self, xs, out):
        """|S(xs[k]) - target_k|^2 - |target_k|^2 (the reference's fun_obj) into out[k]."""
        engine.geometry_batch(self.shapes, xs)
        jobs = [self._energy_job(k, s.centers, s.surfel, s.weights, out[k:k + 1])
                for k, s in enumerate(self.shapes)]
        self._run(jobs)

    def endpoint_covectors(self, xs):
        """px1_k = -gradient of the k-th data term / sigmaError^2 at the vertices xs[k]."""
        out = []
        for k, (s, x) in enumerate(zip(self.shapes, xs)):
            s.update(x)
            px = torch.empty_like(x)
            scale = -2.0 / self.s2
            if self.kind == 'current':
                surfaces.current_gradient_dev(s, self.c[k], self.cr[k], self.K, scale, px)
            elif self.kind == 'measure':
                surfaces.measure_gradient_dev(s, self.c[k], self.a[k], self.K, scale, px)
            else:
                surfaces.varifold_gradient_dev(s, self.c[k], self.a[k], self.cr[k], self.K,
                                               scale, px)
            out.append(px)
        return out


def _surface(surface, filename, what):
    if surface is not None:
        return surfaces.Surface(surf=surface)
    if filename is None:
        raise ValueError('Please provide a %s' % what)
    return surfaces.Surface(filename=filename)


class SurfaceMatching:
    """Spline regression through a time series of surfaces (secondOrderMatching.py:43-509).

    Template: surface (or fileTempl); Targets: list of surfaces (or fileTarg, a list of file
    names); param: SurfaceMatchingParam (kernels, sigmaError, errorType, timeStep = 1/T1);
    initialMomentum: a0 to start from; typeRegression: 'spline2' (default), 'spline',
    'geodesic' or 'affine', taken after the affine burn-in; controlWeight: weight cw of
    |rho|^2; affine / affineWeight / rotWeight / scaleWeight / transWeight as in
    surfaceMatching.  Files: Template.vtk and Target<k>.vtk at construction and, every saveRate
    iterations, the deformed template at every time step with its Jacobian
    (saveFile<t>.vtk); ``outputDir=None`` writes nothing.  subsampleTargetSize > 0 decimates
    the template to that many vertices first (Surface.Simplify) and orients the targets
    against it (secondOrderMatching.py:112-120)."""

    def __init__(self, Template=None, Targets=None, fileTempl=None, fileTarg=None, param=None,
                 initialMomentum=None, maxIter=1000, regWeight=1.0, verb=True,
                 typeRegression='spline2', affine='none', controlWeight=1.0, affineWeight=1.0,
                 rotWeight=None, scaleWeight=None, transWeight=None, subsampleTargetSize=-1,
                 testGradient=True, saveFile='evolution', outputDir='.'):
        self.fv0 = _surface(Template, fileTempl, 'template surface')
        if Targets is None and fileTarg is None:
            raise ValueError('Please provide a list of target surfaces')
        self.fv1 = ([surfaces.Surface(surf=s) for s in Targets] if Targets is not None
                    else [surfaces.Surface(filename=f) for f in fileTarg])
        if subsampleTargetSize > 0:
            self.fv0.Simplify(subsampleTargetSize)
            v0 = self.fv0.surfVolume()
            for s in self.fv1:
                if v0 * s.surfVolume() < 0:
                    s.flipFaces()
        self.param = SurfaceMatchingParam() if param is None else param
        if self.param.errorType not in ('current', 'measure', 'varifold'):
            raise NotImplementedError("the GPU path implements the currents, measure and "
                                      "varifold data terms")
        self.nTarg = len(self.fv1)
        self.maxIter, self.verb, self.testGradient = maxIter, verb, testGradient
        self.regweight = regWeight
        self.saveFile, self.saveRate = saveFile, 10
        self.iter = 0
        self.npt, self.dim = self.fv0.vertices.shape
        self.Tsize1 = int(round(1.0 / self.param.timeStep))
        self.Tsize = self.nTarg * self.Tsize1
        self.gradCoeff = self.npt
        self.controlWeight = float(controlWeight)
        self.typeRegression, self.typeRegressionSave = 'affine', typeRegression
        self.affBurnIn = 50
        self.coeffAff = 1.0
        self.fv0ori = -1 if self.fv0.surfVolume() < 0 else 1

        # affine part: coefficients on the basis and their penalty weights
        self.affine = affine
        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        w = affineWeight * np.ones(self.affineDim)
        for comp, val in ((self.affB.rotComp, rotWeight), (self.affB.simComp, scaleWeight),
                          (self.affB.transComp, transWeight)):
            if len(comp) > 0 and val is not None:
                w[comp] = val
        self.affineWeight = w
        self._w = rt.as_dev(w)
        d2 = self.dim * self.dim + self.dim
        self._basis = rt.as_dev(np.asarray(self.affB.basis, dtype=np.float64).reshape(d2, self.affineDim)
                                if self.affineDim > 0 else np.zeros((d2, 0)))

        # device state
        T, N, d = self.Tsize, self.npt, self.dim
        self.x0 = rt.as_dev(self.fv0.vertices)
        zeros = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=self.x0.device)
        self.a0 = zeros(N, d) if initialMomentum is None else rt.as_dev(np.asarray(initialMomentum, dtype=np.float64)).clone()
        self.rhot = zeros(T, N, d)
        self.Afft = zeros(T, self.affineDim)
        self.a0Try, self.rhotTry, self.AfftTry = self.a0, self.rhot, self.Afft
        self.obj = self.objTry = self.obj0 = None
        self.xt = self.at = None
        # the deformed template at the target times, as the data-term gradient reads it: the
        # template itself until the first objective (the reference's fvDef copies of fv0)
        self._def = [self.x0] * self.nTarg
        self._series = _TargetSeries(self.fv0.faces, self.x0, self.fv1, self.param.errorType,
                                     engine.DevKernel.of(self.param.KparDist),
                                     self.param.sigmaError)

        self.setOutputDir(outputDir)
        if self.outputDir is not None:
            self.fv0.saveVTK(os.path.join(self.outputDir, 'Template.vtk'))
            for k, s in enumerate(self.fv1):
                s.saveVTK(os.path.join(self.outputDir, 'Target' + str(k) + '.vtk'))

    def setOutputDir(self, outputDir):
        self.outputDir = outputDir
        if outputDir is not None and not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.makedirs(outputDir)

    # ---- model --------------------------------------------------------------------------
    def _transforms(self, Afft):
        """(A (T, d, d), b (T, d)): zeros without an affine part, which the evolution still
        takes as given (as the reference does)."""
        T, d = self.Tsize, self.dim
        if self.affineDim == 0:
            return [torch.zeros((T, d, d), dtype=torch.float64, device=self.x0.device),
                    torch.zeros((T, d), dtype=torch.float64, device=self.x0.device)]
        AB = Afft @ self._basis.T
        return [AB[:, :d * d].reshape(T, d, d).contiguous(), AB[:, d * d:].contiguous()]

    def _frames(self, xt):
        """The deformed template's vertices at the target times (k+1) T1."""
        return [xt[(k + 1) * self.Tsize1] for k in range(self.nTarg)]

    def _evaluate(self, a0, rhot, Afft, with_norms=False):
        """Shoot (a0, rhot, Afft) and return (xt, at, device scalars): [1/2 <a0, K a0>,
        dt cw sum |rho|^2 / 2, dt sum <w, Afft^2> / 2, data_0 .. data_{n-1}
        (, target norms)]."""
        n = self.nTarg
        scal = torch.zeros(3 + n + (n if with_norms else 0), dtype=torch.float64,
                           device=self.x0.device)
        xt, at = evol.secondOrderEvolution(self.x0, a0, rhot, self.param.KparDiff,
                                           affine=self._transforms(Afft))
        Ka0 = evol.apply_k_dev(self.x0, self.x0, a0, self.param.KparDiff)
        dt = 1.0 / self.Tsize
        scal[0] = 0.5 * (a0 * Ka0).sum()
        scal[1] = dt * self.controlWeight * (rhot * rhot).sum() / 2
        if self.affineDim > 0:
            scal[2] = dt * (self._w * Afft * Afft).sum() / 2
        self._series.energies(self._frames(xt), scal[3:3 + n])
        if with_norms:
            self._series.target_norms(scal[3 + n:])
        return xt, at, scal

    def _objective(self, vals):
        """(regularisation, data term / sigmaError^2) from the scalars of _evaluate."""
        n = self.nTarg
        s2 = self.param.sigmaError ** 2
        return vals[1] + vals[2] + vals[0], sum(vals[3:3 + n]) / s2

    # ---- optimiser protocol -------------------------------------------------------------
    def objectiveFun(self):
        if self.obj is None:
            self.xt, self.at, scal = self._evaluate(self.a0, self.rhot, self.Afft, with_norms=True)
            self._def = self._frames(self.xt)
            vals = scal.cpu().numpy()
            reg, data = self._objective(vals)
            self.obj0 = sum(vals[3 + self.nTarg:]) / self.param.sigmaError ** 2
            self.obj = reg + self.obj0 + data
        return self.obj

    def getVariable(self):
        return (self.a0, self.rhot, self.Afft)

    def updateTry(self, dir, eps, objRef=None):
        a0Try = self.a0 - eps * dir.a0
        rhotTry = self.rhot - eps * dir.rhot
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        _, _, scal = self._evaluate(a0Try, rhotTry, AfftTry)
        reg, data = self._objective(scal.cpu().numpy())
        objTry = self.obj0 + reg + data
        if np.isnan(objTry):
            logging.warning('nan in updateTry')
            return 1e500
        if objRef is None or objTry < objRef:
            self.a0Try, self.rhotTry, self.AfftTry, self.objTry = a0Try, rhotTry, AfftTry, objTry
        return objTry

    def getGradient(self, coeff=1.0):
        px1 = self._series.endpoint_covectors(self._def)
        pa1 = [torch.zeros_like(self.a0) for _ in range(self.nTarg)]
        g = evol.secondOrderGradient(self.x0, self.a0, self.rhot, px1, pa1, self.param.KparDiff,
                                     affine=self._transforms(self.Afft),
                                     times=(1 + np.arange(self.nTarg)) * self.Tsize1,
                                     controlWeight=self.controlWeight)
        da0, drho = g[0], g[1]
        kind = self.typeRegression
        a0_part = da0 / coeff if kind in ('geodesic', 'spline2') else torch.zeros_like(da0)
        rho_part = drho / coeff if kind in ('spline', 'spline2') else torch.zeros_like(drho)
        aff = torch.zeros_like(self.Afft)
        if self.affineDim > 0 and self.iter < self.affBurnIn:
            T, d = self.Tsize, self.dim
            dA, db = g[2], g[3]
            dAB = torch.cat([dA.reshape(T, d * d), db.reshape(T, d)], dim=1) @ self._basis
            aff = (self._w * self.Afft - dAB) / (self.coeffAff * coeff * T)
        return Direction(a0_part, rho_part, aff)

    def addProd(self, dir1, dir2, beta):
        return dir1.axpy(beta, dir2)

    def copyDir(self, dir0):
        return dir0.clone()

    def randomDir(self):
        """A random direction in the parts the current regression type varies (the affine
        part during the burn-in only)."""
        kind = self.typeRegression
        rnd = lambda *s: torch.randn(s, dtype=torch.float64, device=self.x0.device)
        a0 = rnd(self.npt, self.dim) if kind in ('geodesic', 'spline2') else torch.zeros_like(self.a0)
        rho = rnd(self.Tsize, self.npt, self.dim) if kind in ('spline', 'spline2') else torch.zeros_like(self.rhot)
        aff = rnd(self.Tsize, self.affineDim) if self.iter < self.affBurnIn else torch.zeros_like(self.Afft)
        return Direction(a0, rho, aff)

    def dotProduct(self, g1, g2):
        """Euclidean products, the rho part divided by T (secondOrderMatching.py:371-385)."""
        parts = torch.stack([torch.stack([(g.rhot * g1.rhot).sum(), (g.a0 * g1.a0).sum(),
                                          (g1.aff * g.aff).sum()]) for g in g2]).cpu().numpy()
        res = np.zeros(len(g2))
        for ll, (r, a, f) in enumerate(parts):
            res[ll] = r / self.Tsize
            res[ll] += a
            res[ll] += f * self.coeffAff
        return res

    def acceptVarTry(self):
        self.obj = self.objTry
        self.a0, self.rhot, self.Afft = self.a0Try, self.rhotTry, self.AfftTry

    def endOfIteration(self):
        """Regression type after the burn-in, the trajectory of the accepted variables and,
        every saveRate iterations, the deformed template at every time step with its
        Jacobian."""
        self.iter += 1
        if self.iter >= self.affBurnIn:
            self.typeRegression = self.typeRegressionSave
            self.affine = 'none'
        self.xt, self.at, scal = self._evaluate(self.a0, self.rhot, self.Afft)
        self._def = self._frames(self.xt)
        if self.verb:
            v = scal.cpu().numpy()
            logging.info('deformation terms: init %f, rho %f, aff %f' % (v[0], v[1], v[2]))
        if self.outputDir is not None and self.saveRate > 0 and self.iter % self.saveRate == 0:
            logging.info('Saving surfaces...')
            _, _, ft, Jt = evol.secondOrderEvolution(self.x0, self.a0, self.rhot,
                                                     self.param.KparDiff,
                                                     affine=self._transforms(self.Afft),
                                                     withPointSet=self.x0, withJacobian=True)
            ft, Jt = ft.cpu().numpy(), Jt.cpu().numpy()
            fvDef = surfaces.Surface(surf=self.fv0)
            for kk in range(self.Tsize + 1):
                fvDef.updateVertices(ft[kk])
                vf = surfaces.vtkFields()
                vf.scalars.append('Jacobian')
                vf.scalars.append(np.exp(Jt[kk]) - 1)
                fvDef.saveVTK2(os.path.join(self.outputDir, self.saveFile + str(kk) + '.vtk'), vf)

    def optimizeMatching(self):
        """secondOrderMatching.py:496-507: the gradient bound from the first gradient, cg
        re-initialised at the end of the affine burn-in."""
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        logging.info('Gradient lower bound: %f' % self.gradEps)
        self.cgBurnIn = self.affBurnIn
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient,
              epsInit=0.1)
