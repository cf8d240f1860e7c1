"""Second-order (spline) surface regression through a time series of targets: the drop-in for
py-lddmm/secondOrderMatching.py (SurfaceMatching, secondOrderMatching.py:43-509).

The variables are the initial momentum a0 (N, d), the controls rhot (T, N, d) and the affine
coefficients Afft (T, affineDim); the state follows x' = K(x) a (+ A x + b), a' = -K'(x)(a, a)
+ rho (- a A) (pointEvolution.secondOrderEvolution, on the GPU).  The objective is
1/2 <a0, K(x0) a0> + sum_t dt cw |rho_t|^2 / 2 + sum_t dt <affineWeight, Afft_t^2> / 2 plus the
data terms of the deformed template at times (k+1) T1 against targets k, T = nTarg T1
(secondOrderMatching.py:183-246).  The optimiser is conjugateGradient.cg with the reference's
protocol; gradients come from pointEvolution.secondOrderGradient (GPU adjoint).

Kept as the reference has it (so that a user of the reference finds the same optimisation):
the regression type is 'affine' (zero a0 and rhot gradients) for the first affBurnIn = 50
iterations and then the requested type ('spline2' by default: both a0 and rhot); the affine
branch of the evolution always runs, with zero matrices when there is no affine part; dot
products are Euclidean (rho part divided by T, affine part times coeffAff).
"""
from __future__ import annotations

import logging
import os

import numpy as np

from . import conjugateGradient as cg
from . import pointEvolution as evol
from . import surfaces
from .affineBasis import AffineBasis
from .surfaceMatching import SurfaceMatchingParam


class Direction:
    """secondOrderMatching.py:21-25."""

    def __init__(self):
        self.a0 = []
        self.rhot = []
        self.aff = []


class SurfaceMatching:
    """secondOrderMatching.SurfaceMatching (secondOrderMatching.py:43-509).

    Template: surface; Targets: list of surfaces (or fileTempl / fileTarg file names); param:
    SurfaceMatchingParam (kernels, sigmaError, errorType, timeStep = 1/T1); initialMomentum:
    a0 to start from; typeRegression: 'spline' (rho only), 'geodesic' (a0 only), 'affine'
    (neither) or 'spline2' (both, the default) once the affine burn-in is over;
    controlWeight: weight of |rho|^2; affine and its weights as in surfaceMatching.
    subsampleTargetSize > 0 (mesh decimation of the template) is not supported."""

    def __init__(self, Template=None, Targets=None, fileTempl=None, fileTarg=None, param=None,
                 initialMomentum=None, maxIter=1000, regWeight=1.0, verb=True,
                 typeRegression='spline2', affine='none', controlWeight=1.0, affineWeight=1.0,
                 rotWeight=None, scaleWeight=None, transWeight=None, subsampleTargetSize=-1,
                 testGradient=True, saveFile='evolution', outputDir='.'):
        if Template is None:
            if fileTempl is None:
                raise ValueError('Please provide a template surface')
            self.fv0 = surfaces.Surface(filename=fileTempl)
        else:
            self.fv0 = surfaces.Surface(surf=Template)
        if Targets is None:
            if fileTarg is None:
                raise ValueError('Please provide a list of target surfaces')
            self.fv1 = [surfaces.Surface(filename=f) for f in fileTarg]
        else:
            self.fv1 = [surfaces.Surface(surf=s) for s in Targets]
        if subsampleTargetSize > 0:
            raise NotImplementedError("template decimation (subsampleTargetSize) is not supported")

        self.nTarg = len(self.fv1)
        self.saveRate = 10
        self.iter = 0
        self.outputDir = outputDir
        if not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.mkdir(outputDir)
        self.dim = self.fv0.vertices.shape[1]
        self.maxIter = maxIter
        self.verb = verb
        self.testGradient = testGradient
        self.regweight = regWeight
        self.affine = affine
        self.saveCorrected = affine in ('euclidean', 'translation')

        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        self.affineBasis = self.affB.basis
        self.affineWeight = affineWeight * np.ones(self.affineDim)
        if len(self.affB.rotComp) > 0 and rotWeight is not None:
            self.affineWeight[self.affB.rotComp] = rotWeight
        if len(self.affB.simComp) > 0 and scaleWeight is not None:
            self.affineWeight[self.affB.simComp] = scaleWeight
        if len(self.affB.transComp) > 0 and transWeight is not None:
            self.affineWeight[self.affB.transComp] = transWeight
        self.affw = affineWeight

        self.controlWeight = controlWeight
        self.typeRegression = 'affine'
        self.typeRegressionSave = typeRegression
        self.param = SurfaceMatchingParam() if param is None else param

        self.fv0Fine = surfaces.Surface(surf=self.fv0)
        self.x0 = self.fv0.vertices
        self.fvDef = [surfaces.Surface(surf=self.fv0) for _ in range(self.nTarg)]
        self.npt = self.x0.shape[0]
        self.Tsize1 = int(round(1.0 / self.param.timeStep))
        self.Tsize = self.nTarg * self.Tsize1
        self.rhot = np.zeros([self.Tsize, self.x0.shape[0], self.x0.shape[1]])
        if initialMomentum is None:
            self.xt = np.tile(self.x0, [self.Tsize + 1, 1, 1])
            self.a0 = np.zeros([self.x0.shape[0], self.x0.shape[1]])
            self.at = np.tile(self.a0, [self.Tsize + 1, 1, 1])
        else:
            self.a0 = np.asarray(initialMomentum, dtype=np.float64)
            self.xt, self.at = evol.secondOrderEvolution(self.x0, self.a0, self.rhot,
                                                         self.param.KparDiff)

        self.v = np.zeros([self.Tsize + 1, self.npt, self.dim])
        self.rhotTry = np.zeros([self.Tsize, self.x0.shape[0], self.x0.shape[1]])
        self.a0Try = np.zeros([self.x0.shape[0], self.x0.shape[1]])
        self.Afft = np.zeros([self.Tsize, self.affineDim])
        self.AfftTry = np.zeros([self.Tsize, self.affineDim])

        self.obj = None
        self.objTry = None
        self.gradCoeff = self.fv0.vertices.shape[0]
        self.saveFile = saveFile
        self.fv0.saveVTK(os.path.join(self.outputDir, 'Template.vtk'))
        for k, s in enumerate(self.fv1):
            s.saveVTK(os.path.join(self.outputDir, 'Target' + str(k) + '.vtk'))
        self.affBurnIn = 50
        self.coeffAff1 = 1
        self.coeffAff2 = 10.
        self.coeffAff = self.coeffAff1
        self.fv0ori = -1 if self.fv0.surfVolume() < 0 else 1

    def setOutputDir(self, outputDir):
        self.outputDir = outputDir
        if not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.makedirs(outputDir)

    # ---- objective ----------------------------------------------------------------------
    def _affine(self, Afft):
        """(A (T, d, d), b (T, d)) from the coefficients on the affine basis; zeros without an
        affine part (the evolution still takes its affine branch)."""
        dim2 = self.dim ** 2
        A = [np.zeros([self.Tsize, self.dim, self.dim]), np.zeros([self.Tsize, self.dim])]
        if self.affineDim > 0:
            AB = Afft @ self.affineBasis.T
            A[0] = AB[:, :dim2].reshape(self.Tsize, self.dim, self.dim)
            A[1] = AB[:, dim2:dim2 + self.dim]
        return A

    def dataTerm(self, _fvDef):
        obj = 0
        for k, s in enumerate(_fvDef):
            obj += self.param.fun_obj(s, self.fv1[k], self.param.KparDist) / (self.param.sigmaError ** 2)
        return obj

    def objectiveFunDef(self, a0, rhot, Afft, withTrajectory=False, withJacobian=False, Init=None,
                        display=False):
        """secondOrderMatching.py:183-222."""
        x0 = self.x0 if Init is None else Init[0]
        timeStep = 1.0 / self.Tsize
        A = self._affine(Afft)
        if withJacobian:
            xt, at, Jt = evol.secondOrderEvolution(x0, a0, rhot, self.param.KparDiff,
                                                   withJacobian=True, affine=A)
        else:
            xt, at = evol.secondOrderEvolution(x0, a0, rhot, self.param.KparDiff, affine=A)
        obj0 = 0.5 * (a0 * self.param.KparDiff.applyK(x0, a0)).sum()
        obj1 = timeStep * self.controlWeight * (rhot ** 2).sum() / 2
        obj2 = 0
        if self.affineDim > 0:
            obj2 = timeStep * (self.affineWeight[None, :] * Afft ** 2).sum() / 2
        obj = obj1 + obj2 + obj0
        if display:
            logging.info('deformation terms: init %f, rho %f, aff %f' % (obj0, obj1, obj2))
        if withJacobian:
            return obj, xt, at, Jt
        if withTrajectory:
            return obj, xt, at
        return obj

    def objectiveFun(self):
        """secondOrderMatching.py:235-246."""
        if self.obj is None:
            self.obj, self.xt, self.at = self.objectiveFunDef(self.a0, self.rhot, self.Afft,
                                                              withTrajectory=True)
            self.obj0 = 0
            for k in range(self.nTarg):
                self.obj0 += self.param.fun_obj0(self.fv1[k], self.param.KparDist) / (self.param.sigmaError ** 2)
                self.fvDef[k].updateVertices(np.squeeze(self.xt[(k + 1) * self.Tsize1, :, :]))
            self.obj += self.obj0 + self.dataTerm(self.fvDef)
        return self.obj

    def getVariable(self):
        return (self.a0, self.rhot, self.Afft)

    def updateTry(self, dir, eps, objRef=None):
        """secondOrderMatching.py:251-281."""
        objTry = self.obj0
        a0Try = self.a0 - eps * dir.a0
        rhotTry = self.rhot - eps * dir.rhot
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        foo = self.objectiveFunDef(a0Try, rhotTry, AfftTry, withTrajectory=True)
        objTry += foo[0]
        ff = []
        for k in range(self.nTarg):
            ff.append(surfaces.Surface(surf=self.fvDef[k]))
            ff[k].updateVertices(np.squeeze(foo[1][(k + 1) * self.Tsize1, :, :]))
        objTry += self.dataTerm(ff)
        if np.isnan(objTry):
            logging.warning('nan in updateTry')
            return 1e500
        if objRef is None or objTry < objRef:
            self.a0Try = a0Try
            self.rhotTry = rhotTry
            self.AfftTry = AfftTry
            self.objTry = objTry
        return objTry

    # ---- gradient -----------------------------------------------------------------------
    def endPointGradient(self):
        return [-self.param.fun_objGrad(self.fvDef[k], self.fv1[k], self.param.KparDist)
                / self.param.sigmaError ** 2 for k in range(self.nTarg)]

    def getGradient(self, coeff=1.0):
        """secondOrderMatching.py:290-333: the adjoint gives (da0, drhot, dA, db); the regression
        type selects which of da0, drhot enter the direction; the affine part during the
        burn-in only."""
        px1 = self.endPointGradient()
        pa1 = [np.zeros(self.a0.shape) for _ in range(self.nTarg)]
        A = self._affine(self.Afft)
        foo = evol.secondOrderGradient(self.x0, self.a0, self.rhot, px1, pa1, self.param.KparDiff,
                                       affine=A,
                                       times=(1 + np.arange(self.nTarg)) * self.Tsize1,
                                       controlWeight=self.controlWeight)
        grd = Direction()
        if self.typeRegression == 'spline':
            grd.a0 = np.zeros(foo[0].shape)
            grd.rhot = foo[1] / coeff
        elif self.typeRegression == 'geodesic':
            grd.a0 = foo[0] / coeff
            grd.rhot = np.zeros(foo[1].shape)
        elif self.typeRegression == 'affine':
            grd.a0 = np.zeros(foo[0].shape)
            grd.rhot = np.zeros(foo[1].shape)
        else:
            grd.a0 = foo[0] / coeff
            grd.rhot = foo[1] / coeff
        grd.aff = np.zeros(self.Afft.shape)
        if self.affineDim > 0 and self.iter < self.affBurnIn:
            dA, db = foo[2], foo[3]
            dim2 = self.dim ** 2
            grd.aff = self.affineWeight.reshape([1, self.affineDim]) * self.Afft
            dAB = np.concatenate([dA.reshape(self.Tsize, dim2), db.reshape(self.Tsize, self.dim)], axis=1)
            grd.aff = grd.aff - dAB @ self.affineBasis
            grd.aff /= (self.coeffAff * coeff * self.Tsize)
        return grd

    # ---- cg protocol --------------------------------------------------------------------
    def addProd(self, dir1, dir2, beta):
        dir = Direction()
        dir.a0 = dir1.a0 + beta * dir2.a0
        dir.rhot = dir1.rhot + beta * dir2.rhot
        dir.aff = dir1.aff + beta * dir2.aff
        return dir

    def copyDir(self, dir0):
        dir = Direction()
        dir.a0 = np.copy(dir0.a0)
        dir.rhot = np.copy(dir0.rhot)
        dir.aff = np.copy(dir0.aff)
        return dir

    def randomDir(self):
        dirfoo = Direction()
        zero_a0 = np.zeros([self.npt, self.dim])
        zero_rho = np.zeros([self.Tsize, self.npt, self.dim])
        if self.typeRegression == 'spline':
            dirfoo.a0, dirfoo.rhot = zero_a0, np.random.randn(self.Tsize, self.npt, self.dim)
        elif self.typeRegression == 'geodesic':
            dirfoo.a0, dirfoo.rhot = np.random.randn(self.npt, self.dim), zero_rho
        elif self.typeRegression == 'affine':
            dirfoo.a0, dirfoo.rhot = zero_a0, zero_rho
        else:
            dirfoo.a0 = np.random.randn(self.npt, self.dim)
            dirfoo.rhot = np.random.randn(self.Tsize, self.npt, self.dim)
        # The reference leaves aff unset after the burn-in; a zero part keeps addProd defined.
        dirfoo.aff = (np.random.randn(self.Tsize, self.affineDim) if self.iter < self.affBurnIn
                      else np.zeros([self.Tsize, self.affineDim]))
        return dirfoo

    def dotProduct(self, g1, g2):
        """secondOrderMatching.py:371-385: Euclidean, rho part divided by T."""
        res = np.zeros(len(g2))
        for ll, gr in enumerate(g2):
            res[ll] = (gr.rhot * g1.rhot).sum() / self.Tsize
            res[ll] += (gr.a0 * g1.a0).sum()
            res[ll] += (g1.aff * gr.aff).sum() * self.coeffAff
        return res

    def acceptVarTry(self):
        self.obj = self.objTry
        self.a0 = np.copy(self.a0Try)
        self.rhot = np.copy(self.rhotTry)
        self.Afft = np.copy(self.AfftTry)

    def endOfIteration(self):
        """secondOrderMatching.py:394-494: after the burn-in the requested regression type
        takes over and the affine part is dropped; the trajectory and deformed surfaces are
        refreshed, and every saveRate iterations the deformed template at each time step is
        written with its Jacobian (the reference's affine-corrected copies are not written)."""
        self.iter += 1
        if self.iter >= self.affBurnIn:
            self.typeRegression = self.typeRegressionSave
            self.affine = 'none'
        obj1, self.xt, self.at = self.objectiveFunDef(self.a0, self.rhot, self.Afft,
                                                      withTrajectory=True, display=True)
        for k in range(self.nTarg):
            self.fvDef[k].updateVertices(np.squeeze(self.xt[(k + 1) * self.Tsize1, :, :]))
        if self.iter % self.saveRate == 0:
            logging.info('Saving surfaces...')
            A = self._affine(self.Afft)
            _, _, ft, Jt = evol.secondOrderEvolution(self.x0, self.a0, self.rhot,
                                                     self.param.KparDiff, affine=A,
                                                     withPointSet=self.fv0Fine.vertices,
                                                     withJacobian=True)
            fvDef = surfaces.Surface(surf=self.fv0Fine)
            for kk in range(self.Tsize + 1):
                fvDef.updateVertices(np.squeeze(ft[kk, :, :]))
                vf = surfaces.vtkFields()
                vf.scalars.append('Jacobian')
                vf.scalars.append(np.exp(Jt[kk, :]) - 1)
                fvDef.saveVTK2(os.path.join(self.outputDir, self.saveFile + str(kk) + '.vtk'), vf)

    def optimizeMatching(self):
        """secondOrderMatching.py:496-507."""
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        logging.info('Gradient lower bound: %f' % self.gradEps)
        self.cgBurnIn = self.affBurnIn
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient,
              epsInit=0.1)
