"""Surface matching, drop-in for py-lddmm's surfaceMatching module
(py-lddmm/surfaceMatching.py:19-484).

``SurfaceMatchingParam`` and ``SurfaceMatching`` keep the reference constructors and the
optimiser protocol; the registration lives on the GPU (see matching.DeviceMatching) and
``optimizeMatching`` runs the reference's conjugate gradient over it.  Supported: the
currents, measure and varifold data terms (errorType 'current', 'measure' (the default) or
'varifold'), the global affine part (affine 'affine', 'similitude', 'euclidean',
'translation' or 'none', with affineWeight / rotWeight / scaleWeight / transWeight),
symmetric=False, subsampleTargetSize=-1.  File output (VTK snapshots every saveRate iterations) is not written.
"""
from __future__ import annotations

import inspect
import logging
import os

import numpy as np

from . import conjugateGradient as cg
from . import kernelFunctions as kfun
from . import pointEvolution as evol
from . import surfaces
from .affineBasis import AffineBasis
from .matching import DeviceMatching, Direction  # noqa: F401  (Direction is part of the API)


class SurfaceMatchingParam:
    def __init__(self, timeStep=.1, KparDiff=None, KparDist=None, sigmaKernel=6.5, sigmaDist=2.5,
                 sigmaError=1.0, errorType='measure', typeKernel='gauss'):
        self.timeStep = timeStep
        self.sigmaKernel = sigmaKernel
        self.sigmaDist = sigmaDist
        self.sigmaError = sigmaError
        self.typeKernel = typeKernel
        self.errorType = errorType
        if errorType == 'current':
            self.fun_obj0 = surfaces.currentNorm0
            self.fun_obj = surfaces.currentNormDef
            self.fun_objGrad = surfaces.currentNormGradient
        elif errorType == 'measure':
            self.fun_obj0 = surfaces.measureNorm0
            self.fun_obj = surfaces.measureNormDef
            self.fun_objGrad = surfaces.measureNormGradient
        elif errorType == 'varifold':
            self.fun_obj0 = surfaces.varifoldNorm0
            self.fun_obj = surfaces.varifoldNormDef
            self.fun_objGrad = surfaces.varifoldNormGradient
        else:
            logging.error('Unknown error Type: ' + self.errorType)
        if KparDiff is None:
            self.KparDiff = kfun.Kernel(name=self.typeKernel, sigma=self.sigmaKernel)
        else:
            self.KparDiff = KparDiff
        if KparDist is None:
            self.KparDist = kfun.Kernel(name='gauss', sigma=self.sigmaDist)
        else:
            self.KparDist = KparDist


class SurfaceMatching(DeviceMatching):
    coef = 1.0 / 3.0

    def __new__(cls, *args, **kwargs):
        """symmetric=True selects SymmetricSurfaceMatching (the template's points are a
        variable too); everything else is the batched device driver."""
        if cls is SurfaceMatching:
            bound = inspect.signature(cls.__init__).bind_partial(None, *args, **kwargs)
            if bound.arguments.get('symmetric', False):
                return super().__new__(SymmetricSurfaceMatching)
        return super().__new__(cls)

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, subsampleTargetSize=-1,
                 rotWeight=None, scaleWeight=None, transWeight=None, symmetric=False,
                 testGradient=True, saveFile='evolution', affine='none', outputDir='.'):
        if Template is None or Target is None:
            raise NotImplementedError("Template and Target surfaces are required (no file readers)")
        if subsampleTargetSize > 0:
            raise NotImplementedError("template decimation (subsampleTargetSize) is not supported")
        self.fv0 = surfaces.Surface(surf=Template)
        self.fv1 = surfaces.Surface(surf=Target)
        self.param = SurfaceMatchingParam() if param is None else param
        if self.param.errorType not in ('current', 'measure', 'varifold'):
            raise NotImplementedError("the GPU path implements the currents, measure and "
                                      "varifold data terms")
        self.saveRate = 10
        self.outputDir = outputDir
        self.maxIter = maxIter
        self.verb = verb
        self.symmetric = symmetric
        self.testGradient = testGradient
        self.affine = affine
        self.saveFile = saveFile
        v0 = self.fv0.surfVolume()
        v1 = self.fv1.surfVolume()
        if v0 * v1 < 0:
            self.fv1.flipFaces()
        self.fvInit = surfaces.Surface(surf=self.fv0)
        self.fvDef = surfaces.Surface(surf=self.fvInit)
        x0 = np.copy(self.fvInit.vertices)
        Tsize = int(round(1.0 / self.param.timeStep))
        self._init_device(x0, self.fvInit.faces, self.fv1.centers, self.fv1.surfel,
                          self.param.KparDiff, self.param.KparDist, self.param.sigmaError, Tsize,
                          regWeight, data_term=self.param.errorType, affine=affine,
                          affineWeight=affineWeight, rotWeight=rotWeight, scaleWeight=scaleWeight,
                          transWeight=transWeight)

    def _new_shape(self, vertices, with_weights=False):
        return surfaces.DeviceSurface(self.faces_np, vertices, with_weights=with_weights)


    def endOptim(self):
        self.fvDef.updateVertices(self.xt[-1].cpu().numpy())

    def optimizeMatching(self):
        """surfaceMatching.py:474-484: the gradient bound is taken with coeffAff2 and the
        optimisation starts with coeffAff1."""
        self.coeffAff = self.coeffAff2
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        logging.info('Gradient lower bound: %f' % (self.gradEps))
        self.coeffAff = self.coeffAff1
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient,
              epsInit=0.1)

    def cg_steps(self):
        """The optimiser of optimizeMatching as a generator (see conjugateGradient.cg_steps);
        the initial gradient bound is one more request."""
        self.coeffAff = self.coeffAff2
        grd = yield (cg.GRADIENT, self.gradCoeff)
        [grd2] = yield (cg.DOT, grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        self.coeffAff = self.coeffAff1
        return (yield from cg.cg_steps(self, verb=self.verb, maxIter=self.maxIter,
                                       TestGradient=self.testGradient, epsInit=0.1))


class SymmetricSurfaceMatching(SurfaceMatching):
    """SurfaceMatching with symmetric=True (surfaceMatching.py:74-484): the template's points
    x0 are a variable too, and the data term adds the distance from the moved template to the
    original one, |S(x0) - S0|^2 / sigmaError^2 (its norm |S0|^2 in obj0).  The gradient in x0
    is (d|S(x0) - S0|^2 / dx0 / sigmaError^2 - p_0) / (coeffInitx coeff), p the covector of the
    flow, with coeffInitx = 0.1, which also weighs the x0 part of the dot product.

    Host orchestration in NumPy over the GPU drop-in functions (the Euler flow and its adjoint,
    applyK, the data terms), as the reference's own loop: the template moves with every trial,
    which the batched driver's fixed-x0 trajectory buffers do not serve.  It keeps the
    reference's schedule: the affine scaling coeffAff 1 -> 100 after affBurnIn = 20
    iterations (100 for the initial gradient bound), the affine branch of the flow taken with
    zero matrices when there is no affine part, and the moved template's surface refreshed in
    endOfIteration."""

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, subsampleTargetSize=-1,
                 rotWeight=None, scaleWeight=None, transWeight=None, symmetric=True,
                 testGradient=True, saveFile='evolution', affine='none', outputDir='.'):
        if subsampleTargetSize > 0:
            raise NotImplementedError("template decimation (subsampleTargetSize) is not supported")
        if Template is None:
            if fileTempl is None:
                raise ValueError('Please provide a template surface')
            self.fv0 = surfaces.Surface(filename=fileTempl)
        else:
            self.fv0 = surfaces.Surface(surf=Template)
        if Target is None:
            if fileTarg is None:
                raise ValueError('Please provide a target surface')
            self.fv1 = surfaces.Surface(filename=fileTarg)
        else:
            self.fv1 = surfaces.Surface(surf=Target)
        self.saveRate = 10
        self.iter = 0
        self.outputDir = outputDir
        if not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.makedirs(outputDir)
        self.dim = self.fv0.vertices.shape[1]
        self.maxIter = maxIter
        self.verb = verb
        self.symmetric = True
        self.testGradient = testGradient
        self.regweight = regWeight
        self.affine = affine
        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        self.affineBasis = self.affB.basis
        self.affineWeight = affineWeight * np.ones([self.affineDim, 1])
        if len(self.affB.rotComp) > 0 and rotWeight is not None:
            self.affineWeight[self.affB.rotComp] = rotWeight
        if len(self.affB.simComp) > 0 and scaleWeight is not None:
            self.affineWeight[self.affB.simComp] = scaleWeight
        if len(self.affB.transComp) > 0 and transWeight is not None:
            self.affineWeight[self.affB.transComp] = transWeight
        self.param = SurfaceMatchingParam() if param is None else param
        if self.fv0.surfVolume() * self.fv1.surfVolume() < 0:
            self.fv1.flipFaces()
        self.fvInit = surfaces.Surface(surf=self.fv0)
        self.x0 = np.copy(self.fvInit.vertices)
        self.x0Try = np.copy(self.fvInit.vertices)
        self.fvDef = surfaces.Surface(surf=self.fvInit)
        self.npt = self.fvInit.vertices.shape[0]
        self.fv0ori = -1 if self.fvInit.surfVolume() < 0 else 1
        self.fv1ori = -1 if self.fv1.surfVolume() < 0 else 1
        self.Tsize = int(round(1.0 / self.param.timeStep))
        self.at = np.zeros([self.Tsize, self.npt, self.dim])
        self.atTry = np.zeros_like(self.at)
        self.Afft = np.zeros([self.Tsize, self.affineDim])
        self.AfftTry = np.zeros_like(self.Afft)
        self.xt = np.tile(self.x0, [self.Tsize + 1, 1, 1])
        self.obj = None
        self.objTry = None
        self.gradCoeff = self.npt
        self.saveFile = saveFile
        self.fv0.saveVTK(os.path.join(self.outputDir, 'Template.vtk'))
        self.fv1.saveVTK(os.path.join(self.outputDir, 'Target.vtk'))
        self.coeffAff1 = 1.
        self.coeffAff2 = 100.
        self.coeffAff = self.coeffAff1
        self.coeffInitx = .1
        self.affBurnIn = 20

    def _affine(self, Afft):
        dim2 = self.dim ** 2
        A = [np.zeros([self.Tsize, self.dim, self.dim]), np.zeros([self.Tsize, self.dim])]
        if self.affineDim > 0:
            AB = Afft @ self.affineBasis.T
            A[0] = AB[:, :dim2].reshape(self.Tsize, self.dim, self.dim)
            A[1] = AB[:, dim2:dim2 + self.dim]
        return A

    def dataTerm(self, _fvDef, _fvInit=None):
        obj = self.param.fun_obj(_fvDef, self.fv1, self.param.KparDist) / (self.param.sigmaError ** 2)
        if _fvInit is not None:
            obj += self.param.fun_obj(_fvInit, self.fv0, self.param.KparDist) / (self.param.sigmaError ** 2)
        return obj

    def objectiveFunDef(self, at, Afft, withTrajectory=False, x0=None):
        """surfaceMatching.py:187-221."""
        x0 = self.x0 if x0 is None else x0
        timeStep = 1.0 / self.Tsize
        xt = evol.landmarkDirectEvolutionEuler(x0, at, self.param.KparDiff, affine=self._affine(Afft))
        obj = 0
        for t in range(self.Tsize):
            ra = self.param.KparDiff.applyK(xt[t], at[t])
            obj = obj + self.regweight * timeStep * np.multiply(at[t], ra).sum()
            if self.affineDim > 0:
                obj += timeStep * np.multiply(self.affineWeight.reshape(Afft[t].shape), Afft[t] ** 2).sum()
        return (obj, xt) if withTrajectory else obj

    def objectiveFun(self):
        if self.obj is None:
            self.obj0 = self.param.fun_obj0(self.fv1, self.param.KparDist) / (self.param.sigmaError ** 2)
            self.obj0 += self.param.fun_obj0(self.fv0, self.param.KparDist) / (self.param.sigmaError ** 2)
            self.obj, self.xt = self.objectiveFunDef(self.at, self.Afft, withTrajectory=True)
            self.fvDef.updateVertices(self.xt[-1])
            self.fvInit.updateVertices(self.x0)
            self.obj += self.obj0 + self.dataTerm(self.fvDef, self.fvInit)
        return self.obj

    def getVariable(self):
        return [self.at, self.Afft, self.x0]

    def updateTry(self, dir, eps, objRef=None):
        objTry = self.obj0
        atTry = self.at - eps * dir.diff
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        x0Try = self.x0 - eps * dir.initx
        obj, xt = self.objectiveFunDef(atTry, AfftTry, x0=x0Try, withTrajectory=True)
        objTry += obj
        ff = surfaces.Surface(surf=self.fvDef)
        ff.updateVertices(xt[-1])
        ffI = surfaces.Surface(surf=self.fvInit)
        ffI.updateVertices(x0Try)
        objTry += self.dataTerm(ff, ffI)
        if np.isnan(objTry):
            logging.info('Warning: nan in updateTry')
            return 1e500
        if objRef is None or objTry < objRef:
            self.atTry, self.objTry, self.AfftTry, self.x0Try = atTry, objTry, AfftTry, x0Try
        return objTry

    def endPointGradient(self):
        return self.param.fun_objGrad(self.fvDef, self.fv1, self.param.KparDist) / self.param.sigmaError ** 2

    def initPointGradient(self):
        return self.param.fun_objGrad(self.fvInit, self.fv0, self.param.KparDist) / self.param.sigmaError ** 2

    def getGradient(self, coeff=1.0):
        """surfaceMatching.py:301-327."""
        px1 = -self.endPointGradient()
        foo = evol.landmarkHamiltonianGradient(self.x0, self.at, px1, self.param.KparDiff,
                                               self.regweight, affine=self._affine(self.Afft),
                                               getCovector=True)
        grd = Direction()
        grd.diff = foo[0] / (coeff * self.Tsize)
        grd.aff = np.zeros(self.Afft.shape)
        if self.affineDim > 0:
            dA, db = foo[1], foo[2]
            dim2 = self.dim ** 2
            grd.aff = 2 * np.multiply(self.affineWeight.reshape([1, self.affineDim]), self.Afft)
            dAB = np.concatenate([dA.reshape(self.Tsize, dim2), db.reshape(self.Tsize, self.dim)], axis=1)
            grd.aff = grd.aff - dAB @ self.affineBasis
            grd.aff /= (self.coeffAff * coeff * self.Tsize)
        grd.initx = (self.initPointGradient() - foo[4][0]) / (self.coeffInitx * coeff)
        return grd

    def addProd(self, dir1, dir2, beta):
        d = Direction()
        d.diff = dir1.diff + beta * dir2.diff
        d.aff = dir1.aff + beta * dir2.aff
        d.initx = dir1.initx + beta * dir2.initx
        return d

    def copyDir(self, dir0):
        d = Direction()
        d.diff = np.copy(dir0.diff)
        d.aff = np.copy(dir0.aff)
        d.initx = np.copy(dir0.initx)
        return d

    def randomDir(self):
        d = Direction()
        d.diff = np.random.randn(self.Tsize, self.npt, self.dim)
        d.initx = np.random.randn(self.npt, self.dim)
        d.aff = np.random.randn(self.Tsize, self.affineDim)
        return d

    def dotProduct(self, g1, g2):
        """surfaceMatching.py:353-377: sum_t <g2_t, K(x_t) g1_t> (+ affine, x0 parts)."""
        res = np.zeros(len(g2))
        for t in range(self.Tsize):
            u = self.param.KparDiff.applyK(self.xt[t], g1.diff[t])
            uu = g1.aff[t]
            for ll, gr in enumerate(g2):
                res[ll] = res[ll] + np.multiply(gr.diff[t], u).sum()
                if self.affineDim > 0:
                    res[ll] += np.multiply(uu, gr.aff[t]).sum() * self.coeffAff
        for ll, gr in enumerate(g2):
            res[ll] += (g1.initx * gr.initx).sum() * self.coeffInitx
        return res

    def acceptVarTry(self):
        self.obj = self.objTry
        self.at = np.copy(self.atTry)
        self.Afft = np.copy(self.AfftTry)
        self.x0 = np.copy(self.x0Try)

    def endOfIteration(self):
        """surfaceMatching.py:386-471: coeffAff after the burn-in, the trajectory, the deformed
        and the moved template refreshed (every saveRate iterations the deformed template is
        also written)."""
        self.iter += 1
        if self.iter >= self.affBurnIn:
            self.coeffAff = self.coeffAff2
        obj1, self.xt = self.objectiveFunDef(self.at, self.Afft, withTrajectory=True)
        self.fvDef.updateVertices(self.xt[-1])
        self.fvInit.updateVertices(self.x0)
        if self.iter % self.saveRate == 0:
            logging.info('Saving surfaces...')
            self.fvDef.saveVTK(os.path.join(self.outputDir, self.saveFile + '.vtk'))

    def endOptim(self):
        self.fvDef.updateVertices(self.xt[-1])

    def optimizeMatching(self):
        self.coeffAff = self.coeffAff2
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        logging.info('Gradient lower bound: %f' % self.gradEps)
        self.coeffAff = self.coeffAff1
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient,
              epsInit=0.1)

    def cg_steps(self):
        raise NotImplementedError("symmetric matching runs through optimizeMatching (it is not "
                                  "part of the batched service)")
