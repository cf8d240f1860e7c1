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

import logging

import numpy as np

from . import conjugateGradient as cg
from . import kernelFunctions as kfun
from . import surfaces
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

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, subsampleTargetSize=-1,
                 rotWeight=None, scaleWeight=None, transWeight=None, symmetric=False,
                 testGradient=True, saveFile='evolution', affine='none', outputDir='.'):
        if Template is None or Target is None:
            raise NotImplementedError("Template and Target surfaces are required (no file readers)")
        if symmetric or subsampleTargetSize > 0:
            raise NotImplementedError("the GPU path supports symmetric=False and no target "
                                      "subsampling")
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
