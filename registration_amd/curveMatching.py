"""Curve matching, drop-in for py-lddmm's curveMatching module
(py-lddmm/curveMatching.py:19-352).

``CurveMatchingParam`` and ``CurveMatching`` keep the reference constructors and optimiser
protocol on a GPU-resident registration (matching.DeviceMatching).  Supported: the currents
and measure data terms (errorType 'current' or 'measure', the default) and the global affine
part (affine 'affine', 'similitude', 'euclidean', 'translation' or 'none'; the curve driver
weighs its gradient and products by affineWeight, not coeffAff).  The per-iteration Jacobian flow and the VTK snapshots of
endOfIteration / endOptim only produce output files and are not run.
"""
from __future__ import annotations

import numpy as np

from . import conjugateGradient as cg
from . import curves
from . import matchingParam
from .matching import DeviceMatching, Direction  # noqa: F401


class CurveMatchingParam(matchingParam.MatchingParam):
    def __init__(self, timeStep=.1, KparDiff=None, KparDist=None, sigmaKernel=6.5, sigmaDist=2.5,
                 sigmaError=1.0, errorType='measure', typeKernel='gauss'):
        matchingParam.MatchingParam.__init__(self, timeStep=timeStep, KparDiff=KparDiff,
                                             KparDist=KparDist, sigmaKernel=sigmaKernel,
                                             sigmaDist=sigmaDist, sigmaError=sigmaError,
                                             errorType=errorType, typeKernel=typeKernel)
        if errorType == 'current':
            self.fun_obj0 = curves.currentNorm0
            self.fun_obj = curves.currentNormDef
            self.fun_objGrad = curves.currentNormGradient
        elif errorType == 'measure':
            self.fun_obj0 = curves.measureNorm0
            self.fun_obj = curves.measureNormDef
            self.fun_objGrad = curves.measureNormGradient
        else:
            self.fun_obj0 = self.fun_obj = self.fun_objGrad = None


class CurveMatching(DeviceMatching):
    coef = 0.5
    measure_norm0_eps = 0.0   # curves.measureNorm0 weights are |lenel| (curves.py:624)
    aff_style = 'curve'

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, gradLB=0.001,
                 saveRate=10, rotWeight=None, scaleWeight=None, transWeight=None,
                 testGradient=False, saveFile='evolution', affine='none', outputDir='.'):
        if Template is None or Target is None:
            raise NotImplementedError("Template and Target curves are required (no file readers)")
        self.fv0 = curves.Curve(curve=Template)
        self.fv1 = curves.Curve(curve=Target)
        self.param = CurveMatchingParam() if param is None else param
        if self.param.errorType not in ('current', 'measure'):
            raise NotImplementedError("the GPU path implements the currents and measure data "
                                      "terms (errorType 'current' or 'measure')")
        self.fvDef = curves.Curve(curve=self.fv0)
        self.outputDir = outputDir
        self.maxIter = maxIter
        self.verb = verb
        self.testGradient = testGradient
        self.affine = affine
        self.saveFile = saveFile
        self.gradLB = gradLB
        self.saveRate = saveRate
        Tsize = int(round(1.0 / self.param.timeStep))
        self._init_device(np.copy(self.fv0.vertices), self.fv0.faces, self.fv1.centers,
                          self.fv1.lenel, self.param.KparDiff, self.param.KparDist,
                          self.param.sigmaError, Tsize, regWeight,
                          data_term=self.param.errorType, affine=affine,
                          affineWeight=affineWeight, rotWeight=rotWeight, scaleWeight=scaleWeight,
                          transWeight=transWeight)

    def _new_shape(self, vertices, with_weights=False):
        return curves.DeviceCurve(self.faces_np, vertices, with_weights=with_weights)


    def endOptim(self):
        self.fvDef.updateVertices(self.xt[-1].cpu().numpy())
        self.defCost = None

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient)

    def cg_steps(self):
        grd = yield (cg.GRADIENT, self.gradCoeff)
        [grd2] = yield (cg.DOT, grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        return (yield from cg.cg_steps(self, verb=self.verb, maxIter=self.maxIter,
                                       TestGradient=self.testGradient))
