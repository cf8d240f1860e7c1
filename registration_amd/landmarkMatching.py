"""Landmark-set matching on the GPU path.

py-lddmm has no landmark driver; its C++ sibling matches landmark sets with the squared-distance
objective of LandmarkObjective (diffeoReg/landmarkMatching.h:49-70).  This module gives that
data term the SurfaceMatching call pattern (surfaceMatching.py:74-484), so landmark
registrations use the same optimiser (conjugateGradient.cg), the same device-resident state
and the same batched requests as surfaces and curves:

    E(a) = sum_t lambda dt <a_t, K(x_t) a_t>  +  sum_k |x_T,k - y_k|^2 / sigmaError^2

with the Euler flow of landmarkDirectEvolutionEuler (pointEvolution.py:666-730).  The
gradient is the exact gradient of E: p_T = -2 (x_T - y) / sigmaError^2 enters
landmarkHamiltonianCovector (pointEvolution.py:735-765).  The C++ objective's gradient,
(x - y) / N, is a rescaled direction for its own optimiser, not the derivative of its
energy, and is not used.
"""
from __future__ import annotations

import numpy as np

from . import conjugateGradient as cg
from . import matchingParam
from .matching import DeviceMatching, Direction  # noqa: F401


class LandmarkMatchingParam(matchingParam.MatchingParam):
    """MatchingParam with the landmark data term (errorType='landmark'); KparDist is unused."""

    def __init__(self, timeStep=.1, KparDiff=None, sigmaKernel=6.5, sigmaError=1.0,
                 typeKernel='gauss'):
        matchingParam.MatchingParam.__init__(self, timeStep=timeStep, KparDiff=KparDiff,
                                             KparDist=None, sigmaKernel=sigmaKernel,
                                             sigmaDist=sigmaKernel, sigmaError=sigmaError,
                                             errorType='landmark', typeKernel=typeKernel)


class LandmarkMatching(DeviceMatching):
    """Template and Target are (N, d) arrays of corresponding points (d = 2 or 3)."""

    def __init__(self, Template=None, Target=None, param=None, maxIter=1000, regWeight=1.0,
                 verb=True, gradLB=0.001, testGradient=False, epsInit=0.1):
        if Template is None or Target is None:
            raise ValueError("Template and Target point sets are required")
        x0 = np.array(Template, dtype=np.float64)
        y = np.array(Target, dtype=np.float64)
        if x0.ndim != 2 or x0.shape != y.shape or x0.shape[1] not in (2, 3):
            raise ValueError("Template and Target must be (N, d) arrays of equal shape, d = 2 or 3")
        self.param = LandmarkMatchingParam() if param is None else param
        self.maxIter = maxIter
        self.verb = verb
        self.testGradient = testGradient
        self.gradLB = gradLB
        self.epsInit = epsInit
        self.Template, self.Target = x0, y
        Tsize = int(round(1.0 / self.param.timeStep))
        self._init_device(x0, None, None, None, self.param.KparDiff, None,
                          self.param.sigmaError, Tsize, regWeight, target_points=y)

    def endOptim(self):
        pass

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient,
              epsInit=self.epsInit)

    def cg_steps(self):
        grd = yield (cg.GRADIENT, self.gradCoeff)
        [grd2] = yield (cg.DOT, grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        return (yield from cg.cg_steps(self, verb=self.verb, maxIter=self.maxIter,
                                       TestGradient=self.testGradient, epsInit=self.epsInit))
