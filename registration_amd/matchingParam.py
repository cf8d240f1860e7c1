"""Matching parameters, drop-in for py-lddmm's matchingParam.MatchingParam
(py-lddmm/matchingParam.py:10-25)."""
from __future__ import annotations

from . import kernelFunctions as kfun


class MatchingParam:
    """timeStep: time discretisation; KparDiff: deformation kernel (default typeKernel of width
    sigmaKernel); KparDist: data-term kernel (default Gaussian of width sigmaDist); sigmaError:
    data-term normalisation; errorType: 'measure' (the default) or 'current'; typeKernel:
    'gauss' or 'laplacian'."""

    def __init__(self, timeStep=.1, KparDiff=None, KparDist=None, sigmaKernel=6.5, sigmaDist=2.5,
                 sigmaError=1.0, errorType='measure', typeKernel='gauss'):
        self.timeStep = timeStep
        self.sigmaKernel = sigmaKernel
        self.sigmaDist = sigmaDist
        self.sigmaError = sigmaError
        self.typeKernel = typeKernel
        self.errorType = errorType
        if KparDiff is None:
            self.KparDiff = kfun.Kernel(name=self.typeKernel, sigma=self.sigmaKernel)
        else:
            self.KparDiff = KparDiff
        if KparDist is None:
            self.KparDist = kfun.Kernel(name='gauss', sigma=self.sigmaDist)
        else:
            self.KparDist = KparDist
