"""Surface matching, drop-in for py-lddmm's surfaceMatching module
(py-lddmm/surfaceMatching.py:19-484).

``SurfaceMatchingParam`` and ``SurfaceMatching`` keep the reference constructors and the
optimiser protocol; the registration lives on the GPU (see matching.DeviceMatching) and
``optimizeMatching`` runs the reference's conjugate gradient over it.  Supported: the
currents, measure and varifold data terms (errorType 'current', 'measure' (the default) or
'varifold'), the global affine part (affine 'affine', 'similitude', 'euclidean',
'translation' or 'none', with affineWeight / rotWeight / scaleWeight / transWeight),
symmetric matching (the template's points a variable too), templates and targets given as
surfaces or shape files; subsampleTargetSize=-1 only.
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


def _shape_arg(surface, filename, what):
    """A host Surface from a Surface argument or a shape file (surfaceMatching.py:81-96)."""
    if surface is not None:
        return surfaces.Surface(surf=surface)
    if filename is None:
        raise ValueError('Please provide a %s surface' % what)
    return surfaces.Surface(filename=filename)


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
        """symmetric=True gives a SymmetricSurfaceMatching (the same device driver with the
        template's points as a variable too)."""
        if cls is SurfaceMatching:
            bound = inspect.signature(cls.__init__).bind_partial(None, *args, **kwargs)
            if bound.arguments.get('symmetric', False):
                return super().__new__(SymmetricSurfaceMatching)
        return super().__new__(cls)

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, subsampleTargetSize=-1,
                 rotWeight=None, scaleWeight=None, transWeight=None, symmetric=False,
                 testGradient=True, saveFile='evolution', affine='none', outputDir='.'):
        """surfaceMatching.py:76-169.  Template / Target surfaces, or files (BYU, OFF, VTK, OBJ)
        read through fileio.  The reference writes Template.vtk and Target.vtk into outputDir
        at construction and the deformed template every saveRate iterations;
        ``outputDir=None`` (an addition) writes nothing."""
        self.fv0 = _shape_arg(Template, fileTempl, 'template')
        self.fv1 = _shape_arg(Target, fileTarg, 'target')
        if subsampleTargetSize > 0:
            raise NotImplementedError("template decimation (subsampleTargetSize) is not supported")
        self.param = SurfaceMatchingParam() if param is None else param
        if self.param.errorType not in ('current', 'measure', 'varifold'):
            raise NotImplementedError("the GPU path implements the currents, measure and "
                                      "varifold data terms")
        self.saveRate = 10
        self.setOutputDir(outputDir)
        self.maxIter = maxIter
        self.verb = verb
        self.symmetric = bool(symmetric)
        self.testGradient = testGradient
        self.affine = affine
        self.saveFile = saveFile
        v0 = self.fv0.surfVolume()
        v1 = self.fv1.surfVolume()
        if v0 * v1 < 0:
            self.fv1.flipFaces()
        self.fvInit = surfaces.Surface(surf=self.fv0)
        self.fvDef = surfaces.Surface(surf=self.fvInit)
        self.fv0ori = -1 if self.fvInit.surfVolume() < 0 else 1
        self.fv1ori = -1 if self.fv1.surfVolume() < 0 else 1
        x0 = np.copy(self.fvInit.vertices)
        Tsize = int(round(1.0 / self.param.timeStep))
        self._init_device(x0, self.fvInit.faces, self.fv1.centers, self.fv1.surfel,
                          self.param.KparDiff, self.param.KparDist, self.param.sigmaError, Tsize,
                          regWeight, data_term=self.param.errorType, affine=affine,
                          affineWeight=affineWeight, rotWeight=rotWeight, scaleWeight=scaleWeight,
                          transWeight=transWeight,
                          template_geometry=((self.fv0.centers, self.fv0.surfel)
                                             if self.symmetric else None))
        if self.outputDir is not None:
            self.fv0.saveVTK(os.path.join(self.outputDir, 'Template.vtk'))
            self.fv1.saveVTK(os.path.join(self.outputDir, 'Target.vtk'))

    def setOutputDir(self, outputDir):
        """surfaceMatching.py:171-178: create the directory if it does not exist."""
        self.outputDir = outputDir
        if outputDir is None:
            return
        if not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.makedirs(outputDir)

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
    """SurfaceMatching with symmetric=True (surfaceMatching.py:74-484): the template's points x0
    are a variable too and the data term adds the distance from the moved template to the
    template as given (matching._TemplateTerm).  Same device driver, batched requests and
    optimiser as the plain form; the x0 parts of directions are device tensors (``initx``)."""

    def __init__(self, *args, **kwargs):
        bound = inspect.signature(SurfaceMatching.__init__).bind(None, *args, **kwargs)
        bound.arguments['symmetric'] = True
        super().__init__(*list(bound.args)[1:], **bound.kwargs)
