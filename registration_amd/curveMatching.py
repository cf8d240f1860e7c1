"""Curve matching, drop-in for py-lddmm's curveMatching module
(py-lddmm/curveMatching.py:19-352).

``CurveMatchingParam`` and ``CurveMatching`` keep the reference constructors and optimiser
protocol on a GPU-resident registration (matching.DeviceMatching).  Supported: the currents
and measure data terms (errorType 'current' or 'measure', the default), the global affine
part (affine 'affine', 'similitude', 'euclidean', 'translation' or 'none'; the curve driver
weighs its gradient and products by affineWeight, not coeffAff), templates and targets given
as curves or files, and the reference's output files (Template / Target, and every saveRate
iterations the deformed curve at every step with, for planar curves, the deformed grid);
``outputDir=None`` writes nothing.
"""
from __future__ import annotations

import logging
import os

import numpy as np

from . import conjugateGradient as cg
from . import curves
from . import fileio
from . import matchingParam
from . import pointEvolution as evol
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


def _curve_arg(curve, filename, what):
    """A host Curve from a Curve argument or a curve file (curveMatching.py:63-79)."""
    if curve is not None:
        return curves.Curve(curve=curve)
    if filename is None:
        raise ValueError('Please provide a %s curve' % what)
    return curves.Curve(filename=filename)


def deformation_grid(v0, v1):
    """The planar grid the curve driver deforms for its output (curveMatching.py:84-96): a
    square lattice over the two curves' bounding box with spacing 2.5% of its smaller side and
    a 10-cell margin, and its edges (neighbours along each row, then along each column, as
    grid.Grid builds them).  Returns (points (n m, 2), edges (E, 2))."""
    lo = np.minimum(v0.min(axis=0), v1.min(axis=0))
    hi = np.maximum(v0.max(axis=0), v1.max(axis=0))
    step = min(0.025 * (hi[0] - lo[0]), 0.025 * (hi[1] - lo[1]))
    x, y = np.mgrid[(lo[0] - 10 * step):(hi[0] + 10 * step):step,
                    (lo[1] - 10 * step):(hi[1] + 10 * step):step]
    n, m = x.shape
    pts = np.stack([x.ravel(), y.ravel()], axis=1)
    idx = np.arange(n * m).reshape(n, m)
    rows = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    cols = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    return pts, np.concatenate([rows, cols])


class CurveMatching(DeviceMatching):
    coef = 0.5
    measure_norm0_eps = 0.0   # curves.measureNorm0 weights are |lenel| (curves.py:624)
    aff_style = 'curve'

    def __init__(self, Template=None, Target=None, fileTempl=None, fileTarg=None, param=None,
                 maxIter=1000, regWeight=1.0, affineWeight=1.0, verb=True, gradLB=0.001,
                 saveRate=10, rotWeight=None, scaleWeight=None, transWeight=None,
                 testGradient=False, saveFile='evolution', affine='none', outputDir='.'):
        self.fv0 = _curve_arg(Template, fileTempl, 'template')
        self.fv1 = _curve_arg(Target, fileTarg, 'target')
        self.param = CurveMatchingParam() if param is None else param
        if self.param.errorType not in ('current', 'measure'):
            raise NotImplementedError("the GPU path implements the currents and measure data "
                                      "terms (errorType 'current' or 'measure')")
        self.fvDef = curves.Curve(curve=self.fv0)
        self.outputDir = outputDir
        if outputDir is not None and not os.access(outputDir, os.W_OK):
            if os.access(outputDir, os.F_OK):
                raise ValueError('Cannot save in ' + outputDir)
            os.makedirs(outputDir)
        self.maxIter = maxIter
        self.verb = verb
        self.testGradient = testGradient
        self.affine = affine
        self.saveFile = saveFile
        self.gradLB = gradLB
        self.saveRate = saveRate
        self.gridxy = self.gridEdges = None
        if self.fv0.vertices.shape[1] == 2 and outputDir is not None:
            self.gridxy, self.gridEdges = deformation_grid(self.fv0.vertices, self.fv1.vertices)
        Tsize = int(round(1.0 / self.param.timeStep))
        self._init_device(np.copy(self.fv0.vertices), self.fv0.faces, self.fv1.centers,
                          self.fv1.lenel, self.param.KparDiff, self.param.KparDist,
                          self.param.sigmaError, Tsize, regWeight,
                          data_term=self.param.errorType, affine=affine,
                          affineWeight=affineWeight, rotWeight=rotWeight, scaleWeight=scaleWeight,
                          transWeight=transWeight)
        if outputDir is not None:
            self.fv0.saveVTK(os.path.join(outputDir, 'Template.vtk'))
            self.fv1.saveVTK(os.path.join(outputDir, 'Target.vtk'))

    def _new_shape(self, vertices, with_weights=False):
        return curves.DeviceCurve(self.faces_np, vertices, with_weights=with_weights)


    def endOfIteration(self):
        """curveMatching.py:311-328: every saveRate iterations (when files are written) the
        flow's log-Jacobian (kept in ``jacobian``; the reference evaluates it every iteration
        and writes nothing of it) and the deformed curve at every step, with the deformed
        grid for planar curves."""
        DeviceMatching.endOfIteration(self)
        if self.outputDir is not None and self.saveRate > 0 and self.iter % self.saveRate == 0:
            self._save_steps(jacobian=True)
        else:
            self.fvDef.updateVertices(self.xt[-1].cpu().numpy())

    def _affine_np(self):
        T, d = self.Tsize, self.dim
        if self.affineDim > 0:
            return self.affB.getTransforms(self.Afft.cpu().numpy())
        return [np.zeros((T, d, d)), np.zeros((T, d))]

    def _save_steps(self, jacobian=False):
        """saveFile<t>.vtk for every step t of the accepted flow, and grid<t>.vtk for planar
        curves (the grid carried by the flow as a point set)."""
        A = self._affine_np()
        if jacobian:
            _, Jt = evol.landmarkDirectEvolutionEuler(self.x0, self.at, self.param.KparDiff,
                                                      affine=A, withJacobian=True)
            self.jacobian = Jt.cpu().numpy()
        yt = None
        if self.gridxy is not None:
            _, yt = evol.landmarkDirectEvolutionEuler(self.fv0.vertices, self.at.cpu().numpy(),
                                                      self.param.KparDiff, affine=A,
                                                      withPointSet=self.gridxy)
        xt = self.xt.cpu().numpy()
        for kk in range(self.Tsize + 1):
            self.fvDef.updateVertices(xt[kk])
            self.fvDef.saveVTK(os.path.join(self.outputDir, self.saveFile + str(kk) + '.vtk'))
            if yt is not None:
                fileio.save_vtk_curve(os.path.join(self.outputDir, 'grid' + str(kk) + '.vtk'),
                                      yt[kk], self.gridEdges)

    def endOptim(self):
        """curveMatching.py:331-340: the steps are written unless the last iteration just
        wrote them; defCost = obj - obj0 - data term of the deformed curve."""
        if self.outputDir is not None and (self.saveRate == 0 or self.iter % self.saveRate > 0):
            self._save_steps()
        self.fvDef.updateVertices(self.xt[-1].cpu().numpy())
        self.defCost = self.obj - self.obj0 - (
            self.param.fun_obj(self.fvDef, self.fv1, self.param.KparDist) / self.sigmaError ** 2)

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        logging.info('Gradient bound: %f' % self.gradEps)
        if self.outputDir is not None:
            # the previous run's step files go first (curveMatching.py:348-351)
            kk = 0
            while os.path.isfile(os.path.join(self.outputDir, self.saveFile + str(kk) + '.vtk')):
                os.remove(os.path.join(self.outputDir, self.saveFile + str(kk) + '.vtk'))
                kk += 1
        cg.cg(self, verb=self.verb, maxIter=self.maxIter, TestGradient=self.testGradient)

    def cg_steps(self):
        grd = yield (cg.GRADIENT, self.gradCoeff)
        [grd2] = yield (cg.DOT, grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        return (yield from cg.cg_steps(self, verb=self.verb, maxIter=self.maxIter,
                                       TestGradient=self.testGradient))
