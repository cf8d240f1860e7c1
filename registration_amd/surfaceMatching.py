"""Surface matching, drop-in for py-lddmm's surfaceMatching module
(py-lddmm/surfaceMatching.py:19-484).

``SurfaceMatchingParam`` and ``SurfaceMatching`` keep the reference constructors and the
optimiser protocol; the registration lives on the GPU (see matching.DeviceMatching) and
``optimizeMatching`` runs the reference's conjugate gradient over it.  Supported: the
currents, measure and varifold data terms (errorType 'current', 'measure' (the default) or
'varifold'), the global affine part (affine 'affine', 'similitude', 'euclidean',
'translation' or 'none', with affineWeight / rotWeight / scaleWeight / transWeight),
symmetric matching (the template's points a variable too), templates and targets given as
surfaces or shape files, and template decimation (subsampleTargetSize > 0, by quadric-error
edge collapse in the kernel library's host code).
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
            v1 = -v1  # flipping every face negates each determinant
        # copies of the template whose geometry was just computed from its vertices; with
        # subsampleTargetSize > 0 the registered template is the decimated copy
        # (surfaceMatching.py:131-133; the target was oriented against the full template)
        self.fvInit = self.fv0.clone()
        if subsampleTargetSize > 0:
            self.fvInit.Simplify(subsampleTargetSize)
            logging.info('simplified template %d' % self.fvInit.vertices.shape[0])
            v0 = self.fvInit.surfVolume()
        self.fvDef = self.fvInit.clone()
        # orientations of the registered template and of the target after the flip
        self.fv0ori = -1 if v0 < 0 else 1
        self.fv1ori = -1 if v1 < 0 else 1
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


    def endOfIteration(self):
        """surfaceMatching.py:386-471: the affine scaling after the burn-in and, every
        saveRate iterations when files are written, the snapshot (``_snapshot``)."""
        DeviceMatching.endOfIteration(self)
        if self.outputDir is not None and self.saveRate > 0 and self.iter % self.saveRate == 0:
            self._snapshot()

    def _snapshot(self):
        """The reference's saveRate output: (1) without an affine part, geodesic shooting from
        (x0, a_0) and its end point against the flow's ('EPDiff difference' = sum |x_T -
        x_T^EPDiff|, kept in ``epdiffDifference``; saveFile + 'EPDiff.vtk'); (2) the flow with
        its log-Jacobian J_t (kept in ``jacobian``, (T+1, N)); (3) for 'euclidean' or
        'translation', the trajectory and target in the frame that undoes the affine flow
        (saveFile + 'Corrected<t>.vtk', 'TargetCorrected.vtk'); (4) the deformed template at
        every step with the fields Jacobian (e^J - 1), Jacobian_T (vertex-area ratio - 1),
        Jacobian_N, displacement (the normal velocity integrated in time) and velocity
        (saveFile + '<t>.vtk').  Kernel sums on the GPU; the fields in host NumPy."""
        logging.info('Saving surfaces...')
        K = self.param.KparDiff
        T, N, d = self.Tsize, self.npt, self.dim
        dt = 1.0 / T
        out = lambda name: os.path.join(self.outputDir, name)
        x0 = self.x0.cpu().numpy()
        at = self.at
        xt = self.xt.cpu().numpy()
        # velocities K(x_t) a_t of the accepted trajectory (the reference's self.v; slice T
        # stays zero, as there)
        v = np.zeros((T + 1, N, d))
        for t in range(T):
            v[t] = kfun.apply_k_dev(self.xt[t], self.xt[t], at[t], K).cpu().numpy()
        if self.affineDim <= 0:
            xe, _ = evol.landmarkEPDiff(T, self.x0, at[0].contiguous(), K)
            xe = xe.cpu().numpy()
            self.fvDef.updateVertices(xe[-1])
            self.fvDef.saveVTK(out(self.saveFile + 'EPDiff.vtk'))
            self.epdiffDifference = float(np.fabs(xt[-1] - xe[-1]).sum())
            logging.info('EPDiff difference %f' % self.epdiffDifference)
        self.fvDef.updateVertices(xt[-1])
        self.fvInit.updateVertices(x0)
        if self.affineDim > 0:
            Afft = self.Afft.cpu().numpy()
            A = self.affB.getTransforms(Afft)
        else:
            Afft = np.zeros((T, 0))
            A = [np.zeros((T, d, d)), np.zeros((T, d))]
        xj, Jt = evol.landmarkDirectEvolutionEuler(self.x0, at, K, affine=A, withJacobian=True)
        xj, Jt = xj.cpu().numpy(), Jt.cpu().numpy()
        self.jacobian = Jt
        at_np = at.cpu().numpy()
        if self.affine in ('euclidean', 'translation'):
            f = self.fvInit.clone()
            X = self.affB.integrateFlow(Afft)
            displ = np.zeros(N)
            for t in range(T + 1):
                U = np.linalg.inv(X[0][t])
                yyt = (xt[t] - X[1][t]) @ U.T
                zt = (xj[t] - X[1][t]) @ U.T
                if t < T:
                    vt = K.applyK(yyt, at_np[t] @ U.T, firstVar=zt)
                f.updateVertices(yyt)
                vf = surfaces.vtkFields()
                vf.scalars += ['Jacobian', np.exp(Jt[t]) - 1, 'displacement', displ]
                vf.vectors += ['velocity', vt]
                nu = self.fv0ori * f.computeVertexNormals()
                displ = displ + dt * (vt * nu).sum(axis=1)
                f.saveVTK2(out(self.saveFile + 'Corrected' + str(t) + '.vtk'), vf)
            f = self.fv1.clone()
            f.updateVertices((f.vertices - X[1][-1]) @ U.T)
            f.saveVTK(out('TargetCorrected.vtk'))
        fvDef = self.fvInit.clone()
        AV0 = fvDef.computeVertexArea()[0]
        nu = self.fv0ori * self.fvInit.computeVertexNormals()
        vel = v[0]
        displ = np.zeros(N)
        for kk in range(T + 1):
            fvDef.updateVertices(xj[kk])
            AV = fvDef.computeVertexArea()[0] / AV0 - 1
            vf = surfaces.vtkFields()
            vf.scalars += ['Jacobian', np.exp(Jt[kk]) - 1, 'Jacobian_T', AV[:, 0],
                           'Jacobian_N', np.exp(Jt[kk]) / (AV[:, 0] + 1) - 1,
                           'displacement', displ]
            displ = displ + dt * (vel * nu).sum(axis=1)
            if kk < T:
                nu = self.fv0ori * self.fvDef.computeVertexNormals()
                vel = v[kk]
                kkm = kk
            else:
                kkm = kk - 1
            vf.vectors += ['velocity', v[kkm]]
            fvDef.saveVTK2(out(self.saveFile + str(kk) + '.vtk'), vf)

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
