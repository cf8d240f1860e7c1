"""Symmetric surface matching (SurfaceMatching(..., symmetric=True), surfaceMatching.py:74-484)
against the oracle's restatement: the per-iteration objective sequence of the CG run, without
and with an affine part.  FP64 mode 1e-9 relative; FP32-compensated mode 1e-4."""
import os

import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def pair(prec, tmp_path, affine='none', maxIter=6):
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, surfaces, synthetic
    V, F = synthetic.icosphere(2)
    rng = np.random.default_rng(9)
    W = V * np.array([1.15, 0.9, 1.0]) + 0.02 * rng.standard_normal(V.shape)
    p = sm.SurfaceMatchingParam(timeStep=0.25,
                                KparDiff=kfun.Kernel('gauss', sigma=0.6, precision=prec),
                                KparDist=kfun.Kernel('gauss', sigma=0.4, precision=prec),
                                sigmaError=0.2, errorType='current')
    g = sm.SurfaceMatching(Template=surfaces.Surface(FV=(F, V)), Target=surfaces.Surface(FV=(F, W)),
                           param=p, maxIter=maxIter, symmetric=True, affine=affine,
                           testGradient=False, outputDir=str(tmp_path))
    o = orc.SymmetricSurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, W)),
                                     orc.Kernel('gauss', sigma=0.6), orc.Kernel('gauss', sigma=0.4),
                                     sigmaError=0.2, timeStep=0.25, maxIter=maxIter, affine=affine)
    return g, o


@pytest.mark.parametrize("affine", ["none", "euclidean"])
def test_symmetric_cg_sequence_fp64(tmp_path, affine):
    from registration_amd import surfaceMatching as sm
    g, o = pair("fp64", tmp_path, affine)
    assert isinstance(g, sm.SymmetricSurfaceMatching)
    print(f"symmetric matching fp64 affine={affine}", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 4, (hg, ho)
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)
    gx0 = g.x0.cpu().numpy()
    assert np.abs(gx0 - o.x0).max() <= 1e-8 * np.abs(o.x0).max()
    assert np.abs(gx0 - g.fv0.vertices).max() > 0  # the template's points moved
    assert os.path.exists(os.path.join(str(tmp_path), 'Template.vtk'))


def test_symmetric_cg_fp32c(tmp_path):
    g, o = pair("fp32c", tmp_path)
    print("symmetric matching fp32c", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    n = min(len(hg), len(ho))
    assert n >= 4
    assert np.all(np.abs(hg[:n] - ho[:n]) <= 1e-4 * np.abs(ho[:n])), (hg, ho)


def test_non_symmetric_stays_on_the_batched_driver(tmp_path):
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, surfaces, synthetic
    V, F = synthetic.icosphere(1)
    p = sm.SurfaceMatchingParam(timeStep=0.25, KparDiff=kfun.Kernel('gauss', sigma=0.6),
                                KparDist=kfun.Kernel('gauss', sigma=0.4), sigmaError=0.2,
                                errorType='current')
    g = sm.SurfaceMatching(Template=surfaces.Surface(FV=(F, V)), Target=surfaces.Surface(FV=(F, 1.1 * V)),
                           param=p, symmetric=False, outputDir=None)
    assert type(g) is sm.SurfaceMatching
