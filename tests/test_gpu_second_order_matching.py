"""The second-order (spline) regression driver through a time series of surfaces
(registration_amd.secondOrderMatching.SurfaceMatching, the drop-in for
secondOrderMatching.py:43-509) against the oracle's restatement: the per-iteration objective
sequence of the conjugate-gradient run, through the affine burn-in and the 'spline2'
regression.  FP64 mode: 1e-9 relative; FP32-compensated mode: 1e-4 (its gradients steer the
search slightly differently)."""
import os

import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def shapes():
    from registration_amd import synthetic
    V, F = synthetic.icosphere(2)
    rng = np.random.default_rng(11)
    targets = []
    for k in range(2):
        W = V * (1.0 + 0.1 * (k + 1) * np.array([1.0, -0.5, 0.3])) + 0.02 * rng.standard_normal(V.shape)
        targets.append((F, W))
    return (F, V), targets


def run_pair(prec, tmp_path, affine='euclidean', maxIter=6, burn_in=2, typeRegression='spline2'):
    from registration_amd import kernelFunctions as kfun, secondOrderMatching as som, surfaces
    from registration_amd.surfaceMatching import SurfaceMatchingParam
    (F, V), targets = shapes()
    p = SurfaceMatchingParam(timeStep=1.0 / 3,
                             KparDiff=kfun.Kernel('gauss', sigma=0.6, precision=prec),
                             KparDist=kfun.Kernel('gauss', sigma=0.4, precision=prec),
                             sigmaError=0.2, errorType='current')
    g = som.SurfaceMatching(Template=surfaces.Surface(FV=(F, V)),
                            Targets=[surfaces.Surface(FV=t) for t in targets], param=p,
                            maxIter=maxIter, affine=affine, controlWeight=0.7,
                            typeRegression=typeRegression, testGradient=False,
                            outputDir=str(tmp_path), verb=False)
    o = orc.SecondOrderSurfaceMatching(orc.Surface(FV=(F, V)), [orc.Surface(FV=t) for t in targets],
                                       orc.Kernel('gauss', sigma=0.6), orc.Kernel('gauss', sigma=0.4),
                                       sigmaError=0.2, timeStep=1.0 / 3, maxIter=maxIter,
                                       typeRegression=typeRegression, affine=affine,
                                       controlWeight=0.7)
    g.affBurnIn = o.affBurnIn = burn_in
    return g, o


def test_second_order_cg_sequence_fp64(tmp_path):
    g, o = run_pair("fp64", tmp_path)
    print("second-order regression, fp64", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 4, (hg, ho)
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)
    assert g.typeRegression == 'spline2'
    ga0 = g.a0.cpu().numpy()
    assert np.abs(ga0 - o.a0).max() <= 1e-7 * max(np.abs(o.a0).max(), 1e-12)
    assert os.path.exists(os.path.join(str(tmp_path), 'Template.vtk'))


def test_second_order_cg_fp32c(tmp_path):
    g, o = run_pair("fp32c", tmp_path)
    print("second-order regression, fp32c", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    n = min(len(hg), len(ho))
    assert n >= 4
    assert np.all(np.abs(hg[:n] - ho[:n]) <= 1e-4 * np.abs(ho[:n])), (hg, ho)


def test_second_order_geodesic_and_saves(tmp_path):
    """'geodesic' regression (a0 only) without an affine part, with the periodic output of the
    deformed template and its Jacobian.  The regression type switches from 'affine' only in
    endOfIteration, so without an affine part the first gradient would be zero and the run
    would stop at once (as the reference's does): the test starts in 'geodesic'."""
    g, o = run_pair("fp64", tmp_path, affine='none', maxIter=4, burn_in=0, typeRegression='geodesic')
    g.typeRegression = o.typeRegression = 'geodesic'
    g.saveRate = 2
    print("second-order regression, geodesic", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 3, (hg, ho)
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)
    assert bool((g.rhot == 0).all())
    if g.iter >= 2:
        assert os.path.exists(os.path.join(str(tmp_path), 'evolution0.vtk'))
