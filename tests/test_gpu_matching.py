"""GPU matching drivers (device-resident registration + the reference's CG) against the oracle
restatement: objective, gradient, RKHS products, and the per-iteration objective sequence."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def surface_pair(prec, subdiv=2, T=5):
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, synthetic
    V, F = synthetic.icosphere(subdiv)
    Y = synthetic.bump_target(V)
    p = sm.SurfaceMatchingParam(timeStep=1.0 / T, KparDiff=kfun.Kernel('gauss', sigma=0.5, precision=prec),
                                KparDist=kfun.Kernel('gauss', sigma=0.3, precision=prec),
                                sigmaError=0.1, errorType='current')
    g = sm.SurfaceMatching(Template=sm.surfaces.Surface(FV=(F, V)), Target=sm.surfaces.Surface(FV=(F, Y)),
                           param=p, testGradient=False, verb=False)
    o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, Y)), orc.Kernel('gauss', sigma=0.5),
                            orc.Kernel('gauss', sigma=0.3), sigmaError=0.1, timeStep=1.0 / T)
    return g, o


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32c", 1e-6)])
def test_surface_requests_match_oracle(prec, tol):
    g, o = surface_pair(prec)
    rng = np.random.default_rng(1)
    a = 0.02 * rng.standard_normal(o.at.shape)
    import torch
    g.at = torch.as_tensor(a, device="cuda")
    o.at = a.copy()
    assert abs(g.objectiveFun() - o.objectiveFun()) < tol * abs(o.obj)
    o.endOfIteration()
    gg, go = g.getGradient(g.gradCoeff), o.getGradient(o.gradCoeff)
    err = np.abs(gg.diff.cpu().numpy() - go.diff).max() / np.abs(go.diff).max()
    assert err < 10 * tol
    [d1, d2] = g.dotProduct(gg, [gg, gg])
    [r1] = o.dotProduct(go, [go])
    assert abs(d1 - r1) < 10 * tol * abs(r1) and d1 == d2
    t1 = g.updateTry(gg, 0.05, None)
    t2 = o.updateTry(go, 0.05, None)
    assert abs(t1 - t2) < tol * abs(t2)


def test_surface_cg_sequence_fp64():
    g, o = surface_pair("fp64", subdiv=2, T=5)
    g.maxIter = o.maxIter = 8
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 5
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)


def test_surface_cg_fp32c_energies_and_path():
    """FP32-compensated run: (a) the energy of every accepted iterate equals the FP64 energy of
    the same momenta within 1e-6 relative (the precision claim); (b) the iterate path stays
    within 1e-4 of the FP64 oracle's (FP32 gradients change the search directions slightly,
    so the paths are not identical)."""
    import torch
    g, o = surface_pair("fp32c", subdiv=2, T=5)
    g.maxIter = o.maxIter = 8
    stash = []
    accept = g.acceptVarTry

    def spy():
        accept()
        stash.append((g.at.clone(), g.obj))
    g.acceptVarTry = spy
    g.optimizeMatching()
    o.optimizeMatching()
    g64, _ = surface_pair("fp64", subdiv=2, T=5)
    for at, e in stash:
        g64.at = at
        g64.obj = None
        e64 = g64.objectiveFun()
        assert abs(e - e64) <= 1e-6 * abs(e64), (e, e64)
    hg, ho = np.array(g.history), np.array(o.history)
    n = min(len(hg), len(ho))
    assert n >= 5
    assert np.all(np.abs(hg[:n] - ho[:n]) <= 1e-4 * np.abs(ho[:n])), (hg[:n], ho[:n])


def curve_pair(prec):
    from registration_amd import curveMatching as cm, curves, kernelFunctions as kfun, synthetic
    n = 256
    X, _ = synthetic.circle(n)
    Y = synthetic.star_curve(n, 5, 0.2, noise=0.01, rng=np.random.default_rng(2))
    p = cm.CurveMatchingParam(timeStep=0.2, KparDiff=kfun.Kernel('laplacian', sigma=0.2, order=1, precision=prec),
                              KparDist=kfun.Kernel('gauss', sigma=0.1, precision=prec), sigmaError=0.1,
                              errorType='current')
    g = cm.CurveMatching(Template=curves.Curve(pointSet=X), Target=curves.Curve(pointSet=Y), param=p,
                         maxIter=6, verb=False)
    o = orc.CurveMatching(orc.Curve(pointSet=X), orc.Curve(pointSet=Y), orc.Kernel('laplacian', sigma=0.2, order=1),
                          orc.Kernel('gauss', sigma=0.1), sigmaError=0.1, timeStep=0.2, maxIter=6)
    return g, o


def test_curve_cg_sequence_fp64():
    g, o = curve_pair("fp64")
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 4
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)


def test_curve_cg_fp32c_energies_and_path():
    g, o = curve_pair("fp32c")
    stash = []
    accept = g.acceptVarTry

    def spy():
        accept()
        stash.append((g.at.clone(), g.obj))
    g.acceptVarTry = spy
    g.optimizeMatching()
    o.optimizeMatching()
    g64, _ = curve_pair("fp64")
    for at, e in stash:
        g64.at = at
        g64.obj = None
        e64 = g64.objectiveFun()
        assert abs(e - e64) <= 1e-6 * abs(e64), (e, e64)
    hg, ho = np.array(g.history), np.array(o.history)
    n = min(len(hg), len(ho))
    assert n >= 4
    assert np.all(np.abs(hg[:n] - ho[:n]) <= 1e-4 * np.abs(ho[:n])), (hg[:n], ho[:n])


def test_nan_trial_is_rejected():
    """A trial whose objective is NaN returns +inf and is not stored (surfaceMatching.py:278-280)."""
    import torch
    g, _ = surface_pair("fp32c")
    g.objectiveFun()
    d = g.getGradient(g.gradCoeff)
    d.diff = d.diff.clone()
    d.diff[0, 0, 0] = float('nan')
    before = g.objTry
    assert g.updateTry(d, 0.1, g.obj) == float('inf')
    assert g.objTry is before
