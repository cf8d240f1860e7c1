"""GPU varifold data term of surfaces against the FP64 oracle (which, like the reference,
forms the (F, F) pair weights as dense arrays): norms, gradients, and whole registrations
with errorType='varifold'."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-6)])
@pytest.mark.parametrize("subdiv,sigma,kern", [(2, 0.3, "gauss"), (3, 0.15, "gauss"),
                                               (3, 0.2, "laplacian")])
def test_surface_varifold(prec, tol, subdiv, sigma, kern):
    from registration_amd import kernelFunctions as kfun, surfaces, synthetic
    V, F = synthetic.icosphere(subdiv)
    rng = np.random.default_rng(subdiv + 11)
    X = V + 0.02 * rng.standard_normal(V.shape)
    Y = synthetic.rbf_field(rng, V)
    order = 3 if kern == "laplacian" else 10
    K = kfun.Kernel(kern, sigma=sigma, order=order, precision=prec)
    Ko = orc.Kernel(kern, sigma=sigma, order=order)
    S, T = surfaces.Surface(FV=(F, X)), surfaces.Surface(FV=(F, Y))
    So, To = orc.Surface(FV=(F, X)), orc.Surface(FV=(F, Y))
    print(f"surface varifold s={subdiv} {kern} {prec}", flush=True)
    n0r = orc.surf_varifoldNorm0(To, Ko)
    assert abs(surfaces.varifoldNorm0(T, K) - n0r) < 1e-12 * abs(n0r)
    ndr = orc.surf_varifoldNormDef(So, To, Ko)
    assert abs(surfaces.varifoldNormDef(S, T, K) - ndr) < 1e-12 * abs(n0r)
    g = surfaces.varifoldNormGradient(S, T, K)
    gr = orc.surf_varifoldNormGradient(So, To, Ko)
    assert rel_err(g, gr) < 10 * tol


def test_varifold_self_distance_is_zero_at_size():
    """|S - S|^2 = 0 at 20,480 faces (no dense oracle at this size): the symmetric
    energy-only self sum and the cross sum agree to FP64 rounding."""
    from registration_amd import kernelFunctions as kfun, surfaces, synthetic
    V, F = synthetic.icosphere(5)
    K = kfun.Kernel('gauss', sigma=0.05, precision='fp64')
    S = surfaces.Surface(FV=(F, synthetic.bump_target(V)))
    n0 = surfaces.varifoldNorm0(S, K)
    assert n0 > 0
    assert abs(surfaces.varifoldNormDef(S, S, K) + n0) < 1e-12 * n0


def _surface_pair(prec, T=5, subdiv=2):
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, synthetic
    V, F = synthetic.icosphere(subdiv)
    Y = synthetic.bump_target(V)
    p = sm.SurfaceMatchingParam(timeStep=1.0 / T, KparDiff=kfun.Kernel('gauss', sigma=0.5, precision=prec),
                                KparDist=kfun.Kernel('gauss', sigma=0.3, precision=prec),
                                sigmaError=0.1, errorType='varifold')
    g = sm.SurfaceMatching(outputDir=None, Template=sm.surfaces.Surface(FV=(F, V)),
                           Target=sm.surfaces.Surface(FV=(F, Y)), param=p, testGradient=False)
    o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, Y)),
                            orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3),
                            sigmaError=0.1, timeStep=1.0 / T, errorType='varifold')
    return g, o


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32c", 1e-6)])
def test_varifold_requests_match_oracle(prec, tol):
    import torch
    g, o = _surface_pair(prec)
    a = 0.02 * np.random.default_rng(6).standard_normal(o.at.shape)
    g.at = torch.as_tensor(a, device="cuda")
    o.at = a.copy()
    print(f"varifold requests {prec}", flush=True)
    assert abs(g.objectiveFun() - o.objectiveFun()) < tol * abs(o.obj)
    o.endOfIteration()
    gg, go = g.getGradient(g.gradCoeff), o.getGradient(o.gradCoeff)
    assert rel_err(gg.diff.cpu().numpy(), go.diff) < 10 * tol
    t1, t2 = g.updateTry(gg, 0.05, None), o.updateTry(go, 0.05, None)
    assert abs(t1 - t2) < tol * abs(t2)


def test_varifold_cg_sequence_fp64():
    g, o = _surface_pair("fp64")
    g.maxIter = o.maxIter = 8
    print("varifold cg", flush=True)
    g.optimizeMatching()
    o.optimizeMatching()
    hg, ho = np.array(g.history), np.array(o.history)
    assert len(hg) == len(ho) and len(hg) >= 5
    assert np.all(np.abs(hg - ho) <= 1e-9 * np.abs(ho)), (hg, ho)
    assert hg[-1] < 0.5 * hg[0]


def test_varifold_rejected_for_curves():
    from registration_amd import curveMatching as cm, kernelFunctions as kfun, synthetic
    X, _ = synthetic.circle(50)
    E = synthetic.closed_curve_edges(50)
    with pytest.raises(NotImplementedError):
        p = cm.CurveMatchingParam(KparDiff=kfun.Kernel('gauss', sigma=0.3),
                                  KparDist=kfun.Kernel('gauss', sigma=0.1), errorType='varifold')
        cm.CurveMatching(outputDir=None, Template=cm.curves.Curve(FV=(E, X)), Target=cm.curves.Curve(FV=(E, X)),
                         param=p)
