"""GPU kernel sums (liblddmm_hip.so through the Kernel drop-in) against the reference's own
Fortran outputs (golden fixtures) and the FP64 oracle.

Tolerances: FP64 mode 1e-12 relative to the largest output entry; FP32-compensated mode 2e-6
(the path's 1e-5 trajectory bound with margin)."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32c": 2e-6}


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


def kernel(order, sigma, prec):
    from registration_amd import kernelFunctions as kfun
    if order > 4:
        return kfun.Kernel(name='gauss', sigma=sigma, precision=prec)
    return kfun.Kernel(name='laplacian', sigma=sigma, order=order, precision=prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_golden_kernel_sums(golden, prec):
    for name, c in golden.items():
        d, ny, nx, m, na, order, sigma = c["params"]
        K = kernel(int(order), float(sigma), prec)
        print(f"case {name} {prec}", flush=True)
        f = K.applyK(c["x"], c["w"], firstVar=c["y"])
        assert rel_err(f, c["applyk"]) < TOL[prec], (name, "applyK")
        if "applykdifft" in c:
            g = K.applyDiffKT(c["x"], c["a1"], c["a2"], firstVar=c["y"])
            assert rel_err(g, c["applykdifft"]) < TOL[prec], (name, "applyDiffKT")
        if "applykdiff11and12" in c:
            h = K.applyDDiffK11and12(c["y"], c["dd_a1"], c["dd_a2"], c["dd_p"])
            assert rel_err(h, c["applykdiff11and12"]) < TOL[prec], (name, "applyDDiffK11and12")


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("ny,nx,d,m", [(1, 1, 3, 3), (1, 700, 3, 1), (257, 1, 2, 2), (513, 1031, 3, 5),
                                       (1000, 300, 1, 1), (2049, 2049, 3, 3), (3000, 64, 2, 4)])
def test_apply_k_ragged(prec, ny, nx, d, m):
    rng = np.random.default_rng(ny * 7 + nx)
    y, x, w = rng.standard_normal((ny, d)), rng.standard_normal((nx, d)), rng.standard_normal((nx, m))
    for order, sigma in ((10, 0.6), (1, 0.5), (3, 0.8)):
        K = kernel(order, sigma, prec)
        print(f"applyK ny={ny} nx={nx} d={d} m={m} order={order} {prec}", flush=True)
        got = K.applyK(x, w, firstVar=y)
        ref = orc.apply_k(y, x, w, sigma, order)
        assert rel_err(got, ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("na,m,d", [(1, 3, 3), (2, 3, 3), (3, 3, 3), (1, 1, 3), (3, 2, 2), (4, 3, 3), (5, 2, 2)])
def test_apply_diff_kt(prec, na, m, d):
    rng = np.random.default_rng(na * 31 + m)
    ny, nx = 777, 1300
    y, x = rng.standard_normal((ny, d)), rng.standard_normal((nx, d))
    x[:3] = y[:3]
    a1, a2 = rng.standard_normal((na, ny, m)), rng.standard_normal((na, nx, m))
    for order, sigma in ((10, 0.5), (1, 0.6), (2, 0.7), (4, 0.4)):
        K = kernel(order, sigma, prec)
        print(f"applyDiffKT na={na} m={m} d={d} order={order} {prec}", flush=True)
        got = K.applyDiffKT(x, a1, a2, firstVar=y)
        ref = orc.apply_diff_kt(y, x, a1, a2, sigma, order)
        assert rel_err(got, ref) < TOL[prec] * 5


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("order,sigma,d", [(10, 0.5, 3), (2, 0.6, 3), (3, 0.5, 2), (4, 0.8, 3)])
def test_apply_ddiff_k11and12(prec, order, sigma, d):
    rng = np.random.default_rng(order)
    n = 900
    x, a1, a2, p = (rng.standard_normal((n, d)) for _ in range(4))
    K = kernel(order, sigma, prec)
    print(f"applyDDiffK11and12 order={order} d={d} {prec}", flush=True)
    got = K.applyDDiffK11and12(x, a1, a2, p)
    ref = orc.apply_ddiff_k11and12(x, a1, a2, p, sigma, order)
    assert rel_err(got, ref) < TOL[prec] * 5


def test_device_tensors_stay_on_device():
    import torch
    from registration_amd import kernelFunctions as kfun
    x = torch.randn(100, 3, dtype=torch.float64, device="cuda")
    a = torch.randn(100, 3, dtype=torch.float64, device="cuda")
    z = kfun.Kernel('gauss', sigma=0.5).applyK(x, a)
    assert isinstance(z, torch.Tensor) and z.is_cuda and z.dtype == torch.float64
    ref = orc.apply_k(x.cpu().numpy(), x.cpu().numpy(), a.cpu().numpy(), 0.5, 10)
    assert rel_err(z.cpu().numpy(), ref) < TOL["fp32c"]


def test_kernel_none_and_errors():
    from registration_amd import kernelFunctions as kfun
    x = np.random.default_rng(0).standard_normal((10, 3))
    z = kfun.Kernel(name='foo', sigma=1.0).applyK(x, x)
    assert np.all(z == 0)
    with pytest.raises(ValueError):
        kfun.Kernel('laplacian', sigma=1.0, order=0).applyDiffKT(x, x[None], x[None])
    with pytest.raises(ValueError):
        kfun.Kernel('gauss', sigma=1.0).applyK(np.zeros((5, 4)), np.zeros((5, 1)))


def test_affine_kernel_terms_match_reference_formula():
    from registration_amd import kernelFunctions as kfun
    rng = np.random.default_rng(2)
    x, a, y = rng.standard_normal((50, 3)), rng.standard_normal((50, 3)), rng.standard_normal((20, 3))
    c = np.array([0.1, -0.2, 0.3])
    K = kfun.Kernel('gauss', affine='affine', sigma=0.7, w1=0.3, w2=0.5, center=c, precision='fp64')
    got = K.applyK(x, a, firstVar=y)
    xx, yy = x - c, y - c
    ref = orc.apply_k(y, x, a, 0.7, 10) + 0.3 * yy @ (xx.T @ a) + 0.5 * a.sum(axis=0)
    assert rel_err(got, ref) < 1e-12
    Ke = kfun.Kernel('gauss', affine='euclidean', sigma=0.7, w1=0.3, w2=0.5, center=c, precision='fp64')
    got = Ke.applyK(x, a)
    ref = orc.apply_k(x, x, a, 0.7, 10) + 0.5 * a.sum(axis=0)
    for E in Ke.affine_basis:
        xE = xx @ E.T
        ref = ref + 0.3 * (xE * a).sum() * xE
    assert rel_err(got, ref) < 1e-12


def test_large_rows_sampled_against_oracle():
    """Full-size sum (40,962 points, the large-surface deformation kernel) checked on a row
    sample against the oracle, and FP32-compensated against FP64 on every row."""
    import torch
    from registration_amd import synthetic
    from registration_amd.kernelFunctions import apply_k_dev
    V, _ = synthetic.icosphere(6)
    rng = np.random.default_rng(6)
    a = 1e-2 * rng.standard_normal(V.shape)
    xd = torch.as_tensor(V, device="cuda")
    ad = torch.as_tensor(a, device="cuda")
    z32 = apply_k_dev(xd, xd, ad, kernel(10, 0.25, "fp32c")).cpu().numpy()
    z64 = apply_k_dev(xd, xd, ad, kernel(10, 0.25, "fp64")).cpu().numpy()
    rows = rng.choice(V.shape[0], 256, replace=False)
    ref = orc.apply_k(V[rows], V, a, 0.25, 10)
    assert rel_err(z64[rows], ref) < 1e-12
    assert rel_err(z32[rows], ref) < TOL["fp32c"]
    assert rel_err(z32, z64) < TOL["fp32c"]


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("name,order", [("gauss", 10), ("laplacian", 2)])
def test_tile_culling_matches_full_sums(prec, name, order):
    """Clustered points with a small kernel width: most (wave, tile) pairs are culled because
    every kernel value between them is below 2^-32 (FP32) / 2^-56 (FP64); the sums still
    match the oracle, which evaluates every pair."""
    from registration_amd import kernelFunctions as kfun
    rng = np.random.default_rng(21)
    centers = np.array([[0, 0, 0], [3, 0, 0], [0, 3, 0], [3, 3, 3]], dtype=float)
    pts = np.concatenate([c + 0.15 * rng.standard_normal((700, 3)) for c in centers])
    w = rng.standard_normal((pts.shape[0], 3))
    sigma = 0.05 if name == "gauss" else 0.02
    K = kfun.Kernel(name=name, sigma=sigma, order=order, precision=prec)
    got = K.applyK(pts, w)
    ref = orc.apply_k(pts, pts, w, sigma, K.order)
    assert rel_err(got, ref) < TOL[prec]
    a1, a2 = rng.standard_normal((2, pts.shape[0], 3)), rng.standard_normal((2, pts.shape[0], 3))
    g = K.applyDiffKT(pts, a1, a2)
    gr = orc.apply_diff_kt(pts, pts, a1, a2, sigma, K.order)
    assert rel_err(g, gr) < 5 * TOL[prec]


def test_max_size_landmark_flow_step_rows():
    """The largest landmark set of the mixed-service stream (200,000 points in the unit ball,
    Gaussian sigma 0.1): one fused Euler step, checked on 256 sampled rows against the oracle's
    FP64 sums (FP32-compensated tolerance) and its energy row-sum identity."""
    import torch
    from registration_amd import engine, synthetic
    from registration_amd.kernelFunctions import Kernel
    rng = np.random.default_rng(200)
    X = synthetic.ball_points(rng, 200000)
    a = 1e-3 * rng.standard_normal(X.shape)
    K = engine.DevKernel.of(Kernel('gauss', sigma=0.1, precision='fp32c'))
    x0, at = torch.as_tensor(X, device="cuda"), torch.as_tensor(a[None], device="cuda")
    xt = torch.empty((2,) + X.shape, dtype=torch.float64, device="cuda")
    v = torch.empty((1,) + X.shape, dtype=torch.float64, device="cuda")
    en = torch.zeros(1, dtype=torch.float64, device="cuda")
    engine.shoot([engine.FlowJob(x0, at, K, xt, en, v=v)])
    rows = rng.choice(X.shape[0], 256, replace=False)
    ref = orc.apply_k(X[rows], X, a, 0.1, 10)
    vv = v[0].cpu().numpy()
    assert rel_err(vv[rows], ref) < TOL["fp32c"]
    assert np.allclose(xt[1].cpu().numpy(), X + vv, rtol=0, atol=1e-15)  # dt = 1
    e = float((a * vv).sum())
    assert abs(en.item() - e) <= 1e-12 * abs(e)


def test_max_size_currents_energy_forms_agree():
    """Icosphere s7 (327,680 faces, the largest surface class): the FP64 currents energy
    evaluated in its energy-only form (symmetric self-sum, FP32 exponential for blocks whose
    kernel values are all below 2^-26) equals the plain FP64 form, whose per-row outputs force
    every pair, within 1e-12 relative to the sum of |terms| magnitude."""
    import torch
    from registration_amd import _lib, runtime as rt, synthetic
    V, F = synthetic.icosphere(7)
    Y = synthetic.rbf_field(np.random.default_rng(7), V)
    c1, c2 = V[F].mean(axis=1), Y[F].mean(axis=1)
    n1 = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    n2 = np.cross(Y[F[:, 1]] - Y[F[:, 0]], Y[F[:, 2]] - Y[F[:, 0]])
    # descriptors hold raw addresses: the device tensors must stay referenced until the launch ran
    dc1, dn1, dc2, dn2 = (torch.as_tensor(np.ascontiguousarray(z), device="cuda") for z in (c1, n1, c2, n2))
    tots = []
    for rowdot in (False, True):
        tot = torch.zeros(1, dtype=torch.float64, device="cuda")
        rd = torch.zeros(F.shape[0], dtype=torch.float64, device="cuda") if rowdot else None
        rt.launch(_lib.OP_CURRENTS, _lib.PREC_FP64, 3,
                  rt.seg(y=dc1, w0=dn1, x1=dc2, w1=dn2, total=tot, rowdot=rd,
                         ny=F.shape[0], nx1=F.shape[0], sigma=0.1, kernel=0, order=10, coef=1 / 3))
        tots.append(tot.item())
    print("energy-only %.17e  plain %.17e" % tuple(tots), flush=True)
    scale = float((np.abs(n1) ** 2).sum())   # magnitude of the self-term's diagonal
    assert abs(tots[0] - tots[1]) <= 1e-12 * scale


def test_symmetric_and_mixed_forms_small_cases():
    """Energy-only sums equal the per-row form on curves (d = 2, ragged sizes) and on the
    target norm (KNORM), FP32 and FP64."""
    import torch
    from registration_amd import _lib, runtime as rt, synthetic
    D = lambda z: torch.as_tensor(np.ascontiguousarray(z), device="cuda")
    for n, prec in ((256, _lib.PREC_FP64), (777, _lib.PREC_FP64), (3001, _lib.PREC_FP64),
                    (2500, _lib.PREC_FP32C)):
        X, _ = synthetic.circle(n)
        Y = synthetic.star_curve(n, 5, 0.2)
        E = synthetic.closed_curve_edges(n)
        c1, l1 = (X[E[:, 0]] + X[E[:, 1]]) / 2, X[E[:, 1]] - X[E[:, 0]]
        c2, l2 = (Y[E[:, 0]] + Y[E[:, 1]]) / 2, Y[E[:, 1]] - Y[E[:, 0]]
        # referenced until the launches ran (descriptors hold raw addresses)
        dc1, dl1, dc2, dl2 = D(c1), D(l1), D(c2), D(l2)
        for op in (_lib.OP_CURRENTS, _lib.OP_KNORM):
            out = []
            for rowdot in (False, True):
                tot = torch.zeros(1, dtype=torch.float64, device="cuda")
                rd = torch.zeros(n, dtype=torch.float64, device="cuda") if rowdot else None
                kw = dict(x1=dc2, w1=dl2, nx1=n) if op == _lib.OP_CURRENTS else {}
                rt.launch(op, prec, 2, rt.seg(y=dc1, w0=dl1, total=tot, rowdot=rd, ny=n,
                                              sigma=0.05, kernel=0, order=10, coef=0.5, **kw))
                out.append(tot.item())
            if op == _lib.OP_KNORM:
                ref = float((l1 * orc.apply_k(c1, c1, l1, 0.05, 10)).sum())
                assert abs(out[0] - ref) <= (1e-12 if prec == _lib.PREC_FP64 else 1e-6) * abs(ref)
            tol = 1e-12 if prec == _lib.PREC_FP64 else 1e-6
            assert abs(out[0] - out[1]) <= tol * float((l1 ** 2).sum()), (n, prec, op, out)


def test_fp64_exponential_is_within_a_few_ulp():
    """FP64 applyK with one source and unit weight returns the kernel's exponential itself:
    exp(-r^2 / (2 sigma^2)) = 2^t with t = -log2(e) r^2 / (2 sigma^2), compared with NumPy's exp2
    of the same t.  Beyond the polynomial's ~2 ulp, rounding t costs |t| ln 2 ulp."""
    from registration_amd import kernelFunctions as kfun
    sigma = 0.7
    r = np.concatenate([np.geomspace(1e-3, 9.0, 4000), np.linspace(0.01, 9.0, 4000)]) * sigma
    rng = np.random.default_rng(0)
    u = rng.standard_normal((r.size, 3))
    y = u / np.linalg.norm(u, axis=1, keepdims=True) * r[:, None]
    K = kfun.Kernel('gauss', sigma=sigma, precision='fp64')
    got = K.applyK(np.zeros((1, 3)), np.ones((1, 1)), firstVar=y)[:, 0]
    r2 = (y * y).sum(axis=1)
    t = -(np.log2(np.e) / (2 * sigma * sigma)) * r2
    ref = np.exp2(t)
    rel = np.abs(got - ref) / ref
    assert np.all(rel <= 2.3e-16 * (4 + 2 * np.abs(t))), float((rel / (4 + 2 * np.abs(t))).max())
