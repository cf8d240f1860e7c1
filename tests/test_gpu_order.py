"""Flow steps and RKHS products in a spatial processing order (FP32 Gaussian sums: on
compact 64-source blocks the exponent is 2<x, y> - |y|^2 - |x|^2 in tile coordinates) against
the natural order and the FP64 oracle.

Tolerances: FP32-compensated 2e-6 relative to the largest output entry (as for every FP32
kernel sum); energies and RKHS products 1e-6 relative."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu

TOL = 2e-6
PRESCALE = np.sqrt(np.log2(np.e) / 2)


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


def quarter_fraction(P, order, sigma, bound=2.0):
    """Fraction of 64-source quarters within R'^2 <= bound of their tile origin (the mean of
    the tile's sources 32, 96, 160, 224), i.e. those the kernel takes in dot-product form."""
    Q = P[order] * PRESCALE / sigma
    ok = []
    for base in range(0, len(Q), 256):
        t = Q[base:base + 256]
        org = t[[min(i, len(t) - 1) for i in (32, 96, 160, 224)]].mean(axis=0)
        for q in range(0, len(t), 64):
            ok.append(((t[q:q + 64] - org) ** 2).sum(axis=1).max() <= bound)
    return float(np.mean(ok))


def gauss(sigma):
    from registration_amd import engine
    from registration_amd.kernelFunctions import Kernel
    return engine.DevKernel.of(Kernel('gauss', sigma=sigma, precision='fp32c'))


def flow(X, at, K, perm):
    import torch
    from registration_amd import engine
    T = at.shape[0]
    x0 = torch.as_tensor(X, device="cuda")
    a = torch.as_tensor(at, device="cuda")
    xt = torch.empty((T + 1,) + X.shape, dtype=torch.float64, device="cuda")
    v = torch.empty((T,) + X.shape, dtype=torch.float64, device="cuda")
    en = torch.zeros(T, dtype=torch.float64, device="cuda")
    engine.shoot([engine.FlowJob(x0, a, K, xt, en, v=v, perm=perm)])
    return xt.cpu().numpy(), v.cpu().numpy(), en.cpu().numpy()


def test_ordered_flow_step_large_surface():
    """One Euler step on the large surface (icosphere s6, sigma 0.25): every quarter takes
    the dot-product form; sampled rows against the oracle, all rows against the natural
    order, and the energy row-sum identity."""
    from registration_amd import engine, synthetic
    V, _ = synthetic.icosphere(6)
    rng = np.random.default_rng(61)
    a = 1e-2 * rng.standard_normal(V.shape)
    perm = engine.spatial_order_dev(V)
    assert quarter_fraction(V, perm.cpu().numpy(), 0.25) > 0.99
    K = gauss(0.25)
    print("ordered flow step, icosphere s6", flush=True)
    xo, vo, eo = flow(V, a[None], K, perm)
    xn, vn, en = flow(V, a[None], K, None)
    rows = rng.choice(V.shape[0], 256, replace=False)
    assert rel_err(vo[0][rows], orc.apply_k(V[rows], V, a, 0.25, 10)) < TOL
    assert rel_err(vo, vn) < TOL
    assert np.allclose(xo[1], V + vo[0], rtol=0, atol=1e-15)  # dt = 1
    e = float((a * vo[0]).sum())
    assert abs(eo[0] - e) <= 1e-12 * abs(e)
    assert abs(eo[0] - en[0]) <= 1e-6 * abs(en[0])


def test_ordered_flow_trajectory_against_oracle():
    """T = 5 Euler steps on icosphere s4 (sigma 0.5: most quarters in dot-product form),
    trajectory and energies against the oracle's FP64 flow."""
    from registration_amd import engine, synthetic
    V, _ = synthetic.icosphere(4)
    rng = np.random.default_rng(4)
    at = 0.05 * rng.standard_normal((5,) + V.shape)
    perm = engine.spatial_order_dev(V)
    assert quarter_fraction(V, perm.cpu().numpy(), 0.5) > 0.5
    print("ordered flow, icosphere s4, T=5", flush=True)
    xt, _, en = flow(V, at, gauss(0.5), perm)
    ref = orc.landmarkDirectEvolutionEuler(V, at, orc.Kernel('gauss', sigma=0.5))
    assert np.abs(xt - ref).max() < 1e-6
    for t in range(5):
        e = float((at[t] * orc.apply_k(ref[t], ref[t], at[t], 0.5, 10)).sum())
        assert abs(en[t] - e) <= 1e-6 * abs(e)


def test_ordered_flow_with_outliers():
    """A dense blob with far outliers: blob quarters take the dot-product form, quarters with
    an outlier keep the direct form, and outlier rows (|x|'^2 > 60 from compact tiles) are
    left out of the compact blocks, where all their kernel values are below 2^-41.  All rows
    against the oracle, over three steps."""
    from registration_amd import engine
    rng = np.random.default_rng(17)
    blob = 0.3 * rng.standard_normal((3000, 3))
    d = rng.standard_normal((40, 3))
    out = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(2.5, 4.0, (40, 1))
    X = np.concatenate([blob, out])
    X = X[rng.permutation(len(X))]
    at = 0.02 * rng.standard_normal((3,) + X.shape)
    perm = engine.spatial_order_dev(X)
    frac = quarter_fraction(X, perm.cpu().numpy(), 0.3)
    assert 0.05 < frac < 1.0  # both forms occur
    print(f"ordered flow with outliers, dot-product quarters {frac:.2f}", flush=True)
    xt, v, en = flow(X, at, gauss(0.3), perm)
    assert rel_err(v[0], orc.apply_k(X, X, at[0], 0.3, 10)) < TOL
    ref = orc.landmarkDirectEvolutionEuler(X, at, orc.Kernel('gauss', sigma=0.3))
    assert np.abs(xt - ref).max() < 1e-6
    _, vn, _ = flow(X, at, gauss(0.3), None)
    assert rel_err(v, vn) < TOL


@pytest.mark.parametrize("which", ["spatial", "identity"])
def test_ordered_flow_on_a_curve(which):
    """d = 2: a closed star curve (2,048 points, sigma 0.3), flowed in the spatial order (91% of
    quarters compact) and in its own arc order passed as an identity permutation (all
    compact), against the oracle and the unordered flow."""
    import torch
    from registration_amd import engine, synthetic
    X = synthetic.star_curve(2048, 5, 0.2, noise=0.0, rng=np.random.default_rng(0))
    rng = np.random.default_rng(22)
    at = 0.02 * rng.standard_normal((4,) + X.shape)
    if which == "spatial":
        perm = engine.spatial_order_dev(X)
    else:
        perm = torch.arange(X.shape[0], dtype=torch.int32, device="cuda")
    frac = quarter_fraction(np.c_[X, np.zeros(len(X))], perm.cpu().numpy(), 0.3)
    assert frac > 0.8
    print(f"ordered flow on a curve ({which}), dot-product quarters {frac:.2f}", flush=True)
    K = gauss(0.3)
    xt, v, en = flow(X, at, K, perm)
    ref = orc.landmarkDirectEvolutionEuler(X, at, orc.Kernel('gauss', sigma=0.3))
    assert np.abs(xt - ref).max() < 1e-6
    assert rel_err(v[0], orc.apply_k(X, X, at[0], 0.3, 10)) < TOL
    _, vn, _ = flow(X, at, K, None)
    assert rel_err(v, vn) < TOL


def test_driver_takes_the_more_compact_order():
    """Curves keep their arc order (as an identity permutation); the large surface takes the
    spatial order; small spheres take none."""
    from registration_amd import engine, synthetic
    X = synthetic.star_curve(16384, 5, 0.2, noise=0.0, rng=np.random.default_rng(0))
    o = engine.order_for(X, gauss(0.3))
    assert o is not None and np.array_equal(o.cpu().numpy(), np.arange(len(X)))
    V6, _ = synthetic.icosphere(6)
    o6 = engine.order_for(V6, gauss(0.25))
    assert o6 is not None and not np.array_equal(o6.cpu().numpy(), np.arange(len(V6)))
    V3, _ = synthetic.icosphere(3)
    assert engine.order_for(V3, gauss(0.3)) is None


def test_ordered_and_natural_segments_in_one_launch():
    """A served batch mixes registrations with and without a processing order in one launch;
    each trajectory against the oracle, and against the same trajectory flowed alone."""
    import torch
    from registration_amd import engine, synthetic
    V5, _ = synthetic.icosphere(5)
    V3, _ = synthetic.icosphere(3)
    rng = np.random.default_rng(53)
    a5 = 0.03 * rng.standard_normal((3,) + V5.shape)
    a3 = 0.03 * rng.standard_normal((3,) + V3.shape)
    K = gauss(0.4)
    perm = engine.spatial_order_dev(V5)

    def job(X, a, p):
        xt = torch.empty((4,) + X.shape, dtype=torch.float64, device="cuda")
        en = torch.zeros(3, dtype=torch.float64, device="cuda")
        return engine.FlowJob(torch.as_tensor(X, device="cuda"), torch.as_tensor(a, device="cuda"),
                              K, xt, en, perm=p)
    print("ordered and natural segments in one launch", flush=True)
    j5, j3 = job(V5, a5, perm), job(V3, a3, None)
    engine.shoot([j5, j3])
    alone = job(V3, a3, None)
    engine.shoot([alone])
    for j, X, a in ((j5, V5, a5), (j3, V3, a3)):
        ref = orc.landmarkDirectEvolutionEuler(X, a, orc.Kernel('gauss', sigma=0.4))
        assert np.abs(j.xt.cpu().numpy() - ref).max() < 1e-6
    # A segment's result does not depend on the other segments of its launch.
    assert np.array_equal(j3.xt.cpu().numpy(), alone.xt.cpu().numpy())


def test_ordered_covector_against_oracle():
    """The adjoint (covector) steps in the processing order: trajectory of covectors against
    the oracle's FP64 landmarkHamiltonianCovector and against the natural order."""
    import torch
    from registration_amd import engine, synthetic
    V, _ = synthetic.icosphere(5)
    rng = np.random.default_rng(35)
    T = 4
    at = 0.02 * rng.standard_normal((T,) + V.shape)
    px1 = rng.standard_normal(V.shape)
    K = gauss(0.3)
    perm = engine.spatial_order_dev(V)
    xt, _, _ = flow(V, at, K, perm)
    dev = lambda a: torch.as_tensor(a, device="cuda")
    outs = []
    for p in (perm, None):
        pxt = torch.empty((T + 1,) + V.shape, dtype=torch.float64, device="cuda")
        engine.covector([engine.CovectorJob(dev(xt), dev(at), dev(px1), K, 0.5, pxt, perm=p)])
        outs.append(pxt.cpu().numpy())
    print("ordered covector, icosphere s5, T=4", flush=True)
    ref, _ = orc.landmarkHamiltonianCovector(V, at, px1, orc.Kernel('gauss', sigma=0.3), 0.5)
    assert rel_err(outs[0] - px1, ref - px1) < 5 * TOL
    assert rel_err(outs[0] - px1, outs[1] - px1) < 5 * TOL


def test_ordered_rkhs_products():
    """<g2_j[t], K(x_t) g1[t]> in the processing order against the natural order and the
    oracle (scale: the Cauchy-Schwarz bound sqrt(<g1, K g1><g2, K g2>))."""
    import torch
    from registration_amd import engine, synthetic
    V, _ = synthetic.icosphere(5)
    rng = np.random.default_rng(5)
    T = 3
    at = 0.02 * rng.standard_normal((T,) + V.shape)
    perm = engine.spatial_order_dev(V)
    K = gauss(0.3)
    xt, _, _ = flow(V, at, K, perm)
    g1 = rng.standard_normal((T,) + V.shape)
    g2 = rng.standard_normal((T,) + V.shape)
    dev = lambda a: torch.as_tensor(a, device="cuda")
    outs = []
    for p in (perm, None):
        out = torch.zeros((T, 2), dtype=torch.float64, device="cuda")
        engine.kdot([engine.KDotJob(dev(xt), dev(g1), [dev(g1), dev(g2)], K, out, perm=p)])
        outs.append(out.cpu().numpy())
    print("ordered RKHS products, icosphere s5", flush=True)
    for t in range(T):
        k1 = orc.apply_k(xt[t], xt[t], g1[t], 0.3, 10)
        k2 = orc.apply_k(xt[t], xt[t], g2[t], 0.3, 10)
        ref = [float((g1[t] * k1).sum()), float((g2[t] * k1).sum())]
        scale = np.sqrt(ref[0] * float((g2[t] * k2).sum()))
        for j in range(2):
            assert abs(outs[0][t, j] - ref[j]) <= 1e-6 * scale
            assert abs(outs[0][t, j] - outs[1][t, j]) <= 1e-6 * scale


def test_invalid_order_is_rejected_on_the_host():
    import torch
    from registration_amd import engine
    K = gauss(0.3)
    with pytest.raises(ValueError):
        engine.flow_record(10, 2, K, perm=torch.zeros(9, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        engine.flow_record(10, 2, K, perm=torch.zeros(10, dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_staged_tiles_match_the_natural_order(precision):
    """Ordered segments read their source tiles from the launch's staging area (built once per
    launch by k_stage); segments without an order build each tile in the pair kernel.  A batch
    of ordered and natural segments (staged tiles numbered past a natural segment; ragged last
    tiles; one cloud smaller than a tile) gives the same flows, covectors and RKHS products as
    the same batch with every segment natural: to rounding in FP64, and within the FP32 bar
    (the ordered FP32 sums use the dot-product exponent)."""
    import torch
    from registration_amd import engine, synthetic
    from registration_amd.kernelFunctions import Kernel
    K = engine.DevKernel.of(Kernel('gauss', sigma=0.3, precision=precision))
    rng = np.random.default_rng(77)
    clouds = [synthetic.icosphere(4)[0], synthetic.icosphere(3)[0], synthetic.icosphere(5)[0],
              rng.standard_normal((100, 3))]
    T = 2
    ats = [0.02 * rng.standard_normal((T,) + X.shape) for X in clouds]
    dev = lambda a: torch.as_tensor(a, device="cuda")

    def run(ordered):
        perms = [engine.spatial_order_dev(X) if o else None for X, o in zip(clouds, ordered)]
        jobs = [engine.FlowJob(dev(X), dev(a), K,
                               torch.empty((T + 1,) + X.shape, dtype=torch.float64, device="cuda"),
                               torch.zeros(T, dtype=torch.float64, device="cuda"), perm=p)
                for X, a, p in zip(clouds, ats, perms)]
        engine.shoot(jobs)
        cov = [engine.CovectorJob(j.xt, dev(a), dev(np.ones(X.shape)), K, 0.5,
                                  torch.empty((T + 1,) + X.shape, dtype=torch.float64, device="cuda"),
                                  perm=p)
               for j, X, a, p in zip(jobs, clouds, ats, perms)]
        engine.covector(cov)
        kd = [engine.KDotJob(j.xt, dev(a), [dev(a)], K,
                             torch.zeros((T, 2), dtype=torch.float64, device="cuda"), perm=p)
              for j, a, p in zip(jobs, ats, perms)]
        engine.kdot(kd)
        return ([j.xt.cpu().numpy() - X for j, X in zip(jobs, clouds)] +
                [c.pxt.cpu().numpy() for c in cov] + [k.out.cpu().numpy() for k in kd])

    print(f"staged tiles vs natural order, {precision}", flush=True)
    staged = run([True, False, True, True])
    natural = run([False] * 4)
    tol = 1e-12 if precision == "fp64" else 2e-6
    for a, b in zip(staged, natural):
        assert np.isfinite(a).all() and np.abs(b).max() > 0
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_staged_energy_tiles_match_the_oracle(precision):
    """Energy-only sums (currents, measure, RKHS norm: no per-row outputs) stage the tiles of
    both source sets once per launch and evaluate each unordered self pair once.  Totals in
    batches with ragged sizes, a shape smaller than one tile and a measure without target
    match the oracle's sums (relative to the sum of the magnitudes of the terms)."""
    import torch
    from registration_amd import engine
    from registration_amd.kernelFunctions import Kernel
    K = engine.DevKernel.of(Kernel('gauss', sigma=0.2, precision=precision))
    rng = np.random.default_rng(91)
    dev = lambda a: torch.as_tensor(a, device="cuda")
    sizes = [(3000, 2700), (200, 130), (5000, 5100)]
    host = [(rng.standard_normal((n1, 3)), 0.01 * rng.standard_normal((n1, 3)),
             rng.standard_normal((n2, 3)), 0.01 * rng.standard_normal((n2, 3))) for n1, n2 in sizes]
    shapes = [tuple(dev(a) for a in h) for h in host]
    cj = [engine.CurrentsJob(c1, r1, c2, r2, K, torch.zeros(1, dtype=torch.float64, device="cuda"))
          for c1, r1, c2, r2 in shapes]
    engine.currents(cj)
    mj = [engine.MeasureJob(c1, r1[:, 0].abs().contiguous(), c2 if i != 1 else None,
                            r2[:, 0].abs().contiguous() if i != 1 else None, K,
                            torch.zeros(1, dtype=torch.float64, device="cuda"))
          for i, (c1, r1, c2, r2) in enumerate(shapes)]
    engine.measure(mj)
    kj = [engine.KNormJob(c1, r1, K, torch.zeros(1, dtype=torch.float64, device="cuda"))
          for c1, r1, _, _ in shapes]
    engine.knorm(kj)
    print(f"staged energy tiles vs oracle, {precision}", flush=True)

    def quad(c, w, c2=None, w2=None):
        """<w, K(c, c2) w2> and the sum of |terms|' bound |w| . K |w2|."""
        c2 = c if c2 is None else c2
        w2 = w if w2 is None else w2
        return (np.sum(w * orc.apply_k(c, c2, w2, 0.2, 10)),
                np.sum(np.abs(w) * orc.apply_k(c, c2, np.abs(w2), 0.2, 10)))

    tol = 1e-12 if precision == "fp64" else 2e-6
    for i, (c1, r1, c2, r2) in enumerate(host):
        s, ms = quad(c1, r1)
        x, mx = quad(c1, r1, c2, r2)
        assert abs(cj[i].total.item() - (s - 2 * x)) <= tol * (ms + 2 * mx)
        a1, a2 = np.abs(r1[:, :1]), np.abs(r2[:, :1])
        s, ms = quad(c1, a1)
        x, mx = quad(c1, a1, c2, a2) if i != 1 else (0.0, 0.0)
        assert abs(mj[i].total.item() - (s - 2 * x)) <= tol * (ms + 2 * mx)
        s, ms = quad(c1, r1)
        assert abs(kj[i].total.item() - s) <= tol * ms


@pytest.mark.parametrize("prec", ["fp32c", "fp64"])
def test_stepped_totals_folded_into_staging_equal_single_steps(prec):
    """An ordered flow run as one stepped sequence reduces each step's energy in the next
    step's staging launch (the last step in its own totals launch); every step's energy,
    position and velocity equals the same steps launched one at a time, bit for bit, and
    a served-style batch with an ordered and an unordered segment likewise."""
    import torch
    from registration_amd import _lib, engine, runtime as rt, synthetic
    from registration_amd.kernelFunctions import Kernel
    K = engine.DevKernel.of(Kernel('gauss', sigma=0.3, precision=prec))
    V5, _ = synthetic.icosphere(5)
    V3, _ = synthetic.icosphere(3)
    rng = np.random.default_rng(71)
    T = 5
    jobs, ref = [], []
    for X, ordered in ((V5, True), (V3, False), (V5, False)):
        at = torch.as_tensor(0.03 * rng.standard_normal((T,) + X.shape), device="cuda")
        perm = engine.spatial_order_dev(X) if ordered else None
        x0 = torch.as_tensor(X, device="cuda")
        xt = torch.empty((T + 1,) + X.shape, dtype=torch.float64, device="cuda")
        v = torch.empty((T,) + X.shape, dtype=torch.float64, device="cuda")
        en = torch.full((T,), float("nan"), dtype=torch.float64, device="cuda")
        jobs.append(engine.FlowJob(x0, at, K, xt, en, v=v, perm=perm))
        # the same steps one launch at a time
        yt = torch.empty_like(xt)
        yt[0] = x0
        w = torch.empty_like(v)
        e1 = torch.full((T,), float("nan"), dtype=torch.float64, device="cuda")
        for t in range(T):
            f = dict(y=yt[t], w0=at[t], out0=yt[t + 1], out1=w[t], total=e1[t:t + 1],
                     ny=X.shape[0], sigma=K.sigma, dt=1.0 / T, kernel=K.kernel, order=K.order)
            if perm is not None:
                f["perm"] = perm
            rt.Plan(_lib.OP_FLOW_STEP, K.prec, 3, [rt.seg(**f)]).run()
        ref.append((yt, w, e1))
    print(f"stepped flows ({prec}): one ordered batch of 3 and 3 single-step references",
          flush=True)
    engine.shoot(jobs)
    engine.shoot(jobs[:1])                      # the ordered one alone
    torch.cuda.synchronize()
    for j, (yt, w, e1) in zip(jobs, ref):
        assert torch.equal(j.xt, yt) and torch.equal(j.v, w)
        assert torch.equal(j.energy, e1) and not torch.isnan(e1).any()


@pytest.mark.parametrize("prec,tol", [("fp32c", TOL), ("fp64", 1e-12)])
def test_laplacian_flow_in_spatial_order(prec, tol):
    """Laplacian sums take an order from 32,768 points, the spatial one where it culls more
    (FP32C at sigma 0.05 on the unit sphere; in FP64 the cull radius, 55 sigma, exceeds the
    sphere and the points' own order is kept): a flow step over the 40,962-point sphere in
    the spatial order equals the natural order's and the oracle's velocities on sampled
    rows."""
    import torch
    from registration_amd import engine, synthetic
    from registration_amd.kernelFunctions import Kernel
    V, _ = synthetic.icosphere(6)
    K = engine.DevKernel.of(Kernel('laplacian', sigma=0.05, order=3, precision=prec))
    o = engine.order_choice(V, K)
    if prec == "fp32c":
        assert o is not None and not np.array_equal(o, np.arange(len(V)))
    else:
        assert np.array_equal(o, np.arange(len(V)))
        o = engine.spatial_order(V)
    at = 0.03 * np.random.default_rng(97).standard_normal((1,) + V.shape)
    perm = torch.as_tensor(o, device="cuda")
    print(f"Laplacian flow step in the spatial order ({prec})", flush=True)
    _, v, e = flow(V, at, K, perm)
    _, vn, en = flow(V, at, K, None)
    rows = np.random.default_rng(5).choice(len(V), 256, replace=False)
    ref = orc.apply_k(V[rows], V, at[0], 0.05, 3)
    assert rel_err(v[0][rows], ref) < tol
    assert rel_err(v, vn) < tol
    assert abs(e[0] - en[0]) <= 1e-6 * abs(en[0])
