"""Processing order of a point set (engine.spatial_order over lddmm_spatial_order, host code of
the kernel library): a permutation, deterministic, and spatially compact at the pair kernels'
tile and quarter sizes; lddmm_order_compactness and lddmm_order_live_fraction against NumPy
restatements; the driver's choice of an order.  Host only."""
import numpy as np
import pytest

from registration_amd import engine, synthetic

PRESCALE = np.sqrt(np.log2(np.e) / 2)  # FP32 Gaussian coordinates: x * PRESCALE / sigma


def _K(kind, sigma, prec='fp32c'):
    from registration_amd.kernelFunctions import Kernel
    k = (Kernel('laplacian', sigma=sigma, order=3, precision=prec) if kind == 'lap'
         else Kernel('gauss', sigma=sigma, precision=prec))
    return engine.DevKernel.of(k)


def test_order_is_taken_where_it_pays():
    """Measured on gfx950 (flow and covector steps over icospheres): an order costs a staging
    launch that small sets do not repay, so none below MIN_ORDERED_POINTS (Laplacian:
    MIN_ORDERED_POINTS_LAPLACIAN); above, the spatial order where it culls more or takes
    more dot-product quarters (the large surface at every width, FP32C and FP64, and FP64 at
    a width that culls nothing, where the more compact order wins the tie)."""
    V6, _ = synthetic.icosphere(6)
    V5, _ = synthetic.icosphere(5)
    V4, _ = synthetic.icosphere(4)
    ident6 = np.arange(len(V6), dtype=np.int32)
    for K in (_K('gauss', 0.25), _K('gauss', 0.5), _K('gauss', 0.1), _K('gauss', 0.1, 'fp64'),
              _K('lap', 0.05)):
        o = engine.order_choice(V6, K)
        assert o is not None and not np.array_equal(o, ident6), (K.kernel, K.sigma, K.prec)
    o = engine.order_choice(V6, _K('gauss', 0.5, 'fp64'))
    assert engine.live_fraction(V6, o, _K('gauss', 0.5, 'fp64')) == 1.0
    assert not np.array_equal(o, ident6)
    assert engine.order_choice(V4, _K('gauss', 0.5)) is None       # 2,562 points
    assert engine.order_choice(V5, _K('lap', 0.05)) is None        # Laplacian, 10,242 points
    assert engine.order_choice(V5, _K('gauss', 0.05)) is not None
    assert engine.wants_order(_K('gauss', 0.25)) and not engine.wants_order(_K('gauss', 0.25, 'fp64'))
    assert not engine.wants_order(_K('lap', 0.25))


def test_curves_keep_their_arc_order():
    """A sampled curve is compact in arc order: the choice is the identity permutation."""
    X = synthetic.star_curve(16384, 5, 0.2, noise=0.0, rng=np.random.default_rng(0))
    for K in (_K('gauss', 0.3), _K('gauss', 0.05), _K('gauss', 0.1, 'fp64')):
        assert np.array_equal(engine.order_choice(X, K), np.arange(len(X)))


def live_fraction_ref(P, order, rows, cut2):
    """Restatement of lddmm_order_live_fraction: boxes of 64-point quarters of the order
    (rows past the end repeat the last point), row blocks of rows / 64 quarters, the
    fraction of (row block, quarter) box pairs within squared distance cut2."""
    idx = np.asarray(order, dtype=np.int64)
    nq = -(-len(idx) // 64)
    idx = np.concatenate([idx, np.full(nq * 64 - len(idx), idx[-1])])
    Q = P[idx].reshape(nq, 64, P.shape[1])
    lo, hi = Q.min(axis=1), Q.max(axis=1)
    per = rows // 64
    nr = -(-nq // per)
    rl = np.array([lo[r * per:(r + 1) * per].min(axis=0) for r in range(nr)])
    rh = np.array([hi[r * per:(r + 1) * per].max(axis=0) for r in range(nr)])
    g = np.maximum(np.maximum(lo[None] - rh[:, None], rl[:, None] - hi[None]), 0.0)
    return float(((g * g).sum(axis=2) <= cut2).mean())


@pytest.mark.parametrize("n,d,sigma,prec", [(1, 3, 0.3, 'fp32c'), (65, 2, 0.3, 'fp32c'),
                                            (700, 3, 0.1, 'fp64'), (5000, 3, 0.2, 'fp32c'),
                                            (40962, 3, 0.1, 'fp64')])
def test_live_fraction_matches_a_numpy_restatement(n, d, sigma, prec):
    P = (synthetic.icosphere(6)[0] if n == 40962
         else np.random.default_rng(n).standard_normal((n, d)))
    K = _K('gauss', sigma, prec)
    rows = 128 if prec == 'fp32c' else 64
    for o in (engine.spatial_order(P), np.arange(n, dtype=np.int32)):
        ref = live_fraction_ref(P, o, rows, engine.cull_radius2(K))
        assert engine.live_fraction(P, o, K) == pytest.approx(ref, abs=1e-12)


def test_cull_radius_restates_the_kernel_bounds():
    """The Gaussian at the cull radius is 2^-24 (FP32C) or 2^-56 (FP64) of its peak; the
    Laplacian radius is 30.5 sigma (FP32C) or 55 sigma (FP64)."""
    for prec, bits in (('fp32c', 24), ('fp64', 56)):
        r2 = engine.cull_radius2(_K('gauss', 0.2, prec))
        assert np.exp(-r2 / (2 * 0.2 ** 2)) == pytest.approx(2.0 ** -bits, rel=1e-12)
    assert engine.cull_radius2(_K('lap', 0.2)) == pytest.approx((30.5 * 0.2) ** 2)
    assert engine.cull_radius2(_K('lap', 0.2, 'fp64')) == pytest.approx((55 * 0.2) ** 2)


def quarter_r2(P, order, sigma):
    """Max |y - origin|'^2 of every 64-source quarter, the origin being the mean of the
    tile's sources 32, 96, 160, 224 (clamped), as the pair kernel takes it."""
    Q = P[order] * PRESCALE / sigma
    out = []
    for base in range(0, len(Q), engine.TILE):
        t = Q[base:base + engine.TILE]
        org = t[[min(i, len(t) - 1) for i in (32, 96, 160, 224)]].mean(axis=0)
        for q in range(0, len(t), engine.QUARTER):
            out.append(((t[q:q + engine.QUARTER] - org) ** 2).sum(axis=1).max())
    return np.array(out)


@pytest.mark.parametrize("n", [0, 1, 5, 64, 65, 255, 256, 257, 1000, 4097])
def test_order_is_a_permutation(n):
    rng = np.random.default_rng(n)
    P = rng.standard_normal((n, 3))
    o = engine.spatial_order(P)
    assert o.dtype == np.int32 and o.shape == (n,)
    assert np.array_equal(np.sort(o), np.arange(n))
    assert np.array_equal(o, engine.spatial_order(P.copy()))  # deterministic


def test_tiles_are_subtrees():
    """Split points fall on multiples of 256 above the tile size, so every tile of the order
    is one subtree: no tile straddles two distant patches.  A 256-point patch of the unit
    sphere at 40,962 points has a diameter near 0.32 (box diagonal ~0.45); the icosphere's
    own vertex order has tiles spanning the sphere."""
    V, _ = synthetic.icosphere(6)

    def extents(o):
        return np.array([np.linalg.norm(V[o[b:b + engine.TILE]].max(0) - V[o[b:b + engine.TILE]].min(0))
                         for b in range(0, len(o) - engine.TILE, engine.TILE)])
    ext = extents(engine.spatial_order(V))
    assert np.median(ext) < 0.5 and ext.max() < 0.7
    assert np.median(extents(np.arange(len(V)))) > 0.9


def test_icosphere_quarters_take_the_dot_product_form():
    """On the large surface (icosphere s6, sigma 0.25) every quarter of the order lies within
    R'^2 <= 2 of its tile origin, the bound under which the FP32 kernel uses the dot-product
    exponent; in the icosphere's own vertex order only about half of them do."""
    V, _ = synthetic.icosphere(6)
    r2 = quarter_r2(V, engine.spatial_order(V), 0.25)
    assert (r2 <= 2.0).mean() > 0.99
    assert (quarter_r2(V, np.arange(len(V)), 0.25) <= 2.0).mean() < 0.6


@pytest.mark.parametrize("n,sigma", [(700, 0.3), (5000, 0.2), (40962, 0.25)])
def test_compactness_matches_a_numpy_restatement(n, sigma):
    V = synthetic.icosphere(6)[0] if n == 40962 else np.random.default_rng(n).standard_normal((n, 3))
    o = engine.spatial_order(V)
    ref = float((quarter_r2(V, o, sigma) <= engine.EXP_R2MAX).mean())
    assert engine.compact_fraction(V, o, sigma) == pytest.approx(ref, abs=1e-12)


def test_order_rejects_bad_input():
    with pytest.raises(ValueError):
        engine.spatial_order(np.zeros((10, 4)))
    with pytest.raises(ValueError):
        engine.compact_fraction(np.zeros((10, 3)), np.arange(9), 0.3)
    with pytest.raises(ValueError):
        engine.live_fraction(np.zeros((10, 3)), np.arange(9), _K('gauss', 0.3))
    with pytest.raises(ValueError):
        engine.live_fraction(np.zeros((10, 3)), np.arange(1, 11), _K('gauss', 0.3))


def test_order_ignores_duplicates_and_flat_sets():
    P = np.zeros((300, 2))
    P[:, 0] = np.repeat(np.arange(3), 100)
    o = engine.spatial_order(P)
    assert np.array_equal(np.sort(o), np.arange(300))
    o1 = engine.spatial_order(np.ones((130, 3)))
    assert np.array_equal(np.sort(o1), np.arange(130))


def _bisect_leaves_ref(P, idx, out):
    """Restatement of the bisection: a range of more than 64 points splits at a multiple of
    256 (64 below 256 points) near its middle, along the longest side of its bounding box,
    the low part being the smallest points by (coordinate, index); the leaves as sets."""
    n = len(idx)
    if n <= 64:
        out.append(frozenset(idx.tolist()))
        return
    unit = 256 if n > 256 else 64
    h = int(np.floor(n / (2.0 * unit) + 0.5)) * unit
    h = min(max(h, unit), n - 1)
    Q = P[idx]
    ax = int(np.argmax(Q.max(axis=0) - Q.min(axis=0)))
    srt = idx[np.lexsort((idx, Q[:, ax]))]
    _bisect_leaves_ref(P, srt[:h], out)
    _bisect_leaves_ref(P, srt[h:], out)


@pytest.mark.parametrize("n,d", [(40000, 3), (33000, 2)])
def test_threaded_bisection_matches_a_restatement(n, d):
    """Sets of 32,768 points and more finish their subtrees on host threads; the order's
    64-point leaves equal those of a sequential NumPy restatement, and two calls agree."""
    P = np.random.default_rng(n).standard_normal((n, d))
    o = engine.spatial_order(P)
    assert np.array_equal(o, engine.spatial_order(P.copy()))
    got = [frozenset(o[i:i + 64].tolist()) for i in range(0, n, 64)]
    ref = []
    _bisect_leaves_ref(P, np.arange(n), ref)
    assert got == ref
