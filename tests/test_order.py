"""Processing order of a point set (engine.spatial_order over lddmm_spatial_order, host code of
the kernel library): a permutation, deterministic, and spatially compact at the pair kernels'
tile and quarter sizes; lddmm_order_compactness against a NumPy restatement.  Host only."""
import numpy as np
import pytest

from registration_amd import engine, synthetic

PRESCALE = np.sqrt(np.log2(np.e) / 2)  # FP32 Gaussian coordinates: x * PRESCALE / sigma


def test_order_is_taken_where_it_pays():
    """The driver takes the order for the large surface at sigma 0.25 (every quarter compact)
    and not for the 642-point sphere at sigma 0.3 (no quarter compact: only gathers), nor for
    kernels without the dot-product form."""
    from registration_amd.kernelFunctions import Kernel
    V6, _ = synthetic.icosphere(6)
    V3, _ = synthetic.icosphere(3)
    g = lambda s, p='fp32c': engine.DevKernel.of(Kernel('gauss', sigma=s, precision=p))
    assert engine.compact_fraction(V6, engine.spatial_order(V6), 0.25) > 0.99
    assert engine.compact_fraction(V3, engine.spatial_order(V3), 0.3) < 0.1
    assert engine.wants_order(g(0.25)) and not engine.wants_order(g(0.25, 'fp64'))
    assert not engine.wants_order(engine.DevKernel.of(Kernel('laplacian', sigma=0.25, order=2,
                                                            precision='fp32c')))


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


def test_order_ignores_duplicates_and_flat_sets():
    P = np.zeros((300, 2))
    P[:, 0] = np.repeat(np.arange(3), 100)
    o = engine.spatial_order(P)
    assert np.array_equal(np.sort(o), np.arange(300))
    o1 = engine.spatial_order(np.ones((130, 3)))
    assert np.array_equal(np.sort(o1), np.arange(130))
