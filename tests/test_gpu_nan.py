"""Non-finite inputs propagate as in the reference, whose Fortran sums evaluate exp() of every
pair (py-lddmm/kernelFunctions.f90:1-54): a NaN coordinate makes its own row NaN and, through
every pair it takes part in, every row of a sum whose tiles are all within the culling
radius.  The FP64 exponentials clamp far exponents with a select that keeps NaN (not fmax),
and the expanded-exponent path needs no clamp at all."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("name,order", [("gauss", 10), ("laplacian", 2)])
def test_nan_coordinate_gives_nan_rows(prec, name, order):
    from registration_amd.kernelFunctions import Kernel
    rng = np.random.default_rng(5)
    x = 0.2 * rng.standard_normal((700, 3))      # every pair well inside the culling radius
    a = rng.standard_normal((700, 3))
    K = Kernel(name, sigma=1.0, order=order, precision=prec)
    ok = K.applyK(x, a)
    ref = orc.apply_k(x, x, a, 1.0, order if name == "laplacian" else 10)
    assert np.abs(ok - ref).max() <= (1e-12 if prec == "fp64" else 2e-6) * np.abs(ref).max()
    # a NaN target row (firstVar): that row is NaN, the others are untouched
    y = x[:300].copy()
    y[17, 1] = np.nan
    z = K.applyK(x, a, firstVar=y)
    assert np.isnan(z[17]).all()
    rows = np.arange(300) != 17
    assert np.isfinite(z[rows]).all()
    assert np.abs(z[rows] - ok[:300][rows]).max() <= (1e-12 if prec == "fp64" else 2e-6) * np.abs(ok).max()
    # a NaN source: every pair with it is NaN, so every row is
    xs = x.copy()
    xs[123, 0] = np.nan
    z = K.applyK(xs, a, firstVar=x[:300])
    assert np.isnan(z).all()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_nan_vertex_gives_nan_currents_energy(prec):
    """A NaN vertex of the deformed surface makes its faces' centers and normals NaN; the
    currents energy (FP64 energy-only sums, expanded exponent) is then NaN, which the CG's
    'nan in updateTry' guard relies on."""
    import torch
    from registration_amd import surfaces, synthetic
    from registration_amd.kernelFunctions import Kernel
    V, F = synthetic.icosphere(3)
    K = Kernel('gauss', sigma=2.0, precision=prec)
    S1 = surfaces.Surface(FV=(F, V))
    S2 = surfaces.Surface(FV=(F, 1.1 * V))
    e = surfaces.currentNormDef(S1, S2, K)
    assert np.isfinite(e)
    Vn = V.copy()
    Vn[10, 2] = np.nan
    Sn = surfaces.Surface(FV=(F, Vn))
    assert np.isnan(surfaces.currentNormDef(Sn, S2, K))
    del torch
