"""The rest of the Kernel interface on the GPU against the oracle and the reference's own
Fortran outputs: the applyDiffK family, the dense siblings applyDDiffK11 / applyDDiffK12 and
the matrix-weight sums (applyK(matrixWeights=True), applyDiffKmat)."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


def _kernels(name, order, sigma, prec):
    from registration_amd import kernelFunctions as kfun
    return (kfun.Kernel(name, sigma=sigma, order=order, precision=prec),
            orc.Kernel(name, sigma=sigma, order=order))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-5)])
@pytest.mark.parametrize("name,order,sigma", [("gauss", 10, 0.5), ("laplacian", 1, 0.4),
                                              ("laplacian", 3, 0.5), ("laplacian", 4, 0.6)])
def test_diff_k_family(prec, tol, name, order, sigma):
    rng = np.random.default_rng(order)
    x, a1, a2 = rng.standard_normal((600, 3)) * 0.6, rng.standard_normal((600, 3)), rng.standard_normal((600, 3))
    K, Ko = _kernels(name, order, sigma, prec)
    print(f"diff_k {name}{order} {prec}", flush=True)
    for fn in ("applyDiffK", "applyDiffK2", "applyDiffK1and2"):
        assert rel_err(getattr(K, fn)(x, a1, a2), getattr(Ko, fn)(x, a1, a2)) < tol, fn


def test_diff_k_family_reference_golden(golden_mat):
    from registration_amd import kernelFunctions as kfun
    for name, c in golden_mat.items():
        d, ny, nx, m, order, sigma = c["params"]
        order = int(order)
        K = kfun.Kernel('gauss' if order > 4 else 'laplacian', sigma=sigma, order=order,
                        precision='fp64')
        for form, fn in ((1, "applyDiffK"), (2, "applyDiffK2"), (3, "applyDiffK1and2")):
            got = getattr(K, fn)(c["y"], c["dk_a1"], c["dk_a2"])
            assert rel_err(got, c[f"applykdiff_{form}"]) < 1e-12, (name, fn)
        got = K.applyK(c["x"], c["beta"], firstVar=c["y"], matrixWeights=True)
        assert rel_err(got, c["applykmat"]) < 1e-13, name
        got = K.applyDiffKmat(c["x"], c["beta"][..., 0], firstVar=c["y"])
        assert rel_err(got, c["applykdiffmat"]) < 1e-13, name


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-5)])
@pytest.mark.parametrize("name,order,sigma", [("gauss", 10, 0.7), ("laplacian", 2, 0.5),
                                              ("laplacian", 4, 0.6)])
def test_dense_ddiff_siblings(prec, tol, name, order, sigma):
    rng = np.random.default_rng(order + 1)
    x, n, a, p = (rng.standard_normal((500, 3)) * s for s in (0.5, 1, 1, 1))
    K, Ko = _kernels(name, order, sigma, prec)
    print(f"ddiff siblings {name}{order} {prec}", flush=True)
    assert rel_err(K.applyDDiffK11(x, n, a, p), Ko.applyDDiffK11(x, n, a, p)) < tol
    assert rel_err(K.applyDDiffK12(x, n, a, p), Ko.applyDDiffK12(x, n, a, p)) < tol


@pytest.mark.parametrize("name,order,sigma", [("gauss", 10, 0.5), ("laplacian", 3, 0.4)])
def test_matrix_weight_sums(name, order, sigma):
    rng = np.random.default_rng(12)
    x, y = rng.standard_normal((777, 3)) * 0.5, rng.standard_normal((301, 3)) * 0.5
    beta = rng.standard_normal((301, 777, 3))
    for prec in ("fp64", "fp32c"):   # FP64 arithmetic in both: the sums are bound by beta
        K, Ko = _kernels(name, order, sigma, prec)
        assert rel_err(K.applyK(x, beta, firstVar=y, matrixWeights=True),
                       Ko.applyK(x, beta, firstVar=y, matrixWeights=True)) < 1e-13
        assert rel_err(K.applyDiffKmat(x, beta[..., 1], firstVar=y),
                       Ko.applyDiffKmat(x, beta[..., 1], firstVar=y)) < 1e-13
    bx = rng.standard_normal((777, 777, 2))
    assert rel_err(K.applyK(x, bx, matrixWeights=True), Ko.applyK(x, bx, matrixWeights=True)) < 1e-13
    with pytest.raises(ValueError):
        K.applyK(x, beta, matrixWeights=True)      # beta rows must match the targets


@pytest.mark.parametrize("name,order,sigma", [("gauss", 10, 0.5), ("laplacian", 2, 0.4),
                                              ("laplacian", 4, 0.6)])
def test_precompute_dense_matrices(name, order, sigma):
    from registration_amd import kernelFunctions as kfun
    rng = np.random.default_rng(order)
    x, y = rng.standard_normal((120, 3)) * 0.5, rng.standard_normal((70, 3)) * 0.5
    K = kfun.Kernel(name, sigma=sigma, order=order, weight=1.5)
    o = 10 if name == "gauss" else order
    # K, K' and K'' on the self set (with the reference's diagonal fills) and against firstVar
    assert rel_err(K.precompute(x), 1.5 * orc.kernel_matrix(None, x, sigma, o, True)) < 1e-14
    assert rel_err(K.precompute(x, diff=True), 1.5 * orc.kernel_matrix_diff(None, x, sigma, o, True)) < 1e-14
    assert rel_err(K.precompute(x, diff2=True), 1.5 * orc.kernel_matrix_diff2(x, sigma, o)) < 1e-14
    assert rel_err(K.precompute(x, firstVar=y), 1.5 * orc.kernel_matrix(y, x, sigma, o, False)) < 1e-14
    assert rel_err(K.precompute(x, firstVar=y, diff=True),
                   1.5 * orc.kernel_matrix_diff(y, x, sigma, o, False)) < 1e-14
    # grid points keep their leading shape
    g = rng.standard_normal((4, 5, 3))
    Kg = K.precompute(x, grid=g)
    assert Kg.shape == (4, 5, 120)
    assert rel_err(Kg.reshape(20, 120), 1.5 * orc.kernel_matrix(g.reshape(20, 3), x, sigma, o, False)) < 1e-14
    assert kfun.Kernel('foo', sigma=1.0).precompute(x) is None
