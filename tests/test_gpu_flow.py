"""GPU landmark flow, covector, gradient and EPDiff against the FP64 oracle, and the adjoint
against central finite differences of the GPU objective."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


def kernels(prec, name='gauss', sigma=0.4, order=10):
    from registration_amd import kernelFunctions as kfun
    return kfun.Kernel(name=name, sigma=sigma, order=order, precision=prec), \
        orc.Kernel(name=name, sigma=sigma, order=order)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
@pytest.mark.parametrize("name,order", [("gauss", 10), ("laplacian", 1), ("laplacian", 3)])
def test_flow_and_covector(prec, tol, name, order):
    from registration_amd import pointEvolution as pe
    rng = np.random.default_rng(order)
    N, T = 700, 6
    x0 = rng.standard_normal((N, 3))
    at = 0.05 * rng.standard_normal((T, N, 3))
    px1 = rng.standard_normal((N, 3))
    Kg, Ko = kernels(prec, name, 0.5, order)
    print(f"flow {name}{order} {prec}", flush=True)
    xt = pe.landmarkDirectEvolutionEuler(x0, at, Kg)
    xr = orc.landmarkDirectEvolutionEuler(x0, at, Ko)
    assert rel_err(xt, xr) < tol
    pxt, xt2 = pe.landmarkHamiltonianCovector(x0, at, px1, Kg, 0.7)
    pr, _ = orc.landmarkHamiltonianCovector(x0, at, px1, Ko, 0.7)
    assert rel_err(pxt, pr) < 10 * tol
    dat, _ = pe.landmarkHamiltonianGradient(x0, at, px1, Kg, 0.7)
    dr, _ = orc.landmarkHamiltonianGradient(x0, at, px1, Ko, 0.7)
    assert rel_err(dat, dr) < 10 * tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_flow_options(prec, tol):
    from registration_amd import pointEvolution as pe
    rng = np.random.default_rng(4)
    N, T = 300, 4
    x0 = rng.standard_normal((N, 3))
    at = 0.05 * rng.standard_normal((T, N, 3))
    pts = rng.standard_normal((120, 3))
    nrm = rng.standard_normal((N, 3))
    Kg, Ko = kernels(prec, sigma=0.6)
    xt, yt, nt = pe.landmarkDirectEvolutionEuler(x0, at, Kg, withPointSet=pts, withNormals=nrm)
    xr, yr, nr = orc.landmarkDirectEvolutionEuler(x0, at, Ko, withPointSet=pts, withNormals=nrm)
    assert rel_err(xt, xr) < tol and rel_err(yt, yr) < tol and rel_err(nt, nr) < 10 * tol
    A = 0.1 * rng.standard_normal((T, 3, 3))
    b = 0.1 * rng.standard_normal((T, 3))
    xa = pe.landmarkDirectEvolutionEuler(x0, at, Kg, affine=[A, b])
    xra = orc.landmarkDirectEvolutionEuler(x0, at, Ko, affine=(A, b))
    assert rel_err(xa, xra) < tol
    pa, _ = pe.landmarkHamiltonianCovector(x0, at, nrm, Kg, 1.0, affine=[A, b])
    pra, _ = orc.landmarkHamiltonianCovector(x0, at, nrm, Ko, 1.0, affine=(A, b))
    assert rel_err(pa, pra) < 10 * tol
    # zero affine arrays (what the matching drivers pass) take the fused path unchanged
    xz = pe.landmarkDirectEvolutionEuler(x0, at, Kg, affine=[np.zeros((T, 3, 3)), np.zeros((T, 3))])
    assert np.array_equal(xz, pe.landmarkDirectEvolutionEuler(x0, at, Kg))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32c", 1e-6)])
def test_epdiff(prec, tol):
    from registration_amd import pointEvolution as pe
    rng = np.random.default_rng(7)
    from registration_amd import synthetic
    x0 = synthetic.ball_points(rng, 2000)
    a0 = 0.01 * rng.standard_normal((2000, 3))
    Kg, Ko = kernels(prec, sigma=0.15)
    xt, at = pe.landmarkEPDiff(8, x0, a0, Kg)
    xr, ar = orc.landmarkEPDiff(8, x0, a0, Ko)
    assert rel_err(xt, xr) < tol and rel_err(at, ar) < 10 * tol


def test_gradient_against_central_differences():
    """Landmark shooting (ball points, Gaussian sigma 0.15): the hand-written adjoint against
    central differences (h = 1e-4) of the discrete objective along seeded directions."""
    import torch
    from registration_amd import engine, pointEvolution as pe
    from registration_amd import kernelFunctions as kfun, synthetic
    rng = np.random.default_rng(3)
    N, T = 3000, 5
    x0 = synthetic.ball_points(rng, N)
    a = 0.01 * rng.standard_normal((T, N, 3))
    y = x0 + 0.05 * rng.standard_normal((N, 3))
    K = kfun.Kernel('gauss', sigma=0.15, precision='fp64')
    lam = 1.0

    def J(at):
        xt = pe.landmarkDirectEvolutionEuler(x0, at, K)
        reg = 0.0
        for t in range(T):
            reg += lam / T * (at[t] * K.applyK(xt[t], at[t])).sum()
        return reg + ((xt[-1] - y) ** 2).sum()

    xt = pe.landmarkDirectEvolutionEuler(x0, a, K)
    px1 = -2 * (xt[-1] - y)
    dat, _ = pe.landmarkHamiltonianGradient(x0, a, px1, K, lam)
    h = 1e-4
    for s in range(8):
        v = np.random.default_rng(100 + s).standard_normal((T, N, 3))
        v /= np.linalg.norm(v)  # unit direction: h is the perturbation length
        fd = (J(a + h * v) - J(a - h * v)) / (2 * h)
        # Euclidean gradient of J wrt a_t is dt K(x_t) dat_t
        an = sum((v[t] * K.applyK(xt[t], dat[t])).sum() / T for t in range(T))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (s, fd, an)
