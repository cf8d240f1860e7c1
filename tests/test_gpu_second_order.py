"""GPU second-order (spline) flow, Hamiltonian, covector and gradient against the oracle
restatement of pointEvolution.py:882-1084."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


def case(N=400, T=6, seed=3):
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((N, 3)) * 0.5
    a0 = 0.3 * rng.standard_normal((N, 3))
    rhot = 0.2 * rng.standard_normal((T, N, 3))
    A = 0.05 * rng.standard_normal((T, 3, 3))
    b = 0.05 * rng.standard_normal((T, 3))
    return rng, x0, a0, rhot, (A, b)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32c", 2e-5)])
@pytest.mark.parametrize("name,order", [("gauss", 10), ("laplacian", 3)])
def test_second_order_evolution_and_gradient(prec, tol, name, order):
    from registration_amd import kernelFunctions as kfun, pointEvolution as pe
    rng, x0, a0, rhot, aff = case()
    K = kfun.Kernel(name, sigma=0.6, order=order, precision=prec)
    Ko = orc.Kernel(name, sigma=0.6, order=order)
    print(f"second order {name} {prec}", flush=True)
    z0 = rng.standard_normal((50, 3)) * 0.5
    got = pe.secondOrderEvolution(x0, a0, rhot, K, withJacobian=True, withPointSet=z0, affine=aff)
    ref = orc.secondOrderEvolution(x0, a0, rhot, Ko, withJacobian=True, withPointSet=z0, affine=aff)
    for g, r in zip(got, ref):
        assert rel_err(g, r) < tol
    N = x0.shape[0]
    px1 = [rng.standard_normal((N, 3)), rng.standard_normal((N, 3))]
    pa1 = [np.zeros((N, 3)), 0.1 * rng.standard_normal((N, 3))]
    g = pe.secondOrderGradient(x0, a0, rhot, px1, pa1, K, affine=aff, controlWeight=0.7,
                               getCovector=True)
    r = orc.secondOrderGradient(x0, a0, rhot, px1, pa1, Ko, affine=aff, controlWeight=0.7,
                                getCovector=True)
    assert len(g) == len(r) == 8
    for a, b in zip(g, r):
        assert rel_err(a, b) < 10 * tol
    H = pe.secondOrderHamiltonian(x0, a0, rhot[0], px1[0], pa1[1], K, affine=(aff[0][0], aff[1][0]))
    Hr = orc.secondOrderHamiltonian(x0, a0, rhot[0], px1[0], pa1[1], Ko, affine=(aff[0][0], aff[1][0]))
    assert abs(H - Hr) < 10 * tol * abs(Hr)


def test_second_order_without_affine_leaves_drhot_zero():
    from registration_amd import kernelFunctions as kfun, pointEvolution as pe
    rng, x0, a0, rhot, _ = case(N=100, T=4)
    K = kfun.Kernel('gauss', sigma=0.5, precision='fp64')
    px1, pa1 = [rng.standard_normal(x0.shape)], [np.zeros(x0.shape)]
    da0, drhot, xt, at = pe.secondOrderGradient(x0, a0, rhot, px1, pa1, K)
    r = orc.secondOrderGradient(x0, a0, rhot, px1, pa1, orc.Kernel('gauss', sigma=0.5))
    assert not np.asarray(drhot).any() and rel_err(da0, r[0]) < 1e-11
