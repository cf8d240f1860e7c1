"""GPU applyDivergence and the Jacobian flow against the oracle's dense restatement of the
reference (kernelFunctions.py:510-530, pointEvolution.py:666-730)."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-5)])
@pytest.mark.parametrize("name,order,sigma", [("gauss", 10, 0.5), ("laplacian", 1, 0.4),
                                              ("laplacian", 2, 0.3), ("laplacian", 3, 0.5),
                                              ("laplacian", 4, 0.6)])
@pytest.mark.parametrize("d", [2, 3])
def test_apply_divergence(prec, tol, name, order, sigma, d):
    from registration_amd import kernelFunctions as kfun
    rng = np.random.default_rng(order * 10 + d)
    x = rng.standard_normal((700, d)) * 0.6
    a = rng.standard_normal((700, d))
    y = rng.standard_normal((333, d)) * 0.6
    K = kfun.Kernel(name, sigma=sigma, order=order, precision=prec)
    Ko = orc.Kernel(name, sigma=sigma, order=order)
    print(f"divergence {name}{order} d={d} {prec}", flush=True)
    assert rel_err(K.applyDivergence(x, a), Ko.applyDivergence(x, a)) < tol
    assert rel_err(K.applyDivergence(x, a, firstVar=y), Ko.applyDivergence(x, a, firstVar=y)) < tol


def test_apply_divergence_edge_cases():
    from registration_amd import kernelFunctions as kfun
    K = kfun.Kernel('laplacian', sigma=0.3, order=0, precision='fp64')
    x = np.zeros((3, 3))
    with pytest.raises(ValueError):
        K.applyDivergence(x, x)               # order 0 has no derivative
    Kn = kfun.Kernel('foo', sigma=0.3)         # unknown names become 'none'
    assert Kn.applyDivergence(x, x).shape == (1, 3)
    # the dense matrices carry Kernel.weight (kernelFunctions.py:298)
    rng = np.random.default_rng(0)
    xs, As = rng.standard_normal((40, 3)), rng.standard_normal((40, 3))
    Kw = kfun.Kernel('gauss', sigma=0.5, weight=2.5, precision='fp64')
    ref = orc.Kernel('gauss', sigma=0.5, weight=2.5).applyDivergence(xs, As)
    assert rel_err(Kw.applyDivergence(xs, As), ref) < 1e-12
    # one point: the velocity field's divergence at its own source is zero
    Kg = kfun.Kernel('gauss', sigma=0.3, precision='fp64')
    assert float(np.abs(Kg.applyDivergence(np.ones((1, 3)), np.ones((1, 3)))).max()) == 0.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-5)])
def test_jacobian_flow(prec, tol):
    from registration_amd import kernelFunctions as kfun, pointEvolution as pe
    rng = np.random.default_rng(2)
    N, T = 500, 6
    x0 = rng.standard_normal((N, 3)) * 0.5
    at = 0.2 * rng.standard_normal((T, N, 3))
    y0 = rng.standard_normal((90, 3)) * 0.5
    A = 0.1 * rng.standard_normal((T, 3, 3))
    b = 0.1 * rng.standard_normal((T, 3))
    K = kfun.Kernel('gauss', sigma=0.4, precision=prec)
    Ko = orc.Kernel('gauss', sigma=0.4)
    print(f"jacobian flow {prec}", flush=True)
    xt, Jt = pe.landmarkDirectEvolutionEuler(x0, at, K, withJacobian=True)
    xo, Jo = orc.landmarkDirectEvolutionEuler(x0, at, Ko, withJacobian=True)
    assert rel_err(xt, xo) < tol and rel_err(Jt, Jo) < tol
    out = pe.landmarkDirectEvolutionEuler(x0, at, K, affine=(A, b), withPointSet=y0,
                                          withJacobian=True)
    ref = orc.landmarkDirectEvolutionEuler(x0, at, Ko, affine=(A, b), withPointSet=y0,
                                           withJacobian=True)
    assert len(out) == 3 and out[2].shape == (T + 1, 90)
    for g, r in zip(out, ref):
        assert rel_err(g, r) < tol
