"""The hand-written adjoint behind torch.autograd against central finite differences
(h = 1e-4) along 64 seeded momentum directions, with trajectory checkpoints every 5 steps.

(a) 2D closed curve: 4,096-point unit circle, Laplacian order 1, sigma 0.2, T = 30, loss =
    E + |x_T - y|^2 with y the 5-lobe star target (r = 1 + 0.2 cos 5t + N(0, 0.01^2)).
(b) 3D landmarks: 100,000 points uniform in the unit ball, momenta N(0, 0.01^2), Gaussian
    sigma 0.15, T = 10 (recorded choice), the 64-direction sweep at full size, and the
    checkpoint interval's invariance.
Directions are unit vectors, so h is the length of the perturbation.  The FD checks run in FP64
(an FP32 objective is too noisy for h = 1e-4); the FP32-compensated backward is compared with
the FP64 one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def loss_fn(x0, y, K, lam, ck):
    from registration_amd import autograd as ag

    def L(a):
        xT, E = ag.shoot(x0, a, K, lam, ck)
        return E + ((xT - y) ** 2).sum()
    return L


def fd_check(L, a, ndir, h=1e-4, seed=0):
    import torch
    a = a.clone().requires_grad_(True)
    val = L(a)
    val.backward()
    g = a.grad.detach()
    worst = 0.0
    for s in range(ndir):
        gen = torch.Generator(device="cpu").manual_seed(seed + s)
        v = torch.randn(a.shape, generator=gen, dtype=torch.float64).to(a.device)
        v /= v.norm()  # unit direction: h is the length of the perturbation
        with torch.no_grad():
            fd = (L(a + h * v) - L(a - h * v)).item() / (2 * h)
        an = float((g * v).sum())
        worst = max(worst, abs(fd - an) / max(abs(fd), 1e-12))
    return g, worst


def curve_case(prec):
    import torch
    from registration_amd import kernelFunctions as kfun, synthetic
    n, T = 4096, 30
    X, _ = synthetic.circle(n)
    Y = synthetic.star_curve(n, 5, 0.2, noise=0.01, rng=np.random.default_rng(3))
    K = kfun.Kernel('laplacian', sigma=0.2, order=1, precision=prec)
    a = 1e-3 * torch.as_tensor(np.random.default_rng(4).standard_normal((T, n, 2)), device="cuda")
    x0 = torch.as_tensor(X, device="cuda")
    y = torch.as_tensor(Y, device="cuda")
    return x0, y, K, a


def ball_case(prec, n, T=10):
    import torch
    from registration_amd import kernelFunctions as kfun, synthetic
    rng = np.random.default_rng(11)
    X = synthetic.ball_points(rng, n)
    a = torch.as_tensor(0.01 * rng.standard_normal((T, n, 3)), device="cuda")
    y = torch.as_tensor(X + 0.02 * rng.standard_normal(X.shape), device="cuda")
    K = kfun.Kernel('gauss', sigma=0.15, precision=prec)
    return torch.as_tensor(X, device="cuda"), y, K, a


def test_curve_backward_vs_central_differences():
    x0, y, K, a = curve_case('fp64')
    g, worst = fd_check(loss_fn(x0, y, K, 1.0, 5), a, 64)
    assert worst < 1e-6, worst
    # FP32-compensated backward vs FP64
    x0, y, K32, a = curve_case('fp32c')
    a32 = a.clone().requires_grad_(True)
    loss_fn(x0, y, K32, 1.0, 5)(a32).backward()
    err = float((a32.grad - g).abs().max() / g.abs().max())
    assert err < 1e-5, err


def test_ball_full_size_backward_and_checkpoint_invariance():
    import torch
    x0, y, K, a = ball_case('fp64', 100000)
    print("ball: 100,000 points, 64 directions", flush=True)
    g5, worst = fd_check(loss_fn(x0, y, K, 1.0, 5), a, 64)
    assert worst < 1e-6, worst
    # FP32-compensated backward vs FP64
    x32, y32, K32, _ = ball_case('fp32c', 100000)
    a32 = a.clone().requires_grad_(True)
    loss_fn(x32, y32, K32, 1.0, 5)(a32).backward()
    err = float((a32.grad - g5).abs().max() / g5.abs().max())
    assert err < 1e-5, err
    for ck in (1, 10):
        b = a.clone().requires_grad_(True)
        loss_fn(x0, y, K, 1.0, ck)(b).backward()
        assert torch.equal(b.grad, g5), ck


def test_autograd_matches_hamiltonian_gradient():
    """dL/da_t = dt K(x_t) dat_t with dat from landmarkHamiltonianGradient (px1 = -dD/dx_T)."""
    import torch
    from registration_amd import pointEvolution as pe
    x0, y, K, a = ball_case('fp64', 5000, T=6)
    b = a.clone().requires_grad_(True)
    loss_fn(x0, y, K, 0.7, 2)(b).backward()
    xt = pe.landmarkDirectEvolutionEuler(x0, a, K)
    dat, _ = pe.landmarkHamiltonianGradient(x0, a, -2 * (xt[-1] - y), K, 0.7)
    ref = torch.stack([K.applyK(xt[t], dat[t]) for t in range(a.shape[0])]) / a.shape[0]
    err = float((b.grad - ref).abs().max() / ref.abs().max())
    assert err < 1e-12, err
