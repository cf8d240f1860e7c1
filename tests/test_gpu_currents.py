"""GPU currents data term (surfaces and curves) against the FP64 oracle."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return float(np.abs(np.asarray(got) - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-6)])
@pytest.mark.parametrize("subdiv,sigma", [(2, 0.3), (4, 0.1)])
def test_surface_currents(prec, tol, subdiv, sigma):
    from registration_amd import kernelFunctions as kfun, surfaces, synthetic
    V, F = synthetic.icosphere(subdiv)
    rng = np.random.default_rng(subdiv)
    X = V + 0.02 * rng.standard_normal(V.shape)
    Y = synthetic.rbf_field(rng, V)
    K = kfun.Kernel('gauss', sigma=sigma, precision=prec)
    Ko = orc.Kernel('gauss', sigma=sigma)
    S, T = surfaces.Surface(FV=(F, X)), surfaces.Surface(FV=(F, Y))
    So, To = orc.Surface(FV=(F, X)), orc.Surface(FV=(F, Y))
    print(f"surface currents s={subdiv} {prec}", flush=True)
    # currents norms are FP64 in both precisions (energies), the gradient follows the kernel's
    n0 = surfaces.currentNorm0(T, K)
    n0r = orc.surf_currentNorm0(To, Ko)
    assert abs(n0 - n0r) < 1e-12 * abs(n0r)
    nd = surfaces.currentNormDef(S, T, K)
    ndr = orc.surf_currentNormDef(So, To, Ko)
    assert abs(nd - ndr) < 1e-12 * abs(n0r)
    g = surfaces.currentNormGradient(S, T, K)
    gr = orc.surf_currentNormGradient(So, To, Ko)
    assert rel_err(g, gr) < 10 * tol
    # geometry is bit-identical to the reference's NumPy expressions
    import torch
    ds = surfaces.DeviceSurface(F, X)
    assert np.array_equal(ds.centers.cpu().numpy(), So.centers)
    assert np.array_equal(ds.surfel.cpu().numpy(), So.surfel)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 2e-6)])
def test_curve_currents(prec, tol):
    from registration_amd import curves, kernelFunctions as kfun, synthetic
    n = 1500
    X, _ = synthetic.circle(n)
    Y = synthetic.star_curve(n, 5, 0.2, noise=0.01, rng=np.random.default_rng(1))
    K = kfun.Kernel('gauss', sigma=0.1, precision=prec)
    Ko = orc.Kernel('gauss', sigma=0.1)
    C, D = curves.Curve(pointSet=X), curves.Curve(pointSet=Y)
    Co, Do = orc.Curve(pointSet=X), orc.Curve(pointSet=Y)
    n0r = orc.curve_currentNorm0(Do, Ko)
    assert abs(curves.currentNorm0(D, K) - n0r) < 1e-12 * abs(n0r)
    assert abs(curves.currentNormDef(C, D, K) - orc.curve_currentNormDef(Co, Do, Ko)) < 1e-12 * abs(n0r)
    assert rel_err(curves.currentNormGradient(C, D, K), orc.curve_currentNormGradient(Co, Do, Ko)) < 10 * tol


def test_gradient_gather_is_bit_exact_in_fp64_order():
    """With identical z1/dz the vertex gather reproduces the reference's sequential scatter
    (corner-major, face order) bit for bit."""
    import torch
    from registration_amd import engine, synthetic
    V, F = synthetic.icosphere(3)
    rng = np.random.default_rng(0)
    z1, dz = rng.standard_normal((F.shape[0], 3)), rng.standard_normal((F.shape[0], 3))
    px = np.zeros_like(V)
    x1, x2, x3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    for I, crs in ((F[:, 0], np.cross(x3 - x2, z1)), (F[:, 1], np.cross(x1 - x3, z1)),
                   (F[:, 2], np.cross(x2 - x1, z1))):
        for k in range(I.size):
            px[I[k]] = px[I[k]] + dz[k] - crs[k]
    ptr, inc = engine.incidence(F, V.shape[0])
    dev = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device="cuda")
    out = torch.empty((V.shape[0], 3), dtype=torch.float64, device="cuda")
    engine.surface_vertex_gradient(dev(V), dev(F.astype(np.int32), torch.int32), dev(z1), dev(dz),
                                   dev(ptr, torch.int32), dev(inc, torch.int32), 2.0, out)
    assert np.array_equal(out.cpu().numpy(), 2 * px)


def test_batched_geometry_and_gather_equal_single_shape_kernels():
    import torch
    from registration_amd import curves, engine, surfaces, synthetic
    rng = np.random.default_rng(3)
    shapes, verts, singles = [], [], []
    for s in (1, 2, 3):
        V, F = synthetic.icosphere(s)
        X = V + 0.05 * rng.standard_normal(V.shape)
        shapes.append(surfaces.DeviceSurface(F, V))
        verts.append(torch.as_tensor(X, device="cuda"))
        singles.append(surfaces.DeviceSurface(F, X))
    for n in (50, 333):
        X, _ = synthetic.circle(n)
        X = X + 0.01 * rng.standard_normal(X.shape)
        E = synthetic.closed_curve_edges(n)
        shapes.append(curves.DeviceCurve(E, np.zeros_like(X)))
        verts.append(torch.as_tensor(X, device="cuda"))
        singles.append(curves.DeviceCurve(E, X))
    engine.geometry_batch(shapes, verts)
    for a, b in zip(shapes, singles):
        assert torch.equal(a.centers, b.centers) and torch.equal(a.surfel, b.surfel)
    z1s = [torch.randn_like(sh.centers) for sh in shapes]
    dzs = [torch.randn_like(sh.centers) for sh in shapes]
    pxs = [torch.empty((sh.nv, sh.centers.shape[1]), dtype=torch.float64, device="cuda") for sh in shapes]
    engine.gather_batch(shapes, z1s, dzs, [0.5 + i for i in range(len(shapes))], pxs)
    for i, sh in enumerate(shapes):
        ref = torch.empty_like(pxs[i])
        if sh.faces.shape[1] == 3:
            engine.surface_vertex_gradient(sh.vertices, sh.faces, z1s[i], dzs[i], sh.ptr, sh.inc, 0.5 + i, ref)
        else:
            engine.curve_vertex_gradient(z1s[i], dzs[i], sh.ptr, sh.inc, sh.nv, 0.5 + i, ref)
        assert torch.equal(pxs[i], ref)
