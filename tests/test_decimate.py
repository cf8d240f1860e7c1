"""Template decimation (Surface.Simplify -> lddmm_decimate_surface; the reference's
surfaces.py:244-259 calls VTK's vtkQuadricDecimation, absent here, so the result is checked
by properties rather than against VTK's output: parity unpinned).  Host code only, no GPU:
the mesh stays a closed 2-manifold of the same genus, reaches the vertex target, keeps its
volume and surface, is deterministic, and is oriented as the reference orients it."""
import numpy as np
import pytest

from registration_amd import _lib, surfaces, synthetic


def _edges(F):
    e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    return np.unique(e, axis=0, return_counts=True)


@pytest.mark.parametrize("subdiv,target", [(3, 200), (4, 642), (5, 1000)])
def test_closed_manifold_at_target(subdiv, target):
    V, F = synthetic.icosphere(subdiv)
    s = surfaces.Surface(FV=(F, V))
    vol0 = abs(s.surfVolume())
    area0 = np.linalg.norm(s.surfel, axis=1).sum() / 2
    s.Simplify(target)
    nv, nf = s.vertices.shape[0], s.faces.shape[0]
    assert nv == target
    u, c = _edges(s.faces)
    assert set(c.tolist()) == {2}                   # every edge on exactly two faces
    assert nv - len(u) + nf == 2                    # sphere: Euler characteristic 2
    assert s.faces.min() >= 0 and s.faces.max() < nv
    assert abs(abs(s.surfVolume()) - vol0) < 0.02 * vol0
    area = np.linalg.norm(s.surfel, axis=1).sum() / 2
    assert abs(area - area0) < 0.02 * area0
    r = np.linalg.norm(s.vertices, axis=1)
    assert np.abs(r - 1).max() < 0.02               # stays on the sphere
    assert s.surfVolume() < 0                       # flipped when positive, as the reference does


def test_deterministic_and_no_op_above_size():
    V, F = synthetic.icosphere(3)
    a, b = surfaces.Surface(FV=(F, V)), surfaces.Surface(FV=(F, V))
    a.Simplify(300)
    b.Simplify(300)
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.faces, b.faces)
    c = surfaces.Surface(FV=(F, V))
    c.Simplify(10 ** 6)
    assert c.vertices.shape == V.shape and c.faces.shape == F.shape


def test_open_border_stays():
    """A flat grid: decimation keeps it flat and keeps its border on the square's outline."""
    n = 21
    x, y = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    V = np.stack([x.ravel(), y.ravel(), np.zeros(n * n)], axis=1)
    q = np.arange(n * n).reshape(n, n)
    F = np.concatenate([np.stack([q[:-1, :-1].ravel(), q[:-1, 1:].ravel(), q[1:, 1:].ravel()], 1),
                        np.stack([q[:-1, :-1].ravel(), q[1:, 1:].ravel(), q[1:, :-1].ravel()], 1)])
    s = surfaces.Surface(FV=(F, V))
    s.Simplify(60)
    assert s.vertices.shape[0] <= 100
    assert np.abs(s.vertices[:, 2]).max() < 1e-9
    assert s.vertices[:, :2].min() > -1e-9 and s.vertices[:, :2].max() < 1 + 1e-9
    u, c = _edges(s.faces)
    border = u[c == 1]
    on_outline = lambda p: min(p[0], p[1], 1 - p[0], 1 - p[1]) < 1e-9
    assert all(on_outline(s.vertices[i]) for i in np.unique(border))
    area = np.linalg.norm(s.surfel, axis=1).sum() / 2
    assert abs(area - 1.0) < 1e-9


def test_rejects_bad_input():
    lib = _lib.load()
    V, F = synthetic.icosphere(2)
    s = surfaces.Surface(FV=(F, V))
    s.faces = s.faces.copy()
    s.faces[0, 0] = V.shape[0]
    with pytest.raises(ValueError):
        s.Simplify(50)
    V2 = V.copy()
    V2[3, 1] = np.nan
    with pytest.raises(ValueError):
        surfaces.Surface(FV=(F, V2)).Simplify(50)
