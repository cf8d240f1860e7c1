"""Host geometry of the saveRate snapshots: Surface.computeVertexArea / computeVertexNormals
(surfaces.py:77-139) against the oracle's face-by-face restatement, the oracle pinned by
closed forms (equilateral, right and obtuse triangles; vertex areas summing to the surface
area; radial normals on a sphere), and the planar deformation grid of the curve driver
(curveMatching.py:84-96, grid.py:18-30)."""
import numpy as np
import pytest

from oracle import lddmm_oracle as orc
from registration_amd import curveMatching as cm, surfaces, synthetic


def test_vertex_area_closed_forms():
    h = np.sqrt(3) / 2
    tri = np.array([[0, 1, 2]])
    AV, AF = orc.vertex_area(np.array([[0., 0, 0], [1, 0, 0], [0.5, h, 0]]), tri)
    assert np.allclose(AF, h / 2) and np.allclose(AV[:, 0], h / 6)        # equilateral: thirds
    AV, AF = orc.vertex_area(np.array([[0., 0, 0], [1, 0, 0], [-0.5, 0.5, 0]]), tri)
    assert np.allclose(AV[:, 0], AF[0, 0] * np.array([0.5, 0.25, 0.25]))  # obtuse at vertex 0
    AV, AF = orc.vertex_area(np.array([[0., 0, 0], [4, 0, 0], [0.2, 0.5, 0]]), tri)
    assert np.allclose(AV[:, 0], AF[0, 0] * np.array([0.25, 0.25, 0.5]))  # obtuse at vertex 2
    # right angle at vertex 0 (c1 = 0: the Voronoi split with cot 0): legs 1 and 2
    AV, AF = orc.vertex_area(np.array([[0., 0, 0], [1, 0, 0], [0, 2, 0]]), tri)
    assert np.isclose(AF[0, 0], 1.0)
    assert np.allclose(AV[:, 0], [(1 * 2 + 4 * 0.5) / 8, (1 * 2 + 5 * 0) / 8, (4 * 0.5 + 5 * 0) / 8])


@pytest.mark.parametrize("subdiv", [1, 3])
def test_vertex_fields_match_oracle(subdiv):
    V, F = synthetic.icosphere(subdiv)
    rng = np.random.default_rng(subdiv)
    V = V * (1 + 0.2 * rng.random((V.shape[0], 1)))      # mixed obtuse and acute faces
    s = surfaces.Surface(FV=(F, V))
    AV, AF = s.computeVertexArea()
    oAV, oAF = orc.vertex_area(V, F)
    assert np.allclose(AV, oAV, rtol=1e-13, atol=0) and np.allclose(AF, oAF, rtol=1e-13, atol=0)
    assert np.isclose(oAV.sum(), oAF.sum(), rtol=0.05)   # Voronoi/obtuse split keeps the area
    n = s.computeVertexNormals()
    on = orc.vertex_normals(V, F)
    assert np.abs(n - on).max() < 1e-14
    r = np.linalg.norm(V, axis=1)
    sph = surfaces.Surface(FV=(F, V / r[:, None]))
    nr = sph.computeVertexNormals()
    assert np.abs(np.abs((nr * (V / r[:, None])).sum(axis=1)) - 1).max() < 0.02


def test_deformation_grid_layout():
    X, _ = synthetic.circle(64)
    Y = 1.3 * X + 0.1
    pts, edges = cm.deformation_grid(X, Y)
    lo = np.minimum(X.min(0), Y.min(0))
    hi = np.maximum(X.max(0), Y.max(0))
    step = 0.025 * min(hi - lo)
    xs = np.arange(lo[0] - 10 * step, hi[0] + 10 * step, step)
    ys = np.arange(lo[1] - 10 * step, hi[1] + 10 * step, step)
    n, m = len(xs), len(ys)
    assert pts.shape == (n * m, 2)
    assert np.allclose(pts[:m, 0], xs[0]) and np.allclose(pts[:m, 1], ys)   # row-major, x outer
    assert np.array_equal(edges, orc.grid_lines(n, m))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_centers_and_surfels_are_the_numpy_expressions(seed):
    """Surface.computeCentersAreas on coordinate rows gives the bits of the reference's
    expressions (surfaces.py:62-70: gathered vertex blocks, np.cross), and surfVolume agrees
    with det[x1; x2; x3] / 6 summed over the faces; closed form: the unit cube's volume."""
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((500, 3)) * 10.0 ** rng.integers(-3, 4)
    F = rng.integers(0, 500, size=(997, 3))
    S = surfaces.Surface(FV=(F, V))
    x1, x2, x3 = V[F[:, 0], :], V[F[:, 1], :], V[F[:, 2], :]
    assert np.array_equal(S.centers, (x1 + x2 + x3) / 3)
    assert np.array_equal(S.surfel, np.cross(x2 - x1, x3 - x1))
    det = np.linalg.det(np.stack([x1, x2, x3], axis=1))
    assert S.surfVolume() == pytest.approx(det.sum() / 6, rel=1e-9, abs=1e-12 * np.abs(det).sum())
    cube_v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    cube_f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                       [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]])
    assert surfaces.Surface(FV=(cube_f, cube_v)).surfVolume() == pytest.approx(1.0, abs=1e-15)


def test_host_geometry_indexes_as_numpy_does():
    """A negative face index counts from the end (NumPy indexing); one past the vertices
    raises IndexError; non-FP64 vertices take the NumPy expressions with the same results."""
    V = np.random.default_rng(3).standard_normal((10, 3))
    S = surfaces.Surface(FV=(np.array([[0, 1, -1], [2, -3, 4]]), V))
    F = np.array([[0, 1, 9], [2, 7, 4]])
    x1, x2, x3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    assert np.array_equal(S.centers, (x1 + x2 + x3) / 3)
    assert np.array_equal(S.surfel, np.cross(x2 - x1, x3 - x1))
    with pytest.raises(IndexError):
        surfaces.Surface(FV=(np.array([[0, 1, 10]]), V))
    S32 = surfaces.Surface(FV=(F, V))
    S32.vertices = V.astype(np.float32)
    S32.computeCentersAreas()
    v = V.astype(np.float32)
    assert np.array_equal(S32.centers, (v[F[:, 0]] + v[F[:, 1]] + v[F[:, 2]]) / 3)
