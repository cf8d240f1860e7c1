"""Shape files (host only): the reference's BYU / OFF / VTK / OBJ / curve formats, read back
what the writers produce and parse hand-written files of each format."""
import numpy as np
import pytest

from registration_amd import curves, fileio, surfaces, synthetic


def test_byu_writer_layout_of_one_triangle(tmp_path):
    """The exact text the reference's savebyu format strings produce (surfaces.py:807-849):
    ' d' / ' f' fields, two vertices per line, last face index negated."""
    S = surfaces.Surface(FV=(np.array([[0, 1, 2]]), np.array([[0., 0, 0], [1, 0, 0], [0, 1.5, 0]])))
    p = tmp_path / "t.byu"
    S.savebyu(str(p))
    assert p.read_text() == (" 1  3  1  3 0\n1  1\n"
                             " 0.000000  0.000000  0.000000  1.000000  0.000000  0.000000\n"
                             " 0.000000  1.500000  0.000000 \n"
                             " 1  2 -3 ")


def test_byu_round_trip(tmp_path):
    V, F = synthetic.icosphere(2)
    p = str(tmp_path / "s.byu")
    surfaces.Surface(FV=(F, V)).savebyu(p)   # 480 faces: index lines wrap at 16
    S = surfaces.Surface(filename=p)
    assert np.array_equal(S.faces, F) and np.abs(S.vertices - V).max() < 5e-7
    np.testing.assert_allclose(S.surfel, np.cross(S.vertices[F[:, 1]] - S.vertices[F[:, 0]],
                                                  S.vertices[F[:, 2]] - S.vertices[F[:, 0]]))


def test_off_with_comments_and_errors(tmp_path):
    p = tmp_path / "a.off"
    p.write_text("# made by hand\nOFF\n# counts\n4 2 0\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n"
                 "3 0 1 2\n3 0 2 3\n")
    S = surfaces.Surface(filename=str(p))
    assert S.vertices.shape == (4, 3) and np.array_equal(S.faces, [[0, 1, 2], [0, 2, 3]])
    q = tmp_path / "b.off"
    q.write_text("OFF\n4 1 0\n0 0 0\n1 0 0\n0 1 0\n1 1 0\n4 0 1 3 2\n")
    with pytest.raises(ValueError, match="triangulated"):
        surfaces.Surface(filename=str(q))
    with pytest.raises(ValueError, match="extension"):
        surfaces.Surface(filename=str(tmp_path / "c.ply"))


def test_obj_reader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# obj\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvn 0 0 1\n"
                 "f 1/1/1 2/2/1 3/3/1\nf -4 -2 -1\n")
    S = surfaces.Surface(filename=str(p))
    assert np.array_equal(S.faces, [[0, 1, 2], [0, 2, 3]])
    assert S.surfVolume() == pytest.approx(0.0)


def test_vtk_surface_round_trip_with_fields(tmp_path):
    V, F = synthetic.icosphere(1)
    S = surfaces.Surface(FV=(F, V))
    p = str(tmp_path / "s.vtk")
    S.saveVTK(p, scalars=np.arange(len(V), dtype=float), normals=V, scal_name='idx')
    text = open(p).read()
    assert "POINT_DATA  42" in text and "SCALARS idx float 1" in text and "NORMALS normals float" in text
    R = surfaces.Surface(filename=p)
    assert np.array_equal(R.faces, F) and np.abs(R.vertices - V).max() < 5e-7


def test_curve_formats(tmp_path):
    X, _ = synthetic.circle(37)
    C = curves.Curve(pointSet=X)
    p = str(tmp_path / "c.dat")
    C.saveCurve(p)
    D = curves.Curve(filename=p)
    assert np.array_equal(D.faces, C.faces) and np.abs(D.vertices - X).max() < 5e-7
    t = tmp_path / "c.txt"
    t.write_text("2 3\n0 0\n1 0\n0 1\n")
    T = curves.Curve(filename=str(t))
    assert np.array_equal(T.faces, [[0, 1], [1, 2], [2, 0]]) and T.lenel.shape == (3, 2)
    v = str(tmp_path / "c.vtk")
    C.saveVTK(v, scalars=np.ones(37))
    Cv = curves.Curve(filename=v)
    assert np.array_equal(Cv.faces, C.faces) and np.abs(Cv.vertices - X).max() < 5e-7


def test_save_evolution_and_read_multiple(tmp_path):
    V, F = synthetic.icosphere(1)
    S = surfaces.Surface(FV=(F, V))
    xt = np.stack([V * (1 + 0.1 * k) for k in range(3)])
    base = str(tmp_path / "evo")
    surfaces.saveEvolution(base, S, xt)
    # the reference's '{0: 02d}' names: a sign space, then the index
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["evo 0.byu", "evo 1.byu", "evo 2.byu"]
    got = surfaces.readMultipleByu(str(tmp_path / "evo*.byu"))
    assert len(got) == 3 and all(np.array_equal(g.faces, F) for g in got)
    assert len(surfaces.readMultipleByu(str(tmp_path / "evo*.byu"), Nmax=2)) == 2
