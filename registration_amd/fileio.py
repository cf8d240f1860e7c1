"""Shape files, drop-in for the readers and writers of py-lddmm's surfaces and curves modules
(py-lddmm/surfaces.py:720-1005,1279-1284; curves.py:269-436,518-535).

Host-side text I/O only; the GPU path starts from the arrays these return.  Formats:
  .byu  BYU: header "ncomp npoints nfaces nedges 0", one component line, vertices two per line,
        face indices 1-based with each face's last index negated (written 16 per line);
  .off  OFF: "OFF", then "npoints nfaces nedges", vertices, then "3 i j k" faces (0-based);
        '#' comment lines are skipped;
  .vtk  legacy ASCII POLYDATA: POINTS, POLYGONS (surfaces) or LINES (curves), optional
        POINT_DATA fields on writing;
  .obj  Wavefront: "v x y z" and "f i j k" lines (1-based; "i/t/n" forms read the vertex);
  .dat  curve BYU (2-D vertices, edges with the second index negated);
  .txt  curve point list: "dim npoints", then one point per line; the curve is closed.
The reference reads .vtk and .obj through the vtk package; these readers parse the files
directly.  Where the reference prints a message and returns a partly filled shape (unknown
header, non-triangular OFF faces), these raise ValueError.
"""
from __future__ import annotations

import os

import numpy as np


def _skip(f, c='#'):
    """Next line of f that does not start with c (readskip, surfaces.py:1279-1284)."""
    ln = f.readline()
    while len(ln) > 0 and ln[0] == c:
        ln = f.readline()
    return ln


def _tokens(f):
    """Remaining whitespace-separated tokens of a text file."""
    return f.read().split()


# ---------------------------------------------------------------- surfaces
def read_byu(path):
    """(vertices (N,3), faces (F,3) 0-based) of a .byu file (surfaces.py:757-805)."""
    with open(path) as f:
        head = f.readline().split()
        ncomp, npoints, nfaces = int(head[0]), int(head[1]), int(head[2])
        for _ in range(ncomp):
            f.readline()
        tok = _tokens(f)
    nv = 3 * npoints
    vertices = np.array(tok[:nv], dtype=np.float64).reshape(npoints, 3)
    idx = np.abs(np.array(tok[nv:nv + 3 * nfaces], dtype=np.int64))
    if idx.size != 3 * nfaces:
        raise ValueError(f"{path}: expected {nfaces} triangles, found {idx.size // 3}")
    return vertices, idx.reshape(nfaces, 3) - 1


def save_byu(path, vertices, faces):
    """.byu writer in the reference's layout (surfaces.py:807-849)."""
    V, F = np.asarray(vertices), np.asarray(faces)
    npoints, nfaces = V.shape[0], F.shape[0]
    out = ['{0: d} {1: d} {2: d} {3: d} 0\n'.format(1, npoints, nfaces, 3 * nfaces),
           '1 {0: d}\n'.format(nfaces)]
    for k in range(0, npoints, 2):
        out.append('{0: f} {1: f} {2: f} '.format(V[k, 0], V[k, 1], V[k, 2]))
        if k + 1 < npoints:
            out.append('{0: f} {1: f} {2: f}\n'.format(V[k + 1, 0], V[k + 1, 1], V[k + 1, 2]))
        else:
            out.append('\n')
    j = 0
    for k in range(nfaces):
        for v in (F[k, 0] + 1, F[k, 1] + 1, -F[k, 2] - 1):
            out.append('{0: d} '.format(int(v)))
            j += 1
            if j == 16:
                out.append('\n')
                j = 0
    with open(path, 'w') as f:
        f.write(''.join(out))


def read_off(path):
    """(vertices, faces) of an .off file with triangular faces (surfaces.py:720-753)."""
    with open(path) as f:
        if _skip(f).split()[0].lower() != 'off':
            raise ValueError(f"{path}: not OFF format")
        head = _skip(f).split()
        npoints, nfaces = int(head[0]), int(head[1])
        vertices = np.empty((npoints, 3))
        for k in range(npoints):
            ln = _skip(f).split()
            vertices[k] = [float(ln[0]), float(ln[1]), float(ln[2])]
        faces = np.empty((nfaces, 3), dtype=np.int64)
        for k in range(nfaces):
            ln = _skip(f).split()
            if int(ln[0]) != 3:
                raise ValueError(f"{path}: reading only triangulated surfaces")
            faces[k] = [int(ln[1]), int(ln[2]), int(ln[3])]
    return vertices, faces


def read_obj(path):
    """(vertices, faces) of a Wavefront .obj file; polygons keep their first three vertices,
    as the reference's vtk-based reader does (surfaces.py:958-985)."""
    verts, faces = [], []
    with open(path) as f:
        for ln in f:
            t = ln.split()
            if not t:
                continue
            if t[0] == 'v':
                verts.append([float(t[1]), float(t[2]), float(t[3])])
            elif t[0] == 'f':
                faces.append([int(s.split('/')[0]) for s in t[1:4]])
    F = np.array(faces, dtype=np.int64)
    n = len(verts)
    F = np.where(F < 0, F + n, F - 1)    # negative indices count back from the last vertex
    return np.array(verts, dtype=np.float64).reshape(-1, 3), F.reshape(-1, 3)


def _read_vtk_polydata(path, cell_kind):
    """Points and the cells of one kind (POLYGONS or LINES) of a legacy ASCII VTK file."""
    with open(path) as f:
        tok = f.read().split()
    try:
        i = tok.index('POINTS')
    except ValueError:
        raise ValueError(f"{path}: no POINTS section") from None
    npoints = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * npoints], dtype=np.float64).reshape(npoints, 3)
    cells = []
    if cell_kind in tok:
        j = tok.index(cell_kind)
        ncells = int(tok[j + 1])
        p = j + 3
        for _ in range(ncells):
            m = int(tok[p])
            cells.append([int(v) for v in tok[p + 1:p + 1 + m]])
            p += 1 + m
    return pts, cells


def read_vtk_surface(path):
    """(vertices, faces) of a legacy VTK POLYDATA surface; polygons keep their first three
    point ids (surfaces.py:928-955)."""
    pts, cells = _read_vtk_polydata(path, 'POLYGONS')
    return pts, np.array([c[:3] for c in cells], dtype=np.int64).reshape(-1, 3)


class vtkFields:
    """Point-data fields for save_vtk_surface: lists alternating name, array
    (surfaces.py:15-20)."""

    def __init__(self):
        self.scalars = []
        self.vectors = []
        self.normals = []
        self.tensors = []


def save_vtk_surface(path, vertices, faces, fields: vtkFields | None = None):
    """Legacy ASCII POLYDATA writer in the reference's layout (surfaces.py:869-925)."""
    V, F = np.asarray(vertices), np.asarray(faces)
    out = ['# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET POLYDATA\n',
           '\nPOINTS {0: d} float'.format(V.shape[0])]
    out += ['\n{0: f} {1: f} {2: f}'.format(*V[k, :3]) for k in range(V.shape[0])]
    out.append('\nPOLYGONS {0:d} {1:d}'.format(F.shape[0], 4 * F.shape[0]))
    out += ['\n3 {0: d} {1: d} {2: d}'.format(*(int(v) for v in F[k])) for k in range(F.shape[0])]
    if fields is not None:
        wrote = False

        def header():
            nonlocal wrote
            if not wrote:
                out.append('\nPOINT_DATA {0: d}'.format(V.shape[0]))
                wrote = True
        for k in range(len(fields.scalars) // 2):
            header()
            out.append('\nSCALARS ' + fields.scalars[2 * k] + ' float 1\nLOOKUP_TABLE default')
            vals = np.asarray(fields.scalars[2 * k + 1]).reshape(-1)
            out += ['\n {0: .5f}'.format(vals[ll]) for ll in range(V.shape[0])]
        for name, lst in (('VECTORS', fields.vectors), ('NORMALS', fields.normals)):
            for k in range(len(lst) // 2):
                header()
                out.append('\n' + name + ' ' + lst[2 * k] + ' float')
                a = np.asarray(lst[2 * k + 1])
                out += ['\n {0: .5f} {1: .5f} {2: .5f}'.format(*a[ll, :3]) for ll in range(V.shape[0])]
        for k in range(len(fields.tensors) // 2):
            header()
            out.append('\nTENSORS ' + fields.tensors[2 * k] + ' float')
            t = np.asarray(fields.tensors[2 * k + 1])
            for ll in range(V.shape[0]):
                for kk in range(2):
                    out.append('\n {0: .5f} {1: .5f} {2: .5f}'.format(*t[ll, kk, :3]))
        out.append('\n')
    with open(path, 'w') as f:
        f.write(''.join(out))


def read_surface(path):
    """(vertices, faces) by extension: .byu, .off, .vtk, .obj (surfaces.py:28-46)."""
    ext = os.path.splitext(path)[1]
    readers = {'.byu': read_byu, '.off': read_off, '.vtk': read_vtk_surface, '.obj': read_obj}
    if ext not in readers:
        raise ValueError(f"unknown surface extension: {ext}")
    return readers[ext](path)


# ---------------------------------------------------------------- curves
def read_curve_dat(path):
    """(vertices (N,2), edges (E,2) 0-based) of a curve .dat file (curves.py:297-341)."""
    with open(path) as f:
        head = f.readline().split()
        ncomp, npoints, nfaces = int(head[0]), int(head[1]), int(head[2])
        for _ in range(ncomp):
            f.readline()
        tok = _tokens(f)
    nv = 2 * npoints
    vertices = np.array(tok[:nv], dtype=np.float64).reshape(npoints, 2)
    idx = np.abs(np.array(tok[nv:nv + 2 * nfaces], dtype=np.int64))
    if idx.size != 2 * nfaces:
        raise ValueError(f"{path}: expected {nfaces} edges, found {idx.size // 2}")
    return vertices, idx.reshape(nfaces, 2) - 1


def save_curve_dat(path, vertices, edges):
    """Curve .dat writer in the reference's layout (curves.py:344-385)."""
    V, F = np.asarray(vertices), np.asarray(edges)
    npoints, nfaces = V.shape[0], F.shape[0]
    out = ['{0: d} {1: d} {2: d} {3: d} 0\n'.format(1, npoints, nfaces, 2 * nfaces),
           '1 {0: d}\n'.format(nfaces)]
    for k in range(0, npoints, 2):
        out.append('{0: f} {1: f} '.format(V[k, 0], V[k, 1]))
        if k + 1 < npoints:
            out.append('{0: f} {1: f}\n'.format(V[k + 1, 0], V[k + 1, 1]))
        else:
            out.append('\n')
    j = 0
    for k in range(nfaces):
        for v in (F[k, 0] + 1, -F[k, 1] - 1):
            out.append('{0: d} '.format(int(v)))
            j += 1
            if j == 16:
                out.append('\n')
                j = 0
    with open(path, 'w') as f:
        f.write(''.join(out))


def read_curve_txt(path):
    """(vertices (N,dim), closed-loop edges) of a curve .txt file (curves.py:269-294)."""
    with open(path) as f:
        head = f.readline().split()
        dim, npoints = int(head[0]), int(head[1])
        vertices = np.zeros((npoints, dim))
        for k in range(npoints):
            ln = f.readline().split()
            vertices[k] = [float(v) for v in ln[:dim]]
    edges = np.empty((npoints, 2), dtype=np.int64)
    edges[:, 0] = np.arange(npoints)
    edges[:npoints - 1, 1] = np.arange(1, npoints)
    edges[npoints - 1, 1] = 0
    return vertices, edges


def read_vtk_curve(path):
    """(vertices (N,2), edges) of a legacy VTK POLYDATA curve: the first two coordinates of
    the points and the first two ids of each line (curves.py:414-433)."""
    pts, cells = _read_vtk_polydata(path, 'LINES')
    return pts[:, :2].copy(), np.array([c[:2] for c in cells], dtype=np.int64).reshape(-1, 2)


def save_vtk_curve(path, vertices, edges, scalars=None, normals=None, scal_name='scalars'):
    """Legacy ASCII POLYDATA curve writer in the reference's layout (curves.py:388-411); a
    third coordinate is written as the reference's 0 for planar curves and kept for 3-D ones."""
    V, F = np.asarray(vertices), np.asarray(edges)
    z = (lambda k: V[k, 2]) if V.shape[1] > 2 else (lambda k: 0)
    out = ['# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET POLYDATA\n',
           '\nPOINTS {0: d} float'.format(V.shape[0])]
    out += ['\n{0: f} {1: f} {2: f}'.format(V[k, 0], V[k, 1], z(k)) for k in range(V.shape[0])]
    out.append('\nLINES {0:d} {1:d}'.format(F.shape[0], 3 * F.shape[0]))
    out += ['\n2 {0: d} {1: d}'.format(int(F[k, 0]), int(F[k, 1])) for k in range(F.shape[0])]
    if scalars is not None or normals is not None:
        out.append('\nPOINT_DATA {0: d}'.format(V.shape[0]))
    if scalars is not None:
        out.append('\nSCALARS ' + scal_name + ' float 1\nLOOKUP_TABLE default')
        s = np.asarray(scalars).reshape(-1)
        out += ['\n {0: .5f}'.format(s[ll]) for ll in range(V.shape[0])]
    if normals is not None:
        out.append('\nNORMALS normals float')
        n = np.asarray(normals)
        out += ['\n {0: .5f} {1: .5f} {2: .5f}'.format(n[ll, 0], n[ll, 1], 0) for ll in range(V.shape[0])]
    out.append('\n')
    with open(path, 'w') as f:
        f.write(''.join(out))


def read_curve(path):
    """(vertices, edges) by extension: .dat, .txt, .vtk (curves.py:29-42)."""
    ext = os.path.splitext(path)[1]
    readers = {'.dat': read_curve_dat, '.txt': read_curve_txt, '.vtk': read_vtk_curve}
    if ext not in readers:
        raise ValueError(f"unknown curve extension: {ext}")
    return readers[ext](path)
