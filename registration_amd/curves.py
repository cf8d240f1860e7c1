"""Curves and the currents data term, drop-in for py-lddmm's curves module.

``Curve`` keeps the reference layout (vertices (N,d), faces = edges (E,2), centers = edge
midpoints, lenel = edge vectors; py-lddmm/curves.py:15-73) and the constructors
``Curve(curve=...)``, ``Curve(FV=(edges, vertices))``, ``Curve(pointSet=..., isOpen=...)``,
``Curve(filename=...)`` (.dat, .txt, .vtk, read by registration_amd.fileio), with the
reference's readers and writers (readCurve, readTxt, readVTK, saveCurve, saveVTK,
readMultipleCurves, saveEvolution; curves.py:269-436,518-535).
``currentNorm0`` / ``currentNormDef`` / ``currentNorm`` / ``currentNormGradient`` keep the
reference signatures (curves.py:539-613) and run on the GPU, and so do the measure-term
``measureNorm0`` / ``measureNormDef`` / ``measureNorm`` / ``measureNormGradient``
(curves.py:621-701).
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from . import fileio
from . import runtime as rt


class Curve:
    def __init__(self, curve=None, filename=None, FV=None, pointSet=None, isOpen=False):
        if curve is None:
            if FV is None:
                if filename is not None:
                    self.vertices, self.faces = fileio.read_curve(filename)
                    self.computeCentersLengths()
                elif pointSet is None:
                    self.vertices = np.empty(0)
                    self.centers = np.empty(0)
                    self.faces = np.empty(0)
                    self.lenel = np.empty(0)
                else:
                    n = pointSet.shape[0]
                    self.vertices = np.copy(pointSet).astype(np.float64)
                    self.faces = np.int_(np.zeros([n, 2]))
                    self.faces[:n - 1, 0] = np.arange(n - 1)
                    self.faces[:n - 1, 1] = np.arange(1, n)
                    if not isOpen:
                        self.faces[n - 1, :] = (n - 1, 0)
                    self.computeCentersLengths()
            else:
                self.vertices = np.copy(FV[1]).astype(np.float64)
                self.faces = np.int_(FV[0])
                self.computeCentersLengths()
        else:
            self.vertices = np.copy(curve.vertices)
            self.lenel = np.copy(curve.lenel)
            self.faces = np.copy(curve.faces)
            self.centers = np.copy(curve.centers)

    def computeCentersLengths(self):
        xDef1 = self.vertices[self.faces[:, 0], :]
        xDef2 = self.vertices[self.faces[:, 1], :]
        self.centers = (xDef1 + xDef2) / 2
        self.lenel = xDef2 - xDef1

    def updateVertices(self, x0):
        self.vertices = np.copy(x0)
        self.computeCentersLengths()

    # ---- files (curves.py:269-436) ------------------------------------------------------
    def _read(self, reader, path):
        self.vertices, self.faces = reader(path)
        self.computeCentersLengths()

    def readTxt(self, infile):
        self._read(fileio.read_curve_txt, infile)

    def readCurve(self, infile):
        self._read(fileio.read_curve_dat, infile)

    def readVTK(self, fileName):
        self._read(fileio.read_vtk_curve, fileName)

    def saveCurve(self, outfile):
        fileio.save_curve_dat(outfile, self.vertices, self.faces)

    def saveVTK(self, fileName, scalars=None, normals=None, scal_name='scalars'):
        fileio.save_vtk_curve(fileName, self.vertices, self.faces, scalars, normals, scal_name)


def readMultipleCurves(regexp, Nmax=0):
    """Curves of the files matching a glob pattern, at most Nmax (curves.py:518-527)."""
    import glob
    files = glob.glob(regexp)
    nm = min(Nmax, len(files)) if Nmax > 0 else len(files)
    return [Curve(filename=files[k]) for k in range(nm)]


def saveEvolution(fileName, fv0, xt):
    """One curve .dat file per time point, named as the reference names them
    (curves.py:530-534)."""
    fv = Curve(curve=fv0)
    xt = xt.cpu().numpy() if isinstance(xt, torch.Tensor) else np.asarray(xt)
    for k in range(xt.shape[0]):
        fv.vertices = np.squeeze(xt[k, :, :])
        fv.saveCurve(fileName + '{0: 02d}'.format(k) + '.byu')


class DeviceCurve:
    """A curve resident on the GPU (vertices, edges, centers, edge vectors, incidence CSR;
    with ``with_weights``, the edge weights sqrt(|lenel|^2 + 1e-10), else ``weights`` is
    None)."""

    def __init__(self, faces_np, vertices, with_weights: bool = False):
        faces_np = np.asarray(faces_np)
        self.nv, self.d = int(vertices.shape[0]), int(vertices.shape[1])
        self.faces, self.ptr, self.inc = engine.topology(faces_np, self.nv)
        E = faces_np.shape[0]
        dev = rt.device()
        self.centers = torch.empty((E, self.d), dtype=torch.float64, device=dev)
        self.lenel = torch.empty((E, self.d), dtype=torch.float64, device=dev)
        self.weights = torch.empty(E, dtype=torch.float64, device=dev) if with_weights else None
        self.vertices = None
        self.update(vertices)

    def update(self, vertices):
        if self.weights is not None:
            engine.geometry_batch([self], [rt.as_dev(vertices)])
            return
        self.vertices = rt.as_dev(vertices)
        engine.curve_geometry(self.vertices, self.faces, self.centers, self.lenel)

    # uniform names with DeviceSurface
    @property
    def surfel(self):
        return self.lenel


def _kernel(K):
    return engine.DevKernel.of(K)


def currentNorm0(fv1, KparDist):
    """curves.py:539-543."""
    s = DeviceCurve(fv1.faces, fv1.vertices)
    tot = torch.empty(1, dtype=torch.float64, device=rt.device())
    engine.knorm([engine.KNormJob(s.centers, s.lenel, _kernel(KparDist).energy(), tot)])
    return float(tot.item())


def currentNormDef(fvDef, fv1, KparDist):
    """curves.py:549-559."""
    a = DeviceCurve(fvDef.faces, fvDef.vertices)
    b = DeviceCurve(fv1.faces, fv1.vertices)
    tot = torch.empty(1, dtype=torch.float64, device=rt.device())
    engine.currents([engine.CurrentsJob(a.centers, a.lenel, b.centers, b.lenel,
                                        _kernel(KparDist).energy(), tot, coef=0.5)])
    return float(tot.item())


def currentNorm(fvDef, fv1, KparDist):
    return currentNormDef(fvDef, fv1, KparDist) + currentNorm0(fv1, KparDist)


def current_gradient_dev(s: DeviceCurve, c2, cr2, K: engine.DevKernel, scale: float,
                         px: torch.Tensor, total: torch.Tensor | None = None) -> torch.Tensor:
    E, d = s.centers.shape
    dev = s.centers.device
    z1 = torch.empty((E, d), dtype=torch.float64, device=dev)
    dz = torch.empty((E, d), dtype=torch.float64, device=dev)
    if total is None:
        total = torch.empty(1, dtype=torch.float64, device=dev)
    engine.currents([engine.CurrentsJob(s.centers, s.lenel, c2, cr2, K, total, z1=z1, dz=dz,
                                        coef=0.5)])
    engine.curve_vertex_gradient(z1, dz, s.ptr, s.inc, s.nv, scale, px)
    return px


def currentNormGradient(fvDef, fv1, KparDist):
    """curves.py:566-613."""
    a = DeviceCurve(fvDef.faces, fvDef.vertices)
    b = DeviceCurve(fv1.faces, fv1.vertices)
    px = torch.empty((a.nv, a.d), dtype=torch.float64, device=rt.device())
    current_gradient_dev(a, b.centers, b.lenel, _kernel(KparDist), 2.0, px)
    return px.cpu().numpy()


# ---- measure data term (curves.py:621-701) ------------------------------------------------
def measureNorm0(fv1, KparDist):
    """<a, K(c, c) a> with a = |lenel| (curves.py:621-628: no 1e-10 here, unlike
    measureNormDef and measureNormGradient)."""
    tot = torch.empty(1, dtype=torch.float64, device=rt.device())
    engine.measure([engine.MeasureJob(rt.as_dev(fv1.centers),
                                      rt.as_dev(engine.measure_weights(fv1.lenel, 0.0)),
                                      None, None, _kernel(KparDist).energy(), tot)])
    return float(tot.item())


def measureNormDef(fvDef, fv1, KparDist):
    """curves.py:632-645."""
    tot = torch.empty(1, dtype=torch.float64, device=rt.device())
    engine.measure([engine.MeasureJob(rt.as_dev(fvDef.centers),
                                      rt.as_dev(engine.measure_weights(fvDef.lenel)),
                                      rt.as_dev(fv1.centers), rt.as_dev(engine.measure_weights(fv1.lenel)),
                                      _kernel(KparDist).energy(), tot, coef=0.5)])
    return float(tot.item())


def measureNorm(fvDef, fv1, KparDist):
    return measureNormDef(fvDef, fv1, KparDist) + measureNorm0(fv1, KparDist)


def measure_gradient_dev(s: DeviceCurve, c2, a2, K: engine.DevKernel, scale: float,
                         px: torch.Tensor, total: torch.Tensor | None = None) -> torch.Tensor:
    E, d = s.centers.shape
    dev = s.centers.device
    z1 = torch.empty((E, d), dtype=torch.float64, device=dev)
    dz = torch.empty((E, d), dtype=torch.float64, device=dev)
    if total is None:
        total = torch.empty(1, dtype=torch.float64, device=dev)
    engine.measure([engine.MeasureJob(s.centers, s.weights, c2, a2, K, total, n1=s.lenel,
                                      z1=z1, dz=dz, coef=0.5)])
    engine.curve_vertex_gradient(z1, dz, s.ptr, s.inc, s.nv, scale, px)
    return px


def measureNormGradient(fvDef, fv1, KparDist):
    """curves.py:653-701."""
    a = DeviceCurve(fvDef.faces, fvDef.vertices, with_weights=True)
    px = torch.empty((a.nv, a.d), dtype=torch.float64, device=rt.device())
    measure_gradient_dev(a, rt.as_dev(fv1.centers), rt.as_dev(engine.measure_weights(fv1.lenel)),
                         _kernel(KparDist), 2.0, px)
    return px.cpu().numpy()
