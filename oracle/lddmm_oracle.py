"""FP64 CPU oracle of py-lddmm's point-set LDDMM hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker the GPU product path
(registration_amd/) is compared against.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it; the product package never does.

It restates, in Python 3 / NumPy float64, the executed reference path:
  Kernel                  py-lddmm/kernelFunctions.py:238-454 (sums in oracle/kernel_sums.c,
                          restating kernelFunctions.f90:1-114,430-496)
  landmark flow/adjoint   py-lddmm/pointEvolution.py:599-817 (second order: :882-1084)
  surfaces / curves       py-lddmm/surfaces.py:61-75,711-717,1009-1276; curves.py:58-73,539-701
                          (currents, measure and varifold data terms)
  SurfaceMatching         py-lddmm/surfaceMatching.py:19-484 (SymmetricSurfaceMatching: symmetric=True)
  CurveMatching           py-lddmm/curveMatching.py:19-352 (affine='none')
  SecondOrderSurfaceMatching  py-lddmm/secondOrderMatching.py:43-509 (time series of targets)
  cg                      py-lddmm/conjugateGradient.py:33-217
Python-2 idioms are restated with their Python-2 meaning (``x == None`` on arrays -> ``is None``).
File output (VTK) and the saveRate branches are left out: they do not feed back into the
optimisation.  cg keeps the reference's log lines (its call traces are pinned against the
reference's own conjugateGradient.cg by tests/golden/cg_trace.npz).

With ``sums='reference'`` the kernel sums call the reference's own Fortran subroutines
compiled by oracle/Makefile into oracle/_ref/libkff_ref.so (present only where that build ran).
"""
from __future__ import annotations

import ctypes
import logging
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_sums.so")
_REF = os.path.join(_HERE, "_ref", "libkff_ref.so")

_P = ctypes.c_void_p
_lib = None
_ref = None


def build():
    """Compile oracle/kernel_sums.c (gcc) into oracle/liboracle_sums.so."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        i64, i, d = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        _lib.oracle_apply_k.argtypes = [_P, _P, _P, i64, i64, i, i, i, d, _P]
        _lib.oracle_apply_diff_kt.argtypes = [_P, _P, _P, _P, i64, i64, i, i, i, i, d, _P]
        _lib.oracle_apply_ddiff_k11and12.argtypes = [_P, _P, _P, _P, i64, i, i, d, _P]
        _lib.oracle_apply_k_mat.argtypes = [_P, _P, _P, i64, i64, i, i, i, d, _P]
        _lib.oracle_apply_diff_k.argtypes = [_P, _P, _P, i64, i, i, i, d, _P]
        _lib.oracle_apply_diff_k_mat.argtypes = [_P, _P, _P, i64, i64, i, i, d, _P]
        for f in (_lib.oracle_apply_k, _lib.oracle_apply_diff_kt, _lib.oracle_apply_ddiff_k11and12,
                  _lib.oracle_apply_k_mat, _lib.oracle_apply_diff_k_mat, _lib.oracle_apply_diff_k):
            f.restype = ctypes.c_int
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref():
    """The reference's Fortran kernel sums (column-major, all arguments by reference)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libkff_ref.so is not built (make -C oracle ref)")
        _ref = ctypes.CDLL(_REF)
    return _ref


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_P)


# --------------------------------------------------------------------------------------
# Kernel sums
# --------------------------------------------------------------------------------------
def apply_k(y, x, w, sigma, order, sums="oracle"):
    """f = sum_l K(|y_k - x_l|) w_l (kernelFunctions.f90:1-54)."""
    y, x, w = _c(y), _c(x), _c(w)
    if w.ndim == 1:
        w = w[:, None]
    ny, d = y.shape
    nx, m = w.shape
    if sums == "reference":
        f = np.zeros((ny, m), order="F")
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        ref().applyk_(_p(np.asfortranarray(y)), _p(np.asfortranarray(x)), _p(np.asfortranarray(w)),
                      ctypes.byref(ctypes.c_double(sigma)), ci(order), ci(ny), ci(nx), ci(d), ci(m),
                      _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((ny, m))
    rc = lib().oracle_apply_k(_p(y), _p(x), _p(w), ny, nx, d, m, order, sigma, _p(f))
    if rc:
        raise ValueError("oracle_apply_k: invalid arguments")
    return f


def apply_diff_kt(y, x, a1, a2, sigma, order, sums="oracle"):
    """f = sum_l 2K'(r^2)(y_k - x_l) sum_{j,c} a1[j,k,c] a2[j,l,c] (kernelFunctions.f90:56-114)."""
    y, x, a1, a2 = _c(y), _c(x), _c(a1), _c(a2)
    ny, d = y.shape
    na, _, m = a1.shape
    nx = x.shape[0]
    if sums == "reference":
        f = np.zeros((ny, d), order="F")
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        ref().applykdifft_(_p(np.asfortranarray(y)), _p(np.asfortranarray(x)),
                           _p(np.asfortranarray(a1)), _p(np.asfortranarray(a2)),
                           ctypes.byref(ctypes.c_double(sigma)), ci(order), ci(ny), ci(nx), ci(d),
                           ci(m), ci(na), _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((ny, d))
    rc = lib().oracle_apply_diff_kt(_p(y), _p(x), _p(a1), _p(a2), ny, nx, d, m, na, order, sigma,
                                    _p(f))
    if rc:
        raise ValueError("oracle_apply_diff_kt: invalid arguments")
    return f


def apply_ddiff_k11and12(x, a1, a2, p, sigma, order, sums="oracle"):
    """kernelFunctions.f90:430-496."""
    x, a1, a2, p = _c(x), _c(a1), _c(a2), _c(p)
    n, d = x.shape
    if sums == "reference":
        f = np.zeros((n, d), order="F")
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        ref().applykdiff11and12_(_p(np.asfortranarray(x)), _p(np.asfortranarray(a1)),
                                 _p(np.asfortranarray(a2)), _p(np.asfortranarray(p)),
                                 ctypes.byref(ctypes.c_double(sigma)), ci(order), ci(n), ci(d), _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((n, d))
    rc = lib().oracle_apply_ddiff_k11and12(_p(x), _p(a1), _p(a2), _p(p), n, d, order, sigma, _p(f))
    if rc:
        raise ValueError("oracle_apply_ddiff_k11and12: invalid arguments")
    return f


_DIFF_K_REF = {1: "applykdiff1_", 2: "applykdiff2_", 3: "applykdiff1and2_"}


def apply_diff_k(x, a1, a2, sigma, order, form, sums="oracle"):
    """applyDiffK (form 1), applyDiffK2 (2), applyDiffK1and2 (3): sum_l 2K'(r^2)
    <x_k - x_l, b_kl> a2_l (kernelFunctions.f90:116-290)."""
    x, a1, a2 = _c(x), _c(a1), _c(a2)
    n, d = x.shape
    if sums == "reference":
        f = np.zeros((n, d), order="F")   # f2py zero-fills intent(out)
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        getattr(ref(), _DIFF_K_REF[form])(_p(np.asfortranarray(x)), _p(np.asfortranarray(a1)),
                                          _p(np.asfortranarray(a2)),
                                          ctypes.byref(ctypes.c_double(sigma)), ci(order), ci(n),
                                          ci(d), _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((n, d))
    rc = lib().oracle_apply_diff_k(_p(x), _p(a1), _p(a2), n, d, form, order, sigma, _p(f))
    if rc:
        raise ValueError("oracle_apply_diff_k: invalid arguments")
    return f


def apply_k_mat(y, x, beta, sigma, order, sums="oracle"):
    """f = sum_l K(|y_k - x_l|) beta[k, l, :] (kernelFunctions.f90:499-550)."""
    y, x, beta = _c(y), _c(x), _c(beta)
    ny, d = y.shape
    nx = x.shape[0]
    m = beta.shape[2]
    if sums == "reference":
        f = np.zeros((ny, m), order="F")
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        ref().applykmat_(_p(np.asfortranarray(y)), _p(np.asfortranarray(x)),
                         _p(np.asfortranarray(beta)), ctypes.byref(ctypes.c_double(sigma)),
                         ci(order), ci(ny), ci(nx), ci(d), ci(m), _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((ny, m))
    rc = lib().oracle_apply_k_mat(_p(y), _p(x), _p(beta), ny, nx, d, m, order, sigma, _p(f))
    if rc:
        raise ValueError("oracle_apply_k_mat: invalid arguments")
    return f


def apply_diff_k_mat(y, x, beta, sigma, order, sums="oracle"):
    """f = sum_l 2K'(r^2)(y_k - x_l) beta[k, l] (kernelFunctions.f90:552-610)."""
    y, x, beta = _c(y), _c(x), _c(beta)
    ny, d = y.shape
    nx = x.shape[0]
    if sums == "reference":
        f = np.zeros((ny, d), order="F")
        ci = lambda v: ctypes.byref(ctypes.c_int(v))
        ref().applykdiffmat_(_p(np.asfortranarray(y)), _p(np.asfortranarray(x)),
                             _p(np.asfortranarray(beta)), ctypes.byref(ctypes.c_double(sigma)),
                             ci(order), ci(ny), ci(nx), ci(d), _p(f))
        return np.ascontiguousarray(f)
    f = np.empty((ny, d))
    rc = lib().oracle_apply_diff_k_mat(_p(y), _p(x), _p(beta), ny, nx, d, order, sigma, _p(f))
    if rc:
        raise ValueError("oracle_apply_diff_k_mat: invalid arguments")
    return f


def _lap_pol_diff(u, order):
    """lapPolDiff (kernelFunctions.py:75-88): exact double expressions."""
    if order == 1:
        return np.ones_like(u)
    if order == 2:
        return (1 + u) / 3.
    if order == 3:
        return (3 + 3 * u + u * u) / 15.
    if order == 4:
        return (15 + 15 * u + 6 * u * u + u * u * u) / 105.
    raise ValueError("the Laplacian derivative needs order 1..4")


def _lap_pol(u, order):
    """lapPol (kernelFunctions.py:58-72)."""
    if order == 0:
        return np.ones_like(u)
    if order == 1:
        return 1. + u
    if order == 2:
        return (3. + 3 * u + u * u) / 3.
    if order == 3:
        return (15. + 15 * u + 6 * u * u + u * u * u) / 15.
    return (105. + 105 * u + 45 * u * u + 10 * u * u * u + u * u * u * u) / 105.


def kernel_matrix(y, x, sigma, order, self_set):
    """The dense K matrix of precompute() (kernelFunctions.py:10-55,152-183)."""
    from scipy.spatial import distance as dfun
    if order > 4:
        if self_set:
            K = dfun.squareform(np.exp(-dfun.pdist(x, 'sqeuclidean') / (2 * sigma * sigma)),
                                checks=False)
            np.fill_diagonal(K, 1)
            return K
        return np.exp(-dfun.cdist(y, x, 'sqeuclidean') / (2 * sigma * sigma))
    if self_set:
        u = dfun.pdist(x) / sigma
        K = dfun.squareform(_lap_pol(u, order) * np.exp(-u))
        np.fill_diagonal(K, 1)
        return K
    u = dfun.cdist(y, x) / sigma
    return _lap_pol(u, order) * np.exp(-u)


def kernel_matrix_diff(y, x, sigma, order, self_set):
    """The dense K'(r^2) matrix of precompute(diff=True) (kernelFunctions.py:10-55,152-183):
    Gaussian -exp(-r^2/2s^2)/(2s^2); Laplacian -lapPolDiff(u) e^-u/(2s^2) with u = r/s.  With
    ``self_set`` (firstVar None: pdist + squareform) the diagonal is filled as the reference
    does (Gaussian K = 1 before scaling; Laplacian -1/((2 order - 1) 2 s^2))."""
    from scipy.spatial import distance as dfun
    if order > 4:
        if self_set:
            K = dfun.squareform(np.exp(-dfun.pdist(x, 'sqeuclidean') / (2 * sigma * sigma)),
                                checks=False)
            np.fill_diagonal(K, 1)
        else:
            K = np.exp(-dfun.cdist(y, x, 'sqeuclidean') / (2 * sigma * sigma))
        return -K / (2 * sigma * sigma)
    if self_set:
        u = dfun.pdist(x) / sigma
        K = dfun.squareform(-_lap_pol_diff(u, order) * np.exp(-u) / (2 * sigma * sigma))
        np.fill_diagonal(K, -1. / ((2 * order - 1) * 2 * sigma * sigma))
        return K
    u = dfun.cdist(y, x) / sigma
    return -_lap_pol_diff(u, order) * np.exp(-u) / (2 * sigma * sigma)


def _lap_pol_diff2(u, order):
    """lapPolDiff2 (kernelFunctions.py:92-104)."""
    if order == 2:
        return np.full_like(u, 1.0 / 3.)
    if order == 3:
        return (1 + u) / 15.
    if order == 4:
        return (3 + 3 * u + u * u) / 105.
    raise ValueError("the Laplacian second derivative needs order 2..4")


def kernel_matrix_diff2(x, sigma, order):
    """The dense K''(r^2) matrix of precompute(diff2=True) on a self set: Gaussian
    K/(4 s^4) (diagonal 1/(4 s^4)); Laplacian lapPolDiff2(u) e^-u/(4 s^4) with the
    reference's diagonal fill 1/(35 * 4 s^4) (kernelFunctions.py:10-55,152-183)."""
    from scipy.spatial import distance as dfun
    s4 = 4 * sigma ** 4
    if order > 4:
        K = dfun.squareform(np.exp(-dfun.pdist(x, 'sqeuclidean') / (2 * sigma * sigma)),
                            checks=False)
        np.fill_diagonal(K, 1)
        return K / s4
    u = dfun.pdist(x) / sigma
    K = dfun.squareform(_lap_pol_diff2(u, order) * np.exp(-u) / s4)
    np.fill_diagonal(K, 1. / (35 * s4))
    return K


def apply_ddiff_k11(x, n, a, p, sigma, order):
    """Kernel.applyDDiffK11 (kernelFunctions.py:459-478), dense."""
    r1 = kernel_matrix_diff(None, x, sigma, order, True)
    r2 = kernel_matrix_diff2(x, sigma, order)
    xxp = -np.dot(p, x.T) + (x * p).sum(axis=1)[:, np.newaxis]
    na = np.dot(n, a.T)
    u = r2 * xxp * na
    zpx = 4 * ((u.sum(axis=1))[:, np.newaxis] * x - np.dot(u, x))
    zpx += 2 * (r1 * na).sum(axis=1)[:, np.newaxis] * p
    return zpx


def apply_ddiff_k12(x, n, a, p, sigma, order):
    """Kernel.applyDDiffK12 (kernelFunctions.py:481-506), dense, affine='none'."""
    r1 = kernel_matrix_diff(None, x, sigma, order, True)
    r2 = kernel_matrix_diff2(x, sigma, order)
    xxp = np.dot(x, p.T) - (x * p).sum(axis=1)[np.newaxis, :]
    na = np.dot(n, a.T)
    u = r2 * xxp * na
    zpx = -4 * ((u.sum(axis=1))[:, np.newaxis] * x - np.dot(u, x))
    zpx -= 2 * np.dot(r1 * na, p)
    return zpx


def apply_divergence(x, a, sigma, order, firstVar=None):
    """Kernel.applyDivergence (kernelFunctions.py:510-530), dense as in the reference:
    2 (<(r a)_k, y_k> - (r <a, x>)_k) with r the K' matrix."""
    r = kernel_matrix_diff(firstVar, x, sigma, order, firstVar is None)
    y = x if firstVar is None else firstVar
    return 2 * (np.multiply(np.dot(r, a), y).sum(axis=1) - np.dot(r, np.multiply(a, x).sum(axis=1)))


class Kernel:
    """kernelFunctions.Kernel restated (kernelFunctions.py:238-454), affine='none' only."""

    def __init__(self, name='gauss', affine='none', sigma=6.5, order=10, w1=1.0, w2=1.0, dim=3,
                 center=(0, 0, 0), weight=1.0, sums="oracle"):
        self.name, self.sigma, self.order, self.affine = name, sigma, order, affine
        self.weight = weight    # scales the dense matrices only (kernelFunctions.py:298)
        self.sums = sums
        if name == 'laplacian':
            if self.order > 4:
                self.order = 3
        elif name == 'gauss':
            self.order = 10
        else:
            self.name = 'none'
        if affine != 'none':
            raise NotImplementedError("oracle restates affine='none' only")

    def applyK(self, x, a, firstVar=None, matrixWeights=False):
        y = x if firstVar is None else firstVar
        if matrixWeights:   # a: (n_y, n_x, m) per-pair weights (kernelFunctions.py:317-325)
            return apply_k_mat(y, x, a, self.sigma, self.order, self.sums)
        if self.name == 'none':
            return np.zeros((x.shape[0], a.shape[1]))
        return apply_k(y, x, a, self.sigma, self.order, self.sums)

    def applyDiffKmat(self, x, beta, firstVar=None):
        """kernelFunctions.py:405-411: sum_l 2K'(r^2)(y_k - x_l) beta[k, l]."""
        y = x if firstVar is None else firstVar
        return apply_diff_k_mat(y, x, beta, self.sigma, self.order, self.sums)

    def applyDiffKT(self, x, a1, a2, firstVar=None):
        y = x if firstVar is None else firstVar
        if self.name == 'none':
            return np.zeros(y.shape)
        return apply_diff_kt(y, x, np.asarray(a1), np.asarray(a2), self.sigma, self.order, self.sums)

    def applyDDiffK11and12(self, x, n, a, p):
        return apply_ddiff_k11and12(x, n, a, p, self.sigma, self.order, self.sums)

    def applyDiffK(self, x, a1, a2):
        return apply_diff_k(x, a1, a2, self.sigma, self.order, 1, self.sums)

    def applyDiffK2(self, x, a1, a2):
        return apply_diff_k(x, a1, a2, self.sigma, self.order, 2, self.sums)

    def applyDiffK1and2(self, x, a1, a2):
        return apply_diff_k(x, a1, a2, self.sigma, self.order, 3, self.sums)

    def applyDDiffK11(self, x, n, a, p):
        return self.weight * apply_ddiff_k11(x, n, a, p, self.sigma, self.order)

    def applyDDiffK12(self, x, n, a, p):
        return self.weight * apply_ddiff_k12(x, n, a, p, self.sigma, self.order)

    def applyDivergence(self, x, a, firstVar=None):
        """kernelFunctions.py:510-530 with the dense K' matrix times ``weight`` (:298)."""
        return self.weight * apply_divergence(x, a, self.sigma, self.order, firstVar)


# --------------------------------------------------------------------------------------
# Landmark flow and adjoint (pointEvolution.py:599-817), affine=None or A=0 arrays
# --------------------------------------------------------------------------------------
def landmarkEPDiff(T, x0, a0, KparDiff):
    """pointEvolution.py:599-658 (no affine, no extras)."""
    N, dim = x0.shape
    timeStep = 1.0 / T
    at = np.zeros([T, N, dim])
    xt = np.zeros([T + 1, N, dim])
    xt[0] = x0
    at[0] = a0
    for k in range(T):
        z = xt[k]
        a = at[k]
        xt[k + 1] = z + timeStep * KparDiff.applyK(z, a)
        if k < T - 1:
            at[k + 1] = a - timeStep * KparDiff.applyDiffKT(z, a[np.newaxis], a[np.newaxis])
    return xt, at


def landmarkDirectEvolutionEuler(x0, at, KparDiff, affine=None, withPointSet=None,
                                 withNormals=None, withJacobian=False):
    """pointEvolution.py:666-730."""
    N, dim = x0.shape
    M = at.shape[0] + 1
    timeStep = 1.0 / (M - 1)
    xt = np.zeros([M, N, dim])
    xt[0] = x0
    yt = nt = None
    if withPointSet is not None:
        yt = np.zeros([M, withPointSet.shape[0], dim])
        yt[0] = withPointSet
    if withNormals is not None:
        nt = np.zeros([M, N, dim])
        nt[0] = withNormals
    Jt = np.zeros([M, N if yt is None else yt.shape[1]]) if withJacobian else None
    for k in range(M - 1):
        z = xt[k]
        a = at[k]
        xt[k + 1] = z + timeStep * KparDiff.applyK(z, a)
        if affine is not None:
            A, b = affine
            xt[k + 1] += timeStep * (np.dot(z, A[k].T) + b[k])
        if yt is not None:
            zy = yt[k]
            yt[k + 1] = zy + timeStep * KparDiff.applyK(z, a, firstVar=zy)
            if affine is not None:
                yt[k + 1] += timeStep * (np.dot(zy, affine[0][k].T) + affine[1][k])
            if withJacobian:
                Jt[k + 1] = Jt[k] + timeStep * KparDiff.applyDivergence(z, a, firstVar=zy)
                if affine is not None:
                    Jt[k + 1] += timeStep * np.trace(affine[0][k])
        elif withJacobian:
            Jt[k + 1] = Jt[k] + timeStep * KparDiff.applyDivergence(z, a)
            if affine is not None:
                Jt[k + 1] += timeStep * np.trace(affine[0][k])
        if nt is not None:
            zn = nt[k]
            nt[k + 1] = zn - timeStep * KparDiff.applyDiffKT(z, zn[np.newaxis], a[np.newaxis])
            if affine is not None:
                nt[k + 1] += timeStep * np.dot(zn, affine[0][k])
    if yt is None and nt is None and Jt is None:
        return xt
    out = [xt]
    if yt is not None:
        out.append(yt)
    if nt is not None:
        out.append(nt)
    if Jt is not None:
        out.append(Jt)
    return out


def landmarkHamiltonianCovector(x0, at, px1, KparDiff, regweight, affine=None):
    """pointEvolution.py:735-765."""
    N, dim = x0.shape
    M = at.shape[0]
    timeStep = 1.0 / M
    xt = landmarkDirectEvolutionEuler(x0, at, KparDiff, affine=affine)
    pxt = np.zeros([M + 1, N, dim])
    pxt[M] = px1
    for t in range(M):
        px = pxt[M - t]
        z = xt[M - t - 1]
        a = at[M - t - 1]
        a1 = np.concatenate((px[np.newaxis], a[np.newaxis], -2 * regweight * a[np.newaxis]))
        a2 = np.concatenate((a[np.newaxis], px[np.newaxis], a[np.newaxis]))
        zpx = KparDiff.applyDiffKT(z, a1, a2)
        pxt[M - t - 1] = px + timeStep * zpx
        if affine is not None:
            pxt[M - t - 1] += timeStep * np.dot(px, affine[0][M - t - 1])
    return pxt, xt


def landmarkHamiltonianGradient(x0, at, px1, KparDiff, regweight, getCovector=False, affine=None):
    """pointEvolution.py:793-817."""
    pxt, xt = landmarkHamiltonianCovector(x0, at, px1, KparDiff, regweight, affine=affine)
    dat = np.zeros(at.shape)
    if affine is not None:
        dA = np.zeros(affine[0].shape)
        db = np.zeros(affine[1].shape)
    for k in range(at.shape[0]):
        dat[k] = 2 * regweight * at[k] - pxt[k + 1]
        if affine is not None:
            dA[k] = np.dot(pxt[k + 1].T, xt[k])
            db[k] = pxt[k + 1].sum(axis=0)
    if affine is None:
        return (dat, xt, pxt) if getCovector else (dat, xt)
    return (dat, dA, db, xt, pxt) if getCovector else (dat, dA, db, xt)


class AffineBasis:
    """affineBasis.AffineBasis (affineBasis.py:3-79); the entry the reference indexes with the
    float (i+1.) (:41) is indexed with the integer i+1."""

    def __init__(self, dim=3, affine='affine'):
        u = 1.0 / np.sqrt(2.0)
        dimSym = (dim * (dim - 1)) // 2
        affCode = {'affine': 1, 'similitude': 2, 'euclidean': 3, 'translation': 4}.get(affine, 5)
        self.dim = dim
        self.rotComp, self.simComp, self.affComp, self.transComp = [], [], [], []
        if affCode == 1:
            self.affineDim = 2 * (dimSym + dim)
        elif affCode == 2:
            self.affineDim = dimSym + 1 + dim
        elif affCode == 3:
            self.affineDim = dimSym + dim
        elif affCode == 4:
            self.affineDim = dim
        else:
            self.basis = []
            self.affineDim = 0
            return
        self.basis = np.zeros([2 * (dimSym + dim), self.affineDim])
        k = 0
        if affCode <= 1:
            for i in range(dimSym):
                for j in range(i + 1, dim):
                    self.basis[i * dim + j, k] = u
                    self.basis[j * dim + i, k] = u
                    k += 1
            for i in range(dim - 1):
                uu = np.sqrt((1 - 1.0 / (i + 2))) / (i + 1.)
                self.basis[(i + 1) * dim + (i + 1), k] = np.sqrt(1 - 1.0 / (i + 2))
                for j in range(i + 1):
                    self.basis[j * dim + j, k] = -uu
                k += 1
            self.affComp = list(range(k))
        if affCode <= 2:
            k0 = k
            for i in range(dim):
                self.basis[i * dim + i, k] = 1. / np.sqrt(dim)
            k += 1
            self.simComp = list(range(k0, k))
        if affCode <= 3:
            k0 = k
            for i in range(dim):
                for j in range(i + 1, dim):
                    self.basis[i * dim + j, k] = u
                    self.basis[j * dim + i, k] = -u
                    k += 1
            self.rotComp = list(range(k0, k))
        k0 = k
        for i in range(dim):
            self.basis[i + dim ** 2, k] = 1
            k += 1
        self.transComp = list(range(k0, k))

    def getTransforms(self, Afft):
        Tsize = Afft.shape[0]
        dim2 = self.dim ** 2
        A = [np.zeros([Tsize, self.dim, self.dim]), np.zeros([Tsize, self.dim])]
        if self.affineDim > 0:
            for t in range(Tsize):
                AB = np.dot(self.basis, Afft[t])
                A[0][t] = AB[0:dim2].reshape([self.dim, self.dim])
                A[1][t] = AB[dim2:dim2 + self.dim]
        return A


# --------------------------------------------------------------------------------------
# Second-order (spline) flow and its adjoint (pointEvolution.py:882-1066)
# --------------------------------------------------------------------------------------
def secondOrderEvolution(x0, a0, rhot, KparDiff, withJacobian=False, withPointSet=None,
                         affine=None):
    """pointEvolution.py:882-948: x' = K(x)a, a' = -applyDiffKT(x,[a],[a]) + rho, with the
    affine step U = I + dt A_k (x <- (x + dt zx) U^T + dt b_k, a <- (a + dt za) U^-1).  The
    Jacobian is taken at the point set when one is given, else at the points."""
    T = rhot.shape[0]
    N, dim = x0.shape
    dt = 1.0 / T
    at = np.zeros([T + 1, N, dim])
    xt = np.zeros([T + 1, N, dim])
    xt[0], at[0] = x0, a0
    zt = Jt = None
    if withPointSet is not None:
        zt = np.zeros([T + 1, withPointSet.shape[0], dim])
        zt[0] = withPointSet
    if withJacobian:
        Jt = np.zeros([T + 1, N if zt is None else zt.shape[1]])
    for k in range(T):
        x, a, rho = xt[k], at[k], rhot[k]
        zx = KparDiff.applyK(x, a)
        za = -KparDiff.applyDiffKT(x, a[np.newaxis], a[np.newaxis]) + rho
        if affine is not None:
            U = np.eye(dim) + dt * affine[0][k]
            xt[k + 1] = np.dot(x + dt * zx, U.T) + dt * affine[1][k]
            at[k + 1] = np.dot(a + dt * za, np.linalg.inv(U))
        else:
            xt[k + 1] = x + dt * zx
            at[k + 1] = a + dt * za
        if zt is not None:
            z = zt[k]
            zz = KparDiff.applyK(x, a, firstVar=z)
            zt[k + 1] = (np.dot(z + dt * zz, U.T) + dt * affine[1][k]) if affine is not None \
                else z + dt * zz
        if Jt is not None:
            at_pts = zt[k] if zt is not None else x
            Jt[k + 1] = Jt[k] + dt * KparDiff.applyDivergence(x, a, firstVar=None if zt is None else at_pts)
            if affine is not None:
                Jt[k + 1] += dt * np.trace(affine[0][k])
    if zt is None and Jt is None:
        return xt, at
    out = [xt, at]
    if zt is not None:
        out.append(zt)
    if Jt is not None:
        out.append(Jt)
    return out


def secondOrderHamiltonian(x, a, rho, px, pa, KparDiff, affine=None):
    """pointEvolution.py:952-962."""
    Ht = (px * KparDiff.applyK(x, a)).sum()
    Ht += (pa * (-KparDiff.applyDiffKT(x, a[np.newaxis], a[np.newaxis]) + rho)).sum()
    Ht -= (rho ** 2).sum() / 2
    if affine is not None:
        A, b = affine
        Ht += (px * (np.dot(x, A.T) + b)).sum() - (pa * np.dot(a, A)).sum()
    return Ht


def secondOrderCovector(x0, a0, rhot, px1, pa1, KparDiff, affine=None, times=None):
    """pointEvolution.py:965-1037: backward adjoint of secondOrderEvolution with the
    time-series end points px1[j], pa1[j] added at steps (j+1) T / len(px1)."""
    T = rhot.shape[0]
    nTarg = len(px1)
    Tsize1 = T // nTarg
    N, dim = x0.shape
    dt = 1.0 / T
    xt, at = secondOrderEvolution(x0, a0, rhot, KparDiff)
    pxt = np.zeros([T + 1, N, dim])
    pat = np.zeros([T + 1, N, dim])
    pxt[T], pat[T] = px1[nTarg - 1], pa1[nTarg - 1]
    for t in range(T):
        px, pa = pxt[T - t], pat[T - t]
        x, a = xt[T - t - 1], at[T - t - 1]
        if affine is not None:
            U = np.eye(dim) + dt * affine[0][T - t - 1]
            px_ = np.dot(px, U)
            pa_ = np.dot(pa, np.linalg.inv(U).T)
        else:
            px_, pa_ = px, pa
        a1 = np.concatenate((px_[np.newaxis], a[np.newaxis]))
        a2 = np.concatenate((a[np.newaxis], px_[np.newaxis]))
        zpx = KparDiff.applyDiffKT(x, a1, a2) - KparDiff.applyDDiffK11and12(x, a, a, pa_)
        zpa = KparDiff.applyK(x, px_) - KparDiff.applyDiffK1and2(x, pa_, a)
        pxt[T - t - 1] = px_ + dt * zpx
        pat[T - t - 1] = pa_ + dt * zpa
        if t < T - 1 and (T - 1 - t) % Tsize1 == 0:
            pxt[T - t - 1] += px1[(T - t - 1) // Tsize1 - 1]
            pat[T - t - 1] += pa1[(T - t - 1) // Tsize1 - 1]
    return pxt, pat, xt, at


def secondOrderGradient(x0, a0, rhot, px1, pa1, KparDiff, times=None, getCovector=False,
                        affine=None, controlWeight=1.0):
    """pointEvolution.py:1040-1084.  drhot and dA, db are computed in the affine branch only
    (drhot stays zero without affine, as in the reference)."""
    pxt, pat, xt, at = secondOrderCovector(x0, a0, rhot, px1, pa1, KparDiff, affine=affine,
                                           times=times)
    T = rhot.shape[0]
    dt = 1.0 / T
    drhot = np.zeros(rhot.shape)
    if affine is not None:
        dim = x0.shape[1]
        dA = np.zeros(affine[0].shape)
        for k in range(T):
            x, a, rho = xt[k], at[k], rhot[k]
            px, pa = pxt[k + 1], pat[k + 1]
            zx = x + dt * KparDiff.applyK(x, a)
            za = a + dt * (-KparDiff.applyDiffKT(x, a[np.newaxis], a[np.newaxis]) + rho)
            Ui = np.linalg.inv(np.eye(dim) + dt * affine[0][k])
            pa = np.dot(pa, Ui.T)
            za = np.dot(za, Ui)
            dA[k] = ((px[:, :, np.newaxis] * zx[:, np.newaxis, :]).sum(axis=0)
                     - (za[:, :, np.newaxis] * pa[:, np.newaxis, :]).sum(axis=0))
            drhot[k] = rho * controlWeight - pa
        db = pxt[1:T + 1].sum(axis=1)
    da0 = KparDiff.applyK(x0, a0) - pat[0]
    if affine is None:
        return (da0, drhot, xt, at, pxt, pat) if getCovector else (da0, drhot, xt, at)
    return (da0, drhot, dA, db, xt, at, pxt, pat) if getCovector else (da0, drhot, dA, db, xt, at)


# --------------------------------------------------------------------------------------
# Surfaces and curves (geometry and currents data term)
# --------------------------------------------------------------------------------------
class Surface:
    """surfaces.Surface restated for the FV / copy constructors (surfaces.py:24-75,261-263,711-717)."""

    def __init__(self, surf=None, FV=None):
        if surf is None:
            self.vertices = np.copy(FV[1]).astype(np.float64)
            self.faces = np.int_(FV[0])
        else:
            self.vertices = np.copy(surf.vertices)
            self.faces = np.copy(surf.faces)
        self.computeCentersAreas()

    def computeCentersAreas(self):
        x1 = self.vertices[self.faces[:, 0]]
        x2 = self.vertices[self.faces[:, 1]]
        x3 = self.vertices[self.faces[:, 2]]
        self.centers = (x1 + x2 + x3) / 3
        self.surfel = np.cross(x2 - x1, x3 - x1)

    def updateVertices(self, x0):
        self.vertices = np.copy(x0)
        self.computeCentersAreas()

    def flipFaces(self):
        self.faces = self.faces[:, [0, 2, 1]]
        self.computeCentersAreas()

    def surfVolume(self):
        f, v = self.faces, self.vertices
        z = 0
        for c in f:
            z += np.linalg.det(v[c[:], :]) / 6
        return z


def vertex_area(V, F):
    """surfaces.Surface.computeVertexArea (surfaces.py:77-126), face by face: (AV (nv, 1),
    AF (nf, 1)); an obtuse face gives half its area to the obtuse vertex and a quarter to the
    others, a non-obtuse face each vertex's Voronoi part."""
    nv, nf = V.shape[0], F.shape[0]
    AF = np.zeros([nf, 1])
    AV = np.zeros([nv, 1])
    for k in range(nf):
        i, j, l = F[k]
        x12, x13, x23 = V[j] - V[i], V[l] - V[i], V[l] - V[j]
        n12, n13, n23 = (np.sqrt((e ** 2).sum()) for e in (x12, x13, x23))
        c1 = (x12 * x13).sum() / (n12 * n13)
        c2 = -(x12 * x23).sum() / (n12 * n23)
        c3 = (x13 * x23).sum() / (n13 * n23)
        AF[k] = np.sqrt((np.cross(x12, x13) ** 2).sum()) / 2
        if c1 < 0:
            AV[i] += AF[k] / 2; AV[j] += AF[k] / 4; AV[l] += AF[k] / 4
        elif c2 < 0:
            AV[i] += AF[k] / 4; AV[j] += AF[k] / 2; AV[l] += AF[k] / 4
        elif c3 < 0:
            AV[i] += AF[k] / 4; AV[j] += AF[k] / 4; AV[l] += AF[k] / 2
        else:
            cot1, cot2, cot3 = (c / np.sqrt(1 - c ** 2) for c in (c1, c2, c3))
            AV[i] += ((x12 ** 2).sum() * cot3 + (x13 ** 2).sum() * cot2) / 8
            AV[j] += ((x12 ** 2).sum() * cot3 + (x23 ** 2).sum() * cot1) / 8
            AV[l] += ((x13 ** 2).sum() * cot2 + (x23 ** 2).sum() * cot1) / 8
    return AV, AF


def vertex_normals(V, F):
    """surfaces.Surface.computeVertexNormals (surfaces.py:128-139): the normalised sum of the
    surfels of each vertex's faces, face by face."""
    normals = np.zeros(V.shape)
    for k in range(F.shape[0]):
        s = np.cross(V[F[k, 1]] - V[F[k, 0]], V[F[k, 2]] - V[F[k, 0]])
        for c in range(3):
            normals[F[k, c]] += s
    return normals / np.sqrt((normals ** 2).sum(axis=1)).reshape(-1, 1)


def grid_lines(n, m):
    """grid.Grid's edges for an (n, m) lattice (grid.py:18-30): along each row, then along
    each column, vertices numbered row-major."""
    out = []
    for k in range(n):
        for l in range(m - 1):
            out.append((k * m + l, k * m + l + 1))
    for k in range(n - 1):
        for l in range(m):
            out.append((k * m + l, (k + 1) * m + l))
    return np.array(out, dtype=np.int64)


def surf_currentNorm0(fv1, K):
    return np.multiply(fv1.surfel, K.applyK(fv1.centers, fv1.surfel)).sum()


def surf_currentNormDef(fvDef, fv1, K):
    c1, cr1, c2, cr2 = fvDef.centers, fvDef.surfel, fv1.centers, fv1.surfel
    return (np.multiply(cr1, K.applyK(c1, cr1)).sum()
            - 2 * np.multiply(cr1, K.applyK(c2, cr2, firstVar=c1)).sum())


def surf_currentNormGradient(fvDef, fv1, K):
    """surfaces.py:1040-1089."""
    xDef = fvDef.vertices
    c1, cr1, c2, cr2 = fvDef.centers, fvDef.surfel, fv1.centers, fv1.surfel
    dim = c1.shape[1]
    z1 = K.applyK(c1, cr1) - K.applyK(c2, cr2, firstVar=c1)
    dz1 = (1. / 3.) * (K.applyDiffKT(c1, cr1[np.newaxis], cr1[np.newaxis])
                       - K.applyDiffKT(c2, cr1[np.newaxis], cr2[np.newaxis], firstVar=c1))
    xDef1 = xDef[fvDef.faces[:, 0]]
    xDef2 = xDef[fvDef.faces[:, 1]]
    xDef3 = xDef[fvDef.faces[:, 2]]
    px = np.zeros([xDef.shape[0], dim])
    for I, crs in ((fvDef.faces[:, 0], np.cross(xDef3 - xDef2, z1)),
                   (fvDef.faces[:, 1], np.cross(xDef1 - xDef3, z1)),
                   (fvDef.faces[:, 2], np.cross(xDef2 - xDef1, z1))):
        for k in range(I.size):
            px[I[k]] = px[I[k]] + dz1[k] - crs[k]
    return 2 * px


def _measure_weights(cr, eps=1e-10):
    """Face weights of the measure data term: sqrt(|cr|^2 + eps) per face."""
    return np.sqrt((cr ** 2).sum(axis=1) + eps)


def surf_measureNorm0(fv1, K):
    """surfaces.py:1092-1099: <a2, K(c2, c2) a2>, a2 = sqrt(|surfel|^2 + 1e-10)."""
    cr2 = _measure_weights(fv1.surfel)[:, np.newaxis]
    return np.multiply(cr2, K.applyK(fv1.centers, cr2)).sum()


def surf_measureNormDef(fvDef, fv1, K):
    """surfaces.py:1104-1117."""
    c1, c2 = fvDef.centers, fv1.centers
    cr1 = _measure_weights(fvDef.surfel)[:, np.newaxis]
    cr2 = _measure_weights(fv1.surfel)[:, np.newaxis]
    return (np.multiply(cr1, K.applyK(c1, cr1)).sum()
            - 2 * np.multiply(cr1, K.applyK(c2, cr2, firstVar=c1)).sum())


def surf_measureNormGradient(fvDef, fv1, K):
    """surfaces.py:1125-1182: z1 = (K(c1,c1) a1 - K(c1,c2) a2) * surfel/a1 and
    dz1 = (1/3)[applyDiffKT(c1,[a1],[a1]) - applyDiffKT(c2,[a1],[a2],c1)], then the
    currents vertex scatter."""
    xDef = fvDef.vertices
    c1, c2 = fvDef.centers, fv1.centers
    dim = c1.shape[1]
    a1 = _measure_weights(fvDef.surfel)
    a2 = _measure_weights(fv1.surfel)
    cr1 = fvDef.surfel / a1[:, np.newaxis]
    z1 = K.applyK(c1, a1[:, np.newaxis]) - K.applyK(c2, a2[:, np.newaxis], firstVar=c1)
    z1 = np.multiply(z1, cr1)
    w1 = a1[np.newaxis, :, np.newaxis]
    dz1 = (1. / 3.) * (K.applyDiffKT(c1, w1, w1)
                       - K.applyDiffKT(c2, w1, a2[np.newaxis, :, np.newaxis], firstVar=c1))
    xDef1 = xDef[fvDef.faces[:, 0]]
    xDef2 = xDef[fvDef.faces[:, 1]]
    xDef3 = xDef[fvDef.faces[:, 2]]
    px = np.zeros([xDef.shape[0], dim])
    for I, crs in ((fvDef.faces[:, 0], np.cross(xDef3 - xDef2, z1)),
                   (fvDef.faces[:, 1], np.cross(xDef1 - xDef3, z1)),
                   (fvDef.faces[:, 2], np.cross(xDef2 - xDef1, z1))):
        for k in range(I.size):
            px[I[k]] = px[I[k]] + dz1[k] - crs[k]
    return 2 * px


def surf_varifoldNorm0(fv1, K):
    """surfaces.py:1184-1197 (d = 1): sum_kl K(c_k, c_l) a_k a_l (1 + <u_k, u_l>^2), with
    a = sqrt(|surfel|^2 + 1e-10) and u = surfel / a, through dense (F, F) pair weights."""
    d = 1
    c2 = fv1.centers
    a2 = _measure_weights(fv1.surfel)
    cr2 = fv1.surfel / a2[:, np.newaxis]
    cr2cr2 = (cr2[:, np.newaxis, :] * cr2[np.newaxis, :, :]).sum(axis=2)
    a2a2 = a2[:, np.newaxis] * a2[np.newaxis, :]
    beta2 = (1 + d * cr2cr2 ** 2) * a2a2
    return K.applyK(c2, beta2[..., np.newaxis], matrixWeights=True).sum()


def surf_varifoldNormDef(fvDef, fv1, K):
    """surfaces.py:1200-1221."""
    d = 1
    c1, c2 = fvDef.centers, fv1.centers
    a1 = _measure_weights(fvDef.surfel)
    a2 = _measure_weights(fv1.surfel)
    cr1 = fvDef.surfel / a1[:, np.newaxis]
    cr2 = fv1.surfel / a2[:, np.newaxis]
    cr1cr1 = (cr1[:, np.newaxis, :] * cr1[np.newaxis, :, :]).sum(axis=2)
    a1a1 = a1[:, np.newaxis] * a1[np.newaxis, :]
    cr1cr2 = (cr1[:, np.newaxis, :] * cr2[np.newaxis, :, :]).sum(axis=2)
    a1a2 = a1[:, np.newaxis] * a2[np.newaxis, :]
    beta1 = (1 + d * cr1cr1 ** 2) * a1a1
    beta2 = (1 + d * cr1cr2 ** 2) * a1a2
    return (K.applyK(c1, beta1[..., np.newaxis], matrixWeights=True).sum()
            - 2 * K.applyK(c2, beta2[..., np.newaxis], firstVar=c1, matrixWeights=True).sum())


def surf_varifoldNormGradient(fvDef, fv1, K):
    """surfaces.py:1228-1276: z1 from the matrix-weighted sums of
    u1[k,l] = (2 g_kl u_l - g_kl^2 u_k + u_k) a_l (g = <u_k, u_l>), dz1 from applyDiffKmat of
    the pair weights, then the currents vertex scatter."""
    d = 1
    xDef = fvDef.vertices
    c1, c2 = fvDef.centers, fv1.centers
    dim = c1.shape[1]
    a1 = _measure_weights(fvDef.surfel)
    a2 = _measure_weights(fv1.surfel)
    cr1 = fvDef.surfel / a1[:, np.newaxis]
    cr2 = fv1.surfel / a2[:, np.newaxis]
    cr1cr1 = (cr1[:, np.newaxis, :] * cr1[np.newaxis, :, :]).sum(axis=2)
    cr1cr2 = (cr1[:, np.newaxis, :] * cr2[np.newaxis, :, :]).sum(axis=2)
    beta1 = a1[:, np.newaxis] * a1[np.newaxis, :] * (1 + d * cr1cr1 ** 2)
    beta2 = a1[:, np.newaxis] * a2[np.newaxis, :] * (1 + d * cr1cr2 ** 2)
    u1 = (2 * d * cr1cr1[..., np.newaxis] * cr1[np.newaxis, ...]
          - d * (cr1cr1 ** 2)[..., np.newaxis] * cr1[:, np.newaxis, :]
          + cr1[:, np.newaxis, :]) * a1[np.newaxis, :, np.newaxis]
    u2 = (2 * d * cr1cr2[..., np.newaxis] * cr2[np.newaxis, ...]
          - d * (cr1cr2 ** 2)[..., np.newaxis] * cr1[:, np.newaxis, :]
          + cr1[:, np.newaxis, :]) * a2[np.newaxis, :, np.newaxis]
    z1 = K.applyK(c1, u1, matrixWeights=True) - K.applyK(c2, u2, firstVar=c1, matrixWeights=True)
    dz1 = (1. / 3.) * (K.applyDiffKmat(c1, beta1) - K.applyDiffKmat(c2, beta2, firstVar=c1))
    xDef1 = xDef[fvDef.faces[:, 0]]
    xDef2 = xDef[fvDef.faces[:, 1]]
    xDef3 = xDef[fvDef.faces[:, 2]]
    px = np.zeros([xDef.shape[0], dim])
    for I, crs in ((fvDef.faces[:, 0], np.cross(xDef3 - xDef2, z1)),
                   (fvDef.faces[:, 1], np.cross(xDef1 - xDef3, z1)),
                   (fvDef.faces[:, 2], np.cross(xDef2 - xDef1, z1))):
        for k in range(I.size):
            px[I[k]] = px[I[k]] + dz1[k] - crs[k]
    return 2 * px


class Curve:
    """curves.Curve restated for the pointSet / FV / copy constructors (curves.py:15-73)."""

    def __init__(self, curve=None, FV=None, pointSet=None, isOpen=False):
        if curve is not None:
            self.vertices = np.copy(curve.vertices)
            self.faces = np.copy(curve.faces)
        elif FV is not None:
            self.vertices = np.copy(FV[1]).astype(np.float64)
            self.faces = np.int_(FV[0])
        else:
            n = pointSet.shape[0]
            self.vertices = np.copy(pointSet).astype(np.float64)
            self.faces = np.int_(np.zeros([n, 2]))
            for k in range(n - 1):
                self.faces[k] = (k, k + 1)
            if not isOpen:
                self.faces[n - 1] = (n - 1, 0)
        self.computeCentersLengths()

    def computeCentersLengths(self):
        x1 = self.vertices[self.faces[:, 0]]
        x2 = self.vertices[self.faces[:, 1]]
        self.centers = (x1 + x2) / 2
        self.lenel = x2 - x1

    def updateVertices(self, x0):
        self.vertices = np.copy(x0)
        self.computeCentersLengths()


def curve_currentNorm0(fv1, K):
    return np.multiply(fv1.lenel, K.applyK(fv1.centers, fv1.lenel)).sum()


def curve_currentNormDef(fvDef, fv1, K):
    c1, cr1, c2, cr2 = fvDef.centers, fvDef.lenel, fv1.centers, fv1.lenel
    return (np.multiply(cr1, K.applyK(c1, cr1)).sum()
            - 2 * np.multiply(cr1, K.applyK(c2, cr2, firstVar=c1)).sum())


def curve_currentNormGradient(fvDef, fv1, K):
    """curves.py:566-613."""
    xDef = fvDef.vertices
    c1, cr1, c2, cr2 = fvDef.centers, fvDef.lenel, fv1.centers, fv1.lenel
    dim = c1.shape[1]
    z1 = K.applyK(c1, cr1) - K.applyK(c2, cr2, firstVar=c1)
    dz1 = .5 * (K.applyDiffKT(c1, cr1[np.newaxis], cr1[np.newaxis])
                - K.applyDiffKT(c2, cr1[np.newaxis], cr2[np.newaxis], firstVar=c1))
    px = np.zeros([xDef.shape[0], dim])
    I = fvDef.faces[:, 0]
    for k in range(I.size):
        px[I[k]] = px[I[k]] + dz1[k] - z1[k]
    I = fvDef.faces[:, 1]
    for k in range(I.size):
        px[I[k]] = px[I[k]] + dz1[k] + z1[k]
    return 2 * px


def curve_measureNorm0(fv1, K):
    """curves.py:621-628: <a2, K(c2, c2) a2> with a2 = |lenel| (no 1e-10 here, unlike
    measureNormDef and the gradient)."""
    cr2 = _measure_weights(fv1.lenel, 0.0)[:, np.newaxis]
    return np.multiply(cr2, K.applyK(fv1.centers, cr2)).sum()


def curve_measureNormDef(fvDef, fv1, K):
    """curves.py:632-645."""
    c1, c2 = fvDef.centers, fv1.centers
    cr1 = _measure_weights(fvDef.lenel)[:, np.newaxis]
    cr2 = _measure_weights(fv1.lenel)[:, np.newaxis]
    return (np.multiply(cr1, K.applyK(c1, cr1)).sum()
            - 2 * np.multiply(cr1, K.applyK(c2, cr2, firstVar=c1)).sum())


def curve_measureNormGradient(fvDef, fv1, K):
    """curves.py:653-701: as the surface form with dz1 = (1/2)[...] and the edge scatter
    px[f0] += dz1 - z1, px[f1] += dz1 + z1."""
    xDef = fvDef.vertices
    c1, c2 = fvDef.centers, fv1.centers
    dim = c1.shape[1]
    a1 = _measure_weights(fvDef.lenel)
    a2 = _measure_weights(fv1.lenel)
    cr1 = fvDef.lenel / a1[:, np.newaxis]
    z1 = K.applyK(c1, a1[:, np.newaxis]) - K.applyK(c2, a2[:, np.newaxis], firstVar=c1)
    z1 = np.multiply(z1, cr1)
    w1 = a1[np.newaxis, :, np.newaxis]
    dz1 = .5 * (K.applyDiffKT(c1, w1, w1)
                - K.applyDiffKT(c2, w1, a2[np.newaxis, :, np.newaxis], firstVar=c1))
    px = np.zeros([xDef.shape[0], dim])
    I = fvDef.faces[:, 0]
    for k in range(I.size):
        px[I[k]] = px[I[k]] + dz1[k] - z1[k]
    I = fvDef.faces[:, 1]
    for k in range(I.size):
        px[I[k]] = px[I[k]] + dz1[k] + z1[k]
    return 2 * px


_DATA_TERMS = {
    'surface': {'current': (surf_currentNorm0, surf_currentNormDef, surf_currentNormGradient),
                'measure': (surf_measureNorm0, surf_measureNormDef, surf_measureNormGradient),
                'varifold': (surf_varifoldNorm0, surf_varifoldNormDef, surf_varifoldNormGradient)},
    'curve': {'current': (curve_currentNorm0, curve_currentNormDef, curve_currentNormGradient),
              'measure': (curve_measureNorm0, curve_measureNormDef, curve_measureNormGradient)},
}


# --------------------------------------------------------------------------------------
# Matching drivers and the CG optimiser
# --------------------------------------------------------------------------------------
class Direction:
    def __init__(self):
        self.diff = []
        self.aff = []
        self.initx = []


class _Matching:
    """Shared body of SurfaceMatching (surfaceMatching.py:74-484) and CurveMatching
    (curveMatching.py:58-352) for symmetric=False, with the global affine part
    (affine 'affine' / 'similitude' / 'euclidean' / 'translation' / 'none').  ``aff_style``
    selects the two drivers' affine scalings: 'surface' (coeffAff, surfaceMatching.py:315-322,
    367-369, 388-389, 475-481) or 'curve' (affineWeight, curveMatching.py:246-252,288-296)."""

    aff_style = 'surface'

    def _setup(self, x0, KparDiff, KparDist, sigmaError, timeStep, regWeight, affine='none',
               affineWeight=1.0, rotWeight=None, scaleWeight=None, transWeight=None):
        self.x0 = np.copy(x0)
        self.npt, self.dim = x0.shape
        self.KparDiff, self.KparDist = KparDiff, KparDist
        self.sigmaError = sigmaError
        self.regweight = regWeight
        self.Tsize = int(round(1.0 / timeStep))
        self.at = np.zeros([self.Tsize, self.npt, self.dim])
        self.atTry = np.zeros_like(self.at)
        self.xt = np.tile(self.x0, [self.Tsize + 1, 1, 1])
        self.obj = None
        self.objTry = None
        self.gradCoeff = self.npt
        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        self.affineBasis = self.affB.basis
        self.affineWeight = affineWeight * np.ones([self.affineDim, 1])
        if self.affB.rotComp and rotWeight is not None:
            self.affineWeight[self.affB.rotComp] = rotWeight
        if self.affB.simComp and scaleWeight is not None:
            self.affineWeight[self.affB.simComp] = scaleWeight
        if self.affB.transComp and transWeight is not None:
            self.affineWeight[self.affB.transComp] = transWeight
        self.Afft = np.zeros([self.Tsize, self.affineDim])
        self.AfftTry = np.zeros([self.Tsize, self.affineDim])
        self.coeffAff1, self.coeffAff2 = 1., 100.
        self.coeffAff = self.coeffAff1
        self.affBurnIn = 20
        self.nobj = 0  # objective evaluations (updateTry + objectiveFun)

    def _transforms(self, Afft):
        return self.affB.getTransforms(Afft) if self.affineDim > 0 else None

    def dataTerm(self, fvDef):
        return self.fun_obj(fvDef, self.fv1, self.KparDist) / (self.sigmaError ** 2)

    def objectiveFunDef(self, at, Afft=None, withTrajectory=False):
        timeStep = 1.0 / self.Tsize
        Afft = self.Afft if Afft is None else Afft
        xt = landmarkDirectEvolutionEuler(self.x0, at, self.KparDiff, affine=self._transforms(Afft))
        obj = 0
        for t in range(self.Tsize):
            z = xt[t]
            a = at[t]
            ra = self.KparDiff.applyK(z, a)
            obj = obj + self.regweight * timeStep * np.multiply(a, ra).sum()
            if self.affineDim > 0:
                obj += timeStep * np.multiply(self.affineWeight.reshape(Afft[t].shape), Afft[t] ** 2).sum()
        if withTrajectory:
            return obj, xt
        return obj

    def objectiveFun(self):
        if self.obj is None:
            self.obj0 = self.fun_obj0(self.fv1, self.KparDist) / (self.sigmaError ** 2)
            self.obj, self.xt = self.objectiveFunDef(self.at, self.Afft, withTrajectory=True)
            self.fvDef.updateVertices(self.xt[-1])
            self.obj += self.obj0 + self.dataTerm(self.fvDef)
            self.nobj += 1
        return self.obj

    def getVariable(self):
        return [self.at, self.Afft]

    def updateTry(self, dir, eps, objRef=None):
        objTry = self.obj0
        atTry = self.at - eps * dir.diff
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        foo = self.objectiveFunDef(atTry, AfftTry, withTrajectory=True)
        objTry += foo[0]
        ff = self._copy_shape(self.fvDef)
        ff.updateVertices(foo[1][-1])
        objTry += self.dataTerm(ff)
        self.nobj += 1
        if np.isnan(objTry):
            return 1e500
        if (objRef is None) or (objTry < objRef):
            self.atTry = atTry
            self.objTry = objTry
            self.AfftTry = AfftTry
        return objTry

    def endPointGradient(self):
        return self.fun_objGrad(self.fvDef, self.fv1, self.KparDist) / self.sigmaError ** 2

    def getGradient(self, coeff=1.0):
        px1 = -self.endPointGradient()
        A = self._transforms(self.Afft)
        foo = landmarkHamiltonianGradient(self.x0, self.at, px1, self.KparDiff, self.regweight,
                                          affine=A)
        grd = Direction()
        grd.diff = foo[0] / (coeff * self.Tsize)
        grd.aff = np.zeros(self.Afft.shape)
        if self.affineDim > 0:
            dA, db = foo[1], foo[2]
            dim2 = self.dim ** 2
            w = self.affineWeight.reshape([1, self.affineDim])
            grd.aff = 2 * np.multiply(w, self.Afft) if self.aff_style == 'surface' else 2 * self.Afft
            for t in range(self.Tsize):
                dAff = np.dot(self.affineBasis.T, np.vstack([dA[t].reshape([dim2, 1]),
                                                             db[t].reshape([self.dim, 1])]))
                if self.aff_style == 'surface':
                    grd.aff[t] -= dAff.reshape(grd.aff[t].shape)
                else:
                    grd.aff[t] -= np.divide(dAff.reshape(grd.aff[t].shape),
                                            self.affineWeight.reshape(grd.aff[t].shape))
            if self.aff_style == 'surface':
                grd.aff /= (self.coeffAff * coeff * self.Tsize)
            else:
                grd.aff /= (coeff * self.Tsize)
        return grd

    def addProd(self, dir1, dir2, beta):
        d = Direction()
        d.diff = dir1.diff + beta * dir2.diff
        d.aff = dir1.aff + beta * dir2.aff
        return d

    def copyDir(self, dir0):
        d = Direction()
        d.diff = np.copy(dir0.diff)
        d.aff = np.copy(dir0.aff)
        return d

    def dotProduct(self, g1, g2):
        res = np.zeros(len(g2))
        for t in range(self.Tsize):
            z = self.xt[t]
            u = self.KparDiff.applyK(z, g1.diff[t])
            if self.aff_style == 'curve':
                uu = np.multiply(g1.aff[t], self.affineWeight.reshape(g1.aff[t].shape))
            else:
                uu = g1.aff[t]
            for ll, gr in enumerate(g2):
                res[ll] = res[ll] + np.multiply(gr.diff[t], u).sum()
                if self.affineDim > 0:
                    if self.aff_style == 'curve':
                        res[ll] += np.multiply(uu, gr.aff[t]).sum()
                    else:
                        res[ll] += np.multiply(uu, gr.aff[t]).sum() * self.coeffAff
        return res

    def acceptVarTry(self):
        self.obj = self.objTry
        self.at = np.copy(self.atTry)
        self.Afft = np.copy(self.AfftTry)

    def endOfIteration(self):
        self.iter += 1
        if self.aff_style == 'surface' and self.iter >= self.affBurnIn:
            self.coeffAff = self.coeffAff2
        obj1, self.xt = self.objectiveFunDef(self.at, self.Afft, withTrajectory=True)
        self.fvDef.updateVertices(self.xt[-1])


class SurfaceMatching(_Matching):
    """surfaceMatching.SurfaceMatching with errorType 'current', 'measure' or 'varifold'
    (surfaceMatching.py:29-40,74-484)."""

    def __init__(self, Template, Target, KparDiff, KparDist, sigmaError=1.0, timeStep=.1,
                 regWeight=1.0, maxIter=1000, errorType='current', affine='none', affineWeight=1.0,
                 rotWeight=None, scaleWeight=None, transWeight=None):
        self.fv0 = Surface(surf=Template)
        self.fv1 = Surface(surf=Target)
        v0 = self.fv0.surfVolume()
        v1 = self.fv1.surfVolume()
        if v0 * v1 < 0:
            self.fv1.flipFaces()
        self.fvInit = Surface(surf=self.fv0)
        self.fvDef = Surface(surf=self.fvInit)
        self.fun_obj0, self.fun_obj, self.fun_objGrad = _DATA_TERMS['surface'][errorType]
        self._setup(self.fvInit.vertices, KparDiff, KparDist, sigmaError, timeStep, regWeight,
                    affine, affineWeight, rotWeight, scaleWeight, transWeight)
        self.maxIter = maxIter
        self.iter = 0
        self.epsInit = 0.1
        self.gradLB = 0.001

    @staticmethod
    def _copy_shape(s):
        return Surface(surf=s)

    def optimizeMatching(self, cgfun=None):
        """surfaceMatching.py:474-484: the gradient bound with coeffAff2, the optimisation
        starting with coeffAff1."""
        self.coeffAff = self.coeffAff2
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        self.coeffAff = self.coeffAff1
        cg(self, maxIter=self.maxIter, epsInit=self.epsInit)
        return self.history


class SymmetricSurfaceMatching(SurfaceMatching):
    """surfaceMatching.SurfaceMatching with symmetric=True (surfaceMatching.py:181-185,
    232-289,297-299,325-327,373-376,383,466): the template's points x0 are a variable, the data
    term adds |S(x0) - S0|^2/sigmaError^2 (obj0 adds |S0|^2), the x0 gradient is
    (initPointGradient - p_0)/(coeffInitx coeff) with coeffInitx = 0.1, which also weighs the
    x0 part of the dot product; the moved template's surface is refreshed in endOfIteration.
    The flow takes the affine branch with zero matrices when there is no affine part, as the
    reference's objectiveFunDef and getGradient always pass one."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.x0Try = np.copy(self.x0)
        self.coeffInitx = .1

    def _transforms(self, Afft):
        if self.affineDim > 0:
            return self.affB.getTransforms(Afft)
        return [np.zeros([self.Tsize, self.dim, self.dim]), np.zeros([self.Tsize, self.dim])]

    def dataTerm(self, fvDef, fvInit=None):
        obj = self.fun_obj(fvDef, self.fv1, self.KparDist) / (self.sigmaError ** 2)
        if fvInit is not None:
            obj += self.fun_obj(fvInit, self.fv0, self.KparDist) / (self.sigmaError ** 2)
        return obj

    def objectiveFunDef(self, at, Afft=None, withTrajectory=False, x0=None):
        saved = self.x0
        if x0 is not None:
            self.x0 = x0
        try:
            return super().objectiveFunDef(at, Afft, withTrajectory=withTrajectory)
        finally:
            self.x0 = saved

    def objectiveFun(self):
        if self.obj is None:
            self.obj0 = self.fun_obj0(self.fv1, self.KparDist) / (self.sigmaError ** 2)
            self.obj0 += self.fun_obj0(self.fv0, self.KparDist) / (self.sigmaError ** 2)
            self.obj, self.xt = self.objectiveFunDef(self.at, self.Afft, withTrajectory=True)
            self.fvDef.updateVertices(self.xt[-1])
            self.fvInit.updateVertices(self.x0)
            self.obj += self.obj0 + self.dataTerm(self.fvDef, self.fvInit)
        return self.obj

    def updateTry(self, dir, eps, objRef=None):
        objTry = self.obj0
        atTry = self.at - eps * dir.diff
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        x0Try = self.x0 - eps * dir.initx
        obj, xt = self.objectiveFunDef(atTry, AfftTry, withTrajectory=True, x0=x0Try)
        objTry += obj
        ff = Surface(surf=self.fvDef)
        ff.updateVertices(xt[-1])
        ffI = Surface(surf=self.fvInit)
        ffI.updateVertices(x0Try)
        objTry += self.dataTerm(ff, ffI)
        if np.isnan(objTry):
            return 1e500
        if (objRef is None) or (objTry < objRef):
            self.atTry, self.objTry, self.AfftTry, self.x0Try = atTry, objTry, AfftTry, x0Try
        return objTry

    def getGradient(self, coeff=1.0):
        px1 = -self.endPointGradient()
        foo = landmarkHamiltonianGradient(self.x0, self.at, px1, self.KparDiff, self.regweight,
                                          affine=self._transforms(self.Afft), getCovector=True)
        grd = Direction()
        grd.diff = foo[0] / (coeff * self.Tsize)
        grd.aff = np.zeros(self.Afft.shape)
        if self.affineDim > 0:
            dA, db = foo[1], foo[2]
            dim2 = self.dim ** 2
            grd.aff = 2 * np.multiply(self.affineWeight.reshape([1, self.affineDim]), self.Afft)
            for t in range(self.Tsize):
                dAff = np.dot(self.affineBasis.T, np.vstack([dA[t].reshape([dim2, 1]),
                                                             db[t].reshape([self.dim, 1])]))
                grd.aff[t] -= dAff.reshape(grd.aff[t].shape)
            grd.aff /= (self.coeffAff * coeff * self.Tsize)
        initGrad = self.fun_objGrad(self.fvInit, self.fv0, self.KparDist) / self.sigmaError ** 2
        grd.initx = (initGrad - foo[4][0]) / (self.coeffInitx * coeff)
        return grd

    def addProd(self, dir1, dir2, beta):
        d = super().addProd(dir1, dir2, beta)
        d.initx = dir1.initx + beta * dir2.initx
        return d

    def copyDir(self, dir0):
        d = super().copyDir(dir0)
        d.initx = np.copy(dir0.initx)
        return d

    def dotProduct(self, g1, g2):
        res = super().dotProduct(g1, g2)
        for ll, gr in enumerate(g2):
            res[ll] += (g1.initx * gr.initx).sum() * self.coeffInitx
        return res

    def acceptVarTry(self):
        super().acceptVarTry()
        self.x0 = np.copy(self.x0Try)

    def endOfIteration(self):
        super().endOfIteration()
        self.fvInit.updateVertices(self.x0)


class CurveMatching(_Matching):
    """curveMatching.CurveMatching with errorType 'current' or 'measure'
    (curveMatching.py:24-33,58-352)."""

    aff_style = 'curve'

    def __init__(self, Template, Target, KparDiff, KparDist, sigmaError=1.0, timeStep=.1,
                 regWeight=1.0, maxIter=1000, gradLB=0.001, errorType='current', affine='none',
                 affineWeight=1.0, rotWeight=None, scaleWeight=None, transWeight=None):
        self.fv0 = Curve(curve=Template)
        self.fv1 = Curve(curve=Target)
        self.fvDef = Curve(curve=self.fv0)
        self.fun_obj0, self.fun_obj, self.fun_objGrad = _DATA_TERMS['curve'][errorType]
        self._setup(self.fv0.vertices, KparDiff, KparDist, sigmaError, timeStep, regWeight,
                    affine, affineWeight, rotWeight, scaleWeight, transWeight)
        self.maxIter = maxIter
        self.iter = 0
        self.epsInit = 10.0
        self.gradLB = gradLB

    @staticmethod
    def _copy_shape(s):
        return Curve(curve=s)

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        cg(self, maxIter=self.maxIter, epsInit=self.epsInit)
        return self.history


class PointSet:
    """Landmark set: the deformed 'shape' of landmark matching is its point array."""

    def __init__(self, points):
        self.vertices = np.array(points, dtype=np.float64)

    def updateVertices(self, x0):
        self.vertices = np.copy(x0)


def landmark_norm0(fv1, K):
    return 0.0


def landmark_normDef(fvDef, fv1, K):
    """sum_k |x_k - y_k|^2, the energy of LandmarkObjective (diffeoReg/landmarkMatching.h:52-62)."""
    u = fvDef.vertices - fv1.vertices
    return float((u * u).sum())


def landmark_normGradient(fvDef, fv1, K):
    """d/dx of landmark_normDef: 2 (x - y)."""
    return 2.0 * (fvDef.vertices - fv1.vertices)


class LandmarkMatching(_Matching):
    """Landmark-set matching: the surface driver's call pattern (surfaceMatching.py:74-484)
    with the squared-distance data term above; epsInit 0.1 and gradient bound
    max(gradLB, |g|/10^4) as in SurfaceMatching.optimizeMatching (surfaceMatching.py:474-484)."""

    def __init__(self, Template, Target, KparDiff, sigmaError=1.0, timeStep=.1, regWeight=1.0,
                 maxIter=1000, gradLB=0.001, epsInit=0.1):
        self.fv0 = PointSet(Template)
        self.fv1 = PointSet(Target)
        self.fvDef = PointSet(Template)
        self.fun_obj0 = landmark_norm0
        self.fun_obj = landmark_normDef
        self.fun_objGrad = landmark_normGradient
        self._setup(self.fv0.vertices, KparDiff, None, sigmaError, timeStep, regWeight)
        self.maxIter = maxIter
        self.iter = 0
        self.epsInit = epsInit
        self.gradLB = gradLB

    @staticmethod
    def _copy_shape(s):
        return PointSet(s.vertices)

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(self.gradLB, np.sqrt(grd2) / 10000)
        cg(self, maxIter=self.maxIter, epsInit=self.epsInit)
        return self.history


def cg(opt, verb=True, maxIter=1000, TestGradient=False, epsInit=10.):
    """conjugateGradient.cg restated (conjugateGradient.py:33-217), with its optional hooks
    (startOptim, startOfIteration, endOfIteration, endOptim: :40-41,72-73,211-212,214-215), the
    gradient test (:76-92), the plain-array fallbacks when the object has no dotProduct /
    copyDir / addProd (:95-131) and its log lines (root logger, :66,143-144,161-162,185-186,
    198-199).  Returns opt.getVariable() as the reference does; the accepted objectives are
    kept on ``opt.history`` and the stop on ``opt.stop_reason`` ('maxIter' when none fired)."""
    if hasattr(opt, 'startOptim'):
        opt.startOptim()
    gradEps = getattr(opt, 'gradEps', 1.0)
    cgBurnIn = getattr(opt, 'cgBurnIn', -1)
    gradCoeff = getattr(opt, 'gradCoeff', 1.0)
    restartRate = getattr(opt, 'restartRate', 100)
    hasDot = hasattr(opt, 'dotProduct')

    def dot(g, gs):
        return opt.dotProduct(g, gs) if hasDot else [np.multiply(g, h).sum() for h in gs]

    def copy(d):
        return opt.copyDir(d) if hasattr(opt, 'copyDir') else np.copy(d)

    def itlog(it, obj, eps, beta, grd2):
        logging.info('iteration {0:d}: obj = {1:.5f}, eps = {2:.5f}, beta = {3:.5f}, '
                     'gradient: {4:.5f}'.format(it + 1, obj, eps, beta, np.sqrt(grd2)))

    eps = epsInit
    epsMin = 1e-10
    opt.converged = False
    obj = opt.objectiveFun()
    logging.info('iteration 0: obj = {0: .5f}'.format(obj))
    opt.history = [obj]
    opt.stop_reason = 'maxIter'
    skipCG = 0
    for it in range(maxIter):
        if it % restartRate == 0:
            skipCG = 1
        if hasattr(opt, 'startOfIteration'):
            opt.startOfIteration()
        grd = opt.getGradient(gradCoeff)
        if TestGradient:
            dirfoo = opt.randomDir() if hasattr(opt, 'randomDir') else np.random.normal(size=grd.shape)
            epsfoo = 1e-7
            objfoo = opt.updateTry(dirfoo, epsfoo, obj - 1e10)
            [grdfoo] = dot(grd, [dirfoo])
            logging.info('Test Gradient: %.4f %.4f' % ((objfoo - obj) / epsfoo, -grdfoo * gradCoeff))
        if it == 0 or it == cgBurnIn:
            [grdOld2] = dot(grd, [grd])
            grd2 = grdOld2
            grdTry = np.sqrt(max(1e-20, grdOld2))
            oldDir = copy(grd)
            grdOld = copy(grd)
            dir0 = copy(grd)
            beta = 0
        else:
            [grd2, grd12] = dot(grd, [grd, grdOld])
            if skipCG == 0:
                beta = max(0, (grd2 - grd12) / grdOld2)
            else:
                beta = 0
            grdOld2 = grd2
            grdTry = np.sqrt(max(1e-20, grd2 + beta * grd12))
            dir0 = opt.addProd(grd, oldDir, beta) if hasattr(opt, 'addProd') else grd + beta * oldDir
            oldDir = copy(dir0)
            grdOld = copy(grd)
        objTry = opt.updateTry(dir0, eps, obj)
        noUpdate = 0
        if objTry > obj:
            epsSmall = 0.000001 / grdTry
            objTry0 = opt.updateTry(dir0, epsSmall, obj)
            if objTry0 > obj:
                if (skipCG == 1) | (beta < 1e-10):
                    itlog(it, obj, eps, beta, grd2)
                    logging.info('Stopping Gradient Descent: bad direction')
                    opt.stop_reason = 'bad direction'
                    break
                else:
                    # the reference disables CG here only when verb is set (:151-154)
                    if verb:
                        logging.info('Disabling CG: bad direction')
                        skipCG = 1
                        noUpdate = 1
            else:
                while (objTry > obj) and (eps > epsMin):
                    eps = eps / 2
                    objTry = opt.updateTry(dir0, eps, obj)
        if noUpdate == 0:
            contt = 1
            while contt == 1:
                objTry2 = opt.updateTry(dir0, .5 * eps, objTry)
                if objTry > objTry2:
                    eps = eps / 2
                    objTry = objTry2
                else:
                    contt = 0
            contt = 1
            while contt == 1:
                objTry2 = opt.updateTry(dir0, 1.25 * eps, objTry)
                if objTry > objTry2:
                    eps *= 1.25
                    objTry = objTry2
                else:
                    contt = 0
            if np.fabs(obj - objTry) < .000001:
                if (skipCG == 1) | (beta < 1e-10):
                    itlog(it, obj, eps, beta, grd2)
                    logging.info('Stopping Gradient Descent: small variation')
                    opt.converged = True
                    opt.stop_reason = 'small variation'
                    break
                else:
                    if verb:
                        logging.info('Disabling CG: small variation')
                    skipCG = 1
                    eps = 1.0
            else:
                skipCG = 0
            opt.acceptVarTry()
            obj = objTry
            opt.history.append(obj)
            if verb | (it == maxIter):
                itlog(it, obj, eps, beta, grd2)
            if np.sqrt(grd2) < gradEps:
                logging.info('Stopping Gradient Descent: small gradient')
                opt.converged = True
                opt.stop_reason = 'small gradient'
                break
            eps = 100 * eps
            if hasattr(opt, 'endOfIteration'):
                opt.endOfIteration()
    if hasattr(opt, 'endOptim'):
        opt.endOptim()
    return opt.getVariable()


# --------------------------------------------------------------------------------------
# Second-order (spline) regression through a time series of surfaces
# --------------------------------------------------------------------------------------
class SecondOrderSurfaceMatching:
    """secondOrderMatching.SurfaceMatching (secondOrderMatching.py:43-509) restated without
    file output: variables (a0, rhot, Afft); objective 1/2 <a0, K a0> + sum_t dt cw |rho_t|^2/2
    + sum_t dt <w, Afft_t^2>/2 + data terms at times (k+1) T1; regression type 'affine' for
    the first affBurnIn iterations, then typeRegression; Euclidean dot products."""

    def __init__(self, Template, Targets, KparDiff, KparDist, sigmaError=1.0, timeStep=.1,
                 errorType='current', maxIter=1000, typeRegression='spline2', affine='none',
                 controlWeight=1.0, affineWeight=1.0, initialMomentum=None):
        self.fv0 = Surface(surf=Template)
        self.fv1 = [Surface(surf=s) for s in Targets]
        self.nTarg = len(self.fv1)
        self.iter = 0
        self.dim = self.fv0.vertices.shape[1]
        self.maxIter = maxIter
        self.KparDiff, self.KparDist, self.sigmaError = KparDiff, KparDist, sigmaError
        self.fun_obj0, self.fun_obj, self.fun_objGrad = _DATA_TERMS['surface'][errorType]
        self.affB = AffineBasis(self.dim, affine)
        self.affineDim = self.affB.affineDim
        self.affineBasis = self.affB.basis
        self.affineWeight = affineWeight * np.ones(self.affineDim)
        self.controlWeight = controlWeight
        self.typeRegression = 'affine'
        self.typeRegressionSave = typeRegression
        self.x0 = self.fv0.vertices
        self.fvDef = [Surface(surf=self.fv0) for _ in range(self.nTarg)]
        self.npt = self.x0.shape[0]
        self.Tsize1 = int(round(1.0 / timeStep))
        self.Tsize = self.nTarg * self.Tsize1
        self.rhot = np.zeros([self.Tsize, self.npt, self.dim])
        self.a0 = (np.zeros([self.npt, self.dim]) if initialMomentum is None
                   else np.asarray(initialMomentum, dtype=np.float64))
        self.rhotTry = np.zeros_like(self.rhot)
        self.a0Try = np.zeros_like(self.a0)
        self.Afft = np.zeros([self.Tsize, self.affineDim])
        self.AfftTry = np.zeros_like(self.Afft)
        self.obj = None
        self.objTry = None
        self.gradCoeff = self.npt
        self.affBurnIn = 50
        self.coeffAff = 1

    def _affine(self, Afft):
        dim2 = self.dim ** 2
        A = [np.zeros([self.Tsize, self.dim, self.dim]), np.zeros([self.Tsize, self.dim])]
        if self.affineDim > 0:
            for t in range(self.Tsize):
                AB = np.dot(self.affineBasis, Afft[t])
                A[0][t] = AB[0:dim2].reshape([self.dim, self.dim])
                A[1][t] = AB[dim2:dim2 + self.dim]
        return A

    def dataTerm(self, fvDef):
        return sum(self.fun_obj(s, self.fv1[k], self.KparDist) for k, s in enumerate(fvDef)) \
            / self.sigmaError ** 2

    def objectiveFunDef(self, a0, rhot, Afft):
        timeStep = 1.0 / self.Tsize
        xt, at = secondOrderEvolution(self.x0, a0, rhot, self.KparDiff, affine=self._affine(Afft))
        obj0 = 0.5 * (a0 * self.KparDiff.applyK(self.x0, a0)).sum()
        obj1 = obj2 = 0
        for t in range(self.Tsize):
            obj1 += timeStep * self.controlWeight * (rhot[t] ** 2).sum() / 2
            if self.affineDim > 0:
                obj2 += timeStep * (self.affineWeight * Afft[t] ** 2).sum() / 2
        return obj1 + obj2 + obj0, xt, at

    def objectiveFun(self):
        if self.obj is None:
            self.obj, self.xt, self.at = self.objectiveFunDef(self.a0, self.rhot, self.Afft)
            self.obj0 = 0
            for k in range(self.nTarg):
                self.obj0 += self.fun_obj0(self.fv1[k], self.KparDist) / self.sigmaError ** 2
                self.fvDef[k].updateVertices(self.xt[(k + 1) * self.Tsize1])
            self.obj += self.obj0 + self.dataTerm(self.fvDef)
        return self.obj

    def getVariable(self):
        return (self.a0, self.rhot, self.Afft)

    def updateTry(self, dir, eps, objRef=None):
        a0Try = self.a0 - eps * dir.a0
        rhotTry = self.rhot - eps * dir.rhot
        AfftTry = self.Afft - eps * dir.aff if self.affineDim > 0 else self.Afft
        obj, xt, _ = self.objectiveFunDef(a0Try, rhotTry, AfftTry)
        ff = []
        for k in range(self.nTarg):
            ff.append(Surface(surf=self.fvDef[k]))
            ff[k].updateVertices(xt[(k + 1) * self.Tsize1])
        objTry = self.obj0 + obj + self.dataTerm(ff)
        if np.isnan(objTry):
            return 1e500
        if objRef is None or objTry < objRef:
            self.a0Try, self.rhotTry, self.AfftTry, self.objTry = a0Try, rhotTry, AfftTry, objTry
        return objTry

    def getGradient(self, coeff=1.0):
        px1 = [-self.fun_objGrad(self.fvDef[k], self.fv1[k], self.KparDist) / self.sigmaError ** 2
               for k in range(self.nTarg)]
        pa1 = [np.zeros(self.a0.shape) for _ in range(self.nTarg)]
        foo = secondOrderGradient(self.x0, self.a0, self.rhot, px1, pa1, self.KparDiff,
                                  affine=self._affine(self.Afft),
                                  times=(1 + np.array(range(self.nTarg))) * self.Tsize1,
                                  controlWeight=self.controlWeight)
        grd = Direction()
        if self.typeRegression == 'spline':
            grd.a0, grd.rhot = np.zeros(foo[0].shape), foo[1] / coeff
        elif self.typeRegression == 'geodesic':
            grd.a0, grd.rhot = foo[0] / coeff, np.zeros(foo[1].shape)
        elif self.typeRegression == 'affine':
            grd.a0, grd.rhot = np.zeros(foo[0].shape), np.zeros(foo[1].shape)
        else:
            grd.a0, grd.rhot = foo[0] / coeff, foo[1] / coeff
        grd.aff = np.zeros(self.Afft.shape)
        if self.affineDim > 0 and self.iter < self.affBurnIn:
            dim2 = self.dim ** 2
            dA, db = foo[2], foo[3]
            grd.aff = np.multiply(self.affineWeight.reshape([1, self.affineDim]), self.Afft)
            for t in range(self.Tsize):
                dAff = np.dot(self.affineBasis.T, np.vstack([dA[t].reshape([dim2, 1]),
                                                             db[t].reshape([self.dim, 1])]))
                grd.aff[t] -= dAff.reshape(grd.aff[t].shape)
            grd.aff /= (self.coeffAff * coeff * self.Tsize)
        return grd

    def addProd(self, dir1, dir2, beta):
        d = Direction()
        d.a0 = dir1.a0 + beta * dir2.a0
        d.rhot = dir1.rhot + beta * dir2.rhot
        d.aff = dir1.aff + beta * dir2.aff
        return d

    def copyDir(self, dir0):
        d = Direction()
        d.a0, d.rhot, d.aff = np.copy(dir0.a0), np.copy(dir0.rhot), np.copy(dir0.aff)
        return d

    def dotProduct(self, g1, g2):
        res = np.zeros(len(g2))
        for ll, gr in enumerate(g2):
            res[ll] = (gr.rhot * g1.rhot).sum() / self.Tsize
            res[ll] += (gr.a0 * g1.a0).sum()
            res[ll] += (g1.aff * gr.aff).sum() * self.coeffAff
        return res

    def acceptVarTry(self):
        self.obj = self.objTry
        self.a0, self.rhot, self.Afft = np.copy(self.a0Try), np.copy(self.rhotTry), np.copy(self.AfftTry)

    def endOfIteration(self):
        self.iter += 1
        if self.iter >= self.affBurnIn:
            self.typeRegression = self.typeRegressionSave
        _, self.xt, self.at = self.objectiveFunDef(self.a0, self.rhot, self.Afft)
        for k in range(self.nTarg):
            self.fvDef[k].updateVertices(self.xt[(k + 1) * self.Tsize1])

    def optimizeMatching(self):
        grd = self.getGradient(self.gradCoeff)
        [grd2] = self.dotProduct(grd, [grd])
        self.gradEps = max(0.001, np.sqrt(grd2) / 10000)
        self.cgBurnIn = self.affBurnIn
        cg(self, maxIter=self.maxIter, epsInit=0.1)
        return self.history
