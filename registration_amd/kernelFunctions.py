"""Kernel specification and kernel sums, drop-in for py-lddmm's kernelFunctions.Kernel.

Mirrors ``KernelSpec`` / ``Kernel`` of py-lddmm/kernelFunctions.py:238-530 (same constructor
arguments, same method names, argument roles and shapes).  The kernel sums run in the gfx950
library liblddmm_hip.so; there is no CPU path.

Arguments may be NumPy arrays (as in the reference; results come back as NumPy float64) or
torch tensors (results stay on the device as float64 tensors).

Extensions over the reference signature, all keyword-only with reference-compatible defaults:
``precision`` selects the pairwise arithmetic, ``'fp32c'`` (FP32 with Kahan-compensated
accumulation, FP64 totals; the default) or ``'fp64'``.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _lib
from . import runtime as rt

DEFAULT_PRECISION = os.environ.get("LDDMM_PRECISION", "fp32c")
_PREC = {"fp64": _lib.PREC_FP64, "fp32c": _lib.PREC_FP32C}


def precision_code(p: str) -> int:
    try:
        return _PREC[p]
    except KeyError:
        raise ValueError(f"unknown precision {p!r} (expected 'fp64' or 'fp32c')") from None


class KernelSpec:
    """Kernel specification (py-lddmm/kernelFunctions.py:238-280).

    name = 'gauss' or 'laplacian' (anything else -> 'none', a zero kernel); sigma: width;
    order: Laplacian order (forced to 10 for the Gaussian, orders above 4 clamp to 3);
    affine = 'none' | 'affine' | 'euclidean' kernel-level linear part with weights w1, w2
    and origin center.
    """

    def __init__(self, name='gauss', affine='none', sigma=6.5, order=10, w1=1.0, w2=1.0, dim=3,
                 center=(0, 0, 0), weight=1.0, *, precision=None):
        self.name = name
        self.sigma = sigma
        self.order = order
        self.weight = weight
        self.w1 = w1
        self.w2 = w2
        self.constant_plane = False
        self.center = np.array(center)
        self.affine_basis = []
        self.dim = dim
        self.precomp = []
        self._hold = False
        self._state = False
        self.affine = affine
        self.precision = precision or DEFAULT_PRECISION
        precision_code(self.precision)
        if name == 'laplacian':
            if self.order > 4:
                self.order = 3
            self.par = [sigma, self.order]
        elif name == 'gauss':
            self.order = 10
            self.par = [sigma]
        else:
            self.name = 'none'
            self.par = []
        if self.affine == 'euclidean':
            s2 = np.sqrt(2.0)
            if self.dim == 3:
                self.affine_basis.append(np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 0]]) / s2)
                self.affine_basis.append(np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]]) / s2)
                self.affine_basis.append(np.array([[0, 0, 0], [0, 0, 1], [0, -1, 0]]) / s2)
            elif self.dim == 2:
                self.affine_basis.append(np.array([[0, 1], [-1, 0]]) / s2)
            else:
                raise ValueError('euclidian kernels only available in dimension 2 or 3')

    # -- device-side description -------------------------------------------------------
    @property
    def kernel_code(self) -> int:
        return _lib.KERNEL_GAUSS if self.name == 'gauss' else _lib.KERNEL_LAPLACIAN

    @property
    def prec_code(self) -> int:
        return precision_code(self.precision)


def _wrap(x):
    """(device tensor, was_numpy) for a NumPy array or a tensor."""
    if isinstance(x, torch.Tensor):
        return rt.as_dev(x), False
    return rt.as_dev(np.asarray(x, dtype=np.float64)), True


def _unwrap(t: torch.Tensor, numpy_out: bool):
    return t.cpu().numpy() if numpy_out else t


def _pad_dim(*arrs):
    """Points of dimension 1 are padded with a zero coordinate (distances unchanged)."""
    out = []
    for a in arrs:
        if a is not None and a.shape[-1] == 1:
            a = torch.cat([a, torch.zeros_like(a)], dim=-1).contiguous()
        out.append(a)
    return out


class Kernel(KernelSpec):
    """Kernel sums (py-lddmm/kernelFunctions.py:284-530) on the GPU."""

    def precompute(self, x, firstVar=None, grid=None, diff=False, diff2=False):
        """The dense kernel matrix K(y_k, x_l) (or K'(r^2) with diff, K''(r^2) with diff2) times
        ``weight``, y = firstVar, grid points (..., d) or x (kernelFunctions.py:285-298 with
        kernelMatrixGauss / kernelMatrixLaplacian, :10-55,152-198): exact-double polynomials,
        the reference's diagonal fills when y is x.  None for a kernel named 'none'.  The
        reference's hold() reuses the previous call's distances whatever x is; here every call
        recomputes, which is the same for the same x."""
        if self.name == 'none':
            return None
        xd, np_out = _wrap(x)
        lead = None
        if grid is not None:
            gd = _wrap(grid)[0]
            lead = tuple(gd.shape[:-1])
            yd = gd.reshape(-1, gd.shape[-1]).contiguous()
        else:
            yd = xd if firstVar is None else _wrap(firstVar)[0]
        self_set = firstVar is None and grid is None
        deriv = 1 if diff else (2 if diff2 else 0)
        out = torch.empty((yd.shape[0], xd.shape[0]), dtype=torch.float64, device=xd.device)
        lib = _lib.load()
        _lib.check(lib.lddmm_kernel_matrix(yd.data_ptr(), xd.data_ptr(), yd.shape[0], xd.shape[0],
                                           xd.shape[1], self.kernel_code, int(self.order),
                                           float(self.sigma), deriv, int(self_set), out.data_ptr(),
                                           rt.stream_handle()), "lddmm_kernel_matrix")
        if self.weight != 1.0:
            out = out * self.weight
        if lead is not None:
            out = out.reshape(lead + (xd.shape[0],))
        return _unwrap(out, np_out)

    def hold(self):
        self._state = self._hold
        self._hold = True

    def release(self):
        self._state = False
        self._hold = False

    def reset(self):
        self._hold = self._state

    # ---------------------------------------------------------------------------------
    def applyK(self, x, a, firstVar=None, grid=None, matrixWeights=False):
        """K(x,x)a or K(firstVar,x)a (kernelFunctions.py:310-369 -> kernelFunctions.f90:1-54);
        with ``matrixWeights`` a is (n_y, n_x, m), one weight per pair (applykmat,
        kernelFunctions.f90:499-550)."""
        if matrixWeights:
            return self._apply_k_mat(x, a, firstVar)
        xd, np_out = _wrap(x)
        ad, _ = _wrap(a)
        yd = xd if firstVar is None else _wrap(firstVar)[0]
        if ad.dim() == 1:
            ad = ad[:, None]
        ny, m = yd.shape[0], ad.shape[1]
        if xd.shape[0] != ad.shape[0] or xd.shape[1] != yd.shape[1]:
            raise ValueError("applyK: shape mismatch between x, a and firstVar")
        if self.name == 'none':
            z = torch.zeros((ny, m), dtype=torch.float64, device=xd.device)
        else:
            z = apply_k_dev(yd, xd, ad, self)
        if self.affine == 'affine':
            c = torch.as_tensor(self.center, dtype=torch.float64, device=xd.device)
            xx = xd - c
            yy = xx if firstVar is None else yd - c
            if grid is not None:
                yy = _wrap(grid)[0]
            z = z + self.w1 * (yy @ (xx.T @ ad)) + self.w2 * ad.sum(dim=0)
        elif self.affine == 'euclidean':
            c = torch.as_tensor(self.center, dtype=torch.float64, device=xd.device)
            xx = xd - c
            yy = None if firstVar is None else yd - c
            gg = None if grid is None else _wrap(grid)[0] - c
            z = z + self.w2 * ad.sum(dim=0)
            for E in self.affine_basis:
                Et = torch.as_tensor(E, dtype=torch.float64, device=xd.device)
                xE = xx @ Et.T
                s = (xE * ad).sum()
                if firstVar is None:
                    z = z + self.w1 * s * (xE if gg is None else gg @ Et.T)
                else:
                    z = z + self.w1 * s * (yy @ Et.T)
        return _unwrap(z, np_out)

    def _apply_k_mat(self, x, beta, firstVar):
        xd, np_out = _wrap(x)
        bd, _ = _wrap(beta)
        yd = xd if firstVar is None else _wrap(firstVar)[0]
        if bd.dim() != 3 or bd.shape[0] != yd.shape[0] or bd.shape[1] != xd.shape[0]:
            raise ValueError("applyK(matrixWeights=True): weights must be (n_y, n_x, m)")
        if self.affine != 'none':
            raise NotImplementedError("kernel-level affine terms with matrix weights are not "
                                      "defined by the reference (its affine branch expects (n, m) "
                                      "weights)")
        if self.name == 'none':
            return _unwrap(torch.zeros((yd.shape[0], bd.shape[2]), dtype=torch.float64,
                                       device=xd.device), np_out)
        return _unwrap(apply_k_mat_dev(yd, xd, bd, self), np_out)

    def applyDiffKmat(self, x, beta, firstVar=None):
        """sum_l 2K'(r^2) (y_k - x_l) beta[k, l] at y = firstVar or x (kernelFunctions.py:405-411
        -> kernelFunctions.f90:552-610)."""
        xd, np_out = _wrap(x)
        bd, _ = _wrap(beta)
        yd = xd if firstVar is None else _wrap(firstVar)[0]
        if bd.dim() != 2 or bd.shape[0] != yd.shape[0] or bd.shape[1] != xd.shape[0]:
            raise ValueError("applyDiffKmat: beta must be (n_y, n_x)")
        return _unwrap(apply_diff_k_mat_dev(yd, xd, bd, self), np_out)

    def applyDiffKT(self, x, a1, a2, firstVar=None):
        """sum_l 2 K'(|y_k-x_l|^2) (y_k - x_l) sum_j <a1[j,k], a2[j,l]>
        (kernelFunctions.py:413-449 -> kernelFunctions.f90:56-114)."""
        xd, np_out = _wrap(x)
        a1d, _ = _wrap(a1)
        a2d, _ = _wrap(a2)
        yd = xd if firstVar is None else _wrap(firstVar)[0]
        if a1d.dim() == 2:
            a1d = a1d[None]
        if a2d.dim() == 2:
            a2d = a2d[None]
        if self.name == 'none':
            zpx = torch.zeros_like(yd)
        else:
            zpx = apply_diff_kt_dev(yd, xd, a1d, a2d, self)
        if self.affine == 'affine':
            raise NotImplementedError(
                "the reference's affine branch of applyDiffKT reads an undefined name "
                "(kernelFunctions.py:437)")
        elif self.affine == 'euclidean':
            c = torch.as_tensor(self.center, dtype=torch.float64, device=xd.device)
            xx = xd - c
            for E in self.affine_basis:
                Et = torch.as_tensor(E, dtype=torch.float64, device=xd.device)
                yy = xx @ Et.T
                for k in range(a1d.shape[0]):
                    bb = a1d[k] @ Et
                    zpx = zpx + self.w1 * (yy * a2d[k]).sum() * bb
        return _unwrap(zpx, np_out)

    def applyDivergence(self, x, a, firstVar=None):
        """sum_l div_1(K(y_k, x_l) a_l) = sum_l 2K'(r^2) <a_l, y_k - x_l> at y = firstVar or x
        (kernelFunctions.py:510-530), shape (n_y,).  The reference builds the dense (n_y, n_x)
        K' matrix (precompute(diff=True)); here it is a tiled pair sum with the same exact-double
        Laplacian polynomials (kernelFunctions.py:75-88), scaled by ``weight`` as the dense
        matrices are (:298).  A kernel named 'none' returns zeros of shape (1, n_y), as the
        reference's transposed (n_y, 1) array."""
        xd, np_out = _wrap(x)
        ad, _ = _wrap(a)
        yd = xd if firstVar is None else _wrap(firstVar)[0]
        if xd.shape != ad.shape or xd.shape[1] != yd.shape[1]:
            raise ValueError("applyDivergence: shape mismatch between x, a and firstVar")
        if self.name == 'none':
            return _unwrap(torch.zeros((1, yd.shape[0]), dtype=torch.float64, device=xd.device),
                           np_out)
        z = apply_divergence_dev(yd, xd, ad, self)
        if self.weight != 1.0:
            z = z * self.weight
        if self.affine == 'affine':
            c = torch.as_tensor(self.center, dtype=torch.float64, device=xd.device)
            z = z + self.w1 * ((yd - c) * ad).sum(dim=1)
        return _unwrap(z, np_out)

    def applyDiffK(self, x, a1, a2):
        """sum_l 2K'(r^2) <x_k - x_l, a1_k> a2_l (kernelFunctions.py:387-389 ->
        kernelFunctions.f90:116-173)."""
        return self._diff_k(x, a1, a2, 1)

    def applyDiffK2(self, x, a1, a2):
        """-sum_l 2K'(r^2) <x_k - x_l, a1_l> a2_l (kernelFunctions.py:392-394 ->
        kernelFunctions.f90:175-231)."""
        return self._diff_k(x, a1, a2, 2)

    def applyDiffK1and2(self, x, a1, a2):
        """sum_l 2K'(r^2) <x_k - x_l, a1_k - a1_l> a2_l (kernelFunctions.py:396-398 ->
        kernelFunctions.f90:234-290)."""
        return self._diff_k(x, a1, a2, 3)

    def _diff_k(self, x, a1, a2, form):
        xd, np_out = _wrap(x)
        a1d, a2d = _wrap(a1)[0], _wrap(a2)[0]
        if a1d.shape != xd.shape or a2d.shape != xd.shape:
            raise ValueError("applyDiffK: x, a1 and a2 must have one shape (n, d)")
        if self.name == 'none':
            return _unwrap(torch.zeros_like(xd), np_out)
        return _unwrap(apply_diff_k_dev(xd, a1d, a2d, self, form), np_out)

    def applyDDiffK11(self, x, n, a, p):
        """sum_l 2<n_k, a_l> (2K''<dx, p_k> dx + K' p_k), dx = x_k - x_l: the derivative of
        applyDiffKT(x, [n], [a]) along p at the targets (kernelFunctions.py:459-478).  The
        reference's dense K', K'' matrices (exact-double Laplacian polynomials, scaled by
        ``weight``) become a tiled pair sum."""
        return self._ddiff_dense(x, n, a, p, 1)

    def applyDDiffK12(self, x, n, a, p):
        """-sum_l 2<n_k, a_l> (2K''<dx, p_l> dx + K' p_l): the same derivative at the sources
        (kernelFunctions.py:481-506), plus w1 (n a^T) p for a kernel-level affine part."""
        z = self._ddiff_dense(x, n, a, p, 2)
        if self.affine == 'affine':
            xd, np_out = _wrap(x)
            nd, ad, pd = (_wrap(v)[0] for v in (n, a, p))
            zd = _wrap(z)[0] + self.w1 * (nd @ (ad.T @ pd))
            return _unwrap(zd, np_out)
        if self.affine == 'euclidean':
            raise NotImplementedError(
                "the reference's euclidean branch of applyDDiffK12 multiplies an (n, d) array by "
                "a (d, d) basis element (kernelFunctions.py:503-505), defined only when n = d")
        return z

    def _ddiff_dense(self, x, n, a, p, form):
        xd, np_out = _wrap(x)
        nd, ad, pd = (_wrap(v)[0] for v in (n, a, p))
        if self.name == 'none':
            return _unwrap(torch.zeros_like(xd), np_out)
        z = apply_ddiff_k11and12_dev(xd, nd, ad, pd, self, form=form)
        if self.weight != 1.0:
            z = z * self.weight
        return _unwrap(z, np_out)

    def applyDDiffK11and12(self, x, n, a, p):
        """kernelFunctions.py:452-454 -> kernelFunctions.f90:430-496."""
        xd, np_out = _wrap(x)
        nd, ad, pd = (_wrap(v)[0] for v in (n, a, p))
        if self.name == 'none':
            return _unwrap(torch.zeros_like(xd), np_out)
        return _unwrap(apply_ddiff_k11and12_dev(xd, nd, ad, pd, self), np_out)


# --------------------------------------------------------------------------------------
# Device-tensor forms (float64 contiguous device tensors in, device tensor out).
# --------------------------------------------------------------------------------------
def _check_dim(d: int) -> None:
    if d < 1 or d > 3:
        raise ValueError(f"point dimension {d} is not supported by the GPU kernels (1..3)")


def apply_k_dev(y: torch.Tensor, x: torch.Tensor, w: torch.Tensor, K: KernelSpec,
                out: torch.Tensor | None = None) -> torch.Tensor:
    lib = _lib.load()
    _check_dim(x.shape[1])
    y, x = _pad_dim(y, x)
    w = w.contiguous()
    ny, nx, d, m = y.shape[0], x.shape[0], x.shape[1], w.shape[1]
    if out is None:
        out = torch.empty((ny, m), dtype=torch.float64, device=x.device)
    nbytes = lib.lddmm_apply_k_workspace_bytes(ny, nx, d, m)
    ws = rt.workspace(nbytes)
    _lib.check(lib.lddmm_apply_k(y.data_ptr(), x.data_ptr(), w.data_ptr(), ny, nx, d, m,
                                 K.kernel_code, int(K.order), float(K.sigma), K.prec_code,
                                 out.data_ptr(), ws.data_ptr(), ws.numel(), rt.stream_handle()),
               "lddmm_apply_k")
    return out


def apply_diff_kt_dev(y, x, a1, a2, K: KernelSpec) -> torch.Tensor:
    lib = _lib.load()
    d0 = x.shape[1]
    _check_dim(d0)
    y, x = _pad_dim(y, x)
    a1 = a1.contiguous()
    a2 = a2.contiguous()
    na, ny, m = a1.shape
    nx, d = x.shape
    if a2.shape[0] != na or a2.shape[1] != nx or a2.shape[2] != m or a1.shape[1] != y.shape[0]:
        raise ValueError("applyDiffKT: shape mismatch between x, a1, a2 and firstVar")
    out = torch.zeros((ny, d), dtype=torch.float64, device=x.device)
    if m != 1 and m != d:
        raise ValueError(f"applyDiffKT: weight width {m} must be 1 or the dimension {d}")
    # The library takes up to 3 weight pairs per call; the sum is linear in the pairs.
    for j0 in range(0, na, 3):
        j1 = min(na, j0 + 3)
        part = torch.empty((ny, d), dtype=torch.float64, device=x.device)
        nbytes = lib.lddmm_apply_diff_kt_workspace_bytes(ny, nx, d, m, j1 - j0)
        ws = rt.workspace(nbytes)
        _lib.check(lib.lddmm_apply_diff_kt(y.data_ptr(), x.data_ptr(),
                                           a1[j0:j1].contiguous().data_ptr(),
                                           a2[j0:j1].contiguous().data_ptr(), ny, nx, d, m,
                                           j1 - j0, K.kernel_code, int(K.order), float(K.sigma),
                                           K.prec_code, part.data_ptr(), ws.data_ptr(),
                                           ws.numel(), rt.stream_handle()),
                   "lddmm_apply_diff_kt")
        out = part if j0 == 0 else out + part
    return out[:, :d0].contiguous()


def apply_ddiff_k11and12_dev(x, a1, a2, p, K: KernelSpec, form: int = 0) -> torch.Tensor:
    """form 0: applyDDiffK11and12 (one-call entry point); 1 / 2: the dense siblings
    applyDDiffK11 / applyDDiffK12 (segmented launch with seg.na = form)."""
    lib = _lib.load()
    d0 = x.shape[1]
    _check_dim(d0)
    x, a1, a2, p = _pad_dim(x, a1, a2, p)
    n, d = x.shape
    out = torch.empty((n, d), dtype=torch.float64, device=x.device)
    if form:
        a1, a2, p = a1.contiguous(), a2.contiguous(), p.contiguous()
        r = rt.seg(y=x, w0=a1, w1=a2, w2=p, out0=out, ny=n, sigma=float(K.sigma),
                   kernel=K.kernel_code, order=int(K.order), na=form)
        rt.launch(_lib.OP_DDIFF_K11_12, K.prec_code, d, r)
        return out[:, :d0].contiguous()
    nbytes = lib.lddmm_apply_ddiff_k11and12_workspace_bytes(n, d)
    ws = rt.workspace(nbytes)
    _lib.check(lib.lddmm_apply_ddiff_k11and12(x.data_ptr(), a1.data_ptr(), a2.data_ptr(),
                                              p.data_ptr(), n, d, K.kernel_code, int(K.order),
                                              float(K.sigma), K.prec_code, out.data_ptr(),
                                              ws.data_ptr(), ws.numel(), rt.stream_handle()),
               "lddmm_apply_ddiff_k11and12")
    return out[:, :d0].contiguous()


def apply_divergence_dev(y, x, a, K: KernelSpec) -> torch.Tensor:
    """(n_y,) divergence sum of Kernel.applyDivergence on device tensors."""
    d0 = x.shape[1]
    _check_dim(d0)
    y, x, a = _pad_dim(y, x, a)
    a = a.contiguous()
    out = torch.empty(y.shape[0], dtype=torch.float64, device=x.device)
    r = rt.seg(y=y, x0=x, w0=a, out0=out, ny=y.shape[0], nx0=x.shape[0], sigma=float(K.sigma),
               kernel=K.kernel_code, order=int(K.order))
    rt.launch(_lib.OP_DIVERGENCE, K.prec_code, y.shape[1], r)
    return out


def apply_diff_k_dev(x, a1, a2, K: KernelSpec, form: int) -> torch.Tensor:
    """applyDiffK (form 1), applyDiffK2 (2), applyDiffK1and2 (3) on device tensors (n, d)."""
    d0 = x.shape[1]
    _check_dim(d0)
    x, a1, a2 = _pad_dim(x, a1, a2)
    a1, a2 = a1.contiguous(), a2.contiguous()
    n, d = x.shape
    out = torch.empty((n, d), dtype=torch.float64, device=x.device)
    r = rt.seg(y=x, w0=a1, w1=a2, out0=out, ny=n, sigma=float(K.sigma), kernel=K.kernel_code,
               order=int(K.order), na=form)
    rt.launch(_lib.OP_DIFF_K, K.prec_code, d, r)
    return out[:, :d0].contiguous()


def apply_k_mat_dev(y, x, beta, K: KernelSpec) -> torch.Tensor:
    """applyK with matrix weights beta (n_y, n_x, m) on device tensors; FP64 in both precisions
    (the sum is bound by reading beta)."""
    lib = _lib.load()
    _check_dim(x.shape[1])
    y, x = _pad_dim(y, x)
    beta = beta.contiguous()
    ny, nx, m = beta.shape
    if m > 3:
        return torch.cat([apply_k_mat_dev(y, x, beta[:, :, j:j + 3], K) for j in range(0, m, 3)],
                         dim=1)
    out = torch.empty((ny, m), dtype=torch.float64, device=x.device)
    _lib.check(lib.lddmm_apply_k_mat(y.data_ptr(), x.data_ptr(), beta.data_ptr(), ny, nx,
                                     x.shape[1], m, K.kernel_code, int(K.order), float(K.sigma),
                                     out.data_ptr(), rt.stream_handle()), "lddmm_apply_k_mat")
    return out


def apply_diff_k_mat_dev(y, x, beta, K: KernelSpec) -> torch.Tensor:
    """applyDiffKmat on device tensors (beta (n_y, n_x)); FP64."""
    lib = _lib.load()
    d0 = x.shape[1]
    _check_dim(d0)
    y, x = _pad_dim(y, x)
    beta = beta.contiguous()
    ny, nx = beta.shape
    out = torch.empty((ny, x.shape[1]), dtype=torch.float64, device=x.device)
    _lib.check(lib.lddmm_apply_diff_k_mat(y.data_ptr(), x.data_ptr(), beta.data_ptr(), ny, nx,
                                          x.shape[1], K.kernel_code, int(K.order),
                                          float(K.sigma), out.data_ptr(), rt.stream_handle()),
               "lddmm_apply_diff_k_mat")
    return out[:, :d0].contiguous()
