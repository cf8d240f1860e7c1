"""Landmark flow, adjoint and geodesic shooting, drop-in for py-lddmm's pointEvolution
landmark functions (py-lddmm/pointEvolution.py:599-817).

Same signatures and return conventions as the reference.  Inputs may be NumPy arrays (results
come back as NumPy float64) or torch tensors (results stay on the GPU).  Each Euler step is one
fused launch (kernel sum + update); the backward covector step is one fused launch of the
covector-weighted derivative sum.
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from . import runtime as rt
from .kernelFunctions import apply_diff_kt_dev, apply_divergence_dev, apply_k_dev


def _in(a):
    if isinstance(a, torch.Tensor):
        return rt.as_dev(a), False
    return rt.as_dev(np.asarray(a, dtype=np.float64)), True


def _out(t, numpy_out):
    return t.cpu().numpy() if numpy_out else t


def _affine_dev(affine, T, dim, dev):
    """(A (T,d,d), b (T,d)) on the device, or None when absent or identically zero."""
    if affine is None:
        return None
    A = torch.as_tensor(np.asarray(affine[0], dtype=np.float64) if not isinstance(affine[0], torch.Tensor)
                        else affine[0], dtype=torch.float64, device=dev)
    b = torch.as_tensor(np.asarray(affine[1], dtype=np.float64) if not isinstance(affine[1], torch.Tensor)
                        else affine[1], dtype=torch.float64, device=dev)
    if not bool(A.any()) and not bool(b.any()):
        return None
    return A.reshape(T, dim, dim), b.reshape(T, dim)


def landmarkDirectEvolutionEuler(x0, at, KparDiff, affine=None, withJacobian=False,
                                 withNormals=None, withPointSet=None):
    """Solves dx/dt = K(x,x) a(t) + A(t) x + b(t), x(0) = x0 by explicit Euler with T = at.shape[0]
    steps (pointEvolution.py:666-730).  Returns xt (T+1, N, d), or [xt, yt?, nt?, Jt?] with the
    optional point set / normals trajectories and, with ``withJacobian``, the log-Jacobian
    J_{t+1} = J_t + dt (div v_t + tr A_t), J_0 = 0, at the points (or at the point set when
    one is given; shape (T+1, n))."""
    x0d, np_out = _in(x0)
    atd, _ = _in(at)
    T, N, dim = atd.shape
    dev = x0d.device
    K = engine.DevKernel.of(KparDiff)
    aff = _affine_dev(affine, T, dim, dev)
    xt = torch.empty((T + 1, N, dim), dtype=torch.float64, device=dev)
    energy = torch.empty(T, dtype=torch.float64, device=dev)
    yt = nt = None
    if withPointSet is not None:
        y0, _ = _in(withPointSet)
        yt = torch.empty((T + 1,) + tuple(y0.shape), dtype=torch.float64, device=dev)
        yt[0] = y0
    if withNormals is not None:
        n0, _ = _in(withNormals)
        nt = torch.empty((T + 1, N, dim), dtype=torch.float64, device=dev)
        nt[0] = n0
    Jt = None
    if withJacobian:
        nj = yt.shape[1] if yt is not None else N
        Jt = torch.zeros((T + 1, nj), dtype=torch.float64, device=dev)
    dt = 1.0 / T
    if aff is None and yt is None and nt is None and Jt is None:
        engine.shoot([engine.FlowJob(x0d, atd, K, xt, energy)])
    else:
        xt[0] = x0d
        for k in range(T):
            z, a = xt[k], atd[k]
            engine.shoot([engine.FlowJob(z, a[None], K, xt[k:k + 2], energy[k:k + 1], dt=dt)])
            if aff is not None:
                xt[k + 1] += dt * (z @ aff[0][k].T + aff[1][k])
            if yt is not None:
                zy = yt[k]
                yt[k + 1] = zy + dt * apply_k_dev(zy, z, a, KparDiff)
                if aff is not None:
                    yt[k + 1] += dt * (zy @ aff[0][k].T + aff[1][k])
            if Jt is not None:
                # the divergence at the point set when there is one, else at the points
                at_pts = yt[k] if yt is not None else z
                Jt[k + 1] = Jt[k] + dt * _divergence(at_pts, z, a, KparDiff)
                if aff is not None:
                    Jt[k + 1] += dt * torch.trace(aff[0][k])
            if nt is not None:
                zn = nt[k]
                nt[k + 1] = zn - dt * apply_diff_kt_dev(z, z, zn[None], a[None], KparDiff)
                if aff is not None:
                    nt[k + 1] += dt * (zn @ aff[0][k])
    if yt is None and nt is None and Jt is None:
        return _out(xt, np_out)
    out = [_out(xt, np_out)]
    if yt is not None:
        out.append(_out(yt, np_out))
    if nt is not None:
        out.append(_out(nt, np_out))
    if Jt is not None:
        out.append(_out(Jt, np_out))
    return out


def _divergence(y, x, a, K):
    """Kernel.applyDivergence on device tensors, including a kernel-level affine term."""
    z = apply_divergence_dev(y, x, a, K)
    if getattr(K, 'weight', 1.0) != 1.0:
        z = z * K.weight
    if getattr(K, 'affine', 'none') == 'affine':
        c = torch.as_tensor(K.center, dtype=torch.float64, device=x.device)
        z = z + K.w1 * ((y - c) * a).sum(dim=1)
    return z


def landmarkHamiltonianCovector(x0, at, px1, KparDiff, regweight, affine=None):
    """Discrete adjoint of the Euler flow for the energy sum_t dt lambda <a_t, K a_t>
    (pointEvolution.py:735-765).  Returns (pxt (T+1,N,d), xt (T+1,N,d))."""
    x0d, np_out = _in(x0)
    atd, _ = _in(at)
    px1d, _ = _in(px1)
    T, N, dim = atd.shape
    dev = x0d.device
    K = engine.DevKernel.of(KparDiff)
    xt = landmarkDirectEvolutionEuler(x0d, atd, KparDiff, affine=affine)
    pxt = torch.empty((T + 1, N, dim), dtype=torch.float64, device=dev)
    aff = _affine_dev(affine, T, dim, dev)
    if aff is None:
        engine.covector([engine.CovectorJob(xt, atd, px1d, K, float(regweight), pxt)])
    else:
        pxt[T] = px1d
        dt = 1.0 / T
        for t in range(T):
            px = pxt[T - t]
            engine.covector([engine.CovectorJob(xt[T - t - 1:T - t + 1], atd[T - t - 1:T - t], px,
                                                K, float(regweight), pxt[T - t - 1:T - t + 1],
                                                dt=dt)])
            pxt[T - t - 1] += dt * (px @ aff[0][T - t - 1])
    return _out(pxt, np_out), _out(xt, np_out)


def landmarkHamiltonianGradient(x0, at, px1, KparDiff, regweight, getCovector=False, affine=None):
    """dat[k] = 2 lambda a_k - pxt[k+1] (+ affine dA, db) (pointEvolution.py:793-817)."""
    pxt, xt = landmarkHamiltonianCovector(x0, at, px1, KparDiff, regweight, affine=affine)
    np_out = not isinstance(pxt, torch.Tensor)
    atd, _ = _in(at)
    pxd = rt.as_dev(pxt) if np_out else pxt
    dat = 2 * regweight * atd - pxd[1:]
    if affine is None:
        if getCovector:
            return _out(dat, np_out), xt, pxt
        return _out(dat, np_out), xt
    xtd = rt.as_dev(xt) if np_out else xt
    dA = torch.einsum('tni,tnj->tij', pxd[1:], xtd[:-1])
    db = pxd[1:].sum(dim=1)
    if getCovector:
        return _out(dat, np_out), _out(dA, np_out), _out(db, np_out), xt, pxt
    return _out(dat, np_out), _out(dA, np_out), _out(db, np_out), xt


def landmarkEPDiff(T, x0, a0, KparDiff, affine=None, withJacobian=False, withNormals=None,
                   withPointSet=None):
    """Geodesic shooting from (x0, a0) (pointEvolution.py:599-658).  Returns (xt, at)."""
    if withJacobian or withNormals is not None or withPointSet is not None or affine is not None:
        raise NotImplementedError("landmarkEPDiff options beyond (xt, at) are not part of the GPU path")
    x0d, np_out = _in(x0)
    a0d, _ = _in(a0)
    N, dim = x0d.shape
    dev = x0d.device
    K = engine.DevKernel.of(KparDiff)
    dt = 1.0 / T
    xt = torch.empty((T + 1, N, dim), dtype=torch.float64, device=dev)
    at = torch.zeros((T, N, dim), dtype=torch.float64, device=dev)
    xt[0] = x0d
    at[0] = a0d
    e = torch.empty(1, dtype=torch.float64, device=dev)
    for k in range(T):
        z, a = xt[k], at[k]
        engine.shoot([engine.FlowJob(z, a[None], K, xt[k:k + 2], e, dt=dt)])
        if k < T - 1:
            at[k + 1] = a - dt * apply_diff_kt_dev(z, z, a[None], a[None], KparDiff)
    return _out(xt, np_out), _out(at, np_out)
