"""Landmark flow, adjoint and geodesic shooting, drop-in for py-lddmm's pointEvolution
landmark functions (py-lddmm/pointEvolution.py:599-817) and its second-order (spline) flow,
Hamiltonian, covector and gradient (:882-1084).

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
from .kernelFunctions import (apply_ddiff_k11and12_dev, apply_diff_k_dev, apply_diff_kt_dev,
                              apply_divergence_dev, apply_k_dev)


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
    A = None if aff is None else aff[0].contiguous()
    bv = None if aff is None else aff[1].contiguous()
    if yt is None and nt is None and Jt is None:
        # all T steps in one sequence of fused launches (affine part in the step's epilogue)
        engine.shoot([engine.FlowJob(x0d, atd, K, xt, energy, A=A, b=bv)])
    else:
        xt[0] = x0d
        for k in range(T):
            z, a = xt[k], atd[k]
            engine.shoot([engine.FlowJob(z, a[None], K, xt[k:k + 2], energy[k:k + 1], dt=dt,
                                         A=None if A is None else A[k:k + 1],
                                         b=None if bv is None else bv[k:k + 1])])
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
    engine.covector([engine.CovectorJob(xt, atd, px1d, K, float(regweight), pxt,
                                        A=None if aff is None else aff[0].contiguous())])
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


# ---------------------------------------------------------------------------------------
# Second-order (spline) flow: x' = K(x) a, a' = -applyDiffKT(x,[a],[a]) + rho
# (pointEvolution.py:882-1084).  Host loops over time steps; every kernel sum is one launch.
# ---------------------------------------------------------------------------------------
def _affine_given(affine, T, dim, dev):
    """(A (T,d,d), b (T,d)) device tensors when affine is given (zeros included: the
    reference branches on presence, not value), else None."""
    if affine is None:
        return None
    A = rt.as_dev(affine[0] if isinstance(affine[0], torch.Tensor)
                  else np.asarray(affine[0], dtype=np.float64)).reshape(T, dim, dim)
    b = rt.as_dev(affine[1] if isinstance(affine[1], torch.Tensor)
                  else np.asarray(affine[1], dtype=np.float64)).reshape(T, dim)
    return A, b


def _second_order_velocity(x, a, rho, K):
    """(K(x) a, -applyDiffKT(x,[a],[a]) + rho) on device tensors."""
    zx = apply_k_dev(x, x, a, K)
    za = rho - apply_diff_kt_dev(x, x, a[None], a[None], K)
    return zx, za


def secondOrderEvolution(x0, a0, rhot, KparDiff, withJacobian=False, withPointSet=None,
                         affine=None):
    """Explicit Euler for x' = K(x) a, a' = -applyDiffKT(x,[a],[a]) + rho with T = rhot.shape[0]
    steps and the affine step U = I + dt A_k (pointEvolution.py:882-948).  Returns (xt, at) or
    [xt, at, zt?, Jt?]; the Jacobian is taken at the point set when one is given, else at the
    points (the reference's withJacobian without a point set reads undefined names)."""
    x0d, np_out = _in(x0)
    a0d, _ = _in(a0)
    rhod, _ = _in(rhot)
    T = rhod.shape[0]
    N, dim = x0d.shape
    dev = x0d.device
    dt = 1.0 / T
    aff = _affine_given(affine, T, dim, dev)
    xt = torch.empty((T + 1, N, dim), dtype=torch.float64, device=dev)
    at = torch.empty_like(xt)
    xt[0], at[0] = x0d, a0d
    zt = Jt = None
    if withPointSet is not None:
        z0, _ = _in(withPointSet)
        zt = torch.empty((T + 1,) + tuple(z0.shape), dtype=torch.float64, device=dev)
        zt[0] = z0
    if withJacobian:
        Jt = torch.zeros((T + 1, N if zt is None else zt.shape[1]), dtype=torch.float64, device=dev)
    eye = torch.eye(dim, dtype=torch.float64, device=dev)
    for k in range(T):
        x, a = xt[k], at[k]
        zx, za = _second_order_velocity(x, a, rhod[k], KparDiff)
        if aff is not None:
            U = eye + dt * aff[0][k]
            xt[k + 1] = (x + dt * zx) @ U.T + dt * aff[1][k]
            at[k + 1] = (a + dt * za) @ torch.linalg.inv(U)
        else:
            xt[k + 1] = x + dt * zx
            at[k + 1] = a + dt * za
        if zt is not None:
            z = zt[k]
            zz = apply_k_dev(z, x, a, KparDiff)
            zt[k + 1] = (z + dt * zz) @ U.T + dt * aff[1][k] if aff is not None else z + dt * zz
        if Jt is not None:
            Jt[k + 1] = Jt[k] + dt * _divergence(zt[k] if zt is not None else x, x, a, KparDiff)
            if aff is not None:
                Jt[k + 1] += dt * torch.trace(aff[0][k])
    if zt is None and Jt is None:
        return _out(xt, np_out), _out(at, np_out)
    out = [_out(xt, np_out), _out(at, np_out)]
    if zt is not None:
        out.append(_out(zt, np_out))
    if Jt is not None:
        out.append(_out(Jt, np_out))
    return out


def secondOrderHamiltonian(x, a, rho, px, pa, KparDiff, affine=None):
    """<px, K(x) a> + <pa, -applyDiffKT(x,[a],[a]) + rho> - |rho|^2/2 (+ affine terms)
    (pointEvolution.py:952-962); a float."""
    xd, ad, rd, pxd, pad = (_in(v)[0] for v in (x, a, rho, px, pa))
    zx, za = _second_order_velocity(xd, ad, rd, KparDiff)
    H = (pxd * zx).sum() + (pad * za).sum() - (rd ** 2).sum() / 2
    if affine is not None:
        A, b = (_in(v)[0] for v in affine)
        H = H + (pxd * (xd @ A.T + b)).sum() - (pad * (ad @ A)).sum()
    return float(H)


def secondOrderCovector(x0, a0, rhot, px1, pa1, KparDiff, affine=None, times=None):
    """Backward adjoint of secondOrderEvolution (pointEvolution.py:965-1037) with the
    time-series end points px1[j], pa1[j] added at steps (j+1) T / len(px1).  ``times`` is
    accepted and unused (the reference computes an unused value from it and, when it is
    None, raises while doing so).  Returns (pxt, pat, xt, at)."""
    x0d, np_out = _in(x0)
    a0d, _ = _in(a0)
    rhod, _ = _in(rhot)
    px1d = [_in(p)[0] for p in px1]
    pa1d = [_in(p)[0] for p in pa1]
    T = rhod.shape[0]
    nTarg = len(px1d)
    Tsize1 = T // nTarg
    N, dim = x0d.shape
    dev = x0d.device
    dt = 1.0 / T
    aff = _affine_given(affine, T, dim, dev)
    xt, at = secondOrderEvolution(x0d, a0d, rhod, KparDiff)
    pxt = torch.zeros((T + 1, N, dim), dtype=torch.float64, device=dev)
    pat = torch.zeros_like(pxt)
    pxt[T], pat[T] = px1d[nTarg - 1], pa1d[nTarg - 1]
    eye = torch.eye(dim, dtype=torch.float64, device=dev)
    for t in range(T):
        px, pa = pxt[T - t], pat[T - t]
        x, a = xt[T - t - 1], at[T - t - 1]
        if aff is not None:
            U = eye + dt * aff[0][T - t - 1]
            px_ = px @ U
            pa_ = pa @ torch.linalg.inv(U).T
        else:
            px_, pa_ = px, pa
        a1 = torch.stack([px_, a])
        a2 = torch.stack([a, px_])
        zpx = (apply_diff_kt_dev(x, x, a1, a2, KparDiff)
               - apply_ddiff_k11and12_dev(x, a, a, pa_.contiguous(), KparDiff))
        zpa = apply_k_dev(x, x, px_.contiguous(), KparDiff) - apply_diff_k_dev(x, pa_, a, KparDiff, 3)
        pxt[T - t - 1] = px_ + dt * zpx
        pat[T - t - 1] = pa_ + dt * zpa
        if t < T - 1 and (T - 1 - t) % Tsize1 == 0:
            pxt[T - t - 1] += px1d[(T - t - 1) // Tsize1 - 1]
            pat[T - t - 1] += pa1d[(T - t - 1) // Tsize1 - 1]
    return _out(pxt, np_out), _out(pat, np_out), _out(xt, np_out), _out(at, np_out)


def secondOrderGradient(x0, a0, rhot, px1, pa1, KparDiff, times=None, getCovector=False,
                        affine=None, controlWeight=1.0):
    """Gradient of 1/2 <a0, K(x0) a0> + sum_t dt cw |rho_t|^2/2 + data terms
    (pointEvolution.py:1040-1084): da0 = K(x0) a0 - pa_0, and, when ``affine`` is given (zeros
    included), drhot = cw rho_t - pa_{t+1} (in the affine frame) with dA, db; without affine
    drhot stays zero, as in the reference."""
    x0d, np_out = _in(x0)
    a0d, _ = _in(a0)
    rhod, _ = _in(rhot)
    pxt, pat, xt, at = secondOrderCovector(x0d, a0d, rhod, px1, pa1, KparDiff, affine=affine,
                                           times=times)
    T = rhod.shape[0]
    dim = x0d.shape[1]
    dt = 1.0 / T
    drhot = torch.zeros_like(rhod)
    aff = _affine_given(affine, T, dim, x0d.device)
    if aff is not None:
        eye = torch.eye(dim, dtype=torch.float64, device=x0d.device)
        dA = torch.zeros_like(aff[0])
        for k in range(T):
            x, a, rho = xt[k], at[k], rhod[k]
            px, pa = pxt[k + 1], pat[k + 1]
            zx, za = _second_order_velocity(x, a, rho, KparDiff)
            zx = x + dt * zx
            za = a + dt * za
            Ui = torch.linalg.inv(eye + dt * aff[0][k])
            pa = pa @ Ui.T
            za = za @ Ui
            dA[k] = px.T @ zx - za.T @ pa
            drhot[k] = rho * controlWeight - pa
        db = pxt[1:T + 1].sum(dim=1)
    da0 = apply_k_dev(x0d, x0d, a0d, KparDiff) - pat[0]
    o = lambda v: _out(v, np_out)
    if aff is None:
        base = (o(da0), o(drhot), o(xt), o(at))
    else:
        base = (o(da0), o(drhot), o(dA), o(db), o(xt), o(at))
    return base + (o(pxt), o(pat)) if getCovector else base
