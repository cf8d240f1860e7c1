"""Differentiable landmark shooting: a torch.autograd.Function whose backward is the
hand-written discrete adjoint (the covector kernel), with trajectory checkpoints.

Forward:  x_{t+1} = x_t + dt K(x_t) a_t,  t < T,  dt = 1/T  (pointEvolution.py:694-697)
          E = lambda dt sum_t <a_t, K(x_t) a_t>             (surfaceMatching.py:206-212)
Outputs:  (x_T, E).

Backward, for upstream gradients (g_x, g_E): L = g_E E + <g_x, x_T> has covector
p_T = -g_x and p_t = p_{t+1} + dt applyDiffKT(x_t, [p, a, -2 lambda' a], [a, p, a]) with
lambda' = g_E lambda (landmarkHamiltonianCovector, pointEvolution.py:735-765); then
dL/da_t = dt K(x_t) (2 lambda' a_t - p_{t+1})  (landmarkHamiltonianGradient :803 mapped from
the K-metric to the Euclidean gradient) and dL/dx_0 = -p_0.

Only every ``checkpoint_every``-th state is kept from the forward pass; the backward pass
recomputes each segment's states from its checkpoint (the recomputation is deterministic and
bit-identical to the forward).  ``checkpoint_every=None`` picks the interval from the free
device memory.
"""
from __future__ import annotations

import math

import torch

from . import _lib
from . import engine
from . import runtime as rt


def auto_checkpoint(T: int, N: int, d: int, budget_bytes: float | None = None) -> int:
    """Keep every state when (T+1) states fit in half the free memory, else ~sqrt(T) spacing."""
    state = N * d * 8
    if budget_bytes is None:
        free, _ = torch.cuda.mem_get_info()
        budget_bytes = 0.5 * free
    if (T + 1) * state <= budget_bytes:
        return 1
    return max(1, int(math.ceil(math.sqrt(T))))


def _flow(x0, at, K, dt, xt_out, energy_out):
    engine.shoot([engine.FlowJob(x0, at, K, xt_out, energy_out, dt=dt)])


def _apply_k_batch(xs, ws, K, outs):
    """outs[i] = K(xs[i], xs[i]) ws[i] for a list of time slices, one segmented launch."""
    segs = [rt.seg(y=x, x0=x, w0=w, out0=o, ny=x.shape[0], nx0=x.shape[0], sigma=K.sigma,
                   kernel=K.kernel, order=K.order, m=x.shape[1]) for x, w, o in zip(xs, ws, outs)]
    rt.launch(_lib.OP_APPLY_K, K.prec, xs[0].shape[1], segs)


class Shoot(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, at, kernel_obj, regweight, checkpoint_every):
        K = engine.DevKernel.of(kernel_obj)
        x0 = x0.detach().contiguous().to(torch.float64)
        at = at.detach().contiguous().to(torch.float64)
        T, N, d = at.shape
        dt = 1.0 / T
        k = checkpoint_every or auto_checkpoint(T, N, d)
        dev = x0.device
        energy = torch.empty(T, dtype=torch.float64, device=dev)
        ckpts = {0: x0}
        x = x0
        t = 0
        while t < T:
            n = min(k, T - t)
            seg = torch.empty((n + 1, N, d), dtype=torch.float64, device=dev)
            _flow(x, at[t:t + n], K, dt, seg, energy[t:t + n])
            x = seg[n].clone()
            t += n
            if t < T:
                ckpts[t] = x
        E = regweight * dt * energy.sum()
        ctx.K, ctx.k, ctx.dt, ctx.lam = K, k, dt, float(regweight)
        ctx.ckpt_times = sorted(ckpts)
        ctx.save_for_backward(at, *[ckpts[s] for s in ctx.ckpt_times])
        return x, E

    @staticmethod
    def backward(ctx, g_x, g_E):
        at, *ck = ctx.saved_tensors
        ckpts = dict(zip(ctx.ckpt_times, ck))
        K, k, dt = ctx.K, ctx.k, ctx.dt
        T, N, d = at.shape
        dev = at.device
        gE = 0.0 if g_E is None else float(g_E)
        lam = ctx.lam * gE
        p = (-g_x if g_x is not None else torch.zeros((N, d), dtype=torch.float64, device=dev))
        p = p.contiguous().to(torch.float64)
        g_at = torch.empty_like(at)
        scratch_e = torch.empty(max(k, 1), dtype=torch.float64, device=dev)
        for t0 in sorted(ckpts, reverse=True):
            n = min(k, T - t0)
            seg = torch.empty((n + 1, N, d), dtype=torch.float64, device=dev)
            _flow(ckpts[t0], at[t0:t0 + n], K, dt, seg, scratch_e[:n])
            pseg = torch.empty((n + 1, N, d), dtype=torch.float64, device=dev)
            engine.covector([engine.CovectorJob(seg, at[t0:t0 + n], p, K, lam, pseg, dt=dt)])
            # K-metric gradient 2 lam a_t - p_{t+1}, mapped to the Euclidean one.
            dat = 2 * lam * at[t0:t0 + n] - pseg[1:]
            _apply_k_batch([seg[s] for s in range(n)], [dat[s] for s in range(n)], K,
                           [g_at[t0 + s] for s in range(n)])
            p = pseg[0].clone()
        g_at.mul_(dt)
        g_x0 = -p
        return g_x0, g_at, None, None, None


def shoot(x0, at, kernel_obj, regweight=1.0, checkpoint_every=None):
    """Differentiable (x_T, E) of the Euler flow; see module docstring."""
    return Shoot.apply(x0, at, kernel_obj, regweight, checkpoint_every)
