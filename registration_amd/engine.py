"""Device-resident operations of the LDDMM hot path, batched over independent problems.

Every function takes a list of jobs and issues one segmented launch per time step (or one
launch in total) covering all of them, so a single registration is simply a batch of one and
a served batch shares the same kernels.  Results of a job never depend on the other jobs of
its batch: each segment keeps its own chunking and reduction order.

Reference counterparts:
  shoot      landmarkDirectEvolutionEuler + the energy loop of objectiveFunDef
             (py-lddmm/pointEvolution.py:666-730, surfaceMatching.py:206-212)
  covector   landmarkHamiltonianCovector   (pointEvolution.py:735-765)
  kdot       SurfaceMatching.dotProduct    (surfaceMatching.py:353-377)
  currents   currentNormDef / currentNormGradient (surfaces.py:1020-1089, curves.py:548-613)
  knorm      currentNorm0                  (surfaces.py:1009-1016, curves.py:539-543)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import runtime as rt


@dataclass(frozen=True)
class DevKernel:
    """What a launch needs to know about a Kernel object."""

    kernel: int
    order: int
    sigma: float
    prec: int

    @staticmethod
    def of(K) -> "DevKernel":
        if K.name == 'none':
            raise ValueError("kernel 'none' has no device form")
        if getattr(K, 'affine', 'none') != 'none':
            raise NotImplementedError("kernel-level affine terms are not fused into the device path")
        return DevKernel(K.kernel_code, int(K.order), float(K.sigma), K.prec_code)

    @property
    def family(self):
        return (self.kernel, self.prec)

    def energy(self) -> "DevKernel":
        """The form used for data-term energies: always FP64 pairwise arithmetic.  A currents
        norm is a small difference of large sums (|S|^2 - 2<S,T> + |T|^2), and per-pair FP32
        rounding leaves an error near 1e-8 of the large sums, i.e. far above 1e-6 of the
        objective once the shapes match.  FP32-compensated kernels take PREC_FP64E (FP64
        arithmetic, pairs below 2^-40 culled and the FP32 exponential below 2^-16), which
        keeps the data term within ~2e-10 of FP64 where the shapes nearly match; FP64 kernels
        keep the full FP64 thresholds."""
        prec = _lib.PREC_FP64 if self.prec == _lib.PREC_FP64 else _lib.PREC_FP64E
        return DevKernel(self.kernel, self.order, self.sigma, prec)


def _groups(jobs, key):
    out = {}
    for j in jobs:
        out.setdefault(key(j), []).append(j)
    return out.values()


def _runs(lengths, n_steps):
    """Maximal runs [t0, t1) of steps with one active set, for jobs of ``lengths`` steps
    (job i is active at steps t < lengths[i]): yields (t0, t1, [job indices])."""
    lengths = np.asarray(lengths)
    bounds = np.unique(np.concatenate([lengths[lengths < n_steps], [0, n_steps]])).tolist()
    for t0, t1 in zip(bounds[:-1], bounds[1:]):
        idx = np.nonzero(lengths > t0)[0]
        if idx.size and t1 > t0:
            yield t0, t1, idx


def _u64(vals):
    return np.asarray(vals, dtype=np.uint64)


# --------------------------------------------------------------------------------------
# Processing order of a point set
# --------------------------------------------------------------------------------------
TILE = 256      # sources per LDS tile of the pair kernels
QUARTER = 64    # sources per culling / exponent-form block


def spatial_order(points) -> np.ndarray:
    """A processing order of a point set for the pair kernels (lddmm_spatial_order, host code
    of the kernel library): recursive bisection along the longest side of the bounding box,
    with split points at multiples of 256 (then 64) so that every LDS tile and every 64-point
    quarter of it is one spatially compact subtree.  Returns int32 indices, order[i] = the
    point processed i-th.  Deterministic.  Sums over the points do not depend on it beyond
    rounding; compact tiles let the FP32 Gaussian sums use the dot-product exponent form."""
    P = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    if P.ndim != 2 or not 1 <= P.shape[1] <= 3:
        raise ValueError("points must be an (n, d) array with d in 1..3")
    n = P.shape[0]
    if n > np.iinfo(np.int32).max:
        raise ValueError("point set too large for an int32 order")
    out = np.empty(n, dtype=np.int32)
    _lib.check(_lib.load().lddmm_spatial_order(P.ctypes.data, n, P.shape[1], out.ctypes.data),
               "lddmm_spatial_order")
    return out


def spatial_order_dev(points, device=None) -> torch.Tensor:
    """spatial_order as a device int32 tensor (the ``perm`` of a pair-sum descriptor)."""
    o = spatial_order(points)
    return torch.from_numpy(o).to(device if device is not None else rt.device())


def wants_order(K: "DevKernel") -> bool:
    """Whether K's sums have the dot-product exponent form, which only a processing order
    enables (the FP32 Gaussian K-sums).  Every kernel's sums cull by tile and stage their
    source tiles once per launch when ordered, so order_choice weighs an order for all."""
    return K.kernel == _lib.KERNEL_GAUSS and K.prec == _lib.PREC_FP32C


# The pair kernel's bound on a quarter's squared prescaled offset from its tile origin for
# the dot-product exponent (EXP_R2MAX in the kernel source), and the FP32 Gaussian prescale.
EXP_R2MAX = 2.0
_PRESCALE = float(np.sqrt(np.log2(np.e) / 2.0))


def compact_fraction(points, order, sigma: float) -> float:
    """Fraction of 64-point quarters of ``order`` that the FP32 Gaussian pair kernel would
    evaluate in dot-product form at kernel width sigma: every point within R'^2 <= EXP_R2MAX
    of its tile origin (the mean of the tile's points 32, 96, 160, 224, clamped), in the
    prescaled coordinates x * sqrt(log2(e) / 2) / sigma (lddmm_order_compactness)."""
    P = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int32))
    if P.shape[0] == 0:
        return 0.0
    if o.shape != (P.shape[0],):
        raise ValueError("order must hold one index per point")
    f = _lib.load().lddmm_order_compactness(P.ctypes.data, P.shape[0], P.shape[1], o.ctypes.data,
                                            _PRESCALE / sigma, EXP_R2MAX)
    if f < 0:
        raise ValueError("invalid points or order")
    return f


def cull_radius2(K: "DevKernel") -> float:
    """Squared distance (in the points' own units) beyond which the pair kernels drop a
    (wave, quarter) block of K's sums: the Gaussian below 2^-24 (FP32) or 2^-56 (FP64) of
    its peak, the Laplacian beyond 30.5 sigma (FP32) or 55 sigma (FP64), as cull_r2 of the
    kernel source bounds them."""
    s2 = K.sigma * K.sigma
    if K.kernel == _lib.KERNEL_GAUSS:
        return (24.0 if K.prec == _lib.PREC_FP32C else 56.0) * 2.0 * s2 / np.log2(np.e)
    return (30.5 if K.prec == _lib.PREC_FP32C else 55.0) ** 2 * s2


def live_fraction(points, order, K: "DevKernel", max_pairs: int = 1 << 22) -> float:
    """Fraction of a K-sum's (wave rows, 64-source quarter) blocks of the self sum over
    ``points`` in ``order`` that the pair kernels evaluate: blocks whose bounding boxes lie
    within cull_radius2 of each other (a wave holds 128 consecutive rows in FP32, 64 in
    FP64; lddmm_order_live_fraction).  Beyond ``max_pairs`` block pairs, every k-th row
    block is taken."""
    P = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int32))
    if P.shape[0] == 0:
        return 0.0
    if o.shape != (P.shape[0],):
        raise ValueError("order must hold one index per point")
    rows = 128 if K.prec == _lib.PREC_FP32C else 64
    f = _lib.load().lddmm_order_live_fraction(P.ctypes.data, P.shape[0], P.shape[1], o.ctypes.data,
                                              rows, cull_radius2(K), max_pairs)
    if f < 0:
        raise ValueError("invalid points or order")
    return f


# Segments with fewer work items than this fit in about one round of workgroups on the chip
# (256 CUs hold 1,024-1,280 pair-kernel workgroups at once): no dispatch order shortens them.
MIN_ORDERED_ITEMS = 1024


def item_count(op: int, prec: int, d: int, rec: np.ndarray) -> int:
    """Work items the plan of one segment ``rec`` has, or -1 when its operation takes no
    dispatch order (lddmm_item_order with no output)."""
    r = np.ascontiguousarray(np.array(rec, copy=True).reshape(1))
    r["item_order"] = 0
    r["n_order"] = 0
    n = _lib.load().lddmm_item_order(op, prec, d, r.ctypes.data_as(ctypes.POINTER(_lib.Seg)),
                                     None, None, None, None, 0)
    if n == _lib.EUNSUPPORTED:   # an operation whose launches run in item order only
        return -1
    if n < 0:
        _lib.check(n, "lddmm_item_order")
    return int(n)


def item_order_host(op: int, prec: int, d: int, rec: np.ndarray, y, x1=None, perm=None,
                    min_items: Optional[int] = None) -> Optional[np.ndarray]:
    """Dispatch order of one segment's work items, the items with the most unculled pairs
    first (lddmm_item_order), as host int32 indices; None when the segment plans fewer than
    ``min_items`` (MIN_ORDERED_ITEMS) items or its operation takes no order (lddmm_item_order
    lists those that do).  ``rec``: the (1,) descriptor as it will be launched (device
    pointers may still be unset); ``y``: host copy of its rows (ny, d); ``x1``: host copy of a
    data-term segment's target points; ``perm``: host copy of its processing order.  Host
    work only (the library call releases the interpreter lock), so it may run on a worker
    thread.  The estimate only steers the schedule, so an order made from a registration's
    template serves every later step and trial without changing any result."""
    n = item_count(op, prec, d, rec)
    if n < (MIN_ORDERED_ITEMS if min_items is None else min_items):
        return None
    r = np.ascontiguousarray(np.array(rec, copy=True).reshape(1))
    r["item_order"] = 0
    r["n_order"] = 0
    Y = np.ascontiguousarray(np.asarray(y, dtype=np.float64))
    X1 = None if x1 is None else np.ascontiguousarray(np.asarray(x1, dtype=np.float64))
    P = None if perm is None else np.ascontiguousarray(np.asarray(perm, dtype=np.int32))
    out = np.empty(n, dtype=np.int32)
    _lib.check(_lib.load().lddmm_item_order(op, prec, d, r.ctypes.data_as(ctypes.POINTER(_lib.Seg)),
                                            Y.ctypes.data,
                                            None if X1 is None else X1.ctypes.data,
                                            None if P is None else P.ctypes.data,
                                            out.ctypes.data, n),
               "lddmm_item_order")
    return out


def upload_order(order: Optional[np.ndarray]) -> Optional[torch.Tensor]:
    """A host dispatch order as a device int32 tensor, through pinned memory and
    stream-ordered (no wait on the work already queued)."""
    if order is None:
        return None
    host = torch.empty(order.shape[0], dtype=torch.int32, pin_memory=True)
    host.numpy()[:] = order
    return host.to(rt.device(), non_blocking=True)


def item_order(op: int, prec: int, d: int, rec: np.ndarray, y, x1=None, perm=None,
               min_items: Optional[int] = None) -> Optional[torch.Tensor]:
    """item_order_host as a device int32 tensor for ``attach_item_order``."""
    return upload_order(item_order_host(op, prec, d, rec, y, x1=x1, perm=perm,
                                        min_items=min_items))


def attach_item_order(rec: np.ndarray, order: Optional[torch.Tensor]) -> None:
    """Point descriptor(s) ``rec`` at a dispatch order (None clears it); the caller keeps
    the tensor alive while the descriptors are launched."""
    rec["item_order"] = 0 if order is None else order.data_ptr()
    rec["n_order"] = 0 if order is None else order.numel()


# Smallest point sets whose sums take a processing order.  An ordered launch stages its
# source tiles in a launch of its own (about 3.5 us on gfx950), which costs more than the
# order saves on small sets (flow steps over icospheres of 642 and 2,562 points ran 10-15%
# slower ordered, at every kernel width); from 10,242 points the Gaussian sums gain up to
# 31% (FP64, sigma 0.05), at 40,962 points 19-48%.  Laplacian sums gain only through culling
# and only from about 40,000 points (13-19% at sigma 0.05, 0-1% at wider kernels).
MIN_ORDERED_POINTS = 8192
MIN_ORDERED_POINTS_LAPLACIAN = 32768
# Relative cost of a pair in the dot-product exponent form against the direct form (three
# FMAs and the exponential against three differences, three FMAs and the exponential).
DOT_FORM_COST = 0.78
# FP64 Gaussian sums take the expanded exponent on blocks close to their tile origin: where
# two orders cull alike, the more compact one was 2% faster (40,962 points, sigma 0.25).
XFORM_TIE = 0.02


def order_cost(points, order, K: "DevKernel") -> float:
    """Estimated work of K's self sums over ``points`` in ``order``, per nominal block pair:
    the evaluated fraction of blocks, the FP32 Gaussian's compact quarters at DOT_FORM_COST,
    and for the FP64 Gaussian a small preference for compact quarters (XFORM_TIE)."""
    c = live_fraction(points, order, K)
    if K.kernel == _lib.KERNEL_GAUSS:
        w = (1.0 - DOT_FORM_COST) if wants_order(K) else XFORM_TIE
        c *= 1.0 - w * compact_fraction(points, order, K.sigma)
    return c


def order_choice(points, K: "DevKernel") -> Optional[np.ndarray]:
    """The processing order a registration of ``points`` under K should use, as host int32
    indices, or None when no order pays (fewer than MIN_ORDERED_POINTS points, or
    MIN_ORDERED_POINTS_LAPLACIAN for the Laplacian).  The spatial order competes with the
    points' own order (the identity permutation) on order_cost: sampled curves are already
    compact in arc order, more so than after bisection, and the identity order still stages
    the source tiles once per launch."""
    P = np.asarray(points, dtype=np.float64)
    n = P.shape[0]
    if n < (MIN_ORDERED_POINTS if K.kernel == _lib.KERNEL_GAUSS else MIN_ORDERED_POINTS_LAPLACIAN):
        return None
    ident = np.arange(n, dtype=np.int32)
    o = spatial_order(P)
    return ident if order_cost(P, ident, K) <= order_cost(P, o, K) else o


def order_for(points, K: "DevKernel"):
    """order_choice as a device int32 tensor (the ``perm`` of a pair-sum descriptor)."""
    o = order_choice(points, K)
    return None if o is None else torch.from_numpy(o).to(rt.device())


def axpy_batch(out, a, d, s) -> None:
    """out[i] = a[i] + s[i] * d[i] for lists of contiguous FP64 device tensors of matching
    sizes, the product rounded before the sum (at - eps dir as the reference's NumPy
    evaluates it, with s = -eps); one launch per 32 items (lddmm_axpy_batch)."""
    n = len(out)
    if n == 0:
        return
    if not (len(a) == len(d) == len(s) == n):
        raise ValueError("axpy_batch: lists of different lengths")
    for o, x, y in zip(out, a, d):
        if not (o.numel() == x.numel() == y.numel() and o.dtype == x.dtype == y.dtype == torch.float64
                and o.is_contiguous() and x.is_contiguous() and y.is_contiguous()
                and o.is_cuda and x.is_cuda and y.is_cuda):
            raise ValueError("axpy_batch: contiguous FP64 device tensors of one size per item")
    po = np.fromiter((t.data_ptr() for t in out), np.uint64, n)
    pa = np.fromiter((t.data_ptr() for t in a), np.uint64, n)
    pd = np.fromiter((t.data_ptr() for t in d), np.uint64, n)
    ps = np.asarray(s, dtype=np.float64)
    pl = np.fromiter((t.numel() for t in out), np.int64, n)
    _lib.check(_lib.load().lddmm_axpy_batch(n, po.ctypes.data, pa.ctypes.data, pd.ctypes.data,
                                            ps.ctypes.data, pl.ctypes.data, rt.stream_handle()),
               "lddmm_axpy_batch")


# --------------------------------------------------------------------------------------
# Shooting (flow + energy)
# --------------------------------------------------------------------------------------
@dataclass
class FlowJob:
    x0: torch.Tensor            # (N, d)
    at: torch.Tensor            # (T, N, d)
    K: DevKernel
    xt: torch.Tensor            # (T+1, N, d) output trajectory
    energy: torch.Tensor        # (T,) output: <a_t, K(x_t) a_t>
    v: Optional[torch.Tensor] = None  # (T, N, d) optional velocities K(x_t) a_t
    dt: Optional[float] = None          # time step; 1/T when None
    init: bool = True                   # copy x0 into xt[0] (False: xt[0] already holds x0)
    A: Optional[torch.Tensor] = None    # (T, d, d) affine part, with b (T, d): x += dt (A x + b)
    b: Optional[torch.Tensor] = None
    perm: Optional[torch.Tensor] = None  # (N,) int32 processing order (spatial_order_dev)

    @property
    def T(self):
        return self.at.shape[0]

    @property
    def step(self):
        return 1.0 / self.T if self.dt is None else self.dt


def flow_record(N: int, T: int, K: DevKernel, dt: Optional[float] = None,
                perm: Optional[torch.Tensor] = None) -> np.ndarray:
    """The pointer-free part of a registration's flow-step descriptor (one record), with the
    optional processing order of its points (a device int32 permutation of 0..N-1)."""
    r = rt.segs(1)
    r["ny"] = N
    r["sigma"] = K.sigma
    r["dt"] = 1.0 / T if dt is None else dt
    r["kernel"] = K.kernel
    r["order"] = K.order
    if perm is not None:
        if perm.dtype != torch.int32 or perm.numel() != N or not perm.is_contiguous():
            raise ValueError("perm must be a contiguous int32 tensor of N entries")
        r["perm"] = perm.data_ptr()
    return r


def shoot_records(base, prec, d, T, xt, at, energy, v=None, A=None, b=None) -> None:
    """Flow launches for n trajectories of one (kernel family, precision, d) from arrays:
    ``base`` (n,) flow_record rows; T (n,) step counts; xt, at, energy (n,) device base
    addresses of the (T+1, N, d) trajectories, (T, N, d) momenta and (T,) energy rows; v
    optional velocity outputs, A, b optional affine parts (T, d, d) and (T, d) (0 = none).
    Time step t of every trajectory with t < T goes into launch t."""
    T = np.asarray(T, dtype=np.int64)
    stride = base["ny"].astype(np.uint64) * np.uint64(d * 8)
    xt, at, en = _u64(xt), _u64(at), _u64(energy)
    vv = _u64(v) if v is not None else None
    AA = _u64(A) if A is not None else None
    bb = _u64(b) if b is not None else None
    for t0, t1, idx in _runs(T, int(T.max())):
        i = np.asarray(idx)
        tt = np.arange(t0, t1, dtype=np.uint64)[:, None]
        steps = np.repeat(base[i][None, :], t1 - t0, axis=0)
        st = stride[i]
        steps["y"] = xt[i] + tt * st
        steps["w0"] = at[i] + tt * st
        steps["out0"] = xt[i] + (tt + 1) * st
        if vv is not None:
            steps["out1"] = np.where(vv[i] != 0, vv[i] + tt * st, 0)
        if AA is not None:
            steps["w1"] = np.where(AA[i] != 0, AA[i] + tt * np.uint64(d * d * 8), 0)
            steps["w2"] = np.where(bb[i] != 0, bb[i] + tt * np.uint64(d * 8), 0)
        steps["total"] = en[i] + tt * np.uint64(8)
        rt.run_steps(_lib.OP_FLOW_STEP, prec, d, steps)


def shoot(jobs: Sequence[FlowJob]) -> None:
    """xt[t+1] = xt[t] + dt K(xt[t]) at[t], energy[t] = <at[t], K(xt[t]) at[t]>, dt = 1/T."""
    init = [j for j in jobs if j.init]
    if init:
        torch._foreach_copy_([j.xt[0] for j in init], [j.x0 for j in init])
    for group in _groups(jobs, lambda j: (j.K.family, j.x0.shape[1])):
        d = group[0].x0.shape[1]
        base = rt.stack([rt.raw(flow_record(j.xt.shape[1], j.T, j.K, j.step, j.perm))
                         for j in group])
        aff = any(j.A is not None for j in group)
        shoot_records(base, group[0].K.prec, d, [j.T for j in group],
                      [j.xt.data_ptr() for j in group], [j.at.data_ptr() for j in group],
                      [j.energy.data_ptr() for j in group], [rt.ptr(j.v) for j in group],
                      [rt.ptr(j.A) for j in group] if aff else None,
                      [rt.ptr(j.b) for j in group] if aff else None)


# --------------------------------------------------------------------------------------
# Covector (adjoint)
# --------------------------------------------------------------------------------------
@dataclass
class CovectorJob:
    xt: torch.Tensor            # (T+1, N, d) trajectory of at
    at: torch.Tensor            # (T, N, d)
    px1: torch.Tensor           # (N, d) end-point covector
    K: DevKernel
    regweight: float
    pxt: torch.Tensor           # (T+1, N, d) output
    dt: Optional[float] = None  # time step; 1/T when None
    A: Optional[torch.Tensor] = None  # (T, d, d) affine part: p_t += dt p_{t+1} A_t
    perm: Optional[torch.Tensor] = None  # (N,) int32 processing order (spatial_order_dev)

    @property
    def T(self):
        return self.at.shape[0]

    @property
    def step(self):
        return 1.0 / self.T if self.dt is None else self.dt


def covector_record(N: int, T: int, K: DevKernel, regweight: float,
                    dt: Optional[float] = None, perm: Optional[torch.Tensor] = None) -> np.ndarray:
    """The pointer-free part of a registration's covector-step descriptor (one record), with
    the optional processing order of its points."""
    r = flow_record(N, T, K, dt, perm)
    r["coef"] = regweight
    return r


def covector_records(base, prec, d, T, xt, at, pxt, A=None) -> None:
    """Covector launches for n trajectories of one (kernel family, precision, d) from arrays:
    ``base`` (n,) covector_record rows; T (n,) step counts; xt, at, pxt (n,) device base
    addresses of the (T+1, N, d) trajectories, (T, N, d) momenta and (T+1, N, d) covectors,
    whose slice T already holds the end-point covector; A optional (T, d, d) affine parts
    (0 = none).  Launch s handles t = T - 1 - s of every trajectory with s < T."""
    T = np.asarray(T, dtype=np.int64)
    stride = base["ny"].astype(np.uint64) * np.uint64(d * 8)
    xt, at, px = _u64(xt), _u64(at), _u64(pxt)
    AA = _u64(A) if A is not None else None
    for s0, s1, idx in _runs(T, int(T.max())):
        i = np.asarray(idx)
        ss = np.arange(s0, s1, dtype=np.uint64)[:, None]
        tt = T[i].astype(np.uint64)[None, :] - np.uint64(1) - ss
        steps = np.repeat(base[i][None, :], s1 - s0, axis=0)
        st = stride[i]
        steps["y"] = xt[i] + tt * st
        steps["w0"] = at[i] + tt * st
        steps["w1"] = px[i] + (tt + 1) * st
        steps["out0"] = px[i] + tt * st
        if AA is not None:
            steps["w2"] = np.where(AA[i] != 0, AA[i] + tt * np.uint64(d * d * 8), 0)
        rt.run_steps(_lib.OP_COVECTOR_STEP, prec, d, steps)


def covector(jobs: Sequence[CovectorJob]) -> None:
    """pxt[T] = px1; pxt[t] = pxt[t+1] + dt * applyDiffKT(xt[t], [p, a, -2 lam a], [a, p, a])."""
    torch._foreach_copy_([j.pxt[j.T] for j in jobs], [j.px1 for j in jobs])
    for group in _groups(jobs, lambda j: (j.K.family, j.xt.shape[2])):
        d = group[0].xt.shape[2]
        base = rt.stack([rt.raw(covector_record(j.xt.shape[1], j.T, j.K, j.regweight, j.step,
                                                j.perm))
                         for j in group])
        aff = any(j.A is not None for j in group)
        covector_records(base, group[0].K.prec, d, [j.T for j in group],
                         [j.xt.data_ptr() for j in group], [j.at.data_ptr() for j in group],
                         [j.pxt.data_ptr() for j in group],
                         [rt.ptr(j.A) for j in group] if aff else None)


# --------------------------------------------------------------------------------------
# RKHS products over a trajectory
# --------------------------------------------------------------------------------------
@dataclass
class KDotJob:
    xt: torch.Tensor            # (T+1, N, d) (the first T slices are used)
    g1: torch.Tensor            # (T, N, d)
    g2: Sequence[torch.Tensor]  # 1..2 tensors (T, N, d)
    K: DevKernel
    out: torch.Tensor           # (T, 2) output totals: out[t, j] = <g2[j][t], K(xt[t]) g1[t]>
    perm: Optional[torch.Tensor] = None  # (N,) int32 processing order (spatial_order_dev)


def kdot_records(base, prec, d, T, xt, g1, g2, na, out) -> None:
    """One launch of RKHS products for n trajectories of one (kernel family, precision, d):
    ``base`` (n,) flow_record rows; T (n,) step counts; xt, g1 (n,) base addresses; g2 (n, 2)
    base addresses of the second factors (na[i] of them used); out (n,) addresses of (T, 2)
    totals.  Every time step of every trajectory is a segment."""
    T = np.asarray(T, dtype=np.int64)
    n = len(T)
    stride = base["ny"].astype(np.uint64) * np.uint64(d * 8)
    rep = np.repeat(np.arange(n), T)                       # trajectory of each segment
    first = np.concatenate([[0], np.cumsum(T)[:-1]])
    tt = (np.arange(int(T.sum())) - first[rep]).astype(np.uint64)
    r = base[rep].copy()
    attach_item_order(r, None)   # a flow step's dispatch order does not fit the KDOT plan
    r["dt"] = 0.0
    r["na"] = np.asarray(na, dtype=np.int32)[rep]
    st = stride[rep]
    r["y"] = _u64(xt)[rep] + tt * st
    r["w0"] = _u64(g1)[rep] + tt * st
    g2 = _u64(g2).reshape(n, 2)
    for q in range(2):
        r["g"][:, q] = np.where(g2[rep, q] != 0, g2[rep, q] + tt * st, 0)
    r["total"] = _u64(out)[rep] + tt * np.uint64(16)
    rt.launch(_lib.OP_KDOT, prec, d, r)


def kdot(jobs: Sequence[KDotJob]) -> None:
    """One launch: every time step of every job is a segment."""
    for group in _groups(jobs, lambda j: (j.K.family, j.xt.shape[2])):
        d = group[0].xt.shape[2]
        base = rt.stack([rt.raw(flow_record(j.xt.shape[1], j.g1.shape[0], j.K, perm=j.perm))
                         for j in group])
        g2 = [[g.data_ptr() for g in j.g2] + [0] * (2 - len(j.g2)) for j in group]
        kdot_records(base, group[0].K.prec, d, [j.g1.shape[0] for j in group],
                     [j.xt.data_ptr() for j in group], [j.g1.data_ptr() for j in group], g2,
                     [len(j.g2) for j in group], [j.out.data_ptr() for j in group])


# --------------------------------------------------------------------------------------
# Currents data term
# --------------------------------------------------------------------------------------
@dataclass
class CurrentsJob:
    c1: torch.Tensor            # (F1, d) centers of the deformed shape
    cr1: torch.Tensor           # (F1, d) surfels / edge vectors of the deformed shape
    c2: torch.Tensor            # (F2, d) target centers
    cr2: torch.Tensor           # (F2, d) target surfels
    K: DevKernel
    total: torch.Tensor         # (1,) output: currentNormDef
    z1: Optional[torch.Tensor] = None   # (F1, d) gradient form outputs
    dz: Optional[torch.Tensor] = None
    coef: float = 1.0 / 3.0


def currents_record(c1, cr1, c2, cr2, K: DevKernel, coef: float) -> np.ndarray:
    """The currents descriptor of (c1, cr1) against (c2, cr2) without its outputs (one record)."""
    r = rt.segs(1)
    r["y"] = c1.data_ptr()
    r["w0"] = cr1.data_ptr()
    r["x1"] = c2.data_ptr()
    r["w1"] = cr2.data_ptr()
    r["ny"] = c1.shape[0]
    r["nx1"] = c2.shape[0]
    r["sigma"] = K.sigma
    r["coef"] = coef
    r["kernel"] = K.kernel
    r["order"] = K.order
    return r


def currents_records(recs, prec, d, totals, op=_lib.OP_CURRENTS) -> None:
    """Data-term energies (currents, or measure with op=OP_MEASURE) of n records of one
    (kernel family, precision, d) into ``totals`` (device addresses), one launch."""
    r = np.array(recs, copy=True)
    r["total"] = _u64(totals)
    rt.launch(op, prec, d, r)


def currents(jobs: Sequence[CurrentsJob]) -> None:
    for group in _groups(jobs, lambda j: (j.K.family, j.c1.shape[1], j.dz is not None)):
        d = group[0].c1.shape[1]
        r = rt.segs(len(group))
        r["y"] = [j.c1.data_ptr() for j in group]
        r["w0"] = [j.cr1.data_ptr() for j in group]
        r["x1"] = [j.c2.data_ptr() for j in group]
        r["w1"] = [j.cr2.data_ptr() for j in group]
        r["out0"] = [rt.ptr(j.z1) for j in group]
        r["out1"] = [rt.ptr(j.dz) for j in group]
        r["total"] = [j.total.data_ptr() for j in group]
        r["ny"] = [j.c1.shape[0] for j in group]
        r["nx1"] = [j.c2.shape[0] for j in group]
        r["sigma"] = [j.K.sigma for j in group]
        r["coef"] = [j.coef for j in group]
        r["kernel"] = [j.K.kernel for j in group]
        r["order"] = [j.K.order for j in group]
        rt.launch(_lib.OP_CURRENTS, group[0].K.prec, d, r)


# --------------------------------------------------------------------------------------
# Measure data term
# --------------------------------------------------------------------------------------
def measure_weights(cr: np.ndarray, eps: float = 1e-10) -> np.ndarray:
    """Host face weights of the measure data term, sqrt(|cr|^2 + eps) in the reference's
    NumPy form (surfaces.py:1098, curves.py:637); curves.measureNorm0 uses eps = 0
    (curves.py:625)."""
    cr = np.asarray(cr, dtype=np.float64)
    return np.sqrt((cr ** 2).sum(axis=1) + eps)


@dataclass
class MeasureJob:
    c1: torch.Tensor            # (F1, d) centers
    a1: torch.Tensor            # (F1,) weights
    c2: Optional[torch.Tensor]  # (F2, d) target centers, or None for <a1, K(c1, c1) a1>
    a2: Optional[torch.Tensor]  # (F2,) target weights
    K: DevKernel
    total: torch.Tensor         # (1,) output: measureNormDef (measureNorm0 without target)
    n1: Optional[torch.Tensor] = None   # (F1, d) normals, for the gradient form
    z1: Optional[torch.Tensor] = None   # (F1, d) gradient form outputs
    dz: Optional[torch.Tensor] = None
    coef: float = 1.0 / 3.0


def measure_record(c1, a1, c2, a2, K: DevKernel, coef: float, n1=None) -> np.ndarray:
    """The measure descriptor of (c1, a1) against (c2, a2) without its outputs (one record);
    c2 None leaves the target out (the shape's own norm)."""
    r = rt.segs(1)
    r["y"] = c1.data_ptr()
    r["w0"] = a1.data_ptr()
    r["w2"] = rt.ptr(n1)
    r["x1"] = rt.ptr(c2)
    r["w1"] = rt.ptr(a2)
    r["ny"] = c1.shape[0]
    r["nx1"] = 0 if c2 is None else c2.shape[0]
    r["sigma"] = K.sigma
    r["coef"] = coef
    r["kernel"] = K.kernel
    r["order"] = K.order
    return r


def measure(jobs: Sequence[MeasureJob]) -> None:
    for group in _groups(jobs, lambda j: (j.K.family, j.c1.shape[1], j.dz is not None)):
        d = group[0].c1.shape[1]
        r = rt.stack([rt.raw(measure_record(j.c1, j.a1, j.c2, j.a2, j.K, j.coef, j.n1))
                      for j in group])
        r["out0"] = [rt.ptr(j.z1) for j in group]
        r["out1"] = [rt.ptr(j.dz) for j in group]
        r["total"] = [j.total.data_ptr() for j in group]
        rt.launch(_lib.OP_MEASURE, group[0].K.prec, d, r)


# --------------------------------------------------------------------------------------
# Varifold data term (surfaces)
# --------------------------------------------------------------------------------------
@dataclass
class VarifoldJob:
    c1: torch.Tensor            # (F1, 3) centers
    a1: torch.Tensor            # (F1,) weights sqrt(|n|^2 + 1e-10)
    n1: torch.Tensor            # (F1, 3) surfels
    c2: Optional[torch.Tensor]  # (F2, 3) target, or None for the shape's own norm
    a2: Optional[torch.Tensor]
    n2: Optional[torch.Tensor]
    K: DevKernel
    total: torch.Tensor         # (1,) varifoldNormDef (varifoldNorm0 without target)
    z1: Optional[torch.Tensor] = None   # (F1, 3) gradient form outputs
    dz: Optional[torch.Tensor] = None
    coef: float = 1.0 / 3.0


def varifold_record(c1, a1, n1, c2, a2, n2, K: DevKernel, coef: float) -> np.ndarray:
    """The varifold descriptor of (c1, a1, n1) against (c2, a2, n2) without its outputs (one
    record); c2 None leaves the target out."""
    r = rt.segs(1)
    r["y"] = c1.data_ptr()
    r["w0"] = a1.data_ptr()
    r["w2"] = n1.data_ptr()
    r["x1"] = rt.ptr(c2)
    r["w1"] = rt.ptr(a2)
    r["sw"][0, 0] = rt.ptr(n2)
    r["ny"] = c1.shape[0]
    r["nx1"] = 0 if c2 is None else c2.shape[0]
    r["sigma"] = K.sigma
    r["coef"] = coef
    r["kernel"] = K.kernel
    r["order"] = K.order
    return r


def varifold(jobs: Sequence[VarifoldJob]) -> None:
    for group in _groups(jobs, lambda j: (j.K.family, j.dz is not None)):
        r = rt.stack([rt.raw(varifold_record(j.c1, j.a1, j.n1, j.c2, j.a2, j.n2, j.K, j.coef))
                      for j in group])
        r["out0"] = [rt.ptr(j.z1) for j in group]
        r["out1"] = [rt.ptr(j.dz) for j in group]
        r["total"] = [j.total.data_ptr() for j in group]
        rt.launch(_lib.OP_VARIFOLD, group[0].K.prec, 3, r)


@dataclass
class KNormJob:
    c: torch.Tensor
    cr: torch.Tensor
    K: DevKernel
    total: torch.Tensor         # (1,)


def knorm(jobs: Sequence[KNormJob]) -> None:
    for group in _groups(jobs, lambda j: (j.K.family, j.c.shape[1])):
        d = group[0].c.shape[1]
        r = rt.segs(len(group))
        r["y"] = [j.c.data_ptr() for j in group]
        r["w0"] = [j.cr.data_ptr() for j in group]
        r["total"] = [j.total.data_ptr() for j in group]
        r["ny"] = [j.c.shape[0] for j in group]
        r["sigma"] = [j.K.sigma for j in group]
        r["kernel"] = [j.K.kernel for j in group]
        r["order"] = [j.K.order for j in group]
        rt.launch(_lib.OP_KNORM, group[0].K.prec, d, r)


# --------------------------------------------------------------------------------------
# Geometry
# --------------------------------------------------------------------------------------
def surface_geometry(vertices: torch.Tensor, faces: torch.Tensor, centers: torch.Tensor,
                     surfel: torch.Tensor) -> None:
    lib = _lib.load()
    _lib.check(lib.lddmm_surface_geometry(vertices.data_ptr(), faces.data_ptr(), faces.shape[0],
                                          centers.data_ptr(), surfel.data_ptr(),
                                          rt.stream_handle()), "lddmm_surface_geometry")


def curve_geometry(vertices, edges, centers, lenel) -> None:
    lib = _lib.load()
    _lib.check(lib.lddmm_curve_geometry(vertices.data_ptr(), edges.data_ptr(), edges.shape[0],
                                        vertices.shape[1], centers.data_ptr(), lenel.data_ptr(),
                                        rt.stream_handle()), "lddmm_curve_geometry")


_TOPOLOGY: "collections.OrderedDict" = None
TOPOLOGY_CACHE = 64


def topology(faces_np, nv: int):
    """Device faces (int32) and vertex-incidence CSR (ptr, inc) of a mesh, shared by every
    shape with the same faces and vertex count (read-only; the last TOPOLOGY_CACHE distinct
    topologies are kept)."""
    global _TOPOLOGY
    import collections
    if _TOPOLOGY is None:
        _TOPOLOGY = collections.OrderedDict()
    faces_np = np.ascontiguousarray(faces_np)
    check_faces(faces_np, nv)
    key = (torch.cuda.current_device(), int(nv), faces_np.shape, faces_np.dtype.str,
           hash(faces_np.tobytes()))
    hit = _TOPOLOGY.get(key)
    if hit is not None:
        _TOPOLOGY.move_to_end(key)
        return hit
    ptr, inc = incidence(faces_np, nv)
    val = (rt.as_dev(faces_np.astype(np.int32), dtype=torch.int32),
           rt.as_dev(ptr, dtype=torch.int32), rt.as_dev(inc, dtype=torch.int32))
    _TOPOLOGY[key] = val
    while len(_TOPOLOGY) > TOPOLOGY_CACHE:
        _TOPOLOGY.popitem(last=False)
    return val


def check_faces(faces_np, nv: int) -> None:
    """Reject face arrays the kernels cannot index safely: they must be integer (F, 2) or
    (F, 3) arrays with every index in [0, nv) and nv below 2^31 (the device copies are int32)."""
    if faces_np.ndim != 2 or faces_np.shape[1] not in (2, 3) or \
            not np.issubdtype(faces_np.dtype, np.integer):
        raise ValueError(f"faces must be an integer (F, 2) or (F, 3) array, got "
                         f"{faces_np.dtype} {faces_np.shape}")
    if not 0 <= int(nv) < 2 ** 31:
        raise ValueError(f"vertex count {nv} is outside [0, 2^31)")
    if faces_np.size and (int(faces_np.min()) < 0 or int(faces_np.max()) >= nv):
        raise ValueError(f"face indices must lie in [0, {nv}), got [{int(faces_np.min())}, "
                         f"{int(faces_np.max())}]")


def incidence(faces_np, nv: int):
    """Vertex -> (face, corner) CSR sorted by (corner, face), i.e. in the order the reference's
    scatter loops visit them (surfaces.py:1074-1087, curves.py:604-610).  Host NumPy."""
    import numpy as np
    F, k = faces_np.shape
    corner = np.repeat(np.arange(k)[None, :], F, axis=0)
    face = np.repeat(np.arange(F)[:, None], k, axis=1)
    v = faces_np.reshape(-1)
    code = (face * k + corner).reshape(-1)
    # Order: by vertex, then corner, then face.
    order = np.lexsort((face.reshape(-1), corner.reshape(-1), v))
    ptr = np.zeros(nv + 1, dtype=np.int64)
    np.add.at(ptr, v + 1, 1)
    ptr = np.cumsum(ptr)
    return ptr.astype(np.int32), code[order].astype(np.int32)


def surface_vertex_gradient(vertices, faces, z1, dz, ptr, inc, scale, px) -> None:
    lib = _lib.load()
    _lib.check(lib.lddmm_surface_vertex_gradient(vertices.data_ptr(), faces.data_ptr(),
                                                 z1.data_ptr(), dz.data_ptr(), ptr.data_ptr(),
                                                 inc.data_ptr(), vertices.shape[0], float(scale),
                                                 px.data_ptr(), rt.stream_handle()),
               "lddmm_surface_vertex_gradient")


def curve_vertex_gradient(z1, dz, ptr, inc, nv, scale, px) -> None:
    lib = _lib.load()
    _lib.check(lib.lddmm_curve_vertex_gradient(z1.data_ptr(), dz.data_ptr(), ptr.data_ptr(),
                                               inc.data_ptr(), nv, z1.shape[1], float(scale),
                                               px.data_ptr(), rt.stream_handle()),
               "lddmm_curve_vertex_gradient")


# --------------------------------------------------------------------------------------
# Batched geometry and vertex gathers (one launch for many shapes)
# --------------------------------------------------------------------------------------
GEOM_DT = np.dtype(_lib.Geom)
GATHER_DT = np.dtype(_lib.Gather)


def _blk0(counts):
    c = (np.asarray(counts, dtype=np.int64) + 255) // 256
    return np.concatenate([[0], np.cumsum(c)[:-1]])


def geometry_batch(shapes, vertices) -> None:
    """shape.update(v) for every (shape, v): face centers and normals (and measure weights,
    for shapes that hold them) of many surfaces and curves in one launch."""
    if not shapes:
        return
    lib = _lib.load()
    r = np.zeros(len(shapes), GEOM_DT)
    for i, (sh, v) in enumerate(zip(shapes, vertices)):
        sh.vertices = v
    r["vertices"] = [sh.vertices.data_ptr() for sh in shapes]
    r["faces"] = [sh.faces.data_ptr() for sh in shapes]
    r["centers"] = [sh.centers.data_ptr() for sh in shapes]
    r["normals"] = [sh.surfel.data_ptr() for sh in shapes]
    r["weights"] = [rt.ptr(getattr(sh, "weights", None)) for sh in shapes]
    r["nf"] = [sh.faces.shape[0] for sh in shapes]
    r["k"] = [sh.faces.shape[1] for sh in shapes]
    r["d"] = [sh.centers.shape[1] for sh in shapes]
    r["blk0"] = _blk0(r["nf"])
    dev = rt.upload(r)
    _lib.check(lib.lddmm_geometry_batch(r.ctypes.data_as(ctypes.POINTER(_lib.Geom)), dev.data_ptr(),
                                        len(shapes), rt.stream_handle()), "lddmm_geometry_batch")


def geometry_record(sh) -> np.ndarray:
    """The geometry descriptor of a device shape without its vertex address (one record)."""
    r = np.zeros(1, GEOM_DT)
    r["faces"] = sh.faces.data_ptr()
    r["centers"] = sh.centers.data_ptr()
    r["normals"] = sh.surfel.data_ptr()
    r["weights"] = rt.ptr(getattr(sh, "weights", None))
    r["nf"] = sh.faces.shape[0]
    r["k"] = sh.faces.shape[1]
    r["d"] = sh.centers.shape[1]
    return r


def geometry_records(recs, vertices) -> None:
    """Face geometry of n shapes from their geometry_record rows and vertex addresses."""
    lib = _lib.load()
    r = np.array(recs, copy=True)
    r["vertices"] = _u64(vertices)
    r["blk0"] = _blk0(r["nf"])
    dev = rt.upload(r)
    _lib.check(lib.lddmm_geometry_batch(r.ctypes.data_as(ctypes.POINTER(_lib.Geom)), dev.data_ptr(),
                                        len(r), rt.stream_handle()), "lddmm_geometry_batch")


def gather_record(sh, z1, dz, scale: float, px_ptr: int) -> np.ndarray:
    """The vertex-gather descriptor of a device shape without its vertex address (one record)."""
    r = np.zeros(1, GATHER_DT)
    r["faces"] = sh.faces.data_ptr()
    r["z1"] = z1.data_ptr()
    r["dz"] = dz.data_ptr()
    r["incid_ptr"] = sh.ptr.data_ptr()
    r["incid"] = sh.inc.data_ptr()
    r["px"] = px_ptr
    r["nv"] = sh.nv
    r["scale"] = scale
    r["k"] = sh.faces.shape[1]
    r["d"] = sh.centers.shape[1]
    return r


def gather_records(recs, vertices) -> None:
    """Vertex gathers of n shapes from their gather_record rows and vertex addresses."""
    lib = _lib.load()
    r = np.array(recs, copy=True)
    r["vertices"] = _u64(vertices)
    r["blk0"] = _blk0(r["nv"])
    dev = rt.upload(r)
    _lib.check(lib.lddmm_vertex_gradient_batch(r.ctypes.data_as(ctypes.POINTER(_lib.Gather)),
                                               dev.data_ptr(), len(r), rt.stream_handle()),
               "lddmm_vertex_gradient_batch")


def gather_batch(shapes, z1s, dzs, scales, pxs) -> None:
    """Vertex gathers of the currents gradient for many shapes in one launch."""
    if not shapes:
        return
    lib = _lib.load()
    r = np.zeros(len(shapes), GATHER_DT)
    r["vertices"] = [sh.vertices.data_ptr() for sh in shapes]
    r["faces"] = [sh.faces.data_ptr() for sh in shapes]
    r["z1"] = [z.data_ptr() for z in z1s]
    r["dz"] = [z.data_ptr() for z in dzs]
    r["incid_ptr"] = [sh.ptr.data_ptr() for sh in shapes]
    r["incid"] = [sh.inc.data_ptr() for sh in shapes]
    r["px"] = [p.data_ptr() for p in pxs]
    r["nv"] = [sh.nv for sh in shapes]
    r["scale"] = scales
    r["k"] = [sh.faces.shape[1] for sh in shapes]
    r["d"] = [sh.centers.shape[1] for sh in shapes]
    r["blk0"] = _blk0(r["nv"])
    dev = rt.upload(r)
    _lib.check(lib.lddmm_vertex_gradient_batch(r.ctypes.data_as(ctypes.POINTER(_lib.Gather)),
                                               dev.data_ptr(), len(shapes), rt.stream_handle()),
               "lddmm_vertex_gradient_batch")


LMK_DT = np.dtype(_lib.Landmark)


def landmark_batch(xs, ys, totals, grads=None, scales=None) -> None:
    """Landmark data terms of many registrations in one launch: totals[i] (a 1-element
    device view) = sum |x_i - y_i|^2 in FP64; grads[i] (optional) = scales[i] (x_i - y_i)."""
    if not xs:
        return
    lib = _lib.load()
    n = len(xs)
    r = np.zeros(n, LMK_DT)
    r["x"] = [x.data_ptr() for x in xs]
    r["y"] = [y.data_ptr() for y in ys]
    r["total"] = [rt.ptr(t) for t in totals]
    if grads is not None:
        r["grad"] = [rt.ptr(g) for g in grads]
        r["scale"] = scales
    r["n"] = [x.numel() for x in xs]
    dev = rt.upload(r)
    _lib.check(lib.lddmm_landmark_batch(r.ctypes.data_as(ctypes.POINTER(_lib.Landmark)),
                                        dev.data_ptr(), n, rt.stream_handle()),
               "lddmm_landmark_batch")
