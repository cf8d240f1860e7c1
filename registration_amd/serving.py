"""Continuous-batching registration service (no counterpart in the reference, whose only
multi-registration mode is a sequential loop over target files,
py-lddmm/surfaceMatchingDirectory.py:46-55).

Requests (surface pairs, curve pairs, landmark sets) arrive over time.  Each admitted request becomes a
GPU-resident registration (matching.DeviceMatching) whose optimiser runs as a generator
(conjugateGradient.cg_steps).  One scheduling round advances every active registration by one
optimiser request: requests of the same kind (line-search trial, gradient, RKHS products,
objective) are answered together by the batch_* functions, which issue one segmented launch
per kernel-sum stage for all of them.  Registrations finish at different times; new arrivals
are admitted as soon as device memory allows (HBM-footprint admission), so the batch is
refilled continuously.

A registration's result does not depend on its batch neighbours: every segment keeps its own
chunking and reduction order, so its kernel sums, and therefore its optimiser path, are the
same alone and in any mix.

Precision guard.  FP32-compensated kernel sums carry per-term rounding (kernel values and
64-term FP32 runs) that is small against the sum of |terms|; a registration whose momenta
are large and oscillating makes the velocity a small difference of large sums, and its
energy error can exceed 1e-6 (a closed curve with 931 points, Laplacian kernel of width
0.48 and momenta up to 1.5: 1.15e-5).  When the optimiser of an FP32-compensated
registration stops, its objective is evaluated once more in FP64 at the final iterate; if
the two differ by more than ``certify_tol`` relative, the registration is run again from
the start in FP64 and that run is its result.  Both paths (served and plain) apply the same
rule, so results stay identical alone and in the mix.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import torch

from . import conjugateGradient as cg
from . import curveMatching as cm
from . import curves
from . import kernelFunctions as kfun
from . import landmarkMatching as lm
from . import matching
from . import surfaceMatching as sm
from . import surfaces


@dataclass
class Request:
    """One registration: ``kind`` 'surface', 'curve' or 'landmark'; template/target are
    (faces, vertices), with faces None for landmark sets (corresponding points); kernel
    names/widths as in SurfaceMatchingParam (sigmaDist unused by landmark sets); errorType
    'current', 'measure' or (surfaces) 'varifold' selects the data term of surfaces and
    curves; affine ('none', 'affine', 'similitude', 'euclidean', 'translation') and
    affineWeight the global affine part of surfaces and curves; ``arrival`` in seconds from
    service start."""

    rid: int
    kind: str
    template: tuple
    target: tuple
    sigmaKernel: float
    sigmaDist: float
    sigmaError: float
    T: int
    maxIter: int
    typeKernel: str = 'gauss'
    order: int = 10
    arrival: float = 0.0
    errorType: str = 'current'
    affine: str = 'none'
    affineWeight: float = 1.0


@dataclass
class Result:
    rid: int
    energy: float
    iterations: int
    stop_reason: str
    arrival: float
    admitted: float
    finished: float
    history: list = field(default_factory=list)
    momenta: Optional[torch.Tensor] = None   # final at (T, N, d) when the server keeps it
    precision: str = 'fp32c'                 # arithmetic of the run that gave the result
    energy_fp64: float = float('nan')        # the FP64 objective at the final iterate
    fp32c_error: float = float('nan')        # |energy - FP64 objective| / |FP64 objective|
                                             # of the FP32-compensated run (before any re-run)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def build_registration(req: Request, precision: str):
    """The device-resident matching object of a request (same as the plain path builds).
    The drivers keep the reference's default verb=True: the reference's cg disables CG on a
    bad direction only when verb is set (conjugateGradient.py:151-154), so verb=False would
    let a line search accept an objective increase there."""
    Kd = kfun.Kernel(name=req.typeKernel, sigma=req.sigmaKernel, order=req.order, precision=precision)
    Kw = kfun.Kernel(name='gauss', sigma=req.sigmaDist, precision=precision)
    if req.kind == 'surface':
        p = sm.SurfaceMatchingParam(timeStep=1.0 / req.T, KparDiff=Kd, KparDist=Kw,
                                    sigmaError=req.sigmaError, errorType=req.errorType)
        return sm.SurfaceMatching(outputDir=None, Template=surfaces.Surface(FV=req.template),
                                  Target=surfaces.Surface(FV=req.target), param=p,
                                  maxIter=req.maxIter, testGradient=False, affine=req.affine,
                                  affineWeight=req.affineWeight)
    if req.kind == 'landmark':
        p = lm.LandmarkMatchingParam(timeStep=1.0 / req.T, KparDiff=Kd, sigmaError=req.sigmaError)
        return lm.LandmarkMatching(Template=req.template[1], Target=req.target[1], param=p,
                                   maxIter=req.maxIter)
    if req.kind == 'curve':
        p = cm.CurveMatchingParam(timeStep=1.0 / req.T, KparDiff=Kd, KparDist=Kw,
                                  sigmaError=req.sigmaError, errorType=req.errorType)
        return cm.CurveMatching(outputDir=None, Template=curves.Curve(FV=req.template),
                                Target=curves.Curve(FV=req.target), param=p, maxIter=req.maxIter,
                                affine=req.affine, affineWeight=req.affineWeight)
    raise ValueError(f"unknown request kind {req.kind!r}")


def footprint_bytes(req: Request) -> int:
    """Device memory of one active registration (float64): trajectories (3 buffers), momenta,
    trial, direction, gradient and CG history (~8 (T,N,d) arrays), covector, and the currents
    buffers of the deformed shapes (landmark sets: the target and the data gradient)."""
    F, V = req.template
    N, d = V.shape
    T = req.T
    traj = 3 * (T + 1) * N * d
    mom = 9 * T * N * d
    cur = 10 * F.shape[0] * d if F is not None else 2 * N * d
    return 8 * (traj + mom + cur) + 64 * 1024


_HOST_ONLY = (cg.ACCEPT, cg.END_ITER)
CERTIFY = 'certify'   # an FP64 objective at the final iterate of an FP32-compensated run
DEFAULT_CERTIFY_TOL = 5e-7


def fp64_twin(req: Request, opt):
    """An FP64 registration of ``req`` holding ``opt``'s current iterate (momenta and affine
    coefficients), for evaluating its objective in FP64."""
    twin = build_registration(req, 'fp64')
    twin.at = opt.at.clone()
    if getattr(opt, 'affineDim', 0) > 0:
        twin.Afft = opt.Afft.clone()
    return twin


def _rel(a, b):
    return abs(a - b) / abs(b) if b != 0 else abs(a - b)

# Errors of a request's own inputs (shapes, indices, options outside the GPU path), raised
# by build_registration before any device work; device errors are not caught.
_REJECT = (ValueError, NotImplementedError, IndexError, TypeError)


class _Active:
    __slots__ = ("req", "opt", "gen", "pending", "admitted", "precision", "twin", "err32")

    def __init__(self, req, opt, admitted, precision):
        self.req = req
        self.admitted = admitted
        self.twin = None
        self.err32 = float('nan')
        self.start(opt, precision)

    def start(self, opt, precision):
        self.opt = opt
        self.precision = precision
        self.gen = opt.cg_steps()
        self.pending = next(self.gen)
        while self.pending[0] in _HOST_ONLY:
            self.pending = self.gen.send(getattr(opt, self.pending[0])())


_LAUNCH = {
    cg.TRY: lambda regs: matching.launch_update_try(
        [(a.opt, a.pending[1], a.pending[2], a.pending[3]) for a in regs]),
    cg.GRADIENT: lambda regs: matching.launch_gradient([(a.opt, a.pending[1]) for a in regs]),
    cg.DOT: lambda regs: matching.launch_dot([(a.opt, a.pending[1], a.pending[2]) for a in regs]),
    cg.OBJECTIVE: lambda regs: matching.launch_objective([a.opt for a in regs]),
    CERTIFY: lambda regs: matching.launch_objective([a.twin for a in regs]),
}


class Server:
    """Continuous-batching service on the current GPU.

    hbm_cap_bytes: admission cap on the summed footprint of active registrations;
    max_active: cap on concurrently optimised registrations; precision: 'fp32c' or 'fp64';
    keep_momenta: results carry the final momenta (device tensors) for checking;
    certify_tol: the precision guard's bound on the FP32-compensated energy error (None:
    no guard; FP64 service: unused).

    Active registrations are split into two cohorts that take turns: a turn collects the
    answers of one cohort's batched requests (launched on its previous turn), advances its
    optimisers and launches their next requests, while the GPU is still working on the
    other cohort's launches.  Each registration's own sequence of operations is unchanged,
    so its result is the same as alone.

    A request whose inputs cannot be registered (unknown kind, mismatched shapes, face
    indices outside the vertex range, options outside the GPU path) is not admitted: its
    Result has energy NaN, no iterations and stop_reason 'rejected: <error>'.
    """

    def __init__(self, hbm_cap_bytes: float = 200e9, max_active: int = 256,
                 precision: str = 'fp32c', clock: Callable[[], float] = time.perf_counter,
                 keep_momenta: bool = False,
                 certify_tol: Optional[float] = DEFAULT_CERTIFY_TOL):
        self.keep_momenta = bool(keep_momenta)
        self.certify_tol = certify_tol if precision == 'fp32c' else None
        self.promoted = 0      # registrations re-run in FP64 by the precision guard
        self.cap = float(hbm_cap_bytes)
        self.max_active = int(max_active)
        self.precision = precision
        self.clock = clock
        self.queue: collections.deque = collections.deque()
        self.cohorts = [[], []]
        self.inflight = [[], []]   # per cohort: [(registrations, pending answers)]
        self._turn = 0
        self.results: dict = {}
        self.used = 0.0
        self.rounds = 0        # turns that launched work
        self.answered = 0      # optimiser requests answered (summed over turns)
        self.max_batch = 0     # most registrations launched in one turn
        self.t0 = None

    @property
    def active(self) -> list:
        return self.cohorts[0] + self.cohorts[1]

    def submit(self, req: Request) -> None:
        if footprint_bytes(req) > self.cap:
            raise ValueError(f"request {req.rid} needs more device memory than the admission cap")
        self.queue.append(req)

    # ---------------------------------------------------------------------------------
    def _now(self) -> float:
        return self.clock() - self.t0

    def _admit(self) -> None:
        now = self._now()
        while self.queue and self.queue[0].arrival <= now and \
                len(self.cohorts[0]) + len(self.cohorts[1]) < self.max_active:
            req = self.queue[0]
            need = footprint_bytes(req)
            if self.used + need > self.cap and (self.cohorts[0] or self.cohorts[1]):
                break
            self.queue.popleft()
            try:
                active = _Active(req, build_registration(req, self.precision), now,
                                 self.precision)
            except _REJECT as exc:
                # Invalid inputs fail this request only; the others keep being served.
                self.results[req.rid] = Result(req.rid, float('nan'), 0,
                                               f"rejected: {type(exc).__name__}: {exc}",
                                               req.arrival, now, now)
                continue
            self.used += need
            c = 0 if len(self.cohorts[0]) <= len(self.cohorts[1]) else 1
            self.cohorts[c].append(active)

    def _finish(self, a: _Active, e64=None) -> None:
        o = a.opt
        if e64 is None:
            e64 = float(o.obj) if a.precision == 'fp64' else float('nan')
        self.results[a.req.rid] = Result(a.req.rid, float(o.obj), len(o.history) - 1, o.stop_reason,
                                         a.req.arrival, a.admitted, self._now(), list(o.history),
                                         o.at if self.keep_momenta else None, a.precision,
                                         float(e64), a.err32)
        self.used -= footprint_bytes(a.req)

    def _stopped(self, a: _Active) -> bool:
        """The optimiser of ``a`` stopped: True when ``a`` is done; False when it stays
        active (its FP64 certification is launched next, or its FP64 re-run started)."""
        if a.precision == 'fp32c' and self.certify_tol is not None:
            a.twin = fp64_twin(a.req, a.opt)
            a.pending = (CERTIFY,)
            return False
        self._finish(a)
        return True

    def _certified(self, a: _Active, e64: float) -> bool:
        """The FP64 objective at the final iterate of ``a``'s FP32-compensated run is in:
        finish within the bound, else re-run in FP64 (True when ``a`` is done)."""
        a.twin = None
        a.err32 = _rel(float(a.opt.obj), e64)
        if a.err32 <= self.certify_tol:
            self._finish(a, e64)
            return True
        self.promoted += 1
        a.start(build_registration(a.req, 'fp64'), 'fp64')
        return False

    def step(self) -> None:
        """One turn of the next cohort: answer its requests launched on its previous turn,
        advance its optimisers, launch their next requests."""
        c = self._turn
        self._turn ^= 1
        answers = {}
        for regs, pend in self.inflight[c]:
            for a, v in zip(regs, pend.result()):
                answers[id(a)] = v
        self.inflight[c] = []
        self.answered += len(answers)
        still = []
        for a in self.cohorts[c]:
            if id(a) not in answers:          # admitted since the last turn: nothing to answer
                still.append(a)
                continue
            if a.pending[0] == CERTIFY:
                if not self._certified(a, answers[id(a)]):
                    still.append(a)
                continue
            try:
                a.pending = a.gen.send(answers[id(a)])
                # Requests without device work are answered at once, not in a later turn.
                while a.pending[0] in _HOST_ONLY:
                    a.pending = a.gen.send(getattr(a.opt, a.pending[0])())
                still.append(a)
            except StopIteration:
                if not self._stopped(a):
                    still.append(a)
        self.cohorts[c] = still
        if not still:
            return
        groups = collections.defaultdict(list)
        for a in still:
            groups[a.pending[0]].append(a)
        for kind in (cg.TRY, cg.GRADIENT, cg.DOT, cg.OBJECTIVE, CERTIFY):
            regs = groups.get(kind)
            if regs:
                self.inflight[c].append((regs, _LAUNCH[kind](regs)))
        self.max_batch = max(self.max_batch, len(still))
        self.rounds += 1

    def run(self, idle_sleep: float = 1e-4, progress=None, progress_every: float = 30.0) -> dict:
        """Serve until every submitted request is finished; returns {rid: Result}.
        ``progress(finished, active, queued)`` is called about every ``progress_every``
        seconds of the service clock."""
        if self.t0 is None:
            self.t0 = self.clock()
        next_report = progress_every
        while self.queue or self.cohorts[0] or self.cohorts[1]:
            if progress is not None and self._now() >= next_report:
                progress(len(self.results), len(self.active), len(self.queue))
                next_report += progress_every
            self._admit()
            if not (self.cohorts[0] or self.cohorts[1]):
                wait = self.queue[0].arrival - self._now()
                if wait > 0:
                    time.sleep(min(wait, 0.05) if wait > idle_sleep else idle_sleep)
                continue
            self.step()
        torch.cuda.synchronize()
        return self.results


def poisson_arrivals(n: int, rate: float, seed: int = 0, burst_factor: float = 1.0,
                     burst_prob: float = 0.0) -> np.ndarray:
    """Arrival times of a Poisson process of ``rate`` requests/s; with burst_prob > 0 each gap
    is shortened by ``burst_factor`` with that probability (bursty arrivals)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    if burst_prob > 0:
        gaps = np.where(rng.random(n) < burst_prob, gaps / burst_factor, gaps)
    return np.cumsum(gaps)


def plain_run(req: Request, precision: str,
              certify_tol: Optional[float] = DEFAULT_CERTIFY_TOL) -> Result:
    """One request through the plain single-registration path (reference call pattern), with
    the service's precision guard (FP32-compensated runs only; None: no guard)."""
    t = time.perf_counter()
    opt = build_registration(req, precision)
    opt.optimizeMatching()
    used, e64, err32 = precision, float('nan'), float('nan')
    if precision == 'fp64':
        e64 = float(opt.obj)
    elif certify_tol is not None:
        e64 = fp64_twin(req, opt).objectiveFun()
        err32 = _rel(float(opt.obj), e64)
        if err32 > certify_tol:
            used = 'fp64'
            opt = build_registration(req, 'fp64')
            opt.optimizeMatching()
            e64 = float(opt.obj)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return Result(req.rid, float(opt.obj), len(opt.history) - 1, opt.stop_reason, 0.0, 0.0,
                  t1 - t, list(opt.history), None, used, e64, err32)
