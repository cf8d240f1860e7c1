"""Device plumbing around the kernel library: workspace, streams, descriptor upload, launches.

PyTorch-ROCm provides device memory and streams only; all arithmetic of the kernel sums runs
in liblddmm_hip.so.  A launch is an array of ``lddmm_seg_t`` descriptors (one independent
kernel sum each), held here as a NumPy structured array (``SEG_DT``) so that the descriptors
of many registrations and time steps are built with vectorised pointer arithmetic.  ``Plan``
plans a launch on the host (C), uploads the planned descriptors through pinned memory and
runs it on the current stream; ``run_steps`` plans a sequence of same-size launches once and
uploads all of their descriptors in one transfer.  Device buffers freed here are reused by
the caching allocator only by later work on the same stream, after the launches that read
them.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import Seg

SEG_DT = np.dtype(Seg)
SEG_SIZE = SEG_DT.itemsize
PLAN_FIELDS = ("nx0", "nx1", "item0", "part0", "fin0", "red0", "row_tiles", "chunks0",
               "chunks1", "chunk_len", "fin_blocks", "stage0")
_SEG_PTR = ctypes.POINTER(Seg)


def _byte_ranges(fields):
    """Contiguous byte ranges of a set of fields of SEG_DT."""
    spans = sorted((SEG_DT.fields[f][1], SEG_DT.fields[f][1] + SEG_DT.fields[f][0].itemsize)
                   for f in fields)
    out = [list(spans[0])]
    for lo, hi in spans[1:]:
        if lo == out[-1][1]:
            out[-1][1] = hi
        else:
            out.append([lo, hi])
    return [tuple(r) for r in out]


_PLAN_BYTES = _byte_ranges(PLAN_FIELDS)

_workspaces: dict = {}

# Pair counters per operation code (target rows x source rows summed over launched segments);
# read by the benchmark to convert times into pair rates.
pair_counts: dict = {}
COUNT_PAIRS = False


_has_gpu = None


def device() -> torch.device:
    global _has_gpu
    if _has_gpu is None:
        _has_gpu = torch.cuda.is_available()
    if not _has_gpu:
        raise RuntimeError("registration_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    """Raw handle of the current stream of the current device (what torch launches on)."""
    return torch._C._cuda_getCurrentRawStream(torch.cuda.current_device())


def workspace(nbytes: int) -> torch.Tensor:
    """Per-device scratch buffer, grown geometrically; reuse is stream-ordered."""
    dev = device()
    ws = _workspaces.get(dev.index)
    if ws is None or ws.numel() < nbytes:
        size = max(int(nbytes), 1 << 20)
        if ws is not None:
            size = max(size, 2 * ws.numel())
        ws = torch.empty(size, dtype=torch.uint8, device=dev)
        _workspaces[dev.index] = ws
    return ws


def ptr(t) -> int:
    if t is None:
        return 0
    if isinstance(t, torch.Tensor):
        return t.data_ptr()
    return int(t)


def as_dev(a, dtype=torch.float64) -> torch.Tensor:
    """A contiguous device tensor of ``dtype`` holding ``a`` (numpy array or tensor)."""
    dev = device()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def segs(n: int) -> np.ndarray:
    """n zeroed descriptors."""
    return np.zeros(n, SEG_DT)


def seg(**kw) -> np.ndarray:
    """One descriptor (a length-1 record array) from keyword fields; tensors become pointers.
    The record keeps raw device addresses only: the caller keeps the tensors referenced until
    the launch has run (a tensor freed earlier can be handed to a later allocation, e.g. the
    descriptor upload, which then overwrites it first)."""
    r = segs(1)
    for k, v in kw.items():
        if k in ("tw", "sw", "g"):
            for i, t in enumerate(v):
                r[k][0, i] = ptr(t)
        elif isinstance(v, torch.Tensor) or v is None:
            r[k] = ptr(v)
        else:
            r[k] = v
    return r


def stack(rows, dtype=SEG_DT) -> np.ndarray:
    """Concatenate cached records kept as raw bytes (``raw(record)``) into one record array.
    Concatenating structured arrays directly promotes every field of every operand, which
    costs tens of microseconds per record."""
    return np.concatenate(rows).view(dtype)


def raw(record: np.ndarray) -> np.ndarray:
    """The bytes of a (1,) record array, for ``stack``."""
    return np.ascontiguousarray(record).view(np.uint8).reshape(-1).copy()


def _as_records(s) -> np.ndarray:
    if isinstance(s, np.ndarray):
        return np.ascontiguousarray(s.reshape(-1))
    return stack([np.ascontiguousarray(np.asarray(x).reshape(-1)).view(np.uint8) for x in s])


def upload(recs: np.ndarray) -> torch.Tensor:
    """Copy descriptor records to the device through pinned memory (async on the current stream)."""
    raw = np.ascontiguousarray(recs).view(np.uint8).reshape(-1)
    host = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True)
    host.numpy()[:] = raw
    return host.to(device(), non_blocking=True)


def _count(op, recs):
    if COUNT_PAIRS:
        n = int((recs["ny"].astype(np.int64) * (recs["nx0"] + recs["nx1"])).sum())
        pair_counts[op] = pair_counts.get(op, 0) + n


class Plan:
    """A planned segmented launch (host descriptors + device copy + sizes)."""

    __slots__ = ("op", "prec", "d", "recs", "nseg", "n_items", "n_fin", "work_bytes", "dev")

    def __init__(self, op: int, prec: int, d: int, recs):
        lib = _lib.load()
        self.op, self.prec, self.d = op, prec, d
        self.recs = _as_records(recs)
        self.nseg = self.recs.shape[0]
        n_items = ctypes.c_int64()
        n_fin = ctypes.c_int64()
        wb = ctypes.c_size_t()
        _lib.check(lib.lddmm_segments_plan(op, prec, d, self.recs.ctypes.data_as(_SEG_PTR), self.nseg,
                                           ctypes.byref(n_items), ctypes.byref(n_fin), ctypes.byref(wb)),
                   "lddmm_segments_plan")
        self.n_items, self.n_fin, self.work_bytes = n_items.value, n_fin.value, wb.value
        self.dev = None

    def run(self) -> None:
        lib = _lib.load()
        ws = workspace(self.work_bytes)
        if self.dev is None:
            self.dev = upload(self.recs)
        _count(self.op, self.recs)
        _lib.check(lib.lddmm_segments_run(self.op, self.prec, self.d, self.recs.ctypes.data_as(_SEG_PTR),
                                          self.dev.data_ptr(), self.nseg, self.n_items, self.n_fin,
                                          ws.data_ptr(), ws.numel(), stream_handle()),
                   "lddmm_segments_run")


def launch(op: int, prec: int, d: int, recs) -> Plan:
    p = Plan(op, prec, d, recs)
    p.run()
    return p


def run_steps(op: int, prec: int, d: int, steps: np.ndarray) -> None:
    """Run launches steps[0], steps[1], ... in order (steps: (S, n) records of equal sizes).

    Scheduling fields depend only on a segment's sizes, so step 0 is planned once and its
    plan copied to the other steps; every step's descriptors go up in one transfer.
    """
    S, n = steps.shape
    plan = Plan(op, prec, d, steps[0])
    steps = np.ascontiguousarray(steps)
    # Copy step 0's planned fields into every step as two byte ranges.
    raw_steps = steps.view(np.uint8).reshape(S, n, SEG_SIZE)
    raw_plan = plan.recs.view(np.uint8).reshape(n, SEG_SIZE)
    for lo, hi in _PLAN_BYTES:
        raw_steps[1:, :, lo:hi] = raw_plan[None, :, lo:hi]
    raw_steps[0] = raw_plan
    dev = upload(steps)
    lib = _lib.load()
    ws = workspace(plan.work_bytes)
    _count(op, steps)
    _lib.check(lib.lddmm_segments_run_steps(op, prec, d, steps.ctypes.data_as(_SEG_PTR),
                                            dev.data_ptr(), S, n, plan.n_items, plan.n_fin,
                                            ws.data_ptr(), ws.numel(), stream_handle()),
               "lddmm_segments_run_steps")
