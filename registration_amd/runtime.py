"""Device plumbing around the kernel library: workspace, streams, descriptor upload, launches.

PyTorch-ROCm provides device memory and streams only; all arithmetic of the kernel sums runs
in liblddmm_hip.so.  A launch is a list of ``Seg`` descriptors (one independent kernel sum
each).  ``launch`` plans it on the host (C), uploads the planned descriptors through a
pinned staging buffer and runs it on the current stream.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import Seg

_workspaces: dict = {}

# Pair counters per operation code (target rows x source rows summed over launched segments);
# read by the benchmark to convert times into pair rates.
pair_counts: dict = {}
COUNT_PAIRS = False


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("registration_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


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


def ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def as_dev(a, dtype=torch.float64) -> torch.Tensor:
    """A contiguous device tensor of ``dtype`` holding ``a`` (numpy array or tensor)."""
    dev = device()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def seg(**kw) -> Seg:
    s = Seg()
    for k, v in kw.items():
        if k in ("tw", "sw", "g"):
            arr = getattr(s, k)
            for i, t in enumerate(v):
                arr[i] = ptr(t) if isinstance(t, torch.Tensor) else t
        elif isinstance(v, torch.Tensor):
            setattr(s, k, v.data_ptr())
        else:
            setattr(s, k, v)
    return s


def upload(arr) -> torch.Tensor:
    """Copy a ctypes array to the device through pinned memory (async on the current stream)."""
    nbytes = ctypes.sizeof(arr)
    host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    ctypes.memmove(host.data_ptr(), ctypes.addressof(arr), nbytes)
    return host.to(device(), non_blocking=True)


class Plan:
    """A planned segmented launch (host descriptors + device copy + sizes)."""

    __slots__ = ("op", "prec", "d", "segs", "nseg", "n_items", "n_fin", "work_bytes", "dev")

    def __init__(self, op: int, prec: int, d: int, segs: Sequence[Seg]):
        lib = _lib.load()
        self.op, self.prec, self.d = op, prec, d
        self.nseg = len(segs)
        self.segs = (Seg * self.nseg)(*segs)
        n_items = ctypes.c_int64()
        n_fin = ctypes.c_int64()
        wb = ctypes.c_size_t()
        _lib.check(lib.lddmm_segments_plan(op, prec, d, self.segs, self.nseg, ctypes.byref(n_items),
                                           ctypes.byref(n_fin), ctypes.byref(wb)),
                   "lddmm_segments_plan")
        self.n_items, self.n_fin, self.work_bytes = n_items.value, n_fin.value, wb.value
        self.dev = None

    def run(self, dev_segs: torch.Tensor | None = None) -> None:
        lib = _lib.load()
        ws = workspace(self.work_bytes)
        if dev_segs is None:
            dev_segs = self.dev if self.dev is not None else upload(self.segs)
        self.dev = dev_segs
        if COUNT_PAIRS:
            n = 0
            for i in range(self.nseg):
                sg = self.segs[i]
                n += sg.ny * (sg.nx0 + sg.nx1)
            pair_counts[self.op] = pair_counts.get(self.op, 0) + n
        _lib.check(lib.lddmm_segments_run(self.op, self.prec, self.d, self.segs,
                                          None if dev_segs is None else dev_segs.data_ptr(),
                                          self.nseg, self.n_items, self.n_fin, ws.data_ptr(),
                                          ws.numel(), stream_handle()),
                   "lddmm_segments_run")


def launch(op: int, prec: int, d: int, segs: Sequence[Seg]) -> Plan:
    p = Plan(op, prec, d, segs)
    p.run()
    return p


_PLAN_FIELDS = ("nx0", "nx1", "item0", "part0", "fin0", "red0", "row_tiles", "chunks0",
                "chunks1", "chunk_len", "fin_blocks")


def replicate(plan: Plan, steps: Sequence[Sequence[Seg]]) -> list[Plan]:
    """Plans for a sequence of launches that have the sizes of ``plan``.

    ``steps[t]`` holds the segments of launch t (same sizes as ``plan.segs``, other
    pointers).  The scheduling fields of a segment depend only on its sizes, so a time loop
    plans once and copies them; all descriptor arrays go up in one transfer.
    """
    n = plan.nseg
    esz = ctypes.sizeof(Seg)
    big = (Seg * (n * len(steps)))()
    for t, segs in enumerate(steps):
        if len(segs) != n:
            raise ValueError("replicate: every step needs the planned number of segments")
        for i, sg in enumerate(segs):
            ref = plan.segs[i]
            for f in _PLAN_FIELDS:
                setattr(sg, f, getattr(ref, f))
            big[t * n + i] = sg
    dev = upload(big)
    out = []
    for t in range(len(steps)):
        p = Plan.__new__(Plan)
        p.op, p.prec, p.d, p.nseg = plan.op, plan.prec, plan.d, n
        p.segs = (Seg * n).from_buffer(big, t * n * esz)
        p.n_items, p.n_fin, p.work_bytes = plan.n_items, plan.n_fin, plan.work_bytes
        p.dev = dev[t * n * esz:(t + 1) * n * esz]
        out.append(p)
    return out
