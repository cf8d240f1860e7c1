"""Chunking-target autotuner of the segmented K-sum launches (no counterpart in the reference,
whose sums are one OpenMP loop over rows, py-lddmm/kernelFunctions.f90:1-54).

lddmm_segments_plan splits a segment of ny rows into (row tiles) x (source chunks) work items,
choosing the chunk length so that their number is near a target.  Too few items leave CUs
idle; too many cost chunk partials (written by the pair kernel, read back by the finalize)
and per-item start-up.  The best target depends on the operation, the precision, whether the
sum is energy-only (symmetric self tiles, uneven items) and the problem size.

``tune`` times one launch of a representative problem per (operation, precision, energy-only
form, size class) at each candidate target with HIP events on the launch stream, and keeps
the fastest when it beats the built-in target by more than ``min_gain``; ``write`` stores
the table (with every measured time) as JSON, which ``_lib.load`` applies through
lddmm_set_items_target.  Run on the GPU the library will serve on:

    python -m registration_amd.autotune --out registration_amd/tuning_gfx950.json

Representative problems, one per size class: icosphere vertices (subdivisions 3-7: 642,
2,562, 10,242, 40,962 and 163,842 points) and the face centres of subdivision 7 (327,680)
for the flow and covector steps, with a Gaussian deformation kernel of width 0.25 on the
unit sphere, in the spatial processing order the registrations use; icosphere faces
(subdivisions 2-7: 320 to 327,680) for the currents sums, width 0.1.  Each launch carries the
dispatch order of its work items that a registration's launch of the same plan carries
(engine.item_order): with the longest items first the tail of a launch is short, which
moves the best target towards fewer, larger items.

An entry may carry ``chosen_by``: a target set from whole registrations timed against each
other where one launch's gain stayed under ``min_gain`` (the large surface pair's flow and
covector steps: 1.4-1.7% per launch, 1.6% per registration).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import _lib, engine, runtime as rt, surfaces, synthetic
from .kernelFunctions import Kernel

CANDIDATES = (768, 1536, 3072, 4096, 6144, 8192, 12288, 16384, 24576, 49152)
# Icosphere subdivision whose vertices (flow, covector) represent each size class; 8 stands
# for the 327,680 face centres of subdivision 7.
CLASS_SUBDIV = (3, 4, 5, 6, 7, 8)
# (operation, precision, energy-only) forms tuned: the registration's hot launches.  The
# allocation-free applyK entry point keeps the built-in targets (its scratch bound).
FORMS = (("FLOW_STEP", "fp32c", False), ("FLOW_STEP", "fp64", False),
         ("COVECTOR_STEP", "fp32c", False), ("COVECTOR_STEP", "fp64", False),
         ("CURRENTS", "fp64", True), ("CURRENTS", "fp32c", False))


def _points(cls):
    """Points of a flow or covector problem of size class ``cls`` (n, 3)."""
    s = CLASS_SUBDIV[cls]
    V, F = synthetic.icosphere(min(s, 7))
    if s <= 7:
        return V
    c = V[F].mean(axis=1)
    return c / np.linalg.norm(c, axis=1)[:, None]


def _runner(op, prec, energy_only, cls):
    """A function that plans and runs one launch of the form (plans read the targets in
    force when they are made), with its buffers kept alive."""
    dev = "cuda"
    rng = np.random.default_rng(cls)
    if op == "CURRENTS":
        V, F = synthetic.icosphere(cls + 2)
        S = surfaces.DeviceSurface(F, torch.as_tensor(V, device=dev), with_weights=False)
        Y = torch.as_tensor(V @ synthetic.random_rotation(rng).T, device=dev)
        T = surfaces.DeviceSurface(F, Y, with_weights=False)
        nf = F.shape[0]
        tot = torch.empty(1, dtype=torch.float64, device=dev)
        fields = dict(y=S.centers, ny=nf, x1=T.centers, nx1=nf, w0=S.surfel, w1=T.surfel,
                      sigma=0.1, coef=1.0 / 3.0, kernel=_lib.KERNEL_GAUSS, order=10, total=tot)
        keep = [S, T, tot]
        hosts = (S.centers.cpu().numpy(), T.centers.cpu().numpy(), None)
        if not energy_only:
            z1 = torch.empty((nf, 3), dtype=torch.float64, device=dev)
            dz = torch.empty((nf, 3), dtype=torch.float64, device=dev)
            fields.update(out0=z1, out1=dz)
            keep += [z1, dz]
        rows = nf
    else:
        K = Kernel('gauss', sigma=0.25, precision=prec)
        V = _points(cls)
        N = V.shape[0]
        x = torch.as_tensor(V, device=dev)
        a = torch.as_tensor(1e-2 * rng.standard_normal(V.shape), device=dev)
        out = torch.empty_like(x)
        tot = torch.empty(1, dtype=torch.float64, device=dev)
        perm = engine.order_for(V, engine.DevKernel.of(K))
        fields = dict(y=x, w0=a, out0=out, ny=N, sigma=K.sigma, dt=0.05, kernel=K.kernel_code,
                      order=K.order)
        if op == "FLOW_STEP":
            fields["total"] = tot
        else:
            px = torch.as_tensor(rng.standard_normal(V.shape), device=dev)
            fields.update(w1=px, coef=1.0)
            keep_px = px
        if perm is not None:
            fields["perm"] = perm
        keep = [x, a, out, tot, perm] + ([keep_px] if op != "FLOW_STEP" else [])
        hosts = (V, None, None if perm is None else perm.cpu().numpy())
        rows = N
    rec = rt.seg(**fields)
    code = _lib.OP_IDS[op]
    pc = _lib.PRECS[prec]

    def run(reps):
        # reps launches from one library call (lddmm_segments_run_steps), so that small
        # problems are timed on the GPU, not on the host's launch loop; with the dispatch
        # order a registration's launch of this plan carries (the plan follows the target)
        r = np.array(rec, copy=True)
        order = engine.item_order(code, pc, 3, r, hosts[0], x1=hosts[1], perm=hosts[2])
        engine.attach_item_order(r, order)
        steps = np.repeat(np.asarray(r).reshape(1, 1), reps, axis=0)
        rt.run_steps(code, pc, 3, steps[:2])
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        rt.run_steps(code, pc, 3, steps)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, rt.Plan(code, pc, 3, [r]).n_items
    run.keep = keep
    return run, rows


def tune(forms=FORMS, classes=range(6), candidates=CANDIDATES, min_gain=0.02,
         budget_ms=150.0, progress=print) -> dict:
    """The table of chosen targets, with every measured time (ms per launch)."""
    lib = _lib.load()
    entries = []
    for op, prec, eo in forms:
        code, pc = _lib.OP_IDS[op], _lib.PRECS[prec]
        for cls in classes:
            run, rows = _runner(op, prec, eo, cls)
            assert lib.lddmm_size_class(rows) == cls, (rows, cls)
            lib.lddmm_set_items_target(code, pc, int(eo), cls, 0)
            builtin = int(lib.lddmm_get_items_target(code, pc, int(eo), cls))
            ms0, _ = run(1)
            reps = int(max(3, min(500, budget_ms / max(ms0, 1e-3))))
            times, items = {}, {}
            for c in sorted(set(candidates) | {builtin}):
                lib.lddmm_set_items_target(code, pc, int(eo), cls, c)
                times[c], items[c] = run(reps)
            lib.lddmm_set_items_target(code, pc, int(eo), cls, 0)
            best = min(times, key=times.get)
            chosen = best if times[best] < (1.0 - min_gain) * times[builtin] else builtin
            entries.append({"op": op, "prec": prec, "energy_only": eo, "size_class": cls,
                            "rows": rows, "items": chosen, "builtin": builtin,
                            "ms": {str(k): round(v, 5) for k, v in times.items()},
                            "work_items": {str(k): int(v) for k, v in items.items()},
                            "gain_vs_builtin": round(1.0 - times[chosen] / times[builtin], 4)})
            progress(f"autotune {op} {prec}{' energy' if eo else ''} class {cls} ({rows} rows): "
                     f"{builtin} -> {chosen} ({times[builtin]:.4f} -> {times[chosen]:.4f} ms)")
            del run
    return {"device": torch.cuda.get_device_name(0), "made": time.strftime("%Y-%m-%d %H:%M:%S"),
            "candidates": list(candidates), "min_gain": min_gain, "entries": entries}


def write(table: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def merge(old: dict, new: dict) -> dict:
    """``old``'s table with the entries of the forms and classes ``new`` measured replaced."""
    key = lambda e: (e["op"], e["prec"], bool(e["energy_only"]), int(e["size_class"]))
    fresh = {key(e): e for e in new["entries"]}
    kept = [e for e in old["entries"] if key(e) not in fresh]
    out = dict(new)
    out["entries"] = sorted(kept + list(fresh.values()),
                            key=lambda e: (next(i for i, f in enumerate(FORMS)
                                                if f == (e["op"], e["prec"], bool(e["energy_only"]))),
                                           e["size_class"]))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=_lib.TUNING_PATH)
    ap.add_argument("--classes", default="0,1,2,3,4,5")
    ap.add_argument("--forms", default="",
                    help="op:prec[:energy],... of FORMS to measure (default all); the other "
                         "entries of an existing --out table are kept")
    args = ap.parse_args(argv)
    import os
    os.environ["LDDMM_TUNING"] = "off"   # measure from the built-in targets
    forms = FORMS
    if args.forms:
        want = set()
        for f in args.forms.split(","):
            parts = f.split(":")
            want.add((parts[0], parts[1], len(parts) > 2 and parts[2] == "energy"))
        forms = tuple(f for f in FORMS if f in want)
        if len(forms) != len(want):
            raise SystemExit(f"unknown form in {args.forms!r}; forms: {FORMS}")
    torch.cuda.set_device(0)
    table = tune(forms=forms, classes=[int(c) for c in args.classes.split(",")])
    if args.forms and os.path.exists(args.out):
        with open(args.out) as f:
            table = merge(json.load(f), table)
    write(table, args.out)
    print(f"wrote {len(table['entries'])} targets to {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
