#!/usr/bin/env python3
"""Benchmark: LDDMM surface registrations per second on one MI355X.

Workload (default): the single large surface pair -- template icosphere subdivision 6
(40,962 vertices, 81,920 triangles), target = template displaced by a seeded sum of 20
Gaussian RBFs (width 0.4, amplitudes N(0, 0.1^2)); Gaussian deformation kernel sigma 0.25,
Gaussian currents kernel sigma 0.1, sigma_err 0.1, T = 20 Euler steps, up to 100 conjugate-
gradient iterations, momenta initialised to zero.  One step = one whole registration
(SurfaceMatching + optimizeMatching) with the template and target already on the device.

Printed JSON line: registrations/s, per-registration latency, the live roofline of the
dominant kernel (the fused Euler flow step: Gaussian K(x,x)a over 40,962 x 40,962 pairs, 16
flops per pair), its energy error against the FP64 form, and the FP64 CPU oracle timed on a
bounded sample of the same sums on this host.
"""
from __future__ import annotations

import argparse
import collections
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (AMD published)
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64 (AMD published)
FLOPS_PER_PAIR = {"apply_k_gauss_d3_m3": 16}
# HBM bytes per fused flow step at 40,962 points (tile staging + pair kernel + finalize + totals),
# measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on `bench.py --roofline-only
# --roofline-reps 3` and scaled by the counters' reading for 256 MiB device copies (FETCH x2.0,
# WRITE x1.0), in the template's spatial processing order with the items longest first, at
# the 8,192-item plan (6,561 items): 138.4 MB read + 95.7 MB written, most of it the FP64 chunk
# partials (81 chunks x 40,962 rows x 3), written by the pair kernel and read back by the
# finalize (profiles/flow_step_s6_fp32c.md).
FLOW_STEP_TRAFFIC_S6 = 234.1e6


def _warm(plan, seconds=0.25):
    """Run a planned launch until ``seconds`` have passed: the GPU's clocks ramp up over the
    first ~100 ms of work, and a leg that starts cold would time the ramp."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            plan.run()
        torch.cuda.synchronize()


def _progress(msg):
    """One progress line per finished leg on stderr (the JSON result stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32c", choices=["fp32c", "fp64"])
    ap.add_argument("--workload", default="surface-s6", choices=["surface-s6", "surface-s3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-reps", type=int, default=20)
    ap.add_argument("--served", type=int, default=16,
                    help="requests of the 512-request served-surfaces stream to serve (0 = skip)")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="requests of the served-surfaces stream also timed on the plain path "
                         "(0 = all that are served)")
    ap.add_argument("--small-pairs", type=int, default=512,
                    help="small surface pairs served all at once (0 = skip)")
    ap.add_argument("--mixed-service", type=int, default=20,
                    help="requests of the mixed-service stream to serve (0 = skip)")
    ap.add_argument("--data-term-reps", type=int, default=10,
                    help="launches per data-term kernel in the data_terms leg (0 = skip)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the flow-step launches and a calibration copy (for rocprof passes)")
    ap.add_argument("--energy-only", action="store_true",
                    help="only the FP64 energy launches of the data_terms leg (for rocprof "
                         "passes)")
    ap.add_argument("--no-fp64-form", action="store_true",
                    help="skip the FP64-form registration (profiled runs)")
    ap.add_argument("--graph-trial", action="store_true",
                    help="only the HIP-graph trial comparison (graph_trial_leg)")
    ap.add_argument("--cpu-only", metavar="BENCH_JSON",
                    help="only the CPU baseline, extrapolated with the per-operation pair "
                         "counts of an earlier bench line (no GPU used)")
    return ap.parse_args()


def make_workload(name):
    from registration_amd import synthetic
    cfg = synthetic.surface_pair_config2() if name == "surface-s6" else synthetic.surface_pair_config1()
    cfg["name"] = name
    return cfg


def registration(cfg, prec):
    from registration_amd import kernelFunctions as kfun, surfaceMatching as sm, surfaces
    p = sm.SurfaceMatchingParam(timeStep=1.0 / cfg["T"],
                                KparDiff=kfun.Kernel('gauss', sigma=cfg["sigmaKernel"], precision=prec),
                                KparDist=kfun.Kernel('gauss', sigma=cfg["sigmaDist"], precision=prec),
                                sigmaError=cfg["sigmaError"], errorType='current')
    F, V = cfg["template"]
    _, Y = cfg["target"]
    return sm.SurfaceMatching(outputDir=None, Template=surfaces.Surface(FV=(F, V)), Target=surfaces.Surface(FV=(F, Y)),
                              param=p, maxIter=cfg["maxIter"], testGradient=False)


# The FP64 oracle's whole registration of the large surface pair (the reference's
# optimizeMatching path restated on the CPU), written by tests/golden/make_config2_oracle.py.
CONFIG2_ORACLE = os.path.join(ROOT, "tests", "golden", "config2_oracle.npz")


def fp64_form_leg(cfg, final_fp32c):
    """One registration of the workload in the FP64 form (every kernel sum in FP64), timed
    like the headline ones, and the FP32-compensated run's final energy against its final
    energy and, where its record is present, the oracle's."""
    import numpy as np
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m = registration(cfg, "fp64")
    m.optimizeMatching()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    out = {"seconds": secs, "registrations_per_s": 1.0 / secs, "final_energy": float(m.obj),
           "cg_iterations": len(m.history) - 1, "stop_reason": m.stop_reason,
           "fp32c_final_energy_rel_err": abs(final_fp32c - m.obj) / abs(m.obj)}
    if cfg.get("name") == "surface-s6" and os.path.exists(CONFIG2_ORACLE):
        ref = np.load(CONFIG2_ORACLE)
        e = float(ref["history"][-1])
        out.update({"oracle_final_energy": e,
                    "fp64_final_energy_rel_err_vs_oracle": abs(m.obj - e) / abs(e),
                    "fp32c_final_energy_rel_err_vs_oracle": abs(final_fp32c - e) / abs(e),
                    "oracle_stop_reason": str(ref["stop_reason"])})
    return out


def flow_step_roofline(cfg, prec, reps, natural=True):
    """Average duration of the fused flow-step launch at the workload's size, in the
    processing order the registration uses (the template's spatial order), timed with HIP
    events on the stream the launches go to; the natural vertex order is timed beside it
    unless ``natural`` is False (profiled runs, which then hold one launch form only)."""
    import numpy as np
    import torch
    from registration_amd import _lib, engine, runtime as rt
    from registration_amd.kernelFunctions import Kernel
    F, V = cfg["template"]
    N = V.shape[0]
    rng = np.random.default_rng(0)
    x = torch.as_tensor(V, device="cuda")
    a = torch.as_tensor(1e-2 * rng.standard_normal(V.shape), device="cuda")
    out = torch.empty_like(x)
    tot = torch.empty(1, dtype=torch.float64, device="cuda")
    K = Kernel('gauss', sigma=cfg["sigmaKernel"], precision=prec)
    order = engine.order_for(V, engine.DevKernel.of(K))

    def timed(perm):
        fields = dict(y=x, w0=a, out0=out, total=tot, ny=N, sigma=K.sigma, dt=0.05,
                      kernel=K.kernel_code, order=K.order)
        if perm is not None:
            fields["perm"] = perm
        rec = rt.seg(**fields)
        # the dispatch order a registration's flow steps carry (estimated from the template)
        items = engine.item_order(_lib.OP_FLOW_STEP, K.prec_code, 3, rec, V,
                                  perm=None if perm is None else perm.cpu().numpy())
        engine.attach_item_order(rec, items)
        plan = rt.Plan(_lib.OP_FLOW_STEP, K.prec_code, 3, [rec])
        _warm(plan)  # untimed: the clocks settle over the first launches
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            plan.run()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms = timed(order) if order is not None else timed(None)
    ms_natural = timed(None) if (natural and order is not None) else (ms if order is None else None)
    pairs = N * N
    flops = FLOPS_PER_PAIR["apply_k_gauss_d3_m3"] * pairs
    tflops = flops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
            "traffic": FLOW_STEP_TRAFFIC_S6 if N == 40962 else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
            "kernel": "k_pairs<float,gauss,3,3> + finalize (Euler flow step)",
            "launch_ms": round(ms, 4), "pairs_per_launch": pairs, "flops_per_pair": 16,
            "pairs_per_s": pairs / (ms * 1e-3), "spatial_order": order is not None,
            "launch_ms_natural_order": round(ms_natural, 4) if ms_natural is not None else None}


# Algorithmic flops per pair (add/mul 1, FMA 2, exp/sqrt 1; Gaussian d = 3).  Energies:
# kernel value 3d+1 = 10, plus 2m for the weighted sum (currents m = 3, measure m = 1) or, for
# the varifold, g = <u_k, u_l> (5), g^2 (1), K a_l (1) and a_l K (1 + g^2) accumulated (3).
DATA_TERM_FLOPS = {"currents": 16, "measure": 12, "varifold": 20}
# FP64 FLOP one energy launch evaluates on the large surface pair (81,920 x 81,920 triangle
# pairs per source set), read with rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP64 on
# `bench.py --energy-only --data-term-reps 3` (profiles/currents_energy_s6_fp64.md).  Fewer
# than the nominal pairs x flops: far tiles are culled, each unordered block pair of the self
# sum is evaluated once, and far blocks take the FP32 exponential.
EVALUATED_FP64_FLOP_S6 = {"currents": 5.781e10, "measure": 4.686e10, "varifold": 1.011e11}
# The same counter on the FP64E energies (cull 2^-40, FP32 exponential below 2^-16) of that
# run, which the line-search trials of FP32-compensated registrations evaluate.
EVALUATED_FP64E_FLOP_S6 = {"currents": 4.197e10, "measure": 3.379e10, "varifold": 7.444e10}


def data_terms_leg(cfg, reps, names=("currents", "measure", "varifold"), gradients=True):
    """Per-launch time of the data-term pair kernels on the workload's surface pair (deformed
    template against target, F x F pairs for each of the two source sets), timed with HIP
    events on the launch stream: FP64 energies (what every line-search trial runs) and FP32
    gradient forms, plus the divergence of the Jacobian flow at the template's points."""
    import numpy as np
    import torch
    from registration_amd import _lib, engine, runtime as rt, surfaces
    F, V = cfg["template"]
    _, Y = cfg["target"]
    sig = cfg["sigmaDist"]
    dev = "cuda"
    S = surfaces.DeviceSurface(F, torch.as_tensor(V, device=dev), with_weights=True)
    T = surfaces.DeviceSurface(F, torch.as_tensor(Y, device=dev), with_weights=True)
    nf = F.shape[0]
    tot = torch.empty(1, dtype=torch.float64, device=dev)
    z1 = torch.empty((nf, 3), dtype=torch.float64, device=dev)
    dz = torch.empty((nf, 3), dtype=torch.float64, device=dev)
    base = dict(y=S.centers, ny=nf, x1=T.centers, nx1=nf, sigma=sig, coef=1.0 / 3.0,
                kernel=_lib.KERNEL_GAUSS, order=10, total=tot)
    weights = {
        "currents": dict(w0=S.surfel, w1=T.surfel),
        "measure": dict(w0=S.weights, w1=T.weights, w2=S.surfel),
        "varifold": dict(w0=S.weights, w1=T.weights, w2=S.surfel, sw=[T.surfel]),
    }
    ops = {"currents": _lib.OP_CURRENTS, "measure": _lib.OP_MEASURE, "varifold": _lib.OP_VARIFOLD}

    c1_np, c2_np = S.centers.cpu().numpy(), T.centers.cpu().numpy()
    orders = []

    def ordered(op, prec, rec):
        """rec with the dispatch order a registration's data-term launches carry."""
        o = engine.item_order(op, prec, 3, rec, c1_np, x1=c2_np)
        engine.attach_item_order(rec, o)
        orders.append(o)
        return rec

    def timed(plan):
        _warm(plan)
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            plan.run()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    pairs = 2 * nf * nf
    for name in names:
        op = ops[name]
        ms = timed(rt.Plan(op, _lib.PREC_FP64, 3,
                           [ordered(op, _lib.PREC_FP64, rt.seg(**base, **weights[name]))]))
        # nominal: all 2 F^2 pairs of the two sums (target-target and target-template)
        rec = {"ms": round(ms, 4), "nominal_pairs": pairs,
               "nominal_pairs_per_s": pairs / (ms * 1e-3),
               "flops_per_pair": DATA_TERM_FLOPS[name]}
        flop = EVALUATED_FP64_FLOP_S6.get(name) if cfg.get("name") == "surface-s6" else None
        if flop is not None:
            tf = flop / (ms * 1e-3) / 1e12
            rec.update({"evaluated_fp64_flop": flop, "evaluated_fp64_tflops": round(tf, 3),
                        "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                        "frac_of_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 4)})
        out[f"{name}_energy_fp64"] = rec
        # the same energy in FP64E, the form every line-search trial of an FP32-compensated
        # registration runs (FP64 arithmetic, cull 2^-40, FP32 exponential below 2^-16)
        tot64 = tot.item()
        ms = timed(rt.Plan(op, _lib.PREC_FP64E, 3,
                           [ordered(op, _lib.PREC_FP64E, rt.seg(**base, **weights[name]))]))
        rec = {"ms": round(ms, 4), "nominal_pairs": pairs,
               "nominal_pairs_per_s": pairs / (ms * 1e-3), "total_minus_fp64": tot.item() - tot64}
        flop = EVALUATED_FP64E_FLOP_S6.get(name) if cfg.get("name") == "surface-s6" else None
        if flop is not None:
            tf = flop / (ms * 1e-3) / 1e12
            rec.update({"evaluated_fp64_flop": flop, "evaluated_fp64_tflops": round(tf, 3),
                        "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                        "frac_of_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 4)})
        out[f"{name}_energy_fp64e"] = rec
        if not gradients:
            continue
        ms = timed(rt.Plan(op, _lib.PREC_FP32C, 3,
                           [ordered(op, _lib.PREC_FP32C,
                                    rt.seg(**base, **weights[name], out0=z1, out1=dz))]))
        out[f"{name}_gradient_fp32c"] = {"ms": round(ms, 4), "nominal_pairs": pairs,
                                         "nominal_pairs_per_s": pairs / (ms * 1e-3)}
    if not gradients:
        return out
    # divergence of K(., x) a at the template's points (one Jacobian-flow step)
    x = torch.as_tensor(V, device=dev)
    a = torch.as_tensor(1e-2 * np.random.default_rng(1).standard_normal(V.shape), device=dev)
    div = torch.empty(V.shape[0], dtype=torch.float64, device=dev)
    ms = timed(rt.Plan(_lib.OP_DIVERGENCE, _lib.PREC_FP32C, 3,
                       [rt.seg(y=x, w0=a, out0=div, ny=V.shape[0], sigma=cfg["sigmaKernel"],
                               kernel=_lib.KERNEL_GAUSS, order=10)]))
    n2 = V.shape[0] ** 2
    out["divergence_fp32c"] = {"ms": round(ms, 4), "pairs": n2, "pairs_per_s": n2 / (ms * 1e-3)}
    return out


def _host():
    """The CPU this runs on: model name, logical CPUs, and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}


def _omp_threads(n=None):
    """Set (n given) and return the OpenMP thread count of the CPU sums' parallel loops."""
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    if n is not None:
        gomp.omp_set_num_threads(int(n))
    return int(gomp.omp_get_max_threads())


def _cpu_rates(cfg, rows, sums="oracle"):
    """Pairs/s of the two CPU sum kinds (applyK and applyDiffKT with three momentum sets) on
    the first ``rows`` template points against all of them, and the seconds taken."""
    import numpy as np
    from oracle import lddmm_oracle as orc
    F, V = cfg["template"]
    N = V.shape[0]
    rows = min(N, rows)
    rng = np.random.default_rng(1)
    a = 1e-2 * rng.standard_normal(V.shape)
    a1 = rng.standard_normal((3, rows, 3))
    a2 = rng.standard_normal((3, N, 3))
    t0 = time.perf_counter()
    orc.apply_k(V[:rows], V, a, cfg["sigmaKernel"], 10, sums=sums)
    tk = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.apply_diff_kt(V[:rows], V, a1, a2, cfg["sigmaKernel"], 10, sums=sums)
    td = time.perf_counter() - t0
    return rows * N / tk, rows * N / td, rows, tk + td


def cpu_baseline(cfg, pair_counts):
    """FP64 CPU oracle (C/OpenMP restatement of the reference's Fortran sums) timed on a row
    sample of the workload's two sum kinds, on all usable cores and on one; registrations/s
    extrapolated with the per-operation pair counts of the GPU run.  Where the reference's
    own Fortran sums are built (oracle/_ref), they are timed the same way."""
    from oracle import lddmm_oracle as orc
    from registration_amd import _lib
    N = cfg["template"][1].shape[0]
    k_pairs = sum(v for op, v in pair_counts.items() if op != _lib.OP_COVECTOR_STEP)
    d_pairs = pair_counts.get(_lib.OP_COVECTOR_STEP, 0)
    threads = _omp_threads()
    # Samples of about 10 s of CPU work in all: every row of both sums on all cores (about
    # 3 s at the large pair on 16 cores), 8,192 rows on one core (about 6 s).
    kr, dr, rows, took = _cpu_rates(cfg, N)
    secs = k_pairs / kr + d_pairs / dr
    out = {"value": 1.0 / secs, "unit": "registrations/s", "cores": threads, "kind": "port",
           "sample": f"oracle applyK and applyDiffKT(na=3), {rows} x {N} pairs each "
                     f"({took:.1f} s); extrapolated to the GPU run's {k_pairs + d_pairs:.3e} pairs",
           "extrapolated": True, "host": _host(), "pairs_per_s_applyk": kr,
           "pairs_per_s_diffkt": dr, "seconds_per_registration": secs}
    _omp_threads(1)
    kr1, dr1, rows1, took1 = _cpu_rates(cfg, 8192)
    secs1 = k_pairs / kr1 + d_pairs / dr1
    out["single_thread"] = {"value": 1.0 / secs1, "cores": 1,
                            "sample": f"{rows1} x {N} pairs each ({took1:.1f} s)",
                            "pairs_per_s_applyk": kr1, "pairs_per_s_diffkt": dr1,
                            "seconds_per_registration": secs1}
    _omp_threads(threads)
    if orc.ref_available():
        rk, rd, rr, _ = _cpu_rates(cfg, 2048, sums="reference")
        rsecs = k_pairs / rk + d_pairs / rd
        out["reference"] = {"value": 1.0 / rsecs, "unit": "registrations/s", "cores": threads,
                            "kind": "reference",
                            "sample": f"the reference's Fortran applyk / applykdifft (na=3), "
                                      f"{rr} x {N} pairs each; extrapolated likewise",
                            "pairs_per_s_applyk": rk, "pairs_per_s_diffkt": rd,
                            "seconds_per_registration": rsecs}
    return out


def cpu_only(path, workload):
    """The CPU baseline alone, with the pair counts per operation of the bench line in
    ``path`` (its "pairs_by_op"); no GPU is touched."""
    from registration_amd import _lib
    with open(path) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    names = {v: k for k, v in vars(_lib).items() if k.startswith("OP_") and isinstance(v, int)}
    inv = {v: k for k, v in names.items()}
    counts = {inv[k]: v for k, v in line["pairs_by_op"].items()}
    return cpu_baseline(make_workload(workload), counts)


def _served_run(reqs, prec, max_active=256, keep_momenta=False):
    """The requests through the continuous-batching service (arrival times as given), with
    the kernel-sum pairs it launched counted."""
    import torch
    from registration_amd import runtime as rt, serving
    srv = serving.Server(hbm_cap_bytes=200e9, precision=prec, max_active=max_active,
                         keep_momenta=keep_momenta)
    for r in reqs:
        srv.submit(r)
    torch.cuda.synchronize()
    rt.pair_counts.clear()
    rt.COUNT_PAIRS = True
    t0 = time.perf_counter()
    res = srv.run(progress=lambda f, a, q: _progress(f"served {f}/{len(reqs)} ({a} active, {q} queued)"))
    wall = time.perf_counter() - t0
    rt.COUNT_PAIRS = False
    counts = dict(rt.pair_counts)
    return srv, res, wall, counts


def _pair_rates(counts, wall):
    """Pairs/s of a served run, and the flow-step pairs' FP32 rate (16 flops per pair, the
    flow kernel's convention) as a fraction of the FP32 vector peak, over the run's wall time
    (every other kernel shares that time)."""
    from registration_amd import _lib
    pairs = sum(counts.values())
    flow = counts.get(_lib.OP_FLOW_STEP, 0)
    tf = FLOPS_PER_PAIR["apply_k_gauss_d3_m3"] * flow / wall / 1e12
    return {"pairs": pairs, "pairs_per_s": pairs / wall, "flow_step_pairs": flow,
            "flow_step_tflops_over_wall": round(tf, 3),
            "flow_step_frac_of_fp32_peak_over_wall": round(tf / FP32_PEAK_TFLOPS, 4)}


def _percentiles(ms):
    import numpy as np
    ms = np.asarray(ms)
    return {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))}


def _served_compare(reqs, prec, label, plain_sample=None, max_active=256, groups=None):
    """The requests through the service, and the first ``plain_sample`` of them (all by
    default) one at a time on the plain path; the plain rate is that sample's, and every
    sampled result is compared with its served one.  ``groups`` (rid -> label) adds per-group
    latency percentiles."""
    import collections as _c
    import torch
    from registration_amd import serving
    sample = reqs if plain_sample is None else reqs[:plain_sample]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain, last = [], t0
    for r in sample:
        plain.append(serving.plain_run(r, prec))
        if time.perf_counter() - last > 30.0:
            last = time.perf_counter()
            _progress(f"plain {len(plain)}/{len(sample)}")
    plain_wall = time.perf_counter() - t0
    _progress(f"{label}: plain sample of {len(sample)} done in {plain_wall:.1f} s")
    srv, res, wall, counts = _served_run(reqs, prec, max_active)
    n = len(reqs)
    lat = [res[r.rid].latency * 1e3 for r in reqs]
    plat = [p.finished * 1e3 for p in plain]
    same = all(res[r.rid].history == p.history for r, p in zip(sample, plain))
    plain_rate = len(sample) / plain_wall
    out = {"workload": label, "requests": n,
           "served_registrations_per_s": n / wall, "served_wall_s": wall,
           "served_latency_ms": _percentiles(lat),
           "plain_requests_timed": len(sample),
           "plain_registrations_per_s": plain_rate, "plain_wall_s": plain_wall,
           "plain_latency_ms": _percentiles(plat),
           "speedup_served_vs_plain": (n / wall) / plain_rate,
           "identical_results": same, "rerun_in_fp64": srv.promoted, "turns": srv.rounds,
           "mean_batch_per_turn": srv.answered / max(srv.rounds, 1), "max_batch": srv.max_batch,
           "served_pair_rates": _pair_rates(counts, wall)}
    if groups is not None:
        by = _c.defaultdict(list)
        for r in reqs:
            by[groups[r.rid]].append(res[r.rid].latency * 1e3)
        out["served_latency_ms_by_group"] = {k: dict(count=len(v), **_percentiles(v))
                                             for k, v in sorted(by.items())}
        if plain_sample is None:
            pby = _c.defaultdict(float)
            for r, p in zip(sample, plain):
                pby[groups[r.rid]] += p.finished
            out["plain_seconds_by_group"] = dict(sorted(pby.items()))
    return out


def served_leg(n, prec, plain_sample):
    """The served-surfaces stream: 512 registration pairs (icosphere s3-s6 log-uniform in
    vertices, random rotations, RBF targets, sigma_V ~ U[0.2, 0.5], T 10, 30 CG iterations,
    Poisson arrivals at 50/s), its first n requests served; the plain rate is timed on the
    first ``plain_sample`` of them (all when None)."""
    import collections as _c
    from registration_amd import serving, synthetic
    specs = synthetic.served_surface_requests(512)[:n]
    arr = serving.poisson_arrivals(512, 50.0, seed=4)[:n]
    arr = arr - arr[0]
    reqs = [_surface_request(i, c, arr[i]) for i, c in enumerate(specs)]
    out = _served_compare(reqs, prec,
                          "first %d of the 512-request served-surfaces stream (s3-s6, T 10, 30 CG, "
                          "Poisson 50/s)" % n, plain_sample=plain_sample,
                          groups={i: "s%d" % c["subdiv"] for i, c in enumerate(specs)})
    out["subdiv_histogram"] = {str(k): v for k, v in sorted(_c.Counter(c["subdiv"] for c in specs).items())}
    return out


def _surface_request(i, c, arrival):
    from registration_amd import serving
    return serving.Request(rid=i, kind='surface', template=c["template"], target=c["target"],
                           sigmaKernel=c["sigmaKernel"], sigmaDist=c["sigmaDist"],
                           sigmaError=c["sigmaError"], T=c["T"], maxIter=c["maxIter"],
                           arrival=float(arrival))


def small_pairs_leg(m, prec, plain_sample):
    """m small surface pairs (icosphere s3, 642 vertices, otherwise drawn as the served
    stream: T 10, 30 CG), all arriving at t = 0: the host-bound regime, where batching pays."""
    from registration_amd import synthetic
    specs = synthetic.served_surface_requests(m, seed=5, subdivs=(3,))
    return _served_compare([_surface_request(i, c, 0.0) for i, c in enumerate(specs)], prec,
                           "%d small surface pairs (icosphere s3, 642 vertices; T 10, 30 CG), all "
                           "arriving at t=0; plain rate timed on the first %d" % (m, plain_sample),
                           plain_sample=plain_sample, max_active=m)


class _Iterate:
    """A final iterate (momenta, no affine part) in the form serving.fp64_twin reads."""
    affineDim = 0

    def __init__(self, at):
        self.at = at


def mixed_service_leg(k, prec):
    """The first k requests of the mixed-service stream (2,000 requests: 40% closed curves,
    40% surfaces s3-s7, 20% landmark sets 1k-200k points; Gaussian or Laplacian kernels,
    sigma ~ U[0.1, 0.5], T in {10, 20, 40}; bursty Poisson arrivals at 200/s) through the
    service: per-class latency percentiles, and each result's energy checked against the FP64
    form of the same objective at the same final momenta (per class)."""
    import collections as _c
    import numpy as np
    import torch
    from registration_amd import serving, synthetic
    specs = synthetic.mixed_service_requests(k)
    t0a = specs[0]["arrival"]
    reqs = [serving.Request(rid=i, kind=c["kind"], template=c["template"], target=c["target"],
                            sigmaKernel=c["sigmaKernel"], sigmaDist=c["sigmaDist"],
                            sigmaError=c["sigmaError"], T=c["T"], maxIter=c["maxIter"],
                            typeKernel=c["typeKernel"], order=c["order"], arrival=c["arrival"] - t0a)
            for i, c in enumerate(specs)]
    srv, res, wall, counts = _served_run(reqs, prec, keep_momenta=True)
    _progress(f"mixed service: {k} requests served in {wall:.1f} s")
    errs = _c.defaultdict(list)
    lat = _c.defaultdict(list)
    err32 = _c.defaultdict(list)
    promoted = _c.Counter()
    for r in reqs:
        got = res[r.rid]
        # an FP64 evaluation of the returned iterate, independent of the service's own
        e64 = serving.fp64_twin(r, _Iterate(got.momenta)).objectiveFun()
        errs[r.kind].append(abs(got.energy - e64) / abs(e64))
        lat[r.kind].append(got.latency * 1e3)
        if got.fp32c_error == got.fp32c_error:
            err32[r.kind].append(got.fp32c_error)
        promoted[r.kind] += got.precision == "fp64" and prec == "fp32c"
    torch.cuda.synchronize()
    all_lat = [res[r.rid].latency * 1e3 for r in reqs]
    return {"workload": "first %d requests of the mixed-service stream (curves 256-8192 pts, "
                        "surfaces s3-s7, landmark sets 1k-200k; gauss/laplacian sigma U[0.1,0.5]; "
                        "T 10/20/40; 30 CG; bursty Poisson 200/s)" % k,
            "requests": k, "served_registrations_per_s": k / wall, "served_wall_s": wall,
            "served_latency_ms": _percentiles(all_lat),
            "classes": {c: {"count": len(v), "max_energy_rel_err_vs_fp64": float(max(v)),
                            "latency_ms": _percentiles(lat[c]),
                            "rerun_in_fp64": int(promoted[c]),
                            "fp32c_run_error_max": float(max(err32[c])) if err32[c] else None,
                            "fp32c_run_error_p50": float(np.median(err32[c])) if err32[c] else None}
                        for c, v in sorted(errs.items())},
            "precision_guard": {"certify_tol": srv.certify_tol, "rerun_in_fp64": srv.promoted},
            "sizes": [[r.kind, int(r.template[1].shape[0]), r.T, r.typeKernel] for r in reqs],
            "turns": srv.rounds, "mean_batch_per_turn": srv.answered / max(srv.rounds, 1),
            "max_batch": srv.max_batch, "served_pair_rates": _pair_rates(counts, wall)}


class _HipGraph:
    """Stream capture into a HIP graph through the HIP runtime's C API (libamdhip64)."""

    def __init__(self):
        import ctypes
        self.c = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.graph = ctypes.c_void_p()
        self.exec = ctypes.c_void_p()

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: hipError {rc}")

    def begin(self, stream):
        self._ok(self.hip.hipStreamBeginCapture(self.c.c_void_p(stream), 1), "hipStreamBeginCapture")

    def end(self, stream):
        c = self.c
        self._ok(self.hip.hipStreamEndCapture(c.c_void_p(stream), c.byref(self.graph)), "hipStreamEndCapture")
        self._ok(self.hip.hipGraphInstantiate(c.byref(self.exec), self.graph, None, None, c.c_size_t(0)),
                 "hipGraphInstantiate")

    def launch(self, stream):
        self._ok(self.hip.hipGraphLaunch(self.exec, self.c.c_void_p(stream)), "hipGraphLaunch")

    def nodes(self):
        n = self.c.c_size_t(0)
        self._ok(self.hip.hipGraphGetNodes(self.graph, None, self.c.byref(n)), "hipGraphGetNodes")
        return int(n.value)

    def close(self):
        if self.exec:
            self.hip.hipGraphExecDestroy(self.exec)
        if self.graph:
            self.hip.hipGraphDestroy(self.graph)


def graph_trial_leg(trials=200):
    """One line-search trial of a small registration (icosphere s3: 642 vertices, 1,280
    faces; Gaussian deformation kernel sigma 0.3, T 10, FP32-compensated; FP64 currents
    energy, width 0.3): T fused flow steps, the deformed template's face geometry, its
    currents energy against the target and the target's norm.  Issued three ways, each timed
    over ``trials`` trials by the wall clock with one synchronisation at the end:
    (a) the registration path (matching.launch_objective: descriptors built and uploaded per
    call, the T steps through lddmm_segments_run_steps); (b) the same kernels from descriptors
    planned once, one library call per launch with the descriptors passed by value
    (segs_dev NULL); (c) (b) captured once into a HIP graph, one hipGraphLaunch per trial.
    Every replay's energies are compared bit for bit with (b)'s, with eager launches between
    replays."""
    import ctypes
    import numpy as np
    import torch
    from registration_amd import _lib, engine, matching, runtime as rt, synthetic
    lib = _lib.load()
    cfg = dict(synthetic.surface_pair_config1(), sigmaKernel=0.3, T=10, maxIter=1)
    cfg["name"] = "surface-s3"
    m = registration(cfg, "fp32c")
    T, N = m.Tsize, m.npt
    rng = np.random.default_rng(3)
    m.at = torch.as_tensor(1e-2 * rng.standard_normal((T, N, 3)), device="cuda")
    # (a): the registration path
    matching.launch_objective([m]).result()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(trials):
        matching.launch_objective([m])
    torch.cuda.synchronize()
    a_ms = (time.perf_counter() - t0) / trials * 1e3

    # (b): pre-planned descriptors on a capture-capable stream
    dev = "cuda"
    K = m.Kdiff
    Kd = m.Kdist.energy()
    F = torch.as_tensor(np.ascontiguousarray(cfg["template"][0]), dtype=torch.int32, device=dev)
    xt = torch.zeros((T + 1, N, 3), dtype=torch.float64, device=dev)
    xt[0] = m.x0
    at = m.at.clone()
    en = torch.zeros(T + 2, dtype=torch.float64, device=dev)   # T flow energies, data, norm
    nf = F.shape[0]
    c1 = torch.empty((nf, 3), dtype=torch.float64, device=dev)
    cr1 = torch.empty_like(c1)
    base = engine.flow_record(N, T, K, perm=m._order)
    steps = np.repeat(base[None, :], T, axis=0).reshape(T, 1)
    for t in range(T):
        steps["y"][t, 0] = xt[t].data_ptr()
        steps["w0"][t, 0] = at[t].data_ptr()
        steps["out0"][t, 0] = xt[t + 1].data_ptr()
        steps["total"][t, 0] = en[t].data_ptr()
    fplan = rt.Plan(_lib.OP_FLOW_STEP, K.prec, 3, steps[0])
    raw = steps.view(np.uint8).reshape(T, 1, rt.SEG_SIZE)
    for lo, hi in rt._PLAN_BYTES:
        raw[1:, :, lo:hi] = fplan.recs.view(np.uint8).reshape(1, rt.SEG_SIZE)[None, :, lo:hi]
    raw[0] = fplan.recs.view(np.uint8).reshape(1, rt.SEG_SIZE)
    crec = engine.currents_record(c1, cr1, m.c2, m.cr2, Kd, m.coef)
    crec["total"] = en[T].data_ptr()
    cplan = rt.Plan(_lib.OP_CURRENTS, Kd.prec, 3, crec)
    nrec = rt.seg(y=m.c2, w0=m.cr2, ny=m.c2.shape[0], sigma=Kd.sigma, kernel=Kd.kernel,
                  order=Kd.order, total=en[T + 1])
    nplan = rt.Plan(_lib.OP_KNORM, Kd.prec, 3, nrec)
    works = [torch.empty(p.work_bytes, dtype=torch.uint8, device=dev) for p in (fplan, cplan, nplan)]
    SP = rt._SEG_PTR

    def issue(st):
        for t in range(T):
            _lib.check(lib.lddmm_segments_run(_lib.OP_FLOW_STEP, K.prec, 3, steps[t].ctypes.data_as(SP),
                                              None, 1, fplan.n_items, fplan.n_fin, works[0].data_ptr(),
                                              works[0].numel(), st), "flow step")
        _lib.check(lib.lddmm_surface_geometry(xt[T].data_ptr(), F.data_ptr(), nf, c1.data_ptr(),
                                              cr1.data_ptr(), st), "geometry")
        for p, w in ((cplan, works[1]), (nplan, works[2])):
            _lib.check(lib.lddmm_segments_run(p.op, p.prec, 3, p.recs.ctypes.data_as(SP), None, 1,
                                              p.n_items, p.n_fin, w.data_ptr(), w.numel(), st), "energy")

    s = torch.cuda.Stream()
    sh = s.cuda_stream
    with torch.cuda.stream(s):
        issue(sh)
        torch.cuda.synchronize()
        ref = en.clone()
        t0 = time.perf_counter()
        for _ in range(trials):
            issue(sh)
        torch.cuda.synchronize()
        b_ms = (time.perf_counter() - t0) / trials * 1e3
        # (c): capture (b) once, replay
        g = _HipGraph()
        try:
            g.begin(sh)
            issue(sh)
            g.end(sh)
            nodes = g.nodes()
            same = True
            for k in range(3):
                en.zero_()
                g.launch(sh)
                torch.cuda.synchronize()
                same = same and bool(torch.equal(en, ref))
                issue(sh)                      # eager work between replays
                torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(trials):
                g.launch(sh)
            torch.cuda.synchronize()
            c_ms = (time.perf_counter() - t0) / trials * 1e3
            same = same and bool(torch.equal(en, ref))
        finally:
            g.close()
    return {"workload": "one line-search trial of an icosphere-s3 registration (642 vertices, "
                        "1,280 faces; sigma 0.3, T 10; FP32-compensated flow, FP64 currents "
                        "energy)",
            "trials": trials, "graph_nodes": nodes,
            "registration_path_ms": round(a_ms, 4), "preplanned_calls_ms": round(b_ms, 4),
            "hip_graph_replay_ms": round(c_ms, 4), "graph_replays_identical": same}


def main():
    args = parse()
    if args.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_only(args.cpu_only, args.workload)}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (registrations "
              "are independent, so several GPUs would each serve their own requests with "
              "unchanged kernels)", file=sys.stderr)
        return 2
    import numpy as np
    import torch
    from registration_amd import _lib, runtime as rt
    _lib.load()
    torch.cuda.set_device(0)
    cfg = make_workload(args.workload)

    if args.roofline_only:
        roof = flow_step_roofline(cfg, args.precision, args.roofline_reps, natural=False)
        # Calibration copy of a known size: 256 MiB read + 256 MiB written.
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"roofline": roof, "calibration_copy_bytes": 256 << 20}))
        return 0

    if args.graph_trial:
        print(json.dumps({"graph_trial": graph_trial_leg()}))
        return 0

    if args.energy_only:
        print(json.dumps(data_terms_leg(cfg, args.data_term_reps, gradients=False)))
        return 0

    for _ in range(args.warmup):
        m = registration(cfg, args.precision)
        m.optimizeMatching()
    torch.cuda.synchronize()
    _progress(f"{args.warmup} warmup registration(s) done")

    rt.COUNT_PAIRS = True
    rt.pair_counts.clear()
    lat = []
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        m = registration(cfg, args.precision)
        m.optimizeMatching()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        iters.append(len(m.history) - 1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _progress(f"{args.steps} timed registration(s): {wall:.3f} s")
    rt.COUNT_PAIRS = False
    counts = dict(rt.pair_counts)
    pairs_per_reg = sum(counts.values()) / args.steps

    # Energy error of the final state against the FP64 form of the same objective.
    final_obj = m.obj
    m64 = registration(cfg, "fp64")
    m64.at = m.at.clone()
    obj64 = m64.objectiveFun()
    energy_err = abs(final_obj - obj64) / abs(obj64)
    del m64
    fp64_form = (fp64_form_leg(cfg, final_obj)
                 if args.precision == "fp32c" and not args.no_fp64_form else None)

    # (--roofline-reps 0 leaves the roofline leg out: profiled runs of the registration alone)
    roof = (flow_step_roofline(cfg, args.precision, args.roofline_reps)
            if args.roofline_reps > 0 else None)
    data_terms = data_terms_leg(cfg, args.data_term_reps) if args.data_term_reps > 0 else None
    _progress("energy check, roofline and data-term legs done")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, {k: v / args.steps for k, v in counts.items()})
        _progress("cpu baseline done")
    op_names = {v: k for k, v in vars(_lib).items() if k.startswith("OP_") and isinstance(v, int)}

    served = {}
    if args.served > 0:
        served["mixed"] = served_leg(args.served, args.precision,
                                     None if args.plain_sample == 0 else min(args.plain_sample, args.served))
        _progress("served-surfaces leg done")
    if args.small_pairs > 0:
        served["small"] = small_pairs_leg(args.small_pairs, args.precision, 32)
        _progress("small-pairs leg done")
    if args.mixed_service > 0:
        served["mixed_service"] = mixed_service_leg(args.mixed_service, args.precision)
        _progress("mixed-service leg done")

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "registrations/s (plain single-pair GPU path)",
        "value": args.steps / wall,
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32c" else "f64",
        "data": "synthetic (seeded icosphere + RBF target)",
        "config": {"workload": "single large surface pair: icosphere s6 (40962 V / 81920 F), "
                               "sigma_V 0.25, sigma_W 0.1, T 20, <=100 CG iterations"
                   if args.workload == "surface-s6" else "surface pair icosphere s3, T 10, <=50 CG",
                   "precision": args.precision, "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "cg_iterations": iters,
        "stop_reason": m.stop_reason,
        "final_energy": final_obj,
        "energy_rel_err_vs_fp64": energy_err,
        "fp64_form": fp64_form,
        "pairs_per_registration": pairs_per_reg,
        "pairs_by_op": {op_names[k]: v / args.steps for k, v in sorted(counts.items())},
        "pair_rate_overall": pairs_per_reg / (wall / args.steps),
        "roofline": roof,
        "data_terms": data_terms,
        "cpu_baseline": cpu,
        "served": served or None,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
