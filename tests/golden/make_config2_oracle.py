"""Writes tests/golden/config2_oracle.npz: the FP64 oracle's whole registration of the large
surface pair (icosphere subdivision 6, 40,962 vertices / 81,920 faces, seeded 20-RBF target,
Gaussian sigma_V 0.25, currents sigma_W 0.1, sigma_err 0.1, T 20, at most 100 CG iterations,
momenta from zero), the reference's SurfaceMatching.optimizeMatching path
(py-lddmm/surfaceMatching.py:474-484 driving conjugateGradient.cg) restated by
oracle/lddmm_oracle.py with the C kernel sums of oracle/kernel_sums.c.

Recorded, in call order: every objective evaluation and line-search trial (eps, returned
objective), every RKHS dot-product result, the gradient threshold gradEps, the accepted
objective history, the stop reason and the iteration count.  The GPU test of the same
registration in FP64 compares its own sequence against these numbers.

Run (about an hour on 8 cores):  python tests/golden/make_config2_oracle.py [cache_dir]
With a cache directory, every kernel sum's result is also stored there under a hash of its
inputs, so an interrupted run replays the finished part from disk and goes on from there
(the optimisation path is deterministic, so the replayed calls are the same calls).
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import lddmm_oracle as orc  # noqa: E402
from registration_amd import synthetic  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_oracle.npz")


def cached(fn, cache_dir):
    """fn(*arrays_and_scalars) -> ndarray, memoised on disk by a hash of its arguments."""
    def wrapper(*args, **kw):
        h = hashlib.sha256(fn.__name__.encode())
        for a in list(args) + sorted(kw.items()):
            a = np.ascontiguousarray(a) if isinstance(a, np.ndarray) else np.asarray(repr(a))
            h.update(str(a.shape).encode() + str(a.dtype).encode())
            h.update(a.tobytes())
        path = os.path.join(cache_dir, h.hexdigest() + ".npy")
        if os.path.exists(path):
            return np.load(path)
        r = fn(*args, **kw)
        np.save(path + ".tmp.npy", r)
        os.replace(path + ".tmp.npy", path)
        return r
    return wrapper


def main():
    if len(sys.argv) > 1:
        os.makedirs(sys.argv[1], exist_ok=True)
        orc.apply_k = cached(orc.apply_k, sys.argv[1])
        orc.apply_diff_kt = cached(orc.apply_diff_kt, sys.argv[1])
    c = synthetic.surface_pair_config2()
    (F, V), (_, Y) = c["template"], c["target"]
    o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, Y)),
                            orc.Kernel('gauss', sigma=c["sigmaKernel"]),
                            orc.Kernel('gauss', sigma=c["sigmaDist"]),
                            sigmaError=c["sigmaError"], timeStep=1.0 / c["T"], maxIter=c["maxIter"])
    t0 = time.time()
    trials, dots = [], []
    upd, dot = o.updateTry, o.dotProduct

    def log(msg):
        print(f"[{time.time() - t0:8.1f} s] {msg}", flush=True)

    def updateTry(d, eps, objRef=None):
        r = upd(d, eps, objRef)
        trials.append((eps, r))
        log(f"trial eps={eps:.6g} obj={r:.17g}")
        return r

    def dotProduct(g1, g2):
        r = dot(g1, g2)
        dots.extend(float(v) for v in r)
        log("dot " + " ".join(f"{v:.17g}" for v in r))
        return r

    o.updateTry, o.dotProduct = updateTry, dotProduct
    o.optimizeMatching()
    log(f"done: {len(o.history) - 1} iterations, stop '{o.stop_reason}', obj {o.history[-1]:.17g}")
    np.savez(OUT, history=np.array(o.history), trials=np.array(trials).reshape(-1, 2),
             dots=np.array(dots), gradEps=np.array(o.gradEps), stop_reason=np.array(o.stop_reason),
             iterations=np.array(len(o.history) - 1), seconds=np.array(time.time() - t0),
             threads=np.array(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))))
    log(f"wrote {OUT}")


if __name__ == "__main__":
    main()
