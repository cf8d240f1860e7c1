"""Writes tests/golden/cg_trace.npz: the call traces of the reference's own nonlinear conjugate
gradient (py-lddmm/conjugateGradient.py:33-217) on the seeded objectives of
tests/golden/cg_cases.py, which together reach every branch of it (small-gradient,
small-variation, bad-direction and maxIter stops; CG disabled on a bad direction with verb
True, and the verb=False quirk that accepts an increased objective; CG disabled on a small
variation; the epsSmall probe and the halving loop; restarts at restartRate; cgBurnIn; the
gradient test; the plain-array fallbacks; the optional hooks).

The reference module is imported by path from /root/reference, in this container only.  The
fixture holds data only (traces of floats, log lines, final variables): nothing of the
reference's code is copied into this repository.

Run:  python tests/golden/make_cg_golden.py
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cg_cases  # noqa: E402

REF = "/root/reference/py-lddmm/conjugateGradient.py"
OUT = os.path.join(HERE, "cg_trace.npz")

# Every log line family the reference's cg can write; the cases must produce each of them.
BRANCHES = ["iteration 0: obj =", "Stopping Gradient Descent: small gradient",
            "Stopping Gradient Descent: small variation", "Stopping Gradient Descent: bad direction",
            "Disabling CG: bad direction", "Disabling CG: small variation", "Test Gradient:"]


def load_reference_cg():
    spec = importlib.util.spec_from_file_location("_reference_conjugateGradient", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.cg


def main():
    cg = load_reference_cg()
    out, seen, accepts = {}, set(), {}
    for case in cg_cases.CASES:
        name = case[0]
        trace, lines, v, conv = cg_cases.run(cg, case)
        out[f"{name}/trace"] = trace
        out[f"{name}/log"] = np.array(lines, dtype=np.str_)
        out[f"{name}/v"] = v
        out[f"{name}/converged"] = np.array(conv)
        codes = trace[:, 0].astype(int)
        n_acc = int(np.sum(codes == cg_cases.ACC))
        n_try = int(np.sum(codes == cg_cases.TRY))
        reason = cg_cases.stop_reason(lines)
        out[f"{name}/stop_reason"] = np.array(reason)
        for b in BRANCHES:
            if any(s.startswith(b) for s in lines):
                seen.add(b)
        accepts[name] = n_acc
        print(f"{name:26s} {reason:16s} accepts {n_acc:3d} trials {n_try:4d} "
              f"log {len(lines)} lines", flush=True)
        for s in lines:
            if not s.startswith("iteration"):
                print("    ", s)
    # verb=False: the uphill iteration is accepted (the reference's quirk), so that run
    # accepts one iteration more than the verb=True run of the same object
    assert accepts["bad_direction_verb_false"] == accepts["bad_direction_disable"] + 1, accepts
    missing = [b for b in BRANCHES if b not in seen]
    if missing:
        raise SystemExit(f"branches not reached: {missing}")
    np.savez(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
