"""The oracle's cg and the product's cg against the reference's own conjugateGradient.cg
(py-lddmm/conjugateGradient.py:33-217): on the seeded objectives of tests/golden/cg_cases.py,
which together reach every branch of the reference optimiser, each must make exactly the
calls the reference made (every line-search step length and threshold, every returned
objective, dot product and PR+ beta, the hooks), write the same log lines, stop for the same
reason and end on the same variable (tests/golden/cg_trace.npz, written by
tests/golden/make_cg_golden.py from the reference's code)."""
import os
import sys

import numpy as np
import pytest

from oracle import lddmm_oracle as orc
from registration_amd import conjugateGradient as cgm

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import cg_cases  # noqa: E402

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_trace.npz"))
IMPLS = {"oracle": orc.cg, "product": cgm.cg}


@pytest.mark.parametrize("impl", sorted(IMPLS))
@pytest.mark.parametrize("case", cg_cases.CASES, ids=[c[0] for c in cg_cases.CASES])
def test_cg_reproduces_reference_trace(case, impl):
    name = case[0]
    trace, lines, v, conv = cg_cases.run(IMPLS[impl], case)
    ref = GOLDEN[f"{name}/trace"]
    assert trace.shape == ref.shape, (trace.shape, ref.shape)
    assert np.array_equal(trace, ref, equal_nan=True)
    assert lines == [str(s) for s in GOLDEN[f"{name}/log"]]
    assert np.array_equal(v, GOLDEN[f"{name}/v"])
    assert conv == bool(GOLDEN[f"{name}/converged"])
    assert cg_cases.stop_reason(lines) == str(GOLDEN[f"{name}/stop_reason"])


def test_fixture_reaches_every_branch():
    """The fixture itself spans the reference's branches (so a wrong constant or a dropped
    branch in either implementation changes at least one trace)."""
    reasons = {str(GOLDEN[f"{c[0]}/stop_reason"]) for c in cg_cases.CASES}
    assert reasons == {"small gradient", "small variation", "bad direction", "maxIter"}
    logs = [str(s) for c in cg_cases.CASES for s in GOLDEN[f"{c[0]}/log"]]
    for line in ("Disabling CG: bad direction", "Disabling CG: small variation"):
        assert line in logs
    assert any(s.startswith("Test Gradient:") for s in logs)
    t = GOLDEN["max_iter_restart/trace"]
    betas = t[t[:, 0] == cg_cases.ADD, 1]
    assert np.sum(betas == 0) >= 2 and np.sum(betas > 0) >= 4  # restarts and PR+ steps


def test_a_planted_constant_changes_the_trace():
    """A copy of the oracle's cg with the post-acceptance step growth changed from x100 to x10
    must fail the comparison: the traces pin the constants, not only the stopping point."""
    import inspect
    src = inspect.getsource(orc.cg).replace("eps = 100 * eps", "eps = 10 * eps")
    assert src != inspect.getsource(orc.cg)
    ns = dict(orc.__dict__)
    exec(compile(src, "<planted>", "exec"), ns)
    case = next(c for c in cg_cases.CASES if c[0] == "small_gradient")
    trace, _, _, _ = cg_cases.run(ns["cg"], case)
    ref = GOLDEN["small_gradient/trace"]
    assert trace.shape != ref.shape or not np.array_equal(trace, ref, equal_nan=True)
