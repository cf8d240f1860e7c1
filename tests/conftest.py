import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def _load_cases(name):
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", name)
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    cases = sorted({k.split("/")[0] for k in data})
    return {c: {k.split("/", 1)[1]: v for k, v in data.items() if k.startswith(c + "/")}
            for c in cases}


@pytest.fixture(scope="session")
def golden():
    """Reference Fortran outputs of applyk, applykdifft, applykdiff11and12."""
    return _load_cases("kernel_sums.npz")


@pytest.fixture(scope="session")
def golden_mat():
    """Reference Fortran outputs of the matrix-weight sums applykmat, applykdiffmat."""
    return _load_cases("kernel_mat_sums.npz")
