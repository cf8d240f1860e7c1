"""lddmm_axpy_batch (the line search's trial momenta at - eps dir of a batch of
registrations): bit for bit the two elementwise passes a + (s * d), for item counts over one
launch's pack, sizes across block boundaries, empty items and in place."""
import numpy as np
import pytest
import torch

from registration_amd import engine

pytestmark = pytest.mark.gpu


def test_axpy_batch_equals_two_passes():
    rng = np.random.default_rng(11)
    sizes = [0, 1, 255, 256, 257, 12_345, 20 * 642 * 3] + list(rng.integers(1, 5000, 66))
    a = [torch.as_tensor(rng.standard_normal(n), device="cuda") for n in sizes]
    d = [torch.as_tensor(rng.standard_normal(n) * 10.0 ** rng.integers(-8, 3), device="cuda")
         for n in sizes]
    s = [-float(x) for x in rng.standard_normal(len(sizes)) * 10.0 ** rng.integers(-6, 2, len(sizes))]
    print(f"axpy batch: {len(sizes)} items", flush=True)
    out = [torch.empty_like(x) for x in a]
    engine.axpy_batch(out, a, d, s)
    ref = [x + (si * y) for x, y, si in zip(a, d, s)]
    assert all(torch.equal(o, r) for o, r in zip(out, ref))
    # the reference's NumPy form at - eps * dir with eps = -s: the same bits
    for o, x, y, si in zip(out[:8], a[:8], d[:8], s[:8]):
        assert np.array_equal(o.cpu().numpy(), x.cpu().numpy() - (-si) * y.cpu().numpy())
    engine.axpy_batch(a, a, d, s)                     # in place
    assert all(torch.equal(x, r) for x, r in zip(a, ref))
