"""lddmm_apply_k_noalloc, the allocation-free form of lddmm_apply_k (its scratch is the
library's fixed per-GPU array): called through the C ABI, its output must equal
lddmm_apply_k's bit for bit (same plan, same kernels), on two calls in a row, the second with
a smaller input and a different kernel, so nothing left in the scratch by the first call can
leak into the second; and a request larger than the scratch is refused before anything is
launched (a host-side check, run without a GPU)."""
import numpy as np
import pytest

from registration_amd import _lib


def test_noalloc_refuses_a_request_larger_than_its_scratch():
    lib = _lib.load()
    fake = 4096  # never dereferenced: the size check comes before any HIP call
    # 20,000,000 target rows against 100 sources: 480 MB of chunk partials at m = 3
    rc = lib.lddmm_apply_k_noalloc(fake, fake, fake, 20_000_000, 100, 3, 3, _lib.KERNEL_GAUSS,
                                   10, 0.5, _lib.PREC_FP32C, fake, None)
    assert rc == _lib.ESCRATCH
    with pytest.raises(RuntimeError, match="fixed scratch"):
        _lib.check(rc, "lddmm_apply_k_noalloc")
    rc = lib.lddmm_apply_k_noalloc(fake, fake, fake, 10, 10, 4, 3, _lib.KERNEL_GAUSS, 10, 0.5,
                                   _lib.PREC_FP64, fake, None)
    assert rc == _lib.EINVAL  # d = 4


@pytest.mark.gpu
def test_noalloc_matches_apply_k_on_two_calls_in_a_row():
    import torch
    from registration_amd import runtime as rt
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(11)
    cases = [
        # (ny, nx, d, m, kernel, order, sigma): the second call smaller than the first
        (30_000, 24_000, 3, 3, _lib.KERNEL_GAUSS, 10, 0.3),
        (1_500, 2_100, 2, 2, _lib.KERNEL_LAPLACIAN, 2, 0.4),
    ]
    for prec in (_lib.PREC_FP32C, _lib.PREC_FP64):
        for ny, nx, d, m, kern, order, sigma in cases:
            print(f"noalloc case prec={prec} ny={ny} nx={nx} d={d} m={m} kernel={kern}", flush=True)
            y = torch.as_tensor(rng.standard_normal((ny, d)), device=dev)
            x = torch.as_tensor(rng.standard_normal((nx, d)), device=dev)
            w = torch.as_tensor(rng.standard_normal((nx, m)), device=dev)
            ref = torch.full((ny, m), np.nan, dtype=torch.float64, device=dev)
            got = torch.full((ny, m), np.nan, dtype=torch.float64, device=dev)
            st = rt.stream_handle()
            nbytes = lib.lddmm_apply_k_workspace_bytes(ny, nx, d, m)
            ws = rt.workspace(nbytes)
            _lib.check(lib.lddmm_apply_k(y.data_ptr(), x.data_ptr(), w.data_ptr(), ny, nx, d, m,
                                         kern, order, sigma, prec, ref.data_ptr(), ws.data_ptr(),
                                         ws.numel(), st), "lddmm_apply_k")
            _lib.check(lib.lddmm_apply_k_noalloc(y.data_ptr(), x.data_ptr(), w.data_ptr(), ny, nx,
                                                 d, m, kern, order, sigma, prec, got.data_ptr(),
                                                 st), "lddmm_apply_k_noalloc")
            torch.cuda.synchronize()
            r, g = ref.cpu().numpy(), got.cpu().numpy()
            assert np.isfinite(r).all() and np.abs(r).max() > 0
            assert np.array_equal(r, g), np.abs(r - g).max()
