"""Writes tests/golden/kernel_sums.npz and tests/golden/kernel_mat_sums.npz: inputs and
outputs of the reference's own Fortran kernel sums (py-lddmm/kernelFunctions.f90 applyk,
applykdifft, applykdiff11and12; the matrix-weight sums applykmat, applykdiffmat of the
varifold data term; applykdiff1, applykdiff2, applykdiff1and2), compiled by oracle/Makefile (`make -C oracle ref`) into
oracle/_ref/libkff_ref.so in this container.

Run:  python tests/golden/make_golden.py
The fixture holds data only (seeded inputs, reference outputs), so the GPU tests and the
oracle tests can check against the reference without its sources or binaries.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import lddmm_oracle as orc  # noqa: E402

CASES = [
    # (name, d, ny, nx, m, na, order, sigma)
    ("gauss_d3", 3, 70, 53, 3, 3, 10, 0.6),
    ("gauss_d2", 2, 61, 61, 2, 1, 10, 0.4),
    ("gauss_m1", 3, 33, 40, 1, 2, 10, 1.3),
    ("lap0_d3", 3, 50, 45, 3, 1, 0, 0.7),
    ("lap1_d2", 2, 64, 64, 2, 2, 1, 0.5),
    ("lap2_d3", 3, 48, 37, 3, 3, 2, 0.8),
    ("lap3_d3", 3, 52, 52, 3, 1, 3, 0.6),
    ("lap4_d2", 2, 40, 44, 2, 3, 4, 0.9),
]


def main():
    if not orc.ref_available():
        raise SystemExit("oracle/_ref/libkff_ref.so missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20261203)
    out = {}
    for name, d, ny, nx, m, na, order, sigma in CASES:
        y = rng.standard_normal((ny, d)) * 0.8
        x = rng.standard_normal((nx, d)) * 0.8
        x[:5] = y[:5]            # exact coincidences exercise the diagonal cut-off
        x[5] = y[6] + 1e-10      # below the cut-off
        w = rng.standard_normal((nx, m))
        a1 = rng.standard_normal((na, ny, m))
        a2 = rng.standard_normal((na, nx, m))
        out[f"{name}/y"], out[f"{name}/x"], out[f"{name}/w"] = y, x, w
        out[f"{name}/a1"], out[f"{name}/a2"] = a1, a2
        out[f"{name}/params"] = np.array([d, ny, nx, m, na, order, sigma])
        out[f"{name}/applyk"] = orc.apply_k(y, x, w, sigma, order, sums="reference")
        if order >= 1:
            out[f"{name}/applykdifft"] = orc.apply_diff_kt(y, x, a1, a2, sigma, order,
                                                           sums="reference")
        if order >= 2:
            n = ny
            xs = y
            b1 = rng.standard_normal((n, d))
            b2 = rng.standard_normal((n, d))
            p = rng.standard_normal((n, d))
            out[f"{name}/dd_a1"], out[f"{name}/dd_a2"], out[f"{name}/dd_p"] = b1, b2, p
            out[f"{name}/applykdiff11and12"] = orc.apply_ddiff_k11and12(xs, b1, b2, p, sigma,
                                                                         order, sums="reference")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_sums.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")
    mat_cases()


MAT_CASES = [
    # (name, d, ny, nx, m, order, sigma)
    ("gauss_d3", 3, 37, 29, 1, 10, 0.6),
    ("gauss_m3", 3, 30, 41, 3, 10, 0.9),
    ("lap1_d3", 3, 33, 33, 1, 1, 0.5),
    ("lap3_d3", 3, 28, 35, 3, 3, 0.7),
    ("lap4_d2", 2, 31, 26, 2, 4, 0.8),
]


def mat_cases():
    """Golden vectors of the matrix-weight sums and of the applyDiffK family (own seed, so
    kernel_sums.npz is unchanged)."""
    rng = np.random.default_rng(20261204)
    out = {}
    for name, d, ny, nx, m, order, sigma in MAT_CASES:
        y = rng.standard_normal((ny, d)) * 0.8
        x = rng.standard_normal((nx, d)) * 0.8
        x[:4] = y[:4]            # coincidences: the diagonal cut-off
        beta = rng.standard_normal((ny, nx, m))
        out[f"{name}/y"], out[f"{name}/x"], out[f"{name}/beta"] = y, x, beta
        out[f"{name}/params"] = np.array([d, ny, nx, m, order, sigma])
        out[f"{name}/applykmat"] = orc.apply_k_mat(y, x, beta, sigma, order, sums="reference")
        out[f"{name}/applykdiffmat"] = orc.apply_diff_k_mat(y, x, beta[..., 0], sigma, order,
                                                            sums="reference")
        # the applyDiffK family on the target set (n = ny, d, with a1, a2 of that shape)
        a1 = rng.standard_normal((ny, d))
        a2 = rng.standard_normal((ny, d))
        out[f"{name}/dk_a1"], out[f"{name}/dk_a2"] = a1, a2
        for form in (1, 2, 3):
            out[f"{name}/applykdiff_{form}"] = orc.apply_diff_k(y, a1, a2, sigma, order, form,
                                                                sums="reference")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_mat_sums.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
