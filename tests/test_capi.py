"""The C ABI library loads, exports every symbol include/lddmm.h declares, and its host-only
planning behaves (no device calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from registration_amd import _lib
from registration_amd._lib import Seg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lddmm.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lddmm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"
    assert lib.lddmm_version().startswith(b"lddmm-mi355x")


def _plan(op, prec, d, segs):
    lib = _lib.load()
    arr = (Seg * len(segs))(*segs)
    n_items, n_fin, wb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
    rc = lib.lddmm_segments_plan(op, prec, d, arr, len(segs), ctypes.byref(n_items),
                                 ctypes.byref(n_fin), ctypes.byref(wb))
    return rc, arr, n_items.value, n_fin.value, wb.value


def _seg(ny, **kw):
    s = Seg()
    s.y = 4096  # planning never dereferences pointers
    s.w0 = s.w1 = s.w2 = 4096
    s.out0 = 4096
    s.ny = ny
    s.sigma = kw.pop("sigma", 0.5)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_plan_layout_and_chunking():
    rc, arr, items, fins, wb = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, [_seg(40962), _seg(642)])
    assert rc == 0
    a, b = arr[0], arr[1]
    assert a.row_tiles == -(-40962 // 512) and b.row_tiles == 2  # 256 threads x 2 rows
    assert a.chunk_len % 256 == 0 and a.chunks0 * a.chunk_len >= 40962
    assert (a.chunks0 - 1) * a.chunk_len < 40962
    assert b.item0 == a.row_tiles * a.chunks0
    assert items == b.item0 + b.row_tiles * b.chunks0
    # 64 rows per finalize workgroup
    assert a.fin_blocks == -(-40962 // 64) and b.fin0 == a.fin_blocks and fins == a.fin_blocks + b.fin_blocks
    # The chunking of a segment depends only on its own sizes (result independent of the batch).
    rc2, alone, *_ = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, [_seg(642)])
    assert (alone[0].chunk_len, alone[0].chunks0, alone[0].row_tiles) == (b.chunk_len, b.chunks0, b.row_tiles)
    assert wb > 0


def test_plan_rejects_invalid():
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 4, [_seg(10)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_FLOW_STEP, 7, 3, [_seg(10)])[0] == _lib.EINVAL
    # Laplacian order 0 has no derivative; order 1 has no second derivative.
    lap0 = _seg(10, kernel=_lib.KERNEL_LAPLACIAN, order=0)
    assert _plan(_lib.OP_COVECTOR_STEP, _lib.PREC_FP64, 3, [lap0])[0] == _lib.EINVAL
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 3, [lap0])[0] == 0
    lap1 = _seg(10, kernel=_lib.KERNEL_LAPLACIAN, order=1)
    assert _plan(_lib.OP_DDIFF_K11_12, _lib.PREC_FP64, 3, [lap1])[0] == _lib.EINVAL
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 3, [_seg(10, kernel=_lib.KERNEL_LAPLACIAN, order=5)])[0] == _lib.EINVAL
    # One kernel family per launch.
    mixed = [_seg(10), _seg(10, kernel=_lib.KERNEL_LAPLACIAN, order=1)]
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 3, mixed)[0] == _lib.EINVAL
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 3, [_seg(0)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 3, [_seg(10, sigma=0.0)])[0] == _lib.EINVAL
    # Currents need a target set.
    assert _plan(_lib.OP_CURRENTS, _lib.PREC_FP64, 3, [_seg(10)])[0] == _lib.EINVAL
    # Measure: the target set is optional (its own norm without one); z1 needs the normals.
    meas = _seg(10, w2=0, out0=0)
    assert _plan(_lib.OP_MEASURE, _lib.PREC_FP64, 3, [meas])[0] == 0
    assert _plan(_lib.OP_MEASURE, _lib.PREC_FP64, 3, [_seg(10, w2=0)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_MEASURE, _lib.PREC_FP64, 3, [_seg(10, out0=0, nx1=5)])[0] == _lib.EINVAL
    # Varifold: surfaces only (d = 3), normals required, z1 only with the gradient form, and a
    # target needs its normals too (sw[0]).
    vf = _seg(10, out0=0)
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [vf])[0] == 0
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 2, [_seg(10, out0=0)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [_seg(10, out0=0, w2=0)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [_seg(10)])[0] == _lib.EINVAL
    tgt = _seg(10, out0=0, nx1=7, x1=4096)
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [tgt])[0] == _lib.EINVAL
    tgt.sw[0] = 4096
    rc, arr, *_ = _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [tgt])
    assert rc == 0 and arr[0].nx1 == 7 and arr[0].chunks1 == 1
    grad = _seg(10, out1=4096)
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3, [grad])[0] == 0
    assert _plan(_lib.OP_VARIFOLD, _lib.PREC_FP64, 3,
                 [_seg(10, out1=4096, kernel=_lib.KERNEL_LAPLACIAN, order=0)])[0] == _lib.EINVAL


def test_fp64e_plans_energy_only_segments_alone():
    """FP64E (the energies of FP32C registrations) plans like FP64 and only for sums whose
    total alone is read: KNORM, and CURRENTS / MEASURE / VARIFOLD without per-row outputs."""
    E = _lib.PREC_FP64E
    assert _plan(_lib.OP_KNORM, E, 3, [_seg(10, out0=0)])[0] == 0
    assert _plan(_lib.OP_KNORM, E, 3, [_seg(10, out0=0, rowdot=4096)])[0] == _lib.EINVAL
    cur = _seg(300, out0=0, nx1=7, x1=4096)
    rc, arr, items, fins, wb = _plan(_lib.OP_CURRENTS, E, 3, [cur])
    rc64, arr64, items64, fins64, wb64 = _plan(_lib.OP_CURRENTS, _lib.PREC_FP64, 3, [cur])
    assert rc == 0 and rc64 == 0 and (items, fins, wb) == (items64, fins64, wb64)
    assert arr[0].row_tiles == arr64[0].row_tiles == 2    # one FP64 row per thread
    assert _plan(_lib.OP_CURRENTS, E, 3, [_seg(300, nx1=7, x1=4096, out1=4096)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_MEASURE, E, 3, [_seg(10, w2=0, out0=0)])[0] == 0
    assert _plan(_lib.OP_VARIFOLD, E, 3, [_seg(10, out0=0)])[0] == 0
    assert _plan(_lib.OP_VARIFOLD, E, 3, [_seg(10, out1=4096)])[0] == _lib.EINVAL
    for op in (_lib.OP_FLOW_STEP, _lib.OP_APPLY_K, _lib.OP_KDOT, _lib.OP_DIVERGENCE):
        assert _plan(op, E, 3, [_seg(10, m=3, na=1, g=(4096, 0))])[0] == _lib.EINVAL, op
    assert _plan(_lib.OP_KNORM, 3, 3, [_seg(10, out0=0)])[0] == _lib.EINVAL


def test_plan_processing_order():
    """A processing order is accepted for the flow and covector steps and the RKHS products
    only."""
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, [_seg(100, perm=4096)])[0] == 0
    kd = _seg(100, perm=4096, na=1)
    kd.g[0] = 4096
    assert _plan(_lib.OP_KDOT, _lib.PREC_FP32C, 3, [kd])[0] == 0
    assert _plan(_lib.OP_APPLY_K, _lib.PREC_FP32C, 3, [_seg(100, perm=4096, m=3)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_COVECTOR_STEP, _lib.PREC_FP32C, 3, [_seg(100, perm=4096)])[0] == 0
    assert _plan(_lib.OP_CURRENTS, _lib.PREC_FP64, 3,
                 [_seg(100, perm=4096, x1=4096, nx1=50)])[0] == _lib.EINVAL
    assert _plan(_lib.OP_KNORM, _lib.PREC_FP64, 3, [_seg(100, perm=4096, out0=0)])[0] == _lib.EINVAL


def test_workspace_queries():
    lib = _lib.load()
    assert lib.lddmm_apply_k_workspace_bytes(1000, 2000, 3, 5) > 1000 * 5 * 8
    assert lib.lddmm_apply_diff_kt_workspace_bytes(100, 100, 3, 3, 2) > 0
    assert lib.lddmm_apply_ddiff_k11and12_workspace_bytes(100, 3) > 0
    assert lib.lddmm_apply_k_workspace_bytes(0, 10, 3, 3) == 0


def test_product_never_imports_the_oracle():
    import registration_amd
    pkg = os.path.dirname(registration_amd.__file__)
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            src = open(os.path.join(pkg, f)).read()
            assert not re.search(r"(import\s+oracle|from\s+oracle|lddmm_oracle|liboracle)", src), f


def test_struct_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives the header's
    struct (a layout slip would hand the kernels wrong addresses)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"lddmm_seg_t": _lib.Seg, "lddmm_geom_t": _lib.Geom,
               "lddmm_gather_t": _lib.Gather, "lddmm_lmk_t": _lib.Landmark}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lddmm.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, *_ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('  return 0;\n}')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, *_ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_plan_numbers_staged_tiles_of_ordered_segments():
    """Segments with a processing order get consecutive staged source tiles (stage0), one
    256-source record each appended to the workspace; a run whose descriptors do not carry
    that numbering is rejected on the host, before any device call."""
    tiles = lambda n: -(-n // 256)
    rec = 256 + (3 + 3 + 1) * 256 * 4  # FP32, d = 3: origin/boxes head, pos, weights, -|y|^2
    rc, arr, items, fins, wb_plain = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3,
                                           [_seg(1000), _seg(300), _seg(5000)])
    assert rc == 0 and [a.stage0 for a in arr] == [0, 0, 0]
    rc, arr, items, fins, wb = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3,
                                     [_seg(1000, perm=4096), _seg(300), _seg(5000, perm=4096)])
    assert rc == 0
    assert [a.stage0 for a in arr] == [0, tiles(1000), tiles(1000)]
    assert wb - wb_plain >= (tiles(1000) + tiles(5000)) * rec
    assert wb - wb_plain < (tiles(1000) + tiles(5000)) * rec + 256
    # FP64 covector records: two weight pairs of 8-byte values.
    rc, arr64, _, _, wb64 = _plan(_lib.OP_COVECTOR_STEP, _lib.PREC_FP64, 3, [_seg(1000, perm=4096)])
    rc0, _, _, _, wb64_plain = _plan(_lib.OP_COVECTOR_STEP, _lib.PREC_FP64, 3, [_seg(1000)])
    assert rc == 0 and rc0 == 0
    assert wb64 - wb64_plain >= tiles(1000) * (256 + (3 + 6 + 1) * 256 * 8)
    # A perm added after planning breaks the numbering: EINVAL, nothing launched.
    lib = _lib.load()
    arr[1].perm = 4096
    rc = lib.lddmm_segments_run(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, arr, None, 3, items, fins,
                                4096, wb, None)
    assert rc == _lib.EINVAL


def test_exp2_table_entries_are_correctly_rounded():
    """The FP64 pair kernel's table of 2^(j/256), j = 0..255 (hex literals in the kernel
    source), holds the double nearest to each exact value (200-bit arithmetic)."""
    mpmath = pytest.importorskip("mpmath")
    src = open(os.path.join(ROOT, "registration_amd", "csrc", "lddmm_kernels.hip")).read()
    block = re.search(r"EXP2_TAB256\[256\]\s*=\s*\{(.*?)\};", src, flags=re.S).group(1)
    vals = [float.fromhex(t) for t in re.findall(r"0x[0-9a-fA-Fp.+-]+", block)]
    assert len(vals) == 256
    mpmath.mp.prec = 200
    for j, v in enumerate(vals):
        exact = mpmath.power(2, mpmath.mpf(j) / 256)
        assert v == float(exact), j
        # and no other double is closer
        assert abs(mpmath.mpf(v) - exact) <= mpmath.mpf(2) ** -53 * v


def test_axpy_batch_rejects_bad_arguments_without_launching():
    import ctypes
    import numpy as np
    from registration_amd import _lib
    lib = _lib.load()
    p = np.zeros(1, np.uint64)
    s = np.zeros(1)
    bad = np.array([-1], np.int64)
    assert lib.lddmm_axpy_batch(1, p.ctypes.data, p.ctypes.data, p.ctypes.data, s.ctypes.data,
                                bad.ctypes.data, None) == _lib.EINVAL
    one = np.array([5], np.int64)   # a non-empty item with null pointers
    assert lib.lddmm_axpy_batch(1, p.ctypes.data, p.ctypes.data, p.ctypes.data, s.ctypes.data,
                                one.ctypes.data, None) == _lib.EINVAL
    assert lib.lddmm_axpy_batch(-1, None, None, None, None, None, None) == _lib.EINVAL
    assert lib.lddmm_axpy_batch(0, None, None, None, None, None, None) == _lib.OK
