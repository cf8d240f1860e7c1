"""Chunking targets (host side, no GPU): the per-(operation, precision, energy-only form,
size class) table behind lddmm_segments_plan, its validation, the size classes, the tuning
file that _lib.load applies, and that a plan's work-item count follows the target."""
import json

import numpy as np
import pytest

from registration_amd import _lib, runtime as rt


def _builtin_targets(lib):
    for op in _lib.OP_IDS.values():
        for prec in (0, 1):
            for eo in (0, 1):
                for c in range(6):
                    lib.lddmm_set_items_target(op, prec, eo, c, 0)


@pytest.fixture()
def lib():
    lib = _lib.load()
    _builtin_targets(lib)
    yield lib
    _builtin_targets(lib)
    _lib.apply_tuning(lib)  # the targets every later test of the process plans with


def test_size_classes(lib):
    bounds = [(1, 0), (1023, 0), (1024, 1), (4095, 1), (4096, 2), (16383, 2), (16384, 3),
              (65535, 3), (65536, 4), (262143, 4), (262144, 5), (10 ** 7, 5)]
    for n, c in bounds:
        assert lib.lddmm_size_class(n) == c, (n, c)
    assert lib.lddmm_size_class(0) == _lib.EINVAL


def test_set_get_and_validation(lib):
    op, prec = _lib.OP_FLOW_STEP, _lib.PREC_FP32C
    assert lib.lddmm_get_items_target(op, prec, 0, 3) == 6144
    assert lib.lddmm_get_items_target(_lib.OP_CURRENTS, _lib.PREC_FP64, 1, 4) == 24576
    assert lib.lddmm_set_items_target(op, prec, 0, 3, 12288) == _lib.OK
    assert lib.lddmm_get_items_target(op, prec, 0, 3) == 12288
    assert lib.lddmm_get_items_target(op, prec, 0, 2) == 6144      # other classes untouched
    assert lib.lddmm_get_items_target(op, _lib.PREC_FP64, 0, 3) == 6144
    assert lib.lddmm_set_items_target(op, prec, 0, 3, 0) == _lib.OK  # back to built-in
    assert lib.lddmm_get_items_target(op, prec, 0, 3) == 6144
    for bad in [(99, prec, 0, 0, 10), (op, 7, 0, 0, 10), (op, prec, 2, 0, 10), (op, prec, 0, 6, 10),
                (op, prec, 0, -1, 10), (op, prec, 0, 0, 1 << 40)]:
        assert lib.lddmm_set_items_target(*bad) == _lib.EINVAL, bad


def _items(op, prec, n):
    """Planned work items of one flow-step segment of n points (host planning only)."""
    r = rt.seg(ny=n, sigma=0.25, dt=0.1, kernel=_lib.KERNEL_GAUSS, order=10)
    r["y"] = r["w0"] = r["out0"] = 4096          # non-null placeholders: planning reads no memory
    return rt.Plan(op, prec, 3, [r]).n_items


def test_plan_follows_the_target(lib):
    n = 40962
    base = _items(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, n)
    lib.lddmm_set_items_target(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 0, 3, 24576)
    more = _items(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, n)
    lib.lddmm_set_items_target(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 0, 3, 768)
    fewer = _items(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, n)
    rows = -(-n // 512)
    assert fewer < base < more
    # whole source tiles per chunk: the count lands between half the target and the target
    assert 6144 / 2 < base <= 6144 + rows and 24576 / 2 < more <= 24576 + rows
    assert fewer >= rows                                            # at least one chunk per row tile
    assert _items(_lib.OP_FLOW_STEP, _lib.PREC_FP64, n) == _items(_lib.OP_FLOW_STEP, _lib.PREC_FP64, n)


def test_tuning_file_is_applied(lib, tmp_path):
    path = tmp_path / "t.json"
    table = {"entries": [{"op": "FLOW_STEP", "prec": "fp64", "energy_only": False, "size_class": 4,
                          "items": 3072},
                         {"op": "CURRENTS", "prec": "fp64", "energy_only": True, "size_class": 5,
                          "items": 49152}]}
    path.write_text(json.dumps(table))
    assert _lib.apply_tuning(lib, str(path)) == 2
    assert lib.lddmm_get_items_target(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 0, 4) == 3072
    assert lib.lddmm_get_items_target(_lib.OP_CURRENTS, _lib.PREC_FP64, 1, 5) == 49152
    table["entries"][0]["size_class"] = 9
    path.write_text(json.dumps(table))
    with pytest.raises(RuntimeError):
        _lib.apply_tuning(lib, str(path))


def test_committed_tuning_file_is_well_formed():
    import os
    if not os.path.exists(_lib.TUNING_PATH):
        pytest.skip("no tuning file committed")
    t = json.load(open(_lib.TUNING_PATH))
    for e in t["entries"]:
        assert e["op"] in _lib.OP_IDS and e["prec"] in _lib.PRECS and 0 <= e["size_class"] < 6
        assert str(e["items"]) in e["ms"] and e["ms"][str(e["items"])] <= e["ms"][str(e["builtin"])]
        assert e["op"] != "APPLY_K"     # the allocation-free entry point's scratch bound


def test_merge_replaces_only_the_measured_forms():
    """autotune.merge: entries of the forms and classes a partial run measured replace the
    table's; the rest are kept, in FORMS order."""
    from registration_amd import autotune
    e = lambda op, prec, eo, cls, items: {"op": op, "prec": prec, "energy_only": eo,
                                          "size_class": cls, "items": items}
    old = {"entries": [e("FLOW_STEP", "fp32c", False, 3, 6144), e("FLOW_STEP", "fp64", False, 3, 6144),
                       e("CURRENTS", "fp64", True, 4, 4096)]}
    new = {"device": "x", "made": "now", "candidates": [], "min_gain": 0.02,
           "entries": [e("FLOW_STEP", "fp64", False, 3, 12288)]}
    m = autotune.merge(old, new)
    got = [(x["op"], x["prec"], x["size_class"], x["items"]) for x in m["entries"]]
    assert got == [("FLOW_STEP", "fp32c", 3, 6144), ("FLOW_STEP", "fp64", 3, 12288),
                   ("CURRENTS", "fp64", 4, 4096)]
    assert m["made"] == "now"
