"""Dispatch orders of work items (lddmm_item_order, host side, no GPU): a permutation of the
planned items, deterministic, longest first on geometries where the cost of every item is
known, and checked against the plan by lddmm_segments_plan."""
import ctypes

import numpy as np
import pytest

from registration_amd import _lib, engine, runtime as rt, synthetic

QP = 256  # stand-in device address: planning only checks pointers for presence


def _plan(op, prec, d, rec):
    r = np.ascontiguousarray(np.array(rec, copy=True).reshape(1))
    for f in ("y", "w0", "out0"):
        if op != _lib.OP_CURRENTS or f != "out0":
            r[f] = r[f] or QP
    if op == _lib.OP_CURRENTS:
        r["x1"] = r["x1"] or QP
        r["w1"] = r["w1"] or QP
    n_items, n_fin, wb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
    rc = _lib.load().lddmm_segments_plan(op, prec, d, r.ctypes.data_as(ctypes.POINTER(_lib.Seg)),
                                         1, ctypes.byref(n_items), ctypes.byref(n_fin),
                                         ctypes.byref(wb))
    return rc, r, n_items.value


def _order(op, prec, d, rec, y, x1=None, perm=None):
    lib = _lib.load()
    r = np.ascontiguousarray(np.array(rec, copy=True).reshape(1))
    sp = r.ctypes.data_as(ctypes.POINTER(_lib.Seg))
    n = lib.lddmm_item_order(op, prec, d, sp, None, None, None, None, 0)
    assert n > 0
    out = np.empty(n, np.int32)
    Y = np.ascontiguousarray(y, dtype=np.float64)
    X1 = None if x1 is None else np.ascontiguousarray(x1, dtype=np.float64)
    P = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
    rc = lib.lddmm_item_order(op, prec, d, sp, Y.ctypes.data,
                              None if X1 is None else X1.ctypes.data,
                              None if P is None else P.ctypes.data, out.ctypes.data, n)
    assert rc == _lib.OK
    return out


def _flow_rec(n, sigma, prec, perm=True):
    r = rt.segs(1)
    r["ny"], r["sigma"], r["dt"], r["kernel"], r["order"] = n, sigma, 0.05, _lib.KERNEL_GAUSS, 10
    if perm:
        r["perm"] = QP
    return r


def test_size_query_matches_the_plan_and_the_order_is_a_permutation():
    V, _ = synthetic.icosphere(5)
    o = engine.spatial_order(V)
    rec = _flow_rec(V.shape[0], 0.25, _lib.PREC_FP32C)
    rc, planned, n_items = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec)
    assert rc == _lib.OK
    lib = _lib.load()
    q = np.ascontiguousarray(rec.copy())
    assert lib.lddmm_item_order(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3,
                                q.ctypes.data_as(ctypes.POINTER(_lib.Seg)), None, None, None,
                                None, 0) == n_items
    a = _order(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec, V, perm=o)
    b = _order(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec, V, perm=o)
    assert np.array_equal(np.sort(a), np.arange(n_items))
    assert np.array_equal(a, b)


def test_far_apart_clusters_run_their_culled_items_last():
    # Two clusters of 4,096 points, 100 kernel widths apart: an item whose rows lie in one
    # cluster and whose sources lie in the other has every block culled.
    rng = np.random.default_rng(3)
    A = 0.3 * rng.standard_normal((4096, 3))
    P = np.concatenate([A, A[::-1] + [100.0, 0.0, 0.0]])
    sigma = 1.0
    perm = np.arange(P.shape[0], dtype=np.int32)
    for prec in (_lib.PREC_FP32C, _lib.PREC_FP64):
        rec = _flow_rec(P.shape[0], sigma, prec, perm=False)
        rc, planned, n_items = _plan(_lib.OP_FLOW_STEP, prec, 3, rec)
        assert rc == _lib.OK
        rows = 512 if prec == _lib.PREC_FP32C else 256
        row_tiles, cl = int(planned["row_tiles"][0]), int(planned["chunk_len"][0])
        it = np.arange(n_items)
        rt_, cg = it % row_tiles, it // row_tiles
        r_lo, r_hi = rt_ * rows, np.minimum((rt_ + 1) * rows, P.shape[0]) - 1
        s_lo, s_hi = cg * cl, np.minimum((cg + 1) * cl, P.shape[0]) - 1
        side = lambda i: i >= 4096  # noqa: E731
        cross = (side(r_lo) == side(r_hi)) & (side(s_lo) == side(s_hi)) & (side(r_lo) != side(s_lo))
        same = (side(r_lo) == side(r_hi)) & (side(s_lo) == side(s_hi)) & (side(r_lo) == side(s_lo))
        assert cross.any() and same.any()
        o = _order(_lib.OP_FLOW_STEP, prec, 3, rec, P, perm=perm if prec == 1 else None)
        pos = np.empty(n_items, np.int64)
        pos[o] = np.arange(n_items)
        assert pos[cross].min() > pos[same].max()
        # ties keep item order
        assert np.all(np.diff(pos[cross][np.argsort(it[cross])]) > 0)


def test_symmetric_energy_runs_skipped_tiles_last():
    # FP64 currents energy of a shape against a far target: the self sum evaluates each
    # unordered tile pair once, so its items whose chunk lies wholly below the row tile
    # (and the cross items, all culled) cost only their fixed part.
    V, F = synthetic.icosphere(5)
    C = V[F].mean(axis=1)
    rec = rt.segs(1)
    rec["ny"], rec["nx1"], rec["sigma"], rec["coef"] = C.shape[0], C.shape[0], 0.1, 1.0 / 3.0
    rec["kernel"], rec["order"] = _lib.KERNEL_GAUSS, 10
    rc, planned, n_items = _plan(_lib.OP_CURRENTS, _lib.PREC_FP64, 3, rec)
    assert rc == _lib.OK
    row_tiles, cl, c0 = (int(planned[f][0]) for f in ("row_tiles", "chunk_len", "chunks0"))
    o = _order(_lib.OP_CURRENTS, _lib.PREC_FP64, 3, rec, C, x1=C + [50.0, 0.0, 0.0])
    it = np.arange(n_items)
    rt_, cg = it % row_tiles, it // row_tiles
    below = (cg < c0) & ((cg + 1) * cl <= rt_ * 256)       # every source tile below the rows
    cross = cg >= c0
    live = (cg < c0) & (cg * cl <= rt_ * 256) & ((cg + 1) * cl > rt_ * 256)  # the diagonal
    assert below.any() and live.any()
    pos = np.empty(n_items, np.int64)
    pos[o] = np.arange(n_items)
    assert pos[below | cross].min() > pos[live].max()


def test_plan_checks_the_order_against_its_items():
    V, _ = synthetic.icosphere(4)
    rec = _flow_rec(V.shape[0], 0.3, _lib.PREC_FP32C, perm=False)
    rc, planned, n_items = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec)
    assert rc == _lib.OK
    rec["item_order"] = QP
    rec["n_order"] = n_items
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec)[0] == _lib.OK
    rec["n_order"] = n_items + 1
    assert _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec)[0] == _lib.EINVAL


def test_unsupported_operations_and_bad_arguments():
    lib = _lib.load()
    V, _ = synthetic.icosphere(3)
    rec = np.ascontiguousarray(_flow_rec(V.shape[0], 0.3, _lib.PREC_FP32C, perm=False))
    sp = rec.ctypes.data_as(ctypes.POINTER(_lib.Seg))
    assert lib.lddmm_item_order(_lib.OP_APPLY_K, 1, 3, sp, None, None, None, None, 0) == \
        _lib.EUNSUPPORTED
    n = lib.lddmm_item_order(_lib.OP_FLOW_STEP, 1, 3, sp, None, None, None, None, 0)
    out = np.empty(n + 1, np.int32)
    Y = np.ascontiguousarray(V)
    # wrong length, missing rows, an out-of-range processing order
    assert lib.lddmm_item_order(_lib.OP_FLOW_STEP, 1, 3, sp, Y.ctypes.data, None, None,
                                out.ctypes.data, n + 1) == _lib.EINVAL
    assert lib.lddmm_item_order(_lib.OP_FLOW_STEP, 1, 3, sp, None, None, None,
                                out.ctypes.data, n) == _lib.EINVAL
    bad = np.full(V.shape[0], V.shape[0], np.int32)
    assert lib.lddmm_item_order(_lib.OP_FLOW_STEP, 1, 3, sp, Y.ctypes.data, None,
                                bad.ctypes.data, out.ctypes.data, n) == _lib.EINVAL


def test_engine_helper_skips_small_launches():
    V, _ = synthetic.icosphere(3)
    rec = _flow_rec(V.shape[0], 0.3, _lib.PREC_FP32C, perm=False)
    assert engine.item_order(_lib.OP_FLOW_STEP, _lib.PREC_FP32C, 3, rec, V) is None


@pytest.mark.parametrize("n", [1, 255, 256, 257, 5000])
def test_odd_sizes(n):
    rng = np.random.default_rng(n)
    P = rng.standard_normal((n, 2))
    rec = _flow_rec(n, 0.5, _lib.PREC_FP64, perm=False)
    rc, planned, n_items = _plan(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 2, rec)
    assert rc == _lib.OK
    o = _order(_lib.OP_FLOW_STEP, _lib.PREC_FP64, 2, rec, P)
    assert np.array_equal(np.sort(o), np.arange(n_items))
