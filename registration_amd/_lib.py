"""ctypes binding of liblddmm_hip.so, the gfx950 kernel library (C ABI in include/lddmm.h).

The library is the only compute path of this package: there is no CPU fallback.  If the
shared object is missing or cannot be loaded, every entry point raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblddmm_hip.so")
# Chunking targets chosen by registration_amd.autotune on an MI355X, applied when the
# library loads; LDDMM_TUNING names another file, or "off" for the built-in targets.
TUNING_PATH = os.path.join(_HERE, "tuning_gfx950.json")

OK = 0
EINVAL = -1
EWORKSPACE = -2
EHIP = -3
EUNSUPPORTED = -4
ESCRATCH = -5

KERNEL_GAUSS = 0
KERNEL_LAPLACIAN = 1

PREC_FP64 = 0
PREC_FP32C = 1
PREC_FP64E = 2   # data-term energies of FP32C registrations (energy-only segments)

OP_APPLY_K = 0
OP_DIFF_KT = 1
OP_FLOW_STEP = 2
OP_COVECTOR_STEP = 3
OP_KDOT = 4
OP_CURRENTS = 5
OP_DDIFF_K11_12 = 6
OP_KNORM = 7
OP_MEASURE = 8
OP_VARIFOLD = 9
OP_DIVERGENCE = 10
OP_DIFF_K = 11
OP_IDS = {k[3:]: v for k, v in dict(globals()).items() if k.startswith("OP_")}
PRECS = {"fp64": PREC_FP64, "fp32c": PREC_FP32C}

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t


class Seg(ctypes.Structure):
    """Mirror of ``lddmm_seg_t`` (include/lddmm.h)."""

    _fields_ = [
        ("y", _P), ("x0", _P), ("x1", _P),
        ("w0", _P), ("w1", _P), ("w2", _P),
        ("tw", _P * 4), ("sw", _P * 4), ("g", _P * 2),
        ("out0", _P), ("out1", _P), ("rowdot", _P), ("total", _P),
        ("ny", _I64), ("nx0", _I64), ("nx1", _I64),
        ("sigma", _D), ("dt", _D), ("coef", _D),
        ("kernel", ctypes.c_int32), ("order", ctypes.c_int32),
        ("m", ctypes.c_int32), ("na", ctypes.c_int32),
        ("item0", _I64), ("part0", _I64), ("fin0", _I64), ("red0", _I64),
        ("row_tiles", ctypes.c_int32), ("chunks0", ctypes.c_int32),
        ("chunks1", ctypes.c_int32), ("chunk_len", ctypes.c_int32),
        ("fin_blocks", ctypes.c_int32), ("stage0", ctypes.c_int32),
        ("perm", _P), ("item_order", _P), ("n_order", _I64),
    ]


class Geom(ctypes.Structure):
    """Mirror of ``lddmm_geom_t``."""

    _fields_ = [("vertices", _P), ("faces", _P), ("centers", _P), ("normals", _P),
                ("weights", _P), ("nf", _I64), ("k", ctypes.c_int32), ("d", ctypes.c_int32),
                ("blk0", _I64)]


class Gather(ctypes.Structure):
    """Mirror of ``lddmm_gather_t``."""

    _fields_ = [("vertices", _P), ("faces", _P), ("z1", _P), ("dz", _P), ("incid_ptr", _P),
                ("incid", _P), ("px", _P), ("nv", _I64), ("scale", _D), ("k", ctypes.c_int32),
                ("d", ctypes.c_int32), ("blk0", _I64)]


class Landmark(ctypes.Structure):
    """Mirror of ``lddmm_lmk_t``."""

    _fields_ = [("x", _P), ("y", _P), ("total", _P), ("grad", _P), ("n", _I64), ("scale", _D)]


# Every symbol include/lddmm.h declares, with its ctypes signature.
SIGNATURES = {
    "lddmm_segments_plan": (_I, [_I, _I, _I, ctypes.POINTER(Seg), _I,
                                 ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_SZ)]),
    "lddmm_segments_run": (_I, [_I, _I, _I, ctypes.POINTER(Seg), _P, _I, _I64, _I64, _P, _SZ, _P]),
    "lddmm_segments_run_steps": (_I, [_I, _I, _I, ctypes.POINTER(Seg), _P, _I, _I, _I64, _I64, _P,
                                      _SZ, _P]),
    "lddmm_spatial_order": (_I, [_P, _I64, _I, _P]),
    "lddmm_order_compactness": (ctypes.c_double, [_P, _I64, _I, _P, _D, _D]),
    "lddmm_order_live_fraction": (ctypes.c_double, [_P, _I64, _I, _P, _I, _D, _I64]),
    "lddmm_surface_geometry_host": (_I, [_P, _I64, _P, _I64, _P, _P, _P]),
    "lddmm_axpy_batch": (_I, [_I, _P, _P, _P, _P, _P, _P]),
    "lddmm_item_order": (_I, [_I, _I, _I, ctypes.POINTER(Seg), _P, _P, _P, _P, _I64]),
    "lddmm_apply_k_workspace_bytes": (_SZ, [_I64, _I64, _I, _I]),
    "lddmm_apply_k": (_I, [_P, _P, _P, _I64, _I64, _I, _I, _I, _I, _D, _I, _P, _P, _SZ, _P]),
    "lddmm_apply_k_noalloc": (_I, [_P, _P, _P, _I64, _I64, _I, _I, _I, _I, _D, _I, _P, _P]),
    "lddmm_apply_diff_kt_workspace_bytes": (_SZ, [_I64, _I64, _I, _I, _I]),
    "lddmm_apply_diff_kt": (_I, [_P, _P, _P, _P, _I64, _I64, _I, _I, _I, _I, _I, _D, _I,
                                 _P, _P, _SZ, _P]),
    "lddmm_apply_ddiff_k11and12_workspace_bytes": (_SZ, [_I64, _I]),
    "lddmm_apply_ddiff_k11and12": (_I, [_P, _P, _P, _P, _I64, _I, _I, _I, _D, _I, _P, _P,
                                        _SZ, _P]),
    "lddmm_apply_k_mat": (_I, [_P, _P, _P, _I64, _I64, _I, _I, _I, _I, _D, _P, _P]),
    "lddmm_apply_diff_k_mat": (_I, [_P, _P, _P, _I64, _I64, _I, _I, _I, _D, _P, _P]),
    "lddmm_kernel_matrix": (_I, [_P, _P, _I64, _I64, _I, _I, _I, _D, _I, _I, _P, _P]),
    "lddmm_surface_geometry": (_I, [_P, _P, _I64, _P, _P, _P]),
    "lddmm_curve_geometry": (_I, [_P, _P, _I64, _I, _P, _P, _P]),
    "lddmm_surface_vertex_gradient": (_I, [_P, _P, _P, _P, _P, _P, _I64, _D, _P, _P]),
    "lddmm_curve_vertex_gradient": (_I, [_P, _P, _P, _P, _I64, _I, _D, _P, _P]),
    "lddmm_geometry_batch": (_I, [ctypes.POINTER(Geom), _P, _I, _P]),
    "lddmm_vertex_gradient_batch": (_I, [ctypes.POINTER(Gather), _P, _I, _P]),
    "lddmm_landmark_batch": (_I, [ctypes.POINTER(Landmark), _P, _I, _P]),
    "lddmm_sum_f64": (_I, [_P, _I64, _P, _P]),
    "lddmm_decimate_surface": (_I, [_P, _I64, _P, _I64, _I64, _P, _P, ctypes.POINTER(_I64),
                                    ctypes.POINTER(_I64)]),
    "lddmm_size_class": (_I, [_I64]),
    "lddmm_set_items_target": (_I, [_I, _I, _I, _I, _I64]),
    "lddmm_get_items_target": (_I64, [_I, _I, _I, _I]),
    "lddmm_last_hip_error": (_I, []),
    "lddmm_version": (ctypes.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load the kernel library once; raise RuntimeError if it is not built or broken."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"liblddmm_hip.so not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        apply_tuning(lib)
        _lib = lib
    return _lib


def apply_tuning(lib, path=None) -> int:
    """Set the chunking targets recorded in a tuning file (default: LDDMM_TUNING, else
    TUNING_PATH when present); returns the number of targets set."""
    import json
    path = path or os.environ.get("LDDMM_TUNING", TUNING_PATH)
    if path == "off" or not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    n = 0
    for e in table["entries"]:
        rc = lib.lddmm_set_items_target(OP_IDS[e["op"]], PRECS[e["prec"]], int(e["energy_only"]),
                                        int(e["size_class"]), int(e["items"]))
        if rc != OK:
            raise RuntimeError(f"tuning entry rejected by the library: {e}")
        n += 1
    return n


_ERRORS = {
    EINVAL: "invalid argument (size, dimension, kernel or order)",
    EWORKSPACE: "workspace too small",
    EHIP: "HIP runtime error",
    EUNSUPPORTED: "unsupported shape",
    ESCRATCH: "request larger than the allocation-free entry point's fixed scratch",
}


def check(rc: int, what: str) -> None:
    if rc == OK:
        return
    msg = _ERRORS.get(rc, f"error {rc}")
    if rc == EINVAL or rc == EUNSUPPORTED:
        raise ValueError(f"{what}: {msg}")
    if rc == EHIP:
        raise RuntimeError(f"{what}: {msg} (hipError {load().lddmm_last_hip_error()})")
    raise RuntimeError(f"{what}: {msg}")
