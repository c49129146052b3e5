"""ctypes binding of libseis.so (declarations in include/seis.h)."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_int64, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libseis.so")

CAP_FORWARD = 1
CAP_TANGENT = 2
CAP_ADJOINT = 4
KIND_FORWARD, KIND_TANGENT, KIND_ADJOINT, KIND_BORN, KIND_RECORD = 0, 1, 2, 3, 4
MAX_SHOTS = 16
LABEL_STATIC_MAX_GRIDS = 4096

ERRORS = {-1: "invalid argument", -2: "out of device memory", -3: "HIP error",
          -4: "operation not enabled for this plan"}


class ModelDesc(ctypes.Structure):
    _fields_ = [("nz", c_int32), ("nx", c_int32), ("dx", c_float), ("n_shots", c_int32),
                ("src_z", c_int32), ("src_x", c_int32 * MAX_SHOTS), ("rcv_z", c_int32),
                ("rcv_x0", c_int32), ("rcv_dx", c_int32), ("n_rcv", c_int32)]


class TimeDesc(ctypes.Structure):
    _fields_ = [("nt", c_int32), ("dt", c_float), ("nbc", c_int32), ("wavelet", POINTER(c_double))]


class LbfgsReq(ctypes.Structure):
    _fields_ = [("S", c_void_p), ("Y", c_void_p), ("rho", c_void_p), ("al", c_void_p), ("ring", c_void_p),
                ("hdiag", c_void_p), ("g", c_void_p), ("g_prev", c_void_p), ("d_prev", c_void_p), ("t", c_double),
                ("d", c_void_p), ("scal", c_void_p), ("n", c_int32), ("cap", c_int32), ("first", c_int32),
                ("pad", c_int32)]


LBFGS_MAX_N = 32768

EXPORTS = {
    "seis_plan_create": (c_int32, [POINTER(ModelDesc), c_int32, POINTER(TimeDesc), c_int32, c_int32,
                                   POINTER(c_void_p)]),
    "seis_plan_bytes": (c_int64, [POINTER(ModelDesc), c_int32, POINTER(TimeDesc), c_int32, c_int32]),
    "seis_plan_destroy": (c_int32, [c_void_p]),
    "seis_plan_layout": (c_int32, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "seis_set_active": (c_int32, [c_void_p, POINTER(c_int32)]),
    "seis_set_active_async": (c_int32, [c_void_p, POINTER(c_int32), c_void_p]),
    "seis_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_forward_misfit": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_tangent": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_gradient": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "seis_forward_save": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_adjoint": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    "seis_gn_product": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_tangent_saved": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_gn_product_saved": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "seis_plan_device_bytes": (c_int64, [c_void_p]),
    "seis_plan_set_config": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32]),
    "seis_plan_get_config": (c_int32, [c_void_p, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                       POINTER(c_int32)]),
    "seis_plan_set_stream_threshold": (c_int32, [c_void_p, c_int32]),
    "seis_label_flat": (c_int32, [POINTER(c_int32), POINTER(c_int32), c_int32, c_void_p, c_double, c_void_p,
                                  c_void_p, c_void_p]),
    "seis_lbfgs_direction": (c_int32, [POINTER(LbfgsReq), c_int32, c_void_p]),
    "seis_label_flat_static": (c_int32, [POINTER(c_int32), POINTER(c_int32), c_int32, c_void_p, c_double, c_void_p,
                                         c_void_p, c_void_p]),
    "seis_last_error": (c_char_p, []),
}

_lib = None


def lib():
    """Load libseis.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m seismic_amd.build`")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().seis_last_error().decode(errors="replace")
        kind = ERRORS.get(rc, f"error {rc}")
        if rc == -1:
            raise ValueError(f"{what}: {kind}: {msg}")
        raise RuntimeError(f"{what}: {kind}: {msg}")
