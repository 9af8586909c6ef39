"""ctypes binding of libbmx.so (include/bmx.h).

The library is the product path: there is no CPU fallback. If libbmx.so is missing or a call
fails, this module raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BMX_LIBRARY names another build of the library to load (same C ABI), else the in-tree one.
LIB_PATH = os.environ.get("BMX_LIBRARY") or os.path.join(HERE, "libbmx.so")

_lib = None

ERR_UNSUPPORTED = 2  # BMX_ERR_UNSUPPORTED
ERR_UNSTABLE = 9  # BMX_ERR_UNSTABLE


class BmxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bmx error {code}: {msg}")
        self.code = code


class BeamView(ctypes.Structure):
    """bmx_beam_view"""
    _fields_ = [(n, ctypes.c_void_p) for n in ("x", "px", "y", "py", "z", "pz", "beta",
                                               "state", "loss_ele", "loss_turn")]


_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


class GroupSpec(ctypes.Structure):
    """ctypes mirror of bmx_group (include/bmx.h)."""
    _fields_ = [("lat", ctypes.c_void_p), ("beam", ctypes.c_void_p), ("first", ctypes.c_int64),
                ("count", ctypes.c_int64)]


class ServerConfig(ctypes.Structure):
    """ctypes mirror of bmx_server_config."""
    _fields_ = [("hbm_budget", ctypes.c_int64), ("chunk_turns", ctypes.c_int32),
                ("compact_turns", ctypes.c_int32)]


class RequestStats(ctypes.Structure):
    """ctypes mirror of bmx_request_stats."""
    _fields_ = [("submit_ns", ctypes.c_int64), ("admit_ns", ctypes.c_int64),
                ("done_ns", ctypes.c_int64), ("status", ctypes.c_int32), ("chunks", ctypes.c_int32)]


class RadConfig(ctypes.Structure):
    """ctypes mirror of bmx_rad_config."""
    _fields_ = [("damping_on", ctypes.c_int32), ("fluctuations_on", ctypes.c_int32),
                ("zero_average", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("synch_rad_scale", ctypes.c_double), ("seed", ctypes.c_uint64)]


# numpy mirror of bmx_rad_map (rad_map_struct without xfer_damp_mat, row-major)
RAD_MAP_DTYPE = np.dtype([("ref_orb", "<f8", (6,)), ("damp_dmat", "<f8", (6, 6)),
                          ("xfer_damp_vec", "<f8", (6,)), ("stoc_mat", "<f8", (6, 6))])


# name: (restype, argtypes)
_SIGS = {
    "bmx_version": (_I32, []),
    "bmx_last_error": (ctypes.c_char_p, []),
    "bmx_ctx_create": (_I32, [_I32, _PP]),
    "bmx_ctx_destroy": (_I32, [_P]),
    "bmx_ctx_synchronize": (_I32, [_P]),
    "bmx_ctx_stream": (_P, [_P]),
    "bmx_ctx_take_work": (_I32, [_P, ctypes.POINTER(_I64)]),
    "bmx_ctx_record": (_I32, [_P, _I32]),
    "bmx_ctx_set_maps": (_I32, [_P, _I32]),
    "bmx_ctx_set_run": (_I32, [_P, _I32, _I32, _I32]),
    "bmx_ctx_last_run": (_I32, [_P, _P, _P, _P]),
    "bmx_ctx_profile": (_I32, [_P, _I32]),
    "bmx_ctx_kernel_time": (_I32, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64)]),
    "bmx_ctx_elapsed_ms": (_I32, [_P, _I32, _I32, ctypes.POINTER(ctypes.c_float)]),
    "bmx_lattice_create": (_I32, [_P, _P, _I32, ctypes.c_double, _I32, _PP]),
    "bmx_lattice_destroy": (_I32, [_P]),
    "bmx_lattice_create_rad": (_I32, [_P, _P, _I32, ctypes.c_double, _I32, _P, _P, _PP]),
    "bmx_lattice_set_radiation": (_I32, [_P, _P]),
    "bmx_lattice_validate": (_I32, [_P, _I32, ctypes.c_double, _I32, ctypes.POINTER(_I32)]),
    "bmx_lattice_specialize": (_I32, [_P]),
    "bmx_beam_file_info": (_I32, [ctypes.c_char_p, ctypes.POINTER(_I32), ctypes.POINTER(_I64)]),
    "bmx_beam_file_read": (_I32, [ctypes.c_char_p, _I32, ctypes.POINTER(BeamView), _I64,
                                  ctypes.c_double, ctypes.c_double, ctypes.POINTER(_I64),
                                  ctypes.POINTER(ctypes.c_double)]),
    "bmx_beam_file_write": (_I32, [ctypes.c_char_p, ctypes.POINTER(BeamView), _I64, ctypes.c_char_p,
                                   ctypes.c_double, ctypes.c_double, _I32, _I32]),
    "bmx_closed_orbit": (_I32, [_P, _PP, _I32, _I32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), _I32, ctypes.c_double,
                                ctypes.c_double, ctypes.POINTER(_I32)]),
    "bmx_beam_h5_info": (_I32, [ctypes.c_char_p, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_int64),
                                ctypes.c_char_p, _I32]),
    "bmx_beam_h5_read": (_I32, [ctypes.c_char_p, _I32, _P, ctypes.c_int64, ctypes.c_double,
                                ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
    "bmx_beam_h5_write": (_I32, [ctypes.c_char_p, _P, ctypes.c_int64, ctypes.c_char_p, ctypes.c_double,
                                 ctypes.c_double, _I32, _I32]),
    "bmx_normal_mode_tunes": (_I32, [ctypes.POINTER(ctypes.c_double), _I32,
                                     ctypes.POINTER(ctypes.c_double)]),
    "bmx_chromaticity": (_I32, [_P, _PP, _I32, ctypes.c_double, ctypes.c_double,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(_I32)]),
    "bmx_lattice_info": (_I32, [_P, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_uint64)]),
    "bmx_beam_create": (_I32, [_P, _I64, _PP]),
    "bmx_beam_wrap": (_I32, [_P, ctypes.POINTER(BeamView), _I64, _PP]),
    "bmx_beam_upload": (_I32, [_P, ctypes.POINTER(BeamView), _I64]),
    "bmx_beam_download": (_I32, [_P, ctypes.POINTER(BeamView), _I64]),
    "bmx_beam_destroy": (_I32, [_P]),
    "bmx_beam_count_alive": (_I32, [_P, ctypes.POINTER(_I64)]),
    "bmx_track_turns": (_I32, [_P, _P, _P, _I32, _I32]),
    "bmx_track_turns_from": (_I32, [_P, _P, _P, _I32, _I32, _I32]),
    "bmx_track_turns_noalloc": (_I32, [_P, _P, _P, _I32, _I32]),
    "bmx_track_bunch": (_I32, [_P, _P, _P, _I32, _I32, _I32]),
    "bmx_track_element": (_I32, [_P, _P, _P, _I32, _I32]),
    "bmx_track_turns_plain": (_I32, [_P, _P, _P, _I32, _I32]),
    "bmx_track_orbit": (_I32, [_P, _P, _P, _I32, _I32, _I32, _P]),
    "bmx_server_create": (_I32, [_P, ctypes.POINTER(ServerConfig), _PP]),
    "bmx_server_destroy": (_I32, [_P]),
    "bmx_server_submit": (_I32, [_P, _P, ctypes.POINTER(BeamView), ctypes.POINTER(BeamView), _I64,
                                 _I32, _I32, ctypes.POINTER(_I64)]),
    "bmx_server_poll": (_I32, [_P, _I64, ctypes.POINTER(_I32)]),
    "bmx_server_wait": (_I32, [_P, _I64, ctypes.POINTER(RequestStats)]),
    "bmx_server_totals": (_I32, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "bmx_batch_create": (_I32, [_P, ctypes.POINTER(GroupSpec), _I32, _PP]),
    "bmx_batch_destroy": (_I32, [_P]),
    "bmx_batch_track": (_I32, [_P, _I32, _I32]),
    "bmx_batch_set_run": (_I32, [_P, _I32, _I32, _I32]),
    "bmx_batch_tune": (_I32, [_P, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                              ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_float)]),
}


def load(reload: bool = False):
    """Load libbmx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and not reload:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BmxError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names in _SIGS resolvable in the loaded library."""
    L = load()
    return [n for n in _SIGS if hasattr(L, n)]


def check(rc: int) -> None:
    if rc != 0:
        msg = load().bmx_last_error()
        raise BmxError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
