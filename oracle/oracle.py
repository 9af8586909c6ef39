"""ctypes access to the CPU restatement (libbmad_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmad_oracle.so")
LIB_FMA_PATH = os.path.join(_HERE, "libbmad_oracle_fma.so")
_libs = {}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def build(force: bool = False) -> str:
    """Compile the restatement with its Makefile (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def build_native(out_path: str) -> str:
    """The restatement at -O3 -march=native for this host (the CPU baseline's second leg),
    built into out_path."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "native", f"NATIVE_OUT={out_path}"], check=True)
    return out_path


def lib(fma: bool = False, path: str | None = None):
    """The restatement; fma=True: the build with FMA contraction (same algorithm, other
    rounding); path: another build of the same source (build_native)."""
    if path is None:
        path = LIB_FMA_PATH if fma else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.bmo_track1.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                 _D, _D, _I]
        L.bmo_track1.restype = ctypes.c_int32
        L.bmo_track_bunch.argtypes = ([ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_double, ctypes.c_int64]
                                      + [_D] * 7 + [_I] * 3
                                      + [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int64)])
        L.bmo_track_bunch.restype = ctypes.c_int32
        L.bmo_track_bunch_rad.argtypes = L.bmo_track_bunch.argtypes + [ctypes.c_void_p,
                                                                       ctypes.c_void_p]
        L.bmo_track_bunch_rad.restype = ctypes.c_int32
        L.bmo_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.bmo_philox4x32_10.restype = None
        L.bmo_rad_gauss6.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32, _D]
        L.bmo_rad_gauss6.restype = None
        L.bmo_track_trace.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_int32, ctypes.c_double, _D, _D, _I, ctypes.c_int32,
                                      _D]
        L.bmo_track_trace.restype = ctypes.c_int32
        L.bmo_hwang_bend_edge_kick.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _D]
        L.bmo_check_aperture_limit.argtypes = [ctypes.c_void_p, ctypes.c_int32, _D, _I, _D]
        for f in ("bmo_sqrt_one", "bmo_sinc", "bmo_cosc", "bmo_sincc"):
            getattr(L, f).argtypes = [ctypes.c_double]
            getattr(L, f).restype = ctypes.c_double
        L.bmo_modulo2.argtypes = [ctypes.c_double, ctypes.c_double]
        L.bmo_modulo2.restype = ctypes.c_double
        L.bmo_c_multi.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.bmo_c_multi.restype = ctypes.c_double
        L.bmo_multipole_kt_to_ab.argtypes = [_D, _D, _D, _D, _D]
        L.bmo_multipole_kt_to_ab.restype = None
        _libs[path] = L
    return _libs[path]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _ele_ptr(ele):
    """Address of a BmxEle or of an ELE_DTYPE array."""
    if isinstance(ele, np.ndarray):
        return ele.ctypes.data
    return ctypes.addressof(ele)


class OracleError(RuntimeError):
    pass


def track1(ele, vec, beta, state=1, mc2=0.51099895069e6, charge=-1, p0c=None):
    """One particle through one element. Returns (vec, beta, state)."""
    v = np.array(vec, dtype=np.float64)
    b = ctypes.c_double(beta)
    s = ctypes.c_int32(state)
    if p0c is None:
        p0c = ele.p0c
    err = lib().bmo_track1(_ele_ptr(ele), mc2, charge, p0c, _ptr(v, _D), ctypes.byref(b),
                           ctypes.byref(s))
    if err:
        raise OracleError(f"bmo_track1 error {err}")
    return v, b.value, s.value


class RadConfig(ctypes.Structure):
    """bmx_rad_config (include/bmx.h)."""
    _fields_ = [("damping_on", ctypes.c_int32), ("fluctuations_on", ctypes.c_int32),
                ("zero_average", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("synch_rad_scale", ctypes.c_double), ("seed", ctypes.c_uint64)]


# bmx_rad_map (include/bmx.h), row-major matrices
RAD_MAP_DTYPE = np.dtype([("ref_orb", "<f8", (6,)), ("damp_dmat", "<f8", (6, 6)),
                          ("xfer_damp_vec", "<f8", (6,)), ("stoc_mat", "<f8", (6, 6))])


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block: four 32-bit words of the counter, two of the key."""
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint32 * 4)()
    lib().bmo_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def rad_gauss6(seed: int, pid: int, turn: int, ix1: int, edge: int) -> np.ndarray:
    """The six fluctuation deviates of particle pid at one edge of element ix1 (1-based)."""
    g = np.zeros(6)
    lib().bmo_rad_gauss6(int(seed), int(pid), int(turn), int(ix1), int(edge), _ptr(g, _D))
    return g


def track_bunch(ele_arr: np.ndarray, beam: dict, n_turns: int, turn0: int = 0, p0c=None,
                mc2=0.51099895069e6, charge=-1, n_threads: int = 0, fma: bool = False,
                lib_path: str | None = None, rad_maps: np.ndarray | None = None,
                rad: RadConfig | None = None) -> int:
    """Track a SoA beam dict in place; returns the number of particle-element passes.
    fma=True runs the FMA-contracted build of the same source; lib_path another build.
    rad_maps (shape (n_ele, 2) of RAD_MAP_DTYPE: rm0, rm1 per element) and rad: radiation kicks
    (track1_radiation)."""
    n = beam["x"].shape[0]
    if p0c is None:
        p0c = float(ele_arr[0]["p0c"])
    for k in ("x", "px", "y", "py", "z", "pz", "beta"):
        assert beam[k].dtype == np.float64 and beam[k].flags.c_contiguous
    for k in ("state", "loss_ele", "loss_turn"):
        assert beam[k].dtype == np.int32 and beam[k].flags.c_contiguous
    tracked = ctypes.c_int64(0)
    maps = None
    if rad_maps is not None:
        maps = np.ascontiguousarray(rad_maps, dtype=RAD_MAP_DTYPE)
        assert maps.shape == (len(ele_arr), 2)
        if rad is None:
            raise ValueError("rad_maps needs rad settings")
        rad = RadConfig(rad.damping_on, rad.fluctuations_on, rad.zero_average, 0,
                        rad.synch_rad_scale, rad.seed)
    err = lib(fma, lib_path).bmo_track_bunch_rad(
        ele_arr.ctypes.data, len(ele_arr), mc2, charge, p0c, n,
        *[_ptr(beam[k], _D) for k in ("x", "px", "y", "py", "z", "pz", "beta")],
        *[_ptr(beam[k], _I) for k in ("state", "loss_ele", "loss_turn")],
        n_turns, turn0, n_threads, ctypes.byref(tracked),
        maps.ctypes.data if maps is not None else None,
        ctypes.byref(rad) if maps is not None else None)
    if err:
        raise OracleError(f"bmo_track_bunch error {err}")
    return tracked.value


def track_trace(ele_arr: np.ndarray, vec, beta, n_turns: int, state=1, p0c=None,
                mc2=0.51099895069e6, charge=-1):
    """Coordinates after every element: array (n_turns * n_ele, 6)."""
    v = np.array(vec, dtype=np.float64)
    b = ctypes.c_double(beta)
    s = ctypes.c_int32(state)
    if p0c is None:
        p0c = float(ele_arr[0]["p0c"])
    out = np.zeros((n_turns * len(ele_arr), 6))
    err = lib().bmo_track_trace(ele_arr.ctypes.data, len(ele_arr), mc2, charge, p0c, _ptr(v, _D),
                                ctypes.byref(b), ctypes.byref(s), n_turns, _ptr(out, _D))
    if err:
        raise OracleError(f"bmo_track_trace error {err}")
    return out, v, b.value, s.value


def hwang_edge_kick(ele, vec, first_edge: bool, charge=-1):
    v = np.array(vec, dtype=np.float64)
    lib().bmo_hwang_bend_edge_kick(_ele_ptr(ele), charge, 1 if first_edge else 0, _ptr(v, _D))
    return v


def check_aperture(ele, vec, second_edge=True, state=1):
    v = np.array(vec, dtype=np.float64)
    s = ctypes.c_int32(state)
    u = ctypes.c_double(0.0)
    lib().bmo_check_aperture_limit(_ele_ptr(ele), 1 if second_edge else 0, _ptr(v, _D),
                                   ctypes.byref(s), ctypes.byref(u))
    return s.value, u.value


def multipole_kt_to_ab(knl, knsl, tn):
    """multipole_kt_to_ab (bmad/code/multipole_kt_to_ab.f90): integrated strengths knl, skew
    knsl and tilts tn of orders 0..21 -> (a, b)."""
    k = [np.zeros(22) for _ in range(3)]
    for arr, src in zip(k, (knl, knsl, tn)):
        for n, v in dict(src).items():
            arr[n] = v
    a = np.zeros(22)
    b = np.zeros(22)
    lib().bmo_multipole_kt_to_ab(*[_ptr(x, _D) for x in (*k, a, b)])
    return a, b
