"""Lattice element records and the Bmad bookkeeping needed to fill them.

An element is a ``bmx_ele`` record (include/bmx.h): the raw attribute values of Bmad's
``ele_struct`` (bmad/modules/bmad_struct.f90:1422-1530, ``ele%value`` at :1469). This module
builds those records the way Bmad's parser and bookkeeper would for the attributes a user
writes in a lattice file:

* element defaults -- bmad/parsing/set_ele_defaults.f90:36-37 (fringe_at = both_ends,
  fringe_type = none), :297-319 (sbend fringe_type = basic_bend, rfcavity num_steps = 10);
  fintx/hgapx default to fint/hgap (bmad/parsing/bmad_parser_mod.f90:5238-5239);
  a bare ``tilt`` means pi/4, pi/6, pi/8 for quad/sextupole/octupole
  (bmad/parsing/parser_set_attribute_mod.f90:1816-1822);
  ``x_limit``/``y_limit`` set both x1/x2 and y1/y2 limits
  (parser_set_attribute_mod.f90:2319-2324); aperture_at = exit_end, aperture_type =
  rectangular (bmad_struct.f90:1509-1510), elliptical for ecollimators.
* integration step -- bmad/code/attribute_bookkeeper.f90:292-398 with PTC's check_bend
  (forest/code/So_fitting.f90:3516-3548).
* reference energy -- p0c from E_tot (bmad/code/convert_pc_to.f90:20-45 and its inverse).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

# ---- constants (sim_utils/interfaces/physical_constants.f90:11-33) --------------------------
PI = 3.141592653589793238462643383279
TWOPI = 2 * PI
C_LIGHT = 2.99792458e8
M_ELECTRON = 0.51099895069e6
BMAD_DEFAULT_DS_STEP = 0.2  # bmad_struct.f90:2316

# ---- keys and enums (bmad_struct.f90; mirrored in include/bmx.h) ----------------------------
DRIFT, SBEND, QUADRUPOLE, SEXTUPOLE, RFCAVITY, MARKER, KICKER, OCTUPOLE = 1, 2, 3, 5, 9, 14, 15, 17
MONITOR, INSTRUMENT, HKICKER, VKICKER, RCOLLIMATOR, ECOLLIMATOR, PIPE = 31, 32, 33, 34, 35, 36, 44
THICK_MULTIPOLE = 67
# thin multipoles: multipole$ holds knl in a_pole, tn in b_pole and the skew knsl in
# a_pole_elec (pointer_to_ele_multipole.f90:30-34); ab_multipole$ holds a/b in a_pole/b_pole
MULTIPOLE, AB_MULTIPOLE = 19, 22
KEY_NAMES = {
    DRIFT: "drift", SBEND: "sbend", QUADRUPOLE: "quadrupole", SEXTUPOLE: "sextupole",
    RFCAVITY: "rfcavity", MARKER: "marker", KICKER: "kicker", OCTUPOLE: "octupole",
    MONITOR: "monitor", INSTRUMENT: "instrument", HKICKER: "hkicker", VKICKER: "vkicker",
    RCOLLIMATOR: "rcollimator", ECOLLIMATOR: "ecollimator", PIPE: "pipe",
    THICK_MULTIPOLE: "thick_multipole", MULTIPOLE: "multipole", AB_MULTIPOLE: "ab_multipole",
}
FRINGE_NONE, FRINGE_SOFT_EDGE_ONLY, FRINGE_HARD_EDGE_ONLY, FRINGE_FULL, FRINGE_BASIC_BEND = 1, 2, 3, 4, 7
# exact_multipoles (bmad_struct.f90:256-261)
EXACT_OFF, EXACT_HORIZONTALLY_PURE, EXACT_VERTICALLY_PURE = 1, 2, 3
ENTRANCE_END, EXIT_END, BOTH_ENDS, NO_END, CONTINUOUS = 1, 2, 3, 4, 5
APERTURE_RECTANGULAR, APERTURE_ELLIPTICAL = 2, 3

STATE_ALIVE, STATE_LOST, STATE_LOST_NEG_X, STATE_LOST_POS_X = 1, 2, 3, 4
STATE_LOST_NEG_Y, STATE_LOST_POS_Y, STATE_LOST_Z, STATE_LOST_PZ = 5, 6, 7, 8

N_POLE = 22

_INT_FIELDS = ["key", "is_on", "fringe_type", "fringe_at", "aperture_type", "aperture_at",
               "offset_moves_aperture", "n_multipole"]
_DBL_FIELDS = ["l", "k1", "k2", "k3", "g", "dg", "e1", "e2", "fint", "fintx", "hgap", "hgapx",
               "hkick", "vkick", "kick", "tilt", "ref_tilt", "x_offset", "y_offset",
               "z_offset", "x_pitch", "y_pitch", "voltage", "rf_frequency", "phi0",
               "phi0_multipass", "x1_limit", "x2_limit", "y1_limit", "y2_limit",
               "px_aperture_width2", "px_aperture_center", "py_aperture_width2",
               "py_aperture_center", "z_aperture_width2", "z_aperture_center",
               "pz_aperture_width2", "pz_aperture_center", "ds_step", "p0c", "e_tot", "r0_mag"]


class BmxEle(ctypes.Structure):
    """ctypes mirror of ``bmx_ele`` (include/bmx.h)."""
    _fields_ = ([(n, ctypes.c_int32) for n in _INT_FIELDS]
                + [(n, ctypes.c_double) for n in _DBL_FIELDS]
                + [("a_pole", ctypes.c_double * N_POLE), ("b_pole", ctypes.c_double * N_POLE)]
                + [("r0_elec", ctypes.c_double), ("a_pole_elec", ctypes.c_double * N_POLE),
                   ("b_pole_elec", ctypes.c_double * N_POLE), ("n_multipole_elec", ctypes.c_int32),
                   ("exact_multipoles", ctypes.c_int32), ("fq1", ctypes.c_double),
                   ("fq2", ctypes.c_double)])


ELE_DTYPE = np.dtype(BmxEle)
assert ELE_DTYPE.itemsize == ctypes.sizeof(BmxEle)


def e_tot_to_p0c(e_tot: float, mc2: float = M_ELECTRON) -> float:
    """p0c = sqrt(E^2 - m^2) (inverse of convert_pc_to.f90:37)."""
    return math.sqrt(e_tot ** 2 - mc2 ** 2)


def p0c_to_e_tot(p0c: float, mc2: float = M_ELECTRON) -> float:
    """E_tot = sqrt(pc^2 + m^2) (convert_pc_to.f90:37)."""
    return math.sqrt(p0c ** 2 + mc2 ** 2)


def _check_bend(xl: float, ggi: float, rhoi: float, xbend1: float):
    """PTC check_bend (forest/code/So_fitting.f90:3516-3548): (gf[1..7], met)."""
    gg = float(int(ggi))
    if gg == 0.0:
        gg = 1.0
    co = {2: 1.0 / 12.0, 4: 0.17, 6: 0.17e-1}
    gf = [0.0] * 8
    for i in (3, 5, 7):
        ar = i - 1
        gf[i] = (co[i - 1] / xbend1) ** (1.0 / ar) * abs(rhoi) * xl
        if gf[i] < gg:
            gf[i] = gg
    gf[2] = gf[3]
    gf[4] = gf[5] * 3
    gf[6] = gf[7] * 7
    met = 2
    if gf[4] < gf[2]:
        met = 4
    if gf[6] < gf[4] and gf[6] < gf[2]:
        met = 6
    return gf, met


def _nint(x: float) -> int:
    """Fortran NINT (half away from zero)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def _kick_ab(e) -> tuple[np.ndarray, np.ndarray]:
    """Body-frame a/b strengths incl. kicks, as multipole_ele_to_ab(..., include_kicks$)
    (bmad/code/multipole_ele_to_ab.f90:219-235,326-405; scaling :533-598)."""
    a = np.zeros(N_POLE)
    b = np.zeros(N_POLE)
    key = e.key
    if e.n_multipole > 0:
        a[:] = e.a_pole[:]
        b[:] = e.b_pole[:]
        const, ref_exp = None, 0
        if key == SBEND:
            const = e.l * (e.g + e.dg)
        elif key == KICKER:
            const = e.vkick if e.hkick == 0 else (e.hkick if e.vkick == 0 else math.hypot(e.hkick, e.vkick))
        elif key in (HKICKER, VKICKER):
            const = e.kick
        elif key == QUADRUPOLE:
            const, ref_exp = e.k1 * e.l, 1
        elif key == SEXTUPOLE:
            const, ref_exp = e.k2 * e.l, 2
        elif key == OCTUPOLE:
            const, ref_exp = e.k3 * e.l, 3
        if const is not None:
            if e.r0_mag == 0:
                a *= const
                b *= const
            else:
                factor = const * e.r0_mag ** (ref_exp + 1)
                for n in range(N_POLE):
                    factor /= e.r0_mag
                    a[n] *= factor
                    b[n] *= factor
    if key == HKICKER:
        b[0] += -e.kick
    elif key == VKICKER:
        a[0] += e.kick
    elif key == KICKER:
        b[0] += -e.hkick
        a[0] += e.vkick
    elif e.hkick != 0 or e.vkick != 0:
        t = e.ref_tilt if key == SBEND else e.tilt
        b[0] += -e.hkick * math.cos(t) - e.vkick * math.sin(t)
        a[0] += -e.hkick * math.sin(t) + e.vkick * math.cos(t)
    if key == QUADRUPOLE:
        b[1] += e.k1 * e.l
    elif key == SBEND:
        b[0] += e.dg * e.l
        b[1] += e.k1 * e.l
        b[2] += e.k2 * e.l / 2.0
    elif key == SEXTUPOLE:
        b[2] += e.k2 * e.l / 2.0
    elif key == OCTUPOLE:
        b[3] += e.k3 * e.l / 6.0
    return a, b


def bookkeep_ds_step(e, num_steps: int = 0, ds_step: float = 0.0) -> None:
    """Set e.ds_step as attribute_bookkeeper.f90:292-398 would for this element.

    num_steps / ds_step are the user-set values (0 = not set)."""
    L = e.l
    key = e.key
    if ds_step == 0 and num_steps == 0:
        if key in (DRIFT, PIPE, ECOLLIMATOR, RCOLLIMATOR, INSTRUMENT, MONITOR):
            num_steps, ds_step = 1, L
        elif key in (SBEND, QUADRUPOLE, SEXTUPOLE) and L != 0:
            a, b = _kick_ab(e)
            knl = np.array([math.factorial(n) * math.hypot(a[n], b[n]) for n in range(N_POLE)])
            bend_factor = knl[0] / L
            radius0 = e.r0_mag if e.r0_mag != 0 else 0.01
            if key == SBEND:
                bend_factor += max(abs(e.g), abs(e.g + e.dg))
                quad_factor = knl[1] + knl[2] * radius0 + L * bend_factor ** 2
            elif key == QUADRUPOLE:
                quad_factor = 50 * (knl[1] + L * bend_factor ** 2)
            else:
                quad_factor = knl[2] * radius0 + L * bend_factor ** 2
            gf, ixm = _check_bend(L, quad_factor, bend_factor, 1e-8)
            ixm = min(ixm, 4)
            num_steps = max(_nint(gf[ixm + 1]), 1)
            ds_step = abs(L) / num_steps
        elif key in (KICKER, HKICKER, VKICKER) and L != 0:
            kick = math.hypot(e.hkick, e.vkick) / L if key == KICKER else e.kick / L
            gf, ixm = _check_bend(L, 0.0, kick, 1e-8)
            num_steps = max(_nint(gf[ixm + 1]), 1)
            ds_step = abs(L) / num_steps
    if ds_step <= 0:
        if num_steps <= 0 or abs(L) == 0:
            ds_step = BMAD_DEFAULT_DS_STEP
        else:
            ds_step = abs(L) / num_steps
    e.ds_step = ds_step


_DEFAULT_TILT = {QUADRUPOLE: PI / 4.0, SEXTUPOLE: PI / 6.0, OCTUPOLE: PI / 8.0}


def make_element(key: int, p0c: float, e_tot: float, *, num_steps: int = 0, ds_step: float = 0.0,
                 x_limit: float | None = None, y_limit: float | None = None,
                 default_tilt: bool = False, a_pole=None, b_pole=None, a_pole_elec=None,
                 b_pole_elec=None, knl=None, tn=None, ksl=None, **attrs) -> BmxEle:
    """One element with Bmad defaults applied, then the user attributes, then bookkeeping.

    ``default_tilt=True`` is a bare ``tilt`` in a lattice file. a_pole / b_pole and
    a_pole_elec / b_pole_elec are {order: value} dicts of the magnetic and electric multipoles.
    knl / tn / ksl: a thin multipole's integrated strengths, tilts and skew strengths ({order:
    value}), stored where ele_struct keeps them (a_pole, b_pole, a_pole_elec)."""
    if knl is not None or tn is not None or ksl is not None:
        if key != MULTIPOLE:
            raise AttributeError("knl / tn / ksl belong to a multipole element")
        a_pole, b_pole = dict(knl or {}), dict(tn or {})
        if ksl:
            a_pole_elec = dict(ksl)
    e = BmxEle()
    e.key = key
    e.is_on = 1
    e.fringe_at = BOTH_ENDS
    e.fringe_type = FRINGE_NONE
    if key == SBEND:
        e.fringe_type = FRINGE_BASIC_BEND
    if key == RFCAVITY:
        e.fringe_type = FRINGE_FULL
        if num_steps == 0 and ds_step == 0:
            num_steps = 10
    e.aperture_at = EXIT_END
    e.aperture_type = APERTURE_ELLIPTICAL if key == ECOLLIMATOR else APERTURE_RECTANGULAR
    e.offset_moves_aperture = 1 if key == RCOLLIMATOR else 0
    e.p0c = p0c
    e.e_tot = e_tot
    fintx_set = "fintx" in attrs
    hgapx_set = "hgapx" in attrs
    if default_tilt:
        e.tilt = _DEFAULT_TILT[key]
    for k, v in attrs.items():
        if not hasattr(e, k):
            raise AttributeError(f"unknown element attribute {k!r}")
        setattr(e, k, v)
    if not fintx_set:
        e.fintx = e.fint
    if not hgapx_set:
        e.hgapx = e.hgap
    if x_limit is not None:
        e.x1_limit = e.x2_limit = x_limit
    if y_limit is not None:
        e.y1_limit = e.y2_limit = y_limit
    if a_pole is not None or b_pole is not None:
        ap = np.zeros(N_POLE)
        bp = np.zeros(N_POLE)
        if a_pole is not None:
            for n, v in dict(a_pole).items():
                ap[n] = v
        if b_pole is not None:
            for n, v in dict(b_pole).items():
                bp[n] = v
        e.a_pole[:] = ap
        e.b_pole[:] = bp
        e.n_multipole = N_POLE
    if a_pole_elec is not None or b_pole_elec is not None:
        for n, v in dict(a_pole_elec or {}).items():
            e.a_pole_elec[n] = v
        for n, v in dict(b_pole_elec or {}).items():
            e.b_pole_elec[n] = v
        e.n_multipole_elec = N_POLE
    bookkeep_ds_step(e, num_steps, ds_step)
    return e


def to_array(elements) -> np.ndarray:
    """Pack a sequence of BmxEle into a contiguous ELE_DTYPE array."""
    arr = np.zeros(len(elements), dtype=ELE_DTYPE)
    raw = arr.view(np.uint8).reshape(len(elements), ELE_DTYPE.itemsize)
    for i, e in enumerate(elements):
        raw[i] = np.frombuffer(bytes(e), dtype=np.uint8)
    return arr


def element_view(arr: np.ndarray, i: int) -> BmxEle:
    """BmxEle copy of record i of an ELE_DTYPE array."""
    return BmxEle.from_buffer_copy(arr[i:i + 1].tobytes())
