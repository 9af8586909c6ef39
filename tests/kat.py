"""Known-answer inputs of the reference's regression tests, restated as element records.

The element attributes are those of regression_tests/tracking_method_test/tracking_method_test.bmad
(lines 27-124), fringe_test/fringe_test.bmad and aperture_test/aperture_test.bmad; the
expected outputs are the committed fixture lines under tests/golden/. The comparison rule is
the reference harness's (regression_tests/test_fortran.py:208-288): values are written with the
same Fortran edit descriptor as the reference program and compared as numbers with the
line's ABS tolerance.
"""
from __future__ import annotations

import os
import shlex

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# particle_start of tracking_method_test.bmad:9-14, e_tot of l1 (:141)
TM_START = (0.001, 0.002, 0.003, 0.004, 0.005, 0.006)
TM_E_TOT = 1e6
# the lattice names no particle, so the reference species is Bmad's default, the positron
# (charge +1); only the electric multipole kicks and edges depend on its sign
TM_CHARGE = 1


def read_fixture(name):
    """{label: (type, tol, [values as strings])}"""
    out = {}
    with open(os.path.join(GOLDEN, name)) as f:
        for ln in f:
            ln = ln.strip()
            if not ln:
                continue
            parts = shlex.split(ln)
            out[parts[0]] = (parts[1], float(parts[2]), parts[3:])
    return out


def fortran_es(v: float, digits: int) -> str:
    """Fortran ESw.d output of one value (d digits after the point)."""
    return f"{v:.{digits}E}"


def abs_match(values, expected_strs, tol, digits):
    """test_fortran.py ABS rule on the formatted numbers; returns (ok, worst_diff)."""
    worst = 0.0
    for v, s in zip(values, expected_strs):
        d = abs(float(fortran_es(v, digits)) - float(s))
        worst = max(worst, d)
    return worst <= tol, worst


def tracking_method_elements(E):
    """Element records for the tracking_method_test entries covered by this path."""
    p0c = E.e_tot_to_p0c(TM_E_TOT)
    et = TM_E_TOT
    mk = E.make_element
    off = dict(x_offset=0.05, y_pitch=0.001)
    out = {
        "DRIFT1": mk(E.DRIFT, p0c, et, l=4.5, num_steps=10, tilt=0.1, **off),
        "ECOLLIMATOR1": mk(E.ECOLLIMATOR, p0c, et, l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10,
                           x_offset=0.02, y_pitch=0.001, tilt=0.1, hkick=0.01, vkick=0.002),
        "HKICKER1": mk(E.HKICKER, p0c, et, l=4.5, kick=0.005, num_steps=10, tilt=0.1, **off),
        "INSTRUMENT1": mk(E.INSTRUMENT, p0c, et, l=4.5, num_steps=10),
        "KICKER1": mk(E.KICKER, p0c, et, l=4.5, b_pole={0: 0.002}, vkick=0.003, num_steps=10,
                      tilt=0.1, **off),
        "MONITOR1": mk(E.MONITOR, p0c, et, l=4.5, num_steps=10),
        "OCTUPOLE1": mk(E.OCTUPOLE, p0c, et, l=4.5, k3=1e3, default_tilt=True, num_steps=10,
                        x_offset=0.01, y_pitch=0.001),
        "RCOLLIMATOR1": mk(E.RCOLLIMATOR, p0c, et, l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10,
                           x_offset=0.02, y_pitch=0.001, tilt=0.1),
        "RFCAVITY1": mk(E.RFCAVITY, p0c, et, l=4.5, rf_frequency=500e6, voltage=1e5, phi0=0.1,
                        num_steps=10, tilt=0.1, **off),
        "RFCAVITY3": mk(E.RFCAVITY, p0c, et, l=0.0, rf_frequency=500e6, voltage=1e5, phi0=0.1,
                        num_steps=10, tilt=0.1, a_pole={2: 0.02}, **off),
        "SEXTUPOLE1": mk(E.SEXTUPOLE, p0c, et, l=0.6, k2=1.0, default_tilt=True, num_steps=10,
                         **off),
        "THICK_MULTIPOLE1": mk(E.THICK_MULTIPOLE, p0c, et, l=4.5, a_pole={2: 1.0},
                               b_pole={3: 20.0}, x_offset=0.02, tilt=0.1),
        "VKICKER1": mk(E.VKICKER, p0c, et, l=4.5, kick=0.005, num_steps=10, tilt=0.1, **off),
    }
    # q0: quadrupole, a0 = 0.01, a1 = 0.1, vkick = -0.0001, b0_elec = -2e3, a1_elec = 2e4
    # (tracking_method_test.bmad:62), then quadrupole1-5 (:63-67; 4 and 5 with fringe_type = full)
    q0 = dict(a_pole={0: 0.01, 1: 0.1}, vkick=-0.0001, b_pole_elec={0: -2e3},
              a_pole_elec={1: 2e4}, l=0.6, default_tilt=True, num_steps=10, **off)
    out.update({
        "QUADRUPOLE1": mk(E.QUADRUPOLE, p0c, et, k1=1.0, **q0),
        "QUADRUPOLE2": mk(E.QUADRUPOLE, p0c, et, k1=-1.0, **q0),
        "QUADRUPOLE3": mk(E.QUADRUPOLE, p0c, et, k1=1e-20, **q0),
        "QUADRUPOLE4": mk(E.QUADRUPOLE, p0c, et, k1=1.0, fringe_type=E.FRINGE_FULL, **q0),
        "QUADRUPOLE5": mk(E.QUADRUPOLE, p0c, et, k1=-1.0, fringe_type=E.FRINGE_FULL, **q0),
    })
    # sb0: sbend, a0 = 0.01, a1 = 0.1, vkick = -0.0001, b0_elec = -2e3, a1_elec = 2e4,
    # r0_mag = 2, r0_elec = 3 (tracking_method_test.bmad:79), then sbend1/3/5/7 (:82-91)
    sb0 = dict(a_pole={0: 0.01, 1: 0.1}, vkick=-0.0001, b_pole_elec={0: -2e3},
               a_pole_elec={1: 2e4}, r0_mag=2.0, r0_elec=3.0, l=0.6, num_steps=10,
               ref_tilt=0.1, **off)
    out.update({
        "SBEND1": mk(E.SBEND, p0c, et, k1=0.3, exact_multipoles=E.EXACT_VERTICALLY_PURE, **sb0),
        "SBEND3": mk(E.SBEND, p0c, et, k1=0.3, k2=3.0, **sb0),
        "SBEND5": mk(E.SBEND, p0c, et, g=0.01, e1=0.1, e2=0.2, fringe_type=E.FRINGE_FULL,
                     fint=0.7, hgap=2.0, **sb0),
        "SBEND7": mk(E.SBEND, p0c, et, g=0.01, e1=0.1, e2=0.2, **sb0),
    })
    return out


# tracking_method_test lines in the fixture whose elements are not restated yet
NOT_YET = ()


def tm_start_beta(E):
    p0c = E.e_tot_to_p0c(TM_E_TOT)
    pc = p0c * (1 + TM_START[5])
    return pc / (pc ** 2 + E.M_ELECTRON ** 2) ** 0.5


def fringe_test_element(E):
    """fringe_test.bmad: sbend e1=0.1, e2=-0.2, k1=0.1, g=0.1, fint=0.5, hgap=0.4, 1 MeV e-."""
    p0c = E.e_tot_to_p0c(1e6)
    return E.make_element(E.SBEND, p0c, 1e6, l=0.0, e1=0.1, e2=-0.2, k1=0.1, g=0.1, fint=0.5,
                          hgap=0.4)


FRINGE_START = (0.01, 0.02, 0.03, 0.04, 0.0, 0.5)


def aperture_test_elements(E):
    """aperture_test.bmad ec / rc elements (antimuon, p0c = 3.094353005e9)."""
    p0c = 3.094353005e9
    et = E.p0c_to_e_tot(p0c, 105.6583755e6)
    common = dict(px_aperture_width2=0.01, px_aperture_center=0.01, py_aperture_width2=0.015,
                  z_aperture_width2=0.2, z_aperture_center=-0.6, pz_aperture_width2=0.03,
                  pz_aperture_center=0.02, x1_limit=1e-10, x2_limit=0.02, y_limit=0.015)
    ec = E.make_element(E.ECOLLIMATOR, p0c, et, **common)
    rc = E.make_element(E.RCOLLIMATOR, p0c, et, **common)
    return {"EC": ec, "RC": rc}


def aperture_test_cases():
    """(suffix, list of 9 vec) as aperture_test.f90:104-200 builds them."""
    import numpy as np

    def r4(v):
        # aperture_test.f90 assigns default-real (single precision) literals such as
        # ``xw2 = 0.2`` to real(rp) variables, so the values are float32-rounded.
        return float(np.float32(v))

    cases = []
    for sfx, base, eps_sign in (("Pxy-", "p", -1), ("Pxy+", "p", +1), ("xy-", "x", -1),
                                ("xy+", "x", +1), ("zPz-", "z", -1), ("zPz+", "z", +1)):
        vecs = []
        eps = r4(1e-5)
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                if base == "p":
                    xw2, x0, yw2, y0 = r4(0.01), r4(0.01), r4(0.015), 0.0
                    vecs.append([0.0, x0 + i * (xw2 + eps_sign * eps), 0.0,
                                 y0 + j * (yw2 + eps_sign * eps), -0.6, 0.02])
                elif base == "x":
                    xw2, x0, yw2, y0 = r4(0.01), r4(0.01), r4(0.015), 0.0
                    px = 0.1 if eps_sign < 0 else 0.0
                    vecs.append([x0 + i * (xw2 + eps_sign * eps), px,
                                 y0 + j * (yw2 + eps_sign * eps), 0.0, -0.6, 0.02])
                else:
                    xw2, x0, yw2, y0 = r4(0.2), r4(-0.6), r4(0.03), r4(0.02)
                    vecs.append([0.0, 0.01, 0.0, 0.0, x0 + i * (xw2 + eps_sign * eps),
                                 y0 + j * (yw2 + eps_sign * eps)])
        cases.append((sfx, vecs))
    return cases
