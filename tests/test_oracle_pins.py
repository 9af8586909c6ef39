"""Pins of CPU-restatement functions that no reference fixture covers.

orbit_too_large (bmad/low_level/orbit_too_large.f90:46-118): a branch table written from the
reference text -- the order of the tests (x before y before momentum), the strict comparisons,
the |px| > |py| tie-break towards y, the sign of the lost side, and the unstable factor each
branch leaves in param%unstable_factor. The restatement is reached through its
check_aperture_limit entry (which calls orbit_too_large first, check_aperture_limit.f90:61) on
an element without aperture limits.

convert_this_ab's r0_mag scaling (bmad/code/multipole_ele_to_ab.f90:577-598): with r0_mag = r
the order-n strengths are scaled by const * r^(ref_exp - n), with r0_mag = 0 by const alone
(ref_exp = 0 kickers, 1 quadrupole, 2 sextupole, 3 octupole). An element with r0_mag = r must
therefore track exactly like the same element with r0_mag = 0 and its a_n, b_n pre-multiplied
by r^(ref_exp - n); a wrong exponent, a loop that starts one order off or a division in place
of a multiplication breaks the equality.
"""
import math

import numpy as np
import pytest

LOST_NEG_X, LOST_POS_X, LOST_NEG_Y, LOST_POS_Y, LOST_PZ = 3, 4, 5, 6, 8

# (x, px, y, py, pz) -> (state, unstable factor or None when the particle is not flagged)
_TABLE = [
    # position tests, strict, x first
    ((1001.0, 0, 0, 0, 0), (LOST_POS_X, 1e-3)),
    ((-1001.0, 0, 0, 0, 0), (LOST_NEG_X, 1e-3)),
    ((0, 0, 1001.0, 0, 0), (LOST_POS_Y, 1e-3)),
    ((0, 0, -1001.0, 0, 0), (LOST_NEG_Y, 1e-3)),
    ((1001.0, 0, -2000.0, 0, 0), (LOST_POS_X, 1e-3)),      # x is tested before y
    ((-1500.0, 0, 1500.0, 0, -1.5), (LOST_NEG_X, 0.5)),    # position before momentum
    ((0, 0, 1200.0, 2.0, -1.5), (LOST_POS_Y, 0.2)),
    ((1000.0, 0, -1000.0, 0, 0), (1, None)),               # |x| = limit is not beyond it
    # momentum: rel_p < 0
    ((0, 0, 0, 0, -1.1), (LOST_PZ, 0.1 + 1e-6)),
    ((0, 3.0, 0, 0, -1.25), (LOST_PZ, 0.25 + 1e-6)),       # rel_p before the transverse test
    ((0, 0, 0, 0, -1.0), (1, None)),                        # rel_p = 0, px = py = 0: f = 0
    # momentum: px^2 + py^2 > (1 + pz)^2, side from the larger of |px|, |py|
    ((0, 1.2, 0, 0.5, 0), (LOST_POS_X, math.sqrt(0.69))),
    ((0, -1.2, 0, 0.5, 0), (LOST_NEG_X, math.sqrt(0.69))),
    ((0, 0.5, 0, 1.2, 0), (LOST_POS_Y, math.sqrt(0.69))),
    ((0, 0.5, 0, -1.2, 0), (LOST_NEG_Y, math.sqrt(0.69))),
    ((0, 0.8, 0, 0.8, 0.1), (LOST_POS_Y, math.sqrt(0.07))),    # |px| = |py|: the y branch
    ((0, -0.8, 0, -0.8, 0.1), (LOST_NEG_Y, math.sqrt(0.07))),
    ((0, 0.3, 0, 0.2, -0.7), (LOST_POS_X, 0.2)),            # 0.09 + 0.04 > 0.3^2 off momentum
    ((0, 1.0, 0, 0, 0), (1, None)),                         # f = 0 is not unstable
    ((0, 0.6, 0, 0.79, 0), (1, None)),
    ((5.0, 0.01, -3.0, 0.02, 0.3), (1, None)),
]


def _vec(row):
    x, px, y, py, pz = row
    return np.array([x, px, y, py, 0.0, pz], dtype=np.float64)


@pytest.mark.parametrize("row,expect", _TABLE)
def test_orbit_too_large_branch_table(bmx, oracle, row, expect):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    marker = E.make_element(E.MARKER, p0c, et)
    state, u = oracle.check_aperture(marker, _vec(row), second_edge=True)
    want_state, want_u = expect
    assert state == want_state
    if want_u is not None:
        assert u == pytest.approx(want_u, rel=1e-12)


def test_orbit_too_large_in_track1(bmx, oracle):
    """track1 flags a particle that arrives beyond the limits at the entrance of any element
    (the first-edge check, track1.f90:196) and leaves its coordinates unchanged."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    for key, kw in ((E.DRIFT, dict(l=1.0)), (E.QUADRUPOLE, dict(l=0.3, k1=0.5)),
                    (E.SBEND, dict(l=2.0, g=0.01)), (E.SEXTUPOLE, dict(l=0.2, k2=20.0))):
        ele = E.make_element(key, p0c, et, **kw)
        for row, (want_state, _) in _TABLE:
            if want_state == 1:
                continue
            v0 = _vec(row)
            v, _, st = oracle.track1(ele, v0, 0.99, p0c=p0c)
            assert st == want_state, (key, row)
            assert np.array_equal(v, v0)


def _with_poles(E, p0c, et, key, kw, r0, scale_for_zero_r0=None):
    ele = E.make_element(key, p0c, et, **kw)
    a = [0.0] * E.N_POLE
    b = [0.0] * E.N_POLE
    base_a = {0: 3e-4, 1: -2e-3, 2: 0.7, 3: 4.0}
    base_b = {0: -5e-4, 1: 1e-3, 2: -0.9, 3: 6.0}
    for n in range(4):
        a[n] = base_a[n] * (scale_for_zero_r0(n) if scale_for_zero_r0 else 1.0)
        b[n] = base_b[n] * (scale_for_zero_r0(n) if scale_for_zero_r0 else 1.0)
    ele.a_pole[:] = a
    ele.b_pole[:] = b
    ele.n_multipole = 4
    ele.r0_mag = r0
    return ele


@pytest.mark.parametrize("key_name,kw,ref_exp", [
    ("HKICKER", dict(l=0.4, kick=2e-4), 0),
    ("QUADRUPOLE", dict(l=0.3, k1=0.8), 1),
    ("SEXTUPOLE", dict(l=0.2, k2=20.0), 2),
    ("OCTUPOLE", dict(l=0.25, k3=300.0), 3),
])
@pytest.mark.parametrize("r0", [0.05, 0.3, 2.0])
def test_r0_mag_scaling(bmx, oracle, key_name, kw, ref_exp, r0):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    key = getattr(E, key_name)
    with_r0 = _with_poles(E, p0c, et, key, kw, r0)
    prescaled = _with_poles(E, p0c, et, key, kw, 0.0, lambda n: r0 ** (ref_exp - n))
    rng = np.random.default_rng(17)
    kicked = False
    for _ in range(20):
        v0 = np.concatenate([rng.normal(0, 3e-3, 4), rng.normal(0, 1e-3, 2)])
        v1, b1, s1 = oracle.track1(with_r0, v0, 0.99999, p0c=p0c)
        v2, b2, s2 = oracle.track1(prescaled, v0, 0.99999, p0c=p0c)
        assert s1 == s2 == 1
        scale = np.maximum(np.abs(v2), 1e-9)
        assert np.all(np.abs(v1 - v2) <= 1e-13 * scale), (v1, v2)
        # the poles act: tracking without them gives another orbit
        plain = E.make_element(key, p0c, et, **kw)
        v3, _, _ = oracle.track1(plain, v0, 0.99999, p0c=p0c)
        kicked |= bool(np.any(np.abs(v3 - v1) > 1e-9))
    assert kicked


def test_r0_mag_one_is_unscaled(bmx, oracle):
    """r0_mag = 1 multiplies every order by 1: identical bits to r0_mag = 0."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    a = _with_poles(E, p0c, et, E.SEXTUPOLE, dict(l=0.2, k2=20.0), 1.0)
    b = _with_poles(E, p0c, et, E.SEXTUPOLE, dict(l=0.2, k2=20.0), 0.0)
    v0 = np.array([2e-3, -1e-4, 1e-3, 2e-4, 1e-3, 1e-3])
    assert np.array_equal(oracle.track1(a, v0, 0.99999, p0c=p0c)[0],
                          oracle.track1(b, v0, 0.99999, p0c=p0c)[0])


# ---- soft quadrupole edge (bmad/modules/fringe_mod.f90:556-661) ----------------------------
# No reference fixture has fq1 or fq2 set. Pins: with fq2 = 0 the exit kick (f1 -> -f1)
# undoes the entrance kick exactly (x e^f1 e^-f1, and the z terms -f1 x vx + f1 y vy cancel);
# the entrance-exit pair of a zero-length quadrupole is symplectic in (x, px, y, py, z, pz)
# (a dropped or mis-signed z term breaks it); and it reduces to the closed form of the f2
# terms alone when fq1 = 0 (x += vx f2 at each edge, y -= vy f2, z -= f2 (vx^2 - vy^2)).

def _soft_quad(bmx, fq1, fq2, k1=2.0):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    return E.make_element(E.QUADRUPOLE, p0c, et, l=0.0, k1=k1, fq1=fq1, fq2=fq2,
                          fringe_type=E.FRINGE_SOFT_EDGE_ONLY), p0c


def test_soft_quadrupole_edges_cancel_without_fq2(bmx, oracle):
    ele, p0c = _soft_quad(bmx, 0.3, 0.0)
    v0 = np.array([1e-3, 2e-4, -3e-3, 5e-4, 1e-3, 2e-3])
    v, _, st = oracle.track1(ele, v0, 0.99999, p0c=p0c)
    assert st == 1
    assert np.allclose(v, v0, rtol=0, atol=1e-18), v - v0
    # and a nonzero f1 does act on its own: the entrance half alone is not the identity
    ele.fringe_at = bmx.elements.ENTRANCE_END
    v1, _, _ = oracle.track1(ele, v0, 0.99999, p0c=p0c)
    f1 = 2.0 / (1 + v0[5]) * 0.3
    assert v1[0] == pytest.approx(v0[0] * math.exp(f1), rel=1e-15)
    assert v1[3] == pytest.approx(v0[3] * math.exp(f1), rel=1e-15)


def test_soft_quadrupole_f2_closed_form(bmx, oracle):
    ele, p0c = _soft_quad(bmx, 0.0, 0.05)
    v0 = np.array([1e-3, 2e-4, -3e-3, 5e-4, 1e-3, 2e-3])
    v, _, _ = oracle.track1(ele, v0, 0.99999, p0c=p0c)
    rp = 1 + v0[5]
    vx, vy = v0[1] / rp, v0[3] / rp
    f2 = 2.0 / rp * 0.05
    assert v[0] == pytest.approx(v0[0] + 2 * vx * f2, rel=1e-14)
    assert v[2] == pytest.approx(v0[2] - 2 * vy * f2, rel=1e-14)
    assert v[1] == v0[1] and v[3] == v0[3]
    assert v[4] == pytest.approx(v0[4] - 2 * f2 * (vx * vx - vy * vy), rel=1e-12)


def test_soft_quadrupole_edges_symplectic(bmx, oracle):
    ele, p0c = _soft_quad(bmx, 0.4, 0.07)
    v0 = np.array([2e-3, -1e-3, 1e-3, 2e-3, 0.0, 1e-2])
    h = 1e-7
    J = np.zeros((6, 6))
    for j in range(6):
        a, b = v0.copy(), v0.copy()
        a[j] += h
        b[j] -= h
        va, _, _ = oracle.track1(ele, a, 0.99999, p0c=p0c)
        vb, _, _ = oracle.track1(ele, b, 0.99999, p0c=p0c)
        J[:, j] = (va - vb) / (2 * h)
    S = np.zeros((6, 6))
    for i in range(3):
        S[2 * i, 2 * i + 1] = 1
        S[2 * i + 1, 2 * i] = -1
    err = np.abs(J.T @ S @ J - S).max()
    assert err < 1e-8, err
    # the test has teeth: the same map with its z change negated (z_out -> 2 z_in - z_out) is
    # not symplectic
    Jf = J.copy()
    Jf[4, :] = -J[4, :]
    Jf[4, 4] += 2
    assert np.abs(Jf.T @ S @ Jf - S).max() > 1e-3
