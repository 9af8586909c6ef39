"""The CPU restatement of the thin multipoles (multipole$, ab_multipole$; track1_bmad.f90:142-168,
multipole_mod.f90:203-286, multipole_ele_to_kt.f90) against the reference's own numbers and
closed forms.

* multipole_kt_to_ab: the reference's multipole_test prints the a/b strengths of its thin
  multipole MM (k1l..k6l = 2..6, 1 with tilts t0..t5 = 0.1..0.6); the restatement must print the
  same digits (tests/golden/multipole_test_mm.correct).
* kicks: the thin quadrupole, sextupole and octupole kicks of the textbook field expansion
  (px' = px - Re[k_n (x + i y)^n / n!], py' = py + Im[...]); the skew quadrupole of MAD's
  convention (px' = px + k1s y, py' = py + k1s x); the dipole order of a multipole in the
  reference-orbit form, the first-order change of frame of a reference bent by k0l (px' = px +
  k0l pz, z' = z - k0l x); an offset element kicks at the displaced position.
* the three ways to rotate an order -- its own tilt tn, the element tilt, skew strengths knsl --
  and an ab_multipole rotated by its tilt take different branches of the restatement and must
  give the same kick.
A sign error, a dropped factorial or a swapped x/y in any branch fails one of these.
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _ele(bmx, key, **kw):
    E = bmx.elements
    p0c, et = bmx.lattices.reference(5e9)
    return E.make_element(key, p0c, et, **kw), p0c


def _kick(bmx, oracle, key, vec, **kw):
    ele, p0c = _ele(bmx, key, **kw)
    beta0 = float(bmx.lattices.init_beta(np.array([vec[5]]), p0c, ele.e_tot)[0])
    v, beta, st = oracle.track1(ele, np.array(vec, dtype=np.float64), beta0, p0c=p0c)
    assert st == 1
    return v


V0 = [1.3e-3, 2e-4, -0.7e-3, -1e-4, 3e-4, 2e-3]


def test_kt_to_ab_matches_reference_digits(oracle):
    """multipole_test's MULT lines for MM: a1, a3, a5, b2, b4, b6 as es12.4."""
    line = [ln for ln in open(os.path.join(HERE, "golden", "multipole_test_mm.correct"))
            if ln.startswith('"MULT-11-MM"')][0]
    want = line.split()[3:9]
    a, b = oracle.multipole_kt_to_ab({6: 1, 1: 2, 2: 3, 3: 4, 4: 5, 5: 6}, {},
                                     {0: 0.1, 1: 0.2, 2: 0.3, 3: 0.4, 4: 0.5, 5: 0.6})
    got = [f"{v:.4E}" for v in (a[1], a[3], a[5], b[2], b[4], b[6])]
    assert got == want, (got, want)


@pytest.mark.parametrize("n,k", [(1, 0.7), (2, 35.0), (3, -900.0)])
def test_normal_thin_kicks(bmx, oracle, n, k):
    """Order n of strength knl = k: the kick -k (x + i y)^n / n! (real part to px with a minus,
    imaginary part to py); x, y, z and pz unchanged."""
    E = bmx.elements
    v = _kick(bmx, oracle, E.MULTIPOLE, V0, knl={n: k})
    z = complex(V0[0], V0[2]) ** n * k / math.factorial(n)
    assert v[1] == pytest.approx(V0[1] - z.real, rel=1e-14, abs=1e-20)
    assert v[3] == pytest.approx(V0[3] + z.imag, rel=1e-14, abs=1e-20)
    assert [v[0], v[2], v[4], v[5]] == [V0[0], V0[2], V0[4], V0[5]]


def test_skew_quadrupole_convention(bmx, oracle):
    """k1sl = s: px' = px + s y, py' = py + s x."""
    E = bmx.elements
    s = 0.37
    v = _kick(bmx, oracle, E.MULTIPOLE, V0, ksl={1: s})
    assert v[1] == pytest.approx(V0[1] + s * V0[2], rel=1e-14)
    assert v[3] == pytest.approx(V0[3] + s * V0[0], rel=1e-14)


def test_dipole_reference_orbit_form(bmx, oracle):
    """k0l = th bends the reference orbit: px' = px + th pz, z' = z - th x (py, x, y
    unchanged); tilted by t0 the kick and the path change turn with it."""
    E = bmx.elements
    th = 2e-3
    v = _kick(bmx, oracle, E.MULTIPOLE, V0, knl={0: th})
    assert v[1] == pytest.approx(V0[1] + th * V0[5], rel=1e-14)
    assert v[4] == pytest.approx(V0[4] - th * V0[0], rel=1e-14)
    assert v[3] == V0[3] and v[0] == V0[0]
    t = 0.3
    v = _kick(bmx, oracle, E.MULTIPOLE, V0, knl={0: th}, tn={0: t})
    assert v[1] == pytest.approx(V0[1] + th * math.cos(t) * V0[5], rel=1e-14)
    assert v[3] == pytest.approx(V0[3] + th * math.sin(t) * V0[5], rel=1e-14)
    assert v[4] == pytest.approx(V0[4] - th * (math.cos(t) * V0[0] + math.sin(t) * V0[2]), rel=1e-14)


def test_ab_multipole_dipole_is_a_plain_kick(bmx, oracle):
    """An ab_multipole's order 0 is a kick of the particle, not of the reference: px' = px - b0,
    py' = py + a0, z unchanged."""
    E = bmx.elements
    v = _kick(bmx, oracle, E.AB_MULTIPOLE, V0, a_pole={0: 3e-5}, b_pole={0: -2e-5})
    assert v[1] == pytest.approx(V0[1] + 2e-5, rel=1e-14)
    assert v[3] == pytest.approx(V0[3] + 3e-5, rel=1e-14)
    assert v[4] == V0[4]


def test_rotations_agree(bmx, oracle):
    """A quadrupole order k turned by pi/4 through its tn, through the element tilt, as an
    ab_multipole b1 = k with tilt pi/4, and a skew strength knsl = -k: the same kick,
    px' = px - k y, py' = py - k x."""
    E = bmx.elements
    k = 0.9
    q = math.pi / 4
    kicks = [
        _kick(bmx, oracle, E.MULTIPOLE, V0, knl={1: k}, tn={1: q}),
        _kick(bmx, oracle, E.MULTIPOLE, V0, knl={1: k}, tilt=q),
        _kick(bmx, oracle, E.AB_MULTIPOLE, V0, b_pole={1: k}, tilt=q),
        _kick(bmx, oracle, E.MULTIPOLE, V0, ksl={1: -k}),
    ]
    for v in kicks:
        assert v[1] == pytest.approx(V0[1] - k * V0[2], rel=1e-13)
        assert v[3] == pytest.approx(V0[3] - k * V0[0], rel=1e-13)


def test_sextupole_rotations_agree(bmx, oracle):
    """A sextupole order turned by t three ways (tn, element tilt, ab_multipole tilt) and the
    same field given by normal plus skew strengths: one kick."""
    E = bmx.elements
    k, t = 40.0, 0.21
    a2, b2 = -k * math.sin(3 * t) / 2, k * math.cos(3 * t) / 2
    ref = _kick(bmx, oracle, E.AB_MULTIPOLE, V0, a_pole={2: a2}, b_pole={2: b2})
    z = complex(V0[0], V0[2]) ** 2
    assert ref[1] == pytest.approx(V0[1] - b2 * z.real + a2 * z.imag, rel=1e-13)
    assert ref[3] == pytest.approx(V0[3] + a2 * z.real + b2 * z.imag, rel=1e-13)
    for v in (_kick(bmx, oracle, E.MULTIPOLE, V0, knl={2: k}, tn={2: t}),
              _kick(bmx, oracle, E.MULTIPOLE, V0, knl={2: k}, tilt=t),
              _kick(bmx, oracle, E.AB_MULTIPOLE, V0, b_pole={2: k / 2}, tilt=t),
              _kick(bmx, oracle, E.MULTIPOLE, V0, knl={2: k * math.cos(3 * t)},
                    ksl={2: -k * math.sin(3 * t)})):
        assert v[1] == pytest.approx(ref[1], rel=1e-12)
        assert v[3] == pytest.approx(ref[3], rel=1e-12)


def test_offset_thin_multipole_kicks_at_displaced_position(bmx, oracle):
    """x_offset = dx, y_offset = dy: the kick is evaluated at (x - dx, y - dy) and the particle
    keeps its position."""
    E = bmx.elements
    k, dx, dy = 30.0, 4e-4, -2e-4
    v = _kick(bmx, oracle, E.MULTIPOLE, V0, knl={2: k}, x_offset=dx, y_offset=dy)
    z = complex(V0[0] - dx, V0[2] - dy) ** 2 * k / 2
    assert v[1] == pytest.approx(V0[1] - z.real, rel=1e-13)
    assert v[3] == pytest.approx(V0[3] + z.imag, rel=1e-13)
    assert v[0] == pytest.approx(V0[0], rel=1e-15) and v[2] == pytest.approx(V0[2], rel=1e-15)


def test_high_orders_use_the_reference_factorials(bmx, oracle):
    """Order 20 of the reference's factorial table (a 13-digit literal, 20! rounded): the
    restatement's kick equals k (x + i y)^20 / 2.432902008177e18, not the exact 20!."""
    E = bmx.elements
    x, y = 0.9, 0.2  # large amplitudes so the order-20 term is well above rounding
    vec = [x, 0.0, y, 0.0, 0.0, 0.0]
    k = 1e3
    v = _kick(bmx, oracle, E.AB_MULTIPOLE, vec, b_pole={20: k / 2.432902008177e18})
    z = complex(x, y) ** 20 * k / 2.432902008177e18
    assert v[1] == pytest.approx(-z.real, rel=1e-12)
    w = _kick(bmx, oracle, E.MULTIPOLE, vec, knl={20: k})
    # knl -> kick through c_multi's own 1/n! (computed by multiplication, exact to 1e-16 here)
    assert w[1] == pytest.approx(-(complex(x, y) ** 20).real * k / math.factorial(20), rel=1e-12)
