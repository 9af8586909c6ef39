"""The CPU restatement against closed forms where the reference ships no fixture.

* drift: straight line; dz = L (beta/beta_ref - 1/ps) evaluated in 50-digit decimal.
* quadrupole: transverse map = expm of the linear equations dx/ds = px/(1+pz),
  dpx/ds = -k1 x (scipy Pade expm); z = -1/2 int (dx/ds)^2 + (dy/ds)^2 ds (scipy quad)
  + L (beta - beta0)/beta0 (decimal).
* sbend body: exact circle / exit-face intersection built geometrically in long double
  (a different construction from Bmad's xi / L_c / angle_p formulas); semi-linear branch
  against the same construction at 1e-10 amplitudes.
* the one-turn map of the ring without RF is symplectic (numerical Jacobian).
A dropped term, a sign error or a swapped operand in any of these maps fails these checks.
"""
import decimal
import math

import numpy as np
import pytest
from scipy import integrate, linalg

D = decimal.Decimal
decimal.getcontext().prec = 50
M_E = 0.51099895069e6


def exact_beta_ratio_minus_1(p0c, pz):
    """beta/beta_ref - 1 for momentum p0c (1 + pz), in decimal."""
    p0 = D(float(p0c))
    p = p0 * (1 + D(float(pz)))
    m = D(float(M_E))
    beta = p / (p * p + m * m).sqrt()
    beta0 = p0 / (p0 * p0 + m * m).sqrt()
    return float(beta / beta0 - 1)


@pytest.mark.parametrize("pz", [0.0, 1e-3, -2e-3, 0.02, -0.3])
def test_drift_closed_form(bmx, oracle, pz):
    E = bmx.elements
    p0c, et = bmx.lattices.reference(5e9)
    L = 2.5
    ele = E.make_element(E.DRIFT, p0c, et, l=L)
    v0 = np.array([1e-3, 2e-4, -5e-4, 3e-4, 1e-3, pz])
    beta0 = float(bmx.lattices.init_beta(np.array([pz]), p0c, et)[0])
    v, beta, st = oracle.track1(ele, v0, beta0, p0c=p0c)
    rp = 1 + pz
    pxr, pyr = D(float(v0[1])) / D(float(rp)), D(float(v0[3])) / D(float(rp))
    ps = (1 - pxr * pxr - pyr * pyr).sqrt()
    assert v[0] == pytest.approx(v0[0] + float(D(L) * pxr / ps), rel=1e-14, abs=1e-18)
    assert v[2] == pytest.approx(v0[2] + float(D(L) * pyr / ps), rel=1e-14, abs=1e-18)
    dz = L * exact_beta_ratio_minus_1(p0c, pz) + float(D(L) * (1 - 1 / ps))
    assert v[4] - v0[4] == pytest.approx(dz, rel=1e-12, abs=1e-19)
    assert v[1] == v0[1] and v[3] == v0[3] and v[5] == v0[5]


def _quad_expected(k1, L, rel_p, x0, px0):
    A = np.array([[0.0, 1.0 / rel_p], [-k1, 0.0]])
    M = linalg.expm(A * L)
    xv = M @ np.array([x0, px0])

    def xprime2(s):
        ms = linalg.expm(A * s)
        return ((ms @ np.array([x0, px0]))[1] / rel_p) ** 2

    zint, _ = integrate.quad(xprime2, 0.0, L, epsabs=1e-22, epsrel=1e-13, limit=200)
    return xv, -0.5 * zint


@pytest.mark.parametrize("k1,pz", [(0.5, 0.0), (-0.5, 1e-3), (0.5, -4e-3), (2.0, 0.01), (1e-21, 0.0)])
def test_quadrupole_closed_form(bmx, oracle, k1, pz):
    E = bmx.elements
    p0c, et = bmx.lattices.reference(5e9)
    L = 0.3
    ele = E.make_element(E.QUADRUPOLE, p0c, et, l=L, k1=k1)
    v0 = np.array([1.5e-3, -2e-4, -7e-4, 3e-4, 2e-3, pz])
    beta0 = float(bmx.lattices.init_beta(np.array([pz]), p0c, et)[0])
    v, beta, st = oracle.track1(ele, v0, beta0, p0c=p0c)
    rp = 1 + pz
    (x, px), zx = _quad_expected(k1, L, rp, v0[0], v0[1])
    (y, py), zy = _quad_expected(-k1, L, rp, v0[2], v0[3])
    for got, want in ((v[0], x), (v[1], px), (v[2], y), (v[3], py)):
        assert got == pytest.approx(want, rel=1e-13, abs=1e-18)
    dz_expect = zx + zy + L * exact_beta_ratio_minus_1(p0c, pz)
    # low_energy_z_correction.f90:44-47 truncates beta/beta0 - 1 after pz^2: allow its pz^3 term
    series_err = 2 * L * abs(pz) ** 3 * (M_E / et) ** 2
    assert v[4] - v0[4] == pytest.approx(dz_expect, rel=1e-9, abs=2e-17 + series_err)


def _bend_geometric(g, L, x0, px0, y0, py0, z0, pz, beta, beta_ref):
    """Exact horizontal circle through a sector bend, long double."""
    ld = np.longdouble
    g, L = ld(g), ld(L)
    rho = 1 / g
    theta = g * L
    rel_p = 1 + ld(pz)
    pt = np.sqrt(rel_p * rel_p - ld(py0) ** 2)
    R = pt * rho
    sphi = ld(px0) / pt
    cphi = np.sqrt(1 - sphi * sphi)
    C = np.array([-rho, ld(0)])
    P0 = np.array([ld(x0), ld(0)])
    Cp = P0 + R * np.array([-cphi, sphi])
    u = np.array([np.cos(theta), np.sin(theta)])
    w = C - Cp
    b = u @ w
    c = w @ w - R * R
    disc = np.sqrt(b * b - c)
    roots = [-b + disc, -b - disc]
    t = min(roots, key=lambda r: abs(r - (rho + ld(x0))))
    Q = C + t * u
    n = (Cp - Q) / R
    d = np.array([n[1], -n[0]])
    a0 = P0 - Cp
    a1 = Q - Cp
    cross = a0[0] * a1[1] - a0[1] * a1[0]
    dot = a0 @ a1
    turned = np.arctan2(abs(cross), dot)
    Larc = R * turned
    return dict(x=t - rho, px=pt * (d @ u), y=ld(y0) + ld(py0) / pt * Larc,
                z=ld(z0) + ld(beta) * L / ld(beta_ref) - rel_p * Larc / pt)


@pytest.mark.parametrize("vec", [
    (1e-3, 2e-4, -5e-4, 1e-4, 1e-3, 2e-3),
    (-4e-3, -6e-4, 2e-3, -3e-4, 0.0, -5e-3),
    (1e-2, 1e-3, 0.0, 0.0, -2e-3, 0.01),
    (0.0, 0.0, 0.0, 0.0, 0.0, 1e-3),
])
def test_sbend_body_geometric(bmx, oracle, vec):
    E = bmx.elements
    p0c, et = bmx.lattices.reference(5e9)
    g, L = 0.015, 2.0
    ele = E.make_element(E.SBEND, p0c, et, l=L, g=g, fringe_type=E.FRINGE_NONE)
    pz = vec[5]
    beta = float(bmx.lattices.init_beta(np.array([pz]), p0c, et)[0])
    v, _, st = oracle.track1(ele, vec, beta, p0c=p0c)
    want = _bend_geometric(g, L, *vec[:5], pz, beta, p0c / et)
    assert st == 1
    assert v[0] == pytest.approx(float(want["x"]), rel=1e-12, abs=1e-15)
    assert v[1] == pytest.approx(float(want["px"]), rel=1e-11, abs=1e-16)
    assert v[2] == pytest.approx(float(want["y"]), rel=1e-12, abs=1e-15)
    assert v[4] == pytest.approx(float(want["z"]), rel=1e-10, abs=1e-15)
    assert v[3] == vec[3] and v[5] == vec[5]


def test_sbend_semilinear_branch(bmx, oracle):
    """At |x g|, |px|, |py|, |pz| < 1e-9 the semi-linear map equals the exact geometry to
    first order (second order terms ~1e-19 m)."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference(5e9)
    g, L = 0.015, 2.0
    ele = E.make_element(E.SBEND, p0c, et, l=L, g=g, fringe_type=E.FRINGE_NONE)
    vec = (3e-11, 4e-10, 2e-10, -3e-10, 1e-4, 5e-10)
    beta = float(bmx.lattices.init_beta(np.array([vec[5]]), p0c, et)[0])
    v, _, _ = oracle.track1(ele, vec, beta, p0c=p0c)
    want = _bend_geometric(g, L, *vec[:5], vec[5], beta, p0c / et)
    assert abs(v[0] - float(want["x"])) < 1e-17
    assert abs(v[1] - float(want["px"])) < 1e-17
    assert abs(v[2] - float(want["y"])) < 1e-17
    assert abs(v[4] - float(want["z"])) < 1e-16


def test_ring_one_turn_map_symplectic(bmx, oracle):
    """M^T J M = J for the numerical Jacobian of one turn of the 700-element ring (no RF)."""
    L = bmx.lattices
    arr = bmx.elements.to_array(L.ring0(n_cells=10))
    p0c, et = L.reference()
    v0 = np.array([2e-4, 1e-5, -1e-4, 2e-5, 1e-3, 1e-3])
    h = np.array([1e-7, 1e-8, 1e-7, 1e-8, 1e-6, 1e-7])
    M = np.zeros((6, 6))
    for j in range(6):
        cols = []
        for s in (+1, -1):
            v = v0.copy()
            v[j] += s * h[j]
            beta = float(L.init_beta(np.array([v[5]]), p0c, et)[0])
            _, vo, _, _ = oracle.track_trace(arr, v, beta, 1)
            cols.append(vo)
        M[:, j] = (cols[0] - cols[1]) / (2 * h[j])
    J = np.zeros((6, 6))
    for i in range(3):
        J[2 * i, 2 * i + 1] = 1
        J[2 * i + 1, 2 * i] = -1
    err = np.max(np.abs(M.T @ J @ M - J))
    assert err < 1e-6, err
    assert abs(np.linalg.det(M) - 1) < 1e-6
