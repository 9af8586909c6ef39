"""The CPU restatement against the reference's own known-answer fixtures (no GPU)."""
import pytest

import kat


def test_tracking_method_known_answers(bmx, oracle):
    """tracking_method_test "<ELE>: Bmad_Standard" lines (ABS 2e-14, written es18.10):
    drift, thick multipoles (sextupole, octupole, thick_multipole, kickers, collimators,
    instrument, monitor) with offsets/pitch/tilt, the rf cavity (thick and zero length),
    quadrupoles with magnetic and electric multipoles (stepped body, electric kicks that change
    the energy, electric hard-edge energy kicks, hard and soft edges), and sector bends with
    offsets, pitch and ref_tilt (the bend branch of offset_particle), magnetic and electric
    multipoles, k1 (the k1 body map), k2, exact_multipoles with g = 0 and the full (exact) edge,
    tracked as a positron."""
    E = bmx.elements
    fx = kat.read_fixture("tracking_method_test.correct")
    eles = kat.tracking_method_elements(E)
    beta0 = kat.tm_start_beta(E)
    failures = []
    for name, ele in eles.items():
        label = f"{name}: Bmad_Standard"
        typ, tol, expect = fx[label]
        assert typ == "ABS"
        vec, beta, state = oracle.track1(ele, kat.TM_START, beta0, charge=kat.TM_CHARGE)
        assert state == E.STATE_ALIVE, name
        ok, worst = kat.abs_match(vec, expect[:6], tol, 10)
        if not ok:
            failures.append((name, worst, [kat.fortran_es(v, 10) for v in vec], expect[:6]))
    assert not failures, failures
    # every line of the fixture is restated
    assert set(f"{n}: Bmad_Standard" for n in eles) | set(f"{n}: Bmad_Standard" for n in kat.NOT_YET) == set(fx)


def test_fringe_known_answers(bmx, oracle):
    """fringe_test Up-Orb / Dn-Orb: hwang_bend_edge_kick alone (ABS 1e-12, written es14.6)."""
    E = bmx.elements
    fx = kat.read_fixture("fringe_test.correct")
    ele = kat.fringe_test_element(E)
    for label, first in (("Up-Orb", True), ("Dn-Orb", False)):
        typ, tol, expect = fx[label]
        out = oracle.hwang_edge_kick(ele, kat.FRINGE_START, first)
        d = [o - s for o, s in zip(out, kat.FRINGE_START)]
        ok, worst = kat.abs_match(d, expect, tol, 6)
        assert ok, (label, worst, [kat.fortran_es(v, 6) for v in d], expect)


def test_aperture_known_answers(bmx, oracle):
    """aperture_test EC_S_*/RC_S_* state codes (exact) and EC_U_*/RC_U_* unstable factors
    (es10.2, ABS 1e-8) for elliptical and rectangular apertures with momentum windows."""
    E = bmx.elements
    fx = kat.read_fixture("aperture_test.correct")
    eles = kat.aperture_test_elements(E)
    for ename, ele in eles.items():
        for sfx, vecs in kat.aperture_test_cases():
            states, unst = [], []
            for v in vecs:
                s, u = oracle.check_aperture(ele, v, second_edge=True)
                states.append(s)
                unst.append(u)
            _, _, exp_s = fx[f"{ename}_S_{sfx}"]
            assert states == [int(x) for x in exp_s], (ename, sfx, states, exp_s)
            _, tol, exp_u = fx[f"{ename}_U_{sfx}"]
            got = [float(f"{u:.2E}") for u in unst]
            assert all(abs(g - float(x)) <= tol for g, x in zip(got, exp_u)), (ename, sfx, got, exp_u)


@pytest.mark.parametrize("x", [0.0, 1e-9, 1e-8, 0.099, 0.1, 0.101, -0.3, 1.2])
def test_special_function_branches(oracle, x):
    """sinc/cosc/sincc/sqrt_one near their branch points against direct or series forms."""
    import math
    L = oracle.lib()
    s = L.bmo_sinc(x)
    assert s == (1.0 if abs(x) < 1e-8 else math.sin(x) / x)
    c = L.bmo_cosc(x)
    if abs(x) > 1e-3:
        assert abs(c - (1 - math.cos(x)) / x ** 2) < 1e-12
    sc = L.bmo_sincc(x)
    if abs(x) > 1e-2:
        assert abs(sc - (x - math.sin(x)) / x ** 3) < 1e-10
    else:
        assert abs(sc - 1.0 / 6.0) < 1e-4
    so = L.bmo_sqrt_one(x)
    assert abs(so - (math.sqrt(1 + x) - 1)) <= 1e-15 * max(1.0, abs(x))


def test_c_multi_is_signed_binomial(oracle):
    """c_multi(n, m, no_n_fact) = -+ binomial(n, m), sign - for m mod 4 in {0, 3}."""
    import math
    L = oracle.lib()
    for n in range(0, 12):
        for m in range(0, n + 1):
            sign = -1 if m % 4 in (0, 3) else 1
            assert abs(L.bmo_c_multi(n, m) - sign * math.comb(n, m)) <= 1e-12 * math.comb(n, m)


def test_modulo2(oracle):
    L = oracle.lib()
    import math
    for x in (-7.0, -3.5, -0.1, 0.0, 0.2, 3.3, 12.0):
        m = L.bmo_modulo2(x, math.pi)
        assert -math.pi <= m < math.pi
        k = (x - m) / (2 * math.pi)
        assert abs(k - round(k)) < 1e-12
