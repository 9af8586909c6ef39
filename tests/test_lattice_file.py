"""The lattice-file loader (lattice_file.py, a subset of bmad/parsing/bmad_parser.f90): a test
ring written in the Bmad language, two of the reference's tutorial lattices kept as fixtures,
and overlays, groups, start_branch_at and the attributes of the general element maps."""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RING = os.path.join(HERE, "golden", "lattices", "ring_small.bmad")
TUTORIAL = os.path.join(HERE, "golden", "lattices", "tutorial")


@pytest.fixture(scope="module")
def lf(bmx):
    import importlib
    return importlib.import_module(bmx.__name__ + ".lattice_file")


def _validate(bmx, arr, mc2, charge):
    import ctypes
    L = bmx.native.load()
    n = ctypes.c_int32()
    rc = L.bmx_lattice_validate(arr.ctypes.data, len(arr), mc2, charge, ctypes.byref(n))
    return rc, (L.bmx_last_error().decode() if rc else "")


def test_small_ring(bmx, lf):
    """12 FODO cells with sextupoles, a corrector, a monitor per arc and an rf cavity: element
    count, closed geometry (bend angles sum to 2 pi), the rbend-to-sbend conversion (arc length
    from the chord, edge angles + angle/2), inheritance and attribute references; the table
    compiles for the GPU path."""
    E = bmx.elements
    eles, mc2, charge = lf.load(RING)
    arr = E.to_array(eles)
    assert len(arr) == 148 and charge == -1 and mc2 == E.M_ELECTRON
    bends = arr[arr["key"] == E.SBEND]
    ang = 2 * math.pi / 24
    assert len(bends) == 24
    assert abs(float((bends["g"] * bends["l"]).sum()) - 2 * math.pi) < 1e-12
    assert np.allclose(bends["l"], 2.4 * ang / (2 * math.sin(ang / 2)), rtol=1e-15)
    assert np.allclose(bends["e1"], 0.01 + ang / 2, rtol=1e-15)
    q = arr[arr["key"] == E.QUADRUPOLE]
    assert sorted(set(np.round(q["k1"], 12))) == [round(-0.62 * 0.98, 12), 0.62]
    d = arr[arr["key"] == E.DRIFT]
    assert sorted(set(np.round(d["l"], 12))) == [0.55, 1.0]
    assert (arr["key"] == E.RFCAVITY).sum() == 1 and (arr["key"] == E.HKICKER).sum() == 1
    assert arr[0]["p0c"] == E.e_tot_to_p0c(3e9)
    rc, msg = _validate(bmx, arr, mc2, charge)
    assert rc == 0, msg


@pytest.mark.parametrize("text", [
    "parameter[e_tot] = 1e9\nd: drift, l = 1\nr: line = (d, --d)\nuse, r\n",
    "parameter[e_tot] = 1e9\nd: drift, l = 1\no: overlay = {d[l]: x}\nr: line = (d)\nuse, r\n",
    "parameter[e_tot] = 1e9\nd: drift, l = 1\ng: girder = {d}\nr: line = (d)\nuse, r\n",
    "parameter[e_tot] = 1e9\nd: drift, l = 1, superimpose, ref = d\nr: line = (d)\nuse, r\n",
    "parameter[e_tot] = 1e9\ns: solenoid, l = 1, ks = 0.1\nr: line = (s)\nuse, r\n",
    "d: drift, l = 1\nr: line = (d)\nuse, r\n",
])
def test_outside_the_subset_raises(lf, tmp_path, text):
    p = tmp_path / "lat.bmad"
    p.write_text(text)
    with pytest.raises(lf.LatticeFileError):
        lf.load(str(p))


def test_reflection_and_repetition(lf, tmp_path):
    """'-line' is the line in reverse order (the reference's reflection); 'n*line' repeats."""
    p = tmp_path / "lat.bmad"
    p.write_text("parameter[e_tot] = 1e9\n"
                 "a: drift, l = 1\nb: drift, l = 2\nc: drift, l = 3\nd: drift, l = 4\n"
                 "s: line = (c, d)\nr: line = (a, b, -s, 2*(a, -(b, c)))\nuse, r\n")
    eles, _, _ = lf.load(str(p))
    assert [e.l for e in eles] == [1, 2, 4, 3, 1, 3, 2, 1, 3, 2]


def test_reference_tutorial_rings(bmx, lf):
    """Two of the reference's ring-design tutorial lattices, kept as fixtures
    (bmad-doc/tutorial_ring_design/lattices): 01_FODO's arc cell (pi/66 of bend per cell,
    both files of the call) and 13_LatticeTypes/ring.bmad (rf frequency from its harmonic
    number); the closed ring closes and compiles for the GPU path."""
    E = bmx.elements
    eles, mc2, charge = lf.load(os.path.join(TUTORIAL, "fodoAR0.bmad"))
    arr = E.to_array(eles)
    assert len(arr) == 8
    assert abs(float((arr["g"] * arr["l"]).sum()) - 2 * math.pi / 132) < 1e-15
    eles, mc2, charge = lf.load(os.path.join(TUTORIAL, "ring.bmad"))
    arr = E.to_array(eles)
    assert abs(float((arr["g"] * arr["l"]).sum()) - 2 * math.pi) < 1e-9
    rf = arr[arr["key"] == E.RFCAVITY]
    assert len(rf) >= 1 and np.all(rf["rf_frequency"] > 1e8)
    rc, msg = _validate(bmx, arr, mc2, charge)
    assert rc == 0, msg


def test_overlays_groups_and_branch_start(bmx, lf, tmp_path):
    """Overlays set each attribute they control to the sum of their expressions (wildcard
    slaves, variables set in the definition or later, unset variables 0); a switched-on group
    changes nothing at load; start_branch_at rotates the line; beginning[e_tot] gives the
    reference energy; multipole attributes (a1, b2, b0_elec) land in the element record."""
    E = bmx.elements
    p = tmp_path / "lat.bmad"
    p.write_text(
        "beginning[e_tot] = 2e9\n"
        "sf_1: sextupole, l = 0.2, k2 = 5\nsf_2: sextupole, l = 0.2, k2 = 6\nsd_1: sextupole, l = 0.2, k2 = -4\n"
        "q: quadrupole, l = 0.5, k1 = 0.3, a1 = 1e-3, b2 = 0.5, b0_elec = 2e5, r0_elec = 0.1\n"
        "b: sbend, l = 1, g = 0.01, dg = 1e-4, ref_tilt = 0.1\n"
        "bz: sbend, l = 0.5, k1 = 0.1, exact_multipoles = vertically_pure\n"
        "d: drift, l = 1\nip: marker\n"
        "osf: overlay = {sf_*[k2]: 2 + x}, var = {x}, x = 0.5\n"
        "o2: overlay = {sf_1[k2]: 3*y, q[k1]: y}, var = {y}\n"
        "o2[y] = 0.1\n"
        "gr: group = {sd_1[k2]: 10*z}, var = {z}, z = 3\n"
        "r: line = (d, sf_1, q, ip, sf_2, b, sd_1, bz)\nuse, r\nstart_branch_at ip\n")
    eles, mc2, charge = lf.load(str(p))
    names = [e.key for e in eles]
    assert names == [E.MARKER, E.SEXTUPOLE, E.SBEND, E.SEXTUPOLE, E.SBEND, E.DRIFT, E.SEXTUPOLE,
                     E.QUADRUPOLE]
    sext = [e for e in eles if e.key == E.SEXTUPOLE]
    # sf_2 (after ip): 2 + 0.5; sd_1: group leaves it; sf_1: (2 + 0.5) + 3 * 0.1
    assert sext[0].k2 == pytest.approx(2.5) and sext[1].k2 == -4
    assert sext[2].k2 == pytest.approx(2.8)
    q = [e for e in eles if e.key == E.QUADRUPOLE][0]
    assert q.k1 == pytest.approx(0.1) and q.a_pole[1] == 1e-3 and q.b_pole[2] == 0.5
    assert q.b_pole_elec[0] == 2e5 and q.r0_elec == 0.1 and q.n_multipole_elec == E.N_POLE
    b, bz = [e for e in eles if e.key == E.SBEND]
    assert b.dg == 1e-4 and b.ref_tilt == 0.1 and bz.exact_multipoles == E.EXACT_VERTICALLY_PURE
    assert eles[0].e_tot == 2e9
    rc, msg = _validate(bmx, E.to_array(eles), mc2, charge)
    assert rc == 0, msg


def test_thin_multipoles_and_class_settings(bmx, lf, tmp_path):
    """multipole (KnL, KnSL, Tn, tilt) and ab_multipole (an, bn) elements land in the element
    record where ele_struct keeps them (knl -> a_pole, tn -> b_pole, knsl -> a_pole_elec);
    'type[attr] = v' sets every element of that type, 'type::pattern[attr] = f(%[attr])' the
    matching ones from their own value; an element name shadows a type name it abbreviates."""
    E = bmx.elements
    p = tmp_path / "lat.bmad"
    p.write_text(
        "parameter[e_tot] = 3e9\nparameter[ptc_exact_model] = .false.\n"
        "m1: multipole, k1l = 0.2, t1 = 0.1, k3l = 40, k2sl = 5, tilt = 0.05\n"
        "ab: ab_multipole, a2 = 1.5, b4 = -3, tilt = 0.3\n"
        "sx1: sextupole, l = 0.2, k2 = 10\nsx2: sextupole, l = 0.2, k2 = -8\n"
        "quad1: quadrupole, l = 0.5, k1 = 0.4\n"
        "rfc: rfcavity, l = 0.5, voltage = 1e6, rf_frequency = 5e8\nrfx: rfc\n"
        "sextupole[num_steps] = 4\nsext::sx1[k2] = 0.5 * %[k2]\nquad::*[ds_step] = 0.05\n"
        "r: line = (m1, ab, sx1, sx2, quad1, rfx)\nuse, r\n")
    eles, mc2, charge = lf.load(str(p))
    m1, ab, sx1, sx2, q, rfx = eles
    assert m1.key == E.MULTIPOLE and ab.key == E.AB_MULTIPOLE
    assert (m1.a_pole[1], m1.b_pole[1], m1.a_pole[3], m1.a_pole_elec[2], m1.tilt) == (0.2, 0.1, 40, 5, 0.05)
    assert (ab.a_pole[2], ab.b_pole[4], ab.tilt) == (1.5, -3, 0.3)
    assert (sx1.k2, sx2.k2) == (5.0, -8.0)
    assert sx1.ds_step == pytest.approx(0.05) and sx2.ds_step == pytest.approx(0.05)
    assert q.ds_step == pytest.approx(0.05)
    assert rfx.key == E.RFCAVITY and rfx.voltage == 1e6 and rfx.l == 0.5
    rc, msg = _validate(bmx, E.to_array(eles), mc2, charge)
    assert rc == 0, msg
    p.write_text("parameter[e_tot] = 3e9\nq: quadrupole, l = 0.5, k1l = 0.2\nr: line = (q)\nuse, r\n")
    with pytest.raises(lf.LatticeFileError):
        lf.load(str(p))


def test_closed_orbit_test_lattices(bmx, lf):
    """The SLS-2 lattices of the reference's closed-orbit regression test (closed_orbit_test/
    ad05f_{pos,neg}_RF.lat, kept as fixtures): 3120 elements each, 96 thin octupole
    multipoles, 12 rf cavities at harmonic 96 of the 288 m ring (99.93 MHz), the pos_RF file's
    sextupoles halved by 'sextupole::*[k2] = 0.5 * %[k2]', bends at ds_step 0.005; both compile
    for the GPU path."""
    E = bmx.elements
    d = os.path.join(HERE, "golden", "lattices", "closed_orbit")
    pos, _, _ = lf.load(os.path.join(d, "ad05f_pos_RF.lat"))
    neg, mc2, charge = lf.load(os.path.join(d, "ad05f_neg_RF.lat"))
    for eles in (pos, neg):
        arr = E.to_array(eles)
        assert len(arr) == 3120
        assert (arr["key"] == E.MULTIPOLE).sum() == 96
        rf = arr[arr["key"] == E.RFCAVITY]
        assert len(rf) == 12
        L = float(arr["l"].sum())
        assert L == pytest.approx(288.00264, rel=1e-12)
        beta0 = float(arr[0]["p0c"] / arr[0]["e_tot"])
        assert np.allclose(rf["rf_frequency"], 96 * beta0 * E.C_LIGHT / L, rtol=1e-14)
        assert np.all(arr["ds_step"][arr["key"] == E.SBEND] == 0.005)
        rc, msg = _validate(bmx, arr, mc2, charge)
        assert rc == 0, msg
    kp = np.array([e.k2 for e in pos if e.key == E.SEXTUPOLE])
    kn = np.array([e.k2 for e in neg if e.key == E.SEXTUPOLE])
    assert np.array_equal(kp, 0.5 * kn) and np.all(kn != 0)
    vp = [e.voltage for e in pos if e.key == E.RFCAVITY]
    vn = [e.voltage for e in neg if e.key == E.RFCAVITY]
    assert vp == [-v for v in vn] and vp[0] > 0
