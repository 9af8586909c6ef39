"""The lattice-file loader (lattice_file.py, a subset of bmad/parsing/bmad_parser.f90): a test
ring written in the Bmad language, and the reference's tutorial rings when the reference tree
is present."""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RING = os.path.join(HERE, "golden", "lattices", "ring_small.bmad")
TUTORIAL = "/root/reference/bmad-doc/tutorial_ring_design/lattices"


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


@pytest.mark.skipif(not os.path.isdir(TUTORIAL), reason="reference tree not present")
def test_reference_tutorial_rings(bmx, lf):
    """The reference's ring-design tutorial lattices that stay inside this path's element set
    load, and the closed rings close: 01_FODO's arc cell (pi/66 of bend per cell, both files of
    the call) and 13_LatticeTypes/ring.bmad (rf frequency from its harmonic number)."""
    E = bmx.elements
    eles, mc2, charge = lf.load(os.path.join(TUTORIAL, "01_FODO/a_Arc/fodoAR0.bmad"))
    arr = E.to_array(eles)
    assert len(arr) == 8
    assert abs(float((arr["g"] * arr["l"]).sum()) - 2 * math.pi / 132) < 1e-15
    eles, mc2, charge = lf.load(os.path.join(TUTORIAL, "13_LatticeTypes/ring.bmad"))
    arr = E.to_array(eles)
    assert abs(float((arr["g"] * arr["l"]).sum()) - 2 * math.pi) < 1e-9
    rf = arr[arr["key"] == E.RFCAVITY]
    assert len(rf) >= 1 and np.all(rf["rf_frequency"] > 1e8)
    rc, msg = _validate(bmx, arr, mc2, charge)
    assert rc == 0, msg
