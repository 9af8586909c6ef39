"""openPMD HDF5 beam files on the host (bmx_beam_h5_*; hdf5_read_beam.f90, hdf5_write_beam.f90),
pinned by the reference's hdf5_test: its openPMD file bunch_old.h5 holds, as bunch 1, the same 10
positrons as its ASCII input beam_1.dat (the test prints "Old-Bunch1 TTTTTTTTTT"), and its printed
values of particle 7 (tests/golden/hdf5_test_bunch.correct). A written file reads back to within
a few ulp with every state and loss element, and has the record layout of the reference writer."""
import os
import shutil
import subprocess

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MC2 = 0.51099895069e6


def test_reads_reference_openpmd_bunch(bmx):
    """bunch_old.h5, bunch 1: 10 positrons at p0c = 999999869.44 eV (e_tot = 1 GeV), total
    charge 1e-9 C; every coordinate within 1e-15 of the same particles read from beam_1.dat, and
    particle 7's coordinates and beta equal to the digits the reference test prints."""
    T = bmx.tracking
    b, p0c, q, species = T.read_beam_h5(os.path.join(GOLDEN, "bunch_old.h5"), MC2)
    assert species == "positron" and len(b["x"]) == 10
    assert p0c == pytest.approx(999999869.44, rel=1e-10) and q == pytest.approx(1e-9, rel=1e-15)
    a, _ = T.read_beam_file(os.path.join(GOLDEN, "beam_1.dat"), p0c, MC2)
    for k in ("x", "px", "y", "py", "z", "pz", "beta"):
        assert np.allclose(b[k], a[k], rtol=1e-15, atol=0), k
    assert np.all(b["state"] == 1) and np.all(b["loss_ele"] == 0)
    lines = open(os.path.join(GOLDEN, "hdf5_test_bunch.correct")).read().splitlines()
    vec = [float(v) for v in lines[0].split()[3:9]]
    other = [float(v) for v in lines[1].split()[3:8]]
    got = [b[k][6] for k in ("x", "px", "y", "py", "z", "pz")]
    assert np.allclose(got, vec, rtol=1e-10, atol=0), (got, vec)
    assert b["beta"][6] == pytest.approx(other[4], rel=1e-10)
    assert p0c == pytest.approx(other[3], rel=1e-10)
    assert int(lines[3].split()[2]) == b["state"][6] == 1


def test_photon_bunch_is_refused(bmx):
    """Bunch 2 of bunch_old.h5 holds photons, which this path does not track."""
    T = bmx.tracking
    with pytest.raises(bmx.native.BmxError) as ei:
        T.read_beam_h5(os.path.join(GOLDEN, "bunch_old.h5"), MC2, ix_bunch=2)
    assert ei.value.code == 2 and "photon" in str(ei.value)


def _tracked_looking_beam(bmx):
    L = bmx.lattices
    beam = L.gaussian_beam(313, seed=9, sigma=3 * L.SIGMA0)
    beam["state"][::9] = 4
    beam["state"][5::11] = 8
    lost = beam["state"] != 1
    beam["loss_ele"][lost] = np.arange(lost.sum()) % 900 + 1
    return beam


def test_write_read_round_trip(bmx, tmp_path):
    """A beam with lost particles written and read back: x and y exactly, px, py and pz within 4
    ulp (they pass through px p0c, py p0c and pz p0c), z through -z / (beta c), beta recomputed
    from pz; states and loss elements exact, species, p0c and charge kept; a second bunch
    appended reads back too."""
    T, L = bmx.tracking, bmx.lattices
    p0c, _ = L.reference()
    beam = _tracked_looking_beam(bmx)
    path = str(tmp_path / "beam.h5")
    T.write_beam_h5(path, beam, p0c, species="Electron", charge=2e-12, ix_ele=901)
    back, p0c2, q, species = T.read_beam_h5(path, MC2)
    assert species == "electron" and p0c2 == p0c and q == pytest.approx(313 * 2e-12, rel=1e-15)
    for k in ("x", "px", "y", "py", "z", "pz", "beta"):
        assert np.allclose(back[k], beam[k], rtol=4 * np.finfo(float).eps, atol=0), k
    assert np.array_equal(back["x"], beam["x"]) and np.array_equal(back["y"], beam["y"])
    assert np.array_equal(back["state"], beam["state"])
    lost = beam["state"] != 1
    assert np.array_equal(back["loss_ele"][lost], beam["loss_ele"][lost])
    assert np.all(back["loss_ele"][~lost] == 0)
    second = L.gaussian_beam(50, seed=10, sigma=L.SIGMA0)
    T.write_beam_h5(path, second, p0c, species="electron", charge=1e-12, ix_ele=3, append=True)
    b2, _, q2, _ = T.read_beam_h5(path, MC2, ix_bunch=2)
    assert len(b2["x"]) == 50 and q2 == pytest.approx(50e-12, rel=1e-15)
    assert np.allclose(b2["py"], second["py"], rtol=4 * np.finfo(float).eps, atol=0)


def test_written_layout_is_the_reference_writers(bmx, tmp_path):
    """The records hdf5_write_beam.f90:125-219 writes for a charged bunch, under
    /data/00001/particles/<species>/, with equal values as constant records (groups): here
    spin, the reference momentum, position z and the branch index."""
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        pytest.skip("h5dump is not available")
    T, L = bmx.tracking, bmx.lattices
    p0c, _ = L.reference()
    path = str(tmp_path / "beam.h5")
    T.write_beam_h5(path, _tracked_looking_beam(bmx), p0c, species="electron", charge=1e-12, ix_ele=5)
    out = subprocess.run([h5dump, "-n", path], capture_output=True, text=True, check=True).stdout
    entries = {}
    for ln in out.splitlines():
        parts = ln.split()
        if len(parts) == 2 and parts[0] in ("group", "dataset"):
            entries[parts[1]] = parts[0]
    base = "/data/00001/particles/electron"
    want = {"momentum": "group", "momentum/x": "dataset", "momentum/y": "dataset",
            "momentum/z": "dataset", "totalMomentum": "dataset", "spin": "group",
            "spin/x": "group", "spin/y": "group", "spin/z": "group",
            "totalMomentumOffset": "group", "time": "dataset", "timeOffset": "dataset",
            "position": "group", "position/x": "dataset", "position/y": "dataset",
            "position/z": "group", "sPosition": "group", "weight": "group",
            "particleStatus": "dataset", "branchIndex": "group", "elementIndex": "dataset",
            "locationInElement": "group"}
    for rel, kind in want.items():
        assert entries.get(f"{base}/{rel}") == kind, (rel, entries.get(f"{base}/{rel}"))
    attrs = subprocess.run([h5dump, "-A", "-g", base, path], capture_output=True, text=True,
                           check=True).stdout
    for a in ("speciesType", "totalCharge", "chargeLive", "chargeUnitSI", "numParticles",
              "unitSI", "unitDimension", "unitSymbol", "minValue", "maxValue", "value", "shape"):
        assert f'ATTRIBUTE "{a}"' in attrs, a
