"""Bmad ASCII beam files on the host (bmx_beam_file_*; beam_file_io.f90): the reference's own
beam file fixture, and a bit-identical write/read round trip (a tracking checkpoint)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reads_reference_ascii3_fixture(bmx):
    """regression_tests/hdf5_test/beam_1.dat (the reference's "!ASCII::3" input of its HDF5
    test): 10 positrons, bunch charge 1e-9; particle values and states as the file lists them,
    beta from pz as init_coord sets it."""
    T, L = bmx.tracking, bmx.lattices
    p0c = 1e9
    mc2 = 0.51099895069e6
    beam, q = T.read_beam_file(os.path.join(GOLDEN, "beam_1.dat"), p0c, mc2)
    assert len(beam["x"]) == 10
    assert q == 1e-9
    assert beam["x"][0] == -5.0792635566e-04 and beam["pz"][0] == 1.4862617890e-03
    assert beam["py"][9] == -1.5800206862e-05 and beam["z"][4] == -1.1786695057e-04
    assert np.all(beam["state"] == 1) and np.all(beam["loss_ele"] == 0)
    pc = p0c * (1 + beam["pz"])
    assert np.allclose(beam["beta"], pc / np.sqrt(pc ** 2 + mc2 ** 2), rtol=0, atol=1e-16)


def test_write_read_round_trip_is_bit_identical(bmx, tmp_path):
    """A tracked-looking beam (some particles lost at various elements) written in the column
    layout and read back: every coordinate bit-identical, states and loss elements kept; the
    file has the reference writer's header parameters and column line."""
    T, L = bmx.tracking, bmx.lattices
    p0c, _ = L.reference()
    beam = L.gaussian_beam(257, seed=3, sigma=3 * L.SIGMA0)
    beam["state"][::7] = 4
    beam["loss_ele"][::7] = np.arange(0, 257, 7) % 900 + 1
    path = str(tmp_path / "beam.dat")
    T.write_beam_file(path, beam, p0c, species="Electron", charge=1e-12, ix_ele=901)
    text = open(path).read().splitlines()
    assert text[0].startswith("# ix_bunch") and any(ln.startswith("#!") for ln in text)
    assert any("Lost_Pos_X" in ln for ln in text) and any("Alive" in ln for ln in text)
    back, q = T.read_beam_file(path, p0c)
    for k in ("x", "px", "y", "py", "z", "pz"):
        assert np.array_equal(back[k], beam[k]), k
    assert np.array_equal(back["state"], beam["state"])
    lost = beam["state"] != 1
    assert np.array_equal(back["loss_ele"][lost], beam["loss_ele"][lost])
    assert np.all(back["loss_ele"][~lost] == 0)
    assert abs(q - 257 * 1e-12) < 1e-24
