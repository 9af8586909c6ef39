"""Copy the reference's known-answer lines for this path into small fixture files.

Run in a checkout that has the bmad-ecosystem tree (default /root/reference); the fixtures are
committed, so tests never read the reference at run time. Each fixture line is the verbatim
line of the reference's ``output.correct`` file ("name" TYPE tol values...), i.e. data.

Sources:
  regression_tests/tracking_method_test/output.correct  -- "<ELE>: Bmad_Standard" lines
  regression_tests/fringe_test/output.correct           -- Up-Orb / Dn-Orb (Hwang edge kick)
  regression_tests/aperture_test/output.correct         -- EC_*/RC_* state and unstable lines
  regression_tests/multipole_test/output.correct        -- the thin multipole MM's lines
  regression_tests/closed_orbit_test/ad05f_{pos,neg}_RF.lat -- lattices (lattices/closed_orbit)
  regression_tests/mode3_test/t6.dat and output.correct   -- a one-turn matrix and make_N lines
  regression_tests/hdf5_test/bunch_old.h5 and output.correct -- an openPMD beam file and the
                                                               particle values the test prints
  regression_tests/radiation_test/output.correct        -- the radiation maps (rm0, rm1) of
                                                           a bend and a quadrupole
"""
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

TRACKING_ELEMENTS = ["DRIFT1", "ECOLLIMATOR1", "HKICKER1", "INSTRUMENT1", "KICKER1", "MONITOR1",
                     "OCTUPOLE1", "RCOLLIMATOR1", "RFCAVITY1", "RFCAVITY3", "SEXTUPOLE1",
                     "THICK_MULTIPOLE1", "VKICKER1", "QUADRUPOLE1", "QUADRUPOLE2", "QUADRUPOLE3",
                     "QUADRUPOLE4", "QUADRUPOLE5", "SBEND1", "SBEND3", "SBEND5", "SBEND7"]


def grab(rel, keep):
    path = os.path.join(REF, rel)
    with open(path) as f:
        return [ln.rstrip("\n") for ln in f if keep(ln)]


def main():
    lines = grab("regression_tests/tracking_method_test/output.correct",
                 lambda ln: any(ln.startswith(f'"{e}: Bmad_Standard"') for e in TRACKING_ELEMENTS))
    with open(os.path.join(HERE, "tracking_method_test.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = grab("regression_tests/fringe_test/output.correct",
                 lambda ln: ln.startswith('"Up-Orb"') or ln.startswith('"Dn-Orb"'))
    with open(os.path.join(HERE, "fringe_test.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = grab("regression_tests/aperture_test/output.correct",
                 lambda ln: ln.startswith('"EC_') or ln.startswith('"RC_'))
    with open(os.path.join(HERE, "aperture_test.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = grab("regression_tests/multipole_test/output.correct", lambda ln: '-MM"' in ln)
    with open(os.path.join(HERE, "multipole_test_mm.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(REF, "regression_tests/hdf5_test/bunch_old.h5"), "rb") as f:
        data = f.read()
    with open(os.path.join(HERE, "bunch_old.h5"), "wb") as f:
        f.write(data)
    lines = grab("regression_tests/hdf5_test/output.correct", lambda ln: ln.startswith('"Bunch1-'))
    with open(os.path.join(HERE, "hdf5_test_bunch.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    os.makedirs(os.path.join(HERE, "mode3"), exist_ok=True)
    lines = grab("regression_tests/mode3_test/t6.dat", lambda ln: True)
    with open(os.path.join(HERE, "mode3", "t6.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = grab("regression_tests/mode3_test/output.correct", lambda ln: ln.startswith('"make_N'))
    with open(os.path.join(HERE, "mode3", "make_n.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = grab("regression_tests/radiation_test/output.correct",
                 lambda ln: ln.startswith('"B1-rm') or ln.startswith('"K1-rm'))
    with open(os.path.join(HERE, "radiation_test_maps.correct"), "w") as f:
        f.write("\n".join(lines) + "\n")
    os.makedirs(os.path.join(HERE, "lattices", "closed_orbit"), exist_ok=True)
    for name in ("ad05f_pos_RF.lat", "ad05f_neg_RF.lat"):
        lines = grab("regression_tests/closed_orbit_test/" + name, lambda ln: True)
        with open(os.path.join(HERE, "lattices", "closed_orbit", name), "w") as f:
            f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
