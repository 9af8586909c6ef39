"""The C-ABI library on the host: it loads, exports what include/bmx.h declares, and its table
compiler accepts the benchmark rings and rejects what the GPU path does not implement."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "bmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(bmx_\w+)\s*\(", src, flags=re.M)))


def test_header_symbols_exported(bmx):
    lib = bmx.native.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.bmx_version() == 1


def test_ele_struct_layout_matches_header(bmx, tmp_path):
    """sizeof/offsetof of bmx_ele seen by a C compiler equal the ctypes mirror."""
    E = bmx.elements
    c = tmp_path / "lay.c"
    fields = ["l", "ds_step", "p0c", "a_pole", "b_pole", "pz_aperture_center", "r0_mag",
              "r0_elec", "a_pole_elec", "b_pole_elec", "n_multipole_elec", "exact_multipoles",
              "fq1", "fq2"]
    body = "\n".join(f'printf("%zu\\n", offsetof(bmx_ele, {f}));' for f in fields)
    c.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "{ROOT}/include/bmx.h"\n'
                 f'int main(void){{printf("%zu\\n", sizeof(bmx_ele));{body}return 0;}}\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(E.BmxEle)
    for f, off in zip(fields, vals[1:]):
        assert getattr(E.BmxEle, f).offset == off, f


def _validate(bmx, eles):
    arr = bmx.elements.to_array(eles)
    n = ctypes.c_int32()
    rc = bmx.native.load().bmx_lattice_validate(arr.ctypes.data, len(arr), bmx.elements.M_ELECTRON,
                                                -1, ctypes.byref(n))
    return rc, n.value, bmx.native.load().bmx_last_error().decode()


def test_benchmark_rings_compile(bmx):
    L = bmx.lattices
    for eles in (L.ring0(), L.ring1(), L.ring1_variants(seed=2, n_variants=2)[1]):
        rc, nprm, msg = _validate(bmx, eles)
        assert rc == 0, msg
        assert 0 < nprm * 8 < 160 * 1024  # a whole ring table is far below one CU's LDS


@pytest.mark.parametrize("attrs,code", [
    (dict(key=2, l=1.0, g=0.01, k1=0.1, exact_multipoles=2), 2),  # exact multipoles, curved bend
    (dict(key=2, l=1.0, g=0.01, b_pole={5: 1.0}), 2),     # bend multipole above order 3
    (dict(key=2, l=0.0, g=0.01, fringe_type=4), 2),       # zero-length bend with a full edge
    (dict(key=2, l=1.0, g=0.01, fringe_type=2), 2),       # soft-edge-only bend fringe
    (dict(key=2, l=1.0, g=0.01, x_offset=1e-3, x_limit=0.01, offset_moves_aperture=1), 2),
    (dict(key=3, l=0.3, k1=0.5, x_offset=1e-3, x_limit=0.01, offset_moves_aperture=1), 2),
    (dict(key=9, l=0.5, voltage=1e6, rf_frequency=0.0), 2),
    (dict(key=23, l=1.0), 2),                             # solenoid: not on this path
    (dict(key=1, l=1.0, aperture_at=5, x_limit=0.01), 2),  # continuous aperture
])
def test_outside_path_rejected(bmx, attrs, code):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    key = attrs.pop("key")
    e = E.make_element(key, p0c, et, **attrs)
    rc, _, msg = _validate(bmx, [e])
    assert rc == code, msg
    assert "element 1" in msg


def test_electric_multipoles_only_on_quadrupoles_and_bends(bmx):
    """Electric multipoles compile on quadrupoles (the stepped map) and bends (the general bend)
    and are rejected on every other element kind and above order 3; a switched-off element
    ignores them."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    for key, attrs in ((E.SEXTUPOLE, dict(l=0.2, k2=1.0)), (E.DRIFT, dict(l=1.0)),
                       (E.RFCAVITY, dict(l=0.5, voltage=1e6, rf_frequency=5e8))):
        rc, _, msg = _validate(bmx, [E.make_element(key, p0c, et, b_pole_elec={0: 1e5}, **attrs)])
        assert rc == 2 and "electric" in msg, (key, msg)
    off = E.make_element(E.SEXTUPOLE, p0c, et, l=0.2, k2=1.0, b_pole_elec={0: 1e5}, is_on=0)
    assert _validate(bmx, [off])[0] == 0
    rc, _, msg = _validate(bmx, [E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5,
                                                a_pole_elec={4: 1e5})])
    assert rc == 2 and "above 3" in msg, msg
    # a stepped quadrupole: 32 values (plus the 24-value offset block when misaligned); edge
    # fields alone also take the stepped map (one step)
    for attrs, n in ((dict(b_pole_elec={0: 1e5}), 32), (dict(a_pole={1: 1e-3}), 32),
                     (dict(vkick=1e-5, x_offset=1e-3), 32 + 24),
                     (dict(fringe_type=E.FRINGE_FULL), 32),
                     (dict(fringe_type=E.FRINGE_SOFT_EDGE_ONLY, fq1=0.1), 32),
                     (dict(fringe_type=E.FRINGE_FULL, fringe_at=E.NO_END), 5)):
        rc, nprm, msg = _validate(bmx, [E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5,
                                                       num_steps=3, **attrs)])
        assert rc == 0 and nprm == n, (attrs, nprm, msg)
    rc, _, msg = _validate(bmx, [E.make_element(E.QUADRUPOLE, p0c, et, l=0.0, k1=0.5,
                                                fringe_type=E.FRINGE_FULL)])
    assert rc == 2 and "zero-length" in msg, msg


def test_general_bends_accepted(bmx):
    """Bends with offsets, ref_tilt, k1, k2, multipoles, electric multipoles or the full edge
    take the general bend's 94 values; a plain bend keeps the 29 of the common path."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    rc, nprm, msg = _validate(bmx, [E.make_element(E.SBEND, p0c, et, l=1.0, g=0.01)])
    assert rc == 0 and nprm == 29, msg
    for attrs in (dict(x_offset=1e-3), dict(ref_tilt=0.1), dict(k1=0.1), dict(k2=1.0),
                  dict(a_pole={1: 1e-3}), dict(b_pole_elec={0: 1e5}), dict(fringe_type=4, e1=0.1),
                  dict(vkick=1e-5), dict(dg=1e-4), dict(k1=0.1, exact_multipoles=3, g=0.0)):
        kw = dict(l=1.0, g=0.01)
        kw.update(attrs)
        rc, nprm, msg = _validate(bmx, [E.make_element(E.SBEND, p0c, et, **kw)])
        assert rc == 0 and nprm == 94, (attrs, nprm, msg)


def test_p0c_mismatch_rejected(bmx):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    a = E.make_element(E.DRIFT, p0c, et, l=1.0)
    b = E.make_element(E.DRIFT, p0c * 1.01, et, l=1.0)
    rc, _, msg = _validate(bmx, [a, b])
    assert rc == 4 and "element 2" in msg


def test_misaligned_non_bend_elements_accepted(bmx):
    """Quadrupoles, thick multipoles and rf cavities with offsets, pitches or tilt carry the
    24-value offset_particle block ahead of their parameters (quadrupole: 5 + 24)."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    for attrs in (dict(x_offset=1e-3), dict(y_pitch=1e-3), dict(tilt=0.1), dict(z_offset=0.01)):
        rc, nprm, msg = _validate(bmx, [E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5, **attrs)])
        assert rc == 0 and nprm == 29, (attrs, msg)
    rc, nprm, msg = _validate(bmx, [E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5)])
    assert rc == 0 and nprm == 5


def test_collimator_momentum_windows_accepted(bmx):
    """Collimators with px/py/z/pz windows carry 8 window values after their 4 aperture
    values."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    for key in (E.ECOLLIMATOR, E.RCOLLIMATOR):
        plain = E.make_element(key, p0c, et, l=0.5, x_limit=0.02, y_limit=0.015)
        e = E.make_element(key, p0c, et, l=0.5, x_limit=0.02, y_limit=0.015, px_aperture_width2=0.01,
                           pz_aperture_width2=0.03, pz_aperture_center=0.02)
        rc0, n0, msg0 = _validate(bmx, [plain])
        rc, nprm, msg = _validate(bmx, [e])
        assert rc0 == 0 and rc == 0 and nprm == n0 + 8, (msg0, msg)


def test_misaligned_drift_accepted(bmx):
    """track1_bmad tracks drifts without offset_particle (track1_bmad.f90:93-94)."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    rc, _, msg = _validate(bmx, [E.make_element(E.DRIFT, p0c, et, l=4.5, x_offset=0.05, tilt=0.1)])
    assert rc == 0, msg


def test_bookkeeping_steps(bmx):
    """ds_step bookkeeping (attribute_bookkeeper.f90:292-398 + check_bend) on the rings."""
    L = bmx.lattices
    eles = L.ring1()
    sext = [e for e in eles if e.key == 5][0]
    rf = [e for e in eles if e.key == 9][0]
    sb = [e for e in eles if e.key == 2][0]
    assert sext.ds_step == pytest.approx(0.2)          # num_steps = 1
    assert rf.ds_step == pytest.approx(0.05)           # num_steps = 10 default
    assert sb.ds_step == pytest.approx(1.0)            # check_bend met = 4 -> 2 steps
    assert sb.fringe_type == 7 and sb.fintx == sb.fint


def test_no_device_calls_fail_cleanly(bmx):
    """Without a GPU, creating a context fails with an error code, not a crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = bmx.native.load().bmx_ctx_create(0, ctypes.byref(h))
    assert rc != 0
    assert bmx.native.load().bmx_last_error()


def test_gaussian_beam_is_seeded(bmx):
    L = bmx.lattices
    a = L.gaussian_beam(1000, seed=4)
    b = L.gaussian_beam(1000, seed=4)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert abs(np.std(a["x"]) - 100e-6) < 10e-6
    assert a["beta"].dtype == np.float64 and np.all(a["beta"] < 1)


def test_thin_multipoles_compile(bmx):
    """multipole / ab_multipole compile to the thin-multipole block, 2 + 2 N values for orders
    0 .. N-1 (a multipole's dipole order in the reference-orbit form takes the first two), plus
    the 24-value offset block when offset or pitched (the tilt alone adds none: it is in the
    kicks); electric multipoles on an ab_multipole are rejected."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    mk = E.make_element
    for ele, n in ((mk(E.MULTIPOLE, p0c, et, knl={0: 1e-3, 1: 0.2, 3: 40}), 2 + 2 * 4),
                   (mk(E.MULTIPOLE, p0c, et, knl={0: 1e-3}), 2),
                   (mk(E.MULTIPOLE, p0c, et, knl={21: 1.0}, tilt=0.2), 2 + 2 * 22),
                   (mk(E.MULTIPOLE, p0c, et, knl={2: 3.0}, ksl={1: 0.1}, x_offset=1e-4), 2 + 6 + 24),
                   (mk(E.AB_MULTIPOLE, p0c, et, a_pole={2: 1.0}, tilt=0.3), 2 + 6),
                   (mk(E.AB_MULTIPOLE, p0c, et, b_pole={0: 1e-5}, y_pitch=1e-4), 2 + 2 + 24)):
        rc, nprm, msg = _validate(bmx, [ele])
        assert rc == 0 and nprm == n, (rc, nprm, n, msg)
    rc, _, msg = _validate(bmx, [mk(E.AB_MULTIPOLE, p0c, et, a_pole={1: 1.0}, b_pole_elec={0: 1e5})])
    assert rc == 2 and "electric" in msg, msg
