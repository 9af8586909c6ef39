"""Thin multipoles on the GPU (multipole$ and ab_multipole$: track1_bmad.f90:142-168 with
multipole_kicks, multipole_mod.f90:32-51, 203-286, and multipole_ele_to_kt.f90), against the
CPU restatement (needs an MI355X): single elements of every kind of strength (orders up to 21,
order tilts, skew strengths, the dipole order's reference-orbit form, element tilt and
offsets), a ring with thin multipoles in every cell on the generic and the cell kernels, and
the SLS-2 lattice of the reference's closed-orbit regression test loaded from its file."""
import importlib
import os

import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", params=["fast", "reference"])
def ctx(bmx, request):
    c = bmx.tracking.Context(0)
    c.set_maps(request.param)
    return c


def _variants(bmx):
    """name -> element at the 5 GeV reference; strengths sized so that the kicks reach 1e-5..1e-3
    or more over the test beam's amplitudes (1 mm rms) without losing it."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    mk = E.make_element
    return {
        # the octupoles of the SLS-2 lattice
        "octupole": mk(E.MULTIPOLE, p0c, et, knl={3: 3364.6247}),
        # orders 0..5 with order tilts; the dipole order bends the reference orbit
        "mixed_orders": mk(E.MULTIPOLE, p0c, et, knl={0: 2e-4, 1: 0.3, 2: 20.0, 3: -500.0, 5: 1e6},
                           tn={0: 0.2, 1: 0.1, 3: -0.2}),
        # skew strengths beside normal ones (the combined strength and tilt)
        "skew": mk(E.MULTIPOLE, p0c, et, knl={1: 0.2, 2: 10.0}, ksl={1: 0.05, 2: -4.0, 3: 300.0}),
        # element tilt, offsets and pitches (the frame change without the tilt)
        "tilted_offset": mk(E.MULTIPOLE, p0c, et, knl={2: 15.0, 4: 3e4}, tilt=0.3,
                            x_offset=3e-4, y_offset=-2e-4, x_pitch=1e-4, y_pitch=-5e-5),
        # orders 8, 13 and 21 (the reference's rounded factorials from 19! on), strong at the
        # beam's 3-4 sigma amplitudes
        "high_orders": mk(E.MULTIPOLE, p0c, et, knl={8: 4e22, 13: 6e40, 21: 4.9e68}, tn={13: 0.05}),
        # an ab_multipole rotated by its tilt, with a dipole kick of the particle
        "ab_tilted": mk(E.AB_MULTIPOLE, p0c, et, a_pole={1: 0.01, 3: 50.0},
                        b_pole={0: 1e-5, 2: 8.0}, tilt=0.7, x_offset=-1e-4),
    }


NAMES = ["octupole", "mixed_orders", "skew", "tilted_offset", "high_orders", "ab_tilted"]


@pytest.mark.parametrize("name", NAMES)
def test_thin_multipole_single_element(bmx, oracle, ctx, name):
    """One element, 4096 particles at 10x the base beam sizes and 1e-2 momentum spread: losses
    identical and survivors within 1e-12 of the restatement (the per-element bar)."""
    ele = _variants(bmx)[name]
    arr = bmx.elements.to_array([ele])
    beam = P.random_beam(bmx, 4096, seed=31, scale=10.0, pz_sigma=1e-2)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    assert np.abs(ref["px"] - beam["px"]).max() > 1e-6  # the element kicks
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    print(f"thin multipole {name} maps={ctx.maps}", flush=True)
    bmx.tracking.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    assert alive.sum() > 3000
    err = P.rel_err(out, ref, alive)
    assert err <= 1e-12, err


def _ring_with_thin_multipoles(bmx, n_cells):
    """ring1 of n_cells with a thin multipole (octupole, skew sextupole, a small dipole order)
    after every sextupole and an ab_multipole (skew quadrupole) after every rf cavity."""
    E, L = bmx.elements, bmx.lattices
    p0c, et = L.reference()
    out = []
    for j, e in enumerate(L.ring1(n_cells)):
        out.append(e)
        if e.key == E.SEXTUPOLE:
            out.append(E.make_element(E.MULTIPOLE, p0c, et, knl={0: 1e-6 * (1 + j % 3), 3: 150.0},
                                      ksl={2: 0.5}, tn={3: 0.1}))
        elif e.key == E.RFCAVITY:
            out.append(E.make_element(E.AB_MULTIPOLE, p0c, et, a_pole={1: 2e-3}, tilt=0.05))
    return E.to_array(out)


@pytest.mark.parametrize("maps,specialize", [("reference", False), ("fast", False), ("fast", True)])
def test_ring_with_thin_multipoles(bmx, oracle, maps, specialize):
    """A 10-cell ring with thin multipoles in every cell, 30 turns, on the generic kernel and
    on the cell kernel generated with them: losses identical to the restatement, survivors
    within max(1e-10, 100x) (reference maps) or max(1e-9, 100x) (fast maps) of the
    restatement's own sensitivity envelope."""
    T = bmx.tracking
    arr = _ring_with_thin_multipoles(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=14, scale=15.0, pz_sigma=3e-3)
    ref, env = P.cached(("thin_multipole_ring", 30), lambda: P.envelope(oracle, arr, beam, 30))
    assert (ref["state"] != 1).sum() > 20
    print(f"thin-multipole ring maps={maps} specialize={specialize}", flush=True)
    out = T.track_beam_host(arr, beam, 30, maps=maps, specialize=specialize)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env, floor=1e-10 if maps == "reference" else 1e-9)
    alive = ref["state"] == 1
    print("deviation max", float(dev[alive].max()), "median", float(np.median(dev[alive])), flush=True)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])


def test_thin_multipoles_on_the_plain_path(bmx, oracle):
    """The element-at-a-time kernel carries the thin multipole too: one turn of the ring with
    thin multipoles equals the restatement to the per-element bar."""
    T = bmx.tracking
    arr = _ring_with_thin_multipoles(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=15, scale=5.0)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    ctx = T.Context(0)
    lat = T.Lattice(ctx, arr)
    b = T.Beam.from_host(ctx, beam)
    T.track_beam_plain(ctx, lat, b, 1, 0)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    assert P.rel_err(out, ref, alive) <= 1e-11


@pytest.mark.parametrize("maps", ["reference", "fast"])
def test_closed_orbit_test_lattice_against_oracle(bmx, oracle, maps):
    """The SLS-2 lattice of the reference's closed-orbit regression test (closed_orbit_test/
    ad05f_pos_RF.lat, kept as a fixture): 3120 elements with 96 thin octupoles, 12 rf
    cavities, 2.4 GeV electrons, loaded by the lattice-file loader; 20 turns of 2000 particles
    at 2.5 mm horizontal size (about a third are lost) with both map forms (its period is 1040
    elements, too long for a cell kernel, so it runs on the generic kernel): losses bit-exact and
    survivors within max(1e-10, 100x) (reference maps) or max(1e-9, 100x) (fast maps) of the
    restatement's own sensitivity envelope."""
    lf = importlib.import_module(bmx.__name__ + ".lattice_file")
    eles, mc2, charge = lf.load(os.path.join(HERE, "golden", "lattices", "closed_orbit",
                                             "ad05f_pos_RF.lat"))
    E, T, L = bmx.elements, bmx.tracking, bmx.lattices
    arr = E.to_array(eles)
    assert len(arr) == 3120 and (arr["key"] == E.MULTIPOLE).sum() == 96
    p0c, et = float(arr[0]["p0c"]), float(arr[0]["e_tot"])
    rng = np.random.default_rng(19)
    n = 2000
    sx = 2.5e-3
    beam = {k: s * rng.standard_normal(n) for k, s in
            zip(P.COORDS, (sx, sx / 10, sx / 2, sx / 20, 3e-3, 1e-3))}
    beam["beta"] = L.init_beta(beam["pz"], p0c, et, mc2)
    beam["state"] = np.ones(n, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.zeros(n, dtype=np.int32)
    ref, env = P.cached(("ad05f_pos_rf", 20), lambda: P.envelope(oracle, arr, beam, 20))
    assert 100 < (ref["state"] != 1).sum() < n - 500
    print(f"SLS-2 ad05f ring maps={maps}", flush=True)
    out = T.track_beam_host(arr, beam, 20, mc2=mc2, charge=charge, maps=maps)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env, floor=1e-10 if maps == "reference" else 1e-9)
    alive = ref["state"] == 1
    print("deviation max", float(dev[alive].max()), "median", float(np.median(dev[alive])), flush=True)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])
