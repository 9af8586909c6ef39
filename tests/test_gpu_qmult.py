"""Quadrupoles with multipoles beside k1, with electric multipoles and with edge fields on the
GPU (the stepped quadrupole map: bmad/low_level/track_a_quadrupole.f90:17-148 with n_step > 1,
the electric kicks of multipole_mod.f90:128-143, the electric edge energy kicks of
apply_element_edge_kick.f90:283-326, and the hard multipole and soft quadrupole edges of
fringe_mod.f90:556-865), against the CPU restatement (needs an MI355X).

The tracked species is the electron (charge -1), the opposite sign of the tracking_method_test
known answers (positron), so both signs of the electric strengths' charge factor are run."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["fast", "reference"])
def ctx(bmx, request):
    c = bmx.tracking.Context(0)
    c.set_maps(request.param)
    return c


def _variants(bmx):
    """name -> element, at the 5 GeV reference: fields sized so that the kicks (up to ~1e-3)
    and the net energy changes (entrance and exit edges nearly cancel: ~1e-10 in pz) are well
    above rounding."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    mk = E.make_element
    return {
        # magnetic only: skew dipole and skew quadrupole, a sextupole and a vertical kick
        "magnetic": mk(E.QUADRUPOLE, p0c, et, l=0.5, k1=0.8, a_pole={0: 1e-4, 1: 0.02},
                       b_pole={2: 5.0}, vkick=2e-5, num_steps=4),
        # electric only (dipole and skew quadrupole fields), defocusing k1
        "electric": mk(E.QUADRUPOLE, p0c, et, l=0.4, k1=-0.6, b_pole_elec={0: 2e6},
                       a_pole_elec={1: 1e9}, num_steps=5),
        # both, orders 0..3, electric radius scaling, offsets, pitches and the default tilt
        "mixed_offset": mk(E.QUADRUPOLE, p0c, et, l=0.6, k1=0.5, a_pole={3: 40.0},
                           b_pole={0: -5e-5}, b_pole_elec={0: -1e6, 2: 3e3},
                           a_pole_elec={1: 4e4, 3: 2e2}, r0_elec=0.02, num_steps=6,
                           x_offset=3e-4, y_offset=-2e-4, x_pitch=1e-4, y_pitch=-2e-4,
                           default_tilt=True),
        # k1 = 0 with electric fields: the steps are drifts between the kicks
        "electric_k1_zero": mk(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.0, b_pole_elec={1: 5e8},
                               a_pole_elec={0: 1e6}, num_steps=3),
        # the hard multipole edge alone (fringe_type = full, fq1 = fq2 = 0), a strong quadrupole
        "hard_edge": mk(E.QUADRUPOLE, p0c, et, l=0.4, k1=4.0, fringe_type=E.FRINGE_FULL),
        # the soft quadrupole edge alone
        "soft_edge": mk(E.QUADRUPOLE, p0c, et, l=0.4, k1=-3.0, fq1=0.02, fq2=0.01,
                        fringe_type=E.FRINGE_SOFT_EDGE_ONLY),
        # both edges, multipoles, electric fields and offsets
        "full_edges_mixed": mk(E.QUADRUPOLE, p0c, et, l=0.5, k1=2.5, fq1=-0.01, fq2=0.03,
                               fringe_type=E.FRINGE_FULL, a_pole={1: 0.02}, b_pole_elec={0: 1e6},
                               a_pole_elec={1: 2e8}, num_steps=4, x_offset=2e-4,
                               default_tilt=True),
        # edges at the exit only (the entrance keeps no edge field, electric ones included)
        "exit_edge_only": mk(E.QUADRUPOLE, p0c, et, l=0.5, k1=1.5, fq1=0.02,
                             fringe_type=E.FRINGE_FULL, fringe_at=E.EXIT_END,
                             b_pole_elec={0: 1e6}, num_steps=2),
    }


@pytest.mark.parametrize("name", ["magnetic", "electric", "mixed_offset", "electric_k1_zero",
                                  "hard_edge", "soft_edge", "full_edges_mixed", "exit_edge_only"])
def test_stepped_quadrupole_single_element(bmx, oracle, ctx, name):
    """One element, 4096 particles at 10x the base beam sizes and 1e-2 momentum spread: losses
    identical and survivors within 1e-12 of the restatement (the per-element bar)."""
    ele = _variants(bmx)[name]
    arr = bmx.elements.to_array([ele])
    beam = P.random_beam(bmx, 4096, seed=21, scale=10.0, pz_sigma=1e-2)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    moved = np.abs(ref["pz"] - beam["pz"])
    if ele.n_multipole_elec:
        assert moved.max() > 1e-11, moved.max()  # the electric fields change the energy
    else:
        assert moved.max() == 0
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    print(f"stepped quad {name} maps={ctx.maps}", flush=True)
    bmx.tracking.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    err = P.rel_err(out, ref, alive)
    assert err <= 1e-12, err
    for k in ("beta",):
        rel = np.abs(out[k][alive] - ref[k][alive]) / ref[k][alive]
        assert rel.max() <= 1e-14, (k, rel.max())


def test_electric_edge_loses_the_particle(bmx, oracle, ctx):
    """An electric dipole field so strong that the entrance edge's potential (dE = -charge b0
    x for an electron, apply_energy_kick.f90:43-48) takes particles below their rest energy:
    lost_pz at the element, identical to the restatement, the others tracked on."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    # |dE| = |b0 x| reaches the 5 GeV total energy for x > 5e9 / 5e11 = 1 cm; a 2 mm long
    # field keeps the kicks of the others near 0.1 rad
    ele = E.make_element(E.QUADRUPOLE, p0c, et, l=0.002, k1=0.1, b_pole_elec={0: 5e11},
                         num_steps=2, x_offset=1e-5)
    arr = E.to_array([ele])
    beam = P.random_beam(bmx, 2048, seed=5, scale=100.0)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    lost = ref["state"] != 1
    assert 0 < lost.sum() < lost.size, lost.sum()
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    bmx.tracking.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["state"][bad[:10]], ref["state"][bad[:10]])
    alive = ref["state"] == 1
    assert P.rel_err(out, ref, alive) <= 1e-12


def _ring_with_stepped_quads(bmx, n_cells):
    """ring1 of n_cells whose quadrupoles all carry a skew quadrupole term and electric dipole
    and sextupole fields (so every cell holds stepped quadrupoles)."""
    E, L = bmx.elements, bmx.lattices
    arr = E.to_array(L.ring1(n_cells))
    q = np.nonzero(arr["key"] == E.QUADRUPOLE)[0]
    for j, i in enumerate(q):
        arr["n_multipole"][i] = E.N_POLE
        arr["a_pole"][i][1] = 1e-3 * (1 + 0.1 * (j % 3))
        arr["n_multipole_elec"][i] = E.N_POLE
        arr["b_pole_elec"][i][0] = 5e4
        arr["a_pole_elec"][i][2] = 1e10
        arr["ds_step"][i] = arr["l"][i] / 4
    return arr


@pytest.mark.parametrize("maps,specialize", [("reference", False), ("fast", False), ("fast", True)])
def test_ring_with_stepped_quadrupoles(bmx, oracle, maps, specialize):
    """A 10-cell ring with stepped quadrupoles in every cell, 30 turns, on the generic kernel
    (the span kernel, which carries the stepped map) and on the cell kernel generated with it:
    losses identical to the restatement. Survivors: the reference-form maps within
    max(1e-10, 100x) of the restatement's own sensitivity envelope; the fast maps within
    max(1e-9, 100x), the bar of the 1000-turn tests. This ring loses over half of its beam in 30
    turns, so its survivors sit near the dynamic aperture and amplify per-element differences
    strongly: the fast sector bend's z differs from the restatement by about 4 units in the last
    place of the bend length per pass (1.6e-15 m against 4.4e-16 m for the reference form, which
    rounds in the restatement's order), and over 30 turns that grows to ~1e-10 on the most
    sensitive orbits (up to ~140x their envelope)."""
    T = bmx.tracking
    arr = _ring_with_stepped_quads(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=9, scale=15.0, pz_sigma=3e-3)
    ref, env = P.cached(("stepped_ring", 30), lambda: P.envelope(oracle, arr, beam, 30))
    assert (ref["state"] != 1).sum() > 20
    print(f"stepped-quad ring maps={maps} specialize={specialize}", flush=True)
    out = T.track_beam_host(arr, beam, 30, maps=maps, specialize=specialize)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env, floor=1e-10 if maps == "reference" else 1e-9)
    alive = ref["state"] == 1
    print("deviation max", float(dev[alive].max()), "median", float(np.median(dev[alive])), flush=True)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])


def test_stepped_quadrupole_on_the_plain_path(bmx, oracle):
    """The element-at-a-time kernel (the plain path) carries the stepped map too: one turn of
    the 10-cell ring with stepped quadrupoles equals the restatement to the per-element bar."""
    T = bmx.tracking
    arr = _ring_with_stepped_quads(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=10, scale=5.0)
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
