"""Sector bends beyond the common path on the GPU (the general bend map, bmad/low_level/
track_a_bend.f90:18-333): offsets, pitches and ref_tilt (the bend branch of offset_particle,
bmad/code/offset_particle.f90:206-232,404-430), k1 (sbend_body_with_k1_map), multipoles and
kicks with their b0 part as the field error, electric multipoles, and the full edge
(exact_bend_edge_kick with the hard multipole edge, fringe_mod.f90:692-865,1474-1933), against
the CPU restatement (needs an MI355X)."""
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
    """name -> element at the 5 GeV reference."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    mk = E.make_element
    off = dict(x_offset=4e-4, y_offset=-3e-4, z_offset=2e-4, x_pitch=2e-4, y_pitch=-1e-4)
    return {
        # offsets, pitches and ref_tilt of a curved bend (the bend frame change)
        "offset_curved": mk(E.SBEND, p0c, et, l=1.5, g=0.02, ref_tilt=0.2, **off),
        # ref_tilt alone (a rotation about the entrance axis)
        "tilt_only": mk(E.SBEND, p0c, et, l=1.0, g=0.03, ref_tilt=0.3),
        # offsets of a bend with g = 0 (the straight-element frame change) and a field error
        "offset_straight": mk(E.SBEND, p0c, et, l=0.8, g=0.0, dg=2e-3, **off),
        # a gradient: the k1 body map in steps
        "k1_body": mk(E.SBEND, p0c, et, l=1.2, g=0.015, k1=0.3, num_steps=4),
        # the k1 map of a straight bend, defocusing
        "k1_straight": mk(E.SBEND, p0c, et, l=0.6, g=0.0, k1=-0.4, num_steps=3),
        # multipoles, k2 and kicks (the b0 part becomes dg), scaled by r0_mag
        "multipoles": mk(E.SBEND, p0c, et, l=1.0, g=0.01, k2=2.0, a_pole={1: 0.05, 3: 4.0},
                         b_pole={0: 0.01, 2: 1.0}, vkick=2e-5, hkick=-1e-5, r0_mag=0.5,
                         num_steps=5),
        # electric multipoles (energy-changing kicks and edges)
        "electric": mk(E.SBEND, p0c, et, l=0.8, g=0.01, b_pole_elec={0: 2e6, 2: 1e9},
                       a_pole_elec={1: 5e8}, num_steps=4),
        # the full edge with k1 (hard multipole edge) and fint, hgap
        "exact_edge": mk(E.SBEND, p0c, et, l=1.0, g=0.02, e1=0.1, e2=0.15, k1=0.2,
                         fringe_type=E.FRINGE_FULL, fint=0.5, hgap=0.03, num_steps=2),
        # the full edge at the exit only, with offsets and ref_tilt
        "exact_exit_offset": mk(E.SBEND, p0c, et, l=1.0, g=0.02, e1=0.05, e2=-0.1,
                                fringe_type=E.FRINGE_FULL, fringe_at=E.EXIT_END, fint=0.4,
                                hgap=0.02, ref_tilt=0.1, **off),
        # exact_multipoles with g = 0: k1 stays among the multipoles (kicks, no k1 map)
        "exact_multipoles_straight": mk(E.SBEND, p0c, et, l=0.6, g=0.0, k1=0.3, k2=3.0,
                                        exact_multipoles=E.EXACT_VERTICALLY_PURE, num_steps=6),
    }


NAMES = ["offset_curved", "tilt_only", "offset_straight", "k1_body", "k1_straight", "multipoles",
         "electric", "exact_edge", "exact_exit_offset", "exact_multipoles_straight"]


@pytest.mark.parametrize("name", NAMES)
def test_general_bend_single_element(bmx, oracle, ctx, name):
    """One element, 4096 particles at 10x the base beam sizes and 1e-2 momentum spread: losses
    identical and survivors within 1e-12 of the restatement (the per-element bar)."""
    ele = _variants(bmx)[name]
    arr = bmx.elements.to_array([ele])
    beam = P.random_beam(bmx, 4096, seed=23, scale=10.0, pz_sigma=1e-2)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    print(f"general bend {name} maps={ctx.maps}", flush=True)
    bmx.tracking.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["state"][bad[:10]], ref["state"][bad[:10]])
    alive = ref["state"] == 1
    assert alive.sum() > 3000
    err = P.rel_err(out, ref, alive)
    assert err <= 1e-12, err
    rel = np.abs(out["beta"][alive] - ref["beta"][alive]) / ref["beta"][alive]
    assert rel.max() <= 1e-14, rel.max()


def _ring_with_general_bends(bmx, n_cells, seed=4):
    """ring1 of n_cells whose bends carry offsets, pitches and ref_tilt at random (1e-4 m,
    1e-4 rad, 1e-3 rad), a small gradient and a skew quadrupole multipole."""
    E, L = bmx.elements, bmx.lattices
    rng = np.random.default_rng(seed)
    arr = E.to_array(L.ring1(n_cells))
    sel = np.nonzero(arr["key"] == E.SBEND)[0]
    k = sel.size
    arr["x_offset"][sel] = 1e-4 * rng.standard_normal(k)
    arr["y_offset"][sel] = 1e-4 * rng.standard_normal(k)
    arr["x_pitch"][sel] = 1e-4 * rng.standard_normal(k)
    arr["y_pitch"][sel] = 1e-4 * rng.standard_normal(k)
    arr["ref_tilt"][sel] = 1e-3 * rng.standard_normal(k)
    arr["k1"][sel] = 1e-3
    for i in sel:
        arr["n_multipole"][i] = E.N_POLE
        arr["a_pole"][i][1] = 2e-4
    return arr


@pytest.mark.parametrize("maps,specialize", [("reference", False), ("fast", False), ("fast", True)])
def test_ring_with_general_bends(bmx, oracle, maps, specialize):
    """A 10-cell ring whose bends are all general bends, 30 turns, on the generic kernel built
    with the general bend and on the cell kernel generated with it: losses identical to the
    restatement, survivors within max(1e-10, 100x) (reference maps) or max(1e-9, 100x) (fast
    maps) of the restatement's own sensitivity envelope."""
    T = bmx.tracking
    arr = _ring_with_general_bends(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=12, scale=15.0, pz_sigma=3e-3)
    ref, env = P.cached(("general_bend_ring", 30), lambda: P.envelope(oracle, arr, beam, 30))
    assert (ref["state"] != 1).sum() > 20
    print(f"general-bend ring maps={maps} specialize={specialize}", flush=True)
    out = T.track_beam_host(arr, beam, 30, maps=maps, specialize=specialize)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env, floor=1e-10 if maps == "reference" else 1e-9)
    alive = ref["state"] == 1
    print("deviation max", float(dev[alive].max()), "median", float(np.median(dev[alive])), flush=True)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])


def test_general_bends_on_the_plain_path(bmx, oracle):
    """The element-at-a-time kernel built for such lattices: one turn of the ring with general
    bends equals the restatement to the per-element bar."""
    T = bmx.tracking
    arr = _ring_with_general_bends(bmx, 10)
    beam = P.random_beam(bmx, 2000, seed=13, scale=5.0)
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
