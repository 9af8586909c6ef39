"""Element spans on the GPU: track_bunch from the end of ele1 to the end of ele2
(bmad/multiparticle/beam_mod.f90:103-166, wrapping round the ring when ele2 <= ele1) and turns
that start after an element (LTT's track_bunch(lat, bunch, ele_start, ele_start),
bsim/modules/lt_tracking_mod.f90:3223), against the CPU restatement run on the same elements in
the same order (needs an MI355X)."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[("fast", False), ("fast", True), ("reference", False)],
                ids=["fast", "fast-specialized", "reference"])
def setup(bmx, request):
    maps, special = request.param
    c = bmx.tracking.Context(0)
    c.set_maps(maps)
    return c, special


def _ring_both_ends(bmx, n_cells=10):
    """ring1 with every aperture tested at both ends: every entrance test but element 1's
    repeats the exit test just before it, which the element loop skips, so a span that starts
    mid-ring must make that entrance test itself."""
    E = bmx.elements
    eles = bmx.lattices.ring1(n_cells=n_cells)
    for e in eles:
        e.aperture_at = E.BOTH_ENDS
    return E.to_array(eles)


def _order(n, first, count):
    """0-based element indices of a span of `count` elements from element `first`."""
    return [(first + k) % n for k in range(count)]


def _oracle_span(oracle, arr, beam, idx, n_turns, turn0):
    """The restatement through arr[idx] per turn; loss_ele mapped back to the ring's numbering."""
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr[idx], ref, n_turns, turn0=turn0)
    lost = ref["state"] != 1
    pos = ref["loss_ele"][lost] - 1
    ref["loss_ele"][lost] = np.asarray(idx)[pos] + 1
    return ref


def _lattice(bmx, ctx, arr, special):
    lat = bmx.tracking.Lattice(ctx, arr)
    if special:
        assert lat.specialize()
    return lat


def _compare(out, ref, tol):
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["loss_ele"][bad[:10]], ref["loss_ele"][bad[:10]],
                           out["state"][bad[:10]], ref["state"][bad[:10]])
    alive = ref["state"] == 1
    assert P.rel_err(out, ref, alive) <= tol


@pytest.mark.parametrize("e1,e2", [(0, 91), (17, 60), (60, 17), (33, 33), (91, 5), (90, 91)])
def test_track_bunch_span(bmx, oracle, setup, e1, e2):
    """track_bunch(ele1, ele2) on a 91-element ring: elements ele1+1..ele2, or round the end
    when ele2 <= ele1; losses bit-exact and coordinates within 1e-11 of the restatement."""
    ctx, special = setup
    arr = _ring_both_ends(bmx)
    n = len(arr)
    count = e2 - e1 if e1 < e2 else n - e1 + e2
    idx = _order(n, e1 % n, count)
    beam = P.random_beam(bmx, 3000, seed=31 + e1, scale=40.0, pz_sigma=4e-3)
    ref = _oracle_span(oracle, arr, beam, idx, 1, 7)
    lat = _lattice(bmx, ctx, arr, special)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    print(f"track_bunch {e1} -> {e2} ({count} elements), special={special}", flush=True)
    bmx.tracking.track_bunch(ctx, lat, b, e1, e2, turn=7)
    ctx.synchronize()
    out = b.download()
    # one pass of at most 91 elements: the per-element bar with headroom for its rounding
    _compare(out, ref, 1e-11)
    lost = ref["state"] != 1
    # a 40-sigma beam loses a few percent per element it passes
    assert lost.sum() > (100 if count > 1 else 20), "the span should lose particles"
    assert np.all(ref["loss_turn"][lost] == 7)


@pytest.mark.parametrize("start", [0, 1, 38, 90])
def test_turns_from_start_element(bmx, oracle, setup, start):
    """Turns that start after element `start` (LTT's ele_start), 6 turns in chunks of 2 (the
    particles resume mid-ring at every chunk): equal to the restatement on the rotated ring."""
    ctx, special = setup
    arr = _ring_both_ends(bmx)
    n = len(arr)
    idx = _order(n, start % n, n)
    beam = P.random_beam(bmx, 2500, seed=41 + start, scale=30.0, pz_sigma=4e-3)
    ref = _oracle_span(oracle, arr, beam, idx, 6, 3)
    lat = _lattice(bmx, ctx, arr, special)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    print(f"turns from element {start}, special={special}", flush=True)
    try:
        ctx.set_run(2, 0)
        bmx.tracking.track_beam_from(ctx, lat, b, start, 6, 3)
        ctx.synchronize()
    finally:
        ctx.set_run(5, 0)
    out = b.download()
    # 546 element passes: the multi-turn bar
    _compare(out, ref, 1e-10)
    assert (ref["state"] != 1).sum() > 50


def test_span_pieces_equal_whole_turns(bmx, setup):
    """Four spans that cover two turns (0 -> 30, 30 -> 91, 91 -> 50, 50 -> 91) give the same
    losses as two full turns, and coordinates within 1e-12: the spans compose like the
    reference's element loop. (Spans run the span kernel, full turns the full-turn kernel: the
    same maps, which the compiler may contract to FMAs in other places.)"""
    ctx, special = setup
    arr = _ring_both_ends(bmx)
    beam = P.random_beam(bmx, 2000, seed=5, scale=25.0, pz_sigma=4e-3)
    lat = _lattice(bmx, ctx, arr, special)
    T = bmx.tracking
    whole = T.Beam.from_host(ctx, beam)
    T.track_beam(ctx, lat, whole, 1, 0)
    T.track_beam(ctx, lat, whole, 1, 1)
    parts = T.Beam.from_host(ctx, beam)
    T.track_bunch(ctx, lat, parts, 0, 30, turn=0)
    T.track_bunch(ctx, lat, parts, 30, 91, turn=0)
    T.track_bunch(ctx, lat, parts, 91, 50, turn=1)
    T.track_bunch(ctx, lat, parts, 50, 91, turn=1)
    ctx.synchronize()
    a, b = whole.download(), parts.download()
    assert (a["state"] != 1).sum() > 100
    for k in ("state", "loss_ele", "loss_turn"):
        assert np.array_equal(a[k], b[k]), k
    alive = a["state"] == 1
    assert P.rel_err(b, a, alive) <= 1e-12
    assert np.max(np.abs(b["beta"][alive] / a["beta"][alive] - 1)) <= 1e-12


def test_span_arguments_checked(bmx, setup):
    ctx, _ = setup
    arr = _ring_both_ends(bmx, n_cells=2)
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam(ctx, 10)
    for e1, e2 in ((-1, 3), (0, len(arr) + 1)):
        with pytest.raises(bmx.native.BmxError):
            bmx.tracking.track_bunch(ctx, lat, b, e1, e2)
    with pytest.raises(bmx.native.BmxError):
        bmx.tracking.track_beam_from(ctx, lat, b, len(arr) + 1, 1)
