"""The allocation-free entry point (bmx_track_turns_noalloc) against bmx_track_turns: the same
particles give the same bits, on two calls in a row, the second with a smaller beam (needs an
MI355X)."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


def _run(bmx, ctx, lat, beam, n_turns, turn0, entry):
    T = bmx.tracking
    b = T.Beam.from_host(ctx, beam)
    if entry == "noalloc":
        bmx.native.call("bmx_track_turns_noalloc", ctx.h, lat.h, b.h, int(n_turns), int(turn0))
    else:
        T.track_beam(ctx, lat, b, n_turns, turn0)
    ctx.synchronize()
    out = b.download()
    b.close()
    return out


@pytest.mark.parametrize("special", [False, True], ids=["generic", "specialized"])
def test_noalloc_equals_track_turns(bmx, special):
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    ctx = T.Context(0)
    lat = T.Lattice(ctx, E.to_array(L.ring1(n_cells=10)))
    if special:
        assert lat.specialize()
    big = P.random_beam(bmx, 50000, seed=61, scale=15.0, pz_sigma=3e-3)
    small = P.random_beam(bmx, 3000, seed=62, scale=15.0, pz_sigma=3e-3)
    for beam, n_turns, turn0 in ((big, 23, 4), (small, 17, 0)):
        print(f"noalloc: {beam['x'].size} particles, {n_turns} turns, special={special}", flush=True)
        ref = _run(bmx, ctx, lat, beam, n_turns, turn0, "track_turns")
        got = _run(bmx, ctx, lat, beam, n_turns, turn0, "noalloc")
        assert (ref["state"] != 1).sum() > 10
        for k in P.KEYS:
            assert np.array_equal(got[k], ref[k]), k
    lat.close()
    ctx.close()


def test_noalloc_refuses_beams_beyond_its_scratch(bmx):
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    ctx = T.Context(0)
    lat = T.Lattice(ctx, E.to_array(L.ring1(n_cells=2)))
    b = T.Beam(ctx, (1 << 18) + 1)
    with pytest.raises(bmx.native.BmxError, match="exceed the fixed scratch"):
        bmx.native.call("bmx_track_turns_noalloc", ctx.h, lat.h, b.h, 1, 0)
    b.close()
    lat.close()
    ctx.close()
