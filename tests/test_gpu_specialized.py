"""Cell-specialised tile kernels (bmx_lattice_specialize): run-time compiled per repeated
lattice cell, held to the same bars as the generic kernel against the CPU restatement."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(bmx):
    c = bmx.tracking.Context(0)
    c.set_maps("fast")
    return c


def _track(bmx, ctx, arr, beam, n_turns, specialize, turn0=0):
    T = bmx.tracking
    lat = T.Lattice(ctx, arr)
    if specialize:
        assert lat.specialize()
    b = T.Beam.from_host(ctx, beam)
    T.track_beam(ctx, lat, b, n_turns, turn0)
    ctx.synchronize()
    out = b.download()
    b.close()
    lat.close()
    return out


def _within_envelope(out, ref, env, floor=1e-10):
    """Every survivor within max(floor, 100 x its envelope) of the restatement (parity.envelope:
    the response of the restatement itself to FMA contraction and to 1e-14 rescalings of the
    input, so the bar is what a 1e-12 change of the input moves that particle by)."""
    bad, dev = P.outside_envelope(out, ref, env, floor=floor)
    alive = ref["state"] == 1
    ratio = dev[alive] / np.maximum(env[alive], 1e-300)
    print("deviation percentiles 50, 90, 99, max:", np.percentile(dev[alive], [50, 90, 99, 100]),
          "deviation / envelope max:", float(np.max(ratio)), flush=True)
    assert bad.size == 0, (bad[:10], dev[bad[:10]], env[bad[:10]])


def test_ring0_1000_turns_specialized(bmx, oracle, ctx):
    """700-element ring (cell of 7), 1-sigma beam, 1000 turns: within 1e-9 of the CPU
    restatement for every particle, like the generic kernel."""
    L = bmx.lattices
    arr = bmx.elements.to_array(L.ring0())
    beam = L.gaussian_beam(512, seed=7, sigma=L.SIGMA0)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1000)
    print("ring0 1000 turns, specialised", flush=True)
    out = _track(bmx, ctx, arr, beam, 1000, True)
    assert np.all(out["state"] == 1) and np.all(ref["state"] == 1)
    assert P.rel_err(out, ref) <= 1e-9


def test_ring1_no_rf_1000_turns_specialized(bmx, oracle, ctx):
    """The bench's kernel on the 901-element ring without rf, a 1-sigma beam, 1000 turns: the
    bars of the generic kernel's test (parity.check_norf_1000): the FMA-spread percentiles, the
    median within 1e-9 and every particle within max(1e-9, 100x its envelope)."""
    arr, beam, ref, noise, env = P.ring1_norf_1000(bmx, oracle)
    print("ring1 without rf, 1000 turns, specialised", flush=True)
    out = _track(bmx, ctx, arr, beam, 1000, True)
    P.check_norf_1000(out, ref, noise, env)


def test_ring1_chaotic_1000_turns_specialized(bmx, oracle, ctx):
    """The bench's kernel on the full 901-element ring, 1000 turns: losses of the first 10 turns
    identical to the restatement and the surviving set within 3x of the restatement's own
    sensitivity (parity.check_full_1000)."""
    arr, beam, ref, self_flips = P.ring1_full_1000(bmx, oracle)
    print("ring1 full, 1000 turns, specialised", flush=True)
    out = _track(bmx, ctx, arr, beam, 1000, True)
    P.check_full_1000(out, ref, self_flips)


@pytest.mark.parametrize("waves", [3, 4, 5])
def test_ring1_losses_specialized(bmx, oracle, ctx, waves):
    """901-element ring (cell of 9, rf cavity in the generic tail), a beam losing about half of
    its particles in 30 turns: loss state, element and turn identical to the CPU restatement for
    every particle, for each of the three kernel variants."""
    L = bmx.lattices
    arr = bmx.elements.to_array(L.ring1())
    beam = P.random_beam(bmx, 3000, seed=21, scale=25.0, pz_sigma=4e-3)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=30)
    print(f"ring1 losses, specialised, waves {waves}", flush=True)
    try:
        ctx.set_run(5, 0, waves)
        out = _track(bmx, ctx, arr, beam, 30, True)
    finally:
        ctx.set_run(5, 0, 0)
    lost = ref["state"] != 1
    assert lost.sum() > 500
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["loss_ele"][bad[:10]], ref["loss_ele"][bad[:10]])


def test_head_cell_and_tail(bmx, oracle, ctx):
    """A marker in front of the ring shifts the repeated run: a generic head cell, 99 rotated
    cells and a two-element generic tail. After 20 turns: losses bit-exact; every survivor
    within 100x of its envelope (the restatement's own response to FMA contraction and to 1e-14
    rescalings of its input, parity.envelope): these large-amplitude orbits amplify any
    rounding difference."""
    L, E = bmx.lattices, bmx.elements
    eles = L.ring1()
    p0c, et = L.reference()
    eles = [E.make_element(E.MARKER, p0c, et)] + list(eles)
    arr = E.to_array(eles)
    beam = P.random_beam(bmx, 2000, seed=3, scale=15.0, pz_sigma=3e-3)
    ref, env = P.envelope(oracle, arr, beam, 20)
    print("head cell + tail, specialised", flush=True)
    out = _track(bmx, ctx, arr, beam, 20, True)
    assert (ref["state"] != 1).sum() > 100
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    _within_envelope(out, ref, env)


def test_batch_of_specialized_variants(bmx, oracle, ctx):
    """Eight error variants of a 10-cell ring share one specialised kernel (strengths stay in
    the table); the batch equals the CPU restatement of each variant after 25 turns: losses
    bit-exact, every survivor within 100x of its envelope (the restatement's own response to FMA
    contraction and to 1e-14 rescalings of its input) or 1e-9, the multi-turn bar. The envelope
    can miss a direction along which an orbit amplifies errors: the sector bend computes px from
    sums of angles of the order of the bend angle, whose last-bit rounding (~1e-17 absolute)
    contraction does not change, and four random rescalings need not excite it; one orbit of
    these 8000 large-amplitude ones grows such a difference to ~1.4e-10 in 25 turns (the
    reference-form maps, evaluating the same formulas as the restatement, reach 1.6e-10 on it)."""
    L, T = bmx.lattices, bmx.tracking
    arrs = [bmx.elements.to_array(e) for e in L.ring1_variants(5, 8, n_cells=10)]
    counts = [700, 64, 1, 333, 1000, 129, 64, 500]
    beam = P.random_beam(bmx, sum(counts), seed=9, scale=15.0, pz_sigma=3e-3)
    lats = [T.Lattice(ctx, a) for a in arrs]
    for lt in lats:
        assert lt.specialize()
    b = T.Beam.from_host(ctx, beam)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    batch = T.Batch(ctx, [(lats[i], b, int(starts[i]), counts[i]) for i in range(8)])
    print("batch of specialised variants", flush=True)
    batch.track(25, 0)
    ctx.synchronize()
    out = b.download()
    batch.close()
    b.close()
    lost = 0
    for i in range(8):
        sl = slice(int(starts[i]), int(starts[i]) + counts[i])
        part, env = P.envelope(oracle, arrs[i], {k: v[sl].copy() for k, v in beam.items()}, 25)
        mine = {k: v[sl].copy() for k, v in out.items()}
        bad = P.loss_mismatch(mine, part)
        assert bad.size == 0, (i, bad[:5])
        far, dev = P.outside_envelope(mine, part, env, floor=1e-9)
        assert far.size == 0, (i, far[:5], dev[far[:5]], env[far[:5]])
        lost += int((part["state"] != 1).sum())
    assert lost > 100
    for lt in lats:
        lt.close()


def test_lattice_without_repeated_cell(bmx, ctx):
    """No cell repeats over three quarters of the table: specialize reports it and tracking
    keeps the generic kernel."""
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    p0c, et = L.reference()
    eles = [E.make_element(E.DRIFT, p0c, et, l=1.0 + 0.1 * i) for i in range(3)]
    eles += [E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.2), E.make_element(E.MARKER, p0c, et)]
    lat = T.Lattice(ctx, E.to_array(eles))
    assert lat.specialize() is False
    lat.close()


def test_server_with_specialized_lattices(bmx, ctx):
    """Requests on three specialised variants through the continuous-batching server (37-turn
    chunks): every particle equals tracking its request alone with the specialised kernel, bit
    for bit (losses included)."""
    T, L = bmx.tracking, bmx.lattices
    arrs = [bmx.elements.to_array(e) for e in L.ring1_variants(4, 3, n_cells=10)]
    lats = [T.Lattice(ctx, a) for a in arrs]
    for lt in lats:
        assert lt.specialize()
    reqs = [(i % 3, 500 + 700 * i, 5 + 17 * i) for i in range(6)]
    beams = [P.random_beam(bmx, n, seed=70 + i, scale=12.0, pz_sigma=4e-3)
             for i, (_, n, _) in enumerate(reqs)]
    print("server with specialised lattices", flush=True)
    srv = T.Server(ctx, hbm_budget=1 << 30, chunk_turns=37, compact_turns=10)
    ids = [srv.submit(lats[v], beams[i], t, 3) for i, (v, n, t) in enumerate(reqs)]
    results = [srv.wait(r) for r in ids]
    srv.close()
    lost = 0
    for i, (v, n, t) in enumerate(reqs):
        out, _ = results[i]
        alone = T.track_beam_host(arrs[v], beams[i], t, turn0=3, specialize=True)
        for k in alone:
            assert np.array_equal(out[k], alone[k]), (i, k)
        lost += int((alone["state"] != 1).sum())
    assert lost > 500
    for lt in lats:
        lt.close()


def test_specialized_agrees_with_generic(bmx, ctx):
    """Same maps, other code: on the 700-element ring (every orbit regular) the specialised and
    the generic kernel agree to 1e-12 after 20 turns, and on the 901-element ring their loss
    records are identical for a beam that loses a third of its particles in 20 turns."""
    L = bmx.lattices
    arr0 = bmx.elements.to_array(L.ring0())
    beam = L.gaussian_beam(2000, seed=3, sigma=3 * L.SIGMA0)
    a = _track(bmx, ctx, arr0, beam, 20, True)
    b = _track(bmx, ctx, arr0, beam, 20, False)
    assert P.rel_err(a, b) <= 1e-12
    arr1 = bmx.elements.to_array(L.ring1())
    beam = P.random_beam(bmx, 3000, seed=17, scale=12.0, pz_sigma=4e-3)
    a = _track(bmx, ctx, arr1, beam, 20, True)
    b = _track(bmx, ctx, arr1, beam, 20, False)
    assert (b["state"] != 1).sum() > 300
    assert P.loss_mismatch(a, b).size == 0


def test_run_phase_picks_the_kernel_build(bmx, ctx):
    """With waves chosen automatically, a beam that fills more than 1.45 rounds of the 4-wave
    cell kernel's resident waves runs on the 5-wave build and a thinner one on the 4-wave build
    (the run phase); the two builds of the same maps give bit-identical results."""
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    lat = T.Lattice(ctx, E.to_array(L.ring1(n_cells=10)))
    assert lat.specialize()
    beam = L.gaussian_beam(600_000, seed=2, sigma=3 * L.SIGMA0)
    outs = {}
    for waves in (0, 4):
        ctx.set_run(5, 0, waves)
        try:
            b = T.Beam.from_host(ctx, beam)
            print(f"run phase: 6e5 particles, waves {waves}", flush=True)
            T.track_beam(ctx, lat, b, 3, 0)
            ctx.synchronize()
            outs[waves] = (b.download(), ctx.last_run())
            b.close()
        finally:
            ctx.set_run(5, 0, 0)
    (auto, shape_auto), (four, shape_four) = outs[0], outs[4]
    assert shape_auto[2] == 5 and shape_four[2] == 4, (shape_auto, shape_four)
    for k in P.KEYS:
        assert np.array_equal(auto[k], four[k]), k
    small = L.gaussian_beam(20_000, seed=3, sigma=3 * L.SIGMA0)
    b = T.Beam.from_host(ctx, small)
    T.track_beam(ctx, lat, b, 2, 0)
    ctx.synchronize()
    assert ctx.last_run()[2] == 4
    b.close()
    lat.close()
