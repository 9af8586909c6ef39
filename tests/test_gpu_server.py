"""The continuous-batching server (bmx_server_*): results identical alone and in the mix,
admission by the device-pool budget, requests joining while others run."""
import time

import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(bmx):
    return bmx.tracking.Context(0)


def test_results_same_alone_and_in_the_mix(bmx, ctx):
    """Nine requests of 3 lattice variants with different particle and turn counts, submitted
    at once to a server with 37-turn chunks (so requests finish in different chunks and the
    mix changes while they run): every particle equals tracking its request alone with
    bmx_track_turns, bit for bit (losses included)."""
    T, L = bmx.tracking, bmx.lattices
    arrs = [bmx.elements.to_array(e) for e in L.ring1_variants(4, 3, n_cells=10)]
    lats = [T.Lattice(ctx, a) for a in arrs]
    reqs = [(i % 3, 500 + 700 * i, 5 + 17 * i) for i in range(9)]
    beams = [P.random_beam(bmx, n, seed=50 + i, scale=12.0, pz_sigma=4e-3)
             for i, (_, n, _) in enumerate(reqs)]
    srv = T.Server(ctx, hbm_budget=1 << 30, chunk_turns=37, compact_turns=10)
    ids = [srv.submit(lats[v], beams[i], t, 3) for i, (v, n, t) in enumerate(reqs)]
    results = [srv.wait(r) for r in ids]
    srv.close()
    lost = 0
    for i, (v, n, t) in enumerate(reqs):
        out, stats = results[i]
        alone = T.track_beam_host(arrs[v], beams[i], t, turn0=3)
        for k in alone:
            assert np.array_equal(out[k], alone[k]), (i, k)
        assert stats["chunks"] == -(-t // 37)
        lost += int((alone["state"] != 1).sum())
    assert lost > 1000


def test_budget_admission_and_late_arrivals(bmx, ctx):
    """A pool that holds ~2 of the requests at a time: all complete with the right results,
    a request larger than the pool is refused, and requests submitted while others run are
    admitted at a chunk boundary (their admit time is after their submit time and before the
    first requests finish)."""
    T, L = bmx.tracking, bmx.lattices
    arr = bmx.elements.to_array(L.ring1(n_cells=10))
    lat = T.Lattice(ctx, arr)
    per = 188  # bytes of pool per particle
    srv = T.Server(ctx, hbm_budget=per * 9000, chunk_turns=20, compact_turns=5)
    with pytest.raises(bmx.native.BmxError) as ei:
        srv.submit(lat, P.random_beam(bmx, 10000, seed=1), 10)
    assert ei.value.code == 6  # BMX_ERR_BUDGET
    beams = [P.random_beam(bmx, 4000, seed=10 + i, scale=8.0) for i in range(5)]
    first = [srv.submit(lat, beams[i], 200) for i in range(2)]
    time.sleep(0.05)
    late = [srv.submit(lat, beams[i], 60) for i in range(2, 5)]
    res = {r: srv.wait(r) for r in first + late}
    srv.close()
    for j, r in enumerate(first + late):
        out, st = res[r]
        turns = 200 if j < 2 else 60
        alone = T.track_beam_host(arr, beams[j], turns)
        for k in alone:
            assert np.array_equal(out[k], alone[k]), (j, k)
        assert st["admit_s"] >= st["submit_s"]
    # the pool fits two requests: the third waits for one of the first two to finish
    assert res[late[0]][1]["admit_s"] >= min(res[r][1]["done_s"] for r in first) - 1e-3


def test_served_stream_on_full_ring_variants(bmx, oracle, ctx):
    """The served-stream workload at reduced size: requests on error variants of the full
    901-element ring (rf, sextupoles, apertures), particle counts log-uniform in [1e3, 2e4],
    turn counts log-uniform in [10, 300], 100-turn chunks, submitted in three waves while
    earlier requests run (the turn loop of lt_tracking_mod.f90:1571-1716 per request, served
    together). Each request equals its alone-run bit for bit, and against the CPU restatement:
    losses of its first 10 turns identical, and its surviving set within 3x of what a 1e-15
    rescaling of the input changes in the restatement (the orbits of this ring are chaotic:
    see test_ring1_chaotic_1000_turns)."""
    T, L = bmx.tracking, bmx.lattices
    n_var = 4
    variants = L.ring1_variants(8, n_var)
    arrs = [bmx.elements.to_array(v) for v in variants]
    lats = [T.Lattice(ctx, a) for a in arrs]
    rng = np.random.default_rng(44)
    n_req = 6
    counts = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(2e4), n_req))).astype(int)
    turns = np.floor(np.exp(rng.uniform(np.log(10), np.log(300), n_req))).astype(int)
    turns[0] = 300  # one request long enough to span three chunks
    which = rng.integers(0, n_var, n_req)
    beams = [L.gaussian_beam(int(counts[i]), seed=90 + i, sigma=3 * L.SIGMA0) for i in range(n_req)]
    print("served stream:", list(zip(which.tolist(), counts.tolist(), turns.tolist())), flush=True)
    srv = T.Server(ctx, hbm_budget=1 << 30, chunk_turns=100, compact_turns=5)
    ids = []
    for wave in ((0, 1), (2, 3), (4, 5)):
        ids += [srv.submit(lats[which[i]], beams[i], int(turns[i])) for i in wave]
        time.sleep(0.2)
    results = [srv.wait(r) for r in ids]
    srv.close()
    for i in range(n_req):
        out, stats = results[i]
        assert stats["chunks"] >= -(-int(turns[i]) // 100)
        alone = T.track_beam_host(arrs[which[i]], beams[i], int(turns[i]), specialize=False)
        for k in alone:
            assert np.array_equal(out[k], alone[k]), (i, k)
        ref = P.copy_beam(beams[i])
        oracle.track_bunch(arrs[which[i]], ref, int(turns[i]))
        pert = P.copy_beam(beams[i])
        for k in P.COORDS:
            pert[k] = pert[k] * (1 + 1e-15)
        oracle.track_bunch(arrs[which[i]], pert, int(turns[i]))
        early = (ref["state"] != 1) & (ref["loss_turn"] < 10)
        for k in ("state", "loss_ele", "loss_turn"):
            assert np.array_equal(out[k][early], ref[k][early]), (i, k)
        self_flips = int(np.sum((ref["state"] == 1) ^ (pert["state"] == 1)))
        flips = int(np.sum((ref["state"] == 1) ^ (out["state"] == 1)))
        assert flips <= 3 * max(self_flips, 5), (i, flips, self_flips)
    assert sum(int((results[i][0]["state"] != 1).sum()) for i in range(n_req)) > 1000
    for lt in lats:
        lt.close()


def test_zero_turn_request_returns_its_particles(bmx, ctx):
    """A request of no turns comes back unchanged (copied to its output arrays)."""
    T, L = bmx.tracking, bmx.lattices
    lat = T.Lattice(ctx, bmx.elements.to_array(L.ring1(n_cells=2)))
    beam = P.random_beam(bmx, 777, seed=2)
    srv = T.Server(ctx, hbm_budget=1 << 26, chunk_turns=10)
    r0 = srv.submit(lat, beam, 0)
    r1 = srv.submit(lat, beam, 3)
    out0, st0 = srv.wait(r0)
    out1, _ = srv.wait(r1)
    srv.close()
    for k in beam:
        assert np.array_equal(out0[k], beam[k]), k
    assert st0["chunks"] == 0
    assert not np.array_equal(out1["x"], beam["x"])
