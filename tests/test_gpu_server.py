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
    per = 76  # bytes of pool per particle
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
