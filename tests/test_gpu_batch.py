"""Batched tracking (bmx_batch_*): many lattice variants and uneven particle sets in one
captured launch sequence, against one-lattice-at-a-time tracking and the CPU restatement."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(bmx):
    return bmx.tracking.Context(0)


def _variants(bmx, n_var, seed, n_cells=10):
    L = bmx.lattices
    return [bmx.elements.to_array(e) for e in L.ring1_variants(seed, n_var, n_cells=n_cells)]


def test_batch_equals_single_lattice_runs(bmx, ctx):
    """7 error variants with uneven particle counts (1 .. 3000, some sharing one beam): every
    particle's result equals tracking its variant alone, bit for bit, including losses;
    replaying the captured graph for a second call continues correctly."""
    T = bmx.tracking
    arrs = _variants(bmx, 7, seed=2)
    counts = [1, 63, 64, 65, 1000, 3000, 257]
    lats = [T.Lattice(ctx, a) for a in arrs]
    beams_h = [P.random_beam(bmx, c, seed=100 + i, scale=12.0, pz_sigma=4e-3) for i, c in enumerate(counts)]
    # groups 5 and 6 share one beam (disjoint ranges)
    shared = {k: np.concatenate([beams_h[5][k], beams_h[6][k]]) for k in beams_h[5]}
    b_shared = T.Beam.from_host(ctx, shared)
    beams = [T.Beam.from_host(ctx, beams_h[i]) for i in range(5)]
    groups = [(lats[i], beams[i], 0, counts[i]) for i in range(5)]
    groups += [(lats[5], b_shared, 0, counts[5]), (lats[6], b_shared, counts[5], counts[6])]
    batch = T.Batch(ctx, groups)
    batch.track(6, 0)
    batch.track(5, 6)
    ctx.synchronize()
    outs = [b.download() for b in beams]
    o_sh = b_shared.download()
    outs.append({k: v[: counts[5]] for k, v in o_sh.items()})
    outs.append({k: v[counts[5]:] for k, v in o_sh.items()})
    lost = 0
    for i in range(7):
        alone = T.track_beam_host(arrs[i], beams_h[i], 11)
        for k in alone:
            assert np.array_equal(outs[i][k], alone[k]), (i, k)
        lost += int((alone["state"] != 1).sum())
    assert lost > 300
    batch.close()


def test_batch_matches_oracle(bmx, oracle, ctx):
    """3 variants of the full 901-element ring, 2000 particles each, 20 turns: losses exact;
    the survivors' deviations from the CPU restatement within 3x of the spread that FMA
    contraction alone causes in the CPU restatement (median, 90th, 99th percentile): the
    large-amplitude orbits of this ring are chaotic and amplify any rounding difference."""
    T = bmx.tracking
    arrs = _variants(bmx, 3, seed=7, n_cells=100)
    lats = [T.Lattice(ctx, a) for a in arrs]
    hb = [P.random_beam(bmx, 2000, seed=i, scale=6.0, pz_sigma=3e-3) for i in range(3)]
    beams = [T.Beam.from_host(ctx, h) for h in hb]
    batch = T.Batch(ctx, [(lats[i], beams[i], 0, 2000) for i in range(3)])
    batch.track(20, 0)
    ctx.synchronize()
    for i in range(3):
        out = beams[i].download()
        ref, noise = P.fma_noise(oracle, arrs[i], hb[i], 20)
        assert P.loss_mismatch(out, ref).size == 0
        alive = ref["state"] == 1
        dev = P.rel_dev(out, ref, alive)[alive]
        nz = noise[alive]
        for q in (50, 90, 99):
            assert np.percentile(dev, q) <= 3 * np.percentile(nz, q) + 1e-13, (i, q)
    batch.close()


def test_batch_tune_keeps_beams(bmx, ctx):
    """The autotuner works on a scratch copy: the batch's beams are unchanged by tuning, and
    the tuned run shape gives the same results as the default one."""
    T = bmx.tracking
    arrs = _variants(bmx, 4, seed=3)
    lats = [T.Lattice(ctx, a) for a in arrs]
    hb = [P.random_beam(bmx, 700 * (i + 1), seed=i, scale=10.0) for i in range(4)]
    beams = [T.Beam.from_host(ctx, h) for h in hb]
    batch = T.Batch(ctx, [(lats[i], beams[i], 0, len(hb[i]["x"])) for i in range(4)])
    shape = batch.tune(sample_turns=4)
    assert shape["chunk_turns"] > 0 and shape["grid"] > 0 and shape["waves"] in (3, 4, 5)
    for i in range(4):
        now = beams[i].download()
        for k in now:
            assert np.array_equal(now[k], hb[i][k]), (i, k)
    batch.track(9, 0)
    ctx.synchronize()
    for i in range(4):
        alone = T.track_beam_host(arrs[i], hb[i], 9)
        out = beams[i].download()
        for k in out:
            assert np.array_equal(out[k], alone[k]), (i, k)
    batch.close()


def test_dynamic_aperture_survival_turns(bmx, oracle, ctx):
    """Dynamic-aperture grid (11 angles x 16 radii to 25 mm x 4 pz) on 2 error variants of the
    901-element ring, 50 turns: survival turn and loss element identical to the CPU
    restatement for every grid point."""
    T, L = bmx.tracking, bmx.lattices
    arrs = _variants(bmx, 2, seed=3, n_cells=100)
    grid = L.da_grid(n_radii=16)
    lats = [T.Lattice(ctx, a) for a in arrs]
    surv, loss_ele, _ = T.dynamic_aperture(ctx, lats, grid, 50)
    for i in range(2):
        ref = P.copy_beam(grid)
        oracle.track_bunch(arrs[i], ref, 50)
        ref_surv = np.where(ref["state"] == 1, 50, ref["loss_turn"])
        assert np.array_equal(surv[i], ref_surv)
        assert np.array_equal(loss_ele[i], ref["loss_ele"])
    assert (surv < 50).sum() > 100 and (surv == 50).sum() > 20
