"""The reference's own known-answer vectors run on the GPU (through the C-ABI), and misaligned
rings against the CPU restatement."""
import numpy as np
import pytest

import kat
import parity as P

pytestmark = pytest.mark.gpu

# not on the GPU path: apertures that move with the offsets, rf cavity multipoles
OFF_PATH = {"RCOLLIMATOR1", "RFCAVITY3"}


@pytest.mark.parametrize("maps", ["fast", "reference"])
def test_tracking_method_known_answers_on_gpu(bmx, maps):
    """tracking_method_test "<ELE>: Bmad_Standard" lines (ABS 2e-14, written es18.10): drift,
    thick multipoles (sextupole, octupole, thick_multipole, kickers, collimator, instrument,
    monitor) with offsets, pitch and tilt, the rf cavity, quadrupoles with magnetic and
    electric multipoles and edges, and sector bends with offsets, ref_tilt, multipoles, k1, k2
    and the full edge, each tracked by the GPU kernel from the fixture's particle_start as a
    one-element lattice (the lattice's default species, the positron)."""
    E, T = bmx.elements, bmx.tracking
    fx = kat.read_fixture("tracking_method_test.correct")
    eles = kat.tracking_method_elements(E)
    beta0 = kat.tm_start_beta(E)
    failures = []
    ran = 0
    for name, ele in eles.items():
        if name in OFF_PATH:
            continue
        typ, tol, expect = fx[f"{name}: Bmad_Standard"]
        beam = {k: np.array([v]) for k, v in zip(P.COORDS, kat.TM_START)}
        beam["beta"] = np.array([beta0])
        beam["state"] = np.ones(1, dtype=np.int32)
        beam["loss_ele"] = np.zeros(1, dtype=np.int32)
        beam["loss_turn"] = np.zeros(1, dtype=np.int32)
        out = T.track_beam_host(E.to_array([ele]), beam, 1, maps=maps, charge=kat.TM_CHARGE)
        assert out["state"][0] == E.STATE_ALIVE, name
        vec = [float(out[k][0]) for k in P.COORDS]
        ok, worst = kat.abs_match(vec, expect[:6], tol, 10)
        if not ok:
            failures.append((name, worst, [kat.fortran_es(v, 10) for v in vec], expect[:6]))
        ran += 1
    assert not failures, failures
    assert ran == len(eles) - len(OFF_PATH)


def _misaligned_ring(bmx, seed, n_cells=10):
    """ring1 of n_cells with every quadrupole and sextupole offset, pitched and tilted at random
    (sigma 1e-4 m, 1e-4 rad, 1e-3 rad)."""
    E, L = bmx.elements, bmx.lattices
    rng = np.random.default_rng(seed)
    arr = E.to_array(L.ring1(n_cells))
    sel = (arr["key"] == E.QUADRUPOLE) | (arr["key"] == E.SEXTUPOLE)
    k = int(sel.sum())
    arr["x_offset"][sel] = 1e-4 * rng.standard_normal(k)
    arr["y_offset"][sel] = 1e-4 * rng.standard_normal(k)
    arr["x_pitch"][sel] = 1e-4 * rng.standard_normal(k)
    arr["y_pitch"][sel] = 1e-4 * rng.standard_normal(k)
    arr["tilt"][sel] = 1e-3 * rng.standard_normal(k)
    return arr


@pytest.mark.parametrize("specialize", [False, True])
def test_misaligned_ring_against_oracle(bmx, oracle, specialize):
    """A 10-cell ring with misaligned, pitched and tilted quadrupoles and sextupoles, 20 turns,
    generic and cell-specialised kernels: losses bit-exact and the survivors within 3x of the
    spread FMA contraction alone causes in the CPU restatement (median, 90th, 99th
    percentile)."""
    T = bmx.tracking
    arr = _misaligned_ring(bmx, 3)
    beam = P.random_beam(bmx, 2000, seed=8, scale=15.0, pz_sigma=3e-3)
    ref, noise = P.fma_noise(oracle, arr, beam, 20)
    out = T.track_beam_host(arr, beam, 20, specialize=specialize)
    assert (ref["state"] != 1).sum() > 50
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    dev = P.rel_dev(out, ref, alive)[alive]
    nz = noise[alive]
    for q in (50, 90, 99):
        assert np.percentile(dev, q) <= 3 * np.percentile(nz, q) + 1e-13, (
            q, np.percentile(dev, q), np.percentile(nz, q))


def test_lattice_file_ring_against_oracle(bmx, oracle):
    """The test ring written in the Bmad language (tests/golden/lattices/ring_small.bmad: rbends
    with edge angles, sextupoles, a corrector, monitors, an rf cavity given by frequency), loaded
    by the lattice-file loader and tracked 30 turns: losses bit-exact and the survivors within 3x
    of the oracle's FMA spread."""
    import importlib
    import os
    lf = importlib.import_module(bmx.__name__ + ".lattice_file")
    eles, mc2, charge = lf.load(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                             "golden", "lattices", "ring_small.bmad"))
    E, T, L = bmx.elements, bmx.tracking, bmx.lattices
    arr = E.to_array(eles)
    beam = P.random_beam(bmx, 2000, seed=4, scale=30.0, pz_sigma=3e-3)
    p0c, et = float(arr[0]["p0c"]), float(arr[0]["e_tot"])
    beam["beta"] = L.init_beta(beam["pz"], p0c, et, mc2)
    ref, noise = P.fma_noise(oracle, arr, beam, 30)
    out = T.track_beam_host(arr, beam, 30, mc2=mc2, charge=charge)
    assert (ref["state"] != 1).sum() > 20
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    dev = P.rel_dev(out, ref, alive)[alive]
    nz = noise[alive]
    for q in (50, 90, 99):
        assert np.percentile(dev, q) <= 3 * np.percentile(nz, q) + 1e-13, (
            q, np.percentile(dev, q), np.percentile(nz, q))


@pytest.mark.parametrize("maps", ["fast", "reference"])
def test_aperture_known_answers_on_gpu(bmx, maps):
    """aperture_test EC_S_* / RC_S_* lines: the state codes of 54 particles per collimator
    (x-y apertures with px, py, z, pz windows, elliptical pairs for the ecollimator) after the
    exit-end test of a one-element lattice on the GPU, exactly as the fixture lists them."""
    E, T, L = bmx.elements, bmx.tracking, bmx.lattices
    fx = kat.read_fixture("aperture_test.correct")
    mc2 = 105.6583755e6
    for ename, ele in kat.aperture_test_elements(E).items():
        cases = kat.aperture_test_cases()
        vecs = np.array([v for _, vs in cases for v in vs], dtype=np.float64)
        beam = {k: np.ascontiguousarray(vecs[:, i]) for i, k in enumerate(P.COORDS)}
        beam["beta"] = L.init_beta(beam["pz"], ele.p0c, ele.e_tot, mc2)
        n = len(vecs)
        beam["state"] = np.ones(n, dtype=np.int32)
        beam["loss_ele"] = np.zeros(n, dtype=np.int32)
        beam["loss_turn"] = np.zeros(n, dtype=np.int32)
        out = T.track_beam_host(E.to_array([ele]), beam, 1, mc2=mc2, charge=1, maps=maps)
        k = 0
        for sfx, vs in cases:
            _, _, exp_s = fx[f"{ename}_S_{sfx}"]
            got = [int(s) for s in out["state"][k:k + len(vs)]]
            assert got == [int(x) for x in exp_s], (ename, sfx, got, exp_s)
            k += len(vs)


@pytest.mark.parametrize("specialize", [False, True])
def test_tutorial_storage_ring_against_oracle(bmx, oracle, specialize):
    """A storage ring from the reference's ring-design tutorial (bmad-doc/tutorial_ring_design/
    lattices/12_DynamicAperture/12.1_CalculatingDA/ring.bmad, kept as a fixture): 3332
    elements, 17.8 GeV electrons, 16 rf cavities, overlays on its sextupoles and
    start_branch_at, loaded by the lattice-file loader; 20 turns of 2000 particles on the
    generic and the cell-specialised kernels: losses bit-exact and survivors within
    max(1e-10, 100x) of the restatement's own sensitivity envelope."""
    import importlib
    import os
    lf = importlib.import_module(bmx.__name__ + ".lattice_file")
    eles, mc2, charge = lf.load(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                             "golden", "lattices", "tutorial", "da_ring.bmad"))
    E, T, L = bmx.elements, bmx.tracking, bmx.lattices
    arr = E.to_array(eles)
    assert len(arr) == 3332 and (arr["key"] == E.RFCAVITY).sum() == 16
    p0c, et = float(arr[0]["p0c"]), float(arr[0]["e_tot"])
    rng = np.random.default_rng(17)
    n = 2000
    beam = {k: s * rng.standard_normal(n) for k, s in
            zip(P.COORDS, (5e-4, 5e-5, 5e-4, 5e-5, 3e-3, 1.5e-3))}
    beam["beta"] = L.init_beta(beam["pz"], p0c, et, mc2)
    beam["state"] = np.ones(n, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.zeros(n, dtype=np.int32)
    ref, env = P.cached(("tutorial_ring", 20),
                        lambda: P.envelope(oracle, arr, beam, 20))
    assert 20 < (ref["state"] != 1).sum() < n - 200
    print(f"tutorial ring specialize={specialize}", flush=True)
    out = T.track_beam_host(arr, beam, 20, mc2=mc2, charge=charge, specialize=specialize)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])
