"""Closed orbits and one-turn matrices on the GPU (bmx_closed_orbit), pinned by closed forms
(the linear FODO matrix product, symplecticity) and by the CPU restatement (one tracked turn
of the closed orbit returns it)."""
import math

import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(bmx):
    c = bmx.tracking.Context(0)
    c.set_maps("fast")
    return c


def _linear_ring_matrix(arr, E):
    """4x4 one-turn matrix of drifts, thick quadrupoles and sector bends (edges with e = 0 and
    no fringe integral are identity at linear order), element by element."""
    M = np.eye(4)
    for e in arr:
        L = float(e["l"])
        mx = np.array([[1.0, L], [0.0, 1.0]])
        my = mx.copy()
        if e["key"] == E.QUADRUPOLE:
            k = float(e["k1"])
            r = math.sqrt(abs(k))
            ph = r * L
            foc = np.array([[math.cos(ph), math.sin(ph) / r], [-r * math.sin(ph), math.cos(ph)]])
            dfc = np.array([[math.cosh(ph), math.sinh(ph) / r], [r * math.sinh(ph), math.cosh(ph)]])
            mx, my = (foc, dfc) if k > 0 else (dfc, foc)
        elif e["key"] == E.SBEND:
            g = float(e["g"])
            th = g * L
            mx = np.array([[math.cos(th), math.sin(th) / g], [-g * math.sin(th), math.cos(th)]])
        else:
            assert e["key"] == E.DRIFT
        T = np.zeros((4, 4))
        T[:2, :2] = mx
        T[2:, 2:] = my
        M = T @ M
    return M


def _symplectic_defect(M):
    S = np.zeros((6, 6))
    for i in range(3):
        S[2 * i, 2 * i + 1] = 1
        S[2 * i + 1, 2 * i] = -1
    return float(np.max(np.abs(M.T @ S @ M - S)))


def test_fodo_one_turn_matrix(bmx, ctx):
    """700-element FODO ring: the closed orbit is the axis and the 4x4 one-turn matrix equals the
    product of the elements' linear matrices (closed forms) to the finite-difference accuracy;
    the horizontal and vertical tunes agree to 1e-8."""
    E, L, T = bmx.elements, bmx.lattices, bmx.tracking
    arr = E.to_array(L.ring0())
    lat = T.Lattice(ctx, arr)
    orb, mat, st = T.closed_orbit(ctx, [lat], n_dim=4)
    lat.close()
    assert st[0] == 0
    assert np.max(np.abs(orb[0])) < 1e-15
    ref = _linear_ring_matrix(arr, E)
    M4 = mat[0][:4, :4]
    assert np.max(np.abs(M4 - ref)) <= 1e-7 * max(1.0, np.max(np.abs(ref)))
    for p in (0, 2):
        cos_mu = 0.5 * np.trace(M4[p:p + 2, p:p + 2])
        assert abs(cos_mu - 0.5 * np.trace(ref[p:p + 2, p:p + 2])) < 1e-8
        assert abs(cos_mu) < 1  # stable
    assert abs(np.linalg.det(M4) - 1) < 1e-7


def _kicked_ring(bmx, n_cells=10, kick=2e-5, k1_scale=None, k2_scale=None):
    E, L = bmx.elements, bmx.lattices
    p0c, et = L.reference()
    eles = L.ring1(n_cells, k1_scale=k1_scale, k2_scale=k2_scale)
    eles.insert(1, E.make_element(E.HKICKER, p0c, et, l=0.5, kick=kick))
    return E.to_array(eles)


def _one_turn(oracle, arr, orbit):
    beam = {k: np.array([orbit[i]]) for i, k in enumerate(P.COORDS)}
    pz = beam["pz"]
    p0c, et = float(arr[0]["p0c"]), float(arr[0]["e_tot"])
    pc = p0c * (1 + pz)
    beam["beta"] = np.where(pz == 0, p0c / et, pc / np.sqrt(pc ** 2 + 0.51099895069e6 ** 2))
    beam["state"] = np.ones(1, dtype=np.int32)
    beam["loss_ele"] = np.zeros(1, dtype=np.int32)
    beam["loss_turn"] = np.zeros(1, dtype=np.int32)
    oracle.track_bunch(arr, beam, 1)
    assert beam["state"][0] == 1
    return np.array([beam[k][0] for k in P.COORDS])


def test_closed_orbit_4d_is_a_fixed_point(bmx, oracle, ctx):
    """A 10-cell ring with sextupoles and a horizontal kicker (2e-5 rad), rf off: the 4-D closed
    orbit found on the GPU is a fixed point of the CPU restatement's one-turn map to 1e-13 in
    x, px, y, py, and its one-turn matrix is symplectic to 1e-6."""
    T = bmx.tracking
    arr = _kicked_ring(bmx)
    lat = T.Lattice(ctx, arr)
    orb, mat, st = T.closed_orbit(ctx, [lat], n_dim=4)
    lat.close()
    assert st[0] == 0
    assert 1e-5 < abs(orb[0][0]) < 1e-2, orb[0]
    off = arr.copy()
    off["voltage"][off["key"] == bmx.elements.RFCAVITY] = 0.0
    out = _one_turn(oracle, off, orb[0])
    assert np.max(np.abs(out[:4] - orb[0][:4])) <= 1e-13
    # with the rf off z does not feed back: the z column is (0, 0, 0, 0, 1, 0)
    assert np.max(np.abs(mat[0][:, 4] - np.eye(6)[:, 4])) < 1e-8
    assert _symplectic_defect(mat[0]) < 1e-6


def test_closed_orbit_6d_batch_of_variants(bmx, oracle, ctx):
    """Eight error variants of the kicked 10-cell ring with the rf on, in one call: every 6-D
    closed orbit is a fixed point of the CPU restatement's one-turn map (1e-12 in every
    coordinate) and every one-turn matrix is symplectic to 1e-6."""
    T, L = bmx.tracking, bmx.lattices
    k1, k2 = L.ring1_variant_scales(11, 8, n_cells=10)
    arrs = [_kicked_ring(bmx, k1_scale=k1[v], k2_scale=k2[v]) for v in range(8)]
    lats = [T.Lattice(ctx, a) for a in arrs]
    orb, mat, st = T.closed_orbit(ctx, lats, n_dim=6, max_iter=40)
    for lt in lats:
        lt.close()
    assert np.all(st == 0), st
    for v in range(8):
        out = _one_turn(oracle, arrs[v], orb[v])
        assert np.max(np.abs(out - orb[v])) <= 1e-12, (v, out - orb[v])
        assert _symplectic_defect(mat[v]) < 1e-6
    # the variants differ: their orbits are not all the same
    assert np.ptp(orb[:, 0]) > 0


def test_checkpoint_resume_is_bit_identical(bmx, ctx, tmp_path):
    """Track 12 turns, write the beam file, read it back and track 13 more from turn 12: every
    particle equals 25 turns in one call, bit for bit (losses included)."""
    T, L = bmx.tracking, bmx.lattices
    arr = bmx.elements.to_array(L.ring1(n_cells=10))
    p0c, _ = L.reference()
    beam = P.random_beam(bmx, 3000, seed=12, scale=12.0, pz_sigma=4e-3)
    lat = T.Lattice(ctx, arr)
    b = T.Beam.from_host(ctx, beam)
    T.track_beam(ctx, lat, b, 25)
    ctx.synchronize()
    straight = b.download()
    b.upload(beam)
    T.track_beam(ctx, lat, b, 12)
    ctx.synchronize()
    mid = b.download()
    path = str(tmp_path / "checkpoint.dat")
    T.write_beam_file(path, mid, p0c, ix_ele=len(arr))
    back, _ = T.read_beam_file(path, p0c)
    back["loss_turn"] = mid["loss_turn"]  # the file has no loss turn
    b.upload(back)
    T.track_beam(ctx, lat, b, 13, 12)
    ctx.synchronize()
    resumed = b.download()
    b.close()
    lat.close()
    assert (straight["state"] != 1).sum() > 100
    for k in straight:
        assert np.array_equal(resumed[k], straight[k]), k
