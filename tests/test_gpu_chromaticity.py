"""Tunes and chromaticities of many lattices at once on the GPU (bmx_chromaticity: chrom_calc.f90
for a closed ring, on the batched closed orbits and one-turn matrices of bmx_closed_orbit),
pinned by closed forms (a ring of thick quadrupoles and drifts, whose tunes at any energy are
those of the product of its thick-lens matrices with k1 / (1 + pz)) and by the CPU restatement
(its own closed orbits and one-turn matrices by the same Newton search on tracked probes, and the
tunes by numpy's eigen-decomposition with make_N's convention), for ring1's error variants with
and without coupling (needs an MI355X)."""
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


KF, KD, LQ, LD, NC = 0.8, -0.65, 0.4, 2.0, 40


def _quad_ring(bmx):
    """NC cells of (focusing quadrupole, drift, defocusing quadrupole, drift); the two strengths
    differ so that the horizontal and vertical tunes do."""
    E, L = bmx.elements, bmx.lattices
    p0c, et = L.reference()
    cell = [E.make_element(E.QUADRUPOLE, p0c, et, l=LQ, k1=KF),
            E.make_element(E.DRIFT, p0c, et, l=LD),
            E.make_element(E.QUADRUPOLE, p0c, et, l=LQ, k1=KD),
            E.make_element(E.DRIFT, p0c, et, l=LD)]
    return E.to_array(cell * NC)


def _thick(k, L):
    if k > 0:
        r = math.sqrt(k)
        return np.array([[math.cos(r * L), math.sin(r * L) / r], [-r * math.sin(r * L), math.cos(r * L)]])
    r = math.sqrt(-k)
    return np.array([[math.cosh(r * L), math.sinh(r * L) / r], [r * math.sinh(r * L), math.cosh(r * L)]])


def _quad_ring_tunes(pz):
    """Total tunes (x, y) of the quadrupole-drift ring at momentum offset pz: the cell's
    thick-lens matrices in (x, x') with K = k1 / (1 + pz) (the vertical plane sees -k1),
    cos mu = tr / 2, tune = NC mu / 2 pi."""
    out = []
    for sgn in (1, -1):
        D = np.array([[1.0, LD], [0.0, 1.0]])
        M = D @ _thick(sgn * KD / (1 + pz), LQ) @ D @ _thick(sgn * KF / (1 + pz), LQ)
        mu = math.acos(np.trace(M) / 2)
        out.append(NC * mu / (2 * math.pi))
    return out


def test_quadrupole_ring_closed_form(bmx, ctx):
    """A 40-cell ring of thick quadrupoles and drifts: the fractional tunes at pz = 0 within 1e-9
    and the chromaticities (Q(+1e-4) - Q(-1e-4)) / 2e-4 within 1e-5 of the closed forms (the
    restatement's own closed-orbit matrices give the same to 1e-8)."""
    arr = _quad_ring(bmx)
    lat = bmx.tracking.Lattice(ctx, arr)
    chrom, tune, st = bmx.tracking.chromaticity(ctx, [lat])
    assert st[0] == 0
    q0 = _quad_ring_tunes(0.0)
    qh = _quad_ring_tunes(1e-4)
    ql = _quad_ring_tunes(-1e-4)
    want_chrom = [(h - l) / 2e-4 for h, l in zip(qh, ql)]
    want_tune = [q % 1.0 for q in q0]
    print("tunes", tune[0], want_tune, "chrom", chrom[0], want_chrom, flush=True)
    assert np.allclose(tune[0], want_tune, rtol=0, atol=1e-9)
    assert np.allclose(chrom[0], want_chrom, rtol=0, atol=1e-5)
    assert chrom[0][0] < -4 and chrom[0][1] < -4  # the natural chromaticity of a FODO ring


def _oracle_matrix(oracle, arr, pz, steps=(1e-6, 1e-6, 1e-6, 1e-6, 1e-5, 1e-6)):
    """The restatement's closed orbit with the rf off at momentum pz (Newton steps on the x, px,
    y, py closing, the Jacobian from five-point central differences of 25 tracked probes) and
    the one-turn matrix at the last iterate."""
    a = arr.copy()
    a["voltage"][a["key"] == 9] = 0.0  # BMX_RFCAVITY
    p0c, et = float(a[0]["p0c"]), float(a[0]["e_tot"])
    x = np.zeros(6)
    x[5] = pz
    for _ in range(50):
        v = np.tile(x, (25, 1))
        for d in range(6):
            for j, f in enumerate((1, -1, 2, -2)):
                v[1 + 4 * d + j, d] += f * steps[d]
        beam = {k: v[:, i].copy() for i, k in enumerate(P.COORDS)}
        beam["beta"] = np.array([p0c * (1 + p) / math.sqrt((p0c * (1 + p)) ** 2 + 0.51099895069e6 ** 2)
                                 if p != 0 else p0c / et for p in beam["pz"]])
        beam["state"] = np.ones(25, dtype=np.int32)
        beam["loss_ele"] = np.zeros(25, dtype=np.int32)
        beam["loss_turn"] = np.zeros(25, dtype=np.int32)
        oracle.track_bunch(a, beam, 1)
        assert np.all(beam["state"] == 1)
        out = np.stack([beam[k] for k in P.COORDS], axis=1)
        M = np.stack([(8 * (out[1 + 4 * d] - out[2 + 4 * d]) - (out[3 + 4 * d] - out[4 + 4 * d]))
                      / (12 * steps[d]) for d in range(6)], axis=1)
        step = np.linalg.solve(M[:4, :4] - np.eye(4), -(out[0, :4] - x[:4]))
        x[:4] += step
        if np.all(np.abs(step) < np.abs(x[:4]) * 1e-12 + 1e-15):
            return M
    raise AssertionError("restatement closed orbit did not converge")


def _numpy_tunes(M4):
    """make_N's tune convention on a 4x4 block with numpy's eigen-decomposition."""
    S = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=float)
    w, V = np.linalg.eig(M4)
    modes = []
    for j in range(4):
        if w[j].imag <= 0:
            continue
        v = V[:, j]
        k = np.argmax(np.abs(v))
        v = v * np.conj(v[k]) / abs(v[k])
        lam = w[j]
        if (v @ S @ np.conj(v)).imag < 0:
            lam = np.conj(lam)
        t = math.atan2(-lam.imag, lam.real) % (2 * math.pi)
        modes.append((t, abs(v[0].real)))
    modes.sort(key=lambda m: -m[1])  # x-dominant first
    return np.array([m[0] / (2 * math.pi) for m in modes])


def _oracle_chromaticity(oracle, arr, de=1e-4):
    tl = _numpy_tunes(_oracle_matrix(oracle, arr, -de)[:4, :4])
    th = _numpy_tunes(_oracle_matrix(oracle, arr, de)[:4, :4])
    t0 = _numpy_tunes(_oracle_matrix(oracle, arr, 0.0)[:4, :4])
    d = th - tl
    d = np.where(d > 0.5, d - 1, np.where(d < -0.5, d + 1, d))
    return d / (2 * de), t0


def test_ring1_variants_against_oracle(bmx, oracle, ctx):
    """Six error variants of the 901-element ring (k1 and k2 errors of 1e-3) and one whose
    quadrupoles are tilted at random by 2e-3 (coupled), in one call: each variant's tunes within
    1e-8 and chromaticities within 1e-5 of the restatement's; the variants differ from each
    other by far more than that."""
    E, L, T = bmx.elements, bmx.lattices, bmx.tracking
    arrs = [E.to_array(v) for v in L.ring1_variants(seed=7, n_variants=6)]
    cp = E.to_array(L.ring1())
    q = np.nonzero(cp["key"] == E.QUADRUPOLE)[0]
    cp["tilt"][q] = 2e-3 * np.random.default_rng(5).standard_normal(q.size)
    arrs.append(cp)
    lats = [T.Lattice(ctx, a) for a in arrs]
    chrom, tune, st = T.chromaticity(ctx, lats)
    assert np.all(st == 0), st
    spread = np.ptp(chrom, axis=0)
    print("chromaticities", chrom.tolist(), "tunes", tune.tolist(), flush=True)
    assert spread.min() > 1e-3
    for i, a in enumerate(arrs):
        want_chrom, want_tune = _oracle_chromaticity(oracle, a)
        assert np.allclose(tune[i], want_tune, rtol=0, atol=1e-8), (i, tune[i], want_tune)
        assert np.allclose(chrom[i], want_chrom, rtol=0, atol=1e-5), (i, chrom[i], want_chrom)
