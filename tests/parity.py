"""Shared helpers for comparing GPU results with the CPU restatement."""
from __future__ import annotations

import numpy as np

COORDS = ("x", "px", "y", "py", "z", "pz")
# Relative-error floor per coordinate: far below the beam sizes (>= 1e-5), so the error of a
# coordinate is relative to its value except where that value crosses zero.
FLOOR = {"x": 1e-6, "px": 1e-7, "y": 1e-6, "py": 1e-7, "z": 1e-6, "pz": 1e-7}


def copy_beam(beam):
    return {k: v.copy() for k, v in beam.items()}


def rel_err(got, ref, mask=None):
    """Max over particles and coordinates of |got - ref| / max(|ref|, scale_k), where scale_k
    is the beam's RMS of coordinate k (at least FLOOR[k]).

    The scale keeps the measure relative for particles whose coordinate happens to sit near
    zero: several reference maps compute a coordinate as a difference of terms of order the
    element length (e.g. z in the bend body, track_a_bend.f90:209), so their own rounding is
    absolute, ~1 ulp of those terms, not relative to the result."""
    worst = 0.0
    for k in COORDS:
        g, r = got[k], ref[k]
        if mask is not None:
            g, r = g[mask], r[mask]
        if g.size == 0:
            continue
        scale = max(FLOOR[k], float(np.sqrt(np.mean(r * r))))
        e = np.abs(g - r) / np.maximum(np.abs(r), scale)
        worst = max(worst, float(np.max(e)))
    return worst


def rel_dev(got, ref, mask):
    """Per particle: max over coordinates of the rel_err measure (inf outside mask)."""
    dev = np.full(mask.shape, np.inf)
    if not mask.any():
        return dev
    d = np.zeros(int(mask.sum()))
    for k in COORDS:
        g, r = got[k][mask], ref[k][mask]
        scale = max(FLOOR[k], float(np.sqrt(np.mean(r * r))))
        d = np.maximum(d, np.abs(g - r) / np.maximum(np.abs(r), scale))
    dev[mask] = d
    return dev


def fma_noise(oracle, arr, beam, n_turns, turn0=0, **kw):
    """Per-particle deviation between the two FP64 builds of the CPU restatement (without and
    with FMA contraction) after n_turns: how much rounding alone moves each orbit. Returns
    (plain-build result, deviation; inf where the builds disagree on survival)."""
    ref = copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns, turn0=turn0, **kw)
    fma = copy_beam(beam)
    oracle.track_bunch(arr, fma, n_turns, turn0=turn0, fma=True, **kw)
    both = (ref["state"] == 1) & (fma["state"] == 1)
    return ref, rel_dev(fma, ref, both)


def envelope(oracle, arr, beam, n_turns, turn0=0, delta=1e-14, samples=4, seed=0,
             with_fma_spread=False, **kw):
    """Per-particle sensitivity of the restatement's own result: the largest deviation (rel_dev
    measure) from the plain build's result among (a) the FMA-contracted build, which differs
    from it by rounding in every contracted operation along the whole orbit, and (b) `samples`
    runs whose initial coordinates are each rescaled by (1 +- delta) with random signs, which
    measure how much the orbit amplifies a small error. (a) dominates on regular orbits, where
    rounding noise accumulates; (b) on chaotic ones, where it is amplified exponentially; a
    single ulp-sized rescaling can be absorbed by an addition, hence delta = 1e-14 (about 45
    ulp). Returns (plain-build result, envelope; inf where a run disagrees on survival), and the
    FMA build's deviation alone as a third item if with_fma_spread. kw: further arguments of
    the restatement's track_bunch (radiation maps and settings)."""
    ref, noise = fma_noise(oracle, arr, beam, n_turns, turn0, **kw)
    env = noise
    rng = np.random.default_rng(seed)
    for _ in range(samples):
        b = copy_beam(beam)
        for k in COORDS:
            b[k] = b[k] * (1 + delta * rng.choice([-1.0, 1.0], size=b[k].shape))
        oracle.track_bunch(arr, b, n_turns, turn0=turn0, **kw)
        both = (ref["state"] == 1) & (b["state"] == 1)
        env = np.maximum(env, rel_dev(b, ref, both))
    return (ref, env, noise) if with_fma_spread else (ref, env)


def outside_envelope(got, ref, env, factor=100.0, floor=1e-10):
    """Survivors whose deviation from the restatement exceeds max(floor, factor x envelope):
    with delta = 1e-14 in envelope(), factor 100 is the response to a 1e-12 relative change of
    the input, the per-element parity bar. Returns (indices, per-particle deviation)."""
    alive = ref["state"] == 1
    dev = rel_dev(got, ref, alive)
    bad = alive & (dev > np.maximum(floor, factor * env))
    return np.nonzero(bad)[0], dev


_CACHE = {}


def cached(key, make):
    """make() once per test session for a key (the CPU runs of the long parity tests are shared
    by the tests that compare several GPU kernels with them)."""
    if key not in _CACHE:
        _CACHE[key] = make()
    return _CACHE[key]


def within_noise(got, ref, noise, factor=10.0, floor=1e-10):
    """Indices of surviving particles whose GPU deviation exceeds max(floor, factor x noise)."""
    alive = ref["state"] == 1
    dev = rel_dev(got, ref, alive)
    bad = alive & (dev > np.maximum(floor, factor * noise))
    return np.nonzero(bad)[0]


def loss_mismatch(got, ref):
    """Indices where state / loss element / loss turn differ."""
    bad = ((got["state"] != ref["state"]) | (got["loss_ele"] != ref["loss_ele"]) |
           (got["loss_turn"] != ref["loss_turn"]))
    return np.nonzero(bad)[0]


def random_beam(bmx, n, seed, scale=3.0, pz_sigma=None):
    """Gaussian beam at `scale` times the base sigmas (optionally another pz sigma)."""
    L = bmx.lattices
    sig = L.SIGMA0 * scale
    if pz_sigma is not None:
        sig = sig.copy()
        sig[5] = pz_sigma
    return L.gaussian_beam(n, seed=seed, sigma=sig)

KEYS = COORDS + ("beta", "state", "loss_ele", "loss_turn")


def ring1_norf_1000(bmx, oracle):
    """The 901-element ring with its rf cavity off, a 1-sigma beam of 1024, 1000 turns on the
    CPU: (lattice array, beam, restatement result, FMA spread, envelope)."""
    def make():
        L = bmx.lattices
        arr = bmx.elements.to_array(L.ring1(rf=False))
        beam = L.gaussian_beam(1024, seed=1, sigma=L.SIGMA0)
        ref, env, noise = envelope(oracle, arr, beam, 1000, with_fma_spread=True)
        return arr, beam, ref, noise, env
    return cached("ring1_norf_1000", make)


def ring1_full_1000(bmx, oracle):
    """The full 901-element ring, a 3-sigma beam of 1024, 1000 turns on the CPU: (lattice array,
    beam, restatement result, survival flips that a (1 + 1e-15) rescaling of the input causes)."""
    def make():
        L = bmx.lattices
        arr = bmx.elements.to_array(L.ring1())
        beam = L.gaussian_beam(1024, seed=1, sigma=3 * L.SIGMA0)
        ref = copy_beam(beam)
        oracle.track_bunch(arr, ref, n_turns=1000)
        pert = copy_beam(beam)
        for k in COORDS:
            pert[k] = pert[k] * (1 + 1e-15)
        oracle.track_bunch(arr, pert, n_turns=1000)
        self_flips = int(np.sum((ref["state"] == 1) ^ (pert["state"] == 1)))
        return arr, beam, ref, self_flips
    return cached("ring1_full_1000", make)


def check_norf_1000(out, ref, noise, env):
    """The bars of the 1000-turn test without rf: every particle survives; the deviation's
    median, 90th and 99th percentile within 3x of the FMA spread's, the median within 1e-9, and
    every particle within max(1e-9, 100x its envelope)."""
    assert np.all(out["state"] == 1) and np.all(ref["state"] == 1) and np.all(np.isfinite(noise))
    dev = rel_dev(out, ref, ref["state"] == 1)
    for q in (50, 90, 99):
        assert np.percentile(dev, q) <= 3 * np.percentile(noise, q), (q, np.percentile(dev, q),
                                                                      np.percentile(noise, q))
    assert np.median(dev) <= 1e-9
    bad, _ = outside_envelope(out, ref, env, floor=1e-9)
    print("1000 turns without rf: deviation max", float(dev.max()), "envelope at it",
          float(env[np.argmax(dev)]), flush=True)
    assert bad.size == 0, (bad[:10], dev[bad[:10]], env[bad[:10]])


def check_full_1000(out, ref, self_flips):
    """The bars of the chaotic 1000-turn test: losses in the first 10 turns identical, and the
    surviving sets differ by at most 3x what a 1e-15 rescaling changes in the restatement."""
    early = (ref["state"] != 1) & (ref["loss_turn"] < 10)
    assert early.sum() > 20
    for k in ("state", "loss_ele", "loss_turn"):
        assert np.array_equal(out[k][early], ref[k][early]), k
    flips = int(np.sum((ref["state"] == 1) ^ (out["state"] == 1)))
    assert flips <= 3 * max(self_flips, 5), (flips, self_flips)
