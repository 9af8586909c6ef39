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


def fma_noise(oracle, arr, beam, n_turns, turn0=0):
    """Per-particle deviation between the two FP64 builds of the CPU restatement (without and
    with FMA contraction) after n_turns: how much rounding alone moves each orbit. Returns
    (plain-build result, deviation; inf where the builds disagree on survival)."""
    ref = copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns, turn0=turn0)
    fma = copy_beam(beam)
    oracle.track_bunch(arr, fma, n_turns, turn0=turn0, fma=True)
    both = (ref["state"] == 1) & (fma["state"] == 1)
    return ref, rel_dev(fma, ref, both)


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
