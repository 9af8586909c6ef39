"""Algorithmic FP64 operation counts per particle-element pass.

Convention: add/sub/mul = 1, fma = 2, div = 1, sqrt = 1, each libm transcendental call = 1;
only work that depends on the particle's coordinates is counted (element constants are not).
The per-kind counts in flop_counts.json are measured by running the CPU restatement with a
counting FP64 type (oracle/count_flops.py) over the benchmark beams; they count what the
reference algorithm executes per particle, independent of how the GPU kernel evaluates it.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import elements as E

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flop_counts.json")

# key -> name in flop_counts.json
_KIND = {E.DRIFT: "drift", E.QUADRUPOLE: "quadrupole", E.SEXTUPOLE: "sextupole",
         E.SBEND: "sbend", E.RFCAVITY: "rfcavity", E.MARKER: "marker"}


def table() -> dict:
    with open(_PATH) as f:
        return json.load(f)


def per_element_pass(arr: np.ndarray) -> float:
    """Average FLOPs per particle-element pass over the elements of a lattice array."""
    t = table()["per_kind"]
    total = 0.0
    for e in arr:
        key = int(e["key"])
        name = _KIND.get(key)
        if name is None:
            raise KeyError(f"no FLOP count for element key {key}")
        c = t[name]["flops"]
        if int(e["aperture_at"]) in (E.ENTRANCE_END, E.EXIT_END, E.BOTH_ENDS) and (
                e["x1_limit"] or e["x2_limit"] or e["y1_limit"] or e["y2_limit"]):
            c += t["aperture_elliptical"]["flops"] if int(e["aperture_type"]) == E.APERTURE_ELLIPTICAL \
                else t["aperture_rectangular"]["flops"]
        total += c
    return total / len(arr)
