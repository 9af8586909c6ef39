"""Radiation maps for the radiation tests: the maps the reference's radiation_test prints for a
bend (B1: l = 1, g = 0.1, e1 = 0.1, e2 = 0.2) and a quadrupole (K1) of its sigma.bmad ring
(tests/golden/radiation_test_maps.correct, 7 digits), and lattices carrying scaled copies of
them. The maps are realistic inputs (a lower triangular stoc_mat with the reference's structure,
a damp_dmat that is not symmetric); the tests pin what the kick does with any map."""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RAD_MAP_DTYPE = np.dtype([("ref_orb", "<f8", (6,)), ("damp_dmat", "<f8", (6, 6)),
                          ("xfer_damp_vec", "<f8", (6,)), ("stoc_mat", "<f8", (6, 6))])


def reference_maps() -> dict:
    """name ("B1-rm0", "B1-rm1", "K1-rm0", "K1-rm1") -> one bmx_rad_map record."""
    rows = {}
    with open(os.path.join(HERE, "golden", "radiation_test_maps.correct")) as f:
        for ln in f:
            if not ln.strip():
                continue
            name = ln.split('"')[1]
            vals = [float(v) for v in ln.split('"')[2].split()[2:]]
            rows[name] = vals
    out = {}
    for ele in ("B1", "K1"):
        for half in ("rm0", "rm1"):
            p = f"{ele}-{half}-"
            m = np.zeros((), dtype=RAD_MAP_DTYPE)
            m["ref_orb"] = rows[p + "ref_orb"]
            m["damp_dmat"] = [rows[f"{p}damp_dmat:{i}"] for i in range(1, 7)]
            m["xfer_damp_vec"] = rows[p + "xfer_damp_vec"]
            m["stoc_mat"] = [rows[f"{p}stoc_mat:{i}"] for i in range(1, 7)]
            out[f"{ele}-{half}"] = m
    return out


def scaled(m, damp=1.0, stoc=1.0, ref=True):
    """A copy of map m with damp_dmat and xfer_damp_vec times damp, stoc_mat times stoc, and
    ref_orb kept (ref) or zeroed."""
    c = np.array(m, copy=True)
    c["damp_dmat"] *= damp
    c["xfer_damp_vec"] *= damp
    c["stoc_mat"] *= stoc
    if not ref:
        c["ref_orb"] = 0
    return c


def lattice_maps(arr, bmx, damp=1.0, stoc=1.0):
    """Maps for every element of a lattice table: a bend gets B1's, every other element K1's,
    scaled; elements that do not radiate get them too (the kick must ignore them)."""
    ref = reference_maps()
    E = bmx.elements
    maps = np.zeros((len(arr), 2), dtype=RAD_MAP_DTYPE)
    for i, key in enumerate(arr["key"]):
        name = "B1" if key == E.SBEND else "K1"
        maps[i, 0] = scaled(ref[name + "-rm0"], damp, stoc, ref=False)
        maps[i, 1] = scaled(ref[name + "-rm1"], damp, stoc, ref=False)
    return maps
