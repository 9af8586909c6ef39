"""Synthetic rings and beams of the benchmark workloads.

The rings are the FODO rings of the benchmark definitions, built element by element with the
Bmad defaults of ``elements.make_element``:

* ring 0 (100 cells, 700 elements, 760 m): cell = D(1.0) QF(0.3, k1=+0.5) D(1.0)
  B(2.0, angle 2 pi/200) D(1.0) QD(0.3, k1=-0.5) B(2.0); no apertures.
* ring 1 (901 elements, 800.5 m): cell = D QF SF(0.2, k2=+20) D B D QD SD(0.2, k2=-20) B, plus
  one rfcavity (0.5 m, 10 MV, 500 MHz, phi0 = 0.1/(2 pi) in Bmad's units of 2 pi) at the end of
  the ring, elliptical apertures x_limit = 0.020, y_limit = 0.010 (half-axes, as Bmad's
  x_limit/y_limit) on every element.
* error variants of ring 1: every quadrupole k1 and sextupole k2 scaled by (1 + 1e-3 N(0,1)).

Beams are uncorrelated 6-D Gaussians, zero mean, drawn with numpy's PCG64 generator
(``numpy.random.default_rng(seed).standard_normal``); beta follows init_coord
(bmad/low_level/init_coord.f90:205-207).
"""
from __future__ import annotations

import math

import numpy as np

from . import elements as E

E_TOT_5GEV = 5e9
BEND_ANGLE = 2 * E.PI / 200
SIGMA0 = np.array([100e-6, 10e-6, 100e-6, 10e-6, 1e-3, 1e-3])
APERTURE_X, APERTURE_Y = 0.020, 0.010


def reference(e_tot: float = E_TOT_5GEV, mc2: float = E.M_ELECTRON):
    """(p0c, e_tot) of the electron reference particle."""
    return E.e_tot_to_p0c(e_tot, mc2), e_tot


def _ap(with_aperture: bool):
    if not with_aperture:
        return {}
    return dict(x_limit=APERTURE_X, y_limit=APERTURE_Y, aperture_type=E.APERTURE_ELLIPTICAL)


def ring0(n_cells: int = 100, e_tot: float = E_TOT_5GEV):
    """The 700-element FODO ring without sextupoles, RF or apertures."""
    p0c, et = reference(e_tot)
    g = BEND_ANGLE / 2.0
    cell = [
        E.make_element(E.DRIFT, p0c, et, l=1.0),
        E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5),
        E.make_element(E.DRIFT, p0c, et, l=1.0),
        E.make_element(E.SBEND, p0c, et, l=2.0, g=g),
        E.make_element(E.DRIFT, p0c, et, l=1.0),
        E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=-0.5),
        E.make_element(E.SBEND, p0c, et, l=2.0, g=g),
    ]
    return [e for _ in range(n_cells) for e in cell]


def ring1(n_cells: int = 100, e_tot: float = E_TOT_5GEV, k1_scale=None, k2_scale=None,
          aperture: bool = True, rf: bool = True, rf_phi0: float = 0.1 / (2 * E.PI)):
    """The 901-element ring: FODO cells with sextupoles, one rfcavity, apertures everywhere.

    k1_scale / k2_scale: optional per-instance multipliers (len = 2 * n_cells) for the
    quadrupole k1 and sextupole k2 of each cell, in ring order. rf_phi0 is the cavity's phi0
    in units of 2 pi (0.1 rad off the zero crossing by default)."""
    p0c, et = reference(e_tot)
    g = BEND_ANGLE / 2.0
    ap = _ap(aperture)
    out = []
    iq = 0
    for _ in range(n_cells):
        kq = [1.0, 1.0] if k1_scale is None else [k1_scale[iq], k1_scale[iq + 1]]
        ks = [1.0, 1.0] if k2_scale is None else [k2_scale[iq], k2_scale[iq + 1]]
        iq += 2
        out += [
            E.make_element(E.DRIFT, p0c, et, l=1.0, **ap),
            E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=0.5 * kq[0], **ap),
            E.make_element(E.SEXTUPOLE, p0c, et, l=0.2, k2=20.0 * ks[0], **ap),
            E.make_element(E.DRIFT, p0c, et, l=1.0, **ap),
            E.make_element(E.SBEND, p0c, et, l=2.0, g=g, **ap),
            E.make_element(E.DRIFT, p0c, et, l=1.0, **ap),
            E.make_element(E.QUADRUPOLE, p0c, et, l=0.3, k1=-0.5 * kq[1], **ap),
            E.make_element(E.SEXTUPOLE, p0c, et, l=0.2, k2=-20.0 * ks[1], **ap),
            E.make_element(E.SBEND, p0c, et, l=2.0, g=g, **ap),
        ]
    if rf:
        out.append(E.make_element(E.RFCAVITY, p0c, et, l=0.5, voltage=10e6, rf_frequency=500e6,
                                  phi0=rf_phi0, **ap))
    return out


def ring1_variant_scales(seed: int, n_variants: int, n_cells: int = 100, sigma: float = 1e-3):
    """Relative strength errors per variant: arrays (n_variants, 2*n_cells) for k1 and k2."""
    rng = np.random.default_rng(seed)
    k1 = 1.0 + sigma * rng.standard_normal((n_variants, 2 * n_cells))
    k2 = 1.0 + sigma * rng.standard_normal((n_variants, 2 * n_cells))
    return k1, k2


def ring1_variants(seed: int, n_variants: int, n_cells: int = 100, sigma: float = 1e-3):
    """n_variants error variants of ring1 (list of element lists)."""
    k1, k2 = ring1_variant_scales(seed, n_variants, n_cells, sigma)
    return [ring1(n_cells, k1_scale=k1[v], k2_scale=k2[v]) for v in range(n_variants)]


def init_beta(pz: np.ndarray, p0c: float, e_tot: float, mc2: float = E.M_ELECTRON) -> np.ndarray:
    """coord_struct%beta as init_coord sets it (init_coord.f90:205-207, convert_pc_to.f90:37-41)."""
    pc = p0c * (1 + pz)
    beta = pc / np.sqrt(pc ** 2 + mc2 ** 2)
    return np.where(pz == 0, p0c / e_tot, beta)


def gaussian_beam(n: int, seed: int, sigma=SIGMA0, e_tot: float = E_TOT_5GEV):
    """Dict of SoA arrays x..pz, beta, state, loss_ele, loss_turn for an uncorrelated Gaussian."""
    p0c, et = reference(e_tot)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((6, n)) * np.asarray(sigma, dtype=np.float64)[:, None]
    beam = {k: np.ascontiguousarray(v[i]) for i, k in enumerate(("x", "px", "y", "py", "z", "pz"))}
    beam["beta"] = init_beta(beam["pz"], p0c, et)
    beam["state"] = np.full(n, E.STATE_ALIVE, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.full(n, -1, dtype=np.int32)
    return beam


def variant_particle_counts(seed: int, n_variants: int, lo: int = 256, hi: int = 65536):
    """Log-uniform particle counts per variant, floor(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(lo), math.log(hi), n_variants)
    return np.floor(np.exp(u)).astype(np.int64)


def da_grid(n_angles: int = 11, n_radii: int = 64, r_max: float = 0.025,
            pz_values=(-0.01, -0.005, 0.0, 0.005), e_tot: float = E_TOT_5GEV):
    """Dynamic-aperture initial conditions: polar (x, y) grid x pz values, px = py = z = 0."""
    p0c, et = reference(e_tot)
    th = np.linspace(0.0, E.PI / 2, n_angles)
    r = np.linspace(0.0, r_max, n_radii)
    pts = []
    for pz in pz_values:
        for t in th:
            for rr in r:
                pts.append((rr * math.cos(t), 0.0, rr * math.sin(t), 0.0, 0.0, pz))
    a = np.array(pts, dtype=np.float64)
    n = a.shape[0]
    beam = {k: np.ascontiguousarray(a[:, i]) for i, k in enumerate(("x", "px", "y", "py", "z", "pz"))}
    beam["beta"] = init_beta(beam["pz"], p0c, et)
    beam["state"] = np.full(n, E.STATE_ALIVE, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.full(n, -1, dtype=np.int32)
    return beam
