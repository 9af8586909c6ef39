"""The CPU restatement of the radiation kicks (track1_radiation, radiation_mod.f90:61-147, where
track1 calls it, track1.f90:197-202 and 319-324), pinned without retyping its formulas:

* the fluctuation generator: Philox4x32-10 against the known-answer vectors published with
  Random123 (kat_vectors, philox4x32 with 10 rounds), and its Box-Muller deviates against the
  standard normal distribution (moments, independence of the six components, Kolmogorov-Smirnov);
* the kick's place in the element: with a map on one half only, the tracked result equals the
  element's own tracking without radiation followed (downstream half) or preceded (upstream
  half) by the kick computed in numpy from the fixture maps of the reference's radiation_test;
  a transposed damp_dmat, a kick in the element's body frame instead of the laboratory frame,
  the wrong half's map, a missing zero_average switch or scale factor each fail one of these;
* the fluctuation kick is stoc_mat's lower triangle times the deviates of the particle's own
  (index, turn, element, edge) counter, and its covariance over many particles is
  scale^2 L L^T (a transposed L gives L^T L);
* the kinds track1_radiation skips (drift, pipe, marker, thin multipoles, zero length, switched
  off) are tracked as without radiation;
* a kick that takes pz below -1 loses the particle (lost_pz) at that element, at either edge,
  and a kick that makes the orbit too large loses it by orbit_too_large's state."""
import ctypes

import numpy as np
import pytest
from scipy import stats

import parity as P
import rad_fixtures as R


def test_philox_known_answers(oracle):
    """Random123's kat_vectors for philox4x32_10: (counter, key) -> output."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, want in kat:
        assert oracle.philox4x32_10(ctr, key) == want


def test_deviates_are_standard_normal(oracle):
    """60000 kicks' six deviates: each component has mean 0 and variance 1 within 4.5 standard
    errors, the components are uncorrelated, the pooled sample passes Kolmogorov-Smirnov against
    N(0, 1), and deviates of other edges, turns, elements and seeds are other numbers."""
    n = 60000
    g = np.array([oracle.rad_gauss6(12345, i, 3, 17, 1) for i in range(n)])
    se = 1 / np.sqrt(n)
    assert np.all(np.abs(g.mean(axis=0)) < 4.5 * se), g.mean(axis=0)
    assert np.all(np.abs(g.var(axis=0) - 1) < 4.5 * np.sqrt(2) * se), g.var(axis=0)
    c = np.corrcoef(g.T)
    assert np.all(np.abs(c[~np.eye(6, dtype=bool)]) < 4.5 * se), c
    assert stats.kstest(g.ravel(), "norm").pvalue > 1e-4
    base = oracle.rad_gauss6(12345, 7, 3, 17, 1)
    for args in ((12345, 7, 3, 17, 0), (12345, 7, 4, 17, 1), (12345, 7, 3, 18, 1),
                 (12346, 7, 3, 17, 1), (12345, 8, 3, 17, 1)):
        assert np.all(oracle.rad_gauss6(*args) != base), args


def _quad(bmx, **kw):
    """A tilted, offset quadrupole at 5 GeV: its body frame differs from the laboratory frame."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    kw = {"l": 0.6, "k1": 0.4, "tilt": 0.3, "x_offset": 2e-4, "y_offset": -1e-4, **kw}
    return E.to_array([E.make_element(E.QUADRUPOLE, p0c, et, **kw)])


def _cfg(oracle, damping=True, fluct=False, zero_average=False, scale=1.0, seed=99):
    return oracle.RadConfig(int(damping), int(fluct), int(zero_average), 0, scale, seed)


def _maps(m0=None, m1=None):
    maps = np.zeros((1, 2), dtype=R.RAD_MAP_DTYPE)
    if m0 is not None:
        maps[0, 0] = m0
    if m1 is not None:
        maps[0, 1] = m1
    return maps


def _vec(beam):
    return np.stack([beam[k] for k in P.COORDS], axis=1)


def _beam(bmx, n=2000, seed=3):
    return P.random_beam(bmx, n, seed=seed, scale=5.0, pz_sigma=3e-3)


def _track(oracle, arr, beam, **kw):
    b = P.copy_beam(beam)
    oracle.track_bunch(arr, b, 1, **kw)
    return b


@pytest.mark.parametrize("zero_average", [False, True])
def test_damping_at_the_exit_follows_the_body(bmx, oracle, zero_average):
    """Map on the downstream half only (the reference's bend map B1-rm1, scaled by 1e7 so that
    the kick is ~1e-4 of the coordinates): out = body(v) + rr D (body(v) - ref) [+ rr xv] with rr
    the scale; zero_average leaves out rr xv."""
    arr = _quad(bmx)
    m1 = R.scaled(R.reference_maps()["B1-rm1"], damp=1e7)
    m1["ref_orb"] = [1e-4, -2e-5, 3e-5, 1e-5, 2e-4, -1e-4]
    beam = _beam(bmx)
    base = _track(oracle, arr, beam)
    rr = 0.8
    got = _track(oracle, arr, beam, rad_maps=_maps(m1=m1),
                 rad=_cfg(oracle, zero_average=zero_average, scale=rr))
    v = _vec(base)
    want = v + rr * (v - m1["ref_orb"]) @ m1["damp_dmat"].T
    if not zero_average:
        want = want + rr * m1["xfer_damp_vec"]
    assert np.all(got["state"] == 1)
    kick = np.abs(want - v).max(axis=0)
    assert kick[5] > 1e-6  # the kick is far above rounding
    assert np.allclose(_vec(got), want, rtol=1e-14, atol=1e-19)
    # a transposed damp_dmat would be far off
    wrong = v + rr * (v - m1["ref_orb"]) @ m1["damp_dmat"]
    assert np.abs(_vec(got) - wrong).max() > 1e-9


def test_damping_at_the_entrance_precedes_the_body(bmx, oracle):
    """Map on the upstream half only (K1-rm0 scaled by 1e9): out = body(v + D (v - ref) + xv),
    the kick in laboratory coordinates before the element's offsets and tilt."""
    arr = _quad(bmx)
    m0 = R.scaled(R.reference_maps()["K1-rm0"], damp=1e9)
    beam = _beam(bmx, seed=4)
    got = _track(oracle, arr, beam, rad_maps=_maps(m0=m0), rad=_cfg(oracle))
    v = _vec(beam)
    kicked = v + (v - m0["ref_orb"]) @ m0["damp_dmat"].T + m0["xfer_damp_vec"]
    pre = P.copy_beam(beam)
    for i, k in enumerate(P.COORDS):
        pre[k] = kicked[:, i].copy()
    want = _track(oracle, arr, pre)
    assert np.abs(kicked - v).max() > 1e-7
    assert np.allclose(_vec(got), _vec(want), rtol=1e-13, atol=1e-18)
    # the same kick after the body would differ
    after = _vec(_track(oracle, arr, beam))
    after = after + (after - m0["ref_orb"]) @ m0["damp_dmat"].T + m0["xfer_damp_vec"]
    assert np.abs(_vec(got) - after).max() > 1e-9


@pytest.mark.parametrize("edge", [0, 1])
def test_fluctuation_kick(bmx, oracle, edge):
    """Fluctuations only, on one half: the kick is scale L g with L the lower triangle of
    stoc_mat (B1's, scaled by 1e5) and g the particle's own deviates for (index, turn, element 1,
    edge); over 20000 particles its covariance is scale^2 L L^T."""
    arr = _quad(bmx, tilt=0.0, x_offset=0.0, y_offset=0.0)
    m = R.scaled(R.reference_maps()["B1-rm1"], stoc=1e5)
    L = np.tril(m["stoc_mat"])
    n, seed, scale, turn0 = 20000, 2024, 1.3, 5
    beam = _beam(bmx, n=n, seed=5)
    maps = _maps(m0=m) if edge == 0 else _maps(m1=m)
    got = P.copy_beam(beam)
    oracle.track_bunch(arr, got, 1, turn0=turn0, rad_maps=maps,
                       rad=_cfg(oracle, damping=False, fluct=True, scale=scale, seed=seed))
    g = np.array([oracle.rad_gauss6(seed, i, turn0, 1, edge) for i in range(n)])
    kick = scale * g @ L.T
    if edge == 1:
        base = P.copy_beam(beam)
        oracle.track_bunch(arr, base, 1, turn0=turn0)
        # to the rounding of the terms (the sums may cancel)
        tol = 8 * np.finfo(float).eps * (np.abs(_vec(base)) + scale * np.abs(g) @ np.abs(L).T)
        assert np.all(np.abs(_vec(got) - (_vec(base) + kick)) <= tol)
    else:
        pre = P.copy_beam(beam)
        for i, k in enumerate(P.COORDS):
            pre[k] = beam[k] + kick[:, i]
        oracle.track_bunch(arr, pre, 1, turn0=turn0)
        assert np.allclose(_vec(got), _vec(pre), rtol=1e-13, atol=1e-18)
    cov = np.cov(kick.T)
    want = scale ** 2 * L @ L.T
    d = np.sqrt(np.diag(want))
    # correlation-scale agreement to the statistical error of 20000 samples
    assert np.all(np.abs(cov - want) <= 5 * np.outer(d, d) / np.sqrt(n) + 1e-300)
    assert not np.allclose(want, scale ** 2 * L.T @ L, rtol=0.1)


def test_kinds_without_radiation(bmx, oracle):
    """Drift, pipe, marker, thin multipole, a zero-length kicker and a switched-off quadrupole
    get no kick whatever their maps; a quadrupole beside them does."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    eles = [E.make_element(E.DRIFT, p0c, et, l=1.0),
            E.make_element(E.PIPE, p0c, et, l=0.5),
            E.make_element(E.MARKER, p0c, et),
            E.make_element(E.MULTIPOLE, p0c, et, knl={2: 3.0}),
            E.make_element(E.HKICKER, p0c, et, kick=1e-5),
            E.make_element(E.QUADRUPOLE, p0c, et, l=0.5, k1=0.2, is_on=0)]
    arr = E.to_array(eles)
    m = R.scaled(R.reference_maps()["B1-rm1"], damp=1e8, stoc=1e5)
    maps = np.zeros((len(arr), 2), dtype=R.RAD_MAP_DTYPE)
    maps[:, 0] = m
    maps[:, 1] = m
    beam = _beam(bmx, seed=6)
    base = _track(oracle, arr, beam)
    got = _track(oracle, arr, beam, rad_maps=maps, rad=_cfg(oracle, fluct=True))
    for k in P.KEYS:
        assert np.array_equal(got[k], base[k]), k
    quad = E.to_array(eles + [E.make_element(E.QUADRUPOLE, p0c, et, l=0.5, k1=0.2)])
    maps2 = np.concatenate([maps, maps[:1]])
    got2 = _track(oracle, quad, beam, rad_maps=maps2, rad=_cfg(oracle, fluct=True))
    assert np.abs(got2["pz"] - _track(oracle, quad, beam)["pz"]).max() > 1e-8


@pytest.mark.parametrize("edge", [0, 1])
def test_radiation_losses(bmx, oracle, edge):
    """A constant kick of -3 in pz loses every particle with lost_pz at the element (1-based
    index 2 of a drift-quadrupole-drift line), at either edge; at the upstream edge the particle
    keeps the kicked coordinates and is not tracked through the body. A kick of +2 in px makes
    the orbit too large (px^2 > (1 + pz)^2): lost_pos_x."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    arr = E.to_array([E.make_element(E.DRIFT, p0c, et, l=1.0),
                      E.make_element(E.QUADRUPOLE, p0c, et, l=0.5, k1=0.2),
                      E.make_element(E.DRIFT, p0c, et, l=1.0)])
    beam = _beam(bmx, n=500, seed=7)
    for dv, state in (((5, -3.0), 8), ((1, 2.0), 4)):
        maps = np.zeros((3, 2), dtype=R.RAD_MAP_DTYPE)
        maps[1, edge]["xfer_damp_vec"][dv[0]] = dv[1]
        got = _track(oracle, arr, beam, rad_maps=maps, rad=_cfg(oracle))
        assert np.all(got["state"] == state) and np.all(got["loss_ele"] == 2)
        if edge == 0 and state == 8:
            after_drift = _track(oracle, arr[:1], beam)
            assert np.array_equal(got["x"], after_drift["x"])
            assert np.array_equal(got["pz"], after_drift["pz"] - 3.0)


def test_maps_from_the_library_layout(bmx):
    """The library's numpy mirror of bmx_rad_map has the layout of the C struct (84 doubles:
    ref_orb, damp_dmat, xfer_damp_vec, stoc_mat) and of the fixture helper's."""
    assert bmx.native.RAD_MAP_DTYPE == R.RAD_MAP_DTYPE
    assert R.RAD_MAP_DTYPE.itemsize == 84 * 8
    assert ctypes.sizeof(bmx.native.RadConfig) == 32
