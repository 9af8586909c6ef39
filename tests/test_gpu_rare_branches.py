"""The fast maps' rarely taken branches and orbit_too_large's branches on the GPU, against the
CPU restatement (needs an MI355X).

The benchmark beams stay inside the series domains of the fast maps, so these tests build
particles and elements that leave them, and check in numpy that every branch is reached:
  * quadrupole: t = (k1 L^2 / (1+pz))^2 from below 1/16 (4-term series) through (1/16, 1]
    (5-term series) to 100 (argument quartering, two halvings);
  * drift: (px^2 + py^2)/(1+pz)^2 >= 2^-10 (the exact body instead of the series);
  * sbend: (px/pt)^2 >= 2^-10 (the square root of 1 - s1^2), particles turned past 90 degrees
    (cos(angle + phi_1) <= 0), half exit angles outside the short asin series, and particles
    lost inside the body;
  * orbit_too_large (bmad/low_level/orbit_too_large.f90:46-118): every branch at the first
    element's entrance, and orbits that grow past its limits on the 700-element ring.
Both map forms run; the reference form has none of these branches and must agree as well.
"""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["fast", "reference"])
def ctx(bmx, request):
    c = bmx.tracking.Context(0)
    c.set_maps(request.param)
    return c


def _beam(bmx, n, seed, sig):
    L = bmx.lattices
    beam = L.gaussian_beam(n, seed=seed, sigma=sig)
    return beam


def _track_one(bmx, oracle, ctx, eles, beam, n_turns=1, tol=1e-12):
    arr = bmx.elements.to_array(eles)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=n_turns)
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    bmx.tracking.track_beam(ctx, lat, b, n_turns)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["state"][bad[:10]], ref["state"][bad[:10]])
    alive = ref["state"] == 1
    err = P.rel_err(out, ref, alive)
    assert err <= tol, err
    return out, ref


@pytest.mark.parametrize("k1", [1.0, -1.0, 1.3, -1.3, 5.0, -5.0, 12.0, -12.0, 40.0, -40.0])
def test_strong_quadrupole_series_and_quartering(bmx, oracle, ctx, k1):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    L = 0.5
    ele = E.make_element(E.QUADRUPOLE, p0c, et, l=L, k1=k1)
    beam = _beam(bmx, 4096, 13, [1e-4, 1e-5, 1e-4, 1e-5, 1e-3, 2e-2])
    t = (k1 * L * L / (1 + beam["pz"])) ** 2
    if abs(k1) == 1.0:
        # the boundary of the short series: both sides are taken
        assert np.any(t <= 1 / 16) and np.any((t > 1 / 16) & (t <= 1))
    elif abs(k1) == 1.3:
        assert np.all((t > 1 / 16) & (t <= 1))
    else:
        assert np.all(t > 1)
    print(f"quad k1={k1} t in [{t.min():.3g}, {t.max():.3g}] maps={ctx.maps}", flush=True)
    _track_one(bmx, oracle, ctx, [ele], beam)


def test_drift_and_thick_multipole_exact_fallback(bmx, oracle, ctx):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    eles = [E.make_element(E.DRIFT, p0c, et, l=1.0),
            E.make_element(E.SEXTUPOLE, p0c, et, l=0.2, k2=20.0),
            E.make_element(E.OCTUPOLE, p0c, et, l=0.4, k3=50.0, num_steps=3)]
    beam = _beam(bmx, 4096, 17, [1e-3, 0.06, 1e-3, 0.06, 1e-3, 5e-3])
    pxy2 = (beam["px"] ** 2 + beam["py"] ** 2) / (1 + beam["pz"]) ** 2
    assert np.mean(pxy2 >= 2.0 ** -10) > 0.5 and np.any(pxy2 < 2.0 ** -10)
    assert pxy2.max() > 0.04  # |px| up to about 0.2
    print(f"drift fallback, pxy2 max {pxy2.max():.3g} maps={ctx.maps}", flush=True)
    _track_one(bmx, oracle, ctx, eles, beam)


def test_sbend_far_square_root(bmx, oracle, ctx):
    """The ring's bend (angle 2 pi/200) with |px| up to about 0.2: (px/pt)^2 beyond 2^-10."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    ele = E.make_element(E.SBEND, p0c, et, l=2.0, g=bmx.lattices.BEND_ANGLE / 2)
    beam = _beam(bmx, 4096, 19, [1e-3, 0.07, 1e-3, 0.02, 1e-3, 5e-3])
    pt = np.sqrt((1 + beam["pz"]) ** 2 - beam["py"] ** 2)
    s1s = (beam["px"] / pt) ** 2
    assert np.mean(s1s >= 2.0 ** -10) > 0.5
    _track_one(bmx, oracle, ctx, [ele], beam)


def test_sbend_turned_and_far_exit_angles(bmx, oracle, ctx):
    """A strong bend (1 rad) with px up to 0.95: particles whose direction plus the bend angle
    passes 90 degrees (cos(angle + phi_1) <= 0), exit half-angles beyond the short and the long
    asin series, and particles lost in the body (r < 0 or |x| beyond the limit)."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    angle, L = 1.0, 2.0
    ele = E.make_element(E.SBEND, p0c, et, l=L, g=angle / L)
    n = 4096
    rng = np.random.default_rng(23)
    beam = _beam(bmx, n, 23, [1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3])
    beam["px"] = rng.uniform(-0.95, 0.95, n)
    beam["x"] = rng.uniform(-0.5, 0.5, n)
    pt = np.sqrt((1 + beam["pz"]) ** 2 - beam["py"] ** 2)
    phi1 = np.arcsin(beam["px"] / pt)
    turned = np.cos(angle + phi1) <= 0
    assert turned.sum() > 100
    out, ref = _track_one(bmx, oracle, ctx, [ele], beam, tol=1e-11)
    alive = ref["state"] == 1
    assert (turned & alive).sum() > 50, "turned particles that survive the body"
    assert (~alive).sum() > 50, "particles lost in the body"
    # half exit angle h = (angle + phi_1 - psi): the survivors' exit px = pt sin(angle + phi_1 - 2h)
    h = 0.5 * (angle + phi1[alive] - np.arcsin(np.clip(ref["px"][alive] / pt[alive], -1, 1)))
    assert np.any(np.abs(np.sin(h)) > 1 / 8)


def _otl_particles(bmx):
    """One particle per orbit_too_large branch (orbit_too_large.f90:46-118) plus live ones."""
    rows = [
        (1001.0, 0, 0, 0, 0), (-1001.0, 0, 0, 0, 0), (0, 0, 1001.0, 0, 0), (0, 0, -1001.0, 0, 0),
        (1001.0, 0, -2000.0, 0, 0), (0, 0, 0, 0, -1.1), (0, 3.0, 0, 0, -1.25),
        (0, 1.2, 0, 0.5, 0), (0, -1.2, 0, 0.5, 0), (0, 0.5, 0, 1.2, 0), (0, 0.5, 0, -1.2, 0),
        (0, 0.8, 0, 0.8, 0.1), (0, -0.8, 0, -0.8, 0.1), (0, 0.3, 0, 0.2, -0.7),
    ]
    p0c, et = bmx.lattices.reference()
    n = len(rows)
    a = np.array(rows, dtype=np.float64)
    beam = {"x": a[:, 0].copy(), "px": a[:, 1].copy(), "y": a[:, 2].copy(), "py": a[:, 3].copy(),
            "z": np.zeros(n), "pz": a[:, 4].copy()}
    beam["beta"] = np.full(n, 0.99)
    beam["state"] = np.ones(n, dtype=np.int32)
    beam["loss_ele"] = np.zeros(n, dtype=np.int32)
    beam["loss_turn"] = np.full(n, -1, dtype=np.int32)
    return beam


def test_orbit_too_large_at_entrance(bmx, oracle, ctx):
    """Every branch at the entrance of the first element of the 700-element ring (which has no
    apertures: orbit_too_large is its only loss): state codes and loss element 1 as the
    restatement, coordinates untouched."""
    beam = _otl_particles(bmx)
    arr = bmx.elements.to_array(bmx.lattices.ring0(n_cells=2))
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, n_turns=1)
    assert np.all(ref["state"] != 1)
    assert sorted(set(ref["state"].tolist())) == [3, 4, 5, 6, 8]
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    bmx.tracking.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    assert P.loss_mismatch(out, ref).size == 0
    assert np.all(out["loss_ele"] == 1)
    for k in P.COORDS:
        assert np.array_equal(out[k], beam[k]), k


def test_orbit_too_large_on_ring0(bmx, oracle, ctx):
    """Orbits that grow past orbit_too_large's limits inside the 700-element ring (no
    apertures): large |px|, |py| send x or y beyond 1 km or the transverse momentum beyond
    1 + pz after some elements. Loss state, element and turn identical to the restatement;
    survivors within 100x of their envelope (the restatement's own response to FMA contraction
    and to 1e-14 rescalings of its input) or 1e-10: at transverse angles of 0.3 rad the drift's
    1/ps and the bend's arcsine amplify rounding, so the restatement's two builds already
    differ by up to ~1e-8 here."""
    n = 4096
    beam = _beam(bmx, n, 29, [1e-3, 0.3, 1e-3, 0.3, 1e-3, 0.05])
    arr = bmx.elements.to_array(bmx.lattices.ring0(n_cells=20))
    ref, env = P.envelope(oracle, arr, beam, 3)
    lost = ref["state"] != 1
    codes = set(ref["state"][lost].tolist())
    assert {3, 4, 5, 6} <= codes, codes
    assert np.any(ref["loss_ele"][lost] > 1), "losses inside the ring, not only at entry"
    lat = bmx.tracking.Lattice(ctx, arr)
    b = bmx.tracking.Beam.from_host(ctx, beam)
    bmx.tracking.track_beam(ctx, lat, b, 3)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, (bad[:10], out["loss_ele"][bad[:10]], ref["loss_ele"][bad[:10]])
    far, dev = P.outside_envelope(out, ref, env)
    assert far.size == 0, (far[:10], dev[far[:10]], env[far[:10]])
