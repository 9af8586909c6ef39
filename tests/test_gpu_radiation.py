"""Radiation damping and fluctuation kicks on the GPU (track1_radiation, radiation_mod.f90:61-147,
at both ends of every radiating element, track1.f90:197-202 and 319-324), against the CPU
restatement with the same maps, settings and fluctuation stream (needs an MI355X):

* single elements of every radiating kind -- offset and tilted quadrupole, sextupole, a plain
  bend and one with edge angles and fringe integrals (both taken to the general bend), rf cavity,
  kicker, collimator with an aperture -- with damping and fluctuations: losses identical and
  survivors within 1e-12 (the per-element bar), including a case whose fluctuations take pz
  below -1 for some particles at either edge;
* the 10-cell ring with maps on every element for 50 turns on the generic kernel, the
  cell-specialised kernel, the element-at-a-time kernel and a batch of three lattices: losses
  identical, survivors within the restatement's own sensitivity envelope;
* served requests on radiating lattices, bit-identical to tracking each request alone;
* the 6-D closed orbit with damping (fluctuations switched off, as closed_orbit_calc does)
  against the restatement's Newton solution;
* a lattice's settings changed without recompiling it (zero_average, fluctuations off, another
  seed) and the kicks switched off entirely, which leaves the tracking of the lattice without
  maps (to the rounding of the general bend's formulas).
The fluctuation deviates come from log, sin and cos on the device and in the C library, which
may differ in the last bit; the kicks they scale are small beside the coordinates, so that
difference is far inside the per-element bar."""
import numpy as np
import pytest

import parity as P
import rad_fixtures as R

pytestmark = pytest.mark.gpu


def _cfg(bmx, **kw):
    return bmx.tracking.rad_config(**kw)


def _ocfg(oracle, c):
    return oracle.RadConfig(c.damping_on, c.fluctuations_on, c.zero_average, 0, c.synch_rad_scale,
                            c.seed)


def _single(bmx, name):
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    mk = E.make_element
    return {
        "quad_offset": mk(E.QUADRUPOLE, p0c, et, l=0.6, k1=0.4, tilt=0.3, x_offset=2e-4,
                          y_offset=-1e-4),
        "sextupole": mk(E.SEXTUPOLE, p0c, et, l=0.3, k2=20.0),
        "bend_plain": mk(E.SBEND, p0c, et, l=2.0, g=0.005),
        "bend_edges": mk(E.SBEND, p0c, et, l=1.0, g=0.01, e1=0.1, e2=0.2, fint=0.5, hgap=0.02,
                         fintx=0.5, hgapx=0.02),
        "rfcavity": mk(E.RFCAVITY, p0c, et, l=0.5, voltage=2e6, rf_frequency=5e8, phi0=0.1),
        "kicker": mk(E.HKICKER, p0c, et, l=0.2, kick=1e-5),
        "collimator": mk(E.RCOLLIMATOR, p0c, et, l=0.5, x1_limit=3e-3, x2_limit=3e-3,
                         y1_limit=2e-3, y2_limit=2e-3),
    }[name]


NAMES = ["quad_offset", "sextupole", "bend_plain", "bend_edges", "rfcavity", "kicker", "collimator"]


def _run_gpu(bmx, arr, beam, maps, cfg, n_turns, maps_form="fast", specialize=False, turn0=0):
    T = bmx.tracking
    ctx = T.Context(0)
    ctx.set_maps(maps_form)
    lat = T.Lattice(ctx, arr, rad_maps=maps, rad=cfg)
    if specialize:
        assert lat.specialize()
    b = T.Beam.from_host(ctx, beam)
    T.track_beam(ctx, lat, b, n_turns, turn0)
    ctx.synchronize()
    out = b.download()
    b.close()
    lat.close()
    ctx.close()
    return out


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("maps_form", ["fast", "reference"])
def test_single_element(bmx, oracle, name, maps_form):
    """One element with B1's (bends) or K1's (others) maps scaled to visible kicks, damping and
    fluctuations, 4096 particles: losses identical, survivors within 1e-12 of the restatement."""
    arr = bmx.elements.to_array([_single(bmx, name)])
    maps = R.lattice_maps(arr, bmx, damp=1e7, stoc=1e5)
    cfg = _cfg(bmx, damping=True, fluctuations=True, scale=1.0, seed=77)
    beam = P.random_beam(bmx, 4096, seed=21, scale=10.0, pz_sigma=3e-3)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, 1, turn0=3, rad_maps=maps, rad=_ocfg(oracle, cfg))
    norad = P.copy_beam(beam)
    oracle.track_bunch(arr, norad, 1, turn0=3)
    alive = (ref["state"] == 1) & (norad["state"] == 1)
    assert np.abs(ref["pz"] - norad["pz"])[alive].max() > 1e-7  # the kicks show
    print(f"radiation single {name} maps={maps_form}", flush=True)
    out = _run_gpu(bmx, arr, beam, maps, cfg, 1, maps_form, turn0=3)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    alive = ref["state"] == 1
    assert alive.sum() > 3000
    assert P.rel_err(out, ref, alive) <= 1e-12


@pytest.mark.parametrize("edge", [0, 1])
def test_fluctuation_losses(bmx, oracle, edge):
    """A quadrupole whose fluctuation kick has a pz entry of 0.4 on one half: the particles whose
    deviate falls below -2.5 or so end below pz = -1 and are lost (lost_pz, or orbit_too_large's
    state at the exit) at the element, the same particles as in the restatement."""
    E = bmx.elements
    p0c, et = bmx.lattices.reference()
    arr = E.to_array([E.make_element(E.DRIFT, p0c, et, l=1.0),
                      E.make_element(E.QUADRUPOLE, p0c, et, l=0.5, k1=0.2),
                      E.make_element(E.DRIFT, p0c, et, l=1.0)])
    maps = np.zeros((3, 2), dtype=R.RAD_MAP_DTYPE)
    maps[1, edge]["stoc_mat"][5, 5] = 0.4
    maps[1, edge]["stoc_mat"][5, 0] = 1e-3
    cfg = _cfg(bmx, damping=False, fluctuations=True, seed=5)
    beam = P.random_beam(bmx, 8192, seed=22, scale=3.0)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, 1, rad_maps=maps, rad=_ocfg(oracle, cfg))
    lost = ref["state"] != 1
    assert 10 < lost.sum() < 200 and np.all(ref["loss_ele"][lost] == 2)
    print(f"radiation losses edge={edge}", flush=True)
    out = _run_gpu(bmx, arr, beam, maps, cfg, 1)
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    assert P.rel_err(out, ref, ref["state"] == 1) <= 1e-12


def _ring(bmx, n_cells=10):
    return bmx.elements.to_array(bmx.lattices.ring1(n_cells))


RING_CFG = dict(damping=True, fluctuations=True, scale=1.0, seed=4242)


def _ring_ref(bmx, oracle, n_turns=50):
    arr = _ring(bmx)
    maps = R.lattice_maps(arr, bmx, damp=1e4, stoc=1e3)
    beam = P.random_beam(bmx, 2000, seed=23, scale=15.0, pz_sigma=3e-3)
    cfg = _cfg(bmx, **RING_CFG)
    ref, env = P.cached(("radiation_ring", n_turns),
                        lambda: P.envelope(oracle, arr, beam, n_turns, rad_maps=maps,
                                           rad=_ocfg(oracle, cfg)))
    return arr, maps, beam, cfg, ref, env


def _check_ring(out, ref, env, floor):
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    idx, dev = P.outside_envelope(out, ref, env, floor=floor)
    alive = ref["state"] == 1
    print("deviation max", float(dev[alive].max()), "median", float(np.median(dev[alive])), flush=True)
    assert idx.size == 0, (idx[:10], dev[idx[:10]], env[idx[:10]])


@pytest.mark.parametrize("maps_form,specialize", [("reference", False), ("fast", False), ("fast", True)])
def test_ring_with_radiation(bmx, oracle, maps_form, specialize):
    """The 10-cell ring with maps on every element (bends B1's, the others K1's, damping x1e4,
    fluctuations x1e3), 2000 particles for 50 turns: losses identical, survivors within
    max(1e-10 reference maps / 1e-9 fast maps, 100x the envelope)."""
    arr, maps, beam, cfg, ref, env = _ring_ref(bmx, oracle)
    assert (ref["state"] != 1).sum() > 20
    print(f"radiation ring maps={maps_form} specialize={specialize}", flush=True)
    out = _run_gpu(bmx, arr, beam, maps, cfg, 50, maps_form, specialize)
    _check_ring(out, ref, env, 1e-10 if maps_form == "reference" else 1e-9)


def test_ring_on_the_plain_path(bmx, oracle):
    """The element-at-a-time kernel carries the kicks too (one turn, per-element bar)."""
    T = bmx.tracking
    arr = _ring(bmx)
    maps = R.lattice_maps(arr, bmx, damp=1e4, stoc=1e3)
    cfg = _cfg(bmx, **RING_CFG)
    beam = P.random_beam(bmx, 2000, seed=24, scale=5.0)
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, 1, turn0=7, rad_maps=maps, rad=_ocfg(oracle, cfg))
    ctx = T.Context(0)
    lat = T.Lattice(ctx, arr, rad_maps=maps, rad=cfg)
    b = T.Beam.from_host(ctx, beam)
    T.track_beam_plain(ctx, lat, b, 1, 7)
    ctx.synchronize()
    out = b.download()
    bad = P.loss_mismatch(out, ref)
    assert bad.size == 0, bad[:10]
    assert P.rel_err(out, ref, ref["state"] == 1) <= 1e-11


def test_batched_lattices_with_radiation(bmx, oracle):
    """Three ring variants with different maps and seeds in one batch, each group its own beam:
    each within the envelope of its own restatement run (a group's particle indices count from
    its first particle)."""
    T = bmx.tracking
    ctx = T.Context(0)
    arrs, mapss, cfgs, beams, refs, envs = [], [], [], [], [], []
    for v in range(3):
        arr = _ring(bmx)
        maps = R.lattice_maps(arr, bmx, damp=1e4 * (1 + v), stoc=1e3 * (1 + v))
        cfg = _cfg(bmx, damping=True, fluctuations=True, seed=100 + v)
        beam = P.random_beam(bmx, 1000, seed=30 + v, scale=10.0, pz_sigma=2e-3)
        ref, env = P.envelope(oracle, arr, beam, 20, rad_maps=maps, rad=_ocfg(oracle, cfg))
        arrs.append(arr); mapss.append(maps); cfgs.append(cfg); beams.append(beam)
        refs.append(ref); envs.append(env)
    lats = [T.Lattice(ctx, a, rad_maps=m, rad=c) for a, m, c in zip(arrs, mapss, cfgs)]
    bs = [T.Beam.from_host(ctx, b) for b in beams]
    batch = T.Batch(ctx, [(lats[v], bs[v], 0, 1000) for v in range(3)])
    print("radiation batch", flush=True)
    batch.track(20)
    ctx.synchronize()
    for v in range(3):
        _check_ring(bs[v].download(), refs[v], envs[v], 1e-9)


def test_settings_without_recompiling(bmx, oracle):
    """set_radiation: zero_average on, then fluctuations only with another seed, each equal to a
    restatement run with those settings; then both kicks off, which tracks like the lattice
    without maps (its bends on the general bend's formulas: within 1e-11 after one turn of its
    91 elements, as the element-at-a-time kernel's one-turn bar)."""
    T = bmx.tracking
    arr = _ring(bmx)
    maps = R.lattice_maps(arr, bmx, damp=1e5, stoc=1e3)
    beam = P.random_beam(bmx, 2000, seed=25, scale=5.0)
    ctx = T.Context(0)
    lat = T.Lattice(ctx, arr, rad_maps=maps, rad=_cfg(bmx, **RING_CFG))
    for cfg in (_cfg(bmx, damping=True, fluctuations=True, zero_average=True, seed=4242),
                _cfg(bmx, damping=False, fluctuations=True, seed=99, scale=0.5)):
        lat.set_radiation(cfg)
        ref = P.copy_beam(beam)
        oracle.track_bunch(arr, ref, 1, rad_maps=maps, rad=_ocfg(oracle, cfg))
        b = T.Beam.from_host(ctx, beam)
        print("radiation settings", cfg.damping_on, cfg.fluctuations_on, cfg.zero_average, flush=True)
        T.track_beam(ctx, lat, b, 1)
        ctx.synchronize()
        out = b.download()
        assert P.loss_mismatch(out, ref).size == 0
        assert P.rel_err(out, ref, ref["state"] == 1) <= 1e-11
    lat.set_radiation(_cfg(bmx, damping=False, fluctuations=False))
    ref = P.copy_beam(beam)
    oracle.track_bunch(arr, ref, 1)
    b = T.Beam.from_host(ctx, beam)
    T.track_beam(ctx, lat, b, 1)
    ctx.synchronize()
    out = b.download()
    assert P.loss_mismatch(out, ref).size == 0
    assert P.rel_err(out, ref, ref["state"] == 1) <= 1e-11


def _oracle_closed_orbit6(oracle, arr, maps, cfg, steps=(1e-6, 1e-6, 1e-6, 1e-6, 1e-5, 1e-6)):
    """The restatement's 6-D closed orbit: Newton steps on all six coordinates with the Jacobian
    from five-point central differences of 25 tracked probes, as bmx_closed_orbit makes them."""
    p0c, et = float(arr[0]["p0c"]), float(arr[0]["e_tot"])
    x = np.zeros(6)
    for _ in range(50):
        v = np.tile(x, (25, 1))
        for d in range(6):
            for j, f in enumerate((1, -1, 2, -2)):
                v[1 + 4 * d + j, d] += f * steps[d]
        beam = {k: v[:, i].copy() for i, k in enumerate(P.COORDS)}
        pc = p0c * (1 + beam["pz"])
        beam["beta"] = np.where(beam["pz"] != 0, pc / np.sqrt(pc ** 2 + 0.51099895069e6 ** 2), p0c / et)
        beam["state"] = np.ones(25, dtype=np.int32)
        beam["loss_ele"] = np.zeros(25, dtype=np.int32)
        beam["loss_turn"] = np.zeros(25, dtype=np.int32)
        oracle.track_bunch(arr, beam, 1, rad_maps=maps, rad=cfg)
        assert np.all(beam["state"] == 1)
        out = np.stack([beam[k] for k in P.COORDS], axis=1)
        M = np.stack([(8 * (out[1 + 4 * d] - out[2 + 4 * d]) - (out[3 + 4 * d] - out[4 + 4 * d]))
                      / (12 * steps[d]) for d in range(6)], axis=1)
        step = np.linalg.solve(M - np.eye(6), -(out[0] - x))
        x = x + step
        if np.all(np.abs(step) < np.abs(x) * 1e-12 + 1e-15):
            return x
    raise AssertionError("restatement closed orbit did not converge")


def test_closed_orbit_with_damping(bmx, oracle):
    """The 6-D closed orbit of the 10-cell ring whose bends radiate (B1's maps scaled by 1e3,
    damping on, fluctuations on in the settings): closed_orbit_calc keeps the damping and switches
    the fluctuations off (closed_orbit_calc.f90:264), so the orbit is the damped, energy-losing
    one -- pz and z shifted from the orbit without radiation, where the rf restores the energy
    lost -- equal to the restatement's with fluctuations off within 1e-11."""
    T = bmx.tracking
    arr = _ring(bmx)
    maps = R.lattice_maps(arr, bmx, damp=1e3, stoc=1e3)
    maps[arr["key"] != bmx.elements.SBEND] = np.zeros((), dtype=R.RAD_MAP_DTYPE)
    ctx = T.Context(0)
    cfg = _cfg(bmx, damping=True, fluctuations=True, seed=11)
    lat = T.Lattice(ctx, arr, rad_maps=maps, rad=cfg)
    plain = T.Lattice(ctx, arr)
    print("closed orbit with damping", flush=True)
    orbit, mat, st = T.closed_orbit(ctx, [lat, plain], n_dim=6)
    assert np.all(st == 0), st
    want = _oracle_closed_orbit6(oracle, arr, maps, _ocfg(oracle, _cfg(bmx, damping=True, fluctuations=False)))
    print("orbit", orbit[0].tolist(), "restatement", want.tolist(), flush=True)
    assert np.abs(orbit[0] - orbit[1]).max() > 1e-7  # the damping moves the orbit
    assert np.allclose(orbit[0], want, rtol=0, atol=1e-11)


def test_served_requests_with_radiation(bmx):
    """Requests on three radiating lattices (their own maps and seeds) served together with
    37-turn chunks: every particle equals tracking its request alone, bit for bit -- a request's
    fluctuation kicks follow its own particles' indices, wherever the server places them."""
    T = bmx.tracking
    ctx = T.Context(0)
    arr = _ring(bmx)
    mapss = [R.lattice_maps(arr, bmx, damp=1e4 * (1 + v), stoc=1e3) for v in range(3)]
    cfgs = [_cfg(bmx, damping=True, fluctuations=True, seed=500 + v) for v in range(3)]
    lats = [T.Lattice(ctx, arr, rad_maps=m, rad=c) for m, c in zip(mapss, cfgs)]
    reqs = [(i % 3, 400 + 300 * i, 10 + 13 * i) for i in range(6)]
    beams = [P.random_beam(bmx, n, seed=70 + i, scale=12.0, pz_sigma=3e-3)
             for i, (_, n, _) in enumerate(reqs)]
    srv = T.Server(ctx, hbm_budget=1 << 30, chunk_turns=37, compact_turns=10)
    print("served requests with radiation", flush=True)
    ids = [srv.submit(lats[v], beams[i], t, 3) for i, (v, n, t) in enumerate(reqs)]
    results = [srv.wait(r) for r in ids]
    srv.close()
    for i, (v, n, t) in enumerate(reqs):
        out, _ = results[i]
        alone = T.track_beam_host(arr, beams[i], t, turn0=3, rad_maps=mapss[v], rad=cfgs[v])
        for k in alone:
            assert np.array_equal(out[k], alone[k]), (i, k)
