"""Throughput of bmad_standard FP64 multi-particle tracking on one MI355X.

Workload (default): the 901-element ring with sextupoles, one rf cavity and elliptical
apertures on every element; 10^6 electrons at 5 GeV from a seeded 6-D Gaussian with 3x the
base sigmas; 1000 turns. The 1000 turns are split evenly over the --steps steps (one step =
1000/steps turns of the whole beam in one bmx_track_turns call, state resident in HBM between
steps), so every run is the same 1000-turn workload whatever the step count.
Throughput counts particle-element passes of live particles only, as the reference's loop
skips dead particles (bmad/multiparticle/beam_utils.f90:83).

Prints one JSON line: value = particle-element-turns/s, the roofline of the tracking kernel
(FP64 operations counted with the algorithmic convention of flop_counts.json), and the CPU
restatement's OpenMP throughput on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

METRIC = "particle·element·turns/s FP64 + % FP64 peak; p50/p99 latency served vs plain"
UNIT = "particle·element·turns/s"
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--turns", type=int, default=1000,
                    help="turns of the timed run, split evenly over the steps")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time of the cpu_baseline sample (0 = skip)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--maps", choices=["fast", "reference"], default="fast",
                    help="formulation of the element maps")
    ap.add_argument("--chunk-turns", type=int, default=5,
                    help="turns between alive-list compactions (0 = none)")
    ap.add_argument("--extras", default="batched,da,served,plain",
                    help="comma list of the other workloads to run after the main one "
                         "(batched, da, served, plain, radiation; empty = none)")
    ap.add_argument("--main", choices=["on", "off"], default="on",
                    help="off: skip the single-lattice workload and the CPU baseline (runs of "
                         "the other workloads alone)")
    ap.add_argument("--served-requests", type=int, default=200,
                    help="requests per served stream")
    ap.add_argument("--served-streams", default="full,reduced",
                    help="comma list of served streams: full (particles log-uniform [1e3, 1e6], "
                         "turns [10, 1e4]) and reduced (particles [1e3, 1e5], turns [10, 300]: a "
                         "stream the baselines can track in full within minutes)")
    ap.add_argument("--served-full-baselines", default="none",
                    help="baselines tracked over the full stream's arrivals (comma list of "
                         "one_job, plain; none)")
    ap.add_argument("--served-reduced-baselines", default="one_job,plain",
                    help="baselines tracked over the reduced stream's arrivals: one_job "
                         "(persistent kernel, one request at a time), plain (one launch per "
                         "element per turn), each request tracked in full")
    ap.add_argument("--plain-graph", choices=["on", "off"], default="on",
                    help="plain leg: on = graph-replayed turns (bmx_track_turns_plain); off = "
                         "one host launch per element (bmx_track_element), for kernel traces")
    ap.add_argument("--waves", type=int, default=0,
                    help="waves per SIMD of the tracking kernel variant (3, 4 or 5; 0 = the "
                         "library's default for the kernel in use)")
    ap.add_argument("--specialize", choices=["on", "off"], default="on",
                    help="run-time compiled kernels specialised to the ring's repeated cell "
                         "(fast maps only; off = the generic table-driven kernel)")
    return ap.parse_args()


def traffic_file(special: bool) -> str:
    name = "bmx_cell_tiles" if special else "k_track_tiles"
    return os.path.join("profiles", f"{name}_ring1_1e6_traffic.md")


def profiled_traffic(special: bool):
    """HBM bytes per tracking-kernel launch, averaged over the timed launches of the default
    run, from the committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE runs scaled by a
    copy-kernel calibration), or None if it is absent."""
    import re
    path = os.path.join(ROOT, traffic_file(special))
    try:
        m = re.search(r"\*\*([0-9.]+) MB per launch\*\*", open(path).read())
    except OSError:
        return None
    return float(m.group(1)) * 1e6 if m else None


def host_topology() -> dict:
    """Sockets, cores per socket and threads per core of the host (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "Model name": "model", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    return info


def cpu_baseline(bmx, arr, beam, target_s):
    """The CPU restatement (OpenMP over particles per element, as track1_bunch_hom,
    beam_utils.f90:81-86) on a bounded sample of the same workload: a particle subset over a
    prefix of the turns, threads spread over the host cores. Two builds of the same source: -O2
    (the reference's optional optimisation level) and -O3 -march=native (compiled here, for
    this host); value is the -O2 leg."""
    import tempfile
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = 20000
    sample = {k: np.ascontiguousarray(v[:n]) for k, v in beam.items()}
    native_path = os.path.join(tempfile.mkdtemp(prefix="bmx_cpu_"), "libbmad_oracle_native.so")
    try:
        oracle.build_native(native_path)
    except Exception as e:  # reported; the -O2 leg still runs
        native_path, native_err = None, str(e)
    else:
        native_err = None

    def leg(lib_path):
        # calibrate with one turn, then run enough turns for ~target_s
        t0 = time.perf_counter()
        done = oracle.track_bunch(arr, {k: v.copy() for k, v in sample.items()}, 1,
                                  n_threads=threads, lib_path=lib_path)
        rate1 = done / max(time.perf_counter() - t0, 1e-9)
        turns = max(1, int(target_s * rate1 / (n * len(arr))))
        b = {k: v.copy() for k, v in sample.items()}
        t0 = time.perf_counter()
        done = oracle.track_bunch(arr, b, turns, n_threads=threads, lib_path=lib_path)
        dt = time.perf_counter() - t0
        return done / dt, turns, dt, float(np.mean(b["state"] == 1))

    v2, turns, dt, live = leg(None)
    topo = host_topology()
    out = {"value": v2, "unit": UNIT, "cores": threads, "kind": "port",
           "sample": f"{n} of the {beam['x'].shape[0]} particles x {len(arr)} elements x the "
                     f"first {turns} turns ({dt:.1f} s; {live:.3f} of the sample alive at its end), "
                     f"C restatement -O2 -fopenmp -ffp-contract=off",
           "threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "host": topo}
    if native_path:
        v3, turns3, dt3, _ = leg(native_path)
        out["o3_native"] = {"value": v3, "turns": turns3, "s": dt3,
                            "build": "-O3 -march=native -fopenmp -ffp-contract=off"}
    else:
        out["o3_native"] = {"error": native_err}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: tracking here runs on a single MI355X and has no "
              "multi-GPU path (particles are independent; a beam would shard by particle "
              "ranges, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    bmx = load_package()
    if args.main == "off":
        extras_only(bmx, args)
        return
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    eles = L.ring1()
    arr = E.to_array(eles)
    beam = L.gaussian_beam(args.particles, seed=args.seed, sigma=3 * L.SIGMA0)
    flops = bmx.flops.per_element_pass(arr)

    ctx = T.Context(0)
    ctx.set_maps(args.maps)
    ctx.set_run(args.chunk_turns, 0, args.waves)
    lat = T.Lattice(ctx, arr)
    t0 = time.perf_counter()
    special, special_err = False, None
    if args.specialize == "on" and args.maps == "fast":
        try:
            special = lat.specialize()
        except bmx.native.BmxError as e:  # reported in the JSON; the generic kernel runs
            special_err = str(e)
    t_special = time.perf_counter() - t0
    def kernel_of(w):
        return f"bmx_cell_tiles_{w}" if special else f"k_track_tiles<{str(args.maps == 'fast').lower()}, {w}>"
    # the turns of each step: the run's turns split evenly (the first steps take the remainder)
    step_turns = [args.turns // args.steps + (1 if s < args.turns % args.steps else 0)
                  for s in range(args.steps)]
    step_turn0 = [int(x) for x in np.concatenate([[0], np.cumsum(step_turns)[:-1]])]
    if args.warmup > 0:
        wb = T.Beam.from_host(ctx, beam)
        for w in range(args.warmup):
            T.track_beam(ctx, lat, wb, step_turns[0], w * step_turns[0])
        ctx.synchronize()
        wb.close()
    b = T.Beam.from_host(ctx, beam)
    ctx.take_work()
    ctx.profile(True)
    ctx.kernel_time()
    ctx.synchronize()
    t_wall0 = time.perf_counter()
    ctx.record(0)
    shapes = []
    for s in range(args.steps):
        ctx.record(1 + 2 * s)
        T.track_beam(ctx, lat, b, step_turns[s], step_turn0[s])
        ctx.record(2 + 2 * s)
        shapes.append(ctx.last_run())  # the run shape this step used (host-side record)
    ctx.record(63)
    ctx.synchronize()
    t_wall = time.perf_counter() - t_wall0
    total_ms = ctx.elapsed_ms(0, 63)
    launch_ms = [ctx.elapsed_ms(1 + 2 * s, 2 + 2 * s) for s in range(args.steps)]
    passes = ctx.take_work()
    kernel_ms, launches = ctx.kernel_time()
    ctx.profile(False)
    alive = b.count_alive()

    value = passes / (total_ms * 1e-3)
    # the kernel of each step: with waves chosen automatically the library picks the 5-wave
    # build while the live particles fill more than 1.45 rounds of the 4-wave build, then the
    # 4-wave one (the run phase)
    step_waves = [int(w) for _, _, w in shapes]
    kernels = {}
    for w in step_waves:
        kernels[kernel_of(w)] = kernels.get(kernel_of(w), 0) + 1
    # roofline of the tracking kernel: algorithmic FLOPs per launch / average launch duration
    kernel_s = kernel_ms * 1e-3 / max(launches, 1)
    achieved = flops * (passes / max(launches, 1)) / kernel_s / 1e12
    # algorithmic HBM bytes per launch: 60 B read per particle tracked (seven coordinates and
    # its index in the packed buffer), 60 B written per particle (to the other packed buffer or
    # the beam); particles tracked per launch from the passes per launch over the chunk's
    # element passes per particle
    chunk = args.chunk_turns if args.chunk_turns > 0 else step_turns[0]
    bytes_alg = 120.0 * (passes / max(launches, 1)) / (len(arr) * chunk)
    # the profiled traffic is that of the default workload only
    default_run = (args.particles == 1_000_000 and args.turns == 1000 and args.steps == 10
                   and args.chunk_turns == 5 and args.waves == 0 and args.maps == "fast"
                   and args.seed == 1)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"ring of 100 FODO cells + sextupoles + rf cavity, {len(arr)} elements, "
                        f"elliptical apertures; {args.particles} e- at 5 GeV, 3x sigmas; "
                        f"{args.turns} turns in {args.steps} steps",
            "particles": args.particles, "elements": len(arr), "turns": args.turns,
            "turns_per_step": step_turns[0], "alive_at_end": alive, "wall_s": t_wall,
            "maps": args.maps, "cell_specialized": bool(special),
            "waves_per_simd": step_waves if len(set(step_waves)) > 1 else step_waves[0],
            "specialize_s": t_special, "specialize_error": special_err,
        },
        "roofline": {
            "bound": "valu", "compute": "fp64 vector ALU (no MFMA: the maps are not contractions)",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": profiled_traffic(special) if default_run else None,
            "traffic_note": "HBM bytes per launch averaged over the 200 timed launches of the "
                            "default run, " + traffic_file(special),
            "traffic_algorithmic": bytes_alg,
            "flop_per_particle_element": flops,
            "kernel": " then ".join(kernels), "kernel_steps": kernels,
            "kernel_ms_avg": kernel_ms / max(launches, 1),
            "launches": launches, "step_ms": [float(x) for x in launch_ms],
            "step_run_shape": [{"chunk_turns": int(c), "grid": int(g), "waves": int(w)}
                               for c, g, w in shapes],
        },
    }
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(bmx, arr, beam, args.cpu_seconds)
    out["extras"] = run_extras(bmx, ctx, args, flops, bool(special))
    print(json.dumps(out))


def bend_rad_maps(bmx, arr):
    """Synthetic radiation maps of the magnitudes Bmad gives the ring's bends at 5 GeV (each half
    of a bend: the relative energy loss U = 2/3 r_e gamma^3 g^2 L/2 as xfer_damp_vec's pz, damping
    of px, py by U and of pz by 2U, a quantum-excitation kick in pz of variance
    55/(24 sqrt 3) r_e lambda_e gamma^5 |g|^3 L/2); the other elements carry none."""
    E, T = bmx.elements, bmx.tracking
    r_e, lam_e = 2.8179403262e-15, 3.8615926796e-13  # classical radius, reduced Compton wavelength
    maps = T.rad_maps(len(arr))
    for i in np.nonzero(arr["key"] == E.SBEND)[0]:
        g, half = float(arr[i]["g"]), float(arr[i]["l"]) / 2
        gamma = float(arr[i]["e_tot"]) / E.M_ELECTRON
        u = 2.0 / 3.0 * r_e * gamma ** 3 * g * g * half
        sig = math.sqrt(55.0 / (24.0 * math.sqrt(3.0)) * r_e * lam_e * gamma ** 5 * abs(g) ** 3 * half)
        for h in range(2):
            m = maps[i, h]
            m["damp_dmat"][1, 1] = m["damp_dmat"][3, 3] = -u
            m["damp_dmat"][5, 5] = -2 * u
            m["xfer_damp_vec"][5] = -u
            m["stoc_mat"][5, 5] = sig
    return maps


def run_radiation(bmx, ctx, flops, specialize, n=1_000_000, turns=100):
    """The main workload's ring with radiating bends (damping and fluctuation kicks at both ends
    of each, bend_rad_maps), 1e6 particles over 100 turns: particle-element-turns per second by
    the wall clock around the run, and the same ring without radiation."""
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    arr = E.to_array(L.ring1())
    beam = L.gaussian_beam(n, seed=3, sigma=3 * L.SIGMA0)
    maps = bend_rad_maps(bmx, arr)
    out = {"particles": n, "turns": turns, "radiating": int((arr["key"] == E.SBEND).sum())}
    for label, kw in (("none", {}),
                      ("damping", dict(rad_maps=maps, rad=T.rad_config(damping=True))),
                      ("damping_fluctuations", dict(rad_maps=maps, rad=T.rad_config(damping=True, fluctuations=True, seed=5)))):
        lat = T.Lattice(ctx, arr, **kw)
        if specialize:
            lat.specialize()
        b = T.Beam.from_host(ctx, beam)
        ctx.synchronize()
        ctx.take_work()
        t0 = time.perf_counter()
        T.track_beam(ctx, lat, b, turns)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        passes = ctx.take_work()
        out[label] = {"value": passes / dt, "seconds": dt, "alive": int(b.count_alive())}
        print(f"radiation {label}: {passes / dt:.3e} particle-element-turns/s", file=sys.stderr, flush=True)
        b.close()
        lat.close()
    out["unit"] = "particle·element·turns/s"
    return out


def run_extras(bmx, ctx, args, flops, special):
    extras = [e for e in args.extras.split(",") if e and e != "none"]
    out = {}
    if "batched" in extras:
        out["batched"] = run_batched(bmx, ctx, flops, specialize=special)
    if "da" in extras:
        out["dynamic_aperture"] = run_da(bmx, ctx, specialize=special)
    if "served" in extras:
        streams = [x for x in args.served_streams.split(",") if x]
        for stream, bl, key in (("full", args.served_full_baselines, "served"),
                                ("reduced", args.served_reduced_baselines, "served_reduced")):
            if stream in streams:
                baselines = [b for b in bl.split(",") if b and b != "none"]
                out[key] = run_served(bmx, ctx, args.served_requests, specialize=special,
                                      stream=stream, baselines=baselines)
    if "radiation" in extras:
        out["radiation"] = run_radiation(bmx, ctx, flops, specialize=special)
    if "plain" in extras:
        out["plain"] = run_plain(bmx, ctx, graph=args.plain_graph == "on")
    return out


def extras_only(bmx, args):
    """The other workloads alone (--main off): one JSON line with their results."""
    T = bmx.tracking
    ctx = T.Context(0)
    ctx.set_maps(args.maps)
    flops = bmx.flops.per_element_pass(bmx.elements.to_array(bmx.lattices.ring1()))
    special = args.specialize == "on" and args.maps == "fast"
    out = {"metric": METRIC, "unit": UNIT, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "extras only: " + args.extras, "maps": args.maps},
           "extras": run_extras(bmx, ctx, args, flops, special)}
    print(json.dumps(out))


# HBM bytes of one element pass of a live particle on the plain path: 7 FP64 coordinates and
# the state read, 7 coordinates written (the loss record only on the rare lost lanes)
PLAIN_BYTES_PER_PASS = 7 * 8 + 4 + 7 * 8
HBM_PEAK_GBS = 8000.0


def run_plain(bmx, ctx, graph=True):
    """The plain path (one launch per element per turn, state through HBM; the reference's
    loop nest) on the 901-element ring: 1e4 particles x 20 turns and 1e6 particles x 5 turns,
    3x-sigma beams. graph: the turn's launches captured once and replayed per turn
    (bmx_track_turns_plain); else one host launch per element (bmx_track_element)."""
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    lat = T.Lattice(ctx, E.to_array(L.ring1()))
    out = {"path": "bmx_track_turns_plain (graph-replayed turns)" if graph else
           "bmx_track_element per element (host launches)", "legs": []}
    for n, turns in ((10_000, 20), (1_000_000, 5)):
        beam = L.gaussian_beam(n, seed=1, sigma=3 * L.SIGMA0)
        b = T.Beam.from_host(ctx, beam)
        if graph:
            T.track_beam_plain(ctx, lat, b, 1, 0)  # capture and warm-up turn
        ctx.synchronize()
        ctx.take_work()
        ctx.record(20)
        if graph:
            T.track_beam_plain(ctx, lat, b, turns, 1)
        else:
            for t in range(turns):
                for e in range(1, lat.n_ele + 1):
                    T.track1_bunch(ctx, lat, b, e, 1 + t)
        ctx.record(21)
        ctx.synchronize()
        ms = ctx.elapsed_ms(20, 21)
        passes = ctx.take_work()
        b.close()
        gbs = passes * PLAIN_BYTES_PER_PASS / (ms * 1e-3) / 1e9
        out["legs"].append({"particles": n, "turns": turns, "value": passes / (ms * 1e-3),
                            "unit": UNIT, "ms_per_turn": ms / turns, "hbm_gbs": gbs,
                            "hbm_frac": gbs / HBM_PEAK_GBS,
                            "bytes_per_pass": PLAIN_BYTES_PER_PASS})
    lat.close()
    return out


def run_batched(bmx, ctx, flops, n_var=512, turns=500, seed=2, specialize=True):
    """512 error variants of the ring (sigma 1e-3 on every k1, k2), log-uniform particle
    counts in [256, 65536], 500 turns: one batch, autotuned, replayed as one captured graph."""
    L, T = bmx.lattices, bmx.tracking
    variants = L.ring1_variants(seed, n_var)
    counts = L.variant_particle_counts(seed, n_var)
    total = int(counts.sum())
    beam = L.gaussian_beam(total, seed=seed, sigma=3 * L.SIGMA0)
    lats = [T.Lattice(ctx, bmx.elements.to_array(v)) for v in variants]
    if specialize:
        for lt in lats:
            lt.specialize()
    b = T.Beam.from_host(ctx, beam)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    batch = T.Batch(ctx, [(lats[i], b, int(starts[i]), int(counts[i])) for i in range(n_var)])
    t0 = time.perf_counter()
    shape = batch.tune(sample_turns=10)
    t_tune = time.perf_counter() - t0
    ctx.take_work()
    ctx.record(10)
    batch.track(turns, 0)
    ctx.record(11)
    ctx.synchronize()
    ms = ctx.elapsed_ms(10, 11)
    passes = ctx.take_work()
    alive = b.count_alive()
    batch.close()
    b.close()
    # 6-D closed orbit and one-turn matrix of every variant, all in one call
    t0 = time.perf_counter()
    _, _, co_status = T.closed_orbit(ctx, lats, n_dim=6, max_iter=40)
    t_co = time.perf_counter() - t0
    for lt in lats:
        lt.close()
    rate = passes / (ms * 1e-3)
    return {"workload": f"{n_var} error variants x log-uniform[256, 65536] particles "
                        f"({total} total), {turns} turns, one captured graph",
            "value": rate, "unit": UNIT, "ms": ms, "alive_at_end": alive,
            "tflops_fp64": rate * flops / 1e12, "frac_fp64_peak": rate * flops / 1e12 / FP64_PEAK_TFLOPS,
            "tuned": shape, "tune_s": t_tune,
            "closed_orbits": {"variants": n_var, "n_dim": 6, "s": t_co,
                              "converged": int((co_status == 0).sum())}}


def run_da(bmx, ctx, n_var=4, turns=10_000, seed=3, specialize=True):
    """Dynamic-aperture grid (11 angles x 64 radii x 4 pz) on 4 error variants, 1e4 turns:
    survival turn per particle."""
    L, T = bmx.lattices, bmx.tracking
    variants = L.ring1_variants(seed, n_var)
    lats = [T.Lattice(ctx, bmx.elements.to_array(v)) for v in variants]
    if specialize:
        for lt in lats:
            lt.specialize()
    grid = L.da_grid()
    ctx.take_work()
    t0 = time.perf_counter()
    surv, _, _ = T.dynamic_aperture(ctx, lats, grid, turns)
    dt = time.perf_counter() - t0
    passes = ctx.take_work()
    for lt in lats:
        lt.close()
    return {"workload": f"{n_var} variants x {len(grid['x'])} grid points, {turns} turns",
            "value": passes / dt, "unit": UNIT, "s": dt,
            "survivors": int((surv == turns).sum()), "particles": int(surv.size),
            "median_survival_turn": float(np.median(surv))}


def run_served(bmx, ctx, n_requests, seed=4, load=0.8, budget=16 << 30, specialize=True,
               stream="full", baselines=("one_job",)):
    """Served stream: Poisson arrivals of requests on a pool of 32 error variants, 100-turn
    chunks, HBM budget 16 GiB; full stream: particles log-uniform [1e3, 1e6], turns
    log-uniform [10, 1e4] (reduced: [1e3, 1e5] and [10, 300]). The arrival rate is `load`
    times the served capacity measured on the same requests submitted at once. The baselines
    track the same requests over the same arrival times, one at a time, each in full."""
    L, T = bmx.lattices, bmx.tracking
    S = bmx.serving
    lats = [T.Lattice(ctx, bmx.elements.to_array(v)) for v in L.ring1_variants(seed, 32)]
    if specialize:
        for lt in lats:
            lt.specialize()
    n_hi, t_hi = (1_000_000, 10_000) if stream == "full" else (100_000, 300)
    reqs, gaps = S.make_stream(seed, n_requests, n_hi=n_hi, t_hi=t_hi)
    # capacity: everything submitted at t = 0
    cap = S.run_served(ctx, lats, reqs, np.zeros(len(reqs)), budget)
    mean_gap = cap["makespan_s"] / len(reqs) / load
    arrivals = np.cumsum(gaps) * mean_gap
    arrivals -= arrivals[0]
    served = S.run_served(ctx, lats, reqs, arrivals, budget)
    sv = served["passes"] / served["makespan_s"]
    out = {"workload": f"{n_requests} requests, Poisson arrivals at {load} of served capacity "
                       f"(mean gap {mean_gap:.3f} s), 32 variants, N log-uniform[1e3,{n_hi:.0e}], "
                       f"T log-uniform[10,{t_hi:.0e}], 100-turn chunks, budget {budget} B",
           "stream": stream,
           "served": {"value": sv, **S.percentiles(served["latency_s"]),
                      "makespan_s": served["makespan_s"], "chunks": served["chunks"]},
           "served_capacity": cap["passes"] / cap["makespan_s"],
           "passes": served["passes"], "unit": UNIT}
    if "one_job" in baselines:
        one = S.run_plain(ctx, lats, reqs, arrivals)
        on = one["passes"] / one["makespan_s"]
        out["one_job_at_a_time"] = {"path": "bmx_track_turns (persistent, compacted) per request, "
                                            "no batching across requests",
                                    "value": on, **S.percentiles(one["latency_s"]),
                                    "makespan_s": one["makespan_s"], "passes": one["passes"]}
        out["speedup_vs_one_job"] = sv / on
    if "plain" in baselines:
        plain = S.run_plain_elementwise(ctx, lats, reqs, arrivals)
        pl = plain["passes"] / plain["makespan_s"]
        out["plain"] = {"path": "bmx_track_turns_plain: one launch per element per turn, state in "
                                "HBM, one request at a time, each tracked in full",
                        "value": pl, **S.percentiles(plain["latency_s"]),
                        "makespan_s": plain["makespan_s"], "passes": plain["passes"]}
        out["speedup_vs_plain"] = sv / pl
    for lt in lats:
        lt.close()
    return out


if __name__ == "__main__":
    main()
