"""Throughput of bmad_standard FP64 multi-particle tracking on one MI355X.

Workload (default): the 901-element ring with sextupoles, one rf cavity and elliptical
apertures on every element; 10^6 electrons at 5 GeV from a seeded 6-D Gaussian with 3x the
base sigmas; 1000 turns. One step = 100 turns of the whole beam in one persistent launch
(state resident in HBM between steps), so the default 10 steps are the full 1000 turns.
Throughput counts particle-element passes of live particles only, as the reference's loop
skips dead particles (bmad/multiparticle/beam_utils.f90:83).

Prints one JSON line: value = particle-element-turns/s, the roofline of the tracking kernel
(FP64 operations counted with the algorithmic convention of flop_counts.json), and the CPU
restatement's OpenMP throughput on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
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
    ap.add_argument("--turns-per-step", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU time of the cpu_baseline sample (0 = skip)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--maps", choices=["fast", "reference"], default="fast",
                    help="formulation of the element maps")
    return ap.parse_args()


def cpu_baseline(bmx, arr, beam, target_s):
    """The CPU restatement (OpenMP over particles per element, as track1_bunch_hom) on a
    bounded sample of the same workload: a particle subset over a few turns."""
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = 20000
    sample = {k: np.ascontiguousarray(v[:n]) for k, v in beam.items()}
    # calibrate with one turn, then run enough turns for ~target_s
    t0 = time.perf_counter()
    done = oracle.track_bunch(arr, {k: v.copy() for k, v in sample.items()}, 1, n_threads=threads)
    t1 = time.perf_counter()
    rate1 = done / max(t1 - t0, 1e-9)
    turns = max(1, int(target_s * rate1 / (n * len(arr))))
    t0 = time.perf_counter()
    done = oracle.track_bunch(arr, sample, turns, n_threads=threads)
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"{n} of the {beam['x'].shape[0]} particles x {len(arr)} elements x "
                      f"{turns} turns ({dt:.1f} s, C restatement, -O2 -fopenmp)"}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: tracking here runs on a single MI355X and has no "
              "multi-GPU path (particles are independent; a beam would shard by particle "
              "ranges, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    bmx = load_package()
    L, E, T = bmx.lattices, bmx.elements, bmx.tracking
    eles = L.ring1()
    arr = E.to_array(eles)
    beam = L.gaussian_beam(args.particles, seed=args.seed, sigma=3 * L.SIGMA0)
    flops = bmx.flops.per_element_pass(arr)

    ctx = T.Context(0)
    ctx.set_maps(args.maps)
    lat = T.Lattice(ctx, arr)
    tps = args.turns_per_step
    if args.warmup > 0:
        wb = T.Beam.from_host(ctx, beam)
        for w in range(args.warmup):
            T.track_beam(ctx, lat, wb, tps, w * tps)
        ctx.synchronize()
        wb.close()
    b = T.Beam.from_host(ctx, beam)
    ctx.take_work()
    ctx.synchronize()
    t_wall0 = time.perf_counter()
    ctx.record(0)
    for s in range(args.steps):
        ctx.record(1 + 2 * s)
        T.track_beam(ctx, lat, b, tps, s * tps)
        ctx.record(2 + 2 * s)
    ctx.record(63)
    ctx.synchronize()
    t_wall = time.perf_counter() - t_wall0
    total_ms = ctx.elapsed_ms(0, 63)
    launch_ms = [ctx.elapsed_ms(1 + 2 * s, 2 + 2 * s) for s in range(args.steps)]
    passes = ctx.take_work()
    alive = b.count_alive()

    value = passes / (total_ms * 1e-3)
    kernel_s = np.mean(launch_ms) * 1e-3
    achieved = flops * (passes / args.steps) / kernel_s / 1e12
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
                        f"{args.steps * tps} turns in steps of {tps}",
            "particles": args.particles, "elements": len(arr), "turns": args.steps * tps,
            "turns_per_step": tps, "alive_at_end": alive, "wall_s": t_wall,
            "maps": args.maps,
        },
        "roofline": {
            "bound": "mfma", "compute": "fp64 vector ALU", "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None, "flop_per_particle_element": flops,
            "kernel": "k_track_turns", "kernel_ms_avg": float(np.mean(launch_ms)),
        },
    }
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(bmx, arr, beam, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
