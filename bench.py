"""Throughput of the PCISPH step on one MI355X: the 1,048,576-particle dam break.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dam_break_1m]

One JSON line on stdout: particle-steps/s over K timed steps (after W untimed ones)
with every step iterated to max density error < 1% (3..50 correction iterations),
the mean iterations per step, a roofline figure for the dominant kernel from HIP
events taken over the timed region, and a CPU baseline (the oracle, a C++/OpenMP
restatement of the reference step, on the host cores for a bounded sample of the
16,384-particle dam break).
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

METRIC = "particle-steps/s, 1M dam break @≤1% density err; mean PCISPH iters/step"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of one launch, per fluid particle, after the reference
# algorithm's data movement: every candidate's fields are read once per evaluating
# particle (float3 = 12 B, scalars 4 B), plus 27 cell-table entries (8 B) per table
# searched and the particle's own reads and writes.  cf / cb = mean fluid / boundary
# candidates per particle, measured each step.
STAGE_BYTES = {
    "density": lambda cf, cb: 12 * cf + 12 * cb + 2 * 216 + 12 + 4,
    "normal": lambda cf, cb: 16 * cf + 216 + 12 + 12,
    "force_init": lambda cf, cb: 40 * cf + 216 + 40 + 28 + 12,
    "pressure": lambda cf, cb: 12 * cf + 12 * cb + 2 * 216 + 28 + 8,
    "pressure_force": lambda cf, cb: 20 * cf + 16 * cb + 2 * 216 + 20 + 12 + 48 + 12,
}
STAGES = ["boundary_sort", "hash", "sort", "reorder", "density", "normal", "force_init",
          "pressure", "pressure_force"]


def cpu_baseline(seconds: float, max_steps: int) -> dict:
    """The oracle (FP32 restatement of the reference step) on the 16,384-particle dam
    break, from its first step, until `seconds` of CPU work or `max_steps` steps."""
    from oracle import oracle_py
    from pcisph_amd import scenes
    threads = min(os.cpu_count() or 1, 16)
    oracle_py.set_threads(threads)
    scene = scenes.dam_break_16k()
    o = oracle_py.OracleFluid()
    o.load_scene(scene)
    its = []
    t0 = time.perf_counter()
    while len(its) < max_steps and time.perf_counter() - t0 < seconds:
        its.append(o.step()["iterations"])
    dt = time.perf_counter() - t0
    return {"value": scene.fluid_count * len(its) / dt, "unit": "particle-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"dam_break_16k (16,384 fluid + {scene.boundary_count} boundary particles) "
                      f"steps 1..{len(its)} from rest, oracle FP32 C++/OpenMP restatement, "
                      f"{threads} threads, mean {np.mean(its):.2f} iterations/step"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--precision", default="fast", choices=["fast", "ieee", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="no per-stage HIP events")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        return 2

    import pcisph_amd
    from pcisph_amd import scenes

    scene = getattr(scenes, args.workload)()
    prec = {"fast": pcisph_amd.PRECISION_FAST, "ieee": pcisph_amd.PRECISION_IEEE,
            "fp64": pcisph_amd.PRECISION_FP64}[args.precision]
    fluid = pcisph_amd.Fluid(device=0, precision=prec)
    fluid.load_scene(scene)
    print(f"bench: {scene.name} N={scene.fluid_count} Nb={scene.boundary_count} "
          f"warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
    if args.warmup:
        fluid.run(args.warmup, collect=False)
    fluid.sync()
    if not args.no_profile:
        fluid.profile_enable(True)
    t0 = time.perf_counter()
    stats = fluid.run(args.steps, collect=True)
    fluid.sync()
    elapsed = time.perf_counter() - t0
    times = fluid.profile_read() if not args.no_profile else None
    n = scene.fluid_count
    its = np.array([s["iterations"] for s in stats], dtype=np.float64)
    err = np.array([s["max_density_error"] for s in stats], dtype=np.float64)
    conv = np.array([s["converged"] for s in stats])
    cf = float(np.mean([s["fluid_candidates"] for s in stats])) / n
    cb = float(np.mean([s["boundary_candidates"] for s in stats])) / n
    value = n * args.steps / elapsed

    roofline = None
    stage_ms = None
    if times is not None:
        stage_ms = {STAGES[i]: {"ms_total": round(times["ms"][i], 3), "launches": times["launches"][i],
                                "us_per_launch": round(1e3 * times["ms"][i] / times["launches"][i], 2)
                                if times["launches"][i] else None}
                    for i in range(len(STAGES))}
        dom = max(STAGE_BYTES, key=lambda k: times["ms"][STAGES.index(k)])
        i = STAGES.index(dom)
        avg_s = times["ms"][i] / times["launches"][i] / 1e3
        bytes_per_launch = STAGE_BYTES[dom](cf, cb) * n
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None, "bytes_per_launch": int(bytes_per_launch),
                    "avg_launch_us": round(avg_s * 1e6, 2)}

    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_seconds, 200)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": scene.name, "fluid_particles": n, "boundary_particles": scene.boundary_count,
                   "precision": args.precision, "mean_iterations": round(float(its.mean()), 3),
                   "max_iterations_seen": int(its.max()), "steps_converged": int(conv.sum()),
                   "max_density_error_max": round(float(err.max()), 5),
                   "steps_per_s": round(args.steps / elapsed, 2),
                   "mean_fluid_candidates": round(cf, 2), "mean_boundary_candidates": round(cb, 2),
                   "stage_times": stage_ms},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
