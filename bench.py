"""Throughput of the PCISPH step on one MI355X: the 1,048,576-particle dam break.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dam_break_1m]

One JSON line on stdout: particle-steps/s of the reference loop on the 1M dam break
over K timed steps after W untimed ones, the timed steps taken once the start-up
transient has passed (after --settle-steps untimed steps from rest, 400 by default),
where every step iterates to max density error <= 1% (3..50 correction iterations);
the mean iterations per step, a roofline figure for the dominant kernel from HIP
events taken over the timed region, a CPU baseline (the oracle, a C++/OpenMP
restatement of the reference step, on the host cores for a bounded sample of the
16,384-particle dam break), and the other forms -- the same steps from rest (`plain`,
which overshoot the bound), the warm-started, batched and sloshing forms -- beside it.
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


def cpu_opencl_platforms() -> list[str]:
    """OpenCL installable client drivers registered on this host (/etc/OpenCL/vendors);
    a CPU runtime such as PoCL would be listed here.  Nothing is loaded or run."""
    d = "/etc/OpenCL/vendors"
    try:
        return sorted(os.listdir(d))
    except OSError:
        return []


def cpu_baseline(steps: int = 200, warmup: int = 5, repeats: int = 3, single_steps: int = 20) -> dict:
    """The oracle (FP32 restatement of the reference step) on the 16,384-particle dam
    break, as BASELINE.md plans it: `warmup` untimed steps from rest, then `steps` timed
    ones, median of `repeats` runs on the host threads this process may use (OpenMP,
    OMP_NUM_THREADS, else every core), plus one 1-thread run of `single_steps` steps."""
    from oracle import oracle_py
    from pcisph_amd import scenes
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc

    def run(n_threads, n_steps):
        oracle_py.set_threads(n_threads)
        scene = scenes.dam_break_16k()
        o = oracle_py.OracleFluid()
        o.load_scene(scene)
        for _ in range(warmup):
            o.step()
        t0 = time.perf_counter()
        its = [o.step()["iterations"] for _ in range(n_steps)]
        dt = time.perf_counter() - t0
        return scene.fluid_count * n_steps / dt, float(np.mean(its)), scene.boundary_count

    multi = [run(threads, steps) for _ in range(repeats)]
    value = float(np.median([m[0] for m in multi]))
    single = run(1, single_steps)
    icds = cpu_opencl_platforms()
    return {"value": round(value, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "nproc": nproc, "single_thread_value": round(single[0], 1),
            "cpu_opencl_icds": icds,
            "sample": f"dam_break_16k (16,384 fluid + {multi[0][2]} boundary particles): {warmup} untimed steps "
                      f"from rest, then steps {warmup + 1}..{warmup + steps} timed, median of {repeats} runs on "
                      f"{threads} threads (nproc {nproc}; OMP_NUM_THREADS sets this process's share), mean "
                      f"{multi[0][1]:.2f} iterations/step; 1 thread: steps {warmup + 1}..{warmup + single_steps}; "
                      f"oracle FP32 C++/OpenMP restatement of the reference kernels (no CPU OpenCL runtime: "
                      f"{'ICDs ' + ', '.join(icds) if icds else 'no OpenCL ICD registered'})"}


# HBM-side bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by
# the same counters on a copy of known size) for the plain 1M dam break, in two windows:
# steps 1-22 from rest (the start-up transient) and steps 401-422 (settled).
TRAFFIC_FILES = {"rest": os.path.join(ROOT, "profiles", "traffic_dam_break_1m_plain.json"),
                 "settled": os.path.join(ROOT, "profiles", "traffic_dam_break_1m_settled.json")}
# Vector-ALU instructions per dispatch (rocprofv3 SQ_INSTS_VALU) in the same two windows.
VALU_FILE = os.path.join(ROOT, "profiles", "valu_dam_break_1m.json")
# wave64 FP32 VALU issue capacity: 256 CUs x 4 SIMDs, one instruction per SIMD per 4
# cycles at 2.4 GHz (78.6 TFLOP/s of non-packed FP32 FMA)
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4
TRAFFIC_KERNELS = {"density": "k_density", "normal": "k_normal", "force_init": "k_force_init",
                   "pressure": "k_pressure_staged", "pressure_force": "k_pressure_force"}


def _kernel_entry(kernels: dict, stage: str):
    want = TRAFFIC_KERNELS.get(stage)
    for name, v in kernels.items():
        if name.split("::")[-1].split("<")[0] == want:
            return v
    return None


def measured_traffic(stage: str, window: str = "rest"):
    """Measured bytes per launch of a stage's kernel in a window, or None."""
    try:
        with open(TRAFFIC_FILES[window]) as fh:
            v = _kernel_entry(json.load(fh)["kernels"], stage)
    except (OSError, ValueError, KeyError):
        return None
    return v["traffic_bytes"] if v else None


def measured_valu(stage: str, window: str = "rest"):
    """(VALU instructions per dispatch, issue fraction in the counter run) of a stage's
    kernel in a window, or None."""
    try:
        with open(VALU_FILE) as fh:
            v = _kernel_entry(json.load(fh)["windows"][window], stage)
    except (OSError, ValueError, KeyError):
        return None
    return (v["valu_instructions"], v["valu_issue_fraction"]) if v else None


# Compulsory HBM bytes of one launch per fluid particle (and per boundary particle):
# every array the kernel reads or writes touched once -- own position, predicted
# position, pressure in; density variation, pressure, x + p/rho^2 out (the pressure
# pass), the boundary's position, init density and pressure (its boundary mode) -- and
# the cell table read once.
COMPULSORY = {"pressure": (16 + 16 + 4 + 4 + 4 + 16, 16 + 4 + 4 + 16)}


def roofline_of(times: dict, n: int, cf: float, cb: float, traffic: str | None = None, nb: int = 0,
                buckets: int = 0) -> tuple[dict, dict]:
    """Per-stage device times and the HBM roofline of the dominant kernel: measured
    bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a copy) over its
    mean HIP-event duration, against 8 TB/s.  Beside it, the kernel's compulsory bytes
    and the gather model (every candidate's fields read once per evaluating particle,
    cache hits included) over the same duration, and its vector-ALU issue rate
    (rocprofv3 SQ_INSTS_VALU per dispatch over the same duration, against the wave64
    FP32 issue capacity): the bound of a kernel that reuses its gathers from LDS."""
    stage_ms = {STAGES[i]: {"ms_total": round(times["ms"][i], 3), "launches": times["launches"][i],
                            "us_per_launch": round(1e3 * times["ms"][i] / times["launches"][i], 2)
                            if times["launches"][i] else None}
                for i in range(len(STAGES))}
    dom = max(STAGE_BYTES, key=lambda k: times["ms"][STAGES.index(k)])
    i = STAGES.index(dom)
    avg_s = times["ms"][i] / times["launches"][i] / 1e3
    gather = STAGE_BYTES[dom](cf, cb) * n
    measured = measured_traffic(dom, traffic) if traffic else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": measured, "avg_launch_us": round(avg_s * 1e6, 2),
                "bytes_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch (" +
                                (os.path.relpath(TRAFFIC_FILES[traffic], ROOT) if traffic else "none") +
                                ") / HIP-event mean duration",
                "gather_model_bytes": int(gather), "gather_model_GBs": round(gather / avg_s / 1e9, 1)}
    valu = measured_valu(dom, traffic) if traffic else None
    if valu:
        rate = valu[0] / avg_s
        roofline["valu"] = {"instructions_per_launch": valu[0], "achieved_per_s": round(rate, -6),
                            "peak_per_s": VALU_ISSUE_PEAK, "frac": round(rate / VALU_ISSUE_PEAK, 4),
                            "counter_run_frac": valu[1],
                            "source": "rocprofv3 SQ_INSTS_VALU per dispatch (" + os.path.relpath(VALU_FILE, ROOT) +
                                      ") / HIP-event mean duration; wave64 FP32 issue, one per SIMD per 4 cycles"}
    if dom in COMPULSORY:
        per_f, per_b = COMPULSORY[dom]
        comp = per_f * n + per_b * nb + 8 * buckets
        roofline.update({"compulsory_bytes": int(comp), "compulsory_GBs": round(comp / avg_s / 1e9, 1)})
    if measured:
        achieved = measured / avg_s / 1e9
        roofline.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4)})
    return stage_ms, roofline


def measure(fluid, steps: int, warmup: int, n: int, profile: bool, batch=None, traffic: str | None = None,
            scene=None, profile_every: int = 8, settle: int = 0) -> dict:
    """`warmup` untimed then `steps` timed steps (every scene of a batch), HIP-event
    stage times over the timed region when `profile`: every `profile_every`-th step of
    it is launched kernel by kernel between events, the others replay their HIP graphs
    (events cannot be read back from graph replays).  Moving walls (sloshing scenes)
    are translated before every step."""
    from pcisph_amd import scenes
    shift = None
    if batch is not None and any(sc.meta.get("slosh") for sc in batch):
        shift = lambda k: fluid.translate_boundaries([scenes.boundary_shift(sc, k) for sc in batch])  # noqa: E731
    elif scene is not None and scene.meta.get("slosh"):
        shift = lambda k: fluid.translate_boundary(*scenes.boundary_shift(scene, k))  # noqa: E731

    lat = []

    def go(count, start):
        out = []
        for k in range(count):
            t = time.perf_counter()
            if shift:
                shift(start + k)
            out.append(fluid.step_scenes() if batch is not None else [fluid.step()])
            lat.append(time.perf_counter() - t)
        return out

    go(settle + warmup, 0)
    fluid.sync()
    lat.clear()
    if profile:
        fluid.profile_enable(True, every=profile_every)
    t0 = time.perf_counter()
    stats = go(steps, settle + warmup)
    fluid.sync()
    elapsed = time.perf_counter() - t0
    times = fluid.profile_read() if profile else None
    if profile:
        fluid.profile_enable(False)
    its = np.array([[q["iterations"] for q in st] for st in stats], dtype=np.float64)
    err = np.array([[q["max_density_error"] for q in st] for st in stats], dtype=np.float64)
    conv = np.array([[q["converged"] for q in st] for st in stats])
    live = its > 0
    cf = float(np.sum([[q["fluid_candidates"] for q in st] for st in stats])) / (n * steps)
    cb = float(np.sum([[q["boundary_candidates"] for q in st] for st in stats])) / (n * steps)
    out = {"value": round(n * steps / elapsed, 1), "ms_per_step": round(1e3 * elapsed / steps, 4),
           "steps_per_s": round(steps / elapsed, 2), "mean_iterations": round(float(its[live].mean()), 3),
           "max_iterations_seen": int(its.max()), "converged_fraction": round(float(conv[live].mean()), 4),
           "max_density_error_max": round(float(err.max()), 5),
           "meets_bound": bool(conv[live].all() and err.max() <= 0.01),
           "mean_fluid_candidates": round(cf, 2), "mean_boundary_candidates": round(cb, 2)}
    if batch is not None:
        out["mean_iterations_slowest_scene"] = round(float(its.max(axis=1).mean()), 3)
        # every scene of a batch completes its step when the batch's step returns
        out["step_latency_ms_p50_p99"] = [round(1e3 * float(np.percentile(lat, q)), 3) for q in (50, 99)]
    if times is not None:
        nb = sum(sc.boundary_count for sc in batch) if batch is not None else (scene.boundary_count if scene else 0)
        buckets = fluid.get_params().bucket_count if batch is None else 0
        out["stage_times"], out["roofline"] = roofline_of(times, n, cf, cb, traffic, nb, buckets)
    if settle:
        out["settled_after_steps"] = settle
    return out


# The committed network's training samples came from the block with jitter seed
# 20240611 (steps 400-700) and the 64-scene batch (pcisph_amd/data/warm_start.json).
HELD_OUT_SEEDS = (7, 99)


def warm_gain(scene, weights: str, pre: int, steps: int, prec: int) -> dict:
    """Mean correction iterations per step from one saved state (`pre` plain steps in),
    `steps` steps each: plain loop, previous-pressure start, trained network."""
    import pcisph_amd
    from pcisph_amd import scenes
    from pcisph_amd import warm_start as ws
    f = pcisph_amd.Fluid(device=0, precision=prec)
    f.load_scene(scene)
    shift = (lambda k: f.translate_boundary(*scenes.boundary_shift(scene, k))) \
        if scene.meta.get("slosh") else None  # noqa: E731

    def run(n, start, warm):
        f.set_warm_start(warm)
        out = []
        for k in range(n):
            if shift:
                shift(start + k)
            out.append(f.step())
        return out

    run(pre, 0, False)
    f.snapshot()
    res = {}
    for name, p in (("plain", None), ("previous_pressure", ws.zero_params()), ("network", np.load(weights))):
        f.restore()
        if p is not None:
            f.mlp_set_params(p)
        st = run(steps, pre, p is not None)
        its = np.array([s["iterations"] for s in st], float)
        res[name] = {"mean_iterations": round(float(its.mean()), 3),
                     "max_density_error": round(float(max(s["max_density_error"] for s in st)), 5),
                     "steps_converged": int(sum(s["converged"] for s in st))}
    f.close()
    for k in ("network", "previous_pressure"):
        res[f"iteration_reduction_{k}"] = round(1 - res[k]["mean_iterations"] / res["plain"]["mean_iterations"], 4)
    res["window"] = f"steps {pre + 1}..{pre + steps} from one saved state"
    res["seed"] = scene.meta.get("seed")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--precision", default="fast", choices=["fast", "ieee", "fp64"])
    ap.add_argument("--cpu-steps", type=int, default=200, help="timed oracle steps per CPU-baseline run")
    ap.add_argument("--settle-steps", type=int, default=400,
                    help="untimed steps before the timed ones of the plain_settled form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="no per-stage HIP events")
    ap.add_argument("--profile-every", type=int, default=8,
                    help="time every n-th step of the timed region kernel by kernel between HIP events")
    ap.add_argument("--forms", default="plain,plain_settled,warm_start,warm_start_settled,batch_64,batch_256,"
                                       "batch_256_warm,sloshing_plain,sloshing_warm,gain,gain_sloshing",
                    help="comma list of plain (reference loop), plain_settled (the reference loop after "
                         "--settle-steps untimed steps from rest, once the start-up transient has passed), "
                         "warm_start_settled (a held-out jitter seed settled by --settle-steps plain steps, then the timed "
                         "window from one saved state with the plain loop and with the trained network), "
                         "previous_pressure (each step starts from "
                         "the previous pressures), warm_start (trained network), batch_64 (64-scene batch), "
                         "sloshing_plain / sloshing_warm (the 1M block in a tank oscillating along x, "
                         "0.5 m at 0.8 Hz; plain loop / trained network), gain / gain_sloshing (iterations "
                         "per step of plain loop, previous-pressure start and network from one saved state, "
                         "100 steps each: gain on jitter seeds the network was not trained on, 400 steps in, "
                         "and on its training seed 900 steps in; gain_sloshing on the tank), batch_256 / batch_256_warm (256 scenes of 1,024..65,536 "
                         "particles: dam breaks, sphere drops, sloshing tanks; plain / trained network)")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--work-order", type=int, default=None,
                    help="solver option work_order for the single-scene forms (default: the solver's)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "pcisph_amd", "data", "warm_start.npy"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device five times (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        import torch
        n = 64 << 20
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}), flush=True)
        return 0

    import pcisph_amd
    from pcisph_amd import scenes

    forms = [f for f in args.forms.split(",") if f]
    prec = {"fast": pcisph_amd.PRECISION_FAST, "ieee": pcisph_amd.PRECISION_IEEE,
            "fp64": pcisph_amd.PRECISION_FP64}[args.precision]
    scene = getattr(scenes, args.workload)()
    n = scene.fluid_count
    results = {}
    for form in forms:
        if form in ("gain", "gain_sloshing"):
            if not os.path.exists(args.weights):
                continue
            make = getattr(scenes, args.workload)
            if form == "gain_sloshing":  # out of distribution: no moving walls in the training data
                cases = [(scenes.sloshing_tank(0.5, 0.8, make()), 400)]
            else:  # held out: jitter seeds the network was not trained on, and a later window
                cases = [(make(seed=HELD_OUT_SEEDS[0]), 400), (make(seed=HELD_OUT_SEEDS[1]), 400), (make(), 900)]
            results[form] = {}
            for sc, pre in cases:
                key = f"{sc.name}_seed_{sc.meta.get('seed')}_from_step_{pre}"
                print(f"bench: {form} {key}", file=sys.stderr, flush=True)
                results[form][key] = warm_gain(sc, args.weights, pre, 100, prec)
                print(f"bench: {form} {json.dumps(results[form][key])}", file=sys.stderr, flush=True)
            continue
        if form == "warm_start_settled":
            # the chosen (warm-started) form at the bound: a block the network was not
            # trained on, settled by the plain loop, then the same timed window from one
            # saved state with the plain loop and with the trained network
            if not os.path.exists(args.weights):
                continue
            sc = getattr(scenes, args.workload)(seed=HELD_OUT_SEEDS[0])
            fluid = pcisph_amd.Fluid(device=0, precision=prec)
            fluid.load_scene(sc, **({} if args.work_order is None else {"work_order": args.work_order}))
            print(f"bench: {form} {sc.name} seed {sc.meta.get('seed')} settle={args.settle_steps} "
                  f"warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
            for _ in range(args.settle_steps):
                fluid.step()
            fluid.snapshot()
            plain = measure(fluid, args.steps, args.warmup, sc.fluid_count, not args.no_profile, scene=sc,
                            profile_every=args.profile_every)
            fluid.restore()
            fluid.mlp_set_params(np.load(args.weights))
            fluid.set_warm_start(True)
            r = measure(fluid, args.steps, args.warmup, sc.fluid_count, not args.no_profile, scene=sc,
                        profile_every=args.profile_every)
            fluid.close()
            keep = ("value", "ms_per_step", "mean_iterations", "converged_fraction", "max_density_error_max",
                    "meets_bound")
            r["plain_same_window"] = {k: plain[k] for k in keep}
            r["speedup_over_plain"] = round(r["value"] / plain["value"], 4)
            r.update({"seed": sc.meta.get("seed"), "settled_after_steps": args.settle_steps})
            results[form] = r
            print(f"bench: {form} {json.dumps({k: v for k, v in r.items() if k != 'stage_times'})}",
                  file=sys.stderr, flush=True)
            for name, q in (("plain", plain), ("warm", r)):
                if q.get("stage_times"):
                    print(f"bench: {form} {name} stage_times {json.dumps(q['stage_times'])}", file=sys.stderr,
                          flush=True)
            continue
        if form in ("batch_64", "batch_256", "batch_256_warm"):
            batch = scenes.batch_64() if form == "batch_64" else scenes.batch_256()
            total = sum(sc.fluid_count for sc in batch)
            fluid = pcisph_amd.Fluid(device=0, precision=prec)
            fluid.load_scenes(batch)
            tuned = None
            if form == "batch_64":  # block size and radix digit width from the autotuner
                tuned = fluid.autotune(trial_steps=2)
            if form == "batch_256_warm":
                if not os.path.exists(args.weights):
                    continue
                fluid.mlp_set_params(np.load(args.weights))
                fluid.set_warm_start(True)
            print(f"bench: {form} {len(batch)} scenes N={total} steps={args.batch_steps}", file=sys.stderr,
                  flush=True)
            r = measure(fluid, args.batch_steps, args.warmup, total, not args.no_profile, batch=batch,
                        profile_every=args.profile_every)
            # scenes whose loop has ended skip the remaining iterations, so the per-particle
            # byte model of a launch does not hold for a batch: stage times only
            r.pop("roofline", None)
            r.update({"scenes": len(batch), "fluid_particles": total,
                      "boundary_particles": sum(sc.boundary_count for sc in batch)})
            if tuned:
                r["autotuned"] = tuned
        else:
            sc = scenes.sloshing_tank(0.5, 0.8, getattr(scenes, args.workload)()) \
                if form.startswith("sloshing") else scene
            fluid = pcisph_amd.Fluid(device=0, precision=prec)
            fluid.load_scene(sc, **({} if args.work_order is None else {"work_order": args.work_order}))
            if form in ("warm_start", "sloshing_warm"):
                if not os.path.exists(args.weights):
                    print(f"bench: no trained network at {args.weights}; warm_start skipped", file=sys.stderr)
                    continue
                fluid.mlp_set_params(np.load(args.weights))
                fluid.set_warm_start(True)
            elif form == "previous_pressure":
                fluid.mlp_set_params(np.zeros(pcisph_amd.MLP_PARAMS, np.float32))
                fluid.set_warm_start(True)
            print(f"bench: {form} {sc.name} N={n} Nb={sc.boundary_count} warmup={args.warmup} "
                  f"steps={args.steps}", file=sys.stderr, flush=True)
            r = measure(fluid, args.steps, args.warmup, n, not args.no_profile,
                        traffic=({"plain": "rest", "plain_settled": "settled"}.get(form)
                                 if args.workload == "dam_break_1m" else None), scene=sc,
                        profile_every=args.profile_every,
                        settle=args.settle_steps if form == "plain_settled" else 0)
        fluid.close()
        results[form] = r
        print(f"bench: {form} {json.dumps({k: v for k, v in r.items() if k != 'stage_times'})}",
              file=sys.stderr, flush=True)
        if r.get("stage_times"):
            print(f"bench: {form} stage_times {json.dumps(r['stage_times'])}", file=sys.stderr, flush=True)

    timed = [k for k, r in results.items() if "value" in r]
    if not timed:
        print("bench: no timed form ran", file=sys.stderr)
        return 1
    # the headline is the reference loop where it meets the metric's bound: the 1M dam
    # break past its start-up transient (the steps from rest overshoot the bound under the
    # reference algorithm whatever the start; they are reported as the `plain` form)
    head = next((k for k in ("plain_settled", "plain") if k in timed), timed[0])
    big = ("batch_64", "batch_256", "batch_256_warm")
    main_r = results[head]
    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(steps=args.cpu_steps)

    line = {
        "metric": METRIC, "value": main_r["value"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps if head not in big else args.batch_steps, "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": scene.name if head not in big else head, "form": head,
                   "fluid_particles": n, "boundary_particles": scene.boundary_count,
                   "precision": args.precision, "density_error_bound": 0.01,
                   "iterations_min_max": [3, 50],
                   "stage_timing": None if args.no_profile else f"HIP events, every {args.profile_every}th step",
                   **{k: v for k, v in main_r.items() if k not in ("value", "ms_per_step", "roofline")}},
        "roofline": main_r.get("roofline"), "cpu_baseline": cpu,
        "forms": {k: {q: v for q, v in r.items() if q != "stage_times"} for k, r in results.items()},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
