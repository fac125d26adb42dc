"""Train the warm-start network on samples the solver generates, on the GPU.

    python -m pcisph_amd.train_warm_start --out pcisph_amd/data/warm_start.npy

Samples are (features, converged pressure change) pairs written while the solver
runs: every `stride`-th fluid particle of every step in a collection window, from the
1,048,576-particle dam break and from the 64-scene batch (counts log-uniform in
[2,048, 262,144], dam breaks and sphere drops).  With `--collect plain` (default) the
solver runs the plain loop after its start-up transient, so a target is the change
between two converged pressures that each started from zero -- the pressure a step
needs, whatever start reaches it (PCISPH only raises pressures within a step, so
pressures reached from a warm start are biased towards that start).  `--collect
previous` / `--init` collect with the previous-pressure start / a given network.  The network is
then trained with the project's own BF16 MFMA forward/backward kernels and FP32 Adam,
and evaluated from one saved state of the 1M dam break: mean correction iterations
per step of the plain loop, of the previous-pressure start and of the trained network,
all to the same density-error bound.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import Fluid, scenes
from . import warm_start as ws


def collect_from(fluid: Fluid, pre: int, steps: int, stride: int, capacity: int, batch=None,
                 policy: np.ndarray | None = None, warm: bool = True) -> tuple:
    """Run `pre` steps, then `steps` steps storing every `stride`-th particle, all
    warm-started with `policy` (default: the previous-pressure start) or all plain."""
    shift = None
    if batch is not None and any(sc.meta.get("slosh") for sc in batch):
        shift = lambda k: [scenes.boundary_shift(sc, k) for sc in batch]  # noqa: E731

    def advance(n, start):
        for k in range(n):
            if shift:
                fluid.translate_boundaries(shift(start + k))
            fluid.step()

    fluid.mlp_set_params(ws.zero_params() if policy is None else policy)
    fluid.set_warm_start(warm)
    advance(pre, 0)
    fluid.samples_reserve(capacity)
    fluid.samples_collect(stride)
    advance(steps, pre)
    fluid.samples_collect(0)
    return fluid.samples_download()


def evaluate(fluid: Fluid, params: np.ndarray, steps: int) -> dict:
    """Mean iterations from the fluid's current state: plain, previous pressure, network."""
    fluid.snapshot()
    res = {}
    for name, p in (("plain", None), ("previous_pressure", ws.zero_params()), ("network", params)):
        fluid.restore()
        if p is not None:
            fluid.mlp_set_params(p)
        res[name] = ws.mean_iterations(fluid, steps, p is not None)
    fluid.restore()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True, help="where to write the FP32 parameter vector (.npy)")
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None,
                    help="parameters (.npy) to collect with and to continue training from (an on-policy "
                         "round); default: train from scratch on --collect samples")
    ap.add_argument("--collect", choices=["plain", "previous"], default="plain",
                    help="plain: the plain loop after the start-up transient; previous: the "
                         "previous-pressure start from step 50")
    args = ap.parse_args(argv)
    policy = np.load(args.init).astype(np.float32) if args.init else None
    warm = policy is not None or args.collect == "previous"
    # the plain loop's pressures are converged only after the 1M start-up transient
    pre_1m, pre_batch = (400, 60) if not warm else (50, 20)
    t0 = time.time()
    half = args.samples // 2

    # config 1: the 1M dam break, samples over steps 50..349
    big = Fluid()
    big.load_scene(scenes.dam_break_1m(), max_iterations=50)
    x1, t1 = collect_from(big, pre_1m, 300, max(1, 300 * big.field_size(0) // 3 // half), half, policy=policy,
                          warm=warm)
    print(f"1M dam break: {len(t1)} samples, {time.time() - t0:.1f} s", flush=True)

    # config 2: the 64-scene batch, samples over steps 20..59
    batch = scenes.batch_64()
    b = Fluid()
    b.load_scenes(batch, max_iterations=50)
    total = sum(sc.fluid_count for sc in batch)
    x2, t2 = collect_from(b, pre_batch, 40, max(1, 40 * total // half), half, batch, policy=policy, warm=warm)
    print(f"64-scene batch: {len(t2)} samples, {time.time() - t0:.1f} s", flush=True)
    b.close()

    x = np.concatenate([x1, x2])
    t = np.concatenate([t1, t2])
    state_samples = big.samples_count()
    big.samples_reserve(len(t))
    big.samples_upload(x, t)
    losses, params = [], policy
    schedule = ((3e-3, 1 / 3), (1e-3, 1 / 3), (3e-4, 1 / 3)) if policy is None else ((1e-3, 1 / 2), (3e-4, 1 / 2))
    for lr, frac in schedule:
        losses += ws.train(big, steps=max(1, int(args.train_steps * frac)), batch=args.batch, lr=lr,
                           seed=args.seed, params=params, log_every=1000)
        params = big.mlp_get_params()
    print(f"trained: losses {[round(v, 5) for v in losses]}, {time.time() - t0:.1f} s", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)

    res = evaluate(big, params, args.eval_steps)
    meta = {"samples": int(len(t)), "samples_1m": int(len(t1)), "samples_batch": int(len(t2)),
            "collected_1m_before_cap": int(state_samples), "train_steps": args.train_steps, "batch": args.batch,
            "collected_with": args.init or ("plain loop" if not warm else "previous-pressure start"),
            "collection_steps_1m": [pre_1m, pre_1m + 300],
            "final_loss": losses[-1], "target_mean": float(t.mean()), "target_std": float(t.std()),
            "evaluation_1m_after_collection": res, "seconds": round(time.time() - t0, 1)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(meta, fh, indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
