"""Learned warm start of the pressure-correction loop: data generation, training and
evaluation, all on the device through the C ABI (pcisph_samples_*, pcisph_mlp_*).

The network (8 -> 64 SiLU -> 64 SiLU -> 1) predicts, from local features the force
pass gathers, how far each fluid particle's converged pressure moves from its
previous step's value (in units of p_ref = delta * rho0 / 100).  The loop starts from
max(0, p_prev + prediction * p_ref) -- a network that outputs zero is the classic
previous-pressure start -- keeps that start only if its first residual beats the
plain (zero-pressure) start's, and still iterates to the same density-error bound.
Samples are (features, converged change) pairs the solver writes while it runs;
training is Adam on random batches of them with the project's own MFMA
forward/backward kernels.
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 4801  # W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised parameters, output layer scaled down."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2 / 8), (64, 8))
    w2 = rng.normal(0, np.sqrt(2 / 64), (64, 64))
    w3 = rng.normal(0, np.sqrt(1 / 64), 64) * 0.1
    return np.concatenate([w1.ravel(), np.zeros(64), w2.ravel(), np.zeros(64), w3, [0.0]]).astype(np.float32)


def zero_params() -> np.ndarray:
    """The network that outputs 0: the previous-pressure start."""
    return np.zeros(N_PARAMS, np.float32)


def collect(fluid, steps: int, stride: int, capacity: int, warm: bool = False) -> int:
    """Run `steps` steps (warm-started with the loaded network if `warm`) storing every
    `stride`-th particle's sample."""
    fluid.set_warm_start(warm)
    fluid.samples_reserve(capacity)
    fluid.samples_collect(stride)
    fluid.run(steps, collect=False)
    fluid.samples_collect(0)
    return fluid.samples_count()


def train(fluid, steps: int = 2000, batch: int = 8192, lr: float = 1e-3, seed: int = 0,
          params: np.ndarray | None = None, log_every: int = 0) -> list[float]:
    """Adam on the collected samples; returns the loss trace (every `log_every` steps)."""
    fluid.mlp_set_params(init_params(seed) if params is None else params)
    losses = []
    chunk = log_every or steps
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        losses.append(fluid.mlp_train(k, batch=batch, lr=lr, seed=seed + 1))
        done += k
    return losses


def mean_iterations(fluid, steps: int, warm: bool) -> dict:
    fluid.set_warm_start(warm)
    st = fluid.run(steps, collect=True)
    its = np.array([s["iterations"] for s in st], float)
    err = np.array([s["max_density_error"] for s in st], float)
    used = np.array([(s["warm_start"] & 1) and not (s["warm_start"] & 2) for s in st])
    return {"mean_iterations": float(its.mean()), "max_error": float(err.max()),
            "steps_converged": int(sum(s["converged"] for s in st)),
            "warm_start_kept": float(used.mean()) if warm else 0.0}


__all__ = ["init_params", "zero_params", "collect", "train", "mean_iterations", "N_PARAMS"]
