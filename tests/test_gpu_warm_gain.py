"""GPU: what the trained warm-start network gains over the loop without it, on a
trajectory it was not trained on.

The committed network was trained on samples of the 1,048,576-particle dam break with
jitter seed 20240611 (steps 400-700) and of the 64-scene batch.  Here the same block
with jitter seed 7 -- a different trajectory -- is run 400 plain steps past its
start-up transient and saved; from that state, 100 steps each of: the plain loop
(every step from zero pressure, the reference's), the previous-pressure start (a zero
network), and the trained network -- same density-error bound, same iteration
limits.  The network must cut the mean iterations per step against the plain loop by
a fifth, do better than the previous-pressure start, and converge as often.
"""
import json
import os

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd import warm_start as ws

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pcisph_amd", "data",
                       "warm_start.npy")


HELD_OUT_SEED = 7


def test_trained_network_cuts_iterations_held_out():
    import pcisph_amd
    params = np.load(WEIGHTS)
    assert params.shape == (pcisph_amd.MLP_PARAMS,) and np.all(np.isfinite(params))
    print(f"case warm gain 1M dam break seed {HELD_OUT_SEED} from step 400", flush=True)
    f = pcisph_amd.Fluid()
    f.load_scene(scenes.dam_break_1m(seed=HELD_OUT_SEED), max_iterations=50)
    f.run(400, collect=False)
    f.snapshot()
    res = {}
    for name, p in (("plain", None), ("previous_pressure", ws.zero_params()), ("network", params)):
        f.restore()
        if p is not None:
            f.mlp_set_params(p)
        res[name] = ws.mean_iterations(f, 100, p is not None)
        print(name, json.dumps(res[name]), flush=True)
    plain, prev, net = res["plain"], res["previous_pressure"], res["network"]
    assert net["mean_iterations"] <= 0.8 * plain["mean_iterations"], res
    assert net["mean_iterations"] < prev["mean_iterations"], res
    assert net["steps_converged"] >= plain["steps_converged"], res
    assert net["warm_start_kept"] > 0.5, res
