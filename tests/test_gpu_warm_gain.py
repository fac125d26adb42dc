"""GPU: what the trained warm-start network gains over the loop without it.

From one saved state of the 1,048,576-particle dam break (step 400 of the plain loop,
past its start-up transient), 100 steps each of: the plain loop (every step from zero
pressure, the reference's), the previous-pressure start (a zero network), and the
committed trained network -- same density-error bound, same iteration limits.  The
network must cut the mean iterations per step against the plain loop by a fifth, do
at least as well as the previous-pressure start, and converge as often.
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


def test_trained_network_cuts_iterations():
    import pcisph_amd
    params = np.load(WEIGHTS)
    assert params.shape == (pcisph_amd.MLP_PARAMS,) and np.all(np.isfinite(params))
    print("case warm gain 1M dam break from step 400", flush=True)
    f = pcisph_amd.Fluid()
    f.load_scene(scenes.dam_break_1m(), max_iterations=50)
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
    assert net["mean_iterations"] <= prev["mean_iterations"] + 0.05, res
    assert net["steps_converged"] >= plain["steps_converged"], res
    assert net["warm_start_kept"] > 0.5, res
