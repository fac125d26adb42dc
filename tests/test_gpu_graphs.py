"""GPU: steps replayed from HIP graphs equal steps launched one by one.

With graph_iterations > 0 the launches before the correction loop form one graph and
the loop's iterations are replayed `graph_iterations` at a time; a step's tag lives
in device memory, so every replay publishes the step that runs it.  Each test runs
the same scene both ways and compares every step's statistics and, after every step
(downloads and synchronisations between the replays), the particle state bit for bit.
"""
import os

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field
from pcisph_amd.warm_start import init_params
from tests.helpers import FLOAT_FIELDS, INT_FIELDS, run, small_scenes, snapshot

pytestmark = pytest.mark.gpu

EACH_STEP = [Field.FLUID_POSITIONS, Field.FLUID_VELOCITIES, Field.FLUID_PRESSURES, Field.FLUID_KEYS]
STAT_KEYS = ("iterations", "converged", "max_density_variation", "fluid_candidates", "boundary_candidates",
             "warm_start", "list_waves")


@pytest.fixture(scope="module")
def gpu():
    import pcisph_amd
    return pcisph_amd


def stepwise(solver, steps, batch=False):
    stats, states = [], []
    for _ in range(steps):
        stats.append(solver.step_scenes() if batch else [solver.step()])
        states.append(snapshot(solver, EACH_STEP))
    return stats, states


def assert_same(a, b):
    (sa, xa), (sb, xb) = a, b
    for step, (pa, pb) in enumerate(zip(sa, sb)):
        for qa, qb in zip(pa, pb):
            for key in STAT_KEYS:
                assert qa[key] == qb[key], (step, key, qa[key], qb[key])
    for step, (ra, rb) in enumerate(zip(xa, xb)):
        for name in ra:
            np.testing.assert_array_equal(ra[name], rb[name], err_msg=f"step {step} {name}")


@pytest.mark.parametrize("chunk", [1, 3, 8])
def test_graph_steps_equal_eager(gpu, chunk):
    scene = small_scenes()[1]
    print(f"case graphs chunk={chunk} {scene.name}", flush=True)
    eager = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    eager.load_scene(scene, max_iterations=20)
    graph = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    graph.load_scene(scene, max_iterations=20, graph_iterations=chunk)
    assert_same(stepwise(eager, 6), stepwise(graph, 6))
    full_e, full_g = snapshot(eager), snapshot(graph)
    for name in full_e:
        np.testing.assert_array_equal(full_g[name], full_e[name], err_msg=name)


def test_graph_matches_oracle(gpu, oracle_lib):
    scene = small_scenes()[0]
    print(f"case graphs vs oracle {scene.name}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20)
    sg = run(g, scene, 3, max_iterations=20, graph_iterations=4)
    assert [s["iterations"] for s in so] == [s["iterations"] for s in sg]
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


@pytest.mark.parametrize("network", ["random", "trained"])
def test_graph_batch_warm_start_equal_eager(gpu, network):
    a, b = small_scenes()
    batch = [a, b, scenes.scaled_sphere_drop(1500, seed=5)]
    if network == "random":
        params = init_params(3)
        params[-1] = 0.5
    else:  # the committed network, whose starts these scenes keep
        params = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pcisph_amd",
                                      "data", "warm_start.npy"))
    print(f"case graphs batch warm start {network}", flush=True)
    runs = []
    for chunk in (0, 4):
        g = gpu.Fluid(precision=gpu.PRECISION_FAST)
        g.mlp_set_params(params)
        g.set_warm_start(True)
        g.load_scenes(batch, max_iterations=20, graph_iterations=chunk)
        runs.append(stepwise(g, 4, batch=True))
    assert_same(runs[0], runs[1])


def test_graphs_recaptured_after_changes(gpu):
    """Block size, particle count and the warm start switched between steps: the
    graphs follow (they are re-captured) and stay equal to launching eagerly."""
    scene = small_scenes()[0]
    other = small_scenes()[1]
    print("case graphs recapture", flush=True)
    results = []
    for chunk in (0, 2):
        g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        g.load_scene(scene, max_iterations=20, graph_iterations=chunk)
        out = [stepwise(g, 2)]
        g.set_solver_options(max_iterations=20, block_size=128, graph_iterations=chunk)
        out.append(stepwise(g, 2))
        g.load_scene(other, max_iterations=12, block_size=128, graph_iterations=chunk)
        out.append(stepwise(g, 2))
        g.mlp_set_params(init_params(1))
        g.set_warm_start(True)
        out.append(stepwise(g, 2))
        results.append(out)
    for ea, ga in zip(results[0], results[1]):
        assert_same(ea, ga)
