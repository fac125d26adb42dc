"""GPU: batches of scenes stepped together (one launch per stage over scene tables).

A scene's particles never interact with another scene's, the concatenated sort keeps
every scene in its own sort order, and each scene's loop ends on its own test, so a
scene stepped inside a batch must leave exactly the bits it leaves when stepped alone
-- in every precision, with or without neighbour lists and warm start.  Together with
the single-scene parity tests against the oracle this pins every scene of a batch to
the oracle.
"""
import os

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field
from pcisph_amd.warm_start import init_params
from tests.helpers import FLOAT_FIELDS, INT_FIELDS, run, small_scenes, snapshot

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pcisph_amd", "data",
                       "warm_start.npy")

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iterations", "converged", "max_density_variation", "max_density_error", "fluid_candidates",
             "boundary_candidates", "warm_start", "list_waves")


@pytest.fixture(scope="module")
def gpu():
    import pcisph_amd
    return pcisph_amd


def mixed_scenes():
    """Three scenes of different sizes, bucket counts and particle radii, with
    different thresholds so their loops end at different iterations."""
    a, b = small_scenes()
    c = scenes.scaled_sphere_drop(1500, seed=5, name="sphere_drop_small")
    b.threshold = 0.02
    c.threshold = 0.005
    return [a, b, c]


def assert_warm_rule(steps, thresholds, min_iterations=3, max_iterations=50):
    """Each scene's warm-start flags across warm-started steps: every scene starts from
    the network in the first step; a scene is left out of the next step (bit 2) after a
    step that ran to the cap unconverged, or that converged at the minimum with the
    plain start -- or with the network's start while the plain start's first
    iteration was already within the bound -- and a scene left out runs plain."""
    for k, step in enumerate(steps):
        for j, s in enumerate(step):
            if not s["iterations"]:
                continue
            expect = 1 if k == 0 else int(not steps[k - 1][j]["warm_start"] & 4)
            assert s["warm_start"] & 1 == expect, (k, j, s)
            capped = s["iterations"] >= max_iterations and not s["converged"]
            kept = s["warm_start"] & 3 == 1
            at_floor = s["converged"] and s["iterations"] <= min_iterations
            quiet = s["first_density_error"] < thresholds[j]
            assert bool(s["warm_start"] & 4) == (capped or (at_floor and (not kept or quiet))), (k, j, s)


def run_batch(g, batch, steps, shifts=None, **opts):
    g.load_scenes(batch, **opts)
    out = []
    for k in range(steps):
        if shifts is not None:
            g.translate_boundaries(shifts(k))
        out.append(g.step_scenes())
    return out


def compare_scene(g, k, alone, stats_batch, stats_alone, fields):
    for sb, sa in zip(stats_batch, stats_alone):
        for key in STAT_KEYS:
            assert sb[k][key] == sa[key], (k, key, sb[k][key], sa[key])
    g.select_scene(k)
    rb, ra = snapshot(g, fields), snapshot(alone, fields)
    for name in ra:
        np.testing.assert_array_equal(rb[name], ra[name], err_msg=f"scene {k} {name}")


@pytest.mark.parametrize("precision", ["ieee", "fast", "fp64"])
def test_batch_equals_scenes_alone(gpu, precision):
    prec = {"ieee": gpu.PRECISION_IEEE, "fast": gpu.PRECISION_FAST, "fp64": gpu.PRECISION_FP64}[precision]
    batch = mixed_scenes()
    print(f"case batch_equals_alone {precision} N={[s.fluid_count for s in batch]}", flush=True)
    g = gpu.Fluid(precision=prec)
    sb = run_batch(g, batch, 3, max_iterations=20)
    assert len({s["iterations"] for s in sb[0]}) > 1, "thresholds should end the loops at different iterations"
    fields = FLOAT_FIELDS + INT_FIELDS if precision != "fp64" else INT_FIELDS
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=prec)
        sa = run(alone, sc, 3, max_iterations=20)
        compare_scene(g, k, alone, sb, sa, fields)
        if precision == "fp64":
            g.select_scene(k)
            for f in FLOAT_FIELDS:
                np.testing.assert_array_equal(g.download_f64(f), alone.download_f64(f), err_msg=Field(f).name)


def test_batch_equals_scenes_alone_with_layout(gpu):
    """The same with the Morton layout applied at these sizes (work_order bit 3): a
    scene's buckets are laid out within its own range, so batch and alone still agree
    bit for bit."""
    batch = mixed_scenes()
    print("case batch_equals_alone with layout", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    sb = run_batch(g, batch, 3, max_iterations=20, work_order=3 | 8)
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        sa = run(alone, sc, 3, max_iterations=20, work_order=3 | 8)
        compare_scene(g, k, alone, sb, sa, FLOAT_FIELDS + INT_FIELDS)


@pytest.mark.parametrize("graph_iterations", [0, 4])
def test_batch_tail_on_running_scenes_only(gpu, graph_iterations):
    """Four scenes whose loops end at the minimum iteration and one that runs to the
    cap: from the first check on, more than half of the workgroups belong to ended
    loops, so the later iterations run over a block table of the running scene alone
    (launched one by one, after the graphs already enqueued).  Every scene still leaves
    the bits it leaves alone, and the long scene's loop runs to the cap."""
    slow = scenes.small_dam_break(seed=99)
    slow.threshold = 1e-6
    fast = []
    for k in range(4):
        sc = scenes.small_dam_break(seed=100 + k) if k % 2 == 0 else scenes.dam_break(
            (10, 10, 10), 0.02, 0.001, (0.4, 0.4, 0.3), seed=200 + k, name=f"dam_break_1000_{k}", steps=5)
        sc.threshold = 0.5
        fast.append(sc)
    batch = [fast[0], slow, fast[1], fast[2], fast[3]]
    print(f"case batch tail graph_iterations={graph_iterations}", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    sb = run_batch(g, batch, 2, max_iterations=12, graph_iterations=graph_iterations)
    assert [s["iterations"] for s in sb[0]] == [3, 12, 3, 3, 3], [s["iterations"] for s in sb[0]]
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        sa = run(alone, sc, 2, max_iterations=12, graph_iterations=graph_iterations)
        compare_scene(g, k, alone, sb, sa, FLOAT_FIELDS + INT_FIELDS)


def test_nonfinite_flag_per_scene(gpu):
    """A NaN velocity in one scene of a batch makes that scene's step non-finite (a NaN
    position takes cell 0, so every lookup stays in its tables) and flags that scene
    alone; the flags of all scenes are read back together after the step."""
    batch = mixed_scenes()
    print("case nonfinite flag per scene", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_FAST)
    g.load_scenes(batch, max_iterations=20)
    assert [s["nonfinite"] for s in g.step_scenes()] == [0, 0, 0]
    g.select_scene(1)
    v = g.download(Field.FLUID_VELOCITIES)
    v[0] = np.nan
    g.upload(Field.FLUID_VELOCITIES, v)
    assert [s["nonfinite"] for s in g.step_scenes()] == [0, 1, 0]


def test_batch_without_lists_and_with_empty_scene(gpu):
    a, b = small_scenes()
    empty = scenes.Scene(name="empty", fluid=np.zeros(0, np.float32), boundary=np.zeros(0, np.float32),
                         particle_radius=0.01)
    batch = [a, empty, b]
    print("case batch_empty_scene_no_lists", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    sb = run_batch(g, batch, 2, max_iterations=20, neighbor_skin=0.0)
    assert all(s[1]["iterations"] == 0 and s[1]["fluid_candidates"] == 0 for s in sb)
    for k in (0, 2):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        sa = run(alone, batch[k], 2, max_iterations=20, neighbor_skin=0.0)
        compare_scene(g, k, alone, sb, sa, FLOAT_FIELDS + INT_FIELDS)
    # the aggregate of pcisph_step: the largest iteration count of the batch
    agg = g.step()
    assert agg["iterations"] >= 1


@pytest.mark.parametrize("network", ["random", "trained"])
def test_batch_warm_start_equals_scenes_alone(gpu, network):
    """Warm-started scenes in a batch leave the bits they leave alone: with a random
    network (a positive output bias; every start it gives is rejected for the plain
    one here) and with the committed trained network, whose start every scene keeps in
    the first step -- so the inference's row ranges of the batch are covered."""
    batch = mixed_scenes()
    if network == "random":
        params = init_params(3)
        params[-1] = 0.5
    else:
        params = np.load(WEIGHTS)
    print(f"case batch_warm_start {network}", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.mlp_set_params(params)
    g.set_warm_start(True)
    sb = run_batch(g, batch, 3, max_iterations=20)
    assert_warm_rule(sb, [sc.threshold for sc in batch], max_iterations=20)
    if network == "trained":
        assert all(s["warm_start"] & 3 == 1 for s in sb[0]), sb[0]
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        alone.mlp_set_params(params)
        alone.set_warm_start(True)
        sa = run(alone, sc, 3, max_iterations=20)
        compare_scene(g, k, alone, sb, sa, FLOAT_FIELDS + INT_FIELDS)


def test_warm_start_left_out_after_capped_steps(gpu):
    """A scene whose step runs to the iteration cap without converging is left out of
    the next step's warm start: inside a warm-started batch it runs the plain step
    (plain first iteration, no network pressure force), alone its step has no warm
    start at all -- both must leave the same bits and statistics, and the other
    scenes follow the same rule."""
    batch = mixed_scenes()
    batch[2].threshold = 1e-9  # never converges: every step ends at the cap
    params = init_params(3)
    params[-1] = 0.5
    print("case warm start left out after capped steps", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.mlp_set_params(params)
    g.set_warm_start(True)
    sb = run_batch(g, batch, 3, max_iterations=12)
    assert [s[2]["warm_start"] & 1 for s in sb] == [1, 0, 0]
    assert_warm_rule(sb, [sc.threshold for sc in batch], max_iterations=12)
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        alone.mlp_set_params(params)
        alone.set_warm_start(True)
        sa = run(alone, sc, 3, max_iterations=12)
        for step_b, step_a in zip(sb, sa):
            for key in STAT_KEYS + ("first_density_error",):
                assert step_b[k][key] == step_a[key], (k, key, step_b[k][key], step_a[key])
        g.select_scene(k)
        rb, ra = snapshot(g, FLOAT_FIELDS + INT_FIELDS), snapshot(alone, FLOAT_FIELDS + INT_FIELDS)
        for name in ra:
            np.testing.assert_array_equal(rb[name], ra[name], err_msg=f"scene {k} {name}")


def test_batch_moving_walls(gpu):
    """Sloshing tanks: pcisph_translate_boundaries moves each scene's walls by its own
    shift in one launch; each scene equals the scene alone moved with
    pcisph_translate_boundary."""
    a, b = small_scenes()
    a = scenes.sloshing_tank(0.05, 2.0, a)
    b = scenes.sloshing_tank(0.08, 1.5, b)
    batch = [a, b]
    print("case batch_moving_walls", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    sb = run_batch(g, batch, 3, shifts=lambda k: [scenes.boundary_shift(sc, k) for sc in batch],
                   max_iterations=20)
    for k, sc in enumerate(batch):
        alone = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        alone.load_scene(sc, max_iterations=20)
        sa = []
        for step in range(3):
            alone.translate_boundary(*scenes.boundary_shift(sc, step))
            sa.append(alone.step())
        compare_scene(g, k, alone, sb, sa, FLOAT_FIELDS + INT_FIELDS)


def test_scene_count_change_keeps_scenes(gpu):
    a, b = small_scenes()
    print("case scene_count_change", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scenes([a, b], max_iterations=20)
    g.set_scene_count(3)
    assert g.scene_count() == 3
    for k, sc in enumerate([a, b]):
        g.select_scene(k)
        np.testing.assert_array_equal(g.download(Field.FLUID_POSITIONS), sc.fluid)
        np.testing.assert_array_equal(g.download(Field.BOUNDARY_POSITIONS), sc.boundary)
    g.select_scene(2)
    assert g.field_size(Field.FLUID_POSITIONS) == 0
    st = g.step_scenes()
    assert st[2]["iterations"] == 0 and st[0]["iterations"] >= 3 and st[1]["iterations"] >= 3
    with pytest.raises(Exception):
        g.select_scene(3)
