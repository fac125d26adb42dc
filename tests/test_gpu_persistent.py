"""GPU: the correction loop as one cooperative launch (solver option work_order bit 6).

The persistent kernel runs the same per-range code as the launch-per-pass loop, with
a grid barrier between the passes, so every field and every step statistic must be
identical to the launch-per-pass loop's -- in IEEE FP32, fast FP32 and FP64, with
the pressure pass walking lists or buckets or taken from the LDS staging, for single
scenes and for batches whose scenes end their loops at different iterations, and at
1M particles, where each workgroup takes several ranges of the block table.  Against
the oracle the IEEE FP32 loop is bit-exact as well.
"""
import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field
from tests.helpers import FLOAT_FIELDS, INT_FIELDS, run, small_scenes, snapshot

pytestmark = pytest.mark.gpu

PERSISTENT = 64
STAT_KEYS = ("iterations", "converged", "max_density_variation", "max_density_error", "fluid_candidates",
             "boundary_candidates", "warm_start", "list_waves", "first_density_error")


@pytest.fixture(scope="module")
def gpu():
    import pcisph_amd
    return pcisph_amd


def _mixed():
    a, b = small_scenes()
    c = scenes.scaled_sphere_drop(1500, seed=5, name="sphere_drop_small")
    b.threshold = 0.02
    c.threshold = 0.005
    return [a, b, c]


def _steps(g, steps, batch):
    return [g.step_scenes() if batch else [g.step()] for _ in range(steps)]


def _compare(ga, gb, sa, sb, count, fields):
    for step_a, step_b in zip(sa, sb):
        for k in range(count):
            for key in STAT_KEYS:
                assert step_a[k][key] == step_b[k][key], (k, key, step_a[k][key], step_b[k][key])
    for k in range(count):
        if count > 1:
            ga.select_scene(k)
            gb.select_scene(k)
        ra, rb = snapshot(ga, fields), snapshot(gb, fields)
        for name in ra:
            np.testing.assert_array_equal(ra[name], rb[name], err_msg=f"scene {k} {name}")


@pytest.mark.parametrize("precision", ["ieee", "fast", "fp64"])
@pytest.mark.parametrize("order", [3, 3 | 16])
@pytest.mark.parametrize("batch", [False, True])
def test_persistent_loop_equals_launch_per_pass(gpu, precision, order, batch):
    prec = {"ieee": gpu.PRECISION_IEEE, "fast": gpu.PRECISION_FAST, "fp64": gpu.PRECISION_FP64}[precision]
    work = _mixed() if batch else [small_scenes()[0]]
    print(f"case persistent {precision} work_order={order} batch={batch}", flush=True)
    runs = []
    for wo in (order, order | PERSISTENT):
        g = gpu.Fluid(precision=prec)
        if batch:
            g.load_scenes(work, max_iterations=20, work_order=wo)
        else:
            g.load_scene(work[0], max_iterations=20, work_order=wo)
        runs.append((g, _steps(g, 4, batch)))
    if batch:
        assert len({s["iterations"] for s in runs[0][1][0]}) > 1, "the scenes should end at different iterations"
    fields = FLOAT_FIELDS + INT_FIELDS if precision != "fp64" else INT_FIELDS
    _compare(runs[0][0], runs[1][0], runs[0][1], runs[1][1], len(work), fields)
    if precision == "fp64":
        for k in range(len(work)):
            if batch:
                runs[0][0].select_scene(k)
                runs[1][0].select_scene(k)
            for f in FLOAT_FIELDS:
                np.testing.assert_array_equal(runs[0][0].download_f64(f), runs[1][0].download_f64(f),
                                              err_msg=Field(f).name)


def test_persistent_loop_matches_oracle(gpu, oracle_lib):
    scene = small_scenes()[1]
    print(f"case persistent vs oracle {scene.name}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20)
    sg = run(g, scene, 3, max_iterations=20, work_order=3 | PERSISTENT)
    assert [s["iterations"] for s in so] == [s["iterations"] for s in sg]
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_persistent_loop_1m_several_ranges_per_workgroup(gpu):
    """The 1M dam break: ~5,800 ranges over the resident workgroups, two steps from
    rest at a cap of 8 iterations, IEEE FP32."""
    scene = scenes.dam_break_1m()
    print("case persistent 1M", flush=True)
    runs = []
    for wo in (3, 3 | PERSISTENT):
        g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        g.load_scene(scene, max_iterations=8, work_order=wo)
        runs.append((g, _steps(g, 2, False)))
    _compare(runs[0][0], runs[1][0], runs[0][1], runs[1][1], 1,
             [Field.FLUID_POSITIONS, Field.FLUID_VELOCITIES, Field.FLUID_PRESSURES, Field.FLUID_DENSITIES])
