"""GPU parity: the HIP step against the CPU oracle on the same seeded inputs.

IEEE FP32 and FP64 runs use the reference's order of operations and summation, so
they must reproduce the oracle bit for bit.  The fast FP32 run (fused multiply-adds,
hardware reciprocal/sqrt) must stay within a stated relative tolerance.  The sort
products (keys, permutation, cell table) and neighbour counts are integers and are
always compared bit for bit.
"""
import dataclasses

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field
from tests.helpers import (FLOAT_FIELDS, INT_FIELDS, max_rel_err, run, small_scenes, snapshot)

pytestmark = pytest.mark.gpu

# fast FP32 vs the FP32 oracle after one step: max abs difference of a field over the
# field's largest magnitude (velocities after one step are small differences of large
# forces, so a per-element relative test would measure cancellation, not the kernels)
FAST_RTOL = 2e-4


@pytest.fixture(scope="module")
def gpu():
    import pcisph_amd
    return pcisph_amd


@pytest.mark.parametrize("scene_idx", [0, 1])
def test_ieee_fp32_bit_exact(gpu, oracle_lib, scene_idx):
    scene = small_scenes()[scene_idx]
    print(f"case ieee_fp32 {scene.name} N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20)
    sg = run(g, scene, 3, max_iterations=20)
    for a, b in zip(so, sg):
        assert a["iterations"] == b["iterations"]
        assert a["max_density_variation"] == b["max_density_variation"]
        assert a["fluid_candidates"] == b["fluid_candidates"]
        assert a["boundary_candidates"] == b["boundary_candidates"]
    assert o.get_params().raw() == g.get_params().raw()
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_fp64_bit_exact(gpu, oracle_lib):
    scene = small_scenes()[0]
    print(f"case fp64 {scene.name}", flush=True)
    o = oracle_lib.OracleFluid(fp64=True)
    g = gpu.Fluid(precision=gpu.PRECISION_FP64)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20)
    assert [s["iterations"] for s in so] == [s["iterations"] for s in sg]
    for f in FLOAT_FIELDS:
        np.testing.assert_array_equal(g.download_f64(f), o.download_f64(f), err_msg=Field(f).name)
    for f in INT_FIELDS:
        np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)


@pytest.mark.parametrize("scene_idx", [0, 1])
def test_fast_fp32_within_tolerance(gpu, oracle_lib, scene_idx):
    """Fast FP32 after one step, per field: max |fast - oracle| / max |oracle| below
    FAST_RTOL.  Also reports both FP32 runs' error against the FP64 oracle."""
    scene = small_scenes()[scene_idx]
    print(f"case fast_fp32 {scene.name}", flush=True)
    o = oracle_lib.OracleFluid()
    o64 = oracle_lib.OracleFluid(fp64=True)
    g = gpu.Fluid()
    so = run(o, scene, 1, max_iterations=20)
    run(o64, scene, 1, max_iterations=20)
    sg = run(g, scene, 1, max_iterations=20)
    assert so[0]["iterations"] == sg[0]["iterations"]
    ro, rg = snapshot(o), snapshot(g)
    for f in INT_FIELDS:
        np.testing.assert_array_equal(rg[Field(f).name], ro[Field(f).name], err_msg=Field(f).name)
    errs = {}
    for f in FLOAT_FIELDS:
        name = Field(f).name
        ref = ro[name].astype(np.float64)
        r64 = o64.download_f64(f)
        scale = max(float(np.max(np.abs(r64))), 1e-30)
        errs[name] = (float(np.max(np.abs(rg[name] - ref))) / scale,
                      float(np.max(np.abs(rg[name] - r64))) / scale,
                      float(np.max(np.abs(ref - r64))) / scale)
        print(f"  {name:28s} fast-vs-ieee {errs[name][0]:.2e}  fast-vs-fp64 {errs[name][1]:.2e}  "
              f"ieee-vs-fp64 {errs[name][2]:.2e}", flush=True)
    bad = {k: v for k, v in errs.items() if not v[0] < FAST_RTOL}
    assert not bad, bad


def test_cell_box_edges_bit_exact(gpu, oracle_lib):
    """Neighbourhood lookups read a dense grid of bucket ranges inside each scene's
    cell box (set from the particles' extents when positions are uploaded) and the
    hash-indexed bucket table outside it.  Fluid far from its container (so the box
    covers the container alone), fluid just inside and just beyond the box's two-cell
    margin, and walls moved down out of the box after it was set must all match the
    oracle bit for bit."""
    base = small_scenes()[1]
    h = 4.0 * base.particle_radius
    fl = base.fluid.reshape(-1, 3)
    bd = base.boundary.reshape(-1, 3)
    lo, hi = bd.min(0), bd.max(0)
    mid = (lo + hi) / 2
    rng = np.random.default_rng(7)
    far = np.float32(40.0) + rng.uniform(0, h, (6, 3))
    near = np.array([hi[0] + 1.5 * h, mid[1], mid[2]]) + rng.uniform(0, 0.5 * h, (4, 3))
    beyond = np.array([hi[0] + 3.2 * h, mid[1], mid[2]]) + rng.uniform(0, 0.5 * h, (4, 3))
    fluid = np.concatenate([fl, far, near, beyond]).astype(np.float32).ravel()
    scene = dataclasses.replace(base, fluid=fluid, name="box_edges")
    print(f"case cell_box_edges N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 1, max_iterations=20)
    sg = run(g, scene, 1, max_iterations=20)
    g.translate_boundary(0.0, -5.0 * h, 0.0)  # walls leave the box; it is not re-measured
    o.upload(Field.BOUNDARY_POSITIONS, g.download(Field.BOUNDARY_POSITIONS))
    so += [o.step() for _ in range(2)]
    sg += [g.step() for _ in range(2)]
    for a, b in zip(so, sg):
        for key in ("iterations", "max_density_variation", "fluid_candidates", "boundary_candidates"):
            assert a[key] == b[key], (key, a[key], b[key])
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_iteration_bounds(gpu, oracle_lib):
    """min/max iteration semantics (Fluid.cpp:250-306) with the bounds as options."""
    scene = small_scenes()[0]
    for mn, mx in [(3, 7), (1, 1), (5, 5), (2, 50)]:
        print(f"case iterations {mn}-{mx}", flush=True)
        o = oracle_lib.OracleFluid()
        g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        so = run(o, scene, 2, min_iterations=mn, max_iterations=mx)
        sg = run(g, scene, 2, min_iterations=mn, max_iterations=mx)
        assert [s["iterations"] for s in sg] == [s["iterations"] for s in so]
        assert all(mn <= s["iterations"] <= mx or s["iterations"] == mx for s in sg)


@pytest.mark.parametrize("name,idx", [("small_dam_break_3steps", 0), ("dam_break_1000_2steps", 1)])
def test_ieee_matches_golden(gpu, name, idx):
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    scene = small_scenes()[idx]
    print(f"case golden {name}", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    st = run(g, scene, len(ref["iterations"]), max_iterations=20)
    assert [s["iterations"] for s in st] == list(ref["iterations"])
    assert g.get_params().raw() == ref["params"].tobytes()
    for k in ref.files:
        if k in ("iterations", "max_density_variation", "params"):
            continue
        np.testing.assert_array_equal(g.download(Field[k]), ref[k], err_msg=k)


@pytest.mark.parametrize("b3", [-100.0, 100.0])
def test_warm_start_fallback_is_the_plain_loop(gpu, oracle_lib, b3):
    """A network whose start is no better than the plain one (a huge negative change:
    every start clamps to zero) or much worse (a huge uniform rise) is rejected at the
    first iteration, and the step is then exactly the plain step (IEEE FP32, bit for
    bit against the oracle)."""
    from oracle import mlp_oracle as M
    scene = small_scenes()[0]
    print(f"case warm fallback b3={b3}", flush=True)
    p = np.zeros(M.N_PARAMS, np.float32)
    p[-1] = b3
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scene(scene, max_iterations=20)
    g.mlp_set_params(p)
    g.set_warm_start(True)
    o = oracle_lib.OracleFluid()
    o.load_scene(scene, max_iterations=20)
    for k in range(3):
        a, b = g.step(), o.step()
        # the network's start is evaluated and rejected, unless the step before left
        # the scene out (bit 2: it converged at the minimum with the plain start)
        assert a["warm_start"] & 3 == (3 if k == 0 or not prev & 4 else 0), a
        assert a["iterations"] == b["iterations"]
        prev = a["warm_start"]
    for f in (Field.FLUID_POSITIONS, Field.FLUID_VELOCITIES, Field.FLUID_PRESSURES, Field.FLUID_DENSITIES):
        np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)


def test_sample_collection(gpu):
    scene = small_scenes()[0]
    g = gpu.Fluid()
    g.load_scene(scene, max_iterations=20)
    g.samples_reserve(10000)
    g.samples_collect(4)
    for _ in range(3):
        g.step()
    x, t = g.samples_download()
    assert x.shape == (3 * 128, 8) and np.all(np.isfinite(x)) and np.all(np.isfinite(t))
    # targets are the change from the previous pressure (input 2): never below -input 2
    assert np.all(t + x[:, 2] >= -1e-6)
    # feature 1 (fluid neighbours / 32) and 6 (boundary neighbours / 32) are counts
    assert np.allclose(x[:, 1] * 32, np.round(x[:, 1] * 32)) and np.allclose(x[:, 6] * 32, np.round(x[:, 6] * 32))


@pytest.mark.parametrize("skin", [0.0, 1e-4, 0.2, 0.6])
def test_neighbour_lists_are_exact(gpu, oracle_lib, skin):
    """Neighbour lists off (0), on but too tight for most of the pressure passes (1e-4:
    the predicted positions move further than the skin allows, so nearly every
    wavefront walks its buckets), the default, and a wide skin (pressure passes walk
    the lists): the IEEE step is the oracle's, bit for bit."""
    scene = small_scenes()[1]
    print(f"case lists skin={skin}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20, neighbor_skin=skin)
    assert [s["iterations"] for s in so] == [s["iterations"] for s in sg]
    for f in FLOAT_FIELDS:
        np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)
    waves = sum(s["list_waves"] for s in sg)
    if skin == 0.0:
        assert waves == 0
    elif skin >= 0.2:
        assert waves > 0


def test_list_overflow_falls_back_to_buckets(gpu, oracle_lib):
    """A block packed at a quarter of the rest spacing has ~64x the neighbours of the
    list capacity allows: those particles walk their buckets, results unchanged."""
    from pcisph_amd import scenes
    pts = scenes.lattice(8, 8, 8, 0.005, origin=(0.1, 0.1, 0.1))
    scene = scenes.Scene(name="packed", fluid=pts.astype(np.float32).ravel(),
                         boundary=np.zeros(0, np.float32), particle_radius=0.01, gravity=0.0)
    print("case list overflow", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 1, max_iterations=4)
    sg = run(g, scene, 1, max_iterations=4)
    assert so[0]["iterations"] == sg[0]["iterations"]
    for f in (Field.FLUID_DENSITIES, Field.FLUID_POSITIONS, Field.FLUID_VELOCITIES, Field.FLUID_PRESSURES):
        np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)


@pytest.mark.parametrize("tank", ["small", "wide"])
@pytest.mark.parametrize("fp64", [False, True])
def test_far_collision_pruning_bit_exact(gpu, oracle_lib, fp64, tank):
    """Boundary buckets that hold only hash collisions from at least eight cells away
    read as empty for walks that cannot reach them (512 fluid particles, 256 buckets).
    In the small tank (12.5 x 11 x 10 cells, 6558 wall particles) few buckets qualify;
    in the wide one (30 x 15 x 30 cells) most boundary lookups hit such buckets.
    Velocities up to 120 m/s (small) or 500 m/s (wide) put the predicted positions
    anywhere from 0 to 1.5 h or 6 h from x -- in the wide tank on both sides of the
    reach limit (~6 cells), so both the pruned and the full walks run in the same
    pressure passes; every field stays the oracle's bit for bit."""
    if tank == "small":
        scene, vmax = small_scenes()[0], 120.0
    else:
        scene = scenes.dam_break((8, 8, 8), 0.02, 0.001, (1.2, 0.6, 1.2), seed=1234, name="wide_tank", steps=2)
        vmax = 500.0
    print(f"case far collision pruning {tank} fp64={fp64}", flush=True)
    o = oracle_lib.OracleFluid(fp64=fp64)
    g = gpu.Fluid(precision=gpu.PRECISION_FP64 if fp64 else gpu.PRECISION_IEEE)
    o.load_scene(scene, max_iterations=20)
    g.load_scene(scene, max_iterations=20)
    rng = np.random.default_rng(7)
    v = rng.uniform(-vmax, vmax, 3 * scene.fluid_count).astype(np.float32)
    o.upload(Field.FLUID_VELOCITIES, v)
    g.upload(Field.FLUID_VELOCITIES, v)
    for _ in range(2):
        a, b = o.step(), g.step()
        assert a["iterations"] == b["iterations"]
        assert a["max_density_variation"] == b["max_density_variation"]
        assert a["boundary_candidates"] == b["boundary_candidates"]
    for f in FLOAT_FIELDS:
        if fp64:
            np.testing.assert_array_equal(g.download_f64(f), o.download_f64(f), err_msg=Field(f).name)
        else:
            np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)


@pytest.mark.parametrize("work_order", [3 | 8, 3 | 8 | 16])
def test_particles_leaving_the_box_bit_exact(gpu, oracle_lib, work_order):
    """Particles thrown clear of the scene's cell box (measured at upload): velocities
    up to 500 m/s carry most of the 512 particles of the wide tank tens of cells out
    of it within the first step, so from the second step on the storage order (Morton
    layout forced on) and the staged pass's work order (forced on) place them after
    the particles inside, by the bit above their Morton code; four steps stay the
    oracle's bit for bit."""
    scene = scenes.dam_break((8, 8, 8), 0.02, 0.001, (1.2, 0.6, 1.2), seed=1234, name="wide_tank", steps=4)
    print(f"case particles leaving the box work_order={work_order}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    o.load_scene(scene, max_iterations=20)
    g.load_scene(scene, max_iterations=20, work_order=work_order)
    rng = np.random.default_rng(11)
    v = rng.uniform(-500.0, 500.0, 3 * scene.fluid_count).astype(np.float32)
    o.upload(Field.FLUID_VELOCITIES, v)
    g.upload(Field.FLUID_VELOCITIES, v)
    for _ in range(4):
        a, b = o.step(), g.step()
        assert a["iterations"] == b["iterations"]
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_far_cells_unflagged_bit_exact(gpu, oracle_lib):
    """The small dam break moved 2000 m along x: cell coordinates near 50,000, where
    far-collision flags are never set (the rounding of such coordinates is not covered
    by the reach margin), so every walk reads its buckets in full; the oracle's bits
    still hold."""
    scene = small_scenes()[0]
    shift = np.array([2000.0, 0.0, 0.0], np.float32)
    scene.fluid = (scene.fluid.reshape(-1, 3) + shift).astype(np.float32).ravel()
    scene.boundary = (scene.boundary.reshape(-1, 3) + shift).astype(np.float32).ravel()
    print("case far cells unflagged", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20)
    for a, b in zip(so, sg):
        assert a["iterations"] == b["iterations"]
        assert a["boundary_candidates"] == b["boundary_candidates"]
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


@pytest.mark.parametrize("case", ["free_fall", "one_particle"])
@pytest.mark.parametrize("fp64", [False, True])
def test_small_and_wall_free_scenes_bit_exact(gpu, oracle_lib, case, fp64):
    """Scenes at the edges of the inputs: a falling block with no boundary particles at
    all (every boundary walk, list and table empty) and a single particle (one bucket,
    one candidate: itself), with the Morton layout forced on; every field stays the
    oracle's bit for bit."""
    from pcisph_amd import scenes
    if case == "free_fall":
        pts = scenes.lattice(10, 10, 10, 0.02, origin=(0.1, 0.3, 0.1)).astype(np.float32).ravel()
    else:
        pts = np.array([0.2, 0.3, 0.2], np.float32)
    scene = scenes.Scene(name=case, fluid=pts, boundary=np.zeros(0, np.float32), particle_radius=0.01)
    print(f"case {case} fp64={fp64}", flush=True)
    o = oracle_lib.OracleFluid(fp64=fp64)
    g = gpu.Fluid(precision=gpu.PRECISION_FP64 if fp64 else gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20)
    sg = run(g, scene, 3, max_iterations=20, work_order=3 | 8)
    for a, b in zip(so, sg):
        assert a["iterations"] == b["iterations"]
        assert a["fluid_candidates"] == b["fluid_candidates"]
    for f in FLOAT_FIELDS:
        if fp64:
            np.testing.assert_array_equal(g.download_f64(f), o.download_f64(f), err_msg=Field(f).name)
        else:
            np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)
    for f in INT_FIELDS:
        np.testing.assert_array_equal(g.download(f), o.download(f), err_msg=Field(f).name)


def test_list_radius_adapts_bit_exact(gpu, oracle_lib):
    """Particles thrown at 10 m/s move further per step than a 0.05 h skin allows, so
    the pressure passes walk their buckets; the scene then builds near-h lists with no
    displacement tracking, and every 32 steps probes the full skin.  Results stay the
    oracle's bit for bit through both switches."""
    scene = small_scenes()[1]
    print("case list radius adapts", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    o.load_scene(scene, max_iterations=20)
    g.load_scene(scene, max_iterations=20, neighbor_skin=0.05)
    v = np.tile(np.array([0.0, -10.0, 0.0], np.float32), scene.fluid_count)
    o.upload(Field.FLUID_VELOCITIES, v)
    g.upload(Field.FLUID_VELOCITIES, v)
    waves = []
    for _ in range(35):
        a, b = o.step(), g.step()
        assert a["iterations"] == b["iterations"]
        assert a["max_density_variation"] == b["max_density_variation"]
        waves.append(b["list_waves"])
    # steps 2..33 on near lists: no list tests in the pressure passes
    assert all(w == 0 for w in waves[1:33]), waves
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_sampled_stage_profile(gpu):
    """Stage timing of every n-th step: a timed step launches kernel by kernel, the
    others replay their graphs; the density pass runs once per step."""
    scene = small_scenes()[1]
    print("case sampled stage profile", flush=True)
    g = gpu.Fluid()
    g.load_scene(scene, max_iterations=20)
    g.step()
    for every, steps, timed in ((1, 4, 4), (3, 7, 3)):
        g.profile_enable(True, every=every)
        for _ in range(steps):
            g.step()
        t = g.profile_read()
        g.profile_enable(False)
        density = 4  # PCISPH_STAGE_DENSITY
        assert t["launches"][density] == timed, (every, t["launches"])
        assert t["ms"][density] > 0


# bit 3 applies the Morton layout (bit 0) below its size threshold, as these small
# scenes need to exercise it
@pytest.mark.parametrize("work_order", [0, 2, 4, 3 | 8, 7 | 8])
def test_work_orders_are_exact(gpu, oracle_lib, work_order):
    """Every storage order and placement (the reference's order, XCD placement,
    LDS-staged pressure passes, all together with the Morton layout) leaves the
    oracle's bits."""
    scene = small_scenes()[1]
    print(f"case work_order={work_order}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    # the LDS-tiled pressure passes (bit 2) run with 256-thread workgroups only
    sg = run(g, scene, 2, max_iterations=20, work_order=work_order, block_size=256 if work_order & 4 else 0)
    assert [s["iterations"] for s in so] == [s["iterations"] for s in sg]
    ro, rg = snapshot(o), snapshot(g)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=name)


def test_layout_transitions_bit_exact(gpu, oracle_lib):
    """With the Morton layout the fluid arrays stay out of the reference's order
    between steps (downloads and sort products are translated).  Switching the
    layout off and on mid-run, uploading a fluid field while the arrays are in a
    layout, and a snapshot restored after further steps must all leave the oracle's
    bits."""
    scene = small_scenes()[1]
    print("case layout_transitions", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    run(o, scene, 2, max_iterations=20)
    run(g, scene, 2, max_iterations=20, work_order=3 | 8)  # the layout at this size

    def step_both(k):
        for _ in range(k):
            a, b = o.step(), g.step()
            assert a["iterations"] == b["iterations"]
            assert a["max_density_variation"] == b["max_density_variation"]

    def same(what):
        ro, rg = snapshot(o), snapshot(g)
        for name in ro:
            np.testing.assert_array_equal(rg[name], ro[name], err_msg=f"{what}: {name}")

    g.set_solver_options(max_iterations=20, work_order=0)
    step_both(1)
    same("layout off")
    g.set_solver_options(max_iterations=20, work_order=3 | 8)
    step_both(2)
    same("layout on again")
    v = o.download(Field.FLUID_VELOCITIES) * np.float32(0.5)
    o.upload(Field.FLUID_VELOCITIES, v)
    g.upload(Field.FLUID_VELOCITIES, v)
    step_both(1)
    same("upload into a layout")
    g.snapshot()
    for _ in range(2):
        g.step()
    g.restore()
    step_both(1)
    same("restore")


def test_fast_mode_above_one_metre_runs_ieee(gpu, oracle_lib):
    """The fast kernels clamp the poly6 term to [0, 1] in one instruction, which is
    max(0, h^2 - r^2) only while h^2 <= 1; a scene with a kernel radius above 1 m (here
    1.2 m: the 8^3 dam break scaled by 30) runs the IEEE FP32 kernels in fast mode, so
    its steps equal an IEEE run's and the oracle's bit for bit."""
    scene = scenes.dam_break((8, 8, 8), 0.6, 0.03, (12.0, 12.0, 9.0), seed=1234, name="dam_break_x30", steps=2)
    print(f"case fast_above_1m {scene.name} N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    fast, ieee = gpu.Fluid(), gpu.Fluid(precision=gpu.PRECISION_IEEE)
    o = oracle_lib.OracleFluid()
    sf = run(fast, scene, 2, max_iterations=20)
    si = run(ieee, scene, 2, max_iterations=20)
    so = run(o, scene, 2, max_iterations=20)
    assert fast.get_params().kernel_radius > 1.0
    assert [s["iterations"] for s in sf] == [s["iterations"] for s in si] == [s["iterations"] for s in so]
    rf, ri, ro = snapshot(fast), snapshot(ieee), snapshot(o)
    for name in ri:
        np.testing.assert_array_equal(rf[name], ri[name], err_msg=name)
        np.testing.assert_array_equal(rf[name], ro[name], err_msg=name)
    # and a scene at h <= 1 m does run the fast kernels: its fields differ from IEEE
    small = small_scenes()[0]
    run(fast, small, 1, max_iterations=20)
    run(ieee, small, 1, max_iterations=20)
    assert not np.array_equal(fast.download(Field.FLUID_DENSITIES), ieee.download(Field.FLUID_DENSITIES))
