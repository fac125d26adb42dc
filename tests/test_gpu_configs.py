"""GPU parity of the benchmark configurations against the CPU oracle.

  * the 1,048,576-particle dam break: IEEE FP32 bit-exact for two steps (three
    correction iterations each), fast FP32 within FAST_RTOL after one;
  * the 16,384-particle dam break: IEEE FP32 bit-exact over 20 steps, iteration count
    and max density variation of every step included;
  * multi-scene batches drawn from the 64-scene and 256-scene batch generators
    (dam breaks, sphere drops, sloshing tanks with per-scene wall motion): every scene
    of the batch bit-exact against its own oracle run;
  * the sloshing tank at the 1M resolution: a thin slab of it bit-exact against the
    oracle with the walls moved every step, and the full-size scene through
    size-independent properties (sort products, wall positions, finiteness);
  * the boundary-boundary density pass (pcisph.cl:195-210, never launched by the
    reference, Fluid.cpp:166-171) switched on;
  * radix digit widths 4..8 (the sort's products are a stable sort whatever the width),
    for the default reduce-then-scan sort and the one-sweep variant (work_order bit 7).
"""
import os

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field
from tests.helpers import FLOAT_FIELDS, INT_FIELDS, run, small_scenes, snapshot
from tests.test_gpu_large import check_sort_products

pytestmark = pytest.mark.gpu

FAST_RTOL = 2e-4  # as in test_gpu_parity: max |fast - oracle| / max |oracle| per field
STAT_KEYS = ("iterations", "converged", "max_density_variation", "fluid_candidates", "boundary_candidates")
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def gpu():
    import pcisph_amd
    return pcisph_amd


def assert_same_state(g, o, fields=None, tag=""):
    rg, ro = snapshot(g, fields), snapshot(o, fields)
    for name in ro:
        np.testing.assert_array_equal(rg[name], ro[name], err_msg=f"{tag} {name}")


def assert_same_stats(sg, so, tag=""):
    for k, (a, b) in enumerate(zip(sg, so)):
        for key in STAT_KEYS:
            assert a[key] == b[key], (tag, k, key, a[key], b[key])


# ---------------------------------------------------------------- 1M dam break
@pytest.fixture(scope="module")
def oracle_1m(oracle_lib):
    """The oracle's state after steps 1 and 2 of the 1M dam break (3 iterations each)."""
    oracle_lib.set_threads(THREADS)
    scene = scenes.dam_break_1m()
    print(f"oracle 1m N={scene.fluid_count} Nb={scene.boundary_count} threads={THREADS}", flush=True)
    o = oracle_lib.OracleFluid()
    o.load_scene(scene, min_iterations=3, max_iterations=3)
    out = []
    for _ in range(2):
        st = o.step()
        out.append((st, snapshot(o)))
    return scene, out


def test_1m_ieee_two_steps_bit_exact(gpu, oracle_1m):
    scene, ref = oracle_1m
    print("case 1m ieee bit-exact, 2 steps", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scene(scene, min_iterations=3, max_iterations=3)
    for k, (st_o, snap_o) in enumerate(ref):
        st_g = g.step()
        assert_same_stats([st_g], [st_o], f"step {k + 1}")
        snap_g = snapshot(g)
        for name in snap_o:
            np.testing.assert_array_equal(snap_g[name], snap_o[name], err_msg=f"step {k + 1} {name}")
    g.close()


# the fields whose FP32 tolerance is stated for the step's results: fast FP32 within
# FAST_RTOL of the IEEE FP32 oracle; the other float fields (forces, normals, ...) within
# 4x the IEEE FP32 run's own error against FP64, both measured against an FP64 run
RESULT_FIELDS = (Field.FLUID_POSITIONS, Field.FLUID_VELOCITIES, Field.FLUID_DENSITIES, Field.FLUID_PRESSURES,
                 Field.FLUID_PREDICTED_POSITIONS)


def test_1m_fast_within_tolerance(gpu, oracle_1m):
    """Fast FP32 after one step of the 1M dam break against the IEEE FP32 oracle and
    against the FP64 GPU step (bit-exact with the FP64 oracle on the small scenes of
    test_gpu_parity), whose errors are reported per field."""
    scene, ref = oracle_1m
    print("case 1m fast fp32, 1 step", flush=True)
    st_o, snap_o = ref[0]
    runs = {}
    for name, prec in (("fast", gpu.PRECISION_FAST), ("fp64", gpu.PRECISION_FP64)):
        g = gpu.Fluid(precision=prec)
        g.load_scene(scene, min_iterations=3, max_iterations=3)
        st = g.step()
        assert st["iterations"] == st_o["iterations"]
        runs[name] = ({Field(f).name: g.download_f64(f) for f in FLOAT_FIELDS}, snapshot(g, INT_FIELDS))
        g.close()
    fast, ints = runs["fast"]
    f64 = runs["fp64"][0]
    for f in INT_FIELDS:
        name = Field(f).name
        np.testing.assert_array_equal(ints[name], snap_o[name], err_msg=name)
        a, b = runs["fp64"][1][name], snap_o[name]
        if f == Field.FLUID_NEIGHBOR_COUNTS:  # within-h counts depend on the precision of r^2
            a, b = a.reshape(-1, 4)[:, :2], b.reshape(-1, 4)[:, :2]
        np.testing.assert_array_equal(a, b, err_msg=f"fp64 {name}")
    for f in FLOAT_FIELDS:
        name = Field(f).name
        ieee = snap_o[name].astype(np.float64)
        scale = max(float(np.abs(f64[name]).max()), 1e-30)
        e_fi = float(np.abs(fast[name] - ieee).max()) / scale
        e_f64 = float(np.abs(fast[name] - f64[name]).max()) / scale
        e_i64 = float(np.abs(ieee - f64[name]).max()) / scale
        print(f"  {name:28s} fast-vs-ieee {e_fi:.2e}  fast-vs-fp64 {e_f64:.2e}  ieee-vs-fp64 {e_i64:.2e}",
              flush=True)
        if f in RESULT_FIELDS:
            assert e_fi < FAST_RTOL, (name, e_fi)
        else:
            assert e_f64 <= max(FAST_RTOL, 4.0 * e_i64), (name, e_f64, e_i64)


# ---------------------------------------------------------------- 16k dam break
def test_16k_twenty_steps_bit_exact(gpu, oracle_lib):
    oracle_lib.set_threads(THREADS)
    scene = scenes.dam_break_16k()
    print(f"case 16k ieee 20 steps N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 20)
    sg = run(g, scene, 20)
    print(f"  iterations per step {[s['iterations'] for s in sg]}", flush=True)
    assert_same_stats(sg, so, "16k")
    assert_same_state(g, o, tag="16k step 20")


# ---------------------------------------------------------------- batches
def _batch_subset(which):
    """Scenes of the real batch generators, the smaller ones (the oracle steps each)."""
    if which == "batch_64":
        full = scenes.batch_64()
        return [full[i] for i in (1, 2, 4, 5, 7, 9)]
    full = scenes.batch_256()
    return [full[i] for i in (0, 1, 2, 4, 6, 7, 8, 9, 10, 11)]


@pytest.mark.parametrize("which", ["batch_64", "batch_256"])
def test_batch_scenes_match_oracle(gpu, oracle_lib, which):
    """Each scene of a batch (one launch per stage for all of them, sloshing walls moved
    every step in one launch) equals the oracle stepping that scene alone."""
    oracle_lib.set_threads(THREADS)
    batch = _batch_subset(which)
    steps = 3
    print(f"case {which} subset {[s.name for s in batch]} N={[s.fluid_count for s in batch]}", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scenes(batch)
    sloshing = any(sc.meta.get("slosh") for sc in batch)
    sb = []
    for k in range(steps):
        if sloshing:
            g.translate_boundaries([scenes.boundary_shift(sc, k) for sc in batch])
        sb.append(g.step_scenes())
    assert len({s["iterations"] for s in sb[0]}) > 1, "the scenes' loops should end at different iterations"
    for i, sc in enumerate(batch):
        o = oracle_lib.OracleFluid()
        o.load_scene(sc)
        so = []
        for k in range(steps):
            if sloshing and k > 0:
                # a batch re-sorts every scene's walls when any scene's moved: the walls as
                # the oracle holds them (sorted), moved in float32 for a sloshing scene,
                # re-uploaded unchanged otherwise
                shift = np.float32(scenes.boundary_shift(sc, k))
                b = o.download(Field.BOUNDARY_POSITIONS).reshape(-1, 3)
                o.upload(Field.BOUNDARY_POSITIONS, (b + shift).astype(np.float32))
            elif sc.meta.get("slosh"):
                shift = np.float32(scenes.boundary_shift(sc, k))
                o.upload(Field.BOUNDARY_POSITIONS, (sc.boundary.reshape(-1, 3) + shift).astype(np.float32))
            so.append(o.step())
        assert_same_stats([s[i] for s in sb], so, sc.name)
        g.select_scene(i)
        assert_same_state(g, o, tag=sc.name)
    g.close()


# ---------------------------------------------------------------- sloshing tank
def _slab(base_seed=20240611):
    """The sloshing tank at the 1M resolution (spacing 0.01 m, h = 0.02 m, the 2 x 2 m
    tank oscillating 0.5 m at 0.8 Hz along x) cut to a slab four particles deep."""
    base = scenes.dam_break((128, 128, 4), 0.01, 0.0005, (2.0, 2.0, 0.04), seed=base_seed, name="slab")
    return scenes.sloshing_tank(0.5, 0.8, base, name="sloshing_slab")


def test_sloshing_slab_bit_exact(gpu, oracle_lib):
    oracle_lib.set_threads(THREADS)
    scene = _slab()
    steps = 4
    print(f"case sloshing slab N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    o.load_scene(scene)
    g.load_scene(scene)
    walls = scene.boundary.reshape(-1, 3).copy()
    so, sg = [], []
    for k in range(steps):
        shift = np.float32(scenes.boundary_shift(scene, k))
        g.translate_boundary(*[float(v) for v in shift])
        walls = walls if k == 0 else o.download(Field.BOUNDARY_POSITIONS).reshape(-1, 3)
        moved = (walls + shift).astype(np.float32)
        o.upload(Field.BOUNDARY_POSITIONS, moved)
        so.append(o.step())
        sg.append(g.step())
    assert abs(float(np.float32(scenes.boundary_shift(scene, 1))[0])) > 0
    assert_same_stats(sg, so, "slab")
    assert_same_state(g, o, tag="slab")


def test_sloshing_1m_properties(gpu):
    """Full-size sloshing tank, 8 steps with the walls moved before each: the walls are
    the initial walls plus the float32 sum of the shifts, re-sorted; the last step's
    sort products are the stable sort of the hash keys of the positions it started
    from; nothing is non-finite."""
    scene = scenes.sloshing_tank(0.5, 0.8)
    steps = 8
    print(f"case sloshing 1m properties N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    g = gpu.Fluid()
    g.load_scene(scene)
    walls = scene.boundary.reshape(-1, 3).astype(np.float32)
    before = None
    for k in range(steps):
        shift = np.float32(scenes.boundary_shift(scene, k))
        g.translate_boundary(*[float(v) for v in shift])
        walls = (walls + shift).astype(np.float32)
        if k == steps - 1:
            before = g.download(Field.FLUID_POSITIONS)
        st = g.step()
        assert st["nonfinite"] == 0, (k, st)
    got = g.download(Field.BOUNDARY_POSITIONS).reshape(-1, 3)
    order_g = np.lexsort(got.T[::-1])
    order_e = np.lexsort(walls.T[::-1])
    np.testing.assert_array_equal(got[order_g], walls[order_e])
    p = g.get_params()
    check_sort_products(g.download(Field.FLUID_KEYS), g.download(Field.FLUID_SRC_LOCATIONS),
                        g.download(Field.FLUID_CELL_OFFSETS), before, p.cell_size, p.bucket_count)
    g.close()


# ---------------------------------------------------------------- options
@pytest.mark.parametrize("scene_idx", [0, 1])
def test_boundary_init_pass_bit_exact(gpu, oracle_lib, scene_idx):
    scene = small_scenes()[scene_idx]
    print(f"case launch_boundary_init {scene.name}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20, launch_boundary_init=1)
    sg = run(g, scene, 3, max_iterations=20, launch_boundary_init=1)
    assert_same_stats(sg, so, "binit")
    init = g.download(Field.BOUNDARY_INIT_PRED_DENSITIES)
    assert np.count_nonzero(init) == init.size
    assert_same_state(g, o, FLOAT_FIELDS + INT_FIELDS + [Field.BOUNDARY_INIT_PRED_DENSITIES], tag="binit")


@pytest.mark.parametrize("digit", [4, 5, 6, 7, 8])
def test_sort_digit_widths_bit_exact(gpu, oracle_lib, digit):
    scene = small_scenes()[1]
    print(f"case sort_digit_bits {digit}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20, sort_digit_bits=digit)
    assert_same_stats(sg, so, f"digit {digit}")
    assert_same_state(g, o, tag=f"digit {digit}")


@pytest.mark.parametrize("digit", [5, 6, 7])
def test_1m_sort_digit_widths(gpu, digit):
    scene = scenes.dam_break_1m()
    print(f"case 1m sort_digit_bits {digit}", flush=True)
    g = gpu.Fluid()
    g.load_scene(scene, min_iterations=3, max_iterations=3, sort_digit_bits=digit)
    g.step()
    p = g.get_params()
    check_sort_products(g.download(Field.FLUID_KEYS), g.download(Field.FLUID_SRC_LOCATIONS),
                        g.download(Field.FLUID_CELL_OFFSETS), scene.fluid, p.cell_size, p.bucket_count)
    check_sort_products(g.download(Field.BOUNDARY_KEYS), g.download(Field.BOUNDARY_SRC_LOCATIONS),
                        g.download(Field.BOUNDARY_CELL_OFFSETS), scene.boundary, p.cell_size, p.bucket_count)
    g.close()


ONESWEEP = 3 | 128  # default work order, the step's radix sorts as one-sweep


@pytest.mark.parametrize("digit", [4, 5, 6, 7, 8])
def test_onesweep_sort_bit_exact(gpu, oracle_lib, digit):
    """The one-sweep sort (decoupled look-back over tile prefixes) gives the oracle's
    stable sort, hence every field, at every digit width."""
    scene = small_scenes()[1]
    print(f"case one-sweep sort_digit_bits {digit}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20, sort_digit_bits=digit, work_order=ONESWEEP)
    assert_same_stats(sg, so, f"one-sweep digit {digit}")
    assert_same_state(g, o, tag=f"one-sweep digit {digit}")


@pytest.mark.parametrize("digit", [6, 8])
def test_1m_onesweep_sort_products(gpu, digit):
    """At 1M (512 tiles, all resident, each waiting on the look-back of the ones
    before it) the sort products stay the stable sort of the hash keys, and the step
    equals the reduce-then-scan step bit for bit."""
    scene = scenes.dam_break_1m()
    print(f"case 1m one-sweep sort_digit_bits {digit}", flush=True)
    runs = []
    for wo in (3, ONESWEEP):
        g = gpu.Fluid()
        g.load_scene(scene, min_iterations=3, max_iterations=3, sort_digit_bits=digit, work_order=wo)
        runs.append((g, [g.step() for _ in range(2)]))
    g = runs[1][0]
    p = g.get_params()
    keys, src = g.download(Field.FLUID_KEYS), g.download(Field.FLUID_SRC_LOCATIONS)
    assert np.array_equal(np.sort(src), np.arange(src.size, dtype=np.uint32))
    bkeys, bsrc = g.download(Field.BOUNDARY_KEYS), g.download(Field.BOUNDARY_SRC_LOCATIONS)
    check_sort_products(bkeys, bsrc, g.download(Field.BOUNDARY_CELL_OFFSETS), scene.boundary, p.cell_size,
                        p.bucket_count)
    assert_same_stats(runs[1][1], runs[0][1], "1m one-sweep")
    assert_same_state(runs[1][0], runs[0][0], tag="1m one-sweep")
    for r in runs:
        r[0].close()


# ---------------------------------------------------------------- autotuner
def test_autotune_keeps_state_and_parity(gpu, oracle_lib):
    """The autotuner's trials leave the state as it was: steps after the call equal the
    oracle's bit for bit, and its chosen options are among the ones it tries."""
    scene = small_scenes()[1]
    print("case autotune parity", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20)
    chosen = g.autotune(trial_steps=2)
    print(f"  chosen {chosen}", flush=True)
    assert chosen["block_size"] in (64, 128, 256, 512) and chosen["sort_digit_bits"] in (5, 6, 7, 8)
    assert chosen["work_order"] & ~48 == 3 and chosen["work_order"] & 48 in (0, 16, 32)
    so += [o.step() for _ in range(2)]
    sg += [g.step() for _ in range(2)]
    assert_same_stats(sg, so, "autotune")
    assert_same_state(g, o, tag="autotune")


def test_autotune_restores_warm_start_eligibility(gpu):
    """With the warm start on and every step ending at the iteration cap (a scene whose
    last step hit the cap starts its next step plain), a run with an autotuner call in
    the middle equals a run without it that uses the chosen options."""
    from pcisph_amd.warm_start import zero_params
    scene = small_scenes()[0]
    print("case autotune warm-start eligibility", flush=True)

    def make():
        f = gpu.Fluid(precision=gpu.PRECISION_IEEE)
        f.load_scene(scene, min_iterations=2, max_iterations=2)
        f.mlp_set_params(zero_params())
        f.set_warm_start(True)
        return f

    a = make()
    sa = [a.step() for _ in range(2)]
    chosen = a.autotune(trial_steps=2)
    sa += [a.step() for _ in range(3)]
    b = make()
    sb = [b.step() for _ in range(2)]
    b.set_solver_options(min_iterations=2, max_iterations=2, block_size=chosen["block_size"],
                         sort_digit_bits=chosen["sort_digit_bits"], work_order=chosen["work_order"])
    sb += [b.step() for _ in range(3)]
    assert [s["converged"] for s in sa[:2]] == [0, 0]
    for x, y in zip(sa, sb):
        for key in STAT_KEYS + ("warm_start",):
            assert x[key] == y[key], (key, x[key], y[key])
    assert_same_state(a, b, tag="autotune warm")


# ---------------------------------------------------------------- staged pressure pass
STAGED = 3 | 16  # default work order plus the LDS-staged FP32 pressure pass in every step
STAGED_OFF = 3 | 32  # default work order, the LDS staging off


@pytest.mark.parametrize("scene_idx", [0, 1])
@pytest.mark.parametrize("layout", [0, 8])
def test_staged_pressure_bit_exact(gpu, oracle_lib, scene_idx, layout):
    """The pressure pass from the per-step LDS staging (work order of own cells, rows
    walked in LDS) leaves the oracle's bits, with and without the Morton layout."""
    scene = small_scenes()[scene_idx]
    print(f"case staged pressure {scene.name} layout={layout}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 3, max_iterations=20)
    sg = run(g, scene, 3, max_iterations=20, work_order=STAGED | layout)
    assert_same_stats(sg, so, "staged")
    assert_same_state(g, o, tag="staged")


@pytest.mark.parametrize("block", [64, 256, 512])
def test_staged_pressure_block_sizes(gpu, oracle_lib, block):
    scene = small_scenes()[1]
    print(f"case staged pressure block {block}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20, work_order=STAGED, block_size=block)
    assert_same_stats(sg, so, f"staged block {block}")
    assert_same_state(g, o, tag=f"staged block {block}")


@pytest.mark.parametrize("cap", [256, 512, 1024])
def test_staged_pressure_partial_bit_exact(gpu, oracle_lib, cap):
    """Staging caps below the workgroups' candidate counts (~450-600 per 64-particle
    workgroup here): a workgroup with up to twice the cap stages the rows that fit and
    walks its other rows' buckets in global memory, one with more walks global memory
    throughout -- the oracle's bits either way."""
    scene = small_scenes()[1]
    print(f"case staged pressure partial cap {cap}", flush=True)
    o = oracle_lib.OracleFluid()
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    so = run(o, scene, 2, max_iterations=20)
    sg = run(g, scene, 2, max_iterations=20, work_order=STAGED, stage_cap=cap)
    assert_same_stats(sg, so, f"staged cap {cap}")
    assert_same_state(g, o, tag=f"staged cap {cap}")


def test_staged_pressure_partial_fast_equals_full(gpu):
    """Fast FP32: rows walked in global memory add the same terms in the same order as
    rows walked in LDS, so a partly staged pass equals the fully staged one bit for bit."""
    scene = small_scenes()[0]
    print("case staged pressure partial fast", flush=True)
    runs = []
    for cap in (0, 256):
        g = gpu.Fluid(precision=gpu.PRECISION_FAST)
        runs.append((g, run(g, scene, 3, max_iterations=20, work_order=STAGED, stage_cap=cap)))
    assert_same_stats(runs[1][1], runs[0][1], "fast cap 256")
    assert_same_state(runs[1][0], runs[0][0], tag="fast cap 256")


def test_1m_staged_two_steps_bit_exact(gpu, oracle_1m):
    scene, ref = oracle_1m
    print("case 1m staged pressure ieee, 2 steps", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scene(scene, min_iterations=3, max_iterations=3, work_order=STAGED)
    for k, (st_o, snap_o) in enumerate(ref):
        st_g = g.step()
        assert_same_stats([st_g], [st_o], f"staged step {k + 1}")
        snap_g = snapshot(g)
        for name in snap_o:
            np.testing.assert_array_equal(snap_g[name], snap_o[name], err_msg=f"staged step {k + 1} {name}")
    g.close()


def test_1m_partial_staging_two_steps_bit_exact(gpu, oracle_1m):
    """The 1M dam break with a staging cap of 1,024: its ~1,800-candidate workgroups are
    staged in part (rows past the cap walked in global memory)."""
    scene, ref = oracle_1m
    print("case 1m partly staged pressure ieee, 2 steps", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scene(scene, min_iterations=3, max_iterations=3, work_order=STAGED, stage_cap=1024)
    for k, (st_o, snap_o) in enumerate(ref):
        st_g = g.step()
        assert_same_stats([st_g], [st_o], f"partly staged step {k + 1}")
        snap_g = snapshot(g)
        for name in snap_o:
            np.testing.assert_array_equal(snap_g[name], snap_o[name], err_msg=f"partly staged step {k + 1} {name}")
    g.close()


def test_staged_batch_matches_oracle(gpu, oracle_lib):
    """A batch (one staging per fluid range of every scene) against each scene's oracle."""
    oracle_lib.set_threads(THREADS)
    batch = _batch_subset("batch_64")
    print(f"case staged batch N={[s.fluid_count for s in batch]}", flush=True)
    g = gpu.Fluid(precision=gpu.PRECISION_IEEE)
    g.load_scenes(batch, work_order=STAGED)
    sb = [g.step_scenes() for _ in range(2)]
    for i, sc in enumerate(batch):
        o = oracle_lib.OracleFluid()
        so = run(o, sc, 2)
        assert_same_stats([s[i] for s in sb], so, sc.name)
        g.select_scene(i)
        assert_same_state(g, o, tag=sc.name)
    g.close()


@pytest.mark.parametrize("precision", ["ieee", "fast"])
@pytest.mark.parametrize("cap", [0, 256])
def test_staged_dual_pass_equals_global(gpu, precision, cap):
    """The dual first iteration of a warm-started step (both starts' pressure passes
    in one) from the LDS staging, fully and partly staged, leaves the bits of the
    global-memory dual pass: the committed trained network, whose start the scene keeps
    (so every step runs the dual pass and continues from the network's start), three
    steps with the staging forced on and off."""
    params = np.load(os.path.join(os.path.dirname(os.path.dirname(__file__)), "pcisph_amd", "data",
                                  "warm_start.npy"))
    scene = small_scenes()[1]
    prec = gpu.PRECISION_IEEE if precision == "ieee" else gpu.PRECISION_FAST
    print(f"case staged dual pass {precision} cap {cap}", flush=True)
    runs = []
    for order in (STAGED, STAGED_OFF):
        g = gpu.Fluid(precision=prec)
        g.mlp_set_params(params)
        g.set_warm_start(True)
        runs.append((g, run(g, scene, 3, max_iterations=20, work_order=order, stage_cap=cap)))
    (gs, ss), (gg, sg) = runs
    assert any(s["warm_start"] & 3 == 1 for s in ss), ss  # the network's start ran and was kept
    assert [s["warm_start"] for s in ss] == [s["warm_start"] for s in sg]
    assert_same_stats(ss, sg, f"dual {precision} cap {cap}")
    assert_same_state(gs, gg, tag=f"dual {precision} cap {cap}")
