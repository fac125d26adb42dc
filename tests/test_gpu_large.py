"""GPU checks at the full 1,048,576-particle size and at mid sizes, through properties
that do not need the oracle (sort products), and against the oracle where it is fast
enough (one step of a 64,000-particle scene)."""
import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field

pytestmark = pytest.mark.gpu


def np_hash_keys(pos, cell, B):
    """Vectorised grid_utils.cl:9-28 (floor of an IEEE float32 division, 32-bit wrap)."""
    c = np.floor(pos.reshape(-1, 3).astype(np.float32) / np.float32(cell)).astype(np.int64)
    u = (c & 0xFFFFFFFF).astype(np.uint64)
    M = np.uint64(0xFFFFFFFF)
    h = ((u[:, 0] * np.uint64(19349663)) & M) ^ ((u[:, 1] * np.uint64(73856093)) & M) ^ \
        ((u[:, 2] * np.uint64(83492791)) & M)
    v = ((h << np.uint64(6)) + (u[:, 0] & np.uint64(63)) + np.uint64(4) * (u[:, 1] & np.uint64(63)) +
         np.uint64(16) * (u[:, 2] & np.uint64(63))) & M
    return ((v & np.uint64(0x7FFFFFFF)) % np.uint64(B)).astype(np.uint32)


def check_sort_products(keys, src, cells, pos_before, cell, B):
    n = keys.size
    assert np.array_equal(np.sort(src), np.arange(n, dtype=np.uint32)), "src is not a permutation"
    expect = np_hash_keys(pos_before, cell, B)
    np.testing.assert_array_equal(keys, expect[src])
    np.testing.assert_array_equal(src, np.argsort(expect, kind="stable").astype(np.uint32))
    cells = cells.reshape(-1, 2)
    starts = np.nonzero(np.concatenate([[True], keys[1:] != keys[:-1]]))[0]
    ends = np.concatenate([starts[1:], [n]])
    np.testing.assert_array_equal(cells[keys[starts], 0], starts)
    np.testing.assert_array_equal(cells[keys[starts], 1], ends)
    occupied = np.zeros(B, bool)
    occupied[keys[starts]] = True
    assert np.all(cells[~occupied] == 0)


@pytest.mark.parametrize("precision", [0, 1])
def test_1m_sort_products_and_first_step(precision):
    import pcisph_amd
    scene = scenes.dam_break_1m()
    print(f"case 1m sort precision={precision}", flush=True)
    f = pcisph_amd.Fluid(precision=precision)
    f.load_scene(scene, max_iterations=50)
    st = f.step()
    p = f.get_params()
    check_sort_products(f.download(Field.FLUID_KEYS), f.download(Field.FLUID_SRC_LOCATIONS),
                        f.download(Field.FLUID_CELL_OFFSETS), scene.fluid, p.cell_size, p.bucket_count)
    check_sort_products(f.download(Field.BOUNDARY_KEYS), f.download(Field.BOUNDARY_SRC_LOCATIONS),
                        f.download(Field.BOUNDARY_CELL_OFFSETS), scene.boundary, p.cell_size, p.bucket_count)
    rho = f.download(Field.FLUID_DENSITIES)
    # lattice at spacing 2r (1.0098 rho0 inside) with +-5% jitter per axis
    assert 0.15 * p.rest_density < rho.min() and rho.max() < 1.15 * p.rest_density, (rho.min(), rho.max())
    counts = f.download(Field.FLUID_NEIGHBOR_COUNTS).reshape(-1, 4)
    assert st["fluid_candidates"] == int(counts[:, 0].sum())
    assert st["boundary_candidates"] == int(counts[:, 1].sum())
    print(f"  iterations {st['iterations']} max error {st['max_density_error']:.4f}", flush=True)


def test_64k_ieee_matches_oracle(oracle_lib):
    import pcisph_amd
    scene = scenes.dam_break((40, 40, 40), 0.02, 0.001, (1.0, 1.0, 1.0), seed=5, name="dam_break_64k")
    print(f"case 64k ieee N={scene.fluid_count} Nb={scene.boundary_count}", flush=True)
    oracle_lib.set_threads(16)
    o = oracle_lib.OracleFluid()
    o.load_scene(scene, max_iterations=10)
    g = pcisph_amd.Fluid(precision=pcisph_amd.PRECISION_IEEE)
    g.load_scene(scene, max_iterations=10, work_order=3 | 8)  # the Morton layout at this size
    a, b = o.step(), g.step()
    assert a["iterations"] == b["iterations"]
    for fld in (Field.FLUID_KEYS, Field.FLUID_SRC_LOCATIONS, Field.FLUID_CELL_OFFSETS,
                Field.BOUNDARY_SRC_LOCATIONS, Field.FLUID_DENSITIES, Field.FLUID_POSITIONS,
                Field.FLUID_VELOCITIES, Field.FLUID_PRESSURES):
        np.testing.assert_array_equal(g.download(fld), o.download(fld), err_msg=Field(fld).name)


def test_sort_beyond_2p24_keys():
    """2^24 + 2^20 fluid particles: the radix sort's histogram scan has more chunk
    totals than one workgroup scans and recurses.  The sort products must still be the
    stable sort of the hash keys (checked through properties: sorted keys, a
    permutation, keys[c] = hash(src[c]), ascending sources within equal keys, the cell
    table at every key transition)."""
    import pcisph_amd
    nx, ny, nz = 256, 272, 256
    n = nx * ny * nz
    assert n == (1 << 24) + (1 << 20)
    rng = np.random.default_rng(11)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = (g.astype(np.float32) * np.float32(0.02) + np.float32(0.05) +
           rng.uniform(-0.002, 0.002, g.shape).astype(np.float32)).astype(np.float32).ravel()
    del g
    scene = scenes.Scene(name="block_17m", fluid=pos, boundary=np.zeros(0, np.float32), particle_radius=0.01)
    print(f"case sort beyond 2^24 N={n}", flush=True)
    f = pcisph_amd.Fluid()
    f.load_scene(scene, max_iterations=3, neighbor_skin=0.0)
    f.step()
    p = f.get_params()
    keys, src = f.download(Field.FLUID_KEYS), f.download(Field.FLUID_SRC_LOCATIONS)
    assert np.all(keys[1:] >= keys[:-1]), "keys not sorted"
    assert np.array_equal(np.sort(src), np.arange(n, dtype=np.uint32)), "src is not a permutation"
    np.testing.assert_array_equal(keys, np_hash_keys(pos, p.cell_size, p.bucket_count)[src])
    same = keys[1:] == keys[:-1]
    assert np.all(src[1:][same] > src[:-1][same]), "sort is not stable"
    cells = f.download(Field.FLUID_CELL_OFFSETS).reshape(-1, 2)
    starts = np.nonzero(np.concatenate([[True], ~same]))[0]
    ends = np.concatenate([starts[1:], [n]])
    np.testing.assert_array_equal(cells[keys[starts], 0], starts)
    np.testing.assert_array_equal(cells[keys[starts], 1], ends)
    f.close()
