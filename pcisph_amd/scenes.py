"""Seeded synthetic scenes in the reference's particle layout.

Fluid and boundary positions are packed float32[3N] arrays (x, y, z per particle),
y up, gravity along -y, as the reference lays them out and integrates them
(PCISPH/src/scenes.cpp:88-97, PCISPH/data/kernels/pcisph.cl:167).

The tank is made of boundary particles at the fluid spacing, three layers thick,
outside the tank volume and open at the top -- the reference builds its walls the
same way, as three-particle-thick layers of a particle lattice
(PCISPH/src/scenes.cpp:122-135).  Jitter comes from the project's own PCG32
(O'Neill 2014, XSH-RR 64/32), one stream per particle, so the same seed gives the
same bits on every platform (``std::uniform_*`` differs between standard libraries).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_PCG_MULT = np.uint64(6364136223846793005)
_U64_MASK = (1 << 64) - 1


def _pcg_step(state, inc):
    with np.errstate(over="ignore"):
        return state * _PCG_MULT + inc


def _pcg_output(old):
    xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
    rot = (old >> np.uint64(59)).astype(np.uint32)
    return (xorshifted >> rot) | (xorshifted << ((np.uint32(32) - rot) & np.uint32(31)))


def pcg32_streams(seed: int, n: int, draws: int) -> np.ndarray:
    """uint32[n, draws]: `draws` outputs of PCG32 stream i (i = 0..n-1) for `seed`.

    Stream i uses increment 2*i+1 and the reference seeding procedure of pcg32_srandom
    (state = 0; step; state += seed; step)."""
    inc = (np.arange(n, dtype=np.uint64) << np.uint64(1)) | np.uint64(1)
    state = np.zeros(n, dtype=np.uint64)
    state = _pcg_step(state, inc)
    with np.errstate(over="ignore"):
        state = state + np.uint64(seed & _U64_MASK)
    state = _pcg_step(state, inc)
    out = np.empty((n, draws), dtype=np.uint32)
    for d in range(draws):
        old = state
        state = _pcg_step(state, inc)
        out[:, d] = _pcg_output(old)
    return out


def uniform_pm(seed: int, n: int, amplitude: float) -> np.ndarray:
    """float64[n, 3]: per-axis U(-amplitude, +amplitude) from 24-bit PCG32 mantissas."""
    u = (pcg32_streams(seed, n, 3) >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))
    return amplitude * (2.0 * u - 1.0)


def lattice(nx: int, ny: int, nz: int, spacing: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """float64[nx*ny*nz, 3] points (i+0.5)*spacing + origin, x fastest, then y, then z."""
    i = np.arange(nx)
    j = np.arange(ny)
    k = np.arange(nz)
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    pts = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    return (pts + 0.5) * spacing + np.asarray(origin, dtype=np.float64)


def tank_boundary(cells: tuple[int, int, int], spacing: float, layers: int = 3) -> np.ndarray:
    """Open-top box of boundary particles around the tank [0,cx*s] x [0,cy*s] x [0,cz*s]:
    every lattice point (i+0.5)*s with -layers <= i < cx+layers, -layers <= j < cy,
    -layers <= k < cz+layers that lies outside the tank volume."""
    cx, cy, cz = cells
    i = np.arange(-layers, cx + layers)
    j = np.arange(-layers, cy)
    k = np.arange(-layers, cz + layers)
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    inside = (ii >= 0) & (ii < cx) & (jj >= 0) & (kk >= 0) & (kk < cz)
    pts = np.stack([ii[~inside], jj[~inside], kk[~inside]], axis=1).astype(np.float64)
    return (pts + 0.5) * spacing


@dataclass
class Scene:
    """Particles plus the reference setter values that go with them."""

    name: str
    fluid: np.ndarray          # float32[3N]
    boundary: np.ndarray       # float32[3Nb]
    particle_radius: float
    rest_density: float = 1000.0
    viscosity: float = 8e-5    # reference default (PCISPH/src/scenes.cpp:221)
    surface_tension: float = 0.1
    gravity: float = -9.81
    delta_t: float = 5e-4
    threshold: float = 0.01
    min_iterations: int = 3
    max_iterations: int = 50
    steps: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def fluid_count(self) -> int:
        return self.fluid.size // 3

    @property
    def boundary_count(self) -> int:
        return self.boundary.size // 3


def dam_break(block=(32, 32, 16), spacing=0.02, jitter=0.001, tank=(1.0, 1.0, 0.5),
              seed=20240611, name="dam_break", steps=200, **kw) -> Scene:
    """A `block` = (nx, ny, nz) lattice of fluid at `spacing` with per-axis U(+-jitter)
    in the x=0, z=0 corner of an open-top tank of size `tank` (x, y-up, z)."""
    nx, ny, nz = block
    cells = tuple(int(round(t / spacing)) for t in tank)
    if nx > cells[0] or ny > cells[1] or nz > cells[2]:
        raise ValueError(f"block {block} does not fit tank {tank} at spacing {spacing}")
    pts = lattice(nx, ny, nz, spacing)
    if jitter:
        pts = pts + uniform_pm(seed, pts.shape[0], jitter)
    fluid = pts.astype(np.float32).ravel()
    boundary = tank_boundary(cells, spacing).astype(np.float32).ravel()
    return Scene(name=name, fluid=fluid, boundary=boundary, particle_radius=0.5 * spacing,
                 steps=steps, meta={"block": list(block), "spacing": spacing, "jitter": jitter,
                                    "tank": list(tank), "seed": seed}, **kw)


def sphere_drop(count: int, pool_depth_cells: int, spacing: float, tank_cells=(50, 50, 50),
                drop_height: float = 0.3, seed=7, name="sphere_drop", **kw) -> Scene:
    """A still pool `pool_depth_cells` particles deep filling the tank floor, plus a
    jittered sphere of about `count` particles whose bottom is `drop_height` above the
    pool surface."""
    cx, cy, cz = tank_cells
    pool = lattice(cx, pool_depth_cells, cz, spacing)
    radius = spacing * (3.0 * count / (4.0 * np.pi)) ** (1.0 / 3.0)
    n = int(np.ceil(2 * radius / spacing)) + 2
    cube = lattice(n, n, n, spacing, origin=(-n * spacing / 2,) * 3)
    ball = cube[np.einsum("ij,ij->i", cube, cube) <= radius * radius]
    centre = np.array([cx * spacing / 2, pool_depth_cells * spacing + drop_height + radius,
                       cz * spacing / 2])
    ball = ball + centre + uniform_pm(seed, ball.shape[0], 0.05 * spacing)
    fluid = np.concatenate([pool, ball]).astype(np.float32).ravel()
    boundary = tank_boundary((cx, max(cy, int(np.ceil(centre[1] / spacing + n))), cz),
                             spacing).astype(np.float32).ravel()
    return Scene(name=name, fluid=fluid, boundary=boundary, particle_radius=0.5 * spacing,
                 meta={"count": count, "spacing": spacing, "seed": seed}, **kw)


def dam_break_16k(seed=20240611) -> Scene:
    """16,384-particle dam break: 32x32x16 (x, y, z) block at 0.02 m with U(+-1 mm)
    jitter in a 1 x 1 x 0.5 m tank, h = 0.04 m, 200 steps."""
    return dam_break((32, 32, 16), 0.02, 0.001, (1.0, 1.0, 0.5), seed=seed,
                     name="dam_break_16k", steps=200)


def dam_break_1m(seed=20240611) -> Scene:
    """1,048,576-particle dam break: 128x128x64 (x, y, z) block at 0.01 m with
    U(+-0.5 mm) jitter in a 2 x 2 x 1 m tank (100 x 100 x 50 cells of h = 0.02 m),
    1000 steps."""
    return dam_break((128, 128, 64), 0.01, 0.0005, (2.0, 2.0, 1.0), seed=seed,
                     name="dam_break_1m", steps=1000)


def small_dam_break(block=(8, 8, 8), tank=(0.4, 0.4, 0.3), spacing=0.02, jitter=0.001,
                    seed=1234, steps=5) -> Scene:
    """A few hundred particles: sizes the oracle finishes in well under a second."""
    return dam_break(block, spacing, jitter, tank, seed=seed, name="small_dam_break", steps=steps)


def sloshing_tank(amplitude: float = 0.5, frequency: float = 0.8, base: Scene | None = None,
                  name: str = "sloshing_tank") -> Scene:
    """`base` (default: the 1,048,576-particle dam break) in a tank that oscillates
    along x as amplitude * sin(2 pi frequency t); boundary_shift() gives each step's
    displacement of the walls."""
    sc = base if base is not None else dam_break_1m()
    sc.name = name
    sc.meta = dict(sc.meta, slosh={"amplitude": amplitude, "frequency": frequency})
    return sc


def boundary_shift(scene: Scene, step: int) -> tuple[float, float, float]:
    """Wall displacement from the start of step `step` to the start of step + 1 for a
    sloshing scene, (0, 0, 0) otherwise."""
    slosh = scene.meta.get("slosh")
    if not slosh:
        return (0.0, 0.0, 0.0)
    a, f, dt = slosh["amplitude"], slosh["frequency"], scene.delta_t
    w = 2.0 * np.pi * f
    return (a * (np.sin(w * (step + 1) * dt) - np.sin(w * step * dt)), 0.0, 0.0)


def log_uniform_counts(count: int, n_min: int, n_max: int, seed: int) -> np.ndarray:
    """int64[count] particle counts log-uniform in [n_min, n_max] (PCG32 stream per scene)."""
    u = (pcg32_streams(seed, count, 1)[:, 0] >> np.uint32(8)).astype(np.float64) / (1 << 24)
    return np.rint(np.exp(np.log(n_min) + u * (np.log(n_max) - np.log(n_min)))).astype(np.int64)


def scaled_dam_break(target: int, seed: int, name: str = "dam_break") -> Scene:
    """About `target` particles: an (n, 2n, n) block 0.4 x 0.8 x 0.4 m in a 1 x 1 x 1 m
    tank, the spacing set by the count (jitter 5% of it)."""
    n = max(2, int(round((target / 2.0) ** (1.0 / 3.0))))
    spacing = 0.4 / n
    cells = int(round(1.0 / spacing))
    return dam_break((n, 2 * n, n), spacing, 0.05 * spacing, (cells * spacing,) * 3, seed=seed, name=name)


def scaled_sphere_drop(target: int, seed: int, name: str = "sphere_drop") -> Scene:
    """About `target` particles: a 0.1 m deep still pool on the floor of a 1 x 1 x 1 m
    tank and a sphere of radius 0.15 m whose bottom is 0.3 m above it, the spacing set
    by the count."""
    volume = 1.0 * 0.1 * 1.0 + 4.0 / 3.0 * np.pi * 0.15 ** 3
    spacing = (volume / target) ** (1.0 / 3.0)
    cells = int(round(1.0 / spacing))
    depth = max(1, int(round(0.1 / spacing)))
    ball = int(round(4.0 / 3.0 * np.pi * (0.15 / spacing) ** 3))
    return sphere_drop(ball, depth, spacing, (cells, cells, cells), drop_height=0.3, seed=seed, name=name)


def batch_mix(count: int, n_min: int, n_max: int, seed: int,
              kinds=("dam_break", "sphere_drop")) -> list[Scene]:
    """`count` independent scenes with particle counts log-uniform in [n_min, n_max],
    kinds assigned in turn: "dam_break", "sphere_drop", "sloshing" (a dam break whose
    tank oscillates along x with 0.1 m amplitude at 0.8 Hz)."""
    targets = log_uniform_counts(count, n_min, n_max, seed)
    out = []
    for k, t in enumerate(targets):
        kind = kinds[k % len(kinds)]
        if kind == "dam_break":
            sc = scaled_dam_break(int(t), seed + k, name=f"dam_break_{k}")
        elif kind == "sphere_drop":
            sc = scaled_sphere_drop(int(t), seed + k, name=f"sphere_drop_{k}")
        elif kind == "sloshing":
            sc = sloshing_tank(0.1, 0.8, scaled_dam_break(int(t), seed + k), name=f"sloshing_{k}")
        else:
            raise ValueError(f"unknown scene kind {kind}")
        out.append(sc)
    return out


def batch_64(seed: int = 20240612) -> list[Scene]:
    """64 scenes, counts log-uniform in [2,048, 262,144], half dam breaks and half
    sphere drops, 500 steps."""
    return batch_mix(64, 2048, 262144, seed)


def batch_256(seed: int = 20240613) -> list[Scene]:
    """256 scenes, counts log-uniform in [1,024, 65,536]: dam breaks, sphere drops and
    sloshing tanks."""
    return batch_mix(256, 1024, 65536, seed, kinds=("dam_break", "sphere_drop", "sloshing"))
