"""Pins the magnitudes of the oracle's force terms (no GPU).

Each expectation is written from the published formulas, with every constant derived
in float64 from the kernel radius h alone -- not from the oracle's parameter block:

  poly6            W(r)      = 315/(64 pi h^9) (h^2 - r^2)^3                 (Mueller 2003)
  poly6 gradient   grad W(d) = -945/(32 pi h^9) (h^2 - r^2)^2 d
  viscosity        lap W(r)  = 45/(pi h^6) (h - r)
  spiky gradient   grad W(d) = -45/(pi h^6) (h - r)^2 d / r
  Akinci cohesion  C(r)      = 32/(pi h^9) * { 2 (h-r)^3 r^3 - h^6/64   r <= h/2
                                             {   (h-r)^3 r^3            h/2 < r <= h
  normal           n_i = h sum_j m/rho_j grad W(x_i - x_j)             (Akinci 2013)
  surface tension  F_i = sum_j K_ij (-gamma m^2 C(r) (x_i-x_j)/r - gamma m (n_i - n_j)),
                   K_ij = 2 rho0 / (rho_i + rho_j), over 1e-4 < r < h
  viscosity force  F_i = mu m sum_j (v_j - v_i) / rho_j lap W(r)
  pressure force   F_i = -m^2 sum_j (p_i/rho_i^2 + p_j/rho_j^2) grad W_spiky(x_i - x_j),
                   rho_j = rho0 for boundary particles               (Solenthaler 2009)

as the reference applies them (PCISPH/data/kernels/pcisph.cl:121-173 force
initialisation, :265-300 pressure force, :8-61 kernels).  Every test also evaluates
the expectation with a planted mistake (a dropped factor, a wrong power of the mass,
a missing term) and asserts that the oracle is far from it, so the tolerance used
cannot hide such a mistake.
"""
import math

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field

RHO0 = 1000.0


# ---------------------------------------------------------------- textbook kernels
def w_poly6(r2, h):
    return np.where(r2 <= h * h, 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r2) ** 3, 0.0)


def grad_poly6(d, h):
    r2 = (d * d).sum(-1)
    f = np.where(r2 <= h * h, -945.0 / (32.0 * math.pi * h ** 9) * (h * h - r2) ** 2, 0.0)
    return f[..., None] * d


def lap_visc(r, h):
    return np.where(r <= h, 45.0 / (math.pi * h ** 6) * (h - r), 0.0)


def grad_spiky(d, h):
    r = np.sqrt((d * d).sum(-1))
    ok = (r <= h) & (r >= 1e-4)
    f = np.where(ok, -45.0 / (math.pi * h ** 6) * (h - r) ** 2 / np.where(ok, r, 1.0), 0.0)
    return f[..., None] * d


def cohesion_c(r, h, branch_const=1.0):
    inner = 2.0 * (h - r) ** 3 * r ** 3 - branch_const * h ** 6 / 64.0
    outer = (h - r) ** 3 * r ** 3
    return 32.0 / (math.pi * h ** 9) * np.where(r <= 0.5 * h, inner, np.where(r <= h, outer, 0.0))


def physics(r):
    h = 4.0 * r
    return h, RHO0 * (2.0 * r) ** 3


def textbook_other_forces(x, v, b, r, mu, gamma, g, *, plant=None):
    """float64 forces of pcisph.cl:121-173 from the formulas in the module docstring.
    `plant` names one deliberate mistake (for the sensitivity checks)."""
    h, m = physics(r)
    d = x[:, None, :] - x[None, :, :]
    r2 = (d * d).sum(-1)
    dist = np.sqrt(r2)
    rho = m * w_poly6(r2, h).sum(1)
    if len(b):
        db = x[:, None, :] - b[None, :, :]
        rho = rho + m * w_poly6((db * db).sum(-1), h).sum(1)
    n = h * (m / rho[None, :, None] * grad_poly6(d, h)).sum(1)
    mass_v = m * m if plant == "visc_m2" else m
    rho_v = np.ones_like(rho) if plant == "visc_no_rho" else rho
    visc = mu * mass_v * ((v[None, :, :] - v[:, None, :]) / rho_v[None, :, None] *
                          lap_visc(dist, h)[..., None]).sum(1)
    near = (dist > 1e-4) & (dist < h)
    k = 2.0 * RHO0 / (rho[:, None] + rho[None, :])
    if plant == "no_kij":
        k = np.ones_like(k)
    safe = np.where(near, dist, 1.0)
    c = cohesion_c(dist, h, branch_const=-1.0 if plant == "branch_sign" else 1.0)
    mass_c = m if plant == "coh_m1" else m * m
    coh = (np.where(near, k * -gamma * mass_c * c / safe, 0.0)[..., None] * d).sum(1)
    curv = (np.where(near, k * -gamma * m, 0.0)[..., None] * (n[:, None, :] - n[None, :, :])).sum(1)
    if plant == "no_curvature":
        curv = np.zeros_like(curv)
    grav = np.zeros_like(x)
    grav[:, 1] = m * g
    return grav + visc + coh + curv, rho, {"visc": visc, "coh": coh, "curv": curv}


# ---------------------------------------------------------------- oracle drivers
def _scene(fluid, boundary=None, r=0.01, **kw):
    fluid = np.asarray(fluid, np.float32).ravel()
    boundary = np.zeros(0, np.float32) if boundary is None else np.asarray(boundary, np.float32).ravel()
    return scenes.Scene(name="forces", fluid=fluid, boundary=boundary, particle_radius=r, **kw)


def _run_with_state(oracle_lib, scene, v, **opts):
    """One step after a first one: the first step applies the setters (and zeroes the
    velocities, Fluid.cpp:462-463); then the scene's positions and the given
    velocities are uploaded and a second step runs from exactly that state.  Returns
    the oracle and the slot -> input-index map of the second step's sort."""
    o = oracle_lib.OracleFluid()
    o.load_scene(scene, **opts)
    o.step()
    o.upload(Field.FLUID_POSITIONS, scene.fluid)
    o.upload(Field.FLUID_VELOCITIES, np.asarray(v, np.float32).ravel())
    if scene.boundary_count:
        o.upload(Field.BOUNDARY_POSITIONS, scene.boundary)
    o.step()
    return o, o.download(Field.FLUID_SRC_LOCATIONS).astype(np.int64)


def _assert_pinned(got, expect, planted: dict, rtol=2e-5):
    """got within rtol of expect (relative to the largest |expect| component); every
    planted expectation at least 100 times further away."""
    scale = float(np.abs(expect).max())
    assert scale > 0
    err = float(np.abs(got - expect).max()) / scale
    assert err <= rtol, err
    for name, wrong in planted.items():
        miss = float(np.abs(got - wrong).max()) / scale
        assert miss > 100 * rtol, (name, miss)


# ---------------------------------------------------------------- viscosity
@pytest.mark.parametrize("frac", [0.3, 0.6, 0.95])
def test_viscosity_two_particle_closed_form(oracle_lib, frac):
    """F_0 = mu m (v_1 - v_0) / rho_1 * 45/(pi h^6) (h - d), rho_1 = m W(0) + m W(d);
    F_1 = -F_0 (equal densities).  Cohesion and gravity are off."""
    r, mu = 0.01, 0.5
    h, m = physics(r)
    u = np.array([0.6, 0.0, 0.8])
    d = frac * h
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5] + d * u])
    v = np.array([[0.1, -0.2, 0.3], [-0.3, 0.05, 0.2]])
    scene = _scene(x, viscosity=mu, surface_tension=0.0, gravity=0.0)
    o, src = _run_with_state(oracle_lib, scene, v, max_iterations=3)
    xf = scene.fluid.reshape(-1, 3).astype(np.float64)
    dd = float(np.linalg.norm(xf[1] - xf[0]))
    rho = m * (w_poly6(0.0, h) + w_poly6(dd * dd, h))
    f0 = mu * m * (v[1] - v[0]) / rho * 45.0 / (math.pi * h ** 6) * (h - dd)
    expect = np.stack([f0, -f0])[src]
    got = o.download(Field.FLUID_OTHER_FORCES).reshape(-1, 3).astype(np.float64)
    assert o.download(Field.FLUID_DENSITIES) == pytest.approx([rho, rho], rel=1e-5)
    _assert_pinned(got, expect, {"m^2": expect * m, "no 1/rho_j": expect * rho,
                                 "(h-r)^2": expect * (h - dd)})


# ---------------------------------------------------------------- cohesion + curvature
@pytest.mark.parametrize("frac", [0.3, 0.45, 0.55, 0.75])
def test_surface_tension_two_particle_closed_form(oracle_lib, frac):
    """Two particles at rest, viscosity and gravity off (both branches of C(r)):
    F_0 = K (-gamma m^2 C(d) (x_0 - x_1)/d) + K (-gamma m (n_0 - n_1)),
    K = 2 rho0 / (rho_0 + rho_1) = rho0 / rho, n_0 = h m/rho grad W(x_0 - x_1) = -n_1."""
    r, gamma = 0.01, 0.1
    h, m = physics(r)
    u = np.array([0.0, 0.6, -0.8])
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5] + frac * h * u])
    scene = _scene(x, viscosity=0.0, surface_tension=gamma, gravity=0.0)
    o, src = _run_with_state(oracle_lib, scene, np.zeros((2, 3)), max_iterations=3)
    xf = scene.fluid.reshape(-1, 3).astype(np.float64)
    dvec = xf[0] - xf[1]
    dd = float(np.linalg.norm(dvec))
    rho = m * (w_poly6(0.0, h) + w_poly6(dd * dd, h))
    k = 2.0 * RHO0 / (rho + rho)
    n0 = h * m / rho * grad_poly6(dvec, h)
    coh = k * -gamma * m * m * float(cohesion_c(dd, h)) * dvec / dd
    curv = k * -gamma * m * (n0 - (-n0))
    f0 = coh + curv
    expect = np.stack([f0, -f0])[src]
    got = o.download(Field.FLUID_OTHER_FORCES).reshape(-1, 3).astype(np.float64)
    np.testing.assert_allclose(o.download(Field.FLUID_NORMALS).reshape(-1, 3).astype(np.float64),
                               np.stack([n0, -n0])[src], rtol=1e-5, atol=1e-6 * np.abs(n0).max())
    planted = {"no K_ij": expect / k, "no curvature": np.stack([coh, -coh])[src],
               "cohesion m^1": np.stack([coh / m + curv, -(coh / m + curv)])[src]}
    if frac <= 0.5:
        c_wrong = 32.0 / (math.pi * h ** 9) * (2 * (h - dd) ** 3 * dd ** 3 + h ** 6 / 64)
        coh_w = k * -gamma * m * m * c_wrong * dvec / dd
        planted["branch sign"] = np.stack([coh_w + curv, -(coh_w + curv)])[src]
    _assert_pinned(got, expect, planted)


def test_other_forces_cluster_brute_force(oracle_lib):
    """A jittered cluster with random velocities, unequal densities and a few boundary
    particles (which enter the densities but not viscosity, normals or cohesion,
    pcisph.cl:94-173): every particle's force equals the float64 sum of the formulas."""
    rng = np.random.default_rng(17)
    r = 0.01
    h, m = physics(r)
    k = np.arange(4) * 0.021
    gx, gy, gz = np.meshgrid(k, k, k, indexing="ij")
    x = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + 0.5 + rng.uniform(-0.004, 0.004, (64, 3))
    b = np.stack([np.linspace(0.48, 0.56, 12), np.full(12, 0.49), np.full(12, 0.5)], 1)
    v = rng.uniform(-0.5, 0.5, (64, 3))
    mu, gamma, g = 2e-5, 0.1, -9.81
    scene = _scene(x, b, viscosity=mu, surface_tension=gamma, gravity=g)
    o, src = _run_with_state(oracle_lib, scene, v, max_iterations=3)
    xf = scene.fluid.reshape(-1, 3).astype(np.float64)
    bf = scene.boundary.reshape(-1, 3).astype(np.float64)
    vf = np.asarray(v, np.float32).astype(np.float64)
    expect, rho, parts = textbook_other_forces(xf, vf, bf, r, mu, gamma, g)
    assert float(rho.max() / rho.min()) > 1.2  # K_ij differs from 1 and between pairs
    for name, part in parts.items():  # every term matters at the tolerance used
        assert float(np.abs(part).max()) > 1e-2 * float(np.abs(expect).max()), name
    got = o.download(Field.FLUID_OTHER_FORCES).reshape(-1, 3).astype(np.float64)
    np.testing.assert_allclose(o.download(Field.FLUID_DENSITIES).astype(np.float64), rho[src], rtol=1e-5)
    planted = {p: textbook_other_forces(xf, vf, bf, r, mu, gamma, g, plant=p)[0][src]
               for p in ("visc_m2", "visc_no_rho", "no_kij", "branch_sign", "coh_m1", "no_curvature")}
    _assert_pinned(got, expect[src], planted, rtol=5e-5)


# ---------------------------------------------------------------- pressure force
def test_pressure_force_fluid_fluid_brute_force(oracle_lib):
    """A compressed lattice (spacing 0.95 * 2r, one iteration): every particle's
    pressure force is -m^2 sum_j (p_i/rho_i^2 + p_j/rho_j^2) grad W_spiky(x_i - x_j)
    over the step-start positions, with the pressures and densities the step left."""
    r = 0.01
    h, m = physics(r)
    n = 7
    k = (np.arange(n) - (n - 1) / 2) * 0.019
    gx, gy, gz = np.meshgrid(k, k, k, indexing="ij")
    rng = np.random.default_rng(3)
    x = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + 0.5 + rng.uniform(-0.002, 0.002, (n ** 3, 3))
    scene = _scene(x, gravity=0.0, surface_tension=0.0, viscosity=0.0)
    o, src = _run_with_state(oracle_lib, scene, np.zeros_like(x), min_iterations=1, max_iterations=1)
    xs = scene.fluid.reshape(-1, 3).astype(np.float64)[src]
    p = o.download(Field.FLUID_PRESSURES).astype(np.float64)
    rho = o.download(Field.FLUID_DENSITIES).astype(np.float64)
    assert p.min() >= 0 and np.count_nonzero(p) > 100 and p.max() > 1.5 * np.median(p[p > 0])
    d = xs[:, None, :] - xs[None, :, :]
    q = p / rho ** 2
    expect = -m * m * ((q[:, None] + q[None, :])[..., None] * grad_spiky(d, h)).sum(1)
    got = o.download(Field.FLUID_PRESSURE_FORCES).reshape(-1, 3).astype(np.float64)
    qw = p / RHO0 ** 2  # a planted mistake: rest density in place of each particle's own
    wrong_rho = -m * m * ((qw[:, None] + qw[None, :])[..., None] * grad_spiky(d, h)).sum(1)
    own_only = -m * m * ((q[:, None] + 0 * q[None, :])[..., None] * grad_spiky(d, h)).sum(1)
    _assert_pinned(got, expect, {"m^1": expect / m, "rho0 for fluid": wrong_rho, "no p_j term": own_only},
                   rtol=5e-5)


@pytest.mark.parametrize("boundary_init", [0, 1])
def test_pressure_force_fluid_boundary(oracle_lib, boundary_init):
    """One fluid particle inside a dense boundary cluster.  Fluid-boundary terms use
    p_b / rho0^2 (pcisph.cl:282-283).  Without the boundary-boundary density pass the
    boundary pressures stay 0; with it (launch_boundary_init) they rise above 0 and
    their rho0 factor is pinned too."""
    r = 0.01
    h, m = physics(r)
    k = (np.arange(5) - 2) * 0.016
    gx, gy, gz = np.meshgrid(k, k, k, indexing="ij")
    b = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + 0.5
    b = b[np.abs(b - 0.5).max(1) > 1e-9]  # a hole at the centre for the fluid particle
    x = np.array([[0.5 + 0.003, 0.5 - 0.002, 0.5 + 0.001]])
    scene = _scene(x, b, gravity=0.0, surface_tension=0.0, viscosity=0.0)
    o, _ = _run_with_state(oracle_lib, scene, np.zeros_like(x), min_iterations=1, max_iterations=1,
                           launch_boundary_init=boundary_init)
    bs = o.download(Field.BOUNDARY_POSITIONS).reshape(-1, 3).astype(np.float64)
    pb = o.download(Field.BOUNDARY_PRESSURES).astype(np.float64)
    p = float(o.download(Field.FLUID_PRESSURES)[0])
    rho = float(o.download(Field.FLUID_DENSITIES)[0])
    xi = scene.fluid.reshape(-1, 3).astype(np.float64)[0]
    assert p > 0 and rho > RHO0
    if boundary_init:
        assert np.count_nonzero(pb) > 10
    else:
        np.testing.assert_array_equal(pb, 0)
    g = grad_spiky(xi[None, :] - bs, h)
    expect = -m * m * ((p / rho ** 2 + pb / RHO0 ** 2)[:, None] * g).sum(0)
    got = o.download(Field.FLUID_PRESSURE_FORCES).astype(np.float64)
    planted = {"m^1": expect / m,
               "fluid rho for the boundary": -m * m * ((p / rho ** 2 + pb / rho ** 2)[:, None] * g).sum(0)}
    if not boundary_init:
        planted.pop("fluid rho for the boundary")
    _assert_pinned(got[None], expect[None], {kk: vv[None] for kk, vv in planted.items()})
