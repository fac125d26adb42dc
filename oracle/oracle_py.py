"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker.  `OracleFluid` has the same methods as `pcisph_amd.Fluid`, so one
driver function can run both on the same scene.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from pcisph_amd.abi import Field, SimulationParams, SolverBase, field_dtype

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ only)."""
    cmd = ["make", "-C", HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_hash_key.restype = C.c_uint32
        L.oracle_hash_key.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
        L.oracle_cell_pos.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.POINTER(C.c_int32)]
        L.oracle_sort_bit_count.restype = C.c_uint32
        L.oracle_sort_bit_count.argtypes = [C.c_uint32]
        L.oracle_compute_params.argtypes = [C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                            C.c_float, C.c_float, C.c_float, C.c_float,
                                            C.c_void_p]
        L.oracle_download_f64.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        for name in ("set_boundary_count", "set_fluid_count", "set_delta_t", "set_rest_density",
                     "set_particle_radius", "set_gravity", "set_viscosity", "set_surface_tension",
                     "set_density_variation_threshold", "get_params", "set_solver_options",
                     "field_size", "upload", "download", "step"):
            getattr(L, "oracle_" + name).restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def hash_key(cx: int, cy: int, cz: int, bucket_count: int) -> int:
    return lib().oracle_hash_key(cx, cy, cz, bucket_count)


def cell_pos(x: float, y: float, z: float, cell: float):
    out = (C.c_int32 * 3)()
    lib().oracle_cell_pos(x, y, z, cell, out)
    return tuple(out)


def sort_bit_count(n: int) -> int:
    return lib().oracle_sort_bit_count(n)


def compute_params(particle_radius, rest_density, boundary_count, fluid_count, viscosity,
                   surface_tension, gravity, delta_t) -> SimulationParams:
    p = SimulationParams()
    rc = lib().oracle_compute_params(particle_radius, rest_density, boundary_count, fluid_count,
                                     viscosity, surface_tension, gravity, delta_t, C.byref(p))
    assert rc == 0
    return p


class OracleFluid(SolverBase):
    """The oracle behind the `sim::Fluid`-shaped interface (FP32 or FP64)."""

    def __init__(self, fp64: bool = False):
        self._lib = lib()
        self._prefix = "oracle_"
        self._h = C.c_void_p(self._lib.oracle_create(1 if fp64 else 0))
        self.fp64 = fp64

    def close(self):
        if self._h:
            self._lib.oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download_f64(self, field: int) -> np.ndarray:
        assert field_dtype(field) == np.float32
        n = self.field_size(field)
        out = np.empty(n, dtype=np.float64)
        rc = self._lib.oracle_download_f64(self._h, int(field), out.ctypes.data_as(C.c_void_p), n)
        assert rc == 0
        return out


__all__ = ["OracleFluid", "Field", "build", "hash_key", "cell_pos", "sort_bit_count",
           "compute_params", "set_threads"]
