"""ctypes mirror of the data types declared in include/pcisph.h and
include/pcisph_params.h (the C ABI of the step)."""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np


class SimulationParams(C.Structure):
    """The reference's 80-byte `Simulation_Params`
    (PCISPH/data/kernels/Simulation_Params.h:17-45)."""

    _pack_ = 1
    _fields_ = [
        ("particle_radius", C.c_float),
        ("particle_mass", C.c_float),
        ("rest_density", C.c_float),
        ("boundary_count", C.c_uint32),
        ("fluid_count", C.c_uint32),
        ("kernel_radius", C.c_float),
        ("kernel_radius2", C.c_float),
        ("bucket_count", C.c_uint32),
        ("cell_size", C.c_float),
        ("poly6_normalization", C.c_float),
        ("poly6_d1_normalization", C.c_float),
        ("viscosity_d2_normalization", C.c_float),
        ("spiky_d1_normalization", C.c_float),
        ("surface_tension_coefficient", C.c_float),
        ("surface_tension_term", C.c_float),
        ("surface_tension_normalization", C.c_float),
        ("viscosity_constant", C.c_float),
        ("gravity", C.c_float),
        ("delta_t", C.c_float),
        ("density_variation_scaling_factor", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def raw(self) -> bytes:
        return bytes(C.string_at(C.addressof(self), C.sizeof(self)))


class SolverOptions(C.Structure):
    _fields_ = [
        ("min_iterations", C.c_uint32),
        ("max_iterations", C.c_uint32),
        ("launch_boundary_init", C.c_uint32),
        ("resort_boundary", C.c_uint32),
        ("block_size", C.c_uint32),
        ("sort_digit_bits", C.c_uint32),
        ("neighbor_skin", C.c_float),
        ("graph_iterations", C.c_uint32),
        ("work_order", C.c_uint32),
        ("stage_cap", C.c_uint32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("converged", C.c_uint32),
        ("max_density_variation", C.c_float),
        ("max_density_error", C.c_float),
        ("nonfinite", C.c_uint32),
        ("warm_start", C.c_uint32),
        ("fluid_candidates", C.c_uint64),
        ("boundary_candidates", C.c_uint64),
        ("max_density_expansion", C.c_float),
        ("list_waves", C.c_uint32),
        ("first_density_error", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class Field(enum.IntEnum):
    FLUID_POSITIONS = 0
    FLUID_VELOCITIES = 1
    FLUID_NORMALS = 2
    FLUID_PREDICTED_POSITIONS = 3
    FLUID_DENSITIES = 4
    FLUID_OTHER_FORCES = 5
    FLUID_PRESSURES = 6
    FLUID_PRESSURE_FORCES = 7
    FLUID_DENSITY_VARIATIONS = 8
    BOUNDARY_POSITIONS = 9
    BOUNDARY_PRESSURES = 10
    FLUID_KEYS = 11
    FLUID_SRC_LOCATIONS = 12
    FLUID_CELL_OFFSETS = 13
    BOUNDARY_KEYS = 14
    BOUNDARY_SRC_LOCATIONS = 15
    BOUNDARY_CELL_OFFSETS = 16
    BOUNDARY_INIT_PRED_DENSITIES = 17
    FLUID_NEIGHBOR_COUNTS = 18


UINT_FIELDS = {Field.FLUID_KEYS, Field.FLUID_SRC_LOCATIONS, Field.FLUID_CELL_OFFSETS,
               Field.BOUNDARY_KEYS, Field.BOUNDARY_SRC_LOCATIONS, Field.BOUNDARY_CELL_OFFSETS,
               Field.FLUID_NEIGHBOR_COUNTS}

STATUS = {
    0: "PCISPH_OK",
    1: "PCISPH_ERR_INVALID_ARGUMENT",
    2: "PCISPH_ERR_SIZE",
    3: "PCISPH_ERR_DEVICE",
    4: "PCISPH_ERR_NO_DEVICE",
    5: "PCISPH_ERR_OUT_OF_MEMORY",
    6: "PCISPH_ERR_NUMERIC",
    7: "PCISPH_ERR_INTERNAL",
}


class PcisphError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


def field_dtype(field: int):
    return np.uint32 if Field(field) in UINT_FIELDS else np.float32


SETTERS = ("boundary_count", "fluid_count", "delta_t", "rest_density", "particle_radius",
           "gravity", "viscosity", "surface_tension", "density_variation_threshold")


class SolverBase:
    """Python mirror of `sim::Fluid` (PCISPH/src/sim/Fluid.h:13-102) over a C ABI whose
    entry points are `<prefix>set_*`, `<prefix>upload`, `<prefix>step`, ...
    Subclasses bind the library and the handle."""

    _lib: C.CDLL
    _prefix: str
    _h: C.c_void_p

    def _call(self, name: str, *args):
        rc = getattr(self._lib, self._prefix + name)(self._h, *args)
        if rc != 0:
            raise PcisphError(rc, self._error_message())
        return rc

    def _error_message(self) -> str:
        return ""

    # --- setters (Fluid.h:21-29)
    def set_boundary_count(self, n: int):
        self._call("set_boundary_count", C.c_uint32(int(n)))

    def set_fluid_count(self, n: int):
        self._call("set_fluid_count", C.c_uint32(int(n)))

    def set_delta_t(self, v: float):
        self._call("set_delta_t", C.c_float(v))

    def set_rest_density(self, v: float):
        self._call("set_rest_density", C.c_float(v))

    def set_particle_radius(self, v: float):
        self._call("set_particle_radius", C.c_float(v))

    def set_gravity(self, v: float):
        self._call("set_gravity", C.c_float(v))

    def set_viscosity(self, v: float):
        self._call("set_viscosity", C.c_float(v))

    def set_surface_tension(self, v: float):
        self._call("set_surface_tension", C.c_float(v))

    def set_density_variation_threshold(self, v: float):
        self._call("set_density_variation_threshold", C.c_float(v))

    def get_params(self) -> SimulationParams:
        p = SimulationParams()
        self._call("get_params", C.byref(p))
        return p

    def set_solver_options(self, min_iterations=3, max_iterations=7, launch_boundary_init=0,
                           resort_boundary=0, block_size=0, sort_digit_bits=0, neighbor_skin=0.2,
                           graph_iterations=4, work_order=3, stage_cap=0):
        o = SolverOptions(min_iterations, max_iterations, launch_boundary_init, resort_boundary,
                          block_size, sort_digit_bits, neighbor_skin, graph_iterations, work_order, stage_cap)
        self._call("set_solver_options", C.byref(o))

    def field_size(self, field: int) -> int:
        n = C.c_size_t()
        self._call("field_size", C.c_int(int(field)), C.byref(n))
        return n.value

    def upload(self, field: int, data: np.ndarray):
        arr = np.ascontiguousarray(data, dtype=field_dtype(field)).ravel()
        self._call("upload", C.c_int(int(field)), arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.size))

    def download(self, field: int) -> np.ndarray:
        n = self.field_size(field)
        out = np.empty(n, dtype=field_dtype(field))
        self._call("download", C.c_int(int(field)), out.ctypes.data_as(C.c_void_p), C.c_size_t(n))
        return out

    def step(self) -> dict:
        st = StepStats()
        self._call("step", C.byref(st))
        return st.as_dict()

    def load_scene(self, scene, min_iterations=None, max_iterations=None, **options):
        """Reference setter order of scene::load (PCISPH/src/scenes.cpp:217-226,166-168),
        then the particle uploads of scenes.cpp:192-208."""
        self.set_delta_t(scene.delta_t)
        self.set_rest_density(scene.rest_density)
        self.set_viscosity(scene.viscosity)
        self.set_surface_tension(scene.surface_tension)
        self.set_gravity(scene.gravity)
        self.set_density_variation_threshold(scene.threshold)
        self.set_particle_radius(scene.particle_radius)
        self.set_fluid_count(scene.fluid_count)
        self.set_boundary_count(scene.boundary_count)
        self.set_solver_options(
            min_iterations=scene.min_iterations if min_iterations is None else min_iterations,
            max_iterations=scene.max_iterations if max_iterations is None else max_iterations,
            **options)
        self.upload(Field.FLUID_POSITIONS, scene.fluid)
        self.upload(Field.FLUID_VELOCITIES, np.zeros_like(scene.fluid))
        if scene.boundary_count:
            self.upload(Field.BOUNDARY_POSITIONS, scene.boundary)
