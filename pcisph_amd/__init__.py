"""pcisph_amd - the PCISPH simulation step on AMD Instinct MI355X (gfx950).

`Fluid` mirrors the reference's `sim::Fluid` (PCISPH/src/sim/Fluid.h:13-102) over the C
ABI of include/pcisph.h, implemented by libpcisph_amd.so (hand-written HIP kernels).
There is no fallback: if the library is missing or no gfx950 device is present,
constructing a Fluid raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (Field, PcisphError, SimulationParams, SolverBase, SolverOptions, StepStats,
                  field_dtype)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpcisph_amd.so")

PRECISION_FAST = 0
PRECISION_IEEE = 1
PRECISION_FP64 = 2

EXPORTS = (
    "pcisph_create", "pcisph_destroy", "pcisph_last_error",
    "pcisph_set_boundary_count", "pcisph_set_fluid_count", "pcisph_set_delta_t",
    "pcisph_set_rest_density", "pcisph_set_particle_radius", "pcisph_set_gravity",
    "pcisph_set_viscosity", "pcisph_set_surface_tension",
    "pcisph_set_density_variation_threshold", "pcisph_get_params",
    "pcisph_set_solver_options", "pcisph_get_solver_options", "pcisph_set_precision",
    "pcisph_check_buffers_consistent", "pcisph_field_size", "pcisph_upload",
    "pcisph_download", "pcisph_upload_device", "pcisph_download_device",
    "pcisph_download_f64", "pcisph_step", "pcisph_run", "pcisph_sync",
    "pcisph_compute_params", "pcisph_profile_enable", "pcisph_profile_read", "pcisph_prepare",
    "pcisph_mlp_set_params", "pcisph_mlp_get_params", "pcisph_mlp_forward", "pcisph_mlp_gradients",
    "pcisph_mlp_gradients_noalloc",
    "pcisph_samples_reserve", "pcisph_samples_collect", "pcisph_samples_count",
    "pcisph_samples_download", "pcisph_samples_upload", "pcisph_mlp_train", "pcisph_set_warm_start",
    "pcisph_autotune", "pcisph_translate_boundary",
    "pcisph_set_scene_count", "pcisph_get_scene_count", "pcisph_select_scene", "pcisph_step_scenes",
    "pcisph_translate_boundaries", "pcisph_snapshot", "pcisph_restore",
)

MLP_INPUTS = 8
MLP_PARAMS = 4801

STAGE_COUNT = 10


class StageTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 16), ("launches", C.c_uint64 * 16)]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libpcisph_amd.so (built by pcisph_amd.build).  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m pcisph_amd.build`")
    L = C.CDLL(path)
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int
    L.pcisph_last_error.restype = C.c_char_p
    L.pcisph_last_error.argtypes = [C.c_void_p]
    L.pcisph_create.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]
    L.pcisph_destroy.argtypes = [C.c_void_p]
    L.pcisph_compute_params.argtypes = [C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_void_p]
    L.pcisph_run.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    L.pcisph_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.pcisph_mlp_train.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, C.c_void_p]
    L.pcisph_samples_collect.argtypes = [C.c_void_p, C.c_uint32]
    L.pcisph_set_warm_start.argtypes = [C.c_void_p, C.c_int]
    L.pcisph_translate_boundary.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
    L.pcisph_set_scene_count.argtypes = [C.c_void_p, C.c_uint32]
    L.pcisph_select_scene.argtypes = [C.c_void_p, C.c_uint32]
    L.pcisph_step_scenes.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.pcisph_translate_boundaries.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    _lib = L
    return L


def compute_params(particle_radius, rest_density, boundary_count, fluid_count, viscosity,
                   surface_tension, gravity, delta_t) -> SimulationParams:
    """Derived parameters of a parameter set (host only, no device needed)."""
    p = SimulationParams()
    rc = load_library().pcisph_compute_params(particle_radius, rest_density, boundary_count,
                                              fluid_count, viscosity, surface_tension, gravity,
                                              delta_t, C.byref(p))
    if rc:
        raise PcisphError(rc, "pcisph_compute_params")
    return p


class Fluid(SolverBase):
    """The gfx950 step behind the `sim::Fluid` interface.

    precision: PRECISION_FAST (default), PRECISION_IEEE (bit-compatible FP32 order of
    operations) or PRECISION_FP64."""

    def __init__(self, device: int = 0, stream: int | None = None, precision: int = PRECISION_FAST):
        self._lib = load_library()
        self._prefix = "pcisph_"
        h = C.c_void_p()
        rc = self._lib.pcisph_create(int(device), C.c_void_p(stream) if stream else None, C.byref(h))
        if rc:
            msg = self._lib.pcisph_last_error(None).decode()
            raise PcisphError(rc, f"pcisph_create: {msg}")
        self._h = h
        if precision != PRECISION_FAST:
            self.set_precision(precision)

    def _error_message(self) -> str:
        return self._lib.pcisph_last_error(self._h).decode()

    def set_precision(self, precision: int):
        self._call("set_precision", C.c_int(int(precision)))
        self.precision = precision

    def run(self, steps: int, collect: bool = True) -> list[dict]:
        """`steps` consecutive steps; per-step stats when `collect`."""
        if collect:
            arr = (StepStats * steps)()
            self._call("run", C.c_uint32(steps), C.cast(arr, C.c_void_p))
            return [s.as_dict() for s in arr]
        self._call("run", C.c_uint32(steps), None)
        return []

    def sync(self):
        self._call("sync")

    # --- batches of scenes
    def set_scene_count(self, count: int):
        self._call("set_scene_count", C.c_uint32(int(count)))

    def scene_count(self) -> int:
        n = C.c_uint32()
        self._call("get_scene_count", C.byref(n))
        return n.value

    def select_scene(self, index: int):
        """Setters, uploads and downloads address scene `index` from now on."""
        self._call("select_scene", C.c_uint32(int(index)))

    def step_scenes(self) -> list[dict]:
        """One step of every scene; per-scene stats."""
        n = self.scene_count()
        arr = (StepStats * n)()
        self._call("step_scenes", C.cast(arr, C.c_void_p), C.c_size_t(n))
        return [s.as_dict() for s in arr]

    def load_scenes(self, scenes, min_iterations=None, max_iterations=None, **options):
        """A batch: scene k of `scenes` loaded (as load_scene does) into scene k."""
        self.set_scene_count(len(scenes))
        for k, sc in enumerate(scenes):
            self.select_scene(k)
            self.load_scene(sc, min_iterations=min_iterations, max_iterations=max_iterations, **options)
        self.select_scene(0)

    def prepare(self):
        """Apply pending setter changes now (see pcisph_prepare)."""
        self._call("prepare")

    def profile_enable(self, on=True, every: int = 1):
        """Stage timing with HIP events: off, or every `every`-th step (the others replay
        their graphs untimed)."""
        self._call("profile_enable", C.c_int(max(1, int(every)) if on else 0))

    def profile_read(self) -> dict:
        t = StageTimes()
        self._call("profile_read", C.byref(t))
        return {"ms": list(t.ms)[:STAGE_COUNT], "launches": list(t.launches)[:STAGE_COUNT]}

    # --- learned warm start
    def mlp_set_params(self, params: np.ndarray):
        a = np.ascontiguousarray(params, np.float32).ravel()
        self._call("mlp_set_params", a.ctypes.data_as(C.c_void_p), C.c_size_t(a.size))

    def mlp_get_params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        self._call("mlp_get_params", out.ctypes.data_as(C.c_void_p), C.c_size_t(MLP_PARAMS))
        return out

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        y = np.empty(x.shape[0], np.float32)
        self._call("mlp_forward", x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p),
                   C.c_size_t(x.shape[0]))
        return y

    def mlp_gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32).ravel()
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._call("mlp_gradients", x.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                   C.c_size_t(x.shape[0]), g.ctypes.data_as(C.c_void_p), C.byref(loss))
        return g, loss.value

    def samples_reserve(self, capacity: int):
        self._call("samples_reserve", C.c_size_t(capacity))

    def samples_collect(self, stride: int):
        self._call("samples_collect", C.c_uint32(stride))

    def samples_count(self) -> int:
        n = C.c_size_t()
        self._call("samples_count", C.byref(n))
        return n.value

    def samples_download(self):
        n = self.samples_count()
        x = np.empty((n, MLP_INPUTS), np.float32)
        t = np.empty(n, np.float32)
        self._call("samples_download", x.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                   C.c_size_t(n))
        return x, t

    def samples_upload(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32).ravel()
        self._call("samples_upload", x.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                   C.c_size_t(x.shape[0]))

    def mlp_train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        loss = C.c_float()
        self._call("mlp_train", C.c_uint32(steps), C.c_uint32(batch), C.c_float(lr), C.c_uint64(seed),
                   C.byref(loss))
        return loss.value

    def autotune(self, trial_steps: int = 3) -> dict:
        """Pick block size, radix digit width and the pressure pass's tiling (work_order
        bits 4/5) by timing trial steps (state unchanged)."""
        o = SolverOptions()
        self._call("autotune", C.c_uint32(trial_steps), C.byref(o))
        return {"block_size": o.block_size, "sort_digit_bits": o.sort_digit_bits, "work_order": o.work_order}

    def translate_boundary(self, dx: float, dy: float, dz: float):
        """Move the walls rigidly by (dx, dy, dz); the boundary is re-sorted next step."""
        self._call("translate_boundary", dx, dy, dz)

    def translate_boundaries(self, shifts):
        """Move every scene's walls at once: shifts[k] = (dx, dy, dz) of scene k."""
        a = np.ascontiguousarray(shifts, np.float32).ravel()
        self._call("translate_boundaries", a.ctypes.data_as(C.c_void_p), C.c_size_t(a.size))

    def snapshot(self):
        """Keep the state of every scene on the device (see restore)."""
        self._call("snapshot")

    def restore(self):
        """Back to the last snapshot, bit for bit."""
        self._call("restore")

    def set_warm_start(self, on: bool = True):
        self._call("set_warm_start", C.c_int(1 if on else 0))

    def download_f64(self, field: int) -> np.ndarray:
        n = self.field_size(field)
        out = np.empty(n, dtype=np.float64)
        self._call("download_f64", C.c_int(int(field)), out.ctypes.data_as(C.c_void_p), C.c_size_t(n))
        return out

    def upload_device(self, field: int, device_ptr: int, count: int):
        self._call("upload_device", C.c_int(int(field)), C.c_void_p(device_ptr), C.c_size_t(count))

    def download_device(self, field: int, device_ptr: int, count: int):
        self._call("download_device", C.c_int(int(field)), C.c_void_p(device_ptr), C.c_size_t(count))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pcisph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Fluid", "Field", "PcisphError", "SimulationParams", "SolverOptions", "StepStats",
           "load_library", "compute_params", "field_dtype", "PRECISION_FAST", "PRECISION_IEEE",
           "PRECISION_FP64", "EXPORTS", "MLP_INPUTS", "MLP_PARAMS"]
