"""GPU: the headless C++ driver (pcisph_amd/pcisph_headless), the reference's application
loop without a window (PCISPH/src/main.cpp:157-228): its steps against the oracle, the
recording cadence of -o (a frame each 1/60 s of simulated time), -d as the simulated
duration, a resumed run bit-equal to an uninterrupted one, and the reference's own XRAW
scenes (scenes.cpp:76-149,217-245) -- simple_drop among them, whose boundary count is 0
(the reference passes an unallocated boundary table there; here it means no boundary)."""
import os
import subprocess

import numpy as np
import pytest

from pcisph_amd import scenes
from pcisph_amd.abi import Field

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pcisph_amd", "pcisph_headless")
FIXTURES = os.path.join(ROOT, "tests", "golden", "xraw")
THREADS = min(16, os.cpu_count() or 1)


def drive(*args, cwd=None):
    print("case driver", " ".join(args), flush=True)
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300, cwd=cwd)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def read_frame(path):
    """Frame file: "PCSF", u32 version, u32 N, u32 Nb, u64 step, float[3N] positions,
    float[3N] velocities, float[N] densities, float[3Nb] boundary positions."""
    raw = open(path, "rb").read()
    assert raw[:4] == b"PCSF"
    n, nb = np.frombuffer(raw[8:16], np.uint32)
    step = int(np.frombuffer(raw[16:24], np.uint64)[0])
    a = np.frombuffer(raw[24:], np.float32)
    x, v, rho, b = np.split(a, np.cumsum([3 * n, 3 * n, n]))
    return {"step": step, "x": x, "v": v, "rho": rho, "b": b}


def oracle_frame(oracle_lib, scene, steps, **opts):
    oracle_lib.set_threads(THREADS)
    o = oracle_lib.OracleFluid()
    o.load_scene(scene, **opts)
    for _ in range(steps):
        o.step()
    return {"x": o.download(Field.FLUID_POSITIONS), "v": o.download(Field.FLUID_VELOCITIES),
            "rho": o.download(Field.FLUID_DENSITIES), "b": o.download(Field.BOUNDARY_POSITIONS)}


def assert_frame_equal(got, want, tag):
    for key in ("x", "v", "rho", "b"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=f"{tag} {key}")


def test_driver_steps_match_oracle(oracle_lib, tmp_path):
    drive("-i", "dam_break_16k", "-n", "3", "--precision", "ieee", "--final", str(tmp_path / "f.pcsf"))
    got = read_frame(tmp_path / "f.pcsf")
    assert got["step"] == 3
    assert_frame_equal(got, oracle_frame(oracle_lib, scenes.dam_break_16k(), 3), "dam_break_16k")


def test_driver_records_and_resumes_bit_exact(tmp_path):
    out = tmp_path / "out"
    log = drive("-i", "dam_break_16k", "-n", "80", "-o", "--out-dir", str(out), "--precision", "ieee",
                "--final", str(tmp_path / "a.pcsf"))
    # frames after the first step, then whenever 1/60 s of simulated time (float sum of
    # dt, main.cpp:191-222) has passed since the last one
    dt, t, last, expect = np.float32(5e-4), np.float32(0), -np.inf, []
    for k in range(1, 81):
        t = np.float32(t + dt)
        if t - last >= np.float32(1.0 / 60.0):
            expect.append(k)
            last = t
    frames = sorted(out.glob("frame_*.pcsf"), key=lambda p: int(p.stem.split("_")[1]))
    assert [read_frame(p)["step"] for p in frames] == expect, log
    assert len(expect) >= 3
    mid = frames[1]
    s_mid = read_frame(mid)["step"]
    drive("-i", "dam_break_16k", "--resume", str(mid), "-n", str(80 - s_mid), "--precision", "ieee",
          "--final", str(tmp_path / "b.pcsf"))
    a, b = read_frame(tmp_path / "a.pcsf"), read_frame(tmp_path / "b.pcsf")
    assert a["step"] == b["step"] == 80
    assert_frame_equal(b, a, "resumed")


def test_driver_duration_is_simulated_time(tmp_path):
    """-d <ms> stops once the float sum of the time steps reaches ms / 1000 (main.cpp:43-45,224)."""
    log = drive("-i", "dam_break_16k", "-d", "10")
    t, k = np.float32(0), 0
    while not t >= np.float32(0.001) * np.float32(10):
        t = np.float32(t + np.float32(5e-4))
        k += 1
    assert f"{k} steps (simulated time" in log, log


def reference_scene(name, tmp_path):
    """The reference's scene::load (scenes.cpp:217-245) from the fixture scene files."""
    emit = tmp_path / f"{name}.bin"
    drive("-i", f"xraw:{FIXTURES}:{name}", "--emit-scene", str(emit))
    raw = emit.read_bytes()
    n, nb = np.frombuffer(raw[:8], np.uint32)
    data = np.frombuffer(raw[8:], np.float32)
    ppd = {"dambreak": 18, "cube_splash": 19, "simple_drop": 30}[name]
    radius = float((np.float32(0.5) * np.float32(0.7)) / np.float32(ppd))
    return scenes.Scene(name=name, fluid=data[:3 * n].copy(), boundary=data[3 * n:3 * n + 3 * nb].copy(),
                        particle_radius=radius, rest_density=999.972, viscosity=0.00008, surface_tension=1.0,
                        gravity=0.0 if name == "simple_drop" else -9.81, delta_t=0.002, threshold=0.01,
                        min_iterations=3, max_iterations=7)


@pytest.mark.parametrize("name", ["simple_drop", "dambreak", "cube_splash"])
def test_reference_scenes_match_oracle(oracle_lib, tmp_path, name):
    scene = reference_scene(name, tmp_path)
    drive("-i", f"xraw:{FIXTURES}:{name}", "-n", "2", "--precision", "ieee", "--final", str(tmp_path / "f.pcsf"))
    got = read_frame(tmp_path / "f.pcsf")
    assert got["b"].size == scene.boundary.size
    assert_frame_equal(got, oracle_frame(oracle_lib, scene, 2), name)
