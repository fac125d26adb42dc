"""The headless C++ driver (pcisph_amd/pcisph_headless): scene generation bit-equal to
pcisph_amd.scenes, the reference's XRAW scenes, its command line (no GPU)."""
import os
import subprocess

import numpy as np
import pytest

from pcisph_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pcisph_amd", "pcisph_headless")
REF_SCENES = "/root/reference/PCISPH/data/scenes"
FIXTURES = os.path.join(ROOT, "tests", "golden", "xraw")


def _emit(scene, tmp_path):
    out = tmp_path / "scene.bin"
    res = subprocess.run([DRIVER, "-i", scene, "--emit-scene", str(out)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    raw = out.read_bytes()
    n, nb = np.frombuffer(raw[:8], np.uint32)
    data = np.frombuffer(raw[8:], np.float32)
    return data[:3 * n], data[3 * n:3 * n + 3 * nb]


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_cpp_scene_matches_python(tmp_path):
    fluid, boundary = _emit("dam_break_16k", tmp_path)
    ref = scenes.dam_break_16k()
    np.testing.assert_array_equal(fluid, ref.fluid)
    np.testing.assert_array_equal(boundary, ref.boundary)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
@pytest.mark.parametrize("name,nf,nb", [("dambreak", 279936, 154980), ("cube_splash", 438976, 322272),
                                        ("simple_drop", 27000, 0)])
def test_reference_xraw_scenes(tmp_path, name, nf, nb):
    """Particle counts of the reference's scenes as its loader builds them
    (PCISPH/src/scenes.cpp:76-149,229-240), from the fixture copies of its scene files."""
    fluid, boundary = _emit(f"xraw:{FIXTURES}:{name}", tmp_path)
    assert fluid.size == 3 * nf and boundary.size == 3 * nb
    # the scaling of scene::load: cube of 0.7 m per voxel centred on the origin
    assert np.abs(fluid).max() <= 0.7 * 14 / 2 + 1e-5


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference scenes absent")
@pytest.mark.parametrize("name", ["dambreak", "cube_splash", "simple_drop"])
def test_xraw_fixtures_are_the_reference_files(name):
    with open(os.path.join(REF_SCENES, name + ".xraw"), "rb") as a, \
            open(os.path.join(FIXTURES, name + ".xraw"), "rb") as b:
        assert a.read() == b.read()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
@pytest.mark.parametrize("args,message", [(["-x"], "Invalid argument: -x"), ([], "-i <scene_name>"),
                                          (["-i", "dam_break_16k", "-d", "soon"], "Invalid value for -d")])
def test_command_line_errors(args, message):
    """Unknown arguments and a missing scene are refused before any device work, as the
    reference's parser does (PCISPH/src/main.cpp:47-59)."""
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True)
    assert res.returncode == 2 and message in res.stdout
