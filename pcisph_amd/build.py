"""Build recipe of libpcisph_amd.so: hipcc for gfx950, one object per translation
unit, linked into a shared library inside the package directory.

    python -m pcisph_amd.build          # or pcisph_amd.build.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpcisph_amd.so")
DRIVER = os.path.join(HERE, "pcisph_headless")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep v_pk_*_f32 out of the generated code.
COMMON = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall",
          "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]

# translation unit -> extra flags.  IEEE units forbid contraction so FP32 rounds
# once per written operation; the fast unit allows fused multiply-adds.
UNITS = {
    "solver.hip": ["-ffp-contract=off"],
    "radix_sort.hip": ["-ffp-contract=off"],
    "radix_onesweep.hip": ["-ffp-contract=off"],
    "sph_ieee.hip": ["-ffp-contract=off"],
    "sph_fast.hip": ["-ffp-contract=fast"],
    "mlp.hip": ["-ffp-contract=fast"],
}


def _compile(unit: str, flags: list[str]) -> str:
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, *flags, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every HIP translation unit for gfx950 and link libpcisph_amd.so."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # headless driver: C++ host program over the C ABI (rpath to this directory)
    drv = os.path.join(HERE, "driver", "headless.cpp")
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", drv, "-o", DRIVER + ".tmp",
           f"-L{HERE}", "-lpcisph_amd", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {LIB} and {DRIVER}")
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
