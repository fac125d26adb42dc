"""Copies the reference's three scene files (PCISPH/data/scenes/*.xraw: 24-byte header,
X*Y*Z voxel indices, palette -- data, 1-2.3 KB each) into tests/golden/xraw/, so the GPU
tests can build the reference's scenes where the reference is absent.

    python tests/golden/make_xraw_fixtures.py
"""
import os
import shutil

SRC = "/root/reference/PCISPH/data/scenes"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "xraw")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ("dambreak", "cube_splash", "simple_drop"):
        shutil.copyfile(os.path.join(SRC, name + ".xraw"), os.path.join(DST, name + ".xraw"))
        print(name, os.path.getsize(os.path.join(DST, name + ".xraw")), "bytes")
