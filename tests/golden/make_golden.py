"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference ships no fixtures or golden vectors (and its step cannot run here), so
the fixtures are oracle output on small seeded scenes: the inputs are regenerated
from the seeds in pcisph_amd.scenes, the expected outputs are stored.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle_py import OracleFluid  # noqa: E402
from pcisph_amd import scenes  # noqa: E402
from pcisph_amd.abi import Field  # noqa: E402

FIELDS = [Field.FLUID_KEYS, Field.FLUID_SRC_LOCATIONS, Field.FLUID_CELL_OFFSETS,
          Field.FLUID_NEIGHBOR_COUNTS, Field.BOUNDARY_SRC_LOCATIONS, Field.FLUID_POSITIONS,
          Field.FLUID_VELOCITIES, Field.FLUID_DENSITIES, Field.FLUID_PRESSURES,
          Field.FLUID_PRESSURE_FORCES, Field.FLUID_OTHER_FORCES]

CASES = {
    # name: (scene factory, steps, max_iterations)
    "small_dam_break_3steps": (lambda: scenes.small_dam_break(), 3, 20),
    "dam_break_1000_2steps": (lambda: scenes.dam_break((10, 10, 10), 0.02, 0.001, (0.4, 0.4, 0.3),
                                                       seed=99), 2, 20),
}


def generate(name):
    make, steps, max_it = CASES[name]
    scene = make()
    o = OracleFluid()
    o.load_scene(scene, max_iterations=max_it)
    stats = [o.step() for _ in range(steps)]
    out = {Field(f).name: o.download(f) for f in FIELDS}
    out["iterations"] = np.array([s["iterations"] for s in stats], np.uint32)
    out["max_density_variation"] = np.array([s["max_density_variation"] for s in stats], np.float32)
    out["params"] = np.frombuffer(o.get_params().raw(), np.uint8)
    return scene, out


def main():
    for name in CASES:
        _, out = generate(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
