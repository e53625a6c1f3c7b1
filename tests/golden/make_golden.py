"""Writes the golden fixtures of the per-step update: small seeded models with every
feature switched on (tests/cases.py), stepped by oracle/pyref.py -- the whole-array
numpy restatement written from the update equations, independent of the cell-by-cell C
oracle -- and stored as float64 arrays.

    python tests/golden/make_golden.py

The reference (Julia) cannot run here, so these vectors come from the restatement; the
tests hold the C oracle to them at round-off and the GPU engine at 1e-4.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

import cases  # noqa: E402
from pyref import PyRef  # noqa: E402

BC = ("dirichlet", "neumann", "neumann", "periodic", "periodic", "dirichlet")
CASES = {
    "spheres_3d": lambda: cases.spheres_3d(n=14, steps=30, nspheres=3, pml=3),
    "dispersive_3d": lambda: (cases.dispersive_3d(n=14, steps=30), None),
    "nonlinear_3d": lambda: (cases.nonlinear_3d(n=14, steps=30), None),
    "bcs_3d": lambda: (cases.spheres_3d(n=12, steps=20, bc=BC, monitors=False, pml=0)[0], None),
    "lorentz_2d": lambda: cases.lorentz_2d(nx=36, nz=30, steps=40),
    "slab_1d": lambda: cases.slab_1d(nz=150, steps=150),
}
KEEP = {3: ("Hy", "Dz", "Ex", "Ez"), 2: ("Hy", "Dx", "Dz", "Ex", "Ez"), 1: ("Hy", "Dx", "Ex")}
NAMES = "Hx Hy Hz Dx Dy Dz Ex Ey Ez".split()


def build(name):
    model, mons = CASES[name]()
    return model, mons, model.describe(monitors=mons)


def main():
    for name in CASES:
        model, mons, desc = build(name)
        p = PyRef(desc)
        p.run(model.Nt)
        out = {"steps": np.int64(model.Nt)}
        for c in KEEP[desc.ndims]:
            out[c] = np.asarray(p.field(NAMES.index(c)), dtype=np.float64)
        if desc.integrals & 1:
            out["Sa"] = np.asarray(p.Sa.ravel(order="F"), dtype=np.float64)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, model.Nt, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
