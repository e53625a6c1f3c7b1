"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
numpy restatement oracle/pyref.py): the C oracle reproduces them at round-off on the
CPU, and the GPU engine within the 1e-4 FP32 bar."""
import os
import sys

import numpy as np
import pytest

from maxwellfdtd_amd.models import Handle
from oracle_util import oracle, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def run(name, lib):
    model, mons, desc = make_golden.build(name)
    h = Handle(desc, lib)
    h.step(model.Nt)
    return h, desc


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_golden(name):
    g = load(name)
    h, desc = run(name, oracle("f64"))
    assert int(g["steps"]) == desc.Nt
    for c in make_golden.KEEP[desc.ndims]:
        a = h.field(c)
        scale = max(np.abs(g[c]).max(), 1e-300)
        assert np.max(np.abs(a - g[c])) <= 1e-11 * scale, c
    if "Sa" in g:
        assert rel_l2(h.output(0), g["Sa"]) < 1e-12
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_golden(name):
    from maxwellfdtd_amd import _lib
    g = load(name)
    h, desc = run(name, _lib.engine())
    keep = make_golden.KEEP[desc.ndims]
    a = np.concatenate([h.field(c) for c in keep])
    b = np.concatenate([g[c] for c in keep])
    assert rel_l2(a, b) <= 1e-4
    if "Sa" in g:
        assert rel_l2(h.output(0), g["Sa"]) <= 1e-4
    h.close()
