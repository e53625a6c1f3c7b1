"""The 3D workloads at their full sizes for a short prefix, engine against the FP64 oracle
on the same inputs: the 256^3 plane wave through 40 dielectric spheres (100 steps) and
the 384x384x192 dispersive plasmonics model with its 384x384 DFT plane (50 steps).
Bar: relative L2 <= 1e-4 for the H, D and E vector fields and the integrals, the same
for the total field energy, and for the DFT spectra relative L2 <= 1e-4 with the largest
per-spectrum error (max over frequencies of |S - S_ref| over max |S_ref|) <= 1e-3."""
import os

import numpy as np
import pytest

from maxwellfdtd_amd import _lib
from maxwellfdtd_amd.models import Handle
from maxwellfdtd_amd.workloads import dispersive_3d, plane_wave_spheres_3d
from oracle_util import field_energy, field_errors, oracle, rel_l2, spectra_errors

pytestmark = pytest.mark.gpu

TOL = 1e-4
DFT_MAX_REL = 1e-3


def _pair(model, mons, steps):
    desc = model.describe(monitors=mons)
    print(f"full-size prefix: {desc.dims} cells, {steps} steps", flush=True)
    hg = Handle(desc, _lib.engine())
    hg.step(steps)
    orc = oracle("f64")
    orc.set_threads(len(os.sched_getaffinity(0)))
    ho = Handle(desc, orc)
    ho.step(steps)
    return hg, ho, desc


def _check(hg, ho, desc):
    errs = field_errors(hg, ho, 3)
    for k, name in ((0, "Sa"), (1, "E2")):
        if desc.integrals & (1 << k):
            errs[name] = rel_l2(hg.output(k), ho.output(k))
    assert max(errs.values()) <= TOL, errs
    ge, gh = hg.energy()
    oe, oh = field_energy(ho)
    assert oe > 0 and oh > 0
    assert abs(ge / oe - 1) <= TOL and abs(gh / oh - 1) <= TOL, (ge, oe, gh, oh)
    return errs


def test_plane_wave_spheres_256_prefix():
    model = plane_wave_spheres_3d(n=256, steps=100)
    hg, ho, desc = _pair(model, None, 100)
    _check(hg, ho, desc)


def test_dispersive_384x384x192_prefix_with_dft_plane():
    model, mons = dispersive_3d(steps=50)
    hg, ho, desc = _pair(model, mons, 50)
    _check(hg, ho, desc)
    # the DFT plane lies on the TFSF box's upper z face, where the face correction puts
    # the incident pulse's leading tail from the first step
    spec = spectra_errors(hg, ho, desc)
    assert np.max(np.abs(ho.monitor(0, "Ex"))) > 0
    assert spec["dft0_rel_l2"] <= TOL and spec["dft0_max_rel"] <= DFT_MAX_REL, spec


def test_plane_wave_spheres_256_alone_equals_batched():
    """A lone model whose two field sets outgrow the Infinity Cache gets its new fields
    stored non-temporally; the same model batched with another keeps ordinary stores.
    The cache policy changes no bit of the fields or the integrals."""
    from maxwellfdtd_amd.batch import Batch

    steps = 24
    big = plane_wave_spheres_3d(n=256, steps=steps).describe()
    small = plane_wave_spheres_3d(n=64, steps=steps, nspheres=4).describe()
    print(f"256^3 alone and batched with 64^3, {steps} steps", flush=True)
    ha = Handle(big, _lib.engine())
    ha.step(steps)
    b = Batch([big, small], _lib.engine())
    b.run(0)
    hb = b.handles[0]
    for c in ("Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ex", "Ey", "Ez"):
        x, y = ha.field(c), hb.field(c)
        assert np.array_equal(x, y), (c, float(np.max(np.abs(x - y))))
    for k in (0, 1):
        assert np.array_equal(ha.output(k), hb.output(k)), k
    assert np.max(np.abs(ha.field("Ex"))) > 0
    ha.close()
    b.close()
