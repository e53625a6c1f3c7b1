"""1D whole-run launches (MFD_OPT_RUN1D, the default for 1D models without TFSF sources):
a model stepped by one launch per mfd_model_step call gives the same bits as the same
model stepped by the per-step launches, for every field, monitor and integral, with
FP32 and FP64 state, Lorentz, Kerr + plasma and Debye materials, soft and hard sources,
point, summed and spectral monitors, stepped in uneven chunks across partial DFT blocks;
and it matches the FP64 oracle."""
import numpy as np
import pytest

import cases
from maxwellfdtd_amd import FieldMonitor, SpectralMonitor, _lib
from maxwellfdtd_amd.batch import Batch
from maxwellfdtd_amd.models import Handle, MONITOR_COMPONENTS
from oracle_util import field_errors, oracle, rel_l2

pytestmark = pytest.mark.gpu

CHUNKS = (37, 1, 90, 200)    # 328 steps, then the rest of Nt


def _case(name):
    if name == "slab":
        return cases.slab_1d(steps=400)
    if name == "plain":         # simple materials: the register form
        return cases.slab_1d(steps=400, dispersive=False)
    if name == "config":        # the 1D configuration at full length, register form
        from maxwellfdtd_amd.workloads import slab_1d
        return slab_1d()
    if name == "nonlinear":
        return cases.nonlinear_1d(steps=400)
    model, _, _ = cases.ref_1d_test()     # hard source, vacuum
    L = model.grid.z[-1]
    mons = [FieldMonitor(lambda z: (z > 0.6 * L) & (z < 0.7 * L), sum=True),
            FieldMonitor((0.8 * L,)),
            SpectralMonitor(lambda z: (z > 0.5 * L) & (z < 0.52 * L), w=2e15 * np.linspace(0.5, 1.5, 20),
                            sum=True)]
    return model, mons


def _run(desc, run1d, f64):
    h = Handle(desc, _lib.engine())
    h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_RUN1D, run1d), "set_options(RUN1D)")
    if f64:
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_F64, 1), "set_options(F64)")
    left = desc.Nt
    for n in CHUNKS:
        h.step(min(n, left))
        left -= min(n, left)
    if left:
        h.step(left)
    assert h.it == desc.Nt
    return h


def _outputs(desc, plasma):
    return [k for k in range(4) if (k < 2 and desc.integrals & (1 << k)) or (k >= 2 and plasma)]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("f64", [0, 1])
@pytest.mark.parametrize("name", ["plain", "config", "slab", "nonlinear", "ref"])
def test_run1d_equals_per_step_launches(name, f64, mode):
    """mode 1: the register form where the model allows it (plain, config), else the
    L2 form; mode 2: the L2 form"""
    model, mons = _case(name)
    desc = model.describe(monitors=mons)
    print(f"1D {name}, f64={f64}, run1d={mode}: Nz={desc.dims}, Nt={desc.Nt}", flush=True)
    ha = _run(desc, mode, f64)
    hb = _run(desc, 0, f64)
    for c in ("Hy", "Dx", "Ex"):
        x, y = ha.field(c), hb.field(c)
        assert np.array_equal(x, y), (c, float(np.max(np.abs(x - y))))
    for k in _outputs(desc, model.isplasma):
        assert np.array_equal(ha.output(k), hb.output(k)), k
    for q in range(len(desc.monitors)):
        for c in MONITOR_COMPONENTS[1]:
            x, y = ha.monitor(q, c), hb.monitor(q, c)
            assert np.array_equal(x, y), (q, c)
    # and the whole-run launches against the FP64 oracle
    ho = Handle(desc, oracle("f64"))
    ho.step(desc.Nt)
    errs = field_errors(ha, ho, 1)
    # FP32 state over the 8000 steps of the configuration: 1.0e-4 to 1.2e-4, FP32's own
    # drift once the pulse has left (the oracle's FP32 build gives the same)
    tol = 1e-9 if f64 else (1.5e-4 if name == "config" else 1e-4)
    assert max(errs.values()) <= tol, errs
    merrs = {}
    for q in range(len(desc.monitors)):
        for c in MONITOR_COMPONENTS[1]:
            ref = ho.monitor(q, c)
            if np.max(np.abs(ref)) > 0:
                merrs[f"mon{q}{c}"] = rel_l2(ha.monitor(q, c), ref)
    # monitor samples are stored in FP32 whatever the state's precision
    assert max(merrs.values()) <= (1e-6 if f64 else 1e-4), merrs
    for h in (ha, hb, ho):
        h.close()


@pytest.mark.parametrize("dispersive", [False, True])
def test_run1d_batch_of_uneven_models(dispersive):
    """Several 1D models of different lengths and Nt in one batch: one workgroup each
    (the register form for the simple materials); every model equals its own run."""
    descs = []
    for nz, steps in ((300, 120), (600, 260), (450, 200)):
        model, mons = cases.slab_1d(nz=nz, steps=steps, dispersive=dispersive)
        descs.append(model.describe(monitors=mons))
    b = Batch(descs, _lib.engine())
    b.run(0)
    for d, hb in zip(descs, b.handles):
        assert hb.it == d.Nt
        ha = Handle(d, _lib.engine())
        ha.lib.check(ha.lib.model_set_options(ha.h, _lib.MFD_OPT_RUN1D, 0), "set_options(RUN1D)")
        ha.step(d.Nt)
        for c in ("Hy", "Dx", "Ex"):
            assert np.array_equal(ha.field(c), hb.field(c)), c
        for q in range(len(d.monitors)):
            for c in MONITOR_COMPONENTS[1]:
                assert np.array_equal(ha.monitor(q, c), hb.monitor(q, c)), (q, c)
        ha.close()
    b.close()
