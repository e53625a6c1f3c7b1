"""The reference's files written and read by runs on the MI355X engine: the solve! output
file against the FP64 oracle's, and TFSF record files recorded on the engine and replayed
on the engine (src/outputs.jl:71-485, src/sources.jl:264-317,387-466,592-710,779-877).
Bar: relative L2 <= 1e-4 against the FP64 oracle (the FP32 state's tolerance), 1e-9 for
the FP64 1D state."""
import numpy as np
import pytest

import cases
from maxwellfdtd_amd import FieldMonitor, SpectralMonitor, TFSFSource, Model, solve
from maxwellfdtd_amd.hdf5 import H5File
from maxwellfdtd_amd.models import Handle, MI355X
from maxwellfdtd_amd.outputs import COMPONENTS
from oracle_util import field_errors, oracle, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4
F64_TOL = 1e-9


def _cases():
    m1, mons1 = cases.slab_1d(nz=300, steps=300)
    m2, mons2 = cases.lorentz_2d(nx=48, nz=40, steps=100)
    m3 = cases.nonlinear_3d(n=20, steps=80)
    mons3 = [FieldMonitor((4e-7, 4e-7, 4.5e-7)), SpectralMonitor((4e-7, 4e-7, 4.5e-7), w=[2e15, 3e15])]
    return [(m1, mons1), (m2, mons2), (m3, mons3)]


def _compare_files(fa, fb, model, tol):
    """Two output files of one model: same datasets, field frames, integrals and monitors
    within tol (relative L2)."""
    assert fa.keys() == fb.keys()
    assert fa.keys("fields") == fb.keys("fields")
    nd = model.grid.ndims
    for c in COMPONENTS[nd] + (("rho",) if model.isplasma else ()):
        a, b = fa[f"fields/{c}"], fb[f"fields/{c}"]
        assert a.shape == b.shape
        # frames where a component is still tiny (a plasma density growing as I^K from
        # rounding-level fields) are ill-conditioned; they are compared once they carry
        # a thousandth of the component's largest value.        # a density growing as I^K (K = 6 here) carries 2K times the field's relative
        # error: plasma quantities get ten times the field tolerance
        peak = np.max(np.abs(b))
        t = 10 * tol if c == "rho" else tol
        for k in range(a.shape[-1]):
            if np.max(np.abs(b[..., k])) > 1e-3 * peak:
                assert rel_l2(a[..., k], b[..., k]) <= t, (c, k)
    for name in ("Sa", "E2"):
        assert rel_l2(fa[name], fb[name]) <= tol, name
    if model.isplasma:
        assert rel_l2(fa["rho_end"], fb["rho_end"]) <= 10 * tol
        assert rel_l2(fa["JE"], fb["JE"]) <= 10 * tol
    for n in fa.keys("monitors"):
        assert np.array_equal(fa[f"monitors/{n}/inds"], fb[f"monitors/{n}/inds"])
        a = np.concatenate([np.ravel(fa[f"monitors/{n}/{c}"]) for c in COMPONENTS[nd]])
        b = np.concatenate([np.ravel(fb[f"monitors/{n}/{c}"]) for c in COMPONENTS[nd]])
        assert rel_l2(a, b) <= tol, n


@pytest.mark.parametrize("k", [0, 1, 2])
def test_engine_output_file(tmp_path, k):
    model, mons = _cases()[k]
    pg, po = str(tmp_path / "gpu.hdf"), str(tmp_path / "cpu.hdf")
    print(f"output file case {k}: {model.grid.shape}, Nt={model.Nt}", flush=True)
    solve(model, monitors=mons, nframes=4, fname=pg, backend=MI355X())
    solve(model, monitors=mons, nframes=4, fname=po, lib=oracle("f64"))
    fg, fo = H5File(pg), H5File(po)
    assert fg["fields/Ex"].dtype == np.float32
    _compare_files(fg, fo, model, TOL)


def test_engine_output_file_f64(tmp_path):
    model, mons = cases.slab_1d(nz=300, steps=300)
    pg, po = str(tmp_path / "gpu.hdf"), str(tmp_path / "cpu.hdf")
    solve(model, monitors=mons, nframes=4, fname=pg, backend=MI355X(precision="f64"))
    solve(model, monitors=mons, nframes=4, fname=po, lib=oracle("f64"), precision="f64")
    fg, fo = H5File(pg), H5File(po)
    assert fg["fields/Ex"].dtype == np.float64
    for c in ("Hy", "Ex"):
        a, b = fg[f"fields/{c}"], fo[f"fields/{c}"]
        for j in range(a.shape[-1]):
            if np.any(b[..., j]):
                assert rel_l2(a[..., j], b[..., j]) <= F64_TOL, (c, j)
    assert rel_l2(fg["Sa"], fo["Sa"]) <= F64_TOL


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_engine_tfsf_record_replay(tmp_path, nd):
    """Record the plane wave on the engine, replay the file on the engine around a
    scatterer: equal to the engine's own plane-wave TFSF and to the oracle's replay of
    the same file, within the FP32 tolerance."""
    rec, aux, box, grid, tmax, mat = cases.plane_wave_pair(nd, n=24, steps=70)
    path = str(tmp_path / "rec.h5")
    print(f"tfsf record/replay {nd}D", flush=True)
    solve(rec, tfsf_record=True, tfsf_box=box, tfsf_fname=path, backend=MI355X())
    replay = Model(grid, TFSFSource(fname=path), tmax=tmax, CN=0.99, pml=2 * 40e-9, material=mat)
    from maxwellfdtd_amd import _lib
    hg = Handle(replay.describe(), _lib.engine())
    hg.step(replay.Nt)
    ha = Handle(aux.describe(), _lib.engine())
    ha.step(aux.Nt)
    ho = Handle(replay.describe(), oracle("f64"))
    ho.step(replay.Nt)
    e1 = field_errors(hg, ha, nd)
    e2 = field_errors(hg, ho, nd)
    assert max(e1.values()) <= TOL, e1
    assert max(e2.values()) <= TOL, e2
    assert np.max(np.abs(ho.field("Ex"))) > 1e-2


def test_engine_rate_given_as_a_function():
    """An ionization rate given as a plain function of the intensity (the reference's
    `ionrate` closure) is tabulated on the host and runs on the engine within the
    tolerance of the FP64 oracle running the same table."""
    from maxwellfdtd_amd import _lib
    model = cases.nonlinear_3d(n=20, steps=90, rate="function")
    desc = model.describe()
    hg = Handle(desc, _lib.engine())
    hg.step(model.Nt)
    ho = Handle(desc, oracle("f64"))
    ho.step(model.Nt)
    errs = field_errors(hg, ho, 3)
    assert max(errs.values()) <= TOL, errs
    rho = ho.output(3)
    assert np.max(rho) > 0
    assert rel_l2(hg.output(3), rho) <= 10 * TOL
