"""The files the reference exchanges, on the CPU: the HDF5 writer and reader, the solve!
output layout (src/outputs.jl:46-65,71-485, src/monitors.jl:72-78,362-373), the TFSF
record files and their replay on the recording grid and on other grids
(src/sources.jl:264-317,387-466,592-710,779-877), and tabulated ionization rates.

Where the image holds an independent HDF5 implementation (h5py over libhdf5, run in a
separate interpreter), files written here are read back by it and files it writes are
read here; without it those checks are skipped."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import cases
from maxwellfdtd_amd import (C0, EPS0, FieldMonitor, Grid, Material, Model, Plasma,
                             PowerLawRate, SoftSource, SpectralMonitor, TFSFSource, solve)
from maxwellfdtd_amd.hdf5 import H5File, H5Writer
from maxwellfdtd_amd.models import Handle
from maxwellfdtd_amd.tfsf_files import NAMES, interp_flat, record_plan
from oracle_util import field_errors, oracle, rel_l2

H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _h5py_available():
    if not os.path.exists(H5PY_PYTHON):
        return False
    r = subprocess.run([H5PY_PYTHON, "-I", "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


needs_h5py = pytest.mark.skipif(not _h5py_available(), reason="no h5py interpreter in this image")


def h5py_summary(path):
    """Every dataset of a file as h5py sees it: shape, dtype, sum of |values|, values of
    small datasets."""
    code = r'''
import h5py, numpy as np, json, sys
out = {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        v = obj[()]
        e = {"shape": list(obj.shape), "dtype": str(obj.dtype),
             "abssum": float(np.sum(np.abs(v))) if np.size(v) else 0.0}
        if np.size(v) <= 16:
            a = np.ravel(v)
            e["values"] = [[float(np.real(x)), float(np.imag(x))] for x in a]
        out[name] = e
with h5py.File(sys.argv[1], "r") as f:
    f.visititems(visit)
print(json.dumps(out))
'''
    r = subprocess.run([H5PY_PYTHON, "-I", "-c", code, path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


# ---- HDF5 ------------------------------------------------------------------------------
def _sample_arrays(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "x": np.linspace(0, 1, 7),
        "pml": np.array([3, 5, 2, 7], dtype=np.int64),
        "geometry": rng.integers(0, 3, (4, 3, 5)).astype(np.int32),
        "fields/t": np.arange(3.0),
        "Sa": rng.standard_normal((4, 3, 5)),
        "small/f32": rng.standard_normal((2, 3)).astype(np.float32),
        "monitors/1/Ex": rng.standard_normal((2, 6)) + 1j * rng.standard_normal((2, 6)),
        "u8": np.array([0, 255, 7], dtype=np.uint8),
    }


def test_hdf5_round_trip(tmp_path):
    """Arrays keep their Julia shapes, values and types through write and read; frames
    written one at a time along the last dimension land where they belong."""
    path = str(tmp_path / "a.h5")
    arrs = _sample_arrays()
    frames = np.random.default_rng(1).standard_normal((4, 3, 5, 3)).astype(np.float32)
    with H5Writer(path) as w:
        for k, v in arrs.items():
            w.create(k, v)
        w.create("monitors/1/issum", np.bool_(True))
        w.create_dataset("fields/Ex", np.float32, frames.shape)
        for k in range(3):
            w.write_last("fields/Ex", k, frames[..., k])
    f = H5File(path)
    assert f.keys() == ["Sa", "fields", "geometry", "monitors", "pml", "small", "u8", "x"]
    assert f.keys("fields") == ["Ex", "t"]
    for k, v in arrs.items():
        got = f[k]
        assert got.shape == v.shape and got.dtype == v.dtype and np.array_equal(got, v), k
    assert np.array_equal(f["fields/Ex"], frames)
    assert f["monitors/1/issum"] == 1
    assert "fields/Ex" in f and "fields/Hx" not in f


@needs_h5py
def test_hdf5_files_read_by_libhdf5(tmp_path):
    """libhdf5 (through h5py) opens what the writer wrote and sees the HDF5.jl layout: the
    dimensions of a Julia array reversed, complex numbers as {r, i} compounds."""
    path = str(tmp_path / "a.h5")
    arrs = _sample_arrays()
    with H5Writer(path) as w:
        for k, v in arrs.items():
            w.create(k, v)
    s = h5py_summary(path)
    for k, v in arrs.items():
        e = s[k]
        assert e["shape"] == list(reversed(v.shape)), k
        assert np.isclose(e["abssum"], float(np.sum(np.abs(v))), rtol=1e-6), k
    assert s["monitors/1/Ex"]["dtype"] == "complex128"
    assert s["geometry"]["dtype"] == "int32" and s["small/f32"]["dtype"] == "float32"
    assert [v[0] for v in s["x"]["values"]] == pytest.approx(list(arrs["x"]))


@needs_h5py
def test_reader_opens_libhdf5_files(tmp_path):
    """Files libhdf5 writes (h5py defaults: symbol-table groups, here one with 70 links
    spread over several symbol table nodes) load through the reader."""
    path = str(tmp_path / "b.h5")
    code = r'''
import h5py, numpy as np, sys
with h5py.File(sys.argv[1], "w") as g:
    g["a"] = np.arange(24.0).reshape(2, 3, 4)
    g["grp/b"] = np.arange(5, dtype=np.int32)
    g["grp/c"] = np.array([1+2j, 3-4j])
    g["s"] = 3.5
    for k in range(70):
        g["many/d%02d" % k] = np.full(3, k, dtype=np.float32)
'''
    r = subprocess.run([H5PY_PYTHON, "-I", "-c", code, path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    f = H5File(path)
    assert f.keys() == ["a", "grp", "many", "s"]
    assert np.array_equal(f.read("a", fortran=False), np.arange(24.0).reshape(2, 3, 4))
    assert np.array_equal(f["grp/b"], np.arange(5)) and np.array_equal(f["grp/c"], [1 + 2j, 3 - 4j])
    assert f["s"] == 3.5 and len(f.keys("many")) == 70 and np.array_equal(f["many/d69"], [69, 69, 69])


# ---- solve! output file -----------------------------------------------------------------
def _output_cases():
    m1, mons1 = cases.slab_1d(nz=200, steps=150)
    m2, mons2 = cases.lorentz_2d(nx=40, nz=36, steps=60)
    m3 = cases.nonlinear_3d(n=14, steps=40)
    mons3 = [FieldMonitor((5e-7, 5e-7, 3e-7)), SpectralMonitor((5e-7, 5e-7, 3e-7), w=[2e15, 3e15])]
    return [(m1, mons1), (m2, mons2), (m3, mons3)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_output_file_layout(tmp_path, k):
    """solve(model; fname, nframes, monitors) writes the reference's layout, with the
    values the run returns: coordinates, t, pml, geometry, fields/{t, components, rho},
    Sa, E2, rho_end, JE and monitors/n/{issum, inds, w, components}."""
    model, mons = _output_cases()[k]
    path = str(tmp_path / "out.hdf")
    orc = oracle("f64")
    sol = solve(model, monitors=mons, nframes=3, lib=orc)
    solve(model, monitors=mons, nframes=3, fname=path, lib=orc)
    f = H5File(path)
    nd = model.grid.ndims
    axes = {1: ["z"], 2: ["x", "z"], 3: ["x", "y", "z"]}[nd]
    for a in axes:
        assert np.array_equal(f[a], getattr(model.grid, a))
    assert np.array_equal(f["t"], model.t)
    assert len(f["pml"]) == 2 * nd
    assert np.array_equal(f["geometry"], model.geometry)
    assert np.allclose(f["fields/t"], sol.tout)
    names = {1: ("Hy", "Ex"), 2: ("Hy", "Ex", "Ez"), 3: ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")}[nd]
    for c in names:
        data = f[f"fields/{c}"]
        assert data.shape == tuple(model.grid.shape) + (len(sol.tout),)
        for i in range(len(sol.tout)):
            assert np.allclose(data[..., i], sol.frames[c][i].astype(np.float32), rtol=0, atol=0), c
    assert np.array_equal(f["Sa"], sol.Sa)
    assert np.array_equal(f["E2"], sol.E2)
    if model.isplasma:
        assert np.array_equal(f["rho_end"], sol.rho_end) and np.array_equal(f["JE"], sol.JE)
        assert f["fields/rho"].shape[-1] == len(sol.tout)
    assert f.keys("monitors") == [str(i + 1) for i in range(len(mons))]
    for i, mon in enumerate(sol.monitors, start=1):
        assert f[f"monitors/{i}/issum"] == int(mon["issum"])
        assert f[f"monitors/{i}/inds"].shape[0] == nd
        for c in names:
            assert np.array_equal(f[f"monitors/{i}/{c}"], np.asarray(mon[c])), (i, c)
        if mon["kind"] == "spectral":
            assert np.array_equal(f[f"monitors/{i}/w"], mon["w"])
    if _h5py_available():
        s = h5py_summary(path)
        assert s["fields/Ex"]["shape"] == [len(sol.tout)] + list(reversed(model.grid.shape))
        assert "Sa" in s and "monitors/1/Ex" in s


def test_output_component_check():
    model, _ = cases.slab_1d(nz=100, steps=10)
    with pytest.raises(ValueError, match="1D field has only :Hy and :Ex components"):
        solve(model, nframes=2, components=("Ez",), lib=oracle("f64"))


# ---- TFSF record files ------------------------------------------------------------------
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_tfsf_record_and_same_grid_replay(tmp_path, nd):
    """solve(...; tfsf_record=true, tfsf_box) writes prepare_tfsf_record's datasets with
    their [face..., Nt] shapes; TFSFSource(fname) on the recording grid replays the
    recorded tables exactly, so around a scatterer it equals the plane-wave (auxiliary
    run) TFSF of the same wave bit for bit."""
    orc = oracle("f64")
    rec, aux, box, grid, tmax, mat = cases.plane_wave_pair(nd)
    path = str(tmp_path / "rec_tfsf.h5")
    solve(rec, tfsf_record=True, tfsf_box=box, tfsf_fname=path, lib=orc)
    f = H5File(path)
    assert set(NAMES[nd]) <= set(f.keys())
    assert np.allclose(f["box"], box) and np.array_equal(f["t"], rec.t)
    for name, comp, ranges in record_plan(grid, box):
        shape = tuple(len(r) for r in ranges if len(r) > 1) + (rec.Nt,)
        assert f[name].shape == shape, name
    replay = Model(grid, TFSFSource(fname=path), tmax=tmax, CN=0.99, pml=2 * 40e-9, material=mat)
    assert replay.source_structs[0]["box"] == aux.source_structs[0]["box"]
    h1 = Handle(aux.describe(), orc)
    h1.step(aux.Nt)
    h2 = Handle(replay.describe(), orc)
    h2.step(replay.Nt)
    errs = field_errors(h1, h2, nd)
    assert max(errs.values()) == 0.0, errs
    assert np.max(np.abs(h1.field("Ex"))) > 1e-2
    if _h5py_available():
        s = h5py_summary(path)
        assert s[NAMES[nd][0]]["shape"] == list(reversed(f[NAMES[nd][0]].shape))


def test_tfsf_replay_on_another_transverse_grid(tmp_path):
    """A 3D record replayed on a grid with coarser x and y spacing (same z grid and time
    step): the incident plane wave is uniform across x and y, so the interpolated tables
    are exact and the replay equals that grid's own plane-wave TFSF."""
    orc = oracle("f64")
    rec, _, box, grid, tmax, _ = cases.plane_wave_pair(3)
    path = str(tmp_path / "rec.h5")
    solve(rec, tfsf_record=True, tfsf_box=box, tfsf_fname=path, lib=orc)
    d = 40e-9
    n2, dx2 = 13, 1.25 * d
    g2 = Grid(xmin=0, xmax=(n2 - 1) * dx2, Nx=n2, ymin=0, ymax=(n2 - 1) * dx2, Ny=n2,
              zmin=0, zmax=grid.z[-1], Nz=grid.Nz)
    CN2 = rec.dt * C0 * math.sqrt(2 / dx2 ** 2 + 1 / d ** 2)
    w = cases.gauss_waveform(lam0=800e-9, tau0=3e-15, dt0=8e-15)
    replay = Model(g2, TFSFSource(fname=path), tmax=tmax, CN=CN2, pml=2 * d)
    aux = Model(g2, TFSFSource(box=box, waveform=w, component="Ex", z0=2 * d), tmax=tmax, CN=CN2,
                pml=2 * d)
    assert replay.Nt == rec.Nt and replay.source_structs[0]["box"] == aux.source_structs[0]["box"]
    h1 = Handle(aux.describe(), orc)
    h1.step(aux.Nt)
    h2 = Handle(replay.describe(), orc)
    h2.step(replay.Nt)
    errs = field_errors(h1, h2, 3)
    assert max(errs.values()) < 1e-12, errs
    assert np.max(np.abs(h1.field("Ex"))) > 1e-2


def test_interpolation_nodes_linearity_and_flat_ends():
    """linear_interpolation(knots, data; extrapolation_bc=Flat()) semantics: exact at the
    knots, exact for data linear along each axis (non-uniform knots included), and the
    end values beyond the knot range."""
    rng = np.random.default_rng(4)
    kx = np.sort(rng.uniform(0, 1, 7))
    kt = np.linspace(0, 2, 9)
    data = rng.standard_normal((7, 9))
    assert np.array_equal(interp_flat([kx, kt], data, [kx, kt]), data)
    lin = 3.0 + 2.0 * kx[:, None] - 0.5 * kt[None, :] + 0.25 * kx[:, None] * kt[None, :]
    qx = np.linspace(kx[0], kx[-1], 23)
    qt = np.linspace(0, 2, 31)
    want = 3.0 + 2.0 * qx[:, None] - 0.5 * qt[None, :] + 0.25 * qx[:, None] * qt[None, :]
    assert np.allclose(interp_flat([kx, kt], lin, [qx, qt]), want, rtol=0, atol=1e-12)
    out = interp_flat([kx, kt], data, [np.array([kx[0] - 1, kx[-1] + 5]), np.array([-3.0, 9.0])])
    assert np.array_equal(out, data[[0, -1]][:, [0, -1]])


# ---- ionization rate given as a function -------------------------------------------------
def test_rate_function_is_tabulated_like_the_shim():
    """A plain function of the intensity becomes a 2001-point log-log table over
    1e10-1e20 W/m^2 (the Julia shim's tabulation); for a power law that table is exact
    up to rounding, so the plasma density equals the PowerLawRate run's."""
    w0 = 2 * np.pi * C0 / 800e-9
    sigma, K = 1e-83, 6

    def run(rate):
        pl = Plasma(ionrate=rate, rho0=1.0, nuc=0.0, frequency=w0, Uiev=9.0, mr=1.0)
        grid = Grid(zmin=0, zmax=3e-8, Nz=4)
        src = SoftSource(geometry=(0.0,), waveform=lambda z, t: 0 * t, component="Ex")
        m = Model(grid, src, tmax=(300 + 0.5) * 1e-8 / C0, CN=1.0, pml=0,
                  material=Material(geometry=lambda z: z > -1, eps=1.0, plasma=pl))
        h = Handle(m.describe(integrals=0), oracle("f64"))
        E0 = 3e9
        h.set_field("Dx", np.full(4, EPS0 * E0))
        h.set_field("Ex", np.full(4, E0))
        h.step(300)
        return h.output(3)[0]

    exact = run(PowerLawRate(sigma=sigma, K=K))
    tab = run(lambda I: sigma * I ** K)
    assert exact > 1e-6
    assert abs(tab / exact - 1) < 1e-9, (tab, exact)
