"""Host-side model building: the parts of the reference that run on the CPU before the
backend is called (grids, time grid, boundary codes, CPML layers, material and source
coefficients, output times).  Expected values come from the reference's formulas and
messages (files cited per test) or from closed forms."""
import math
import os

import numpy as np
import pytest

from maxwellfdtd_amd.constants import C0, EPS0, MU0
from maxwellfdtd_amd.grids import Grid
from maxwellfdtd_amd.materials import (DebyeSusceptibility, DrudeSusceptibility,
                                       LorentzSusceptibility, Material, material_coefficients,
                                       susceptibility)
from maxwellfdtd_amd.models import Model, bc2int, output_times, time_grid, time_step
from maxwellfdtd_amd.pml import CPML, PML, left_layer, right_layer
from maxwellfdtd_amd.sources import HardSource, SoftSource, TFSFSource
from maxwellfdtd_amd.util import geometry2bool, geometry2indices, unravel
from maxwellfdtd_amd.workloads import bytes_per_cell_update, cpml_slab_cells


def src1d(**kw):
    return SoftSource(geometry=(0.0,), waveform=lambda z, t: 0 * t, component="Ex", **kw)


# ---- grids (src/grids.jl:1-96) --------------------------------------------------------

def test_grid_dispatch_and_steps():
    g1 = Grid(zmin=-1.0, zmax=1.0, Nz=21)
    assert g1.ndims == 1 and g1.shape == (21,) and math.isclose(g1.dz, 0.1)
    g2 = Grid(xmin=0, xmax=1, Nx=11, zmin=0, zmax=2, Nz=5)
    assert g2.ndims == 2 and g2.shape == (11, 5) and g2.dims3 == (11, 1, 5)
    g3 = Grid(xmin=0, xmax=1, Nx=3, ymin=0, ymax=1, Ny=4, zmin=0, zmax=1, Nz=5)
    assert g3.dims3 == (3, 4, 5) and math.isclose(g3.dy, 1 / 3)
    np.testing.assert_allclose(g1.z, np.linspace(-1, 1, 21))


def test_grid_wrong_arguments():
    with pytest.raises(ValueError, match="Wrong set of arguments"):
        Grid(ymin=0, ymax=1, Ny=3, zmin=0, zmax=1, Nz=5)


# ---- time step and time grid (src/models.jl:1181-1196, 57-122) -------------------------

def test_time_step_courant():
    d = 40e-9
    assert math.isclose(time_step(Grid(zmin=0, zmax=d * 9, Nz=10), 0.5), 0.5 * d / C0)
    g3 = Grid(xmin=0, xmax=d * 9, Nx=10, ymin=0, ymax=d * 9, Ny=10, zmin=0, zmax=d * 9, Nz=10)
    assert math.isclose(time_step(g3, 1.0), d / (C0 * math.sqrt(3)))


@pytest.mark.parametrize("tmax_steps,expected", [(10, 11), (10.4, 11), (10.6, 11), (0, 1)])
def test_time_grid_length(tmax_steps, expected):
    dt = 1e-17
    t = time_grid(dt, tmax_steps * dt)
    assert len(t) == expected
    assert t[0] == 0 and np.all(np.diff(t) > 0)
    assert t[-1] <= tmax_steps * dt * (1 + 1e-12)


# ---- boundary conditions (src/models.jl:1199-1216) ------------------------------------

def test_bc_codes():
    assert [bc2int(b) for b in ("periodic", ":dirichlet", "neumann")] == [1, 2, 3]
    with pytest.raises(ValueError, match="Wrong 'bc' value"):
        bc2int("absorbing")


def test_model_bc_layout():
    g = Grid(xmin=0, xmax=1e-6, Nx=11, zmin=0, zmax=1e-6, Nz=11)
    src = SoftSource(geometry=(0.5e-6, 0.5e-6), waveform=lambda x, z, t: 0 * t, component="Ex")
    m = Model(g, src, tmax=1e-15, bc=("periodic", "dirichlet", "neumann", "periodic"))
    assert m.bcs == [1, 2, 1, 1, 3, 1]
    with pytest.raises(ValueError):
        Model(g, src, tmax=1e-15, bc=("periodic", "dirichlet"))


# ---- CPML layers (src/pml.jl:1-304) ----------------------------------------------------

def test_cpml_layer_indices_and_profile():
    x = np.linspace(0, 99, 100) * 1e-8
    dt = 1e-17
    L = left_layer(x, 10e-8, dt)
    R = right_layer(x, CPML(thickness=10e-8, kmax=3), dt)
    assert (L.ib, L.Nb) == (11, 11)            # x[ib] is the interface, 1-based
    assert (R.ib, R.Nb) == (90, 11)
    # at the interface sigma = 0 and kappa = 1: B = exp(-alpha dt/EPS0), A = 0
    assert L.K[-1] == 1 and L.A[-1] == 0
    assert math.isclose(L.B[-1], math.exp(-10e-6 * dt / EPS0))
    # at the outer edge: sigma_max = -(m+1) ln(R0) / (2 eta0 L), kappa = kmax
    eta0 = math.sqrt(MU0 / EPS0)
    smax = -4 * math.log(10e-6) / (2 * eta0 * 10e-8)
    assert math.isclose(L.B[0], math.exp(-(smax + 10e-6) * dt / EPS0))
    assert math.isclose(R.K[-1], 3.0) and math.isclose(R.K[0], 1.0)
    assert np.all(np.diff(L.K) <= 0) and np.all(np.diff(R.B) <= 0)


def test_pml_faces_and_zero_thickness():
    g = Grid(xmin=0, xmax=1e-6, Nx=51, zmin=0, zmax=1e-6, Nz=51)
    layers = PML(g, (0.1e-6, 0, 0.2e-6, 0), 1e-17)
    present = [l.present for l in layers]
    assert present == [True, False, False, False, True, False]
    assert layers[4].Nb == 11 and layers[0].Nb == 6
    with pytest.raises(ValueError):
        PML(g, (0.1e-6, 0.1e-6), 1e-17)


# ---- materials (src/materials.jl:1-577) ------------------------------------------------

def test_negative_eps_rejected():
    with pytest.raises(ValueError, match="negative eps"):
        Material(geometry=lambda z: z > 0, eps=-2.0)


def test_material_coefficients_lossy():
    dt, eps, mu, sigma = 1e-17, 2.25, 1.5, 1e5
    c = material_coefficients(Material(geometry=lambda z: z > 0, eps=eps, mu=mu,
                                       sigma=sigma), dt)
    sd = sigma / (EPS0 * eps)
    assert math.isclose(c["Mh"], dt / (MU0 * mu))
    assert math.isclose(c["Me"], 1 / (EPS0 * eps))
    assert math.isclose(c["Md1"], (1 - sd * dt / 2) / (1 + sd * dt / 2))
    assert math.isclose(c["Md2"], dt / (1 + sd * dt / 2))
    assert c["nq"] == 0 and c["iskerr"] == 0 and c["isplasma"] == 0


def test_susceptibility_forms():
    w = 2 * np.pi * 3e14
    chi_d = DebyeSusceptibility(deps=2.0, tau=1e-15)
    chi_dr = DrudeSusceptibility(wp=1e16, gamma=1e14)
    chi_l = LorentzSusceptibility(deps=1.0, w0=5e15, delta=1e14)
    assert np.isclose(susceptibility(chi_d, w), 2.0 / (1 - 1j * w * 1e-15))
    assert np.isclose(susceptibility(chi_dr, w), -1e32 / (w ** 2 + 1j * w * 1e14))
    assert np.isclose(susceptibility(chi_l, w),
                      1.0 * 25e30 / (25e30 - w ** 2 - 2j * w * 1e14))


def test_kerr_coefficients():
    c = material_coefficients(Material(geometry=lambda z: z > 0, chi3=1e-22), 1e-17)
    assert c["iskerr"] == 1 and c["Mk2"] == 0 and math.isclose(c["Mk3"], EPS0 * 1e-22)


# ---- sources (src/sources.jl:9-244) ----------------------------------------------------

def test_soft_e_source_drives_d():
    g = Grid(zmin=0, zmax=1e-6, Nz=11)
    m = Model(g, src1d(), tmax=1e-15)
    s = m.source_structs[0]
    assert s["comp"] == 3 and s["coeff"] == EPS0 and s["kind"] == 0
    m = Model(g, src1d(explicit=True), tmax=1e-15)
    assert m.source_structs[0]["comp"] == 6 and m.source_structs[0]["coeff"] == 1.0


def test_hard_source_and_waveform_table():
    g = Grid(zmin=0, zmax=1e-6, Nz=11)
    src = HardSource(geometry=lambda z: z > 0.75e-6, waveform=lambda z, t, p: p * z + t,
                     p=2.0, component="Hy")
    m = Model(g, src, tmax=1e-15)
    s = m.source_structs[0]
    assert s["kind"] == 1 and s["comp"] == 1 and list(s["cells"]) == [8, 9, 10]
    np.testing.assert_allclose(s["wave"], 2.0 * g.z[[8, 9, 10]][None, :] + m.t[:, None])


def test_source_without_cells():
    g = Grid(zmin=0, zmax=1e-6, Nz=11)
    src = SoftSource(geometry=lambda z: z > 2e-6, waveform=lambda z, t: t, component="Ex")
    with pytest.raises(ValueError, match="did not find any grid points"):
        Model(g, src, tmax=1e-15)


def test_tfsf_file_missing(tmp_path):
    g = Grid(zmin=0, zmax=1e-6, Nz=11)
    with pytest.raises(FileNotFoundError):
        Model(g, TFSFSource(fname=str(tmp_path / "record.h5")), tmax=1e-15)


# ---- geometry (src/util.jl:1-100) ------------------------------------------------------

def test_geometry_indices_column_major():
    g = Grid(xmin=0, xmax=3, Nx=4, ymin=0, ymax=2, Ny=3, zmin=0, zmax=1, Nz=2)
    idx = geometry2indices(lambda x, y, z: (x == 3) & (y == 1) & (z == 1), g)
    assert list(idx) == [3 + 4 * (1 + 3 * 1)]
    ix, iy, iz = unravel(g, idx)
    assert (ix[0], iy[0], iz[0]) == (3, 1, 1)
    assert list(geometry2indices((2.2, 0.9, 0.1), g)) == [2 + 4 * (1 + 3 * 0)]
    mask = geometry2bool(lambda x, y, z: x + 0 * y + 0 * z > 1.5, g)
    assert mask.shape == (4, 3, 2) and mask.sum() == 2 * 3 * 2


# ---- output times (src/outputs.jl:491-518) ---------------------------------------------

def test_output_times():
    t = np.arange(101) * 1.0
    assert len(output_times(t, 10, None, None)) == 11
    assert len(output_times(t, None, 1000, None)) == 101
    assert np.array_equal(output_times(t, None, None, 25.0)[:3], [0.0, 25.0, 50.0])
    with pytest.raises(ValueError, match="should be specified"):
        output_times(t, None, None, None)
    with pytest.raises(ValueError, match="can be specified"):
        output_times(t, 10, 5, None)


# ---- byte model of the 3D kernels -----------------------------------------------------

def test_bytes_per_cell_update():
    from maxwellfdtd_amd.workloads import plane_wave_spheres_3d
    m = plane_wave_spheres_3d(n=64, steps=4)
    N = 64 ** 3
    s = cpml_slab_cells(m) / N
    bh, be = bytes_per_cell_update(m)
    assert math.isclose(bh, 37 + 16 * s) and math.isclose(be, 49 + 16 * s)
    assert math.isclose(s, 6 * 11 / 64)      # 11-point layers on six faces of 64^2


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with a message (no GPU needed)."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
