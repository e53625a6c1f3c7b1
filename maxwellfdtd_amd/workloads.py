"""Synthetic workloads with fixed seeds, built through the package's Model API.

plane_wave_spheres_3d  3D, 256^3 cells, dx = 40 nm, CN = 0.99, 10-cell CPML on every
                       face, x-polarised +z TFSF plane wave (800 nm, 10 fs FWHM) on the
                       box of cells 16..241, 40 dielectric spheres (seed 2026): centres
                       uniform in the box, radius U[5, 20] cells, eps U[1.5, 4.0], later
                       spheres overwrite earlier ones (src/models.jl:91-98).
slab_1d                1D, Nz = 4000, dz = 40 nm, CN = 0.99, 10-cell CPML, eps = 2.25 on
                       cells 2000..2600, soft Gaussian Ex source at cell 500, point monitor
                       and 32-wavelength DFT (600-1000 nm) at cell 3500.
Byte models per cell-update (algorithmic, compulsory traffic of the FP32 state):
  3D H half 37 B + 16 B per CPML-slab cell, E half 49 B + 16 B per CPML-slab cell;
  2D H half 17 B + 8 B, E half 29 B + 8 B; 1D H half 13 B + 8 B, E half 17 B + 8 B.
"""
import math

import numpy as np

from .constants import C0
from .grids import Grid
from .materials import Material
from .models import Model
from .monitors import FieldMonitor, SpectralMonitor
from .sources import SoftSource, TFSFSource

LAM0 = 800e-9
W0 = 2 * np.pi * C0 / LAM0


def gaussian_pulse(fwhm=10e-15, t0=30e-15, w0=W0):
    """exp(-2 ln2 ((t - t0)/fwhm)^2) cos(w0 (t - t0))."""
    a = 2 * math.log(2) / fwhm ** 2
    return lambda t: np.exp(-a * (t - t0) ** 2) * np.cos(w0 * (t - t0))


def plane_wave_spheres_3d(n=256, steps=2000, nspheres=40, seed=2026, dx=40e-9, pml_cells=10,
                          box_margin=15, rmin=5, rmax=20):
    grid = Grid(xmin=0, xmax=(n - 1) * dx, Nx=n, ymin=0, ymax=(n - 1) * dx, Ny=n,
                zmin=0, zmax=(n - 1) * dx, Nz=n)
    x = grid.x
    lo, hi = x[box_margin], x[n - 1 - box_margin]
    rng = np.random.default_rng(seed)
    mats = []
    for _ in range(nspheres):
        c = rng.uniform(lo, hi, 3)
        r = rng.uniform(rmin, rmax) * dx
        eps = rng.uniform(1.5, 4.0)
        mats.append(Material(
            geometry=lambda X, Y, Z, c=c, r=r: (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r ** 2,
            eps=float(eps)))
    src = TFSFSource(box=(lo, hi, lo, hi, lo, hi), waveform=gaussian_pulse(), component="Ex",
                     z0=x[pml_cells + 2])
    dt = 0.99 / C0 / math.sqrt(3 / dx ** 2)
    return Model(grid, src, tmax=(steps - 1) * dt + 0.5 * dt, CN=0.99, pml=pml_cells * dx,
                 material=mats)


def slab_1d(nz=4000, steps=8000, dz=40e-9):
    grid = Grid(zmin=0, zmax=(nz - 1) * dz, Nz=nz)
    z = grid.z
    slab = Material(geometry=lambda Z: (Z >= z[1999] - dz / 4) & (Z <= z[2599] + dz / 4), eps=2.25)
    w = gaussian_pulse()
    src = SoftSource(geometry=(z[499],), waveform=lambda Z, t: w(t) + 0 * Z, component="Ex")
    dt = 0.99 / C0 * dz
    model = Model(grid, src, tmax=(steps - 1) * dt + 0.5 * dt, CN=0.99, pml=10 * dz, material=slab)
    lam = np.linspace(600e-9, 1000e-9, 32)
    mons = [FieldMonitor((z[3499],)), SpectralMonitor((z[3499],), w=2 * np.pi * C0 / lam)]
    return model, mons


def cpml_slab_cells(model):
    """Cells that lie in at least one CPML slab, summed over faces (a cell in two slabs
    counts twice, as it does two psi updates per half-step)."""
    g = model.grid
    dims = g.dims3
    total = 0
    for f, L in enumerate(model.pml):
        if not L.present:
            continue
        other = int(np.prod(dims)) // dims[f // 2]
        total += L.Nb * other
    return total


def bytes_per_cell_update(model):
    """(H-half, E-half) algorithmic bytes per cell of one step, FP32 state, uint8 id."""
    N = int(np.prod(model.grid.dims3))
    slab = cpml_slab_cells(model) / N
    if model.grid.ndims == 3:
        return 37 + 16 * slab, 49 + 16 * slab
    if model.grid.ndims == 2:
        return 17 + 8 * slab, 29 + 8 * slab
    return 13 + 8 * slab, 17 + 8 * slab
