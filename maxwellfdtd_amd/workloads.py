"""Synthetic workloads with fixed seeds, built through the package's Model API.

plane_wave_spheres_3d  3D, 256^3 cells, dx = 40 nm, CN = 0.99, 10-cell CPML on every
                       face, x-polarised +z TFSF plane wave (800 nm, 10 fs FWHM) on the
                       box of cells 16..241, 40 dielectric spheres (seed 2026): centres
                       uniform in the box, radius U[5, 20] cells, eps U[1.5, 4.0], later
                       spheres overwrite earlier ones (src/models.jl:91-98).
slab_1d                1D, Nz = 4000, dz = 40 nm, CN = 0.99, 10-cell CPML, eps = 2.25 on
                       cells 2000..2600, soft Gaussian Ex source at cell 500, point monitor
                       and 32-wavelength DFT (600-1000 nm) at cell 3500.
dispersive_3d          3D, 384 x 384 x 192 cells, dx = 5 nm, CN = 0.99, 10-cell CPML,
                       broadband TFSF pulse (700 nm carrier, 3 fs FWHM: 500-900 nm), a 6 x 6
                       array of metal spheres (seed 2027, radius U[15, 30] cells, centres
                       jittered on a regular lattice; one material) with a Drude pole (wp = 1.37e16,
                       gamma = 1e14) and two Lorentz poles, on a Debye substrate (deps = 2,
                       tau = 1e-14 s) filling the lowest fifth of the TFSF box; DFT plane
                       monitor of 384 x 384 points at 16 frequencies above the spheres.
batched_2d             512 uneven 2D (x, z) instances (seed 2028): Nx, Nz uniform in
                       {256..2048}, dx = 40 nm, CN = 0.99, 10-cell CPML, 2-6 dielectric slabs
                       and rods with eps U[1, 12], one Lorentz rod (eps_inf U[1.5, 4],
                       deps U[0.5, 3], resonance 400-1000 nm), a soft Gaussian Ex point
                       source, steps uniform in {2000..10000}.
served_3d              request stream (seed 2029): Poisson arrivals at 20 requests/s, cubic
                       grids of N uniform in {64..320} cells, dx = 40 nm, CN = 0.99, 10-cell
                       CPML, TFSF 800 nm 10 fs pulse of peak intensity log-uniform in
                       [1e16, 1e18] W/m^2 onto a sphere of Kerr (n0 = 1.45, n2 = 2.5e-20 m^2/W)
                       plus plasma medium (6-photon rate sigma6 I^6, rho0 = 2.1e28 m^-3,
                       Uiev = 9 eV, mr = 0.64, collision rate 1e15/s), steps uniform in
                       {1000..4000}.
Byte models per cell-update (algorithmic, compulsory traffic of the FP32 state):
  3D H half 37 B + 16 B per CPML-slab cell, E half 49 B + 16 B per CPML-slab cell;
  2D H half 17 B + 8 B, E half 29 B + 8 B; 1D H half 13 B + 8 B, E half 17 B + 8 B.
"""
import math

import numpy as np

from .constants import C0, EPS0
from .grids import Grid
from .materials import (DebyeSusceptibility, DrudeSusceptibility, LorentzSusceptibility,
                        Material, Plasma, PowerLawRate)
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


def dispersive_3d(nx=384, ny=384, nz=192, steps=5000, seed=2027, dx=5e-9, pml_cells=10,
                  box_margin=14, rmin=15, rmax=30, nfreq=16):
    """The 3D dispersive plasmonics workload.  Returns (model, monitors)."""
    grid = Grid(xmin=0, xmax=(nx - 1) * dx, Nx=nx, ymin=0, ymax=(ny - 1) * dx, Ny=ny,
                zmin=0, zmax=(nz - 1) * dx, Nz=nz)
    x, y, z = grid.x, grid.y, grid.z
    rng = np.random.default_rng(seed)
    xb = (x[box_margin], x[nx - 1 - box_margin])
    yb = (y[box_margin], y[ny - 1 - box_margin])
    zb = (z[box_margin], z[nz - 1 - box_margin])
    zsub = zb[0] + (zb[1] - zb[0]) / 5
    drude = DrudeSusceptibility(wp=1.37e16, gamma=1e14)
    lor1 = LorentzSusceptibility(deps=1.09, w0=4.1e15, delta=0.45e15)
    lor2 = LorentzSusceptibility(deps=0.54, w0=5.8e15, delta=0.9e15)
    sub = Material(geometry=lambda X, Y, Z: (X > xb[0]) & (X < xb[1]) & (Y > yb[0]) &
                   (Y < yb[1]) & (Z > zb[0]) & (Z <= zsub),
                   eps=2.0, chi=DebyeSusceptibility(deps=2.0, tau=1e-14))
    pitch_x, pitch_y = (xb[1] - xb[0]) / 6, (yb[1] - yb[0]) / 6
    spheres = []
    for i in range(6):
        for j in range(6):
            r = min(rng.uniform(rmin, rmax) * dx, 0.45 * min(pitch_x, pitch_y))
            spheres.append((xb[0] + (i + 0.5) * pitch_x + rng.uniform(-0.05, 0.05) * pitch_x,
                            yb[0] + (j + 0.5) * pitch_y + rng.uniform(-0.05, 0.05) * pitch_y,
                            zsub + r + dx, r))

    def in_spheres(X, Y, Z):
        inside = np.zeros(np.broadcast(X, Y, Z).shape, dtype=bool)
        for cx, cy, cz, r in spheres:
            inside |= (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r ** 2
        return inside

    metal = Material(geometry=in_spheres, eps=1.0, chi=[drude, lor1, lor2])
    mats = [sub, metal]
    w0 = 2 * np.pi * C0 / 700e-9
    src = TFSFSource(box=(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1]),
                     waveform=gaussian_pulse(fwhm=3e-15, t0=8e-15, w0=w0), component="Ex",
                     z0=z[pml_cells + 2])
    dt = 0.99 / C0 / math.sqrt(1 / dx ** 2 + 1 / dx ** 2 + 1 / dx ** 2)
    model = Model(grid, src, tmax=(steps - 1) * dt + 0.5 * dt, CN=0.99, pml=pml_cells * dx,
                  material=mats)
    izm = nz - 1 - pml_cells - 4
    lam = np.linspace(500e-9, 900e-9, nfreq)
    zm = z[izm]
    mon = SpectralMonitor(lambda X, Y, Z: np.abs(Z - zm) < dx / 2 + 0 * X + 0 * Y,
                          w=2 * np.pi * C0 / lam)
    return model, [mon]


def batched_2d(count=512, seed=2028, nmin=256, nmax=2048, smin=2000, smax=10000, dx=40e-9,
               pml_cells=10):
    """The uneven 2D batch: yields `count` models, one per instance, built lazily."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        Nx, Nz = (int(v) for v in rng.integers(nmin, nmax + 1, 2))
        steps = int(rng.integers(smin, smax + 1))
        nobj = int(rng.integers(2, 7))
        objs = []
        for _ in range(nobj):
            eps = float(rng.uniform(1.0, 12.0))
            if rng.uniform() < 0.5:       # slab: a band in z across the interior
                z0 = rng.uniform(0.3, 0.8) * Nz
                objs.append(("slab", eps, z0, rng.uniform(0.02, 0.1) * Nz))
            else:                         # rod: a disc in the (x, z) plane
                objs.append(("rod", eps, rng.uniform(0.3, 0.7) * Nx, rng.uniform(0.3, 0.8) * Nz,
                             rng.uniform(0.02, 0.08) * min(Nx, Nz)))
        # eps_inf >= 1.5 keeps eps_inf + chi(Nyquist) above 1: the explicit ADE update
        # (P^{n+1} from E^n, src/models.jl:1043-1060) lowers the permittivity seen at the
        # Nyquist frequency by up to ~0.2 for these poles, and below 1 the CN = 0.99 step
        # is unstable (the FP64 oracle blows up too)
        lor = (float(rng.uniform(1.5, 4.0)), float(rng.uniform(0.5, 3.0)),
               2 * np.pi * C0 / float(rng.uniform(400e-9, 1000e-9)), float(rng.uniform(1e13, 1e14)),
               rng.uniform(0.3, 0.7) * Nx, rng.uniform(0.3, 0.8) * Nz,
               rng.uniform(0.02, 0.08) * min(Nx, Nz))
        src = (rng.uniform(0.3, 0.7) * Nx, rng.uniform(0.15, 0.25) * Nz)
        specs.append((Nx, Nz, steps, objs, lor, src))
    for Nx, Nz, steps, objs, lor, src in specs:
        yield _instance_2d(Nx, Nz, steps, objs, lor, src, dx, pml_cells)


def _instance_2d(Nx, Nz, steps, objs, lor, src, dx, pml_cells):
    grid = Grid(xmin=0, xmax=(Nx - 1) * dx, Nx=Nx, zmin=0, zmax=(Nz - 1) * dx, Nz=Nz)
    mats = []
    for o in objs:
        if o[0] == "slab":
            _, eps, zc, hw = o
            mats.append(Material(geometry=lambda X, Z, zc=zc * dx, hw=hw * dx: np.abs(Z - zc) <= hw + 0 * X,
                                 eps=eps))
        else:
            _, eps, xc, zc, r = o
            mats.append(Material(geometry=lambda X, Z, c=(xc * dx, zc * dx), r=r * dx:
                                 (X - c[0]) ** 2 + (Z - c[1]) ** 2 <= r ** 2, eps=eps))
    eps_inf, deps, w0, delta, xc, zc, r = lor
    mats.append(Material(geometry=lambda X, Z, c=(xc * dx, zc * dx), r=r * dx:
                         (X - c[0]) ** 2 + (Z - c[1]) ** 2 <= r ** 2, eps=eps_inf,
                         chi=LorentzSusceptibility(deps=deps, w0=w0, delta=delta)))
    w = gaussian_pulse()
    sx, sz = src
    source = SoftSource(geometry=(grid.x[int(sx)], grid.z[int(sz)]),
                        waveform=lambda X, Z, t: w(t) + 0 * X, component="Ex")
    dt = 0.99 / C0 / math.sqrt(2 / dx ** 2)
    return Model(grid, source, tmax=(steps - 1) * dt + 0.5 * dt, CN=0.99, pml=pml_cells * dx,
                 material=mats)


SIGMA6 = 1e-95     # 6-photon cross-section, m^12 W^-6 s^-1: a rate of 1e13/s at 1e18 W/m^2


def served_3d(count=256, seed=2029, rate=20.0, nmin=64, nmax=320, smin=1000, smax=4000,
              dx=40e-9, pml_cells=10):
    """The served request stream: yields (arrival_time_s, model), built lazily."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(rng.integers(nmin, nmax + 1))
        steps = int(rng.integers(smin, smax + 1))
        I0 = 10.0 ** float(rng.uniform(16.0, 18.0))
        specs.append((t, n, steps, I0))
    for t, n, steps, I0 in specs:
        yield t, _request_3d(n, steps, I0, dx, pml_cells)


def _request_3d(n, steps, I0, dx, pml_cells):
    grid = Grid(xmin=0, xmax=(n - 1) * dx, Nx=n, ymin=0, ymax=(n - 1) * dx, Ny=n,
                zmin=0, zmax=(n - 1) * dx, Nz=n)
    x = grid.x
    n0, n2 = 1.45, 2.5e-20
    chi3 = 4 / 3 * n0 ** 2 * EPS0 * C0 * n2
    plasma = Plasma(ionrate=PowerLawRate(SIGMA6, 6.0), rho0=2.1e28, nuc=1e15, frequency=W0,
                    Uiev=9.0, mr=0.64)
    c = x[n // 2]
    r = 0.25 * (n - 1) * dx
    medium = Material(geometry=lambda X, Y, Z: (X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2 <= r ** 2,
                      eps=n0 ** 2, chi3=chi3, plasma=plasma)
    ksi = n0 * EPS0 * C0 / 2
    E0 = math.sqrt(I0 / ksi)
    w = gaussian_pulse()
    m = pml_cells + 3
    src = TFSFSource(box=(x[m], x[n - 1 - m], x[m], x[n - 1 - m], x[m], x[n - 1 - m]),
                     waveform=lambda t: E0 * w(t), component="Ex", z0=x[pml_cells + 1])
    dt = 0.99 / C0 / math.sqrt(3 / dx ** 2)
    return Model(grid, src, tmax=(steps - 1) * dt + 0.5 * dt, CN=0.99, pml=pml_cells * dx,
                 material=medium)


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


def _listed(chi):
    if chi is None:
        return []
    return list(chi) if isinstance(chi, (list, tuple)) else [chi]


def algorithmic_bytes(model):
    """Algorithmic HBM bytes of one step of `model` (FP32 state, uint8 id, perfect
    neighbour reuse): the base of the H and E halves and the CPML psi pairs
    (bytes_per_cell_update), plus per material cell the E^n read of a stateful cell and
    its pole state (Debye 8 B per component, Drude/Lorentz 16 B: P and P^(n-1) read and
    written) and the plasma state (rho and drho FP64 read and written 32 B, Pp 16 B and
    Pa 8 B per component, the JE accumulator 8 B).  Kerr adds no state.  Returns bytes
    per step for the whole grid."""
    N = int(np.prod(model.grid.dims3))
    bh, be = bytes_per_cell_update(model)
    total = (bh + be) * N
    nc = {1: 1, 2: 2, 3: 3}[model.grid.ndims]
    counts = np.bincount(np.asarray(model.geometry).ravel(), minlength=len(model.materials) + 1)
    for k, mat in enumerate(model.materials, start=1):
        chis = _listed(mat.chi)
        extra = 0.0
        if chis or mat.plasma is not None:
            extra += 4 * nc                                   # E^n, read once
        for c in chis:
            extra += (8 if type(c).__name__ == "DebyeSusceptibility" else 16) * nc
        if mat.plasma is not None:
            extra += 32 + (16 + 8) * nc + 8
        total += extra * counts[k]
    return float(total)
