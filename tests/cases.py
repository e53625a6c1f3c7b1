"""Model builders shared by the CPU (oracle) and GPU (engine) tests.

The first group translates the reference's own test files (test/1d_test.jl,
test/2d_test.jl, test/3d_test.jl, test/*_test_losses.jl) into this package's API.
The second group are small seeded models that switch on every feature of the per-step
update (CPML on all faces, materials, ADE poles, Kerr, plasma, TFSF, monitors,
integrals, every boundary condition) for parity checks of the engine against the oracle.
"""
import math

import numpy as np

from maxwellfdtd_amd import (C0, EPS0, CPML, DebyeSusceptibility, DrudeSusceptibility,
                             FieldMonitor, Grid, HardSource, IonizationRate,
                             LorentzSusceptibility, Material, Model, Plasma, PowerLawRate,
                             SoftSource, SpectralMonitor, TFSFSource, susceptibility)


def gauss_waveform(lam0=2e-6, tau0=20e-15, dt0=None):
    w0 = 2 * np.pi * C0 / lam0
    dt0 = 3 * tau0 if dt0 is None else dt0
    return lambda t: np.exp(-(t - dt0) ** 2 / tau0 ** 2) * np.cos(w0 * (t - dt0))


# ---- translated reference tests ------------------------------------------------------
def ref_1d_test():
    """test/1d_test.jl:1-22 -> (model, Eth, rtol)."""
    grid = Grid(zmin=-5e-6, zmax=55e-6, Nz=2001)
    w = gauss_waveform()
    src = HardSource(geometry=lambda z: np.abs(z) < grid.dz / 2,
                     waveform=lambda z, t: w(t) + 0 * z, component="Ex")
    model = Model(grid, src, tmax=150e-15, pml=4e-6)
    Eth = w(model.t[-1] - grid.z / C0)
    return model, Eth, 1e-2


def ref_2d_test():
    """test/2d_test.jl:1-28."""
    grid = Grid(xmin=-55e-6, xmax=55e-6, Nx=11, zmin=-5e-6, zmax=55e-6, Nz=1101)
    w = gauss_waveform()
    a0 = 10e-6

    def waveform(x, z, t):
        return np.exp(-0.5 * (x / a0) ** 2) * w(t)

    src = HardSource(geometry=lambda x, z: np.abs(z) < grid.dz / 2 + 0 * x,
                     waveform=waveform, component="Ex")
    model = Model(grid, src, tmax=150e-15, pml=4e-6)
    X, Z = np.meshgrid(grid.x, grid.z, indexing="ij")
    Eth = np.exp(-0.5 * (X / a0) ** 2) * w(model.t[-1] - Z / C0)
    return model, Eth, 5e-2


def ref_3d_test():
    """test/3d_test.jl:1-30."""
    grid = Grid(xmin=-55e-6, xmax=55e-6, Nx=11, ymin=-55e-6, ymax=55e-6, Ny=11,
                zmin=-5e-6, zmax=55e-6, Nz=801)
    w = gauss_waveform()
    a0 = 10e-6

    def waveform(x, y, z, t):
        return np.exp(-0.5 * (np.sqrt(x ** 2 + y ** 2) / a0) ** 2) * w(t)

    src = HardSource(geometry=lambda x, y, z: np.abs(z) < grid.dz / 2 + 0 * x + 0 * y,
                     waveform=waveform, component="Ex")
    model = Model(grid, src, tmax=150e-15, pml=4e-6)
    X, Y, Z = np.meshgrid(grid.x, grid.y, grid.z, indexing="ij")
    Eth = np.exp(-0.5 * (np.sqrt(X ** 2 + Y ** 2) / a0) ** 2) * w(model.t[-1] - Z / C0)
    return model, Eth, 1e-1


def _loss_material(kind, geometry):
    lam0 = 2e-6
    w0 = 2 * np.pi * C0 / lam0
    if kind == "sigma":
        sigma, er = 300.0, 1.5 ** 2
        ei = sigma / (EPS0 * w0)
        ni = np.sqrt((np.sqrt(er ** 2 + ei ** 2) - er) / 2)
        ki = ni * w0 / C0
        return Material(geometry=geometry, eps=er, sigma=sigma), ki, 300e-15
    wq = 2 * np.pi * C0 / 1e-6
    chi = LorentzSusceptibility(deps=2.0, w0=wq, delta=0.01 * wq)
    n0 = np.sqrt(1 + susceptibility(chi, w0))
    ki = np.imag(n0) * w0 / C0
    return Material(geometry=geometry, chi=chi), ki, 600e-15


def ref_losses(ndims, kind):
    """test/{1d,2d,3d}_test_losses.jl: (model, ki, z) for the Beer-Lambert check on Sa."""
    w = gauss_waveform()
    if ndims == 1:
        grid = Grid(zmin=-15e-6, zmax=25e-6, Nz=2001)
        src = SoftSource(geometry=lambda z: np.abs(z - -10e-6) <= grid.dz / 2,
                         waveform=lambda z, t: w(t) + 0 * z, component="Ex")
        geom = lambda z: z >= 0
    elif ndims == 2:
        grid = Grid(xmin=-25e-6, xmax=25e-6, Nx=61, zmin=-15e-6, zmax=25e-6, Nz=601)
        src = SoftSource(geometry=lambda x, z: np.abs(z - -10e-6) <= grid.dz / 2 + 0 * x,
                         waveform=lambda x, z, t: w(t) + 0 * x, component="Ex")
        geom = lambda x, z: (z >= 0) + 0 * x > 0
    else:
        grid = Grid(xmin=-25e-6, xmax=25e-6, Nx=11, ymin=-25e-6, ymax=25e-6, Ny=11,
                    zmin=-15e-6, zmax=25e-6, Nz=601)
        src = SoftSource(geometry=lambda x, y, z: np.abs(z - -10e-6) <= grid.dz / 2 + 0 * x + 0 * y,
                         waveform=lambda x, y, z, t: w(t) + 0 * x, component="Ex")
        geom = lambda x, y, z: (z >= 0) + 0 * x + 0 * y > 0
    mat, ki, tmax = _loss_material(kind, geom)
    pml = 5e-6
    model = Model(grid, src, tmax=tmax, pml=pml, material=mat)
    return model, ki


def beer_lambert_ok(Sa, z, ki, rtol=1e-2):
    """test_intensity_losses_*D (test/1d_test_losses.jl:1-17): Sa normalised on
    z in [0, 20 um] against exp(-2 ki z)."""
    Sa = np.asarray(Sa, dtype=float)
    mask = (z < 0) | (z > 20e-6)
    shape = [1] * Sa.ndim
    shape[-1] = len(z)
    m = np.broadcast_to(mask.reshape(shape), Sa.shape)
    S = np.where(m, 0.0, Sa)
    S = S / S.max()
    I = np.where(m, 0.0, np.broadcast_to(np.exp(-2 * ki * z).reshape(shape), Sa.shape))
    den = max(np.linalg.norm(S), np.linalg.norm(I))
    return np.linalg.norm(S - I) <= rtol * den


# ---- seeded all-feature models ---------------------------------------------------------
def _cpml(cells, d, kmax):
    """CPML of `cells` cells (kappa up to kmax), or the plain thickness when kmax is None."""
    return cells * d if kmax is None else CPML(thickness=cells * d, kmax=kmax)


def spheres_3d(n=32, steps=120, seed=7, tfsf=True, soft=True, monitors=True, nspheres=5,
               bc="periodic", pml=4, nx=None, kmax=None):
    """n^3 grid (nx by n by n when nx is given), CPML `pml` cells on every face (kappa
    profile up to kmax when given), seeded dielectric spheres, a TFSF plane wave and a
    soft point source."""
    rng = np.random.default_rng(seed)
    d = 40e-9
    L = (n - 1) * d
    Nx = n if nx is None else nx
    Lx = (Nx - 1) * d
    grid = Grid(xmin=0, xmax=Lx, Nx=Nx, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    mats = []
    for _ in range(nspheres):
        c = rng.uniform(0.3 * L, 0.7 * L, 3)
        if nx is not None:
            c[0] = c[0] * (Lx / L)
        r = rng.uniform(2, 5) * d
        eps = rng.uniform(1.5, 4.0)
        mats.append(Material(geometry=lambda x, y, z, c=c, r=r:
                             (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r ** 2,
                             eps=eps))
    w = gauss_waveform(lam0=800e-9, tau0=5e-15, dt0=12e-15)
    srcs = []
    if tfsf:
        b0, b1 = (pml + 3) * d, L - (pml + 3) * d
        bx1 = Lx - (pml + 3) * d
        srcs.append(TFSFSource(box=(b0, bx1, b0, b1, b0, b1), waveform=w, component="Ex",
                               z0=(pml + 1) * d))
    if soft:
        srcs.append(SoftSource(geometry=(0.45 * Lx, 0.5 * L, 0.55 * L),
                               waveform=lambda x, y, z, t: 1e3 * w(t) + 0 * x, component="Ey"))
    dt = 0.99 / C0 / np.sqrt(3 / d ** 2)
    model = Model(grid, srcs, tmax=(steps - 1) * dt, CN=0.99, pml=_cpml(pml, d, kmax), material=mats, bc=bc)
    mons = None
    if monitors:
        w0 = 2 * np.pi * C0 / 800e-9
        mons = [FieldMonitor((0.5 * Lx, 0.5 * L, 0.7 * L)),
                FieldMonitor(lambda x, y, z: (np.abs(z - 0.7 * L) < d / 2) & (np.abs(x - 0.5 * Lx) < 2.5 * d)
                             & (np.abs(y - 0.5 * L) < 2.5 * d), sum=True),
                SpectralMonitor(lambda x, y, z: (np.abs(z - 0.75 * L) < d / 2) & (x > 0.3 * Lx) & (x < 0.7 * Lx)
                                & (y > 0.4 * L) & (y < 0.6 * L), w=w0 * np.linspace(0.8, 1.2, 5)),
                SpectralMonitor((0.5 * Lx, 0.5 * L, 0.3 * L), w=w0 * np.linspace(0.9, 1.1, 3), sum=True)]
    return model, mons


def dispersive_3d(n=28, steps=150, seed=11):
    """Drude + 2 Lorentz metal sphere over a Debye substrate, TFSF illumination."""
    rng = np.random.default_rng(seed)
    d = 5e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    metal_chi = [DrudeSusceptibility(wp=1.37e16, gamma=1e14),
                 LorentzSusceptibility(deps=1.09, w0=4.2e15, delta=0.9e15),
                 LorentzSusceptibility(deps=0.5, w0=6.0e15, delta=1.2e15)]
    c = np.array([0.5, 0.5, 0.55]) * L + rng.uniform(-1, 1, 3) * d
    metal = Material(geometry=lambda x, y, z: (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
                     <= (6 * d) ** 2, eps=1.0, chi=metal_chi)
    sub = Material(geometry=lambda x, y, z: (z < 0.3 * L) + 0 * x + 0 * y > 0, eps=2.0,
                   chi=DebyeSusceptibility(deps=2.0, tau=1e-14))
    w = gauss_waveform(lam0=700e-9, tau0=2e-15, dt0=6e-15)
    src = TFSFSource(box=(6 * d, L - 6 * d, 6 * d, L - 6 * d, 6 * d, L - 6 * d), waveform=w,
                     component="Ey", z0=5 * d)
    dt = 0.99 / C0 / np.sqrt(3 / d ** 2)
    model = Model(grid, src, tmax=(steps - 1) * dt, CN=0.99, pml=4 * d, material=[sub, metal])
    return model


def nonlinear_3d(n=24, steps=150, I0=1e17, rate="power"):
    """Kerr + plasma medium under an intense TFSF pulse (800 nm)."""
    d = 40e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    n0, n2 = 1.45, 2.5e-20
    chi3 = 4 / 3 * n0 ** 2 * EPS0 * C0 * n2
    w0 = 2 * np.pi * C0 / 800e-9
    if rate == "power":
        ionrate = PowerLawRate(sigma=1e-90, K=6)
    elif rate == "function":     # a plain function of I, tabulated log-log by the package
        ionrate = lambda I: 1e-90 * I ** 6
    else:
        x = np.linspace(13, 19, 61)
        ionrate = IonizationRate(x=x, y=6 * x - 90.0 + 0.01 * np.sin(x))
    plasma = Plasma(ionrate=ionrate, rho0=2.1e28, nuc=1e15, frequency=w0, Uiev=9.0, mr=0.64)
    med = Material(geometry=lambda x, y, z: (z > 0.35 * L) + 0 * x + 0 * y > 0, eps=n0 ** 2,
                   chi3=chi3, plasma=plasma)
    E0 = np.sqrt(I0 / (n0 * EPS0 * C0 / 2))
    tau0 = 4e-15
    w = lambda t: E0 * np.exp(-(t - 10e-15) ** 2 / tau0 ** 2) * np.cos(w0 * (t - 10e-15))
    src = TFSFSource(box=(6 * d, L - 6 * d, 6 * d, L - 6 * d, 6 * d, L - 6 * d), waveform=w,
                     component="Ex", z0=5 * d)
    dt = 0.99 / C0 / np.sqrt(3 / d ** 2)
    model = Model(grid, src, tmax=(steps - 1) * dt, CN=0.99, pml=4 * d, material=med)
    return model


def lorentz_2d(nx=96, nz=80, steps=200, seed=3, tfsf=True, kmax=None, bc="periodic", pml=5):
    """2D (x, z) Lorentz rods and a dielectric slab, TFSF plane wave + soft point source
    (faces `bc`, CPML `pml` cells thick)."""
    rng = np.random.default_rng(seed)
    d = 40e-9
    grid = Grid(xmin=0, xmax=(nx - 1) * d, Nx=nx, zmin=0, zmax=(nz - 1) * d, Nz=nz)
    Lx, Lz = (nx - 1) * d, (nz - 1) * d
    mats = [Material(geometry=lambda x, z: (z > 0.6 * Lz) & (z < 0.7 * Lz),
                     eps=float(rng.uniform(1, 12)))]
    for _ in range(3):
        c = rng.uniform(0.3, 0.7, 2) * np.array([Lx, Lz])
        mats.append(Material(geometry=lambda x, z, c=c: (x - c[0]) ** 2 + (z - c[1]) ** 2 <= (4 * d) ** 2,
                             eps=float(rng.uniform(1, 4)),
                             chi=LorentzSusceptibility(deps=float(rng.uniform(0.5, 3)),
                                                       w0=2 * np.pi * C0 / 600e-9, delta=1e14)))
    w = gauss_waveform(lam0=800e-9, tau0=5e-15, dt0=12e-15)
    srcs = [SoftSource(geometry=(0.3 * Lx, 0.4 * Lz), waveform=lambda x, z, t: 100 * w(t) + 0 * x,
                       component="Ez")]
    if tfsf:
        srcs.append(TFSFSource(box=(8 * d, Lx - 8 * d, 8 * d, Lz - 8 * d), waveform=w,
                               component="Ex", z0=6 * d))
    dt = 0.99 / C0 / np.sqrt(2 / d ** 2)
    model = Model(grid, srcs, tmax=(steps - 1) * dt, CN=0.99, pml=_cpml(pml, d, kmax), material=mats, bc=bc)
    w0 = 2 * np.pi * C0 / 800e-9
    mons = [FieldMonitor((0.5 * Lx, 0.8 * Lz)),
            SpectralMonitor(lambda x, z: (np.abs(z - 0.8 * Lz) < d / 2) & (x > 0.2 * Lx) & (x < 0.8 * Lx),
                            w=w0 * np.linspace(0.8, 1.2, 7))]
    return model, mons


def nonlinear_2d(nx=96, nz=80, steps=180, I0=1e17):
    """2D (x, z): a Kerr + plasma half-space, a Kerr-only rod and a plain slab under an
    intense TFSF pulse, plus a soft point source."""
    d = 40e-9
    grid = Grid(xmin=0, xmax=(nx - 1) * d, Nx=nx, zmin=0, zmax=(nz - 1) * d, Nz=nz)
    Lx, Lz = (nx - 1) * d, (nz - 1) * d
    n0, n2 = 1.45, 2.5e-20
    chi3 = 4 / 3 * n0 ** 2 * EPS0 * C0 * n2
    w0 = 2 * np.pi * C0 / 800e-9
    plasma = Plasma(ionrate=PowerLawRate(sigma=1e-90, K=6), rho0=2.1e28, nuc=1e15, frequency=w0,
                    Uiev=9.0, mr=0.64)
    mats = [Material(geometry=lambda x, z: (z > 0.45 * Lz) + 0 * x > 0, eps=n0 ** 2, chi3=chi3,
                     plasma=plasma),
            Material(geometry=lambda x, z: (x - 0.5 * Lx) ** 2 + (z - 0.3 * Lz) ** 2 <= (5 * d) ** 2,
                     eps=2.0, chi3=chi3),
            Material(geometry=lambda x, z: (z > 0.2 * Lz) & (z < 0.25 * Lz) & (x < 0.4 * Lx), eps=3.0)]
    E0 = np.sqrt(I0 / (n0 * EPS0 * C0 / 2))
    w = lambda t: E0 * np.exp(-(t - 10e-15) ** 2 / (4e-15) ** 2) * np.cos(w0 * (t - 10e-15))
    srcs = [TFSFSource(box=(8 * d, Lx - 8 * d, 8 * d, Lz - 8 * d), waveform=w, component="Ex",
                       z0=6 * d),
            SoftSource(geometry=(0.7 * Lx, 0.35 * Lz), waveform=lambda x, z, t: 1e-3 * w(t) + 0 * x,
                       component="Ez")]
    dt = 0.99 / C0 / np.sqrt(2 / d ** 2)
    return Model(grid, srcs, tmax=(steps - 1) * dt, CN=0.99, pml=5 * d, material=mats)


def nonlinear_1d(nz=500, steps=700, I0=1e17):
    """1D: a Kerr + plasma slab and a Debye slab under an intense soft-source pulse,
    CPML at both ends, point and spectral monitors."""
    d = 20e-9
    grid = Grid(zmin=0, zmax=(nz - 1) * d, Nz=nz)
    L = (nz - 1) * d
    n0, n2 = 1.45, 2.5e-20
    chi3 = 4 / 3 * n0 ** 2 * EPS0 * C0 * n2
    w0 = 2 * np.pi * C0 / 800e-9
    plasma = Plasma(ionrate=PowerLawRate(sigma=1e-90, K=6), rho0=2.1e28, nuc=1e15, frequency=w0,
                    Uiev=9.0, mr=0.64)
    mats = [Material(geometry=lambda z: (z > 0.4 * L) & (z < 0.6 * L), eps=n0 ** 2, chi3=chi3,
                     plasma=plasma),
            Material(geometry=lambda z: (z > 0.7 * L) & (z < 0.8 * L), eps=1.5,
                     chi=DebyeSusceptibility(deps=1.0, tau=2e-15))]
    E0 = np.sqrt(I0 / (n0 * EPS0 * C0 / 2))
    w = lambda t: E0 * np.exp(-(t - 12e-15) ** 2 / (4e-15) ** 2) * np.cos(w0 * (t - 12e-15))
    src = SoftSource(geometry=lambda z: np.abs(z - 0.2 * L) < d / 2, waveform=lambda z, t: w(t) + 0 * z,
                     component="Ex")
    dt = 0.99 / C0 * d
    model = Model(grid, src, tmax=(steps - 1) * dt, CN=0.99, pml=10 * d, material=mats)
    mons = [FieldMonitor((0.9 * L,)),
            SpectralMonitor((0.9 * L,), w=w0 * np.linspace(0.7, 1.3, 16))]
    return model, mons


def slab_1d(nz=600, steps=900, dispersive=True, kmax=None):
    """1D: CPML, a Lorentz + conductive slab and a soft source, plus monitors."""
    d = 40e-9
    grid = Grid(zmin=0, zmax=(nz - 1) * d, Nz=nz)
    L = (nz - 1) * d
    chi = LorentzSusceptibility(deps=1.5, w0=2 * np.pi * C0 / 600e-9, delta=3e13) if dispersive else None
    mats = [Material(geometry=lambda z: (z > 0.5 * L) & (z < 0.65 * L), eps=2.25, sigma=50.0, chi=chi)]
    w = gauss_waveform(lam0=800e-9, tau0=10e-15, dt0=30e-15)
    src = SoftSource(geometry=lambda z: np.abs(z - 0.125 * L) < d / 2, waveform=lambda z, t: w(t) + 0 * z,
                     component="Ex")
    dt = 0.99 / C0 * d
    model = Model(grid, src, tmax=(steps - 1) * dt, CN=0.99, pml=_cpml(10, d, kmax), material=mats)
    w0 = 2 * np.pi * C0 / 800e-9
    mons = [FieldMonitor((0.875 * L,)),
            SpectralMonitor((0.875 * L,), w=2 * np.pi * C0 / np.linspace(600e-9, 1000e-9, 32))]
    return model, mons


def tfsf_faces_3d(n=40, steps=150, pml=4, lo=9, hi=30, stateful=True):
    """A TFSF-only 3D model whose box faces cut through materials: a Lorentz sphere
    centred on the lower x face, a dielectric sphere on the upper x face, a Debye sphere
    on the x/z edge and a Kerr sphere on a y face, so that face cells hold stateless,
    stateful and nonlinear materials (stateful=False: the Lorentz and Debye spheres are
    plain dielectrics).  Box = cells lo..hi (0-based) on every axis."""
    d = 40e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    x = grid.x
    mid = x[(lo + hi) // 2]

    def ball(c, r):
        return lambda X, Y, Z: (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r ** 2

    w0 = 2 * np.pi * C0 / 800e-9
    chi_l = LorentzSusceptibility(deps=1.0, w0=1.2 * w0, delta=1e14) if stateful else None
    chi_d = DebyeSusceptibility(deps=1.5, tau=5e-15) if stateful else None
    mats = [
        Material(geometry=ball((x[lo], mid, mid), 4 * d), eps=1.8, chi=chi_l),
        Material(geometry=ball((x[hi], mid + 3 * d, mid), 4 * d), eps=3.0),
        Material(geometry=ball((x[lo], mid, x[lo]), 3.5 * d), eps=2.0, chi=chi_d),
        Material(geometry=ball((mid, x[lo], mid + 4 * d), 3.5 * d), eps=2.1, chi3=1e-19),
    ]
    w = gauss_waveform(lam0=800e-9, tau0=5e-15, dt0=12e-15)
    src = TFSFSource(box=(x[lo], x[hi], x[lo], x[hi], x[lo], x[hi]), waveform=lambda t: 1e8 * w(t),
                     component="Ex", z0=x[pml + 1])
    dt = 0.99 / C0 / np.sqrt(3 / d ** 2)
    model = Model(grid, src, tmax=(steps - 1) * dt, CN=0.99, pml=pml * d, material=mats)
    mons = [FieldMonitor((x[lo], mid, mid)), FieldMonitor((x[lo - 1], mid, mid)),
            SpectralMonitor((x[hi], mid, mid), w=np.array([w0, 1.2 * w0]))]
    return model, mons


# ---- TFSF replay tables ----------------------------------------------------------------
def tfsf_table_shapes(ndims, box, Nt):
    """Shapes of the replay arrays of TFSFSource1D/2D/3D for a 1-based box
    (src/sources.jl:269-282,387-399,565-591 and the record pass :779-877): per face, the
    tangential components, each [face extent..., Nt] with the extents of prepare_tfsf_record."""
    X1, X2, Y1, Y2, Z1, Z2 = box
    Nxi, Nyi, Nzi = X2 - X1, Y2 - Y1, Z2 - Z1
    if ndims == 1:
        return [(Nt,)] * 4
    if ndims == 2:
        return [(Nzi, Nt)] * 4 + [(Nxi, Nt)] * 4
    out = []
    for face in range(6):
        a, b = ((Nyi, Nzi), (Nxi, Nzi), (Nxi, Nyi))[face // 2]
        for j in range(4):
            big = j in (0, 3)
            out.append((a + 1, b, Nt) if big else (a, b + 1, Nt))
    return out


def synthetic_tfsf_tables(ndims, box, t, seed=5, E0=1.0):
    """Seeded replay tables of the right shapes: a Gaussian pulse in time whose amplitude
    and delay vary smoothly across each face, H tables scaled by 1/eta0.  They need not be
    a physical plane wave: the engine and the oracle replay any table the same way."""
    rng = np.random.default_rng(seed)
    eta0 = np.sqrt(4e-7 * np.pi / EPS0)
    Nt = len(t)
    dt = t[1] - t[0]
    tabs = []
    for k, shp in enumerate(tfsf_table_shapes(ndims, box, Nt)):
        sp = shp[:-1]
        amp = 1 + 0.3 * rng.standard_normal(sp) if sp else 1.0
        delay = 0.3 * Nt * dt + (0.05 * Nt * dt * rng.uniform(-1, 1, sp) if sp else 0.0)
        tt = np.asarray(t).reshape((1,) * len(sp) + (Nt,))
        amp = np.asarray(amp)[..., None] if sp else amp
        delay = np.asarray(delay)[..., None] if sp else delay
        width = 0.08 * Nt * dt
        v = E0 * amp * np.exp(-((tt - delay) / width) ** 2) * np.cos(2 * np.pi * (tt - delay) / (0.1 * Nt * dt))
        is_h = (ndims == 3 and k % 4 in (0, 1)) or (ndims < 3 and k % 2 == 0)
        tabs.append(v / eta0 if is_h else v)
    return tabs


def tfsf_tables_model(ndims, steps=120, pml=4, kmax=None):
    """A model driven by TFSF replay tables (the path the reference's TFSFSource takes):
    a dielectric scatterer inside the box, CPML around it."""
    d = 40e-9
    if ndims == 1:
        nz = 300
        grid = Grid(zmin=0, zmax=(nz - 1) * d, Nz=nz)
        L = (nz - 1) * d
        box = (0.3 * L, 0.7 * L)
        mat = Material(geometry=lambda z: (z > 0.45 * L) & (z < 0.55 * L), eps=2.25)
        dt = 0.99 / C0 * d
    elif ndims == 2:
        nx, nz = 72, 64
        grid = Grid(xmin=0, xmax=(nx - 1) * d, Nx=nx, zmin=0, zmax=(nz - 1) * d, Nz=nz)
        Lx, Lz = (nx - 1) * d, (nz - 1) * d
        box = (10 * d, Lx - 10 * d, 9 * d, Lz - 9 * d)
        mat = Material(geometry=lambda x, z: (x - 0.5 * Lx) ** 2 + (z - 0.5 * Lz) ** 2 <= (6 * d) ** 2, eps=3.0)
        dt = 0.99 / C0 / np.sqrt(2 / d ** 2)
    else:
        n = 36
        L = (n - 1) * d
        grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
        box = (8 * d, L - 8 * d, 9 * d, L - 9 * d, 10 * d, L - 10 * d)
        mat = Material(geometry=lambda x, y, z: (x - L / 2) ** 2 + (y - L / 2) ** 2 + (z - L / 2) ** 2
                       <= (5 * d) ** 2, eps=2.5)
        dt = 0.99 / C0 / np.sqrt(3 / d ** 2)
    tmax = (steps - 1) * dt
    # the box indices the model will use, from a plane-wave placeholder of the same box
    probe = Model(grid, TFSFSource(box=box, waveform=lambda t: 0 * t, component="Ex"), tmax=tmax,
                  CN=0.99, pml=_cpml(pml, d, kmax))
    tabs = synthetic_tfsf_tables(ndims, probe.source_structs[0]["box"], probe.t)
    model = Model(grid, TFSFSource(box=box, tables=tabs), tmax=tmax, CN=0.99,
                  pml=_cpml(pml, d, kmax), material=mat)
    return model


def plane_wave_pair(nd, n=16, steps=40, grid=None, CN=0.99, scatterer=True):
    """(record model, aux-mode model with a scatterer, box coords, grid, tmax) for a
    normal-incidence plane wave."""
    d = 40e-9
    L = (n - 1) * d
    w = gauss_waveform(lam0=800e-9, tau0=3e-15, dt0=8e-15)
    if grid is None:
        if nd == 1:
            grid = Grid(zmin=0, zmax=(4 * n - 1) * d, Nz=4 * n)
        elif nd == 2:
            grid = Grid(xmin=0, xmax=L, Nx=n, zmin=0, zmax=L, Nz=n)
        else:
            grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    zL = grid.z[-1]
    zb = (4 * d, zL - 4 * d)
    box = {1: zb, 2: (4 * d, L - 4 * d) + zb, 3: (4 * d, L - 4 * d, 4 * d, L - 4 * d) + zb}[nd]
    tmax = (steps - 0.5) * 0.99 * d / (C0 * math.sqrt(nd))
    plane = {1: lambda z: np.abs(z - 2 * d) < d / 2,
             2: lambda x, z: (np.abs(z - 2 * d) < d / 2) + 0 * x > 0,
             3: lambda x, y, z: (np.abs(z - 2 * d) < d / 2) + 0 * x * y > 0}[nd]
    wave = {1: lambda z, t: w(t) + 0 * z, 2: lambda x, z, t: w(t) + 0 * x,
            3: lambda x, y, z, t: w(t) + 0 * x}[nd]
    rec = Model(grid, SoftSource(geometry=plane, waveform=wave, component="Ex"), tmax=tmax, CN=CN,
                pml=2 * d)
    c = zL / 2
    geo = {1: lambda z: np.abs(z - c) < 3 * d,
           2: lambda x, z: (x - L / 2) ** 2 + (z - c) ** 2 <= (3 * d) ** 2,
           3: lambda x, y, z: (x - L / 2) ** 2 + (y - L / 2) ** 2 + (z - c) ** 2 <= (3 * d) ** 2}[nd]
    mat = Material(geometry=geo, eps=3.0) if scatterer else None
    aux = Model(grid, TFSFSource(box=box, waveform=w, component="Ex", z0=2 * d), tmax=tmax, CN=CN,
                pml=2 * d, material=mat)
    return rec, aux, box, grid, tmax, mat
