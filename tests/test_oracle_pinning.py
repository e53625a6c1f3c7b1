"""Checks fixed by mathematics for the oracle functions the reference's own tests leave
open: the plasma current and the multiphoton loss of the plasma update
(src/models.jl:1080-1133 and the 1D/2D twins), the Dirichlet and Neumann faces
(src/models.jl:755-802,920-973 and the 1D/2D twins), the CPML update with kappa > 1
(src/models.jl:812-871, src/pml.jl:68-81) and the log-log ionization-rate table
(examples/5_nonlinear_propagation/IonizationRate.jl:30-52, with the reference's own data
file SiO2_keldysh_1.03um.tf as the fixture).

Each expected value comes from an independent route -- the Drude model's frequency
response, energy conservation, a mirror-image extension of the grid, a quadrature of
the CPML memory kernel, the data file's own rows -- so that a plausible transcription
slip (a dropped factor, a wrong sign, a one-sided difference, a shifted profile index,
a division by K in the wrong place) fails."""
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from maxwellfdtd_amd import (C0, EPS0, MU0, CPML, FieldMonitor, Grid, IonizationRate,
                             Material, Model, Plasma, PowerLawRate, SoftSource)
from maxwellfdtd_amd.constants import HBAR, ME, QE
from maxwellfdtd_amd.materials import material_coefficients
from maxwellfdtd_amd.models import Handle
from oracle_util import oracle, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
RATE_FILE = os.path.join(HERE, "golden", "SiO2_keldysh_1.03um.tf")


def _uniform_plasma_1d(plasma, dz, tmax, source=None, nz=4, CN=1.0):
    """A periodic 1D grid filled with one plasma medium (eps = 1): every cell is the same
    0D system, since uniform fields have no curl."""
    grid = Grid(zmin=0, zmax=(nz - 1) * dz, Nz=nz)
    if source is None:
        source = SoftSource(geometry=(0.0,), waveform=lambda z, t: 0 * t, component="Ex")
    mat = Material(geometry=lambda z: z > -1, eps=1.0, plasma=plasma)
    return Model(grid, source, tmax=tmax, CN=CN, pml=0, material=mat)


# ---- plasma current: the Drude response ------------------------------------------------
def test_plasma_current_is_the_drude_current():
    """With the medium fully ionised (rho = 1 from the first strong field on), the plasma
    polarisation Pp obeys Pp'' + nuc Pp' = QE^2/(ME mr) rho0 rho E: the free-electron
    (Drude) current J = Pp' answers E with J(w) = QE^2/(ME mr) rho0 E(w) / (nuc + i w)
    (examples/5_nonlinear_propagation/test_plasma.jl:70-93 makes the same comparison).
    The delivered energy JE = sum J E dt (src/models.jl:1126-1129) of a pulse, driven
    through a uniform soft source and recorded by a field monitor, must equal that
    response's sum J E dt over the recorded field.  rho0 = 1e3 keeps Pp from acting
    back on E (a plasma frequency of 2e9 rad/s).  A one-sided J, a sign slip in Ap/Bp/Cp or a missing rho0 moves JE by > 1e-2;
    the recursion itself is second order (8e-6 here, w0 dt = 8e-3)."""
    dz = 1e-9
    dt = dz / C0
    w0 = 2 * np.pi * C0 / 800e-9
    tau, t0, E0 = 4e-15, 20e-15, 1e11
    g = lambda t: E0 * np.exp(-((t - t0) / tau) ** 2) * np.cos(w0 * (t - t0))
    # D_n = eps0 (g(t_{n+1}) - g(0)), so the E that step n+1 computes is g(t_{n+1})
    src = SoftSource(geometry=lambda z: z > -1, waveform=lambda z, t: g(t + dt) - g(t) + 0 * z,
                     component="Ex")
    nuc, mr = 1e15, 1.0
    # a huge one-photon rate saturates rho at 1 while the field is still 1e-5 of its peak
    rho0 = 1e3
    pl = Plasma(ionrate=PowerLawRate(sigma=1e10, K=1), rho0=rho0, nuc=nuc, frequency=w0,
                Uiev=9.0, mr=mr)
    model = _uniform_plasma_1d(pl, dz, 60e-15, source=src)
    desc = model.describe(monitors=[FieldMonitor((0.0,))], integrals=0)
    h = Handle(desc, oracle("f64"))
    h.step(model.Nt)
    E = h.monitor(0, "Ex")[0]            # E after each step
    assert abs(h.output(3)[0] / rho0 - 1.0) < 1e-9 and np.max(np.abs(E)) > 0.9 * E0
    N = len(E)
    w = 2 * np.pi * np.fft.fftfreq(N, model.dt)
    c = QE ** 2 / (ME * mr) * rho0
    J = np.real(np.fft.ifft(np.fft.fft(E) * c / (nuc + 1j * w)))
    # step n accumulates J E dt at the E it starts from: the samples before the last
    JE_drude = float(np.sum(J[:-1] * E[:-1]) * model.dt)
    JE = h.output(2)[0]
    assert JE > 0
    assert abs(JE / JE_drude - 1) < 1e-4, (JE, JE_drude)


# ---- multiphoton ionisation: energy balance -------------------------------------------
def test_multiphoton_loss_takes_K_photons_per_electron():
    """The loss polarisation Pa += Ma drho rho0 E / |E|^2 with Ma = K hbar w dt
    (src/materials.jl:181-187, src/models.jl:1099-1107) makes the field pay K photon
    energies for every electron freed: in a uniform cell with constant D, eps0 E
    dE/dt = -Ja E = -K hbar w rho0 drho/dt, so the field energy lost, 1/2 eps0 (E0^2 -
    E^2), equals K hbar w rho0 rho at the end.  The plasma current is switched off
    (mr = 1e6 makes Cp and the avalanche rate negligible).  The field loses 10 % of its
    energy; the balance holds to the first-order error of the rate integration (3e-4).
    A missing K, invE2, rho0 or a flipped sign breaks it by far more."""
    dz = 3e-9
    dt = dz / C0
    w0 = 2 * np.pi * C0 / 800e-9
    Uiev = 9.0
    pl = Plasma(ionrate=PowerLawRate(sigma=5e-92, K=6), rho0=2e28, nuc=1e15, frequency=w0,
                Uiev=Uiev, mr=1e6)
    model = _uniform_plasma_1d(pl, dz, (2000 - 0.5) * dt)
    E0 = 1e10
    h = Handle(model.describe(integrals=0), oracle("f64"))
    h.set_field("Dx", np.full(4, EPS0 * E0))
    h.set_field("Ex", np.full(4, E0))
    h.step(model.Nt)
    E = h.field("Ex")[0]
    rho = h.output(3)[0]                 # rho * rho0, electrons per m^3
    K = math.ceil(Uiev * QE / (HBAR * w0))
    assert K == 6
    lost = 0.5 * EPS0 * (E0 ** 2 - E ** 2)
    paid = K * HBAR * w0 * rho
    assert 0.05 < lost / (0.5 * EPS0 * E0 ** 2) < 0.5
    assert abs(lost / paid - 1) < 2e-3, (lost, paid)


# ---- Dirichlet and Neumann faces: mirror images ---------------------------------------
# A field on the bounded grid with Dirichlet / Neumann faces equals, cell for cell, the
# field on a periodic grid holding the mirror-image extension.  Along an axis a, the
# components on half-integer positions are E_a and the two H components across a (E_a
# at i+1/2, H_b at i+1/2 for b != a); the others sit on integer positions.  A face sets
# the field beyond it (src/models.jl:755-802,920-973):
#   left Dirichlet:  H[0] = 0        mirror plane at 1/2,   tangential E even
#   left Neumann:    H[0] = H[1]     mirror plane at 1,     tangential E odd
#   right Dirichlet: E[N+1] = 0      mirror plane at N+1,   tangential E odd
#   right Neumann:   E[N+1] = E[N]   mirror plane at N+1/2, tangential E even
# Under a mirror, E is a polar vector and H an axial one: with tangential E even, normal
# E is odd, normal H even, tangential H odd; with tangential E odd, every sign flips.
def _axis_fold(N, left, right, half, e_or_h, tangential):
    """For a periodic extension of one axis: (source index 0-based or -1, sign) of every
    extended index 1..P, and P."""
    p1, even_l = (0.5, True) if left == "dirichlet" else (1.0, False)
    p2, even_r = (N + 1.0, False) if right == "dirichlet" else (N + 0.5, True)
    base = 1 if (e_or_h == "E") == tangential else -1      # sign when tangential E is even
    s1 = base if even_l else -base
    s2 = base if even_r else -base
    L = p2 - p1
    P = int(round(2 * L)) if even_l == even_r else int(round(4 * L))
    src = np.empty(P, dtype=int)
    sgn = np.empty(P)
    for k in range(1, P + 1):
        q = k + 0.5 * half
        n, r = divmod(q - p1, 2 * L)
        s = (s1 * s2) ** int(n)
        if r <= L:
            qq = p1 + r
        else:
            qq = 2 * p2 - (p1 + r)
            s *= s2
        kk = qq - 0.5 * half
        on_odd_plane = (abs(qq - p1) < 1e-9 and s1 < 0) or (abs(qq - p2) < 1e-9 and s2 < 0)
        if on_odd_plane or not (1 <= round(kk) <= N):
            src[k - 1], sgn[k - 1] = -1, 0.0
        else:
            src[k - 1], sgn[k - 1] = int(round(kk)) - 1, s
    return src, sgn, P


COMPS = {1: [("Ex", "E", 0), ("Hy", "H", 1)],
         2: [("Ex", "E", 0), ("Ez", "E", 2), ("Hy", "H", 1)],
         3: [("Ex", "E", 0), ("Ey", "E", 1), ("Ez", "E", 2),
             ("Hx", "H", 0), ("Hy", "H", 1), ("Hz", "H", 2)]}


def _mirror_pair(ndims, bcs_axes, n=(7, 6, 8), steps=40, seed=0):
    """Runs the bounded model and its periodic mirror extension on the FP64 oracle;
    returns the relative L2 differences of every component over the bounded grid."""
    rng = np.random.default_rng(seed)
    d = 30e-9
    axes = {1: [2], 2: [0, 2], 3: [0, 1, 2]}[ndims]
    dims = [1, 1, 1]
    for a in axes:
        dims[a] = n[a]
    bc = {a: bcs_axes[k] for k, a in enumerate(axes)}
    folds = {}
    ext = list(dims)
    for a in axes:
        for name, kind, dirn in COMPS[ndims]:
            half = (kind == "E") == (dirn == a)
            folds[(name, a)] = _axis_fold(dims[a], bc[a][0], bc[a][1], half, kind, dirn != a)
            ext[a] = folds[(name, a)][2]

    def grid_of(sz):
        kw = {}
        for a, ax in zip(axes, [("x", "Nx"), ("y", "Ny"), ("z", "Nz")] if ndims == 3 else
                         ([("x", "Nx"), ("z", "Nz")] if ndims == 2 else [("z", "Nz")])):
            kw[ax[0] + "min"] = 0.0
            kw[ax[0] + "max"] = (sz[a] - 1) * d
            kw[ax[1]] = sz[a]
        return Grid(**kw)

    def model_of(sz, bcs):
        g = grid_of(sz)
        at = tuple(0.0 for _ in axes)
        src = SoftSource(geometry=at, waveform=lambda *a: 0 * a[-1], component="Ex")
        dt = 0.9 / C0 / math.sqrt(len(axes) / d ** 2)
        return Model(g, src, tmax=(steps - 0.5) * dt, CN=0.9, bc=bcs, pml=0)

    bcs = []
    for a in axes:
        bcs += list(bc[a])
    bounded = model_of(dims, tuple(bcs))
    periodic = model_of(ext, "periodic")
    assert bounded.Nt == periodic.Nt == steps
    hb = Handle(bounded.describe(integrals=0), oracle("f64"))
    hp = Handle(periodic.describe(integrals=0), oracle("f64"))
    for name, kind, dirn in COMPS[ndims]:
        raw = rng.standard_normal(dims) * (1.0 if kind == "E" else 1 / 377.0)
        big = raw
        sgn = np.ones(1)
        idx = []
        signs = []
        for a in range(3):
            if a in axes:
                s_, g_, _ = folds[(name, a)]
                idx.append(s_)
                signs.append(g_)
            else:
                idx.append(np.zeros(1, dtype=int))
                signs.append(np.ones(1))
        safe = [np.where(i_ < 0, 0, i_) for i_ in idx]
        val = raw[np.ix_(*safe)] * signs[0][:, None, None] * signs[1][None, :, None] * signs[2][None, None, :]
        inner = val[:dims[0], :dims[1], :dims[2]]
        hb.set_field(name, inner)
        hp.set_field(name, val)
        if kind == "E":
            hb.set_field("D" + name[1], EPS0 * inner)
            hp.set_field("D" + name[1], EPS0 * val)
    hb.step(steps)
    hp.step(steps)
    errs = {}
    for name, kind, dirn in COMPS[ndims]:
        a_ = hb.field(name).reshape(dims, order="F")
        b_ = hp.field(name).reshape(ext, order="F")[:dims[0], :dims[1], :dims[2]]
        errs[name] = rel_l2(a_, b_)
        assert np.max(np.abs(a_)) > 0
    return errs


BC_PAIRS = [("dirichlet", "dirichlet"), ("neumann", "neumann"), ("dirichlet", "neumann"),
            ("neumann", "dirichlet")]


@pytest.mark.parametrize("pair", BC_PAIRS)
def test_mirror_faces_1d(pair):
    errs = _mirror_pair(1, [pair])
    assert max(errs.values()) < 1e-13, errs


@pytest.mark.parametrize("pairs", [(BC_PAIRS[0], BC_PAIRS[1]), (BC_PAIRS[2], BC_PAIRS[3]),
                                   (BC_PAIRS[1], BC_PAIRS[2])])
def test_mirror_faces_2d(pairs):
    errs = _mirror_pair(2, list(pairs))
    assert max(errs.values()) < 1e-13, errs


@pytest.mark.parametrize("pairs", [(BC_PAIRS[0], BC_PAIRS[1], BC_PAIRS[2]),
                                   (BC_PAIRS[3], BC_PAIRS[0], BC_PAIRS[1]),
                                   (BC_PAIRS[2], BC_PAIRS[3], BC_PAIRS[0])])
def test_mirror_faces_3d(pairs):
    errs = _mirror_pair(3, list(pairs))
    assert max(errs.values()) < 1e-13, errs


# ---- CPML with kappa > 1: the memory kernel -------------------------------------------
def _cpml_cell(z, i, side, L, kmax, m=3, R0=10e-6, alpha=10e-6):
    """kappa, sigma, alpha of the CPML at node z[i] from the layer's definition
    (src/pml.jl:68-81): d = |z - zb| / L, kappa = 1 + (kmax - 1) d^m,
    sigma = sigma_max d^m, sigma_max = -(m + 1) ln R0 / (2 eta0 L)."""
    eta0 = math.sqrt(MU0 / EPS0)
    smax = -(m + 1) * math.log(R0) / (2 * eta0 * L)
    zb = z[0] + L if side == "left" else z[-1] - L
    dd = abs(z[i] - zb) / L
    return 1 + (kmax - 1) * dd ** m, smax * dd ** m, alpha


@pytest.mark.parametrize("kmax", [1.0, 3.0, 5.0])
def test_cpml_update_is_the_stretched_derivative(kmax):
    """Inside a CPML layer a derivative u becomes u/kappa + (zeta * u): the stretched
    coordinate 1/s(w), s = kappa + sigma/(alpha + i w eps0), whose memory term has the
    kernel zeta(t) = -sigma/(eps0 kappa^2) exp(-(sigma/kappa + alpha) t / eps0).  With u
    held over each time step (the recursive convolution of the CPML), step n sees
    u^(n-1)/kappa + sum_k h_k u^(n-k), h_k = integral of zeta over [(k-1) dt, k dt]: here
    h_k comes from numerical quadrature of zeta, not from the A, B closed forms.  Checked
    on both updates (H from dEx/dz, E from dHy/dz) in both layers, from monitors around
    one layer cell of a free-running 1D run.  Driving the memory term by u/kappa, adding
    it divided by kappa, or reading the right layer's profile at the wrong index all
    break it by > 1e-2 at kmax = 5."""
    nz, cells, dz = 80, 10, 20e-9
    grid = Grid(zmin=0, zmax=(nz - 1) * dz, Nz=nz)
    z = grid.z
    w0 = 2 * np.pi * C0 / 800e-9
    g = lambda t: np.exp(-((t - 15e-15) / 3e-15) ** 2) * np.cos(w0 * (t - 15e-15))
    L = cells * dz
    src = SoftSource(geometry=(z[nz // 2],), waveform=lambda z_, t: g(t), component="Ex")
    model = Model(grid, src, tmax=40e-15 + nz * dz / C0, CN=0.5, pml=CPML(thickness=L, kmax=kmax))
    dt = model.dt
    N = model.Nt
    for side, i in (("left", 3), ("right", nz - 4)):
        kap, sig, alpha = _cpml_cell(z, i, side, L, kmax)
        a = (sig / kap + alpha) / EPS0
        zeta = lambda t: -sig / (EPS0 * kap ** 2) * math.exp(-a * t)
        K = int(math.ceil(40.0 / (a * dt))) + 1
        hk = np.array([0.0] + [quad(zeta, (k - 1) * dt, k * dt, epsabs=0, epsrel=1e-13)[0]
                               for k in range(1, K)])
        desc = model.describe(monitors=[FieldMonitor((z[j],)) for j in (i - 1, i, i + 1)],
                              integrals=0)
        h = Handle(desc, oracle("f64"))
        h.step(N)
        Hm, H0 = h.monitor(0, "Hy")[0], h.monitor(1, "Hy")[0]
        E0, Ep = h.monitor(1, "Ex")[0], h.monitor(2, "Ex")[0]

        # H update at step n: u^(n-1) = (E_{i+1} - E_i)/dz after step n-1 (0 at n = 1)
        u = np.concatenate([[0.0], (Ep - E0)[:-1] / dz])
        predH = -(dt / MU0) * (u / kap + _memory(u, hk, N))
        dH = np.diff(np.concatenate([[0.0], H0]))
        assert rel_l2(dH, predH) < 1e-8, (side, "H", rel_l2(dH, predH))
        # E update at step n: v^n = (H_i - H_{i-1})/dz after step n's H update
        v = (H0 - Hm) / dz
        predE = -(dt / EPS0) * (v / kap + _memory(v, hk, N))
        dE = np.diff(np.concatenate([[0.0], E0]))
        assert rel_l2(dE, predE) < 1e-8, (side, "E", rel_l2(dE, predE))


def _memory(u, hk, N):
    """sum_{k>=1} h_k u[j+1-k] for output j = n-1: the memory term of step n, whose
    newest sample u[n-1] (the derivative that step divides by kappa) gets h_1."""
    out = np.zeros(N)
    for k in range(1, len(hk)):
        out[k - 1:] += hk[k] * u[:N - k + 1]
    return out


@pytest.mark.parametrize("kmax", [3.0, 5.0])
def test_cpml_with_kappa_absorbs(kmax):
    """A pulse leaving through a CPML with a kappa profile reflects little: the probe
    time series against the same run in a domain long enough that nothing returns."""
    d = 40e-9
    w0 = 2 * np.pi * C0 / 800e-9
    g = lambda t: np.exp(-((t - 30e-15) / 10e-15) ** 2) * np.cos(w0 * (t - 30e-15))

    def run(n, cells):
        grid = Grid(zmin=0, zmax=(n - 1) * d, Nz=n)
        m = Model(grid, SoftSource(geometry=lambda z: np.abs(z - 150 * d) < d / 2,
                                   waveform=lambda z, t: g(t) + 0 * z, component="Ex"),
                  tmax=3000 * 0.5 * d / C0, pml=CPML(thickness=cells * d, kmax=kmax))
        h = Handle(m.describe(monitors=[FieldMonitor((300 * d,))], integrals=0), oracle("f64"))
        h.step(m.Nt)
        return h.monitor(0, "Ex")[0]

    refl = {}
    for cells in (10, 20):
        big = run(4000, cells)
        refl[cells] = np.max(np.abs(run(400, cells) - big)) / np.max(np.abs(big))
    assert refl[10] < 0.1
    assert refl[20] < refl[10] / 3


# ---- the log-log ionization-rate table -------------------------------------------------
def _rate_rows():
    data = np.loadtxt(RATE_FILE)
    assert data.shape == (100, 2)
    return data[:, 0], data[:, 1]


def test_rate_table_file_is_evenly_spaced_in_log():
    """IonizationRate(fname) asserts sorted, evenly log-spaced intensities
    (IonizationRate.jl:14-27); the reference's data file satisfies that."""
    I, R = _rate_rows()
    rt = IonizationRate.from_file(RATE_FILE)
    assert np.allclose(rt.x, np.log10(I)) and np.allclose(rt.y, np.log10(R))
    assert np.all(np.diff(rt.x) > 0) and np.allclose(np.diff(rt.x), np.diff(rt.x)[0])


def test_rate_table_host_evaluation():
    """At the rows: the rates of the file; between rows: the geometric mean of
    neighbouring rates at the geometric mean of their intensities (linear in log-log);
    beyond the ends: the end segments continued (IonizationRate.jl:30-52)."""
    I, R = _rate_rows()
    rt = IonizationRate.from_file(RATE_FILE)
    assert np.allclose(rt(I), R, rtol=1e-12, atol=0)
    mid = np.sqrt(I[:-1] * I[1:])
    assert np.allclose(rt(mid), np.sqrt(R[:-1] * R[1:]), rtol=1e-12, atol=0)
    s0 = math.log10(R[1] / R[0]) / math.log10(I[1] / I[0])
    s1 = math.log10(R[-1] / R[-2]) / math.log10(I[-1] / I[-2])
    assert np.isclose(rt(I[0] / 7)[0], R[0] * 7.0 ** -s0, rtol=1e-12)
    assert np.isclose(rt(I[-1] * 3)[0], R[-1] * 3.0 ** s1, rtol=1e-12)
    assert rt(0.0)[0] == 0.0


@pytest.mark.parametrize("where", ["node", "between", "below", "above"])
def test_rate_table_in_the_plasma_update(where):
    """The oracle's plasma update with the file's table: for a constant field the
    density rises as rho = 1 - exp(-R t) when the avalanche rate is zero (nuc = 0, so
    sigmaB = 0: src/models.jl:1115-1124), with R the file's rate at I = ksi E^2 -- at a
    row, between rows (log-log linear), below and above the table (end segments).  The
    rows used give R dt >= 1e-5, where FP64 resolves 1 - exp(-R dt); below the table the
    file's rates (1e-24 /s) are too small for that, so that case uses the same table
    with every rate scaled by 1e44."""
    I, R = _rate_rows()
    x, y = np.log10(I), np.log10(R)
    shift = 0.0
    k = 80
    if where == "node":
        It, Rt = I[k], R[k]
    elif where == "between":
        It, Rt = math.sqrt(I[k] * I[k + 1]), math.sqrt(R[k] * R[k + 1])
    elif where == "below":
        shift = 44.0
        s0 = (y[1] - y[0]) / (x[1] - x[0])
        It, Rt = I[0] / 5, R[0] * 1e44 * 5.0 ** -s0
    else:
        s1 = (y[-1] - y[-2]) / (x[-1] - x[-2])
        It, Rt = I[-1] * 1.5, R[-1] * 1.5 ** s1
    w0 = 2 * np.pi * C0 / 1030e-9
    pl = Plasma(ionrate=IonizationRate(x=x, y=y + shift), rho0=1.0, nuc=0.0, frequency=w0,
                Uiev=9.0, mr=1.0)
    dz = 3e-8
    dt = dz / C0
    assert Rt * dt >= 1e-5
    steps = int(min(2000, max(2, round(0.5 / (Rt * dt)))))
    model = _uniform_plasma_1d(pl, dz, (steps + 0.5) * dt, CN=1.0)
    assert model.Nt == steps + 1
    c = material_coefficients(Material(geometry=None, eps=1.0, plasma=pl), model.dt)
    assert c["Rava"] == 0.0
    E0 = math.sqrt(It / c["ksi"])
    h = Handle(model.describe(integrals=0), oracle("f64"))
    h.set_field("Dx", np.full(4, EPS0 * E0))
    h.set_field("Ex", np.full(4, E0))
    h.step(steps)
    assert np.allclose(h.field("Ex"), E0, rtol=1e-12)
    rho = h.output(3)[0]
    expect = -math.expm1(-Rt * steps * model.dt)
    assert 0.01 < expect < 0.9 and abs(rho / expect - 1) < 1e-8, (rho, expect)
