"""Materials, susceptibilities and plasma (src/materials.jl:1-587).

`Material(...)` keeps the reference's keywords.  `material_coefficients` is the FP64
part of MaterialStruct1D/2D/3D (src/materials.jl:115-483): the update coefficients,
ADE recursion coefficients, Kerr and plasma scalars.  The per-cell state arrays are
allocated by the engine, only for the cells of the material.
"""
from dataclasses import dataclass
import math
from typing import Any, Callable, Optional

import numpy as np

from .constants import C0, EPS0, HBAR, ME, MU0, QE


class Susceptibility:
    pass


@dataclass
class DebyeSusceptibility(Susceptibility):
    """chi = deps / (1 - i w tau) (src/materials.jl:494-517)."""
    deps: float
    tau: float


@dataclass
class DrudeSusceptibility(Susceptibility):
    """chi = -wp^2 / (w^2 + i w gamma) (src/materials.jl:523-546)."""
    wp: float
    gamma: float


@dataclass
class LorentzSusceptibility(Susceptibility):
    """chi = deps w0^2 / (w0^2 - w^2 - 2 i w delta) (src/materials.jl:552-577)."""
    deps: float
    w0: float
    delta: float


def susceptibility(chi, w):
    """susceptibility(chi, w) (src/materials.jl:503-506,532-535,562-565,581-587)."""
    if isinstance(chi, (list, tuple)):
        tot = 0.0
        for c in chi:
            tot = tot + susceptibility(c, w)
        return tot
    if isinstance(chi, DebyeSusceptibility):
        return chi.deps / (1 - 1j * w * chi.tau)
    if isinstance(chi, DrudeSusceptibility):
        return -chi.wp ** 2 / (w ** 2 + 1j * w * chi.gamma)
    if isinstance(chi, LorentzSusceptibility):
        return chi.deps * chi.w0 ** 2 / (chi.w0 ** 2 - w ** 2 - 2j * w * chi.delta)
    raise TypeError(f"unknown susceptibility {chi!r}")


def ade_coefficients(chi, dt):
    """(Aq, Bq, Cq) of the ADE recursion P+ = Aq P + Bq P- + Cq E
    (src/materials.jl:509-517,538-546,568-577)."""
    if isinstance(chi, DebyeSusceptibility):
        aq = 1 / chi.tau
        bq = EPS0 * chi.deps / chi.tau
        return 1 - aq * dt, 0.0, bq * dt
    if isinstance(chi, DrudeSusceptibility):
        aq = chi.gamma
        bq = EPS0 * chi.wp ** 2
        return (2 / (aq * dt / 2 + 1), (aq * dt / 2 - 1) / (aq * dt / 2 + 1),
                bq * dt ** 2 / (aq * dt / 2 + 1))
    if isinstance(chi, LorentzSusceptibility):
        aq = 2 * chi.delta
        bq = chi.w0 ** 2
        cq = EPS0 * chi.deps * chi.w0 ** 2
        return ((2 - bq * dt ** 2) / (aq * dt / 2 + 1), (aq * dt / 2 - 1) / (aq * dt / 2 + 1),
                cq * dt ** 2 / (aq * dt / 2 + 1))
    raise TypeError(f"unknown susceptibility {chi!r}")


def ade_plasma_coefficients(nuc, mr, dt):
    """(Ap, Bp, Cp) of the plasma current recursion (src/materials.jl:64-69)."""
    Ap = 2 / (nuc * dt / 2 + 1)
    Bp = (nuc * dt / 2 - 1) / (nuc * dt / 2 + 1)
    Cp = QE ** 2 / (ME * mr) * dt ** 2 / (nuc * dt / 2 + 1)
    return Ap, Bp, Cp


# ---- ionization rates ------------------------------------------------------------------
@dataclass
class PowerLawRate:
    """Multiphoton rate R(I) = sigma * I^K, evaluated in FP64 (0 for I <= 0)."""
    sigma: float
    K: float

    def __call__(self, I):
        I = np.asarray(I, dtype=float)
        with np.errstate(all="ignore"):
            return np.where(I > 0, self.sigma * np.power(np.maximum(I, 0), self.K), 0.0)


@dataclass
class IonizationRate:
    """Tabulated rate with linear interpolation in log10-log10 space and linear
    extrapolation from the end segments (examples/5_nonlinear_propagation/
    IonizationRate.jl:5-52).  x = log10(I), y = log10(R), x evenly spaced."""
    x: np.ndarray
    y: np.ndarray

    @classmethod
    def from_file(cls, fname):
        data = np.loadtxt(fname)
        x = np.log10(data[:, 0])
        y = np.log10(data[:, 1])
        dx = np.diff(x)
        if not np.all(np.diff(x) > 0):
            raise ValueError("ionization table intensities must be sorted")
        if not np.allclose(dx, dx[0]):
            raise ValueError("ionization table intensities must be evenly spaced in log10")
        return cls(x, y)

    def __call__(self, I):
        I = np.atleast_1d(np.asarray(I, dtype=float))
        out = np.zeros_like(I)
        pos = I > 0
        xl = np.log10(I[pos])
        x, y = self.x, self.y
        i = np.searchsorted(x, xl, side="left") - 1
        i = np.where(xl <= x[0], 0, i)
        i = np.where(xl >= x[-1], len(x) - 2, i)
        i = np.clip(i, 0, len(x) - 2)
        dydx = (y[i + 1] - y[i]) / (x[i + 1] - x[i])
        out[pos] = 10.0 ** (y[i] + dydx * (xl - x[i]))
        return out


RATE_TABLE_LOG10_RANGE = (10.0, 20.0)    # intensities 1e10 .. 1e20 W/m^2
RATE_TABLE_POINTS = 2001


def tabulate_rate(fn, lo=RATE_TABLE_LOG10_RANGE[0], hi=RATE_TABLE_LOG10_RANGE[1],
                  n=RATE_TABLE_POINTS):
    """An arbitrary ionization-rate function R(I) as the device evaluates it: an
    IonizationRate table of n points evenly spaced in log10(I) over [lo, hi], linear in
    log-log between them (IonizationRate.jl:30-52).  Rates that are zero or negative at
    a node are floored at 1e-300 (log10 = -300), which the device turns into 0."""
    x = np.linspace(lo, hi, n)
    vals = np.array([float(fn(10.0 ** v)) for v in x])
    y = np.log10(np.maximum(vals, 1e-300))
    return IonizationRate(x=x, y=y)


@dataclass
class Plasma:
    """Plasma(; ionrate, rho0, nuc, frequency, Uiev, mr=1) (src/materials.jl:48-61).
    `ionrate` is a PowerLawRate (evaluated exactly on the device), an IonizationRate
    table, or any function of the intensity, which is tabulated with tabulate_rate
    (2001 log-log points over 1e10-1e20 W/m^2) as the Julia backend shim does."""
    ionrate: Any
    rho0: float
    nuc: float
    frequency: float
    Uiev: float
    mr: float = 1.0


@dataclass
class Material:
    """Material(; geometry, eps=1, mu=1, sigma=0, chi, chi2, chi3, plasma)
    (src/materials.jl:1-44)."""
    geometry: Any
    eps: float = 1.0
    mu: float = 1.0
    sigma: float = 0.0
    chi: Any = None
    chi2: Optional[float] = None
    chi3: Optional[float] = None
    plasma: Optional[Plasma] = None

    def __post_init__(self):
        if self.eps < 0:
            raise ValueError(
                f"You specified negative eps={self.eps}. Negative, frequency-independent "
                "permittivities lead to exponentially growing solutions. To model a "
                "material with negative permittivity, use frequency-dependent "
                "susceptibility chi.")


def material_coefficients(material, dt):
    """FP64 coefficients of MaterialStruct (src/materials.jl:389-483) as a dict."""
    eps, mu, sigma = float(material.eps), float(material.mu), float(material.sigma)
    sigmaD = sigma / (EPS0 * eps)
    c = dict(Mh=dt / (MU0 * mu), Me=1 / (EPS0 * eps),
             Md1=(1 - sigmaD * dt / 2) / (1 + sigmaD * dt / 2),
             Md2=dt / (1 + sigmaD * dt / 2))
    chi = material.chi
    if chi is None:
        c.update(nq=0, Aq=np.zeros(0), Bq=np.zeros(0), Cq=np.zeros(0))
    else:
        chis = chi if isinstance(chi, (list, tuple)) else (chi,)
        coefs = np.array([ade_coefficients(ch, dt) for ch in chis], dtype=float)
        c.update(nq=len(chis), Aq=coefs[:, 0].copy(), Bq=coefs[:, 1].copy(),
                 Cq=coefs[:, 2].copy())
    if material.chi2 is None and material.chi3 is None:
        c.update(iskerr=0, Mk2=0.0, Mk3=0.0)
    else:
        c.update(iskerr=1,
                 Mk2=0.0 if material.chi2 is None else EPS0 * material.chi2,
                 Mk3=0.0 if material.chi3 is None else EPS0 * material.chi3)
    pl = material.plasma
    if pl is None:
        c.update(isplasma=0, Rava=0.0, ksi=0.0, rho0=0.0, Ap=0.0, Bp=0.0, Cp=0.0, Ma=0.0,
                 rate_kind=0, rate_c=0.0, rate_k=0.0, rate_x=np.zeros(0), rate_y=np.zeros(0))
    else:
        chi0 = susceptibility(chi, pl.frequency) if chi is not None else 0.0
        n = math.sqrt((eps + np.real(chi0)) * mu)
        ksi = n * EPS0 * C0 / 2
        Ap, Bp, Cp = ade_plasma_coefficients(pl.nuc, pl.mr, dt)
        Ui = pl.Uiev * QE
        Wph = HBAR * pl.frequency
        K = math.ceil(Ui / Wph)
        sigmaB = QE ** 2 / (ME * pl.mr) * pl.nuc / (pl.nuc ** 2 + pl.frequency ** 2)
        Rava = sigmaB / Ui
        Ma = K * Wph * dt
        r = pl.ionrate
        if isinstance(r, PowerLawRate):
            rate = dict(rate_kind=1, rate_c=float(r.sigma), rate_k=float(r.K),
                        rate_x=np.zeros(0), rate_y=np.zeros(0))
        elif isinstance(r, IonizationRate):
            rate = dict(rate_kind=2, rate_c=0.0, rate_k=0.0,
                        rate_x=np.asarray(r.x, float), rate_y=np.asarray(r.y, float))
        elif callable(r):
            t = tabulate_rate(r)
            rate = dict(rate_kind=2, rate_c=0.0, rate_k=0.0, rate_x=t.x, rate_y=t.y)
        else:
            raise TypeError("Plasma.ionrate must be a PowerLawRate, an IonizationRate or a "
                            "function of the intensity")
        c.update(isplasma=1, Rava=Rava, ksi=ksi, rho0=float(pl.rho0), Ap=Ap, Bp=Bp, Cp=Cp,
                 Ma=Ma, **rate)
    return c
