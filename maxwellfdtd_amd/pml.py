"""Convolutional PML: CPML and the layer profiles (src/pml.jl:1-304).

Each face layer carries the reference's (ib, Nb, K, A, B): ib is the 1-based index of
the layer boundary, K = kappa, B = exp(-(sigma/K + alpha)*dt/EPS0) and
A = sigma/(sigma*K + alpha*K^2)*(B - 1) over the Nb layer points (src/pml.jl:57-127).
A zero thickness means "no layer" (the reference raises UndefVarError for it,
src/pml.jl:58-62,94-98; here the layer is simply absent).
"""
from dataclasses import dataclass
import math

import numpy as np

from .constants import EPS0, MU0


@dataclass
class CPML:
    """CPML(; thickness, kmax=1, alpha=10e-6, R0=10e-6, m=3) (src/pml.jl:37-40)."""
    thickness: float
    kmax: float = 1.0
    alpha: float = 10e-6
    R0: float = 10e-6
    m: int = 3


@dataclass
class PMLLayer:
    ib: int            # 1-based boundary index
    Nb: int
    K: np.ndarray
    A: np.ndarray
    B: np.ndarray
    present: bool = True


def _params(L):
    if isinstance(L, CPML):
        return float(L.thickness), float(L.kmax), float(L.alpha), float(L.R0), int(L.m)
    return float(L), 1.0, 10e-6, 10e-6, 3


def _profile(x, xb, idx, Lx, dt, kmax, alpha, R0, m):
    eta0 = math.sqrt(MU0 / EPS0)
    smax = -(m + 1) * math.log(R0) / (2 * eta0 * Lx)
    d = np.abs(x[idx] - xb) / Lx
    kappa = 1 + (kmax - 1) * d ** m
    sigma = smax * d ** m
    K = kappa
    B = np.exp(-(sigma / K + alpha) * dt / EPS0)
    A = sigma / (sigma * K + alpha * K ** 2) * (B - 1)
    return K, A, B


def left_layer(x, L, dt):
    """LeftPMLLayer (src/pml.jl:57-84)."""
    Lx, kmax, alpha, R0, m = _params(L)
    if Lx == 0:
        return PMLLayer(0, 0, np.ones(1), np.zeros(1), np.zeros(1), present=False)
    xb = x[0] + Lx
    ixb = int(np.argmin(np.abs(x - xb))) + 1          # 1-based
    idx = np.arange(0, ixb)
    K, A, B = _profile(x, xb, idx, Lx, dt, kmax, alpha, R0, m)
    return PMLLayer(ixb, ixb, K, A, B)


def right_layer(x, L, dt):
    """RightPMLLayer (src/pml.jl:94-121)."""
    Lx, kmax, alpha, R0, m = _params(L)
    Nx = len(x)
    if Lx == 0:
        return PMLLayer(Nx + 1, 0, np.ones(1), np.zeros(1), np.zeros(1), present=False)
    xb = x[-1] - Lx
    ixb = int(np.argmin(np.abs(x - xb))) + 1
    idx = np.arange(ixb - 1, Nx)
    K, A, B = _profile(x, xb, idx, Lx, dt, kmax, alpha, R0, m)
    return PMLLayer(ixb, Nx - ixb + 1, K, A, B)


def PML(grid, pml, dt):
    """PML(grid, pml, dt) (src/pml.jl:147-304): six layers in the order
    x-left, x-right, y-left, y-right, z-left, z-right (absent axes give no layer).
    `pml` is a thickness, a CPML, or a per-face tuple (zl,zr) / (xl,xr,zl,zr) /
    (xl,xr,yl,yr,zl,zr)."""
    nd = grid.ndims
    nface = 2 * nd
    if isinstance(pml, (int, float, CPML)):
        Ls = [pml] * nface
    else:
        Ls = list(pml)
        if len(Ls) != nface:
            raise ValueError(f"pml tuple for a {nd}-D grid needs {nface} entries")
    absent = PMLLayer(0, 0, np.ones(1), np.zeros(1), np.zeros(1), present=False)
    layers = [absent] * 6
    if nd == 1:
        layers[4] = left_layer(grid.z, Ls[0], dt)
        layers[5] = right_layer(grid.z, Ls[1], dt)
    elif nd == 2:
        layers[0] = left_layer(grid.x, Ls[0], dt)
        layers[1] = right_layer(grid.x, Ls[1], dt)
        layers[4] = left_layer(grid.z, Ls[2], dt)
        layers[5] = right_layer(grid.z, Ls[3], dt)
    else:
        layers[0] = left_layer(grid.x, Ls[0], dt)
        layers[1] = right_layer(grid.x, Ls[1], dt)
        layers[2] = left_layer(grid.y, Ls[2], dt)
        layers[3] = right_layer(grid.y, Ls[3], dt)
        layers[4] = left_layer(grid.z, Ls[4], dt)
        layers[5] = right_layer(grid.z, Ls[5], dt)
    return layers
