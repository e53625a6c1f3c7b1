"""Spatial grids: Grid, Grid1D, Grid2D, Grid3D (src/grids.jl:65-240).

Coordinates follow `range(min, max, N)`; the spacing is `x[2] - x[1]` as in the
reference.  Julia builds the range with TwicePrecision arithmetic; here the points are
`min + i*step` in FP64, which can differ from Julia's by one ulp.
"""
from dataclasses import dataclass

import numpy as np


def _range(vmin, vmax, n):
    n = int(n)
    if n < 2:
        raise ValueError("a grid axis needs at least 2 points")
    return np.linspace(float(vmin), float(vmax), n)


@dataclass
class Grid1D:
    """1-D grid in z (src/grids.jl:102-129)."""
    Nz: int
    dz: float
    Lz: float
    z: np.ndarray
    ndims = 1

    @classmethod
    def build(cls, zmin, zmax, Nz):
        z = _range(zmin, zmax, Nz)
        return cls(len(z), z[1] - z[0], z[-1] - z[0], z)

    @property
    def shape(self):
        return (self.Nz,)

    @property
    def dims3(self):
        return (1, 1, self.Nz)


@dataclass
class Grid2D:
    """2-D grid in (x, z) (src/grids.jl:135-178)."""
    Nx: int
    Nz: int
    dx: float
    dz: float
    Lx: float
    Lz: float
    x: np.ndarray
    z: np.ndarray
    ndims = 2

    @classmethod
    def build(cls, xmin, xmax, Nx, zmin, zmax, Nz):
        x = _range(xmin, xmax, Nx)
        z = _range(zmin, zmax, Nz)
        return cls(len(x), len(z), x[1] - x[0], z[1] - z[0], x[-1] - x[0], z[-1] - z[0], x, z)

    @property
    def shape(self):
        return (self.Nx, self.Nz)

    @property
    def dims3(self):
        return (self.Nx, 1, self.Nz)


@dataclass
class Grid3D:
    """3-D grid in (x, y, z) (src/grids.jl:184-240)."""
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    Lx: float
    Ly: float
    Lz: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    ndims = 3

    @classmethod
    def build(cls, xmin, xmax, Nx, ymin, ymax, Ny, zmin, zmax, Nz):
        x = _range(xmin, xmax, Nx)
        y = _range(ymin, ymax, Ny)
        z = _range(zmin, zmax, Nz)
        return cls(len(x), len(y), len(z), x[1] - x[0], y[1] - y[0], z[1] - z[0],
                   x[-1] - x[0], y[-1] - y[0], z[-1] - z[0], x, y, z)

    @property
    def shape(self):
        return (self.Nx, self.Ny, self.Nz)

    @property
    def dims3(self):
        return (self.Nx, self.Ny, self.Nz)


def Grid(*, xmin=None, xmax=None, Nx=None, ymin=None, ymax=None, Ny=None, zmin, zmax, Nz):
    """Grid(; ...) dispatcher (src/grids.jl:65-96): 1D needs z, 2D x and z, 3D x, y, z."""
    xargs = (xmin, xmax, Nx)
    yargs = (ymin, ymax, Ny)
    if all(a is None for a in xargs) and all(a is None for a in yargs):
        return Grid1D.build(zmin, zmax, Nz)
    if all(a is not None for a in xargs) and all(a is None for a in yargs):
        return Grid2D.build(xmin, xmax, Nx, zmin, zmax, Nz)
    if all(a is not None for a in xargs) and all(a is not None for a in yargs):
        return Grid3D.build(xmin, xmax, Nx, ymin, ymax, Ny, zmin, zmax, Nz)
    raise ValueError(
        "Wrong set of arguments.\n"
        "       To create a grid you need to define the following variables:\n"
        "       1-D grid:\n           zmin, zmax, Nz\n"
        "       2-D grid:\n           xmin, xmax, Nx\n           zmin, zmax, Nz\n"
        "       3-D grid:\n           xmin, xmax, Nx\n           ymin, ymax, Ny\n"
        "           zmin, zmax, Nz")


def coords(grid):
    """Coordinate axes (x, y, z) with length-1 placeholders for absent axes."""
    if grid.ndims == 1:
        return np.zeros(1), np.zeros(1), grid.z
    if grid.ndims == 2:
        return grid.x, np.zeros(1), grid.z
    return grid.x, grid.y, grid.z
