"""Geometry helpers: geometry2bool / geometry2indices (src/util.jl:1-100).

A geometry is a predicate of the grid coordinates, a boolean array over the grid, or a
tuple of coordinates naming the nearest grid point.  Predicates are first evaluated on
broadcast coordinate arrays (numpy-style functions); a predicate that cannot take
arrays is evaluated point by point.  Indices are returned as 0-based linear cell
indices, x fastest (column-major, the reference's CartesianIndex order).
"""
import numpy as np

from .grids import coords


def _mesh(grid):
    x, y, z = coords(grid)
    if grid.ndims == 1:
        return (z,)
    if grid.ndims == 2:
        return (x[:, None], z[None, :])
    return (x[:, None, None], y[None, :, None], z[None, None, :])


def evaluate_on_grid(fn, grid, *extra):
    """Evaluate fn(coords..., *extra) over the grid; result shape grid.shape."""
    args = _mesh(grid)
    try:
        with np.errstate(all="ignore"):
            out = np.asarray(fn(*args, *extra))
        if out.shape == grid.shape:
            return out
        if out.ndim == 0:
            return np.broadcast_to(out, grid.shape).copy()
        out = np.broadcast_to(out, grid.shape).copy()
        return out
    except Exception:
        pass
    x, y, z = coords(grid)
    res = np.empty(grid.shape, dtype=object)
    for idx in np.ndindex(*grid.shape):
        if grid.ndims == 1:
            pt = (z[idx[0]],)
        elif grid.ndims == 2:
            pt = (x[idx[0]], z[idx[1]])
        else:
            pt = (x[idx[0]], y[idx[1]], z[idx[2]])
        res[idx] = fn(*pt, *extra)
    return res


def geometry2bool(geometry, grid):
    """Boolean mask over the grid (src/util.jl:1-32)."""
    if callable(geometry):
        mask = evaluate_on_grid(geometry, grid)
    else:
        mask = np.asarray(geometry)
        if mask.shape != grid.shape:
            raise ValueError(f"geometry array has shape {mask.shape}, grid is {grid.shape}")
    return np.asarray(mask, dtype=bool)


def geometry2indices(geometry, grid):
    """0-based linear indices (x fastest) of the selected cells (src/util.jl:35-100).
    A tuple selects the nearest grid point with argmin |x - xg| per axis."""
    if isinstance(geometry, tuple):
        x, y, z = coords(grid)
        if grid.ndims == 1:
            (zg,) = geometry
            return np.array([int(np.argmin(np.abs(z - zg)))], dtype=np.int64)
        if grid.ndims == 2:
            xg, zg = geometry
            ix, iz = int(np.argmin(np.abs(x - xg))), int(np.argmin(np.abs(z - zg)))
            return np.array([ix + grid.Nx * iz], dtype=np.int64)
        xg, yg, zg = geometry
        ix = int(np.argmin(np.abs(x - xg)))
        iy = int(np.argmin(np.abs(y - yg)))
        iz = int(np.argmin(np.abs(z - zg)))
        return np.array([ix + grid.Nx * (iy + grid.Ny * iz)], dtype=np.int64)
    mask = geometry2bool(geometry, grid)
    return np.flatnonzero(mask.ravel(order="F")).astype(np.int64)


def unravel(grid, lin):
    """(ix, iy, iz) 0-based arrays for linear indices (absent axes give zeros)."""
    Nx, Ny, Nz = grid.dims3
    lin = np.asarray(lin, dtype=np.int64)
    return lin % Nx, (lin // Nx) % Ny, lin // (Nx * Ny)
