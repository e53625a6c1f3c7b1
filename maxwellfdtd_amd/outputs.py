"""The solve! output file (Output, write_fields, write_monitors, write_integral_variables,
src/outputs.jl:46-65,71-485; write_monitor, src/monitors.jl:72-78,125-135,181-192,
264-274,305-316,362-373), written with this package's HDF5 writer.

Layout, as the reference writes it: the coordinates (z; x, z; x, y, z), t, pml (the
layers' 1-based boundary indices), geometry (when any cell holds a material), the
group fields with t and one [grid..., Ntout] dataset per component (plus rho for
plasma), Sa, E2 (with materials), rho_end and JE (with plasma), and monitors/n/ with
issum, inds, w (spectral) and one array per component.  Field frames are written as
the run reaches each output time, in the precision of the engine's state (FP32, or
FP64 for an FP64 1D model).  inds holds each point's 1-based (ix[, iy], iz) as the
columns of an Int matrix (the reference stores CartesianIndex values).
"""
import numpy as np

from .grids import coords
from .hdf5 import H5Writer

COMPONENTS = {1: ("Hy", "Ex"), 2: ("Hy", "Ex", "Ez"), 3: ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")}


def check_components(ndims, components):
    """The component check of Output (src/outputs.jl:93-106,207-222,367-384)."""
    names = COMPONENTS[ndims]
    if components is None:
        return names
    comps = tuple(str(c).lstrip(":") for c in components)
    for c in comps:
        if c not in names:
            what = {1: ":Hy and :Ex", 2: ":Hy, :Ex, and :Ez",
                    3: ":Hx, :Hy, :Hz, :Ex, :Ey, and :Ez"}[ndims]
            raise ValueError(f"You asked to output :{c} field component, but {ndims}D field "
                             f"has only {what} components.")
    return comps


class OutputFile:
    """The output file of one solve: created before the time loop, fields written at
    every output time, monitors and integral variables written at the end."""

    def __init__(self, fname, model, tout, components, dtype):
        g = model.grid
        self.grid = g
        self.w = H5Writer(fname)
        self.components = components
        self.isfields = tout is not None and len(tout) > 0
        self.isplasma = model.isplasma
        self.isgeometry = model.isgeometry
        x, y, z = coords(g)
        if g.ndims >= 2:
            self.w.create("x", x)
        if g.ndims == 3:
            self.w.create("y", y)
        self.w.create("z", z)
        self.w.create("t", np.asarray(model.t, dtype=float))
        faces = {1: (4, 5), 2: (0, 1, 4, 5), 3: (0, 1, 2, 3, 4, 5)}[g.ndims]
        self.w.create("pml", np.array([model.pml[f].ib for f in faces], dtype=np.int64))
        if self.isgeometry:
            self.w.create("geometry", np.asarray(model.geometry, dtype=np.int64))
        self.shape = g.shape
        if self.isfields:
            self.w.create("fields/t", np.asarray(tout, dtype=float))
            nt = len(tout)
            for c in components:
                self.w.create_dataset(f"fields/{c}", dtype, tuple(self.shape) + (nt,))
            if self.isplasma:
                self.w.create_dataset("fields/rho", dtype, tuple(self.shape) + (nt,))

    def write_fields(self, k, frames):
        """Frame k (0-based) of the field datasets; frames maps component -> array."""
        for c, v in frames.items():
            self.w.write_last(f"fields/{c}", k, v)

    def finish(self, Sa, E2, JE, rho, monitors):
        """write_monitors and write_integral_variables (src/outputs.jl:27-65)."""
        for n, mon in enumerate(monitors, start=1):
            base = f"monitors/{n}"
            self.w.create(f"{base}/issum", np.bool_(mon["issum"]))
            self.w.create(f"{base}/inds", np.asarray(mon["inds_1based"], dtype=np.int64))
            if mon["kind"] == "spectral":
                self.w.create(f"{base}/w", np.asarray(mon["w"], dtype=float))
            for c in COMPONENTS[self.grid.ndims]:
                self.w.create(f"{base}/{c}", mon[c])
        self.w.create("Sa", Sa)
        if self.isgeometry:
            self.w.create("E2", E2)
            if self.isplasma:
                self.w.create("rho_end", rho)
                self.w.create("JE", JE)
        self.w.close()

    def close(self):
        self.w.close()


def cartesian_inds(grid, cells):
    """1-based Cartesian indices of 0-based linear cells: a (ndims, n) Int matrix."""
    Nx, Ny, Nz = grid.dims3
    cells = np.asarray(cells, dtype=np.int64)
    ix, iy, iz = cells % Nx, (cells // Nx) % Ny, cells // (Nx * Ny)
    rows = {1: [iz], 2: [ix, iz], 3: [ix, iy, iz]}[grid.ndims]
    return np.array(rows, dtype=np.int64) + 1
