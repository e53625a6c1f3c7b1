"""TFSF record files: what solve!(...; tfsf_record=true, tfsf_box) writes and what
TFSFSource(fname) replays (src/sources.jl:264-317,349-384,387-466,498-591,592-710,
779-877).

Recording runs on the engine: the face samples the record pass writes each step
(write_tfsf_record) are taken by field monitors on the same cells, and the file is
written once the run has finished, with the reference's datasets (box, x, y, z, t and
the inc* arrays, each [face..., Nt]).

Replay reads such a file onto the model's own grid and time axis the way
SourceStruct(::TFSFSource) does: the box coordinates go to the nearest grid indices,
and every table is resampled with gridded linear interpolation and flat (clamped)
extrapolation, the semantics of Interpolations.jl's
linear_interpolation(knots, data; extrapolation_bc=Flat()) (src/sources.jl:302,436,639).
On the grid and time axis it was recorded on, replay returns the recorded tables
exactly.
"""
import numpy as np

from .grids import coords
from .hdf5 import H5File, H5Writer

# table names in the field order of TFSFSource1D/2D/3D (the engine's replay order)
NAMES = {
    1: ["incHy_1", "incEx_1", "incHy_2", "incEx_2"],
    2: ["incHy_1z", "incEz_1z", "incHy_2z", "incEz_2z",
        "incHy_x1", "incEx_x1", "incHy_x2", "incEx_x2"],
    3: ["incHy_1yz", "incHz_1yz", "incEy_1yz", "incEz_1yz",
        "incHy_2yz", "incHz_2yz", "incEy_2yz", "incEz_2yz",
        "incHx_x1z", "incHz_x1z", "incEx_x1z", "incEz_x1z",
        "incHx_x2z", "incHz_x2z", "incEx_x2z", "incEz_x2z",
        "incHx_xy1", "incHy_xy1", "incEx_xy1", "incEy_xy1",
        "incHx_xy2", "incHy_xy2", "incEx_xy2", "incEy_xy2"],
}


def nearest_box(grid, box):
    """1-based box indices, argmin(abs.(x .- x1)) per face coordinate."""
    x, y, z = coords(grid)
    nd = grid.ndims
    axes = {1: [z], 2: [x, z], 3: [x, y, z]}[nd]
    out = []
    for k, ax in enumerate(axes):
        out.append(int(np.argmin(np.abs(ax - box[2 * k]))) + 1)
        out.append(int(np.argmin(np.abs(ax - box[2 * k + 1]))) + 1)
    return out


def record_plan(grid, box):
    """For every table of the record pass (write_tfsf_record): (name, component, 0-based
    index ranges per axis (x, y, z; a range or a single plane)).  `box` holds the
    coordinates; indices as the reference computes them."""
    nd = grid.ndims
    b = nearest_box(grid, box)
    r = lambda a, c: np.arange(a - 1, c)          # 1-based a..c inclusive -> 0-based
    one = lambda a: np.array([a - 1])
    if nd == 1:
        iz1, iz2 = b
        return [("incHy_1", "Hy", (one(1), one(1), one(iz1 - 1))),
                ("incEx_1", "Ex", (one(1), one(1), one(iz1))),
                ("incHy_2", "Hy", (one(1), one(1), one(iz2))),
                ("incEx_2", "Ex", (one(1), one(1), one(iz2)))]
    if nd == 2:
        ix1, ix2, iz1, iz2 = b
        Z, X = r(iz1, iz2 - 1), r(ix1, ix2 - 1)
        return [("incHy_1z", "Hy", (one(ix1 - 1), one(1), Z)), ("incEz_1z", "Ez", (one(ix1), one(1), Z)),
                ("incHy_2z", "Hy", (one(ix2), one(1), Z)), ("incEz_2z", "Ez", (one(ix2), one(1), Z)),
                ("incHy_x1", "Hy", (X, one(1), one(iz1 - 1))), ("incEx_x1", "Ex", (X, one(1), one(iz1))),
                ("incHy_x2", "Hy", (X, one(1), one(iz2))), ("incEx_x2", "Ex", (X, one(1), one(iz2)))]
    ix1, ix2, iy1, iy2, iz1, iz2 = b
    X, Xm = r(ix1, ix2), r(ix1, ix2 - 1)
    Y, Ym = r(iy1, iy2), r(iy1, iy2 - 1)
    Z, Zm = r(iz1, iz2), r(iz1, iz2 - 1)
    plan = []
    for side, nx_h, nx_e in (("1yz", ix1 - 1, ix1), ("2yz", ix2, ix2)):
        plan += [(f"incHy_{side}", "Hy", (one(nx_h), Y, Zm)), (f"incHz_{side}", "Hz", (one(nx_h), Ym, Z)),
                 (f"incEy_{side}", "Ey", (one(nx_e), Ym, Z)), (f"incEz_{side}", "Ez", (one(nx_e), Y, Zm))]
    for side, ny_h, ny_e in (("x1z", iy1 - 1, iy1), ("x2z", iy2, iy2)):
        plan += [(f"incHx_{side}", "Hx", (X, one(ny_h), Zm)), (f"incHz_{side}", "Hz", (Xm, one(ny_h), Z)),
                 (f"incEx_{side}", "Ex", (Xm, one(ny_e), Z)), (f"incEz_{side}", "Ez", (X, one(ny_e), Zm))]
    for side, nz_h, nz_e in (("xy1", iz1 - 1, iz1), ("xy2", iz2, iz2)):
        plan += [(f"incHx_{side}", "Hx", (X, Ym, one(nz_h))), (f"incHy_{side}", "Hy", (Xm, Y, one(nz_h))),
                 (f"incEx_{side}", "Ex", (Xm, Y, one(nz_e))), (f"incEy_{side}", "Ey", (X, Ym, one(nz_e)))]
    return plan


def plan_cells(grid, ranges):
    """0-based linear cells of an index product, x fastest, and its Julia shape (the
    plane axis dropped)."""
    Nx, Ny, Nz = grid.dims3
    ix, iy, iz = ranges
    I, J, K = np.meshgrid(ix, iy, iz, indexing="ij")
    cells = (I + Nx * (J + Ny * K)).ravel(order="F").astype(np.int32)
    shape = tuple(len(a) for a in ranges if len(a) > 1)
    return cells, shape


def record_monitors(grid, box):
    """Monitor descriptions for the record pass: one field monitor per table."""
    out = []
    for name, comp, ranges in record_plan(grid, box):
        cells, shape = plan_cells(grid, ranges)
        out.append(dict(kind="field", cells=cells, issum=False, tfsf=(name, comp, shape)))
    return out


def write_record_file(fname, grid, t, box, tables):
    """The record file of prepare_tfsf_record/write_tfsf_record: t, box, the box's
    coordinate ranges (2D: x[ix1:ix2-1], z[iz1:iz2-1]; 3D: x[ix1:ix2], y, z) and every
    inc* table ([face..., Nt], the recording engine's precision)."""
    nd = grid.ndims
    x, y, z = coords(grid)
    b = nearest_box(grid, box)
    with H5Writer(fname) as w:
        if nd == 2:
            ix1, ix2, iz1, iz2 = b
            w.create("x", x[ix1 - 1:ix2 - 1])
            w.create("z", z[iz1 - 1:iz2 - 1])
        elif nd == 3:
            ix1, ix2, iy1, iy2, iz1, iz2 = b
            w.create("x", x[ix1 - 1:ix2])
            w.create("y", y[iy1 - 1:iy2])
            w.create("z", z[iz1 - 1:iz2])
        w.create("t", np.asarray(t, dtype=float))
        w.create("box", np.asarray(box, dtype=float))
        for name in NAMES[nd]:
            w.create(name, tables[name])


def interp_flat(knots, data, queries):
    """Gridded multilinear interpolation of data (shape = knot counts) at the tensor
    product of the query axes, with flat extrapolation: queries outside a knot range
    take the end value.  Linear along each axis in turn, which on a tensor grid is
    the multilinear interpolant."""
    out = np.asarray(data, dtype=float)
    for axis, (k, q) in enumerate(zip(knots, queries)):
        k = np.asarray(k, dtype=float)
        q = np.asarray(q, dtype=float)
        if len(k) == 1:
            out = np.repeat(np.take(out, [0], axis=axis), len(q), axis=axis)
            continue
        qc = np.clip(q, k[0], k[-1])
        i = np.clip(np.searchsorted(k, qc, side="right") - 1, 0, len(k) - 2)
        w = (qc - k[i]) / (k[i + 1] - k[i])
        shp = [1] * out.ndim
        shp[axis] = len(q)
        w = w.reshape(shp)
        a = np.take(out, i, axis=axis)
        b = np.take(out, i + 1, axis=axis)
        out = a + w * (b - a)
        exact = (w == 0.0)
        out = np.where(exact, a, out)
        out = np.where(w == 1.0, b, out)
    return out


def replay_tables(fname, grid, t):
    """TFSF replay tables for a model on `grid` with time axis `t`, read from a record
    file (SourceStruct(::TFSFSource), src/sources.jl:284-317,408-466,592-710): the box
    coordinates of the file mapped to this grid, each table resampled onto this grid's
    nodes and time axis.  Returns (box coordinates, tables in the engine's order as
    Fortran-ordered arrays [face..., Nt])."""
    f = H5File(fname)
    nd = grid.ndims
    box = np.asarray(f["box"], dtype=float)
    tex = np.asarray(f["t"], dtype=float)
    t = np.asarray(t, dtype=float)
    x, y, z = coords(grid)
    b = nearest_box(grid, box)
    tabs = []
    if nd == 1:
        for name in NAMES[1]:
            tabs.append(interp_flat([tex], f[name], [t]))
        return box, tabs
    if nd == 2:
        ix1, ix2, iz1, iz2 = b
        xex, zex = np.asarray(f["x"]), np.asarray(f["z"])
        zq, xq = z[iz1 - 1:iz2 - 1], x[ix1 - 1:ix2 - 1]
        for name in NAMES[2]:
            if name.endswith("z"):
                tabs.append(interp_flat([zex, tex], f[name], [zq, t]))
            else:
                tabs.append(interp_flat([xex, tex], f[name], [xq, t]))
        return box, tabs
    ix1, ix2, iy1, iy2, iz1, iz2 = b
    xex, yex, zex = np.asarray(f["x"]), np.asarray(f["y"]), np.asarray(f["z"])
    full = {"x": (xex, x[ix1 - 1:ix2]), "y": (yex, y[iy1 - 1:iy2]), "z": (zex, z[iz1 - 1:iz2])}
    short = {k: (v[0][:-1], v[1][:-1]) for k, v in full.items()}
    for name in NAMES[3]:
        comp, face = name[3:5], name[6:]
        tang = {"1yz": "yz", "2yz": "yz", "x1z": "xz", "x2z": "xz", "xy1": "xy", "xy2": "xy"}[face]
        u, v = tang
        # one tangential axis runs over the short range (N - 1 nodes): an E component's
        # own axis, an H component's other tangential axis (prepare_tfsf_record shapes)
        own = comp[1]
        short_ax = own if comp[0] == "E" else (v if own == u else u)
        ku, qu = short[u] if short_ax == u else full[u]
        kv, qv = short[v] if short_ax == v else full[v]
        tabs.append(interp_flat([ku, kv, tex], f[name], [qu, qv, t]))
    return box, tabs
