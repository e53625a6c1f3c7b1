"""Sources: SoftSource, HardSource, TFSFSource (src/sources.jl:1-877).

Waveforms are Python functions of the grid coordinates and time (plus an optional
parameter tuple `p`), as in the reference.  They are evaluated once on the host in
FP64 at t[it] = (it-1)*dt for every source cell, and the device reads one row of that
table per step.
"""
from dataclasses import dataclass, field
from typing import Any, Callable, Optional, Sequence

import numpy as np

from .constants import EPS0
from .grids import coords
from .util import geometry2indices, unravel

COMPONENTS = ("Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ex", "Ey", "Ez")
FIELD_COMPONENTS = {
    1: ("Hy", "Dx", "Ex"),
    2: ("Hy", "Dx", "Dz", "Ex", "Ez"),
    3: COMPONENTS,
}


def comp_index(name):
    name = str(name).lstrip(":")
    if name not in COMPONENTS:
        raise ValueError(f"unknown field component {name!r}")
    return COMPONENTS.index(name)


class Source:
    pass


@dataclass
class SoftSource(Source):
    """SoftSource(; geometry, waveform, p=nothing, component, explicit=false)
    (src/sources.jl:9-37)."""
    geometry: Any
    waveform: Callable
    component: str
    p: Any = None
    explicit: bool = False


@dataclass
class HardSource(Source):
    """HardSource(; geometry, waveform, p=nothing, component) (src/sources.jl:137-161)."""
    geometry: Any
    waveform: Callable
    component: str
    p: Any = None


@dataclass
class TFSFSource(Source):
    """Total-field / scattered-field source (src/sources.jl:250-877).

    `box` holds the face coordinates (z1,z2) / (x1,x2,z1,z2) / (x1,x2,y1,y2,z1,z2),
    mapped to the nearest grid indices as SourceStruct(::TFSFSource) does.  The incident
    field comes either from
      - `tables`: replay arrays in the field order of TFSFSource1D/2D/3D, each shaped like
        the reference's [face..., Nt] array (what a same-grid record file holds); or
      - `waveform` (+ `component`, `z0`, `p`, `explicit`): a normal-incidence plane wave
        produced by the x/y-uniform auxiliary run of the record pass, i.e. a soft
        `component` source on the plane nearest z0 with waveform(t[, p]), computed on the
        device in FP64 every step.
      - `fname`: a record file as solve(...; tfsf_record=True) or the reference's
        prepare_tfsf_record/write_tfsf_record write it (HDF5: box, t, x, y, z, inc*),
        resampled onto the model's grid and time axis with linear interpolation and
        flat extrapolation (src/sources.jl:264-317,387-466,592-710); `box` then comes
        from the file.
    """
    box: Sequence[float] = ()
    tables: Optional[Sequence[np.ndarray]] = None
    waveform: Optional[Callable] = None
    component: str = "Ex"
    z0: Optional[float] = None
    p: Any = None
    explicit: bool = False
    fname: Optional[str] = None



def eval_waveform(fn, p, grid, cells, t):
    """Table w[it, k] = fn(coords of cell k, t[it] [, p]) in FP64, shape (Nt, n)."""
    x, y, z = coords(grid)
    ix, iy, iz = unravel(grid, cells)
    extra = () if p is None else (p,)
    tt = np.asarray(t, dtype=float)[:, None]
    if grid.ndims == 1:
        cargs = (z[iz][None, :],)
    elif grid.ndims == 2:
        cargs = (x[ix][None, :], z[iz][None, :])
    else:
        cargs = (x[ix][None, :], y[iy][None, :], z[iz][None, :])
    n = len(cells)
    try:
        with np.errstate(all="ignore"):
            out = np.asarray(fn(*cargs, tt, *extra), dtype=float)
        return np.ascontiguousarray(np.broadcast_to(out, (len(t), n)), dtype=float)
    except Exception:
        pass
    out = np.empty((len(t), n), dtype=float)
    for k in range(n):
        pt = tuple(a[0, k] for a in cargs)
        for it, tv in enumerate(t):
            out[it, k] = fn(*pt, tv, *extra)
    return out


def point_source_struct(src, grid, t):
    """SourceStruct(::SoftSource/::HardSource, field, t) (src/sources.jl:52-73,175-186)
    -> dict(kind, comp, coeff, cells, wave)."""
    cells = geometry2indices(src.geometry, grid)
    if len(cells) == 0:
        raise ValueError("I did not find any grid points which satisfy your source geometry.")
    comp = str(src.component).lstrip(":")
    coeff = 1.0
    kind = 1 if isinstance(src, HardSource) else 0
    if kind == 0 and comp in ("Ex", "Ey", "Ez") and not src.explicit:
        comp = "D" + comp[1]
        coeff = EPS0
    if comp not in FIELD_COMPONENTS[grid.ndims]:
        raise ValueError(f"a {grid.ndims}-D field has no component {comp}")
    wave = eval_waveform(src.waveform, src.p, grid, cells, t)
    return dict(kind=kind, comp=comp_index(comp), coeff=coeff,
                cells=cells.astype(np.int32), wave=wave)


def _nearest(axis, v):
    return int(np.argmin(np.abs(axis - v))) + 1


def tfsf_struct(src, grid, t):
    """SourceStruct(::TFSFSource, field, t) (src/sources.jl:284-317,408-466,592-710):
    1-based box indices and the incident-field description."""
    x, y, z = coords(grid)
    file_tables = None
    if src.fname is not None:
        from .tfsf_files import replay_tables
        fbox, file_tables = replay_tables(src.fname, grid, t)
        b = list(fbox)
    else:
        b = list(src.box)
    box = [1, 1, 1, 1, 1, 1]
    if grid.ndims == 1:
        box[4], box[5] = _nearest(z, b[0]), _nearest(z, b[1])
    elif grid.ndims == 2:
        box[0], box[1] = _nearest(x, b[0]), _nearest(x, b[1])
        box[4], box[5] = _nearest(z, b[2]), _nearest(z, b[3])
    else:
        box[0], box[1] = _nearest(x, b[0]), _nearest(x, b[1])
        box[2], box[3] = _nearest(y, b[2]), _nearest(y, b[3])
        box[4], box[5] = _nearest(z, b[4]), _nearest(z, b[5])
    if grid.ndims == 1:
        need = [(box[4], box[5], grid.Nz)]
    elif grid.ndims == 2:
        need = [(box[0], box[1], grid.Nx), (box[4], box[5], grid.Nz)]
    else:
        need = [(box[0], box[1], grid.Nx), (box[2], box[3], grid.Ny), (box[4], box[5], grid.Nz)]
    for a, bb, n in need:
        if not (2 <= a < bb <= n):
            raise ValueError("TFSF box must lie strictly inside the grid: 2 <= i1 < i2 <= N")
    if file_tables is not None:
        return dict(box=box, mode=0,
                    tables=[np.ascontiguousarray(np.asarray(tb, dtype=float).ravel(order="F"))
                            for tb in file_tables])
    if src.tables is not None:
        tabs = [np.asarray(tb, dtype=float) for tb in src.tables]
        need_n = {1: 4, 2: 8, 3: 24}[grid.ndims]
        if len(tabs) != need_n:
            raise ValueError(f"{grid.ndims}-D TFSF replay needs {need_n} tables")
        return dict(box=box, mode=0,
                    tables=[np.ascontiguousarray(tb.ravel(order="F")) for tb in tabs])
    if src.waveform is None:
        raise ValueError("TFSFSource needs `tables` or a plane-wave `waveform`")
    comp = str(src.component).lstrip(":")
    coeff = 1.0
    if comp in ("Ex", "Ey") and not src.explicit:
        comp = "D" + comp[1]
        coeff = EPS0
    if comp not in ("Ex", "Ey", "Dx", "Dy", "Hx", "Hy"):
        raise ValueError("plane-wave TFSF excites Ex/Ey (or Dx/Dy, Hx/Hy) only")
    if grid.ndims < 3 and comp in ("Ey", "Dy", "Hx"):
        raise ValueError(f"a {grid.ndims}-D field has no y-polarised plane wave")
    z0 = src.z0 if src.z0 is not None else z[box[4] - 1] - 2 * (z[1] - z[0])
    iz0 = int(np.argmin(np.abs(z - z0)))
    extra = () if src.p is None else (src.p,)
    tt = np.asarray(t, dtype=float)
    try:
        with np.errstate(all="ignore"):
            w = np.asarray(src.waveform(tt, *extra), dtype=float)
        w = np.broadcast_to(w, tt.shape).astype(float)
    except Exception:
        w = np.array([src.waveform(tv, *extra) for tv in tt], dtype=float)
    return dict(box=box, mode=1, aux_comp=[comp_index(comp)], aux_iz=[iz0],
                aux_coeff=[coeff], aux_wave=np.ascontiguousarray(w[None, :]))
