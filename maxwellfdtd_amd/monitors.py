"""Monitors: FieldMonitor, SpectralMonitor and the DFT frequency grid
(src/monitors.jl:1-373, src/dft.jl:1-51)."""
from dataclasses import dataclass
from typing import Any, Optional, Sequence

import numpy as np

from .util import geometry2indices


@dataclass
class FieldMonitor:
    """FieldMonitor(geometry; sum=false) (src/monitors.jl:7-30)."""
    geometry: Any
    sum: bool = False


@dataclass
class SpectralMonitor:
    """SpectralMonitor(geometry; wmin, wmax, sum=false) (src/monitors.jl:198-223).
    `w` (an extension) gives explicit angular frequencies instead of the FFT bins."""
    geometry: Any
    wmin: Optional[float] = None
    wmax: Optional[float] = None
    sum: bool = False
    w: Optional[Sequence[float]] = None


def dft_grid(t, wmin=None, wmax=None):
    """DFT(t; wmin, wmax) (src/dft.jl:10-24): FFT-bin frequencies clipped to the bins
    nearest wmin/wmax, and tshift."""
    Nt = len(t)
    dt = t[1] - t[0]
    tshift = (Nt // 2) * dt if Nt % 2 == 0 else ((Nt + 1) // 2) * dt
    Nw = Nt // 2 + 1
    w = np.array([2 * np.pi * iw / (dt * Nt) for iw in range(Nw)])
    iwmin = 0 if wmin is None else int(np.argmin(np.abs(w - wmin)))
    iwmax = Nw - 1 if wmax is None else int(np.argmin(np.abs(w - wmax)))
    return w[iwmin:iwmax + 1], tshift


def monitor_struct(mon, grid, t):
    """Monitor(monitor, grid, t) (src/monitors.jl:43-69,237-259)."""
    cells = geometry2indices(mon.geometry, grid)
    if len(cells) == 0:
        raise ValueError("I did not find any grid points which satisfy your monitor geometry.")
    if isinstance(mon, FieldMonitor):
        return dict(kind="field", cells=cells.astype(np.int32), issum=bool(mon.sum))
    if mon.w is not None:
        Nt = len(t)
        dt = t[1] - t[0]
        tshift = (Nt // 2) * dt if Nt % 2 == 0 else ((Nt + 1) // 2) * dt
        w = np.asarray(mon.w, dtype=float)
    else:
        w, tshift = dft_grid(t, mon.wmin, mon.wmax)
    return dict(kind="spectral", cells=cells.astype(np.int32), issum=bool(mon.sum),
                w=np.ascontiguousarray(w), tshift=float(tshift))
