"""Test helpers: bind the oracle (test infrastructure) through the same ABI call
sequence the engine receives."""
import os
import subprocess

import numpy as np

from maxwellfdtd_amd import _lib
from maxwellfdtd_amd.models import Handle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_ORACLES = {}


def oracle(precision="f64"):
    """ctypes binding of oracle/liboracle_<precision>.so (built on demand with make)."""
    if precision not in _ORACLES:
        path = os.path.join(ORACLE_DIR, f"liboracle_{precision}.so")
        src = os.path.join(ORACLE_DIR, "fdtd_oracle.c")
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
            subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)
        import ctypes as C
        extra = {"set_threads": (C.c_int32, [C.c_int32]), "real_size": (C.c_int32, [])}
        _ORACLES[precision] = _lib.Binding(path, "ofd_", extra)
    return _ORACLES[precision]


def run_oracle(model, monitors=None, integrals=None, precision="f64", steps=None):
    """Run `model` on the oracle; returns the Handle (caller reads fields)."""
    desc = model.describe(monitors=monitors, integrals=integrals)
    h = Handle(desc, oracle(precision))
    h.step(model.Nt if steps is None else steps)
    return h


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(float)
        b = b.astype(float)
    den = max(np.linalg.norm(a), np.linalg.norm(b))
    return 0.0 if den == 0 else float(np.linalg.norm(a - b) / den)


def isapprox(a, b, rtol):
    """Julia isapprox for arrays: norm(a-b) <= rtol*max(norm(a), norm(b))."""
    return rel_l2(a, b) <= rtol


FIELD_GROUPS = {1: {"H": ("Hy",), "D": ("Dx",), "E": ("Ex",)},
                2: {"H": ("Hy",), "D": ("Dx", "Dz"), "E": ("Ex", "Ez")},
                3: {"H": ("Hx", "Hy", "Hz"), "D": ("Dx", "Dy", "Dz"), "E": ("Ex", "Ey", "Ez")}}


def field_errors(ha, hb, ndims):
    """Relative L2 error of the vector fields H, D, E (all components together) of two
    handles: ||a - b|| / max(||a||, ||b||)."""
    out = {}
    for g, comps in FIELD_GROUPS[ndims].items():
        a = np.concatenate([ha.field(c) for c in comps])
        b = np.concatenate([hb.field(c) for c in comps])
        out[g] = rel_l2(a, b)
    return out


def field_energy(h):
    """(W_E, W_H) = 1/2 sum eps0 eps |E|^2 dV, 1/2 sum mu0 mu |H|^2 dV of a handle's
    current fields, in FP64 on the host: eps0*eps = 1/Me and mu0*mu = dt/Mh of each
    cell's material (vacuum for id 0), dV = dx dy dz / dx dz / dz."""
    d = h.desc
    geo = np.asarray(d.geometry)
    Me = np.array([d.vac[1]] + [c["Me"] for c in d.materials], dtype=float)
    Mh = np.array([d.vac[0]] + [c["Mh"] for c in d.materials], dtype=float)
    eps = 1.0 / Me[geo]
    mu = d.dt / Mh[geo]
    names = {1: (("Ex",), ("Hy",)), 2: (("Ex", "Ez"), ("Hy",)),
             3: (("Ex", "Ey", "Ez"), ("Hx", "Hy", "Hz"))}[d.ndims]
    e2 = sum(h.field(c) ** 2 for c in names[0])
    h2 = sum(h.field(c) ** 2 for c in names[1])
    dx, dy, dz = d.d
    dV = {1: dz, 2: dx * dz, 3: dx * dy * dz}[d.ndims]
    return 0.5 * dV * float(np.sum(eps * e2)), 0.5 * dV * float(np.sum(mu * h2))


def spectra_errors(hg, ho, desc):
    """Relative L2 and the max relative error of every spectral monitor's spectra: per
    component and point, max over frequencies of |S_gpu - S_ref| over the largest
    |S_ref| of that spectrum.  Spectra whose largest value is below a thousandth of the
    largest spectrum of their field (E or H, all components and points) are left out of
    the maximum: points the field has not reached and components that vanish by
    symmetry hold only rounding noise there.  They are in the relative L2."""
    from maxwellfdtd_amd.models import MONITOR_COMPONENTS
    out = {}
    for i, mon in enumerate(desc.monitors):
        if mon["kind"] != "spectral":
            continue
        a_all, b_all, spec = [], [], {}
        for c in MONITOR_COMPONENTS[desc.ndims]:
            a = np.atleast_2d(hg.monitor(i, c))
            b = np.atleast_2d(ho.monitor(i, c))
            a_all.append(a.ravel())
            b_all.append(b.ravel())
            spec[c] = (np.max(np.abs(a - b), axis=-1), np.max(np.abs(b), axis=-1))
        peak = {f: max((float(np.max(d)) for c, (_, d) in spec.items() if c[0] == f and d.size),
                       default=0.0) for f in "EH"}
        worst = 0.0
        for c, (err, den) in spec.items():
            ok = den > 1e-3 * peak[c[0]]
            if np.any(ok):
                worst = max(worst, float(np.max(err[ok] / den[ok])))
        out[f"dft{i}_rel_l2"] = float(f"{rel_l2(np.concatenate(a_all), np.concatenate(b_all)):.3e}")
        out[f"dft{i}_max_rel"] = float(f"{worst:.3e}")
    return out
