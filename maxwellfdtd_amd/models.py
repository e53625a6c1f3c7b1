"""Model assembly and the solve entry point (src/models.jl:1-242,1178-1225).

`Model(grid, source; tmax, CN, bc, pml, material)` does the reference's host-side FP64
work (time grid, CPML profiles, geometry ids, material coefficients, source tables).
`solve(model, ...)` hands that description to the MI355X engine through the C ABI,
runs the time loop there and returns the fields, monitors and integral outputs.
"""
import ctypes as C
from dataclasses import dataclass, field
import math
from typing import Any, Dict, List, Optional

import numpy as np

from . import _lib
from .constants import C0, EPS0, MU0
from .grids import coords
from .materials import Material, material_coefficients
from .monitors import FieldMonitor, SpectralMonitor, monitor_struct
from .pml import PML
from .sources import (COMPONENTS, FIELD_COMPONENTS, HardSource, SoftSource, Source,
                      TFSFSource, comp_index, point_source_struct, tfsf_struct)
from .util import geometry2bool

BC_CODES = {"periodic": 1, "dirichlet": 2, "neumann": 3}


def bc2int(bc):
    """bc2int (src/models.jl:1199-1212)."""
    key = str(bc).lstrip(":")
    if key not in BC_CODES:
        raise ValueError(
            "Wrong 'bc' value. It can be either ':periodic', ':dirichlet', or ':neumann'.")
    return BC_CODES[key]


def time_step(grid, CN):
    """time_step (src/models.jl:1181-1196)."""
    if grid.ndims == 1:
        return CN / C0 / math.sqrt(1 / grid.dz ** 2)
    if grid.ndims == 2:
        return CN / C0 / math.sqrt(1 / grid.dx ** 2 + 1 / grid.dz ** 2)
    return CN / C0 / math.sqrt(1 / grid.dx ** 2 + 1 / grid.dy ** 2 + 1 / grid.dz ** 2)


def time_grid(dt, tmax):
    """t = range(start=0, step=dt, stop=tmax): Julia's float-range length rule
    (round((stop-start)/step) + 1, minus one on overshoot); t[it] = (it-1)*dt."""
    lf = tmax / dt
    if lf < 0:
        n = 0
    elif lf == 0:
        n = 1
    else:
        n = int(round(lf)) + 1
        if 0 < tmax < (n - 1) * dt:
            n -= 1
    return np.arange(n, dtype=float) * dt


@dataclass
class ModelDescription:
    """Everything the engine needs, in FP64 / int32 host arrays."""
    ndims: int
    dims: tuple
    d: tuple
    dt: float
    Nt: int
    bcs: list
    vac: tuple
    geometry: np.ndarray
    materials: List[dict]
    layers: list
    sources: List[dict]
    monitors: List[dict] = field(default_factory=list)
    integrals: int = 0


class Model:
    """Model(grid, source; tmax, CN=0.5, bc=:periodic, pml=0, material=nothing)
    (src/models.jl:57-122)."""

    def __init__(self, grid, source, *, tmax, CN=0.5, bc="periodic", pml=0, pml_box=None,
                 material=None):
        self.grid = grid
        nd = grid.ndims
        self.dt = time_step(grid, CN)
        self.t = time_grid(self.dt, tmax)
        self.Nt = len(self.t)
        if self.Nt < 2:
            raise ValueError("tmax must cover at least two time steps")
        srcs = [source] if isinstance(source, Source) else list(source)
        self.sources = srcs
        self.source_structs = []
        for s in srcs:
            if isinstance(s, TFSFSource):
                self.source_structs.append(dict(type="tfsf", **tfsf_struct(s, grid, self.t)))
            else:
                self.source_structs.append(dict(type="point", **point_source_struct(s, grid, self.t)))
        nface = 2 * nd
        if isinstance(bc, (tuple, list)):
            if len(bc) != nface:
                raise ValueError(f"bc tuple for a {nd}-D grid needs {nface} entries")
            codes = [bc2int(b) for b in bc]
        else:
            codes = [bc2int(bc)] * nface
        bcs = [1] * 6
        if nd == 1:
            bcs[4], bcs[5] = codes
        elif nd == 2:
            bcs[0], bcs[1], bcs[4], bcs[5] = codes
        else:
            bcs = codes
        self.bcs = bcs
        if pml_box is not None:
            pml = pml_box
        self.pml = PML(grid, pml, self.dt)
        geometry = np.zeros(grid.shape, dtype=np.int32)
        if material is None:
            mats = []
        elif isinstance(material, Material):
            mats = [material]
        else:
            mats = list(material)
        for imat, mat in enumerate(mats, start=1):
            mask = geometry2bool(mat.geometry, grid)
            geometry[mask] = imat
        self.geometry = geometry
        self.materials = mats
        self.material_coefs = [material_coefficients(m, self.dt) for m in mats]
        self.Mh, self.Me, self.Md1, self.Md2 = self.dt / MU0, 1 / EPS0, 1.0, self.dt
        self.isplasma = any(c["isplasma"] for c in self.material_coefs)

    @property
    def isgeometry(self):
        return bool(np.any(self.geometry > 0))

    def describe(self, monitors=None, integrals=None):
        """ModelDescription for the engine.  integrals defaults to the reference's
        solve! behaviour: Sa always, E2 when any cell holds a material
        (src/outputs.jl:475-485)."""
        g = self.grid
        if integrals is None:
            integrals = 1 | (2 if self.isgeometry else 0)
        mons = []
        for mon in (monitors or []):
            mons.append(monitor_struct(mon, g, self.t))
        d = (g.dx if g.ndims >= 2 else 1.0, g.dy if g.ndims == 3 else 1.0, g.dz)
        return ModelDescription(
            ndims=g.ndims, dims=g.dims3, d=d, dt=self.dt, Nt=self.Nt, bcs=self.bcs,
            vac=(self.Mh, self.Me, self.Md1, self.Md2),
            geometry=np.ascontiguousarray(self.geometry.ravel(order="F"), dtype=np.int32),
            materials=self.material_coefs, layers=self.pml, sources=self.source_structs,
            monitors=mons, integrals=integrals)


def build_handle(desc, lib):
    """Issue the ABI call sequence that builds `desc` in `lib`; returns the handle."""
    md = _lib.mfd_model_desc()
    md.ndims = desc.ndims
    md.Nx, md.Ny, md.Nz = desc.dims
    md.dx, md.dy, md.dz = desc.d
    md.dt = desc.dt
    md.Nt = desc.Nt
    for i in range(6):
        md.bcs[i] = desc.bcs[i]
    md.Mh, md.Me, md.Md1, md.Md2 = desc.vac
    h = C.c_void_p()
    lib.check(lib.model_create(C.byref(md), C.byref(h)), "model_create")
    keep = []
    try:
        lib.check(lib.model_set_geometry(h, _lib.iptr(desc.geometry)), "model_set_geometry")
        for c in desc.materials:
            m = _lib.mfd_material_desc()
            m.Mh, m.Me, m.Md1, m.Md2 = c["Mh"], c["Me"], c["Md1"], c["Md2"]
            arrs = [np.ascontiguousarray(c[k], dtype=float) for k in ("Aq", "Bq", "Cq", "rate_x", "rate_y")]
            keep += arrs
            m.nq = c["nq"]
            m.Aq, m.Bq, m.Cq = (_lib.dptr(a) for a in arrs[:3])
            m.iskerr, m.Mk2, m.Mk3 = c["iskerr"], c["Mk2"], c["Mk3"]
            m.isplasma = c["isplasma"]
            for k in ("Rava", "ksi", "rho0", "Ap", "Bp", "Cp", "Ma"):
                setattr(m, k, c[k])
            m.rate_kind, m.rate_c, m.rate_k = c["rate_kind"], c["rate_c"], c["rate_k"]
            m.rate_n = len(arrs[3])
            m.rate_x, m.rate_y = _lib.dptr(arrs[3]), _lib.dptr(arrs[4])
            lib.check(lib.model_add_material(h, C.byref(m)), "model_add_material")
        for face, L in enumerate(desc.layers):
            if not L.present:
                continue
            K, A, B = (np.ascontiguousarray(v, dtype=float) for v in (L.K, L.A, L.B))
            lib.check(lib.model_set_pml(h, face, L.ib, L.Nb, _lib.dptr(K), _lib.dptr(A),
                                        _lib.dptr(B)), "model_set_pml")
        for s in desc.sources:
            if s["type"] == "point":
                cells = np.ascontiguousarray(s["cells"], dtype=np.int32)
                wave = np.ascontiguousarray(s["wave"], dtype=float)
                lib.check(lib.model_add_source(h, s["kind"], s["comp"], s["coeff"], len(cells),
                                               _lib.iptr(cells), _lib.dptr(wave)),
                          "model_add_source")
            else:
                t = _lib.mfd_tfsf_desc()
                for i in range(6):
                    t.box[i] = s["box"][i]
                t.mode = s["mode"]
                if s["mode"] == 0:
                    tabs = [np.ascontiguousarray(a, dtype=float) for a in s["tables"]]
                    keep += tabs
                    arr = (C.POINTER(C.c_double) * len(tabs))(*[_lib.dptr(a) for a in tabs])
                    keep.append(arr)
                    t.ntables = len(tabs)
                    t.tables = C.cast(arr, C.POINTER(C.POINTER(C.c_double)))
                else:
                    comp = np.ascontiguousarray(s["aux_comp"], dtype=np.int32)
                    iz = np.ascontiguousarray(s["aux_iz"], dtype=np.int32)
                    coeff = np.ascontiguousarray(s["aux_coeff"], dtype=float)
                    wave = np.ascontiguousarray(s["aux_wave"], dtype=float)
                    keep += [comp, iz, coeff, wave]
                    t.aux_n = len(comp)
                    t.aux_comp, t.aux_iz = _lib.iptr(comp), _lib.iptr(iz)
                    t.aux_coeff, t.aux_wave = _lib.dptr(coeff), _lib.dptr(wave)
                lib.check(lib.model_add_tfsf(h, C.byref(t)), "model_add_tfsf")
        for mon in desc.monitors:
            cells = np.ascontiguousarray(mon["cells"], dtype=np.int32)
            if mon["kind"] == "field":
                lib.check(lib.model_add_field_monitor(h, len(cells), _lib.iptr(cells),
                                                      int(mon["issum"])), "add_field_monitor")
            else:
                w = np.ascontiguousarray(mon["w"], dtype=float)
                lib.check(lib.model_add_spectral_monitor(h, len(cells), _lib.iptr(cells), len(w),
                                                         _lib.dptr(w), mon["tshift"],
                                                         int(mon["issum"])),
                          "add_spectral_monitor")
        if desc.integrals:
            lib.check(lib.model_enable_integrals(h, desc.integrals), "enable_integrals")
    except Exception:
        lib.model_destroy(h)
        raise
    return h


MONITOR_COMPONENTS = {1: ("Hy", "Ex"), 2: ("Hy", "Ex", "Ez"),
                      3: ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")}


class Handle:
    """A model living in one ABI implementation; read-back helpers."""

    def __init__(self, desc, lib):
        self.desc = desc
        self.lib = lib
        self.h = build_handle(desc, lib)
        self.N = int(np.prod(desc.dims))

    def step(self, n):
        self.lib.check(self.lib.model_step(self.h, int(n)), "model_step")

    @property
    def it(self):
        v = C.c_int32()
        self.lib.check(self.lib.model_get_it(self.h, C.byref(v)), "model_get_it")
        return v.value

    @property
    def status(self):
        v = C.c_int32()
        self.lib.check(self.lib.model_get_status(self.h, C.byref(v)), "model_get_status")
        return v.value

    def field(self, comp):
        out = np.empty(self.N, dtype=float)
        self.lib.check(self.lib.model_get_field(self.h, comp_index(comp), _lib.dptr(out)), "get_field")
        return out

    def set_field(self, comp, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=float).ravel(order="F"))
        self.lib.check(self.lib.model_set_field(self.h, comp_index(comp), _lib.dptr(v)), "set_field")

    def output(self, which):
        out = np.empty(self.N, dtype=float)
        self.lib.check(self.lib.model_get_output(self.h, which, _lib.dptr(out)), "get_output")
        return out

    def energy(self):
        """(W_E, W_H): total electric and magnetic field energy of the current state."""
        out = np.zeros(2)
        self.lib.check(self.lib.model_get_energy(self.h, _lib.dptr(out)), "get_energy")
        return float(out[0]), float(out[1])

    def monitor(self, imon, comp):
        mon = self.desc.monitors[imon]
        n = 1 if mon["issum"] else len(mon["cells"])
        if mon["kind"] == "field":
            out = np.empty(n * self.desc.Nt, dtype=float)
            self.lib.check(self.lib.model_get_monitor(self.h, imon, comp_index(comp), _lib.dptr(out)),
                           "get_monitor")
            return out if mon["issum"] else out.reshape((n, self.desc.Nt), order="F")
        nw = len(mon["w"])
        out = np.empty(2 * n * nw, dtype=float)
        self.lib.check(self.lib.model_get_monitor(self.h, imon, comp_index(comp), _lib.dptr(out)),
                       "get_monitor")
        z = out[0::2] + 1j * out[1::2]
        return z if mon["issum"] else z.reshape((n, nw), order="F")

    def close(self):
        if self.h is not None:
            self.lib.model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MI355X:
    """The MI355X backend: FP32 fields on one gfx950 GPU (the stand-in for GPU{T},
    src/backends.jl:13-23); MI355X(precision="f64") keeps a 1D model's state in FP64."""

    def __init__(self, precision="f32"):
        self.precision = precision

    def __repr__(self):
        return f"MI355X(precision={self.precision!r})"


@dataclass
class Fields:
    grid: Any
    data: Dict[str, np.ndarray]

    def __getattr__(self, name):
        d = self.__dict__.get("data", {})
        if name in d:
            return d[name]
        raise AttributeError(name)


@dataclass
class Solution:
    """What solve! leaves behind: fields, monitors, integrals and frames."""
    model: Model
    field: Fields
    monitors: List[dict]
    Sa: Optional[np.ndarray]
    E2: Optional[np.ndarray]
    JE: Optional[np.ndarray]
    rho_end: Optional[np.ndarray]
    frames: Dict[str, list]
    tout: np.ndarray


def output_times(t, nstride, nframes, dtout):
    """output_times (src/outputs.jl:491-518)."""
    Nt = len(t)
    if nstride is None and nframes is None and dtout is None:
        raise ValueError("One of 'nstride', 'nframes', or 'dtout' should be specified.")
    if sum(v is not None for v in (nstride, nframes, dtout)) > 1:
        raise ValueError("Only one of 'nstride', 'nframes', or 'dtout' can be specified.")
    if nstride is not None:
        if nstride >= Nt:
            nstride = Nt - 1
    elif nframes is not None:
        nstride = 1 if nframes >= Nt else int(round(Nt / nframes))
    else:
        tmax = t[-1]
        nstride = 1 if dtout >= tmax else int(round(Nt / (t[-1] / dtout)))
    return t[::nstride]


def _shape(grid, flat):
    return flat.reshape(grid.shape, order="F")


def solve(model, *, backend=None, fname=None, nframes=None, nstride=None, dtout=None,
          components=None, monitors=None, integrals=None, chunk=256, precision=None,
          tfsf_record=False, tfsf_box=None, tfsf_fname=None, lib=None):
    """solve!(model; backend, fname, nframes, nstride, dtout, components, monitors,
    tfsf_record, tfsf_box, tfsf_fname) (src/models.jl:166-242) on the MI355X engine.

    Field frames at the output times are returned in memory, and written to `fname` in
    the reference's HDF5 layout when it is given (then they are not kept in memory).
    tfsf_record=True records the incident field on the faces of tfsf_box (face
    coordinates) and writes the record file tfsf_fname (default: fname with "_tfsf"
    before the extension) that TFSFSource(fname=...) replays.  precision="f64" (1D
    models) keeps the state in FP64, as the reference's CPU{Float64} run; the default
    is the FP32 state of its GPU{Float32} run (MI355X(precision=...) sets the same).
    `lib` is the ABI implementation to run on (default: the HIP engine)."""
    import os
    from . import outputs as _out
    from . import tfsf_files as _tf
    if backend is not None and not isinstance(backend, MI355X):
        raise ValueError("this package runs on the MI355X backend only")
    if precision is None:
        precision = getattr(backend, "precision", "f32")
    if precision not in ("f32", "f64"):
        raise ValueError("precision must be 'f32' or 'f64'")
    lib = _lib.engine() if lib is None else lib
    if isinstance(monitors, (FieldMonitor, SpectralMonitor)):
        monitors = [monitors]
    desc = model.describe(monitors=monitors, integrals=integrals)
    nuser = len(desc.monitors)
    grid = model.grid
    if tfsf_record:
        if tfsf_box is None:
            raise ValueError("tfsf_record needs tfsf_box")
        if tfsf_fname is None:
            base = fname if fname is not None else "out.hdf"
            root, ext = os.path.splitext(base)
            tfsf_fname = root + "_tfsf" + ext
        desc.monitors = desc.monitors + _tf.record_monitors(grid, tfsf_box)
    h = Handle(desc, lib)
    if precision == "f64":
        if hasattr(lib, "model_set_options"):
            lib.check(lib.model_set_options(h.h, _lib.MFD_OPT_F64, 1), "set_options(F64)")
        elif not (hasattr(lib, "real_size") and lib.real_size() == 8):
            # an implementation without options computes in one fixed precision
            h.close()
            raise ValueError(f"{lib.path} has no FP64 state")
    names = MONITOR_COMPONENTS[grid.ndims]
    isfields = not (nstride is None and nframes is None and dtout is None)
    comps = _out.check_components(grid.ndims, components) if isfields else ()
    tout = output_times(model.t, nstride, nframes, dtout) if isfields else np.zeros(0)
    keep = fname is None
    frames = {c: [] for c in comps} if isfields and keep else {}
    if model.isplasma and isfields and keep:
        frames["rho"] = []
    ofile = None
    if fname is not None:
        d = os.path.dirname(fname)
        if d:
            os.makedirs(d, exist_ok=True)
        ofile = _out.OutputFile(fname, model, tout if isfields else None, comps,
                                np.float64 if precision == "f64" else np.float32)
    dt = model.dt
    try:
        itout = 0
        it = 0
        while it < model.Nt:
            # the next step whose time matches an output time (|t[it]-tout| <= dt/2)
            target = model.Nt
            if isfields and itout < len(tout):
                k = int(np.ceil(tout[itout] / dt - 0.5))
                target = min(target, max(k + 1, it + 1))
            n = min(chunk, target - it)
            h.step(n)
            it += n
            if isfields and itout < len(tout) and abs(model.t[it - 1] - tout[itout]) <= dt / 2:
                fr = {c: _shape(grid, h.field(c)) for c in comps}
                if model.isplasma:
                    fr["rho"] = _shape(grid, h.output(_lib.MFD_OUT_RHO))
                if ofile is not None:
                    ofile.write_fields(itout, fr)
                else:
                    for c, v in fr.items():
                        frames[c].append(v)
                itout += 1
        data = {c: _shape(grid, h.field(c)) for c in FIELD_COMPONENTS[grid.ndims]}
        mons = []
        for i, mon in enumerate(desc.monitors[:nuser]):
            entry = dict(kind=mon["kind"], issum=mon["issum"], inds=mon["cells"],
                         inds_1based=_out.cartesian_inds(grid, mon["cells"]))
            if mon["kind"] == "spectral":
                entry["w"] = mon["w"]
            for c in names:
                entry[c] = h.monitor(i, c)
            mons.append(entry)
        if tfsf_record:
            tables = {}
            for i, mon in enumerate(desc.monitors[nuser:], start=nuser):
                name, comp, shape = mon["tfsf"]
                rec = h.monitor(i, comp)                       # (n, Nt), x fastest
                tables[name] = np.reshape(rec, tuple(shape) + (model.Nt,), order="F")
            _tf.write_record_file(tfsf_fname, grid, model.t, tfsf_box, tables)
        Sa = _shape(grid, h.output(0)) if desc.integrals & 1 else None
        E2 = _shape(grid, h.output(1)) if desc.integrals & 2 else None
        JE = _shape(grid, h.output(2)) if model.isplasma else None
        rho = _shape(grid, h.output(3)) if model.isplasma else None
        if ofile is not None:
            ofile.finish(Sa if Sa is not None else np.zeros(grid.shape),
                         E2 if E2 is not None else np.zeros(grid.shape), JE, rho, mons)
            ofile = None
    finally:
        if ofile is not None:
            ofile.close()
        h.close()
    return Solution(model, Fields(grid, data), mons, Sa, E2, JE, rho, frames, tout)


def _lib_rho():
    return 3
