"""ctypes binding of the C ABI in include/maxwell_fdtd.h.

`Binding(path, prefix)` wraps any shared library that exports the ABI's functions
under `prefix` ("mfd_" for the engine).  `engine()` loads the in-tree HIP library and
raises if it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_LIB = os.path.join(HERE, "libmaxwellfdtd_amd.so")

MFD_OK, MFD_ERR_ARG, MFD_ERR_STATE, MFD_ERR_HIP = 0, -1, -2, -3
MFD_ERR_NOMEM, MFD_ERR_NAN, MFD_ERR_INF, MFD_ERR_NODEV = -4, -5, -6, -7
# field components, boundary conditions, CPML faces, sources, ionisation rates, integrals
MFD_HX, MFD_HY, MFD_HZ, MFD_DX, MFD_DY, MFD_DZ, MFD_EX, MFD_EY, MFD_EZ = range(9)
MFD_BC_PERIODIC, MFD_BC_DIRICHLET, MFD_BC_NEUMANN = 1, 2, 3
MFD_FACE_XL, MFD_FACE_XR, MFD_FACE_YL, MFD_FACE_YR, MFD_FACE_ZL, MFD_FACE_ZR = range(6)
MFD_SRC_SOFT, MFD_SRC_HARD = 0, 1
MFD_TFSF_TABLES, MFD_TFSF_AUX1D = 0, 1
MFD_RATE_NONE, MFD_RATE_POWER, MFD_RATE_LOGLOG = 0, 1, 2
MFD_INT_SA, MFD_INT_E2 = 1, 2
MFD_OUT_SA, MFD_OUT_E2, MFD_OUT_JE, MFD_OUT_RHO = 0, 1, 2, 3
NAN_MSG = "Something went wrong. I found NaN field values."
INF_MSG = "Something went wrong. I found Inf field values."


class mfd_model_desc(C.Structure):
    _fields_ = [("ndims", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("Nt", C.c_int32), ("bcs", C.c_int32 * 6),
                ("Mh", C.c_double), ("Me", C.c_double), ("Md1", C.c_double),
                ("Md2", C.c_double)]


class mfd_material_desc(C.Structure):
    _fields_ = [("Mh", C.c_double), ("Me", C.c_double), ("Md1", C.c_double),
                ("Md2", C.c_double), ("nq", C.c_int32),
                ("Aq", C.POINTER(C.c_double)), ("Bq", C.POINTER(C.c_double)),
                ("Cq", C.POINTER(C.c_double)), ("iskerr", C.c_int32),
                ("Mk2", C.c_double), ("Mk3", C.c_double), ("isplasma", C.c_int32),
                ("Rava", C.c_double), ("ksi", C.c_double), ("rho0", C.c_double),
                ("Ap", C.c_double), ("Bp", C.c_double), ("Cp", C.c_double),
                ("Ma", C.c_double), ("rate_kind", C.c_int32), ("rate_c", C.c_double),
                ("rate_k", C.c_double), ("rate_n", C.c_int32),
                ("rate_x", C.POINTER(C.c_double)), ("rate_y", C.POINTER(C.c_double))]


class mfd_tfsf_desc(C.Structure):
    _fields_ = [("box", C.c_int32 * 6), ("mode", C.c_int32), ("ntables", C.c_int32),
                ("tables", C.POINTER(C.POINTER(C.c_double))), ("aux_n", C.c_int32),
                ("aux_comp", C.POINTER(C.c_int32)), ("aux_iz", C.POINTER(C.c_int32)),
                ("aux_coeff", C.POINTER(C.c_double)), ("aux_wave", C.POINTER(C.c_double))]


P = C.c_void_p
I32 = C.c_int32
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "version": (I32, []),
    "last_error": (C.c_char_p, []),
    "model_create": (I32, [C.POINTER(mfd_model_desc), C.POINTER(P)]),
    "model_set_geometry": (I32, [P, IP]),
    "model_add_material": (I32, [P, C.POINTER(mfd_material_desc)]),
    "model_set_pml": (I32, [P, I32, I32, I32, DP, DP, DP]),
    "model_add_source": (I32, [P, I32, I32, C.c_double, I32, IP, DP]),
    "model_add_tfsf": (I32, [P, C.POINTER(mfd_tfsf_desc)]),
    "model_add_field_monitor": (I32, [P, I32, IP, I32]),
    "model_add_spectral_monitor": (I32, [P, I32, IP, I32, DP, C.c_double, I32]),
    "model_enable_integrals": (I32, [P, I32]),
    "model_footprint": (I32, [P, C.POINTER(C.c_int64)]),
    "model_upload": (I32, [P]),
    "model_step": (I32, [P, I32]),
    "model_get_it": (I32, [P, IP]),
    "model_get_status": (I32, [P, IP]),
    "model_get_field": (I32, [P, I32, DP]),
    "model_set_field": (I32, [P, I32, DP]),
    "model_get_monitor": (I32, [P, I32, I32, DP]),
    "model_get_output": (I32, [P, I32, DP]),
    "model_destroy": (I32, [P]),
}


def dptr(a):
    return a.ctypes.data_as(DP)


def iptr(a):
    return a.ctypes.data_as(IP)


class EngineError(RuntimeError):
    pass


class Binding:
    """Callable wrapper of one library implementing the ABI under `prefix`."""

    def __init__(self, path, prefix, extra=None):
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        sigs = dict(SIGNATURES)
        sigs.update(extra or {})
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def error(self):
        msg = self.last_error()
        return msg.decode() if msg else ""

    def check(self, status, what=""):
        if status == MFD_OK:
            return
        if status == MFD_ERR_NAN:
            raise EngineError(NAN_MSG)
        if status == MFD_ERR_INF:
            raise EngineError(INF_MSG)
        raise EngineError(f"{self.prefix}{what} failed ({status}): {self.error()}")


class mfd_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_int32), ("queued", C.c_int32), ("active", C.c_int32),
                ("done", C.c_int32), ("failed", C.c_int32), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("cell_updates", C.c_double),
                ("busy_s", C.c_double), ("gcells_per_s", C.c_double), ("hbm_in_use", C.c_int64)]


MFD_OPT_GRAPH_STEPS, MFD_OPT_ZCHUNK, MFD_OPT_PERSISTENT, MFD_OPT_KERNEL_TIMING = 0, 1, 2, 3
MFD_OPT_DFT_BLOCK, MFD_OPT_ROWS, MFD_OPT_KERNEL_FORM, MFD_OPT_STAGGER, MFD_OPT_QUANTUM = 4, 5, 6, 7, 8
MFD_OPT_FILL_MCELLS = 9
MFD_OPT_SNAKE, MFD_OPT_DLESS, MFD_OPT_PML_SPLIT, MFD_OPT_SCHED, MFD_OPT_ZCHUNK_H = 10, 11, 12, 13, 14
MFD_OPT_XCD_H, MFD_OPT_XCD_BLOCK, MFD_OPT_F64, MFD_OPT_TFSF_FOLD, MFD_OPT_FUSED_2D = 15, 16, 17, 18, 19
MFD_OPT_RUN1D = 20
MFD_REQ_QUEUED, MFD_REQ_RUNNING, MFD_REQ_DONE, MFD_REQ_FAILED = 0, 1, 2, 3

ENGINE_EXTRA = {
    "device_info": (I32, [C.POINTER(C.c_int64), C.POINTER(C.c_int64), IP, C.c_char_p, I32]),
    "model_set_options": (I32, [P, I32, I32]),
    "model_timing": (I32, [P, I32, DP, IP]),
    "model_get_energy": (I32, [P, DP]),
    "model_get_energy_noalloc": (I32, [P, P, P]),
    "batch_create": (I32, [C.POINTER(P)]),
    "batch_add": (I32, [P, P]),
    "batch_set_options": (I32, [P, I32, I32]),
    "batch_run": (I32, [P, I32]),
    "batch_autotune": (I32, [P, I32, IP, IP, IP, DP]),
    "batch_active": (I32, [P, IP]),
    "batch_destroy": (I32, [P]),
    "server_create": (I32, [C.c_int64, I32, C.POINTER(P)]),
    "server_set_options": (I32, [P, I32, I32]),
    "server_submit": (I32, [P, P, C.POINTER(C.c_int64)]),
    "server_poll": (I32, [P, C.c_int64, IP, IP, DP]),
    "server_wait": (I32, [P, C.c_int64, I32, IP]),
    "server_get_stats": (I32, [P, C.POINTER(mfd_server_stats)]),
    "server_destroy": (I32, [P]),
}

_engine = None


def engine():
    """The HIP engine library; raises if it has not been built."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise EngineError(
                f"{ENGINE_LIB} is missing: build the HIP engine first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        _engine = Binding(ENGINE_LIB, "mfd_", ENGINE_EXTRA)
    return _engine


PROBE_LIB = os.path.join(os.path.dirname(ENGINE_LIB), "libmfd_hbm_probe.so")
_probe = None


class _Probe:
    """ctypes binding of the HBM probe library (bandwidth measurements)."""

    SIGS = {
        "probe_last_error": (C.c_char_p, []),
        "probe_copy": (I32, [C.c_int64, I32, DP]),
        "probe_copy_fast": (I32, [C.c_int64, I32, DP]),
        "probe_multi": (I32, [C.c_int64, I32, I32, I32, I32, DP]),
    }

    def __init__(self, path):
        self.path = path
        self.lib = C.CDLL(path)
        for name, (res, args) in self.SIGS.items():
            fn = getattr(self.lib, "mfd_" + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def check(self, status, what=""):
        if status != MFD_OK:
            msg = self.probe_last_error()
            raise EngineError(f"mfd_{what} failed ({status}): {msg.decode() if msg else ''}")


def probe():
    """The HBM probe library; raises if it has not been built."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_LIB):
            raise EngineError(f"{PROBE_LIB} is missing: build first "
                              "(python -c 'import __graft_entry__ as g; g.build()')")
        _probe = _Probe(PROBE_LIB)
    return _probe
