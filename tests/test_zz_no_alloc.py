"""The allocation-free energy entry point (mfd_model_get_energy_noalloc) against the
allocating one (mfd_model_get_energy), called through the C ABI: the same kernels with
the same launch shape, so the FP64 sums are bitwise equal.  Two calls in a row, the
second on a smaller model, so the second reuses the device scratch array that the first
filled with more partial sums.  The caller's output buffer and stream come from the HIP
runtime the engine itself links (loaded here through ctypes), as a caller's would."""
import ctypes as C

import numpy as np
import pytest

import cases
from maxwellfdtd_amd import _lib
from maxwellfdtd_amd.models import Handle

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    fn = lib.lib.mfd_model_get_energy_noalloc
    fn.restype = C.c_int32
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return fn


class Hip:
    """The few HIP runtime calls a caller makes: device memory, a stream, copies."""

    def __init__(self):
        _lib.engine()                       # the engine has loaded its HIP runtime
        self.rt = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipMemset", [C.c_void_p, C.c_int, C.c_size_t]),
                           ("hipStreamCreateWithFlags", [C.POINTER(C.c_void_p), C.c_uint]),
                           ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipStreamDestroy", [C.c_void_p]),
                           ("hipDeviceSynchronize", [])):
            f = getattr(self.rt, name)
            f.restype = C.c_int
            f.argtypes = args

    def ok(self, err, what):
        assert err == 0, f"{what} failed with HIP error {err}"

    def malloc(self, n):
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), n), "hipMalloc")
        return p

    def to_host(self, p, n):
        out = np.empty(n, dtype=np.float64)
        self.ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8 * n, 2), "hipMemcpy")
        return out


def test_energy_noalloc_equals_allocating_entry_point():
    lib = _lib.engine()
    fn = _noalloc(lib)
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreateWithFlags(C.byref(stream), 1), "hipStreamCreateWithFlags")
    d_out = hip.malloc(16)
    big, _ = cases.spheres_3d(n=40, steps=60)
    small, _ = cases.spheres_3d(n=20, steps=40, seed=3)
    try:
        for k, model in enumerate((big, small)):
            h = Handle(model.describe(), lib)
            h.step(model.Nt)
            ref = np.array(h.energy())
            assert ref[0] > 0 and ref[1] > 0
            hip.ok(hip.rt.hipMemset(d_out, 0xFF, 16), "hipMemset")   # NaN until written
            if k == 0:      # on the caller's stream
                lib.check(fn(h.h, d_out, stream), "energy_noalloc")
                hip.ok(hip.rt.hipStreamSynchronize(stream), "hipStreamSynchronize")
            else:           # on the model's own copy stream
                lib.check(fn(h.h, d_out, None), "energy_noalloc")
                hip.ok(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
            got = hip.to_host(d_out, 2)
            assert np.array_equal(got, ref), (k, got, ref)
            h.close()
    finally:
        hip.rt.hipFree(d_out)
        hip.rt.hipStreamDestroy(stream)


def test_energy_noalloc_rejects_a_model_not_on_the_device():
    lib = _lib.engine()
    fn = _noalloc(lib)
    hip = Hip()
    d_out = hip.malloc(16)
    model, _ = cases.spheres_3d(n=12, steps=10, monitors=False, tfsf=False, pml=2)
    h = Handle(model.describe(), lib)
    try:
        assert fn(h.h, d_out, None) == _lib.MFD_ERR_STATE
        assert "uploaded" in lib.error()
        assert fn(h.h, None, None) == _lib.MFD_ERR_ARG
    finally:
        h.close()
        hip.rt.hipFree(d_out)
