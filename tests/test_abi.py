"""The C-ABI boundary without a GPU: the engine library loads and exports every entry
point include/maxwell_fdtd.h declares, the oracle exports the same model interface
under its own prefix, the ctypes structs match the C layout, and host-side argument
checks reject bad input before any device work.  No compute call runs here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from maxwellfdtd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "maxwell_fdtd.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mfd_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_model_interface():
    names = declared()
    for base in _lib.SIGNATURES:
        assert "mfd_" + base in names
    for base in _lib.ENGINE_EXTRA:
        assert "mfd_" + base in names


@pytest.fixture(scope="module")
def engine_so():
    if not os.path.exists(_lib.ENGINE_LIB):
        pytest.fail("engine library not built: run __graft_entry__.build()")
    return C.CDLL(_lib.ENGINE_LIB)


def test_engine_exports_every_declared_symbol(engine_so):
    missing = [n for n in declared() if not hasattr(engine_so, n)]
    assert not missing, missing


def test_engine_binding_and_version(engine_so):
    lib = _lib.engine()
    assert lib.version() == 1
    assert isinstance(lib.error(), str)


def test_oracle_exports_model_interface():
    from oracle_util import oracle
    for prec in ("f64", "f32"):
        orc = oracle(prec)
        for base in _lib.SIGNATURES:
            assert hasattr(orc.lib, "ofd_" + base)


def _c_layout():
    """sizeof/offsetof of the ABI structs as the C compiler lays them out."""
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "maxwell_fdtd.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(mfd_model_desc), offsetof(mfd_model_desc, dt),
         offsetof(mfd_model_desc, bcs), offsetof(mfd_model_desc, Mh));
  printf("%zu %zu %zu %zu\n", sizeof(mfd_material_desc), offsetof(mfd_material_desc, Aq),
         offsetof(mfd_material_desc, rate_kind), offsetof(mfd_material_desc, rate_y));
  printf("%zu %zu %zu %zu\n", sizeof(mfd_tfsf_desc), offsetof(mfd_tfsf_desc, tables),
         offsetof(mfd_tfsf_desc, aux_comp), offsetof(mfd_tfsf_desc, aux_wave));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return [list(map(int, line.split())) for line in out.strip().splitlines()]


def test_ctypes_structs_match_c_layout():
    md, mat, tf = _c_layout()
    S = _lib
    assert md == [C.sizeof(S.mfd_model_desc), S.mfd_model_desc.dt.offset,
                  S.mfd_model_desc.bcs.offset, S.mfd_model_desc.Mh.offset]
    assert mat == [C.sizeof(S.mfd_material_desc), S.mfd_material_desc.Aq.offset,
                   S.mfd_material_desc.rate_kind.offset, S.mfd_material_desc.rate_y.offset]
    assert tf == [C.sizeof(S.mfd_tfsf_desc), S.mfd_tfsf_desc.tables.offset,
                  S.mfd_tfsf_desc.aux_comp.offset, S.mfd_tfsf_desc.aux_wave.offset]


def _desc(**kw):
    d = _lib.mfd_model_desc()
    d.ndims, d.Nx, d.Ny, d.Nz = 3, 8, 8, 8
    d.dx = d.dy = d.dz = 1e-8
    d.dt, d.Nt = 1e-17, 10
    for i in range(6):
        d.bcs[i] = 1
    d.Mh, d.Me, d.Md1, d.Md2 = 1.0, 1.0, 1.0, 1e-17
    for k, v in kw.items():
        if k == "bcs":
            for i in range(6):
                d.bcs[i] = v[i]
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("bad", [dict(ndims=4), dict(Nz=1), dict(ndims=1),
                                 dict(ndims=2), dict(Nt=0), dict(dt=0.0), dict(Nz=70000),
                                 dict(Nx=2048, Ny=2048, Nz=512)])
def test_engine_rejects_bad_descriptor(engine_so, bad):
    lib = _lib.engine()
    h = C.c_void_p()
    st = lib.model_create(C.byref(_desc(**bad)), C.byref(h))
    assert st == _lib.MFD_ERR_ARG
    assert "descriptor" in lib.error()


def test_hbm_probes_check_arguments(engine_so):
    """The copy probes refuse bad sizes, counts and a missing output before any device
    call, so these run without a GPU."""
    lib = _lib.probe()
    ms = C.c_double(0.0)
    assert lib.probe_copy(512, 1, C.byref(ms)) == _lib.MFD_ERR_ARG
    assert lib.probe_copy_fast(1 << 20, 0, C.byref(ms)) == _lib.MFD_ERR_ARG
    assert lib.probe_copy(1 << 20, 1, None) == _lib.MFD_ERR_ARG
    assert lib.probe_multi(1 << 20, 2, 3, 16, 1, C.byref(ms)) == _lib.MFD_ERR_ARG
    assert lib.probe_multi(1 << 20, 7, 7, 16, 1, C.byref(ms)) == _lib.MFD_ERR_ARG
    assert b"nin == nout" in lib.probe_last_error()
    assert lib.probe_multi(1 << 20, 2, 2, 16, 1, None) == _lib.MFD_ERR_ARG
    engine = _lib.engine()
    assert not hasattr(engine.lib, "mfd_stream_copy")


def test_engine_rejects_bad_bc(engine_so):
    lib = _lib.engine()
    h = C.c_void_p()
    st = lib.model_create(C.byref(_desc(bcs=[1, 1, 4, 1, 1, 1])), C.byref(h))
    assert st == _lib.MFD_ERR_ARG
    assert "Wrong 'bc' value" in lib.error()


def test_engine_null_arguments(engine_so):
    lib = _lib.engine()
    assert lib.model_create(None, None) == _lib.MFD_ERR_ARG
    assert lib.model_step(None, 1) == _lib.MFD_ERR_ARG
    assert lib.model_destroy(None) in (_lib.MFD_OK, _lib.MFD_ERR_ARG)


def test_engine_without_gpu_fails_loudly(engine_so):
    """On a host without a gfx950 device a valid model is refused with MFD_ERR_NODEV;
    nothing falls back to a CPU path."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.engine()
    h = C.c_void_p()
    st = lib.model_create(C.byref(_desc()), C.byref(h))
    assert st == _lib.MFD_ERR_NODEV
    assert lib.error()


def test_oracle_rejects_like_engine():
    from oracle_util import oracle
    orc = oracle("f64")
    h = C.c_void_p()
    assert orc.model_create(C.byref(_desc(ndims=4)), C.byref(h)) == _lib.MFD_ERR_ARG
    assert orc.model_create(C.byref(_desc(bcs=[1, 1, 4, 1, 1, 1])), C.byref(h)) == _lib.MFD_ERR_ARG
    assert "Wrong 'bc' value" in orc.error()


def test_batch_handle_and_option_checks(engine_so):
    """Batch handles and their option checks are host-side: they work without a GPU."""
    lib = _lib.engine()
    assert lib.batch_create(None) == _lib.MFD_ERR_ARG
    b = C.c_void_p()
    assert lib.batch_create(C.byref(b)) == _lib.MFD_OK
    try:
        assert lib.batch_add(b, None) == _lib.MFD_ERR_ARG
        assert lib.batch_set_options(b, _lib.MFD_OPT_ZCHUNK, 0) == _lib.MFD_ERR_ARG
        assert lib.batch_set_options(b, _lib.MFD_OPT_ROWS, 3) == _lib.MFD_ERR_ARG
        assert "rows per workgroup" in lib.error()
        assert lib.batch_set_options(b, _lib.MFD_OPT_DFT_BLOCK, 16) == _lib.MFD_ERR_ARG
        assert "per-model option" in lib.error()
        assert lib.batch_set_options(b, _lib.MFD_OPT_ZCHUNK, 32) == _lib.MFD_OK
        assert lib.batch_set_options(b, _lib.MFD_OPT_PML_SPLIT, 1) == _lib.MFD_OK
        assert lib.batch_set_options(b, _lib.MFD_OPT_SCHED, 1) == _lib.MFD_ERR_ARG   # server only
        assert lib.batch_set_options(b, 99, 1) == _lib.MFD_ERR_ARG
        n = C.c_int32(-1)
        assert lib.batch_active(b, C.byref(n)) == _lib.MFD_OK and n.value == 0
        assert lib.batch_autotune(b, 4, None, None, None, None) == _lib.MFD_ERR_STATE
    finally:
        assert lib.batch_destroy(b) == _lib.MFD_OK


def test_server_needs_a_gpu(engine_so):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.engine()
    s = C.c_void_p()
    assert lib.server_create(0, 4, None) == _lib.MFD_ERR_ARG
    assert lib.server_create(0, 4, C.byref(s)) != _lib.MFD_OK
    assert lib.error()


def test_header_constants_match_binding():
    """Every MFD_* enumerator the header declares has the same value in the Python
    binding (the ctypes twin of the Julia shim)."""
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "maxwell_fdtd.h")).read(),
                  flags=re.S)
    found = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        nxt = 0
        for item in body.split(","):
            item = item.strip()
            if not item:
                continue
            name, _, val = item.partition("=")
            nxt = int(val) if val.strip() else nxt
            found[name.strip()] = nxt
            nxt += 1
    assert len(found) > 50
    missing = [k for k in found if not hasattr(_lib, k)]
    assert not missing, missing
    wrong = {k: (v, getattr(_lib, k)) for k, v in found.items() if getattr(_lib, k) != v}
    assert not wrong, wrong
