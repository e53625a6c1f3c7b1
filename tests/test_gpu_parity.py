"""GPU parity: the MI355X engine against the FP64 oracle on identical inputs, through the
same C-ABI call sequence.  Bar: relative L2 <= 1e-4 on the H, D and E vector fields, on
monitors, on DFT spectra and on integral outputs (the FP32-vs-FP64 tolerance of the
per-step update), plus the reference's own analytic checks."""
import ctypes

import numpy as np
import pytest

import cases
from maxwellfdtd_amd import EPS0, _lib, solve
from maxwellfdtd_amd.models import Handle, MONITOR_COMPONENTS
from oracle_util import field_energy, field_errors, oracle, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def run_pair(model, mons=None, integrals=None, steps=None, options=None):
    desc = model.describe(monitors=mons, integrals=integrals)
    n = model.Nt if steps is None else steps
    hg = Handle(desc, _lib.engine())
    for k, v in (options or {}).items():
        hg.lib.check(hg.lib.model_set_options(hg.h, k, v), "set_options")
    hg.step(n)
    ho = Handle(desc, oracle("f64"))
    ho.step(n)
    return hg, ho, desc


def assert_parity(hg, ho, desc, model, tol=TOL, out_tol=None):
    """Fields, integrals, plasma outputs and monitors within tol (out_tol for the monitor
    samples of an FP64 model, which are stored in FP32)."""
    out_tol = tol if out_tol is None else out_tol
    errs = field_errors(hg, ho, model.grid.ndims)
    assert max(errs.values()) <= tol, errs
    if desc.integrals & 1:
        assert rel_l2(hg.output(0), ho.output(0)) <= tol
    if desc.integrals & 2:
        assert rel_l2(hg.output(1), ho.output(1)) <= tol
    if model.isplasma:
        assert rel_l2(hg.output(2), ho.output(2)) <= tol * 10, "JE"
        assert rel_l2(hg.output(3), ho.output(3)) <= tol, "rho"
    for i, mon in enumerate(desc.monitors):
        names = MONITOR_COMPONENTS[model.grid.ndims]
        a = np.concatenate([np.ravel(hg.monitor(i, c)) for c in names])
        b = np.concatenate([np.ravel(ho.monitor(i, c)) for c in names])
        assert rel_l2(a, b) <= out_tol, (i, mon["kind"], rel_l2(a, b))
    return errs


def test_engine_loads():
    lib = _lib.engine()
    assert lib.version() == 1


def test_ref_1d_propagation():
    model, Eth, rtol = cases.ref_1d_test()
    hg, ho, desc = run_pair(model)
    assert rel_l2(hg.field("Ex"), Eth) <= rtol
    assert_parity(hg, ho, desc, model)


def test_ref_2d_propagation():
    model, Eth, rtol = cases.ref_2d_test()
    hg, ho, desc = run_pair(model)
    assert rel_l2(hg.field("Ex"), Eth.ravel(order="F")) <= rtol
    assert_parity(hg, ho, desc, model)


def test_ref_3d_propagation():
    model, Eth, rtol = cases.ref_3d_test()
    hg, ho, desc = run_pair(model)
    assert rel_l2(hg.field("Ex"), Eth.ravel(order="F")) <= rtol
    assert_parity(hg, ho, desc, model)


@pytest.mark.parametrize("ndims", [1, 2, 3])
@pytest.mark.parametrize("kind", ["sigma", "lorentz"])
def test_ref_losses(ndims, kind):
    model, ki = cases.ref_losses(ndims, kind)
    sol = solve(model)
    assert cases.beer_lambert_ok(sol.Sa, model.grid.z, ki)


def test_spheres_3d_all_features():
    model, mons = cases.spheres_3d()
    hg, ho, desc = run_pair(model, mons)
    assert_parity(hg, ho, desc, model)


def test_cpml_free_tiles():
    """A grid long in x (160 cells) has 64-cell tiles that no CPML layer reaches; with the
    CPML split they run the slab-free kernels.  The split changes which code serves a
    cell, not the arithmetic it states: the oracle bar holds, and the unsplit run agrees
    to rounding.  (Bitwise equality is not promised: the slab-free instantiation is
    compiled on its own, and its FMA contraction can differ in the last bit.)"""
    model, mons = cases.spheres_3d(n=28, nx=160, steps=100)
    hg, ho, desc = run_pair(model, mons, options={_lib.MFD_OPT_PML_SPLIT: 1})
    assert_parity(hg, ho, desc, model)
    whole = Handle(desc, _lib.engine())
    whole.step(desc.Nt)
    for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez", "Dx"):
        assert rel_l2(hg.field(c), whole.field(c)) <= 1e-5, c
    assert rel_l2(hg.output(0), whole.output(0)) <= 1e-5
    # the two launches side by side on two streams (graph replays and eager steps) run
    # the same kernels on the same tiles: bit for bit the serial split
    for gsteps in (16, 0):
        side = Handle(desc, _lib.engine())
        side.lib.check(side.lib.model_set_options(side.h, _lib.MFD_OPT_PML_SPLIT, 2), "set_options")
        side.lib.check(side.lib.model_set_options(side.h, _lib.MFD_OPT_GRAPH_STEPS, gsteps), "set_options")
        side.step(desc.Nt)
        for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez", "Dx"):
            assert np.array_equal(hg.field(c), side.field(c)), (gsteps, c)
        assert np.array_equal(hg.output(0), side.output(0))
        side.close()


def test_spheres_3d_boundary_conditions():
    bc = ("dirichlet", "neumann", "neumann", "periodic", "periodic", "dirichlet")
    model, mons = cases.spheres_3d(n=24, steps=80, bc=bc, monitors=False, pml=0)
    hg, ho, desc = run_pair(model)
    assert_parity(hg, ho, desc, model)


def test_dispersive_3d():
    model = cases.dispersive_3d()
    hg, ho, desc = run_pair(model)
    assert_parity(hg, ho, desc, model)


@pytest.mark.parametrize("rate", ["power", "table"])
def test_nonlinear_3d(rate):
    model = cases.nonlinear_3d(rate=rate)
    hg, ho, desc = run_pair(model)
    assert_parity(hg, ho, desc, model)


@pytest.mark.parametrize("fused", [0, 1])
def test_nonlinear_2d(fused):
    """2D Kerr + plasma and Kerr-only cells keep their D (the E kernel loads it once the
    id is known); the plain cells around them run without it.  Two-pass kernels and the
    fused 2D step."""
    model = cases.nonlinear_2d()
    hg, ho, desc = run_pair(model, options={_lib.MFD_OPT_FUSED_2D: fused})
    assert_parity(hg, ho, desc, model)
    assert np.max(ho.output(3)) > 0          # the pulse ionised the medium


@pytest.mark.parametrize("fused", [0, 1])
def test_lorentz_2d(fused):
    model, mons = cases.lorentz_2d()
    hg, ho, desc = run_pair(model, mons, options={_lib.MFD_OPT_FUSED_2D: fused})
    assert_parity(hg, ho, desc, model)


def test_slab_1d():
    model, mons = cases.slab_1d()
    hg, ho, desc = run_pair(model, mons)
    assert_parity(hg, ho, desc, model)


@pytest.mark.parametrize("kmax", [3.0, 5.0])
@pytest.mark.parametrize("ndims,kform", [(1, None), (2, None), (3, 3), (3, 2), (3, 0)])
def test_cpml_kappa(ndims, kform, kmax):
    """CPML with a kappa profile (kmax > 1): the slab updates divide the E/H differences by
    K after the psi recursion (src/models.jl:812-871, src/pml.jl:68-81).  With kmax = 1 the
    division is the identity, so only these cases exercise it on the GPU."""
    opts = None
    if ndims == 3:
        model, mons = cases.spheres_3d(n=28, steps=100, kmax=kmax)
        opts = {_lib.MFD_OPT_KERNEL_FORM: kform}
    elif ndims == 2:
        model, mons = cases.lorentz_2d(steps=160, kmax=kmax)
    else:
        # 550 steps: the left-going half of the pulse has crossed the left layer and its
        # reflection has passed the monitors.  (Much later, once the pulse has left, the
        # point monitor holds a small residual whose FP32-vs-FP64 difference alone, in the
        # oracle's own FP32 build, exceeds 1e-4.)
        model, mons = cases.slab_1d(steps=550, kmax=kmax)
    present = [L for L in model.pml if L.present]
    assert present and all(np.isclose(L.K.max(), kmax) for L in present)
    hg, ho, desc = run_pair(model, mons, options=opts)
    assert_parity(hg, ho, desc, model)


@pytest.mark.parametrize("ndims,kform", [(1, None), (2, None), (3, 3), (3, 2), (3, 0)])
def test_tfsf_replay_tables(ndims, kform):
    """TFSF from replay tables (MFD_TFSF_TABLES), the form in which the reference's
    TFSFSource reaches add_source! (src/sources.jl:269-346,400-497,565-776): 4 tables in
    1D, 8 in 2D, 24 in 3D, each [face..., Nt], around a dielectric scatterer."""
    model = cases.tfsf_tables_model(ndims)
    assert model.source_structs[0]["mode"] == _lib.MFD_TFSF_TABLES
    opts = {_lib.MFD_OPT_KERNEL_FORM: kform} if kform is not None else None
    hg, ho, desc = run_pair(model, options=opts)
    assert_parity(hg, ho, desc, model)
    assert np.max(np.abs(ho.field("Ex"))) > 1e-2


@pytest.mark.parametrize("kform", [3, 2, 0])
def test_tfsf_faces_through_materials(kform):
    """TFSF as the only source (one gather launch per step; with the pipelined kernels
    the E kernel finishes x-face cells itself) with face cells holding stateless,
    ADE, Debye and Kerr materials."""
    model, mons = cases.tfsf_faces_3d()
    hg, ho, desc = run_pair(model, mons, options={6: kform})
    assert_parity(hg, ho, desc, model)
    assert np.max(np.abs(ho.field("Ex"))) > 1e6       # the pulse is in the box


@pytest.mark.parametrize("case", ["plane_wave", "plane_wave_monitors", "tables", "kerr_faces", "stateful_faces"])
def test_tfsf_fold_is_bitwise_neutral(case):
    """With MFD_OPT_TFSF_FOLD = 1 (the default) the fused step stores the corrected H of
    the x-face cells it finishes, and the face launch after it walks only the other
    touched cells (y and z faces, edges, cells with ADE state); with 0 the face launch
    also writes those H corrections.  The two give the same bits, with 2 and 4 rows per
    lane and z chunks that put the z faces on tile edges, for plane-wave and table
    sources, Kerr cells on a face, and stateful cells on the faces.  The first run also
    matches the FP64 oracle."""
    if case.startswith("plane_wave"):
        model, mons = cases.spheres_3d(n=38, nx=70, steps=70, soft=False, nspheres=6)
        mons = mons if case == "plane_wave_monitors" else None
    elif case == "tables":
        model, mons = cases.tfsf_tables_model(3), None
    else:
        model, mons = cases.tfsf_faces_3d(n=36, steps=90, lo=8, hi=26, stateful=case == "stateful_faces")
    desc = model.describe(monitors=mons, integrals=3)
    outs = []
    for fold, rows, zc in ((1, 4, 16), (0, 4, 16), (1, 2, 7), (0, 2, 9)):
        h = Handle(desc, _lib.engine())
        for opt, v in ((_lib.MFD_OPT_KERNEL_FORM, 3), (_lib.MFD_OPT_TFSF_FOLD, fold), (_lib.MFD_OPT_ROWS, rows),
                       (_lib.MFD_OPT_ZCHUNK, zc)):
            h.lib.check(h.lib.model_set_options(h.h, opt, v), "set_options")
        h.step(model.Nt)
        outs.append([h.field(c) for c in ("Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ex", "Ey", "Ez")]
                    + [h.output(0), h.output(1)]
                    + [np.ravel(h.monitor(q, c)) for q in range(len(mons or [])) for c in ("Ex", "Hy")])
        if len(outs) == 1:
            ho = Handle(desc, oracle("f64"))
            ho.step(model.Nt)
            assert_parity(h, ho, desc, model)
            assert np.max(np.abs(ho.field("Ex"))) > 0
        h.close()
    for o in outs[1:]:
        for k, (a, b) in enumerate(zip(outs[0], o)):
            assert np.array_equal(a, b), k


def test_plain_tfsf_without_stored_d():
    """TFSF as the only source with plain dielectrics: the E kernel keeps no D for plain
    cells (D = E/Me, rebuilt on read).  Parity against the oracle on H, D and E, the same
    run with stored D agrees to rounding, and reading D in the middle of a run (which
    rebuilds it) does not change the result."""
    model, mons = cases.spheres_3d(n=40, steps=140, soft=False, nspheres=6)
    hg, ho, desc = run_pair(model, mons)
    assert_parity(hg, ho, desc, model)
    full = Handle(desc, _lib.engine())
    full.lib.check(full.lib.model_set_options(full.h, 11, 0), "set_options")
    full.step(desc.Nt)
    mid = Handle(desc, _lib.engine())
    mid.step(desc.Nt // 2)
    mid.field("Dx")
    mid.step(desc.Nt - desc.Nt // 2)
    for c in ("Hx", "Dy", "Ez", "Dz"):
        assert rel_l2(hg.field(c), full.field(c)) <= 1e-5, c
        assert np.array_equal(hg.field(c), mid.field(c)), c


def test_2d_without_stored_d():
    """2D keeps no D for plain cells that no source writes (D = E/Me, rebuilt on read);
    Lorentz rods, TFSF and soft-source cells keep theirs.  Parity against the oracle on
    H, D and E, the run with stored D agrees to rounding, and reading D mid-run does not
    change the result."""
    model, mons = cases.lorentz_2d()
    hg, ho, desc = run_pair(model, mons)
    assert_parity(hg, ho, desc, model)
    full = Handle(desc, _lib.engine())
    full.lib.check(full.lib.model_set_options(full.h, _lib.MFD_OPT_DLESS, 0), "set_options")
    full.step(desc.Nt)
    mid = Handle(desc, _lib.engine())
    mid.step(desc.Nt // 2)
    mid.field("Dz")
    mid.step(desc.Nt - desc.Nt // 2)
    for c in ("Hy", "Dx", "Dz", "Ex", "Ez"):
        assert rel_l2(hg.field(c), full.field(c)) <= 1e-5, c
        assert np.array_equal(hg.field(c), mid.field(c)), c


@pytest.mark.parametrize("case", ["dispersive", "nonlinear", "soft"])
def test_3d_without_stored_d(case):
    """3D keeps no D for plain cells: Drude/Lorentz/Debye, Kerr + plasma and point-source
    cells keep theirs (chosen by the id, D loaded on use), TFSF face cells too.  Parity
    against the oracle, agreement with the stored-D run to rounding, and a mid-run read
    of D that does not change the result."""
    if case == "dispersive":
        model, mons = cases.dispersive_3d(), None
    elif case == "nonlinear":
        model, mons = cases.nonlinear_3d(), None
    else:
        model, mons = cases.spheres_3d(n=32, steps=100)
    hg, ho, desc = run_pair(model, mons)
    assert_parity(hg, ho, desc, model)
    full = Handle(desc, _lib.engine())
    full.lib.check(full.lib.model_set_options(full.h, _lib.MFD_OPT_DLESS, 0), "set_options")
    full.step(desc.Nt)
    mid = Handle(desc, _lib.engine())
    mid.step(desc.Nt // 2)
    mid.field("Dy")
    mid.step(desc.Nt - desc.Nt // 2)
    for c in ("Hx", "Hz", "Dx", "Dy", "Dz", "Ex", "Ey", "Ez"):
        assert rel_l2(hg.field(c), full.field(c)) <= 1e-5, c
        assert np.array_equal(hg.field(c), mid.field(c)), c


def test_no_drift_without_stored_d():
    """A closed, lossless box (periodic, no CPML) keeps the field for good: over 3000
    steps the run without stored D stays within rounding of the run with stored D and of
    the FP64 oracle.  A plain cell whose D were rebuilt as fl(E * fl(1/Me)) every step
    would gain about 6e-8 per step in vacuum and miss the first bound."""
    model, mons = cases.spheres_3d(n=32, steps=3000, soft=False, nspheres=6, pml=0,
                                   monitors=False)
    desc = model.describe()
    a = Handle(desc, _lib.engine())
    a.step(desc.Nt)
    b = Handle(desc, _lib.engine())
    b.lib.check(b.lib.model_set_options(b.h, 11, 0), "set_options")
    b.step(desc.Nt)
    errs = field_errors(a, b, 3)
    assert max(errs.values()) <= 5e-5, errs
    ho = Handle(desc, oracle("f64"))
    ho.step(desc.Nt)
    errs = field_errors(a, ho, 3)
    assert max(errs.values()) <= 1e-4, errs
    assert np.max(np.abs(ho.field("Ex"))) > 0.1


def test_graph_and_eager_agree_bitwise():
    model, mons = cases.spheres_3d(n=24, steps=70)
    desc = model.describe(monitors=mons)
    outs = []
    for g in (0, 16):
        h = Handle(desc, _lib.engine())
        h.lib.check(h.lib.model_set_options(h.h, 0, g), "set_options")
        h.step(model.Nt)
        outs.append([h.field(c) for c in ("Hx", "Ey", "Dz")] + [np.ravel(h.monitor(2, "Ex"))])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kform", [2, 0])
def test_h_tile_list_is_bitwise_neutral(kform):
    """The H half marching its own z chunk, its tiles in XCD chunks or in grid order (then
    the H list and the E list differ in length), and the tiles dealt to the XCDs in
    blocks of any size, change no bit."""
    model, mons = cases.spheres_3d(n=40, nx=144, steps=60)
    desc = model.describe(monitors=mons)
    outs = []
    for zh, xcd, blk in ((0, 1, 0), (8, 0, 64), (4, 1, 7), (32, 0, 1)):
        h = Handle(desc, _lib.engine())
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_KERNEL_FORM, kform), "set_options")
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_ZCHUNK_H, zh), "set_options")
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_XCD_H, xcd), "set_options")
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_XCD_BLOCK, blk), "set_options")
        h.step(model.Nt)
        outs.append([h.field(c) for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")] + [h.output(0)])
        h.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", ["periodic", "dirichlet", "neumann"])
def test_fused_step_tiling_is_bitwise_neutral(bc):
    """The fused step (one z march per step over ping-pong H/E buffers, H recomputed one
    cell below each tile in x, y and z) gives the same bits whatever the tile shape: 2 or
    4 rows per lane, z chunks of 3 to 32 planes, one launch or three CPML-class launches
    (one after the other or side by side), any XCD block.  The grid is 150 x 41 x 37 so
    that tiles are partial in x, y and z, and the faces carry `bc`.  The result also
    matches the FP64 oracle and the two-pass form to rounding."""
    model, mons = cases.spheres_3d(n=37, nx=150, steps=60, bc=bc)
    desc = model.describe(monitors=mons)
    outs = []
    for rows, zc, split, blk in ((4, 16, 0, 64), (2, 3, 1, 7), (4, 32, 2, 1), (2, 8, 0, 0)):
        h = Handle(desc, _lib.engine())
        for opt, v in ((_lib.MFD_OPT_KERNEL_FORM, 3), (_lib.MFD_OPT_ROWS, rows), (_lib.MFD_OPT_ZCHUNK, zc),
                       (_lib.MFD_OPT_PML_SPLIT, split), (_lib.MFD_OPT_XCD_BLOCK, blk)):
            h.lib.check(h.lib.model_set_options(h.h, opt, v), "set_options")
        h.step(model.Nt)
        outs.append([h.field(c) for c in ("Hx", "Hy", "Hz", "Dx", "Ex", "Ey", "Ez")] + [h.output(0), h.output(1)]
                    + [np.ravel(h.monitor(q, "Ez")) for q in range(len(mons))])
        if len(outs) == 1:
            ho = Handle(desc, oracle("f64"))
            ho.step(model.Nt)
            assert_parity(h, ho, desc, model)
            two = Handle(desc, _lib.engine())
            two.lib.check(two.lib.model_set_options(two.h, _lib.MFD_OPT_KERNEL_FORM, 2), "set_options")
            two.step(model.Nt)
            errs = field_errors(h, two, 3)
            assert max(errs.values()) <= 1e-5, errs
        h.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("bc,pml", [("periodic", 5), ("dirichlet", 4), ("neumann", 0)])
def test_fused_2d_step_tiling_is_bitwise_neutral(bc, pml):
    """The fused 2D step (MFD_OPT_FUSED_2D: one z march per step over ping-pong Hy/Ex/Ez buffers, Hy
    recomputed one column left of and one row below each tile) gives the same bits for
    z chunks of 1 to 64 planes, with faces `bc`; the grid (150 x 77) leaves tiles partial
    in x and z.  It matches the FP64 oracle, and the two-pass 2D kernels to rounding."""
    model, mons = cases.lorentz_2d(nx=150, nz=77, steps=140, bc=bc, pml=pml)
    desc = model.describe(monitors=mons, integrals=3)
    outs = []
    for zc in (16, 1, 7, 64):
        h = Handle(desc, _lib.engine())
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_FUSED_2D, 1), "set_options")
        h.lib.check(h.lib.model_set_options(h.h, _lib.MFD_OPT_ZCHUNK, zc), "set_options")
        h.step(model.Nt)
        outs.append([h.field(c) for c in ("Hy", "Dx", "Dz", "Ex", "Ez")] + [h.output(0), h.output(1)]
                    + [np.ravel(h.monitor(q, "Ez")) for q in range(len(mons))])
        if len(outs) == 1:
            ho = Handle(desc, oracle("f64"))
            ho.step(model.Nt)
            assert_parity(h, ho, desc, model)
            two = Handle(desc, _lib.engine())
            two.step(model.Nt)
            errs = field_errors(h, two, 2)
            assert max(errs.values()) <= 1e-5, errs
        h.close()
    for o in outs[1:]:
        for k, (a, b) in enumerate(zip(outs[0], o)):
            assert np.array_equal(a, b), k


def test_fused_step_buffers_follow_the_model():
    """The fused step swaps each model's H/E buffers every step.  A model stepped alone,
    then in a batch, then alone again (three runners, each holding its own copy of the
    model descriptor) and read back between the runs gives the bits of one uninterrupted
    run, for an odd and an even number of steps per leg."""
    model, mons = cases.spheres_3d(n=30, steps=77)
    desc = model.describe(monitors=mons)
    ref = Handle(desc, _lib.engine())
    ref.step(model.Nt)
    h = Handle(desc, _lib.engine())
    h.step(13)
    h.field("Ex")
    lib = h.lib
    b = ctypes.c_void_p()
    lib.check(lib.batch_create(ctypes.byref(b)), "batch_create")
    lib.check(lib.batch_add(b, h.h), "batch_add")
    lib.check(lib.batch_run(b, 24), "batch_run")
    lib.check(lib.batch_destroy(b), "batch_destroy")
    h.field("Hy")
    h.step(model.Nt - 37)
    for c in ("Hx", "Hy", "Hz", "Dz", "Ex", "Ey", "Ez"):
        assert np.array_equal(ref.field(c), h.field(c)), c
    assert np.array_equal(np.ravel(ref.monitor(2, "Ex")), np.ravel(h.monitor(2, "Ex")))


def test_chunked_stepping_is_bitwise_identical():
    model, mons = cases.spheres_3d(n=24, steps=90)
    desc = model.describe(monitors=mons)
    h1 = Handle(desc, _lib.engine())
    h1.step(model.Nt)
    h2 = Handle(desc, _lib.engine())
    for n in (1, 7, 33, 2, 19, 28):
        h2.step(n)
        h2.monitor(2, "Ex")      # forces a partial DFT flush mid-run
    for c in ("Hy", "Ez", "Dx"):
        assert np.array_equal(h1.field(c), h2.field(c))
    for i in range(4):
        for c in ("Ex", "Hz"):
            a, b = h1.monitor(i, c), h2.monitor(i, c)
            assert rel_l2(a, b) <= 1e-12


def test_nonfinite_is_reported():
    from maxwellfdtd_amd import Grid, Model, SoftSource
    grid = Grid(zmin=0, zmax=1e-6, Nz=101)
    src = SoftSource(geometry=(0.5e-6,), waveform=lambda z, t: 1.0 + 0 * t, component="Ex")
    model = Model(grid, src, tmax=4000 * 1.5 * grid.dz / 3e8, CN=1.5, pml=0)
    desc = model.describe()
    h = Handle(desc, _lib.engine())
    with pytest.raises(_lib.EngineError, match="I found (NaN|Inf) field values"):
        h.step(model.Nt)


# ---- total field energy ----------------------------------------------------------------
def test_energy_matches_fp64_fields_along_the_run():
    """The engine's total field energy (W_E, W_H) against the same sums over the FP64
    oracle's fields, at several steps of a run with CPML, TFSF, a soft source and
    dielectric spheres (mu = 1 and eps from the materials)."""
    model, mons = cases.spheres_3d(n=32, steps=160)
    desc = model.describe()
    hg = Handle(desc, _lib.engine())
    ho = Handle(desc, oracle("f64"))
    for n in (40, 40, 80):
        hg.step(n)
        ho.step(n)
        we, wh = hg.energy()
        oe, oh = field_energy(ho)
        assert oe > 0 and oh > 0
        assert abs(we / oe - 1) <= 1e-4 and abs(wh / oh - 1) <= 1e-4, (we, oe, wh, oh)


def test_energy_closed_form_and_conservation():
    """A uniform Ex in a periodic box, half of it dielectric: W_E = 1/2 eps0 eps E0^2 V
    exactly and W_H = 0.  And in a closed lossless periodic box, once the source has
    stopped, the discrete energy of the leapfrog scheme, W_E(E^n) + 1/2 sum mu0 mu
    H^(n-1/2).H^(n+1/2) dV, stays constant (the plain sum of squares swings by several
    per cent: E and H are half a step apart)."""
    from maxwellfdtd_amd import Grid, Material, Model, SoftSource
    n, d = 16, 40e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    src = SoftSource(geometry=(L / 2, L / 2, L / 2), waveform=lambda x, y, z, t: 0 * t,
                     component="Ex")
    mat = Material(geometry=lambda x, y, z: z < L / 2, eps=2.5)
    model = Model(grid, src, tmax=10 * d / 3e8, pml=0, material=mat)
    h = Handle(model.describe(integrals=0), _lib.engine())
    E0 = 3.0e5
    h.set_field("Ex", np.full(grid.shape, E0))
    we, wh = h.energy()
    eps_planes = np.where(np.arange(n) * d < L / 2, 2.5, 1.0)
    expect = 0.5 * EPS0 * E0 ** 2 * d ** 3 * n * n * float(np.sum(eps_planes))
    assert abs(we / expect - 1) < 1e-6 and wh == 0.0

    pm, _ = cases.spheres_3d(n=24, steps=900, tfsf=False, monitors=False, pml=0, nspheres=3)
    desc = pm.describe(integrals=0)
    hg = Handle(desc, _lib.engine())
    Mh = np.array([desc.vac[0]] + [c["Mh"] for c in desc.materials])
    mu = desc.dt / Mh[desc.geometry]
    dV = float(np.prod(desc.d))

    def discrete_energy():
        we_, _ = hg.energy()
        H0 = [hg.field(c) for c in ("Hx", "Hy", "Hz")]
        hg.step(1)
        H1 = [hg.field(c) for c in ("Hx", "Hy", "Hz")]
        return we_ + 0.5 * dV * float(np.sum(mu * sum(a * b for a, b in zip(H0, H1))))

    hg.step(450)                 # the soft source's pulse is over by then
    w = [discrete_energy()]
    for _ in range(4):
        hg.step(99)
        w.append(discrete_energy())
    w = np.array(w)
    assert w[0] > 0 and np.max(np.abs(w / w[0] - 1)) < 1e-5, w / w[0] - 1


# ---- FP64 state for 1D models (the reference's CPU{Float64} configuration) -------------
F64_TOL = 1e-9
F64_OUT_TOL = 1e-6      # monitor samples are stored in FP32


def run_pair_f64(model, mons=None, steps=None):
    desc = model.describe(monitors=mons)
    n = model.Nt if steps is None else steps
    hg = Handle(desc, _lib.engine())
    hg.lib.check(hg.lib.model_set_options(hg.h, _lib.MFD_OPT_F64, 1), "set_options(F64)")
    hg.step(n)
    ho = Handle(desc, oracle("f64"))
    ho.step(n)
    return hg, ho, desc


def _tfsf_plane_wave_1d(nz=300, steps=500):
    from maxwellfdtd_amd import Grid, Material, Model, TFSFSource
    d = 40e-9
    grid = Grid(zmin=0, zmax=(nz - 1) * d, Nz=nz)
    L = (nz - 1) * d
    w = cases.gauss_waveform(lam0=800e-9, tau0=5e-15, dt0=15e-15)
    src = TFSFSource(box=(0.3 * L, 0.7 * L), waveform=w, component="Ex", z0=0.2 * L)
    mat = Material(geometry=lambda z: (z > 0.45 * L) & (z < 0.55 * L), eps=3.0)
    return Model(grid, src, tmax=(steps - 1) * 0.99 * d / 3e8, CN=0.99, pml=8 * d, material=mat)


@pytest.mark.parametrize("case", ["slab", "nonlinear", "tfsf_tables", "tfsf_aux", "kappa", "ref"])
def test_f64_1d_matches_the_fp64_oracle(case):
    """A 1D model with MFD_OPT_F64 keeps fields, CPML, coefficients, ADE / plasma state,
    source and replay tables in FP64 on the GPU: its fields match the FP64 oracle to 1e-9
    (operation order and fused multiply-adds differ), where the FP32 state sits at 1e-5
    to 1e-4; Sa/E2 (FP64 accumulators) likewise, the FP32-stored monitor samples to 1e-6."""
    mons = None
    if case == "slab":
        model, mons = cases.slab_1d()
    elif case == "nonlinear":
        model, mons = cases.nonlinear_1d()
    elif case == "tfsf_tables":
        model = cases.tfsf_tables_model(1)
    elif case == "tfsf_aux":
        model = _tfsf_plane_wave_1d()
    elif case == "kappa":
        model, mons = cases.slab_1d(steps=700, kmax=5.0)
    else:
        model, _, _ = cases.ref_1d_test()
    hg, ho, desc = run_pair_f64(model, mons)
    assert_parity(hg, ho, desc, model, tol=F64_TOL, out_tol=F64_OUT_TOL)
    if model.isplasma:
        assert np.max(ho.output(3)) > 0


def test_f64_1d_config_full_length():
    """The 1D configuration at its full length (Nz = 4000, 8000 steps, eps = 2.25 slab,
    soft Gaussian source, point and 32-wavelength DFT monitors) in FP64 matches the FP64
    oracle far inside 1e-4; the same run in FP32 does not reach 1e-4 on the final field
    (the pulse has left and only a small residual remains)."""
    from maxwellfdtd_amd.workloads import slab_1d
    model, mons = slab_1d()
    assert model.grid.Nz == 4000 and model.Nt == 8000
    hg, ho, desc = run_pair_f64(model, mons)
    errs = assert_parity(hg, ho, desc, model, tol=F64_TOL, out_tol=F64_OUT_TOL)
    assert max(errs.values()) <= F64_TOL
    we, wh = hg.energy()
    oe, oh = field_energy(ho)
    assert abs(we / oe - 1) < 1e-9 and abs(wh / oh - 1) < 1e-9


def test_f64_option_limits():
    model, _ = cases.spheres_3d(n=12, steps=10, monitors=False, tfsf=False, pml=2)
    h = Handle(model.describe(), _lib.engine())
    assert h.lib.model_set_options(h.h, _lib.MFD_OPT_F64, 1) == _lib.MFD_ERR_ARG
    m1, _ = cases.slab_1d(nz=100, steps=20)
    h1 = Handle(m1.describe(), _lib.engine())
    h1.step(2)
    assert h1.lib.model_set_options(h1.h, _lib.MFD_OPT_F64, 1) == _lib.MFD_ERR_ARG
