"""Batched runs and the serving loop on the GPU: every solve's result in a batch or in
the served mix equals the same solve stepped alone (bitwise: the per-cell arithmetic
does not depend on the tile list or launch shape), and matches the FP64 oracle within
1e-4 relative L2."""
import itertools
import time

import numpy as np
import pytest

from maxwellfdtd_amd import _lib, workloads
from maxwellfdtd_amd.batch import Batch, Server, replay
from maxwellfdtd_amd.models import Handle, MONITOR_COMPONENTS
from oracle_util import field_errors, oracle, rel_l2

pytestmark = pytest.mark.gpu

COMPS = {2: ("Hy", "Dx", "Dz", "Ex", "Ez"), 3: ("Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ex", "Ey", "Ez")}


def small_2d(count=6, seed=2028):
    return list(itertools.islice(workloads.batched_2d(count=count, seed=seed, nmin=48, nmax=120,
                                                      smin=60, smax=160), count))


def small_3d(count=4, seed=2029):
    return [m for _, m in itertools.islice(workloads.served_3d(count=count, seed=seed, nmin=32,
                                                               nmax=44, smin=40, smax=90), count)]


def alone(desc):
    h = Handle(desc, _lib.engine())
    h.step(desc.Nt)
    return h


def assert_same(h1, h2, ndims, sa_tol=0.0):
    for c in COMPS[ndims]:
        a, b = h1.field(c), h2.field(c)
        assert np.array_equal(a, b), (c, np.max(np.abs(a - b)))
    if h1.desc.integrals & 1:
        if sa_tol == 0.0:
            assert np.array_equal(h1.output(0), h2.output(0))
        else:
            assert rel_l2(h1.output(0), h2.output(0)) <= sa_tol


@pytest.mark.parametrize("fused", [0, 1])
def test_batch_2d_equals_alone_and_oracle(fused):
    """A batch of uneven 2D solves (two-pass kernels, or the fused 2D step) equals each
    solve run alone, bit for bit, and the FP64 oracle."""
    models = small_2d()
    descs = [m.describe() for m in models]
    assert len({d.Nt for d in descs}) > 1 and len({d.dims for d in descs}) > 1
    b = Batch(descs, options={_lib.MFD_OPT_FUSED_2D: fused})
    b.run(0)
    assert b.active == 0
    for d, hb in zip(descs, b.handles):
        assert hb.it == d.Nt
        ha = Handle(d, _lib.engine())
        ha.lib.check(ha.lib.model_set_options(ha.h, _lib.MFD_OPT_FUSED_2D, fused), "set_options")
        ha.step(d.Nt)
        assert_same(hb, ha, 2)
        ha.close()
    for d, hb in list(zip(descs, b.handles))[:2]:
        ho = Handle(d, oracle("f64"))
        ho.step(d.Nt)
        errs = field_errors(hb, ho, 2)
        assert max(errs.values()) <= 1e-4, errs
        ho.close()
    b.close()


def spectral_models(sizes_steps, seed=41):
    """3D models with field and spectral monitors (DFT over blocks of 32 steps): step
    counts that are not multiples of the block, so blocks of different solves end at
    different times and every solve ends inside a block."""
    import cases
    return [cases.spheres_3d(n=n, steps=s, seed=seed + k)[0].describe(
        monitors=cases.spheres_3d(n=n, steps=s, seed=seed + k)[1])
        for k, (n, s) in enumerate(sizes_steps)]


def monitors_of(h):
    names = MONITOR_COMPONENTS[h.desc.ndims]
    return [[h.monitor(i, c) for c in names] for i in range(len(h.desc.monitors))]


def assert_same_monitors(a, b):
    for i, (ma, mb) in enumerate(zip(a, b)):
        for ca, cb in zip(ma, mb):
            assert np.array_equal(ca, cb), (i, np.max(np.abs(np.asarray(ca) - np.asarray(cb))))


def test_batch_spectral_monitors_uneven_nt():
    """Spectral monitors in a batch of unequal Nt (45, 83, 130 steps) equal the solo
    runs bitwise: each solve contracts its own DFT blocks, and a solve that finished is
    not contracted again."""
    descs = spectral_models([(20, 45), (24, 83), (20, 130)])
    solo = []
    for d in descs:
        h = alone(d)
        solo.append((monitors_of(h), h.field("Ex")))
        h.close()
    b = Batch(descs, options={_lib.MFD_OPT_GRAPH_STEPS: 16})
    b.run(0)
    for (mons, ex), hb in zip(solo, b.handles):
        assert np.array_equal(hb.field("Ex"), ex)
        assert_same_monitors(monitors_of(hb), mons)
    b.close()


def test_server_spectral_monitors_staggered():
    """Spectral monitors of requests admitted at different times (small quanta, so the
    solves' DFT blocks are out of phase) equal the solo runs bitwise."""
    descs = spectral_models([(20, 70), (22, 101), (20, 59), (18, 90)], seed=43)
    solo = []
    for d in descs:
        h = alone(d)
        solo.append(monitors_of(h))
        h.close()
    srv = Server(options={_lib.MFD_OPT_QUANTUM: 5})
    got = {}

    def done(t, h, lat):
        got[t] = monitors_of(h)

    arrivals = [(0.004 * k, d) for k, d in enumerate(descs)]
    tickets = replay(srv, arrivals, describe=lambda d: d, on_done=done)
    assert srv.stats()["done"] == len(descs)
    for t, mons in zip(tickets, solo):
        assert_same_monitors(got[t], mons)
    srv.close()


def test_server_fails_what_free_memory_cannot_hold():
    """A request within the HBM budget but larger than the device's free memory, with
    nothing else running, fails with MFD_ERR_NOMEM instead of waiting for memory that no
    request of the server will free; the server's statistics then show nothing active."""
    from maxwellfdtd_amd import FieldMonitor, Grid, Model, SoftSource
    n, d = 32, 40e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    src = SoftSource(geometry=(L / 2, L / 2, L / 2), waveform=lambda x, y, z, t: 0 * t,
                     component="Ex")
    dt = 0.99 / 3e8 / np.sqrt(3 / d ** 2)
    # a field monitor on every cell for 10^6 steps: 6 x 32768 x 10^6 x 4 B = 786 GB
    model = Model(grid, src, tmax=(10 ** 6 - 1) * dt, CN=0.99, pml=0)
    desc = model.describe(monitors=[FieldMonitor(lambda x, y, z: 0 * x + 0 * y + 0 * z == 0)])
    srv = Server(hbm_budget=10 ** 15)
    t = srv.submit(desc)[0]
    state = srv.wait(t, 60000)
    st, code, _ = srv.poll(t)
    assert state == _lib.MFD_REQ_FAILED and st == _lib.MFD_REQ_FAILED and code == _lib.MFD_ERR_NOMEM
    stats = srv.stats()
    assert stats["failed"] == 1 and stats["active"] == 0 and stats["queued"] == 0
    srv.close()


def test_server_stats_after_retirement():
    """After the last request retires, the statistics report no active request."""
    descs = [m.describe() for m in small_3d(count=2, seed=17)]
    srv = Server(options={_lib.MFD_OPT_QUANTUM: 8})
    for d in descs:
        srv.submit(d)
    assert srv.wait(0, 120000) == _lib.MFD_REQ_DONE
    st = srv.stats()
    assert st["done"] == 2 and st["active"] == 0 and st["queued"] == 0
    srv.close()


def test_batch_autotune_keeps_results():
    models = small_2d(count=4, seed=7)
    descs = [m.describe() for m in models]
    b = Batch(descs)
    best = b.autotune(trial_steps=2)
    assert best["zchunk"] in (8, 16, 32, 64, 128) and best["ms_per_step"] > 0
    b.run(0)
    for d, hb in zip(descs, b.handles):
        ha = alone(d)
        assert_same(hb, ha, 2)
        ha.close()
    b.close()


def test_batch_3d_uneven():
    models = small_3d(count=3)
    descs = [m.describe() for m in models]
    b = Batch(descs, options={_lib.MFD_OPT_GRAPH_STEPS: 8})
    b.run(0)
    for d, hb in zip(descs, b.handles):
        ha = alone(d)
        assert_same(hb, ha, 3)
        ha.close()
    d, hb = descs[0], b.handles[0]
    ho = Handle(d, oracle("f64"))
    ho.step(d.Nt)
    errs = field_errors(hb, ho, 3)
    assert max(errs.values()) <= 1e-4, errs
    assert rel_l2(hb.output(3), ho.output(3)) <= 1e-4     # plasma density
    ho.close()
    b.close()


def test_batch_3d_mixed_integrals():
    """A batch of 3D models asking for Sa and E2, Sa only and E2 only runs the fused step
    built for per-model integral variables: each model's fields and integrals equal its
    run alone, and the Sa-only and E2-only models match the FP64 oracle."""
    models = small_3d(count=3)
    descs = [m.describe(integrals=flags) for m, flags in zip(models, (3, 1, 2))]
    b = Batch(descs, options={_lib.MFD_OPT_GRAPH_STEPS: 8})
    b.run(0)
    for d, hb in zip(descs, b.handles):
        ha = alone(d)
        assert_same(hb, ha, 3)
        if d.integrals & 2:
            assert np.array_equal(hb.output(1), ha.output(1))
        ha.close()
    for d, hb in zip(descs[1:], b.handles[1:]):
        ho = Handle(d, oracle("f64"))
        ho.step(d.Nt)
        assert max(field_errors(hb, ho, 3).values()) <= 1e-4
        q = 0 if d.integrals & 1 else 1
        assert rel_l2(hb.output(q), ho.output(q)) <= 1e-4, q
        ho.close()
    b.close()


def test_server_results_equal_alone():
    models = small_3d(count=5, seed=11)
    descs = [m.describe() for m in models]
    ref = []
    for d in descs:
        h = alone(d)
        ref.append({c: h.field(c) for c in COMPS[3]} | {"Sa": h.output(0)})
        h.close()
    srv = Server(options={_lib.MFD_OPT_QUANTUM: 16})
    got = {}

    def done(t, h, lat):
        assert lat is not None and lat > 0
        got[t] = {c: h.field(c) for c in COMPS[3]} | {"Sa": h.output(0)}

    arrivals = [(0.01 * k, d) for k, d in enumerate(descs)]
    tickets = replay(srv, arrivals, describe=lambda d: d, on_done=done)
    st = srv.stats()
    assert st["done"] == 5 and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["gcells_per_s"] > 0
    for t, r in zip(tickets, ref):
        for k, v in r.items():
            assert np.array_equal(got[t][k], v), k
    srv.close()


def test_server_budget_and_oversize():
    models = small_3d(count=3, seed=5)
    descs = [m.describe() for m in models]
    fps = []
    for d in descs:
        h = Handle(d, _lib.engine())
        fps.append(C_footprint(h))
        h.close()
    # room for the largest request plus a bit: requests run about one at a time
    srv = Server(hbm_budget=int(max(fps) * 1.2), max_active=8)
    # a model whose footprint exceeds the whole budget fails with MFD_ERR_NOMEM
    big = workloads.plane_wave_spheres_3d(n=96, steps=4, nspheres=2, rmin=2, rmax=4, box_margin=14)
    tb, _ = srv.submit(big.describe())
    ts = [srv.submit(d)[0] for d in descs]
    assert srv.wait(0, 120000) == _lib.MFD_REQ_DONE
    st, code, _ = srv.poll(tb)
    assert st == _lib.MFD_REQ_FAILED and code == _lib.MFD_ERR_NOMEM
    for t in ts:
        st, code, _ = srv.poll(t)
        assert st == _lib.MFD_REQ_DONE and code == 0
    srv.close()


def C_footprint(h):
    import ctypes as C
    v = C.c_int64()
    h.lib.check(h.lib.model_footprint(h.h, C.byref(v)), "footprint")
    return v.value


def test_batch_autotune_3d_keeps_results():
    models = small_3d(count=3, seed=21)
    descs = [m.describe() for m in models]
    b = Batch(descs, options={_lib.MFD_OPT_GRAPH_STEPS: 4})
    best = b.autotune(trial_steps=2)
    assert best["zchunk"] in (8, 16, 32) and best["rows"] in (2, 4, 8)
    b.run(0)
    for d, hb in zip(descs, b.handles):
        ha = alone(d)
        assert_same(hb, ha, 3)
        ha.close()
    b.close()


def unstable_3d(n=32, steps=200):
    """A model that blows up: CN = 1.5 breaks the Courant limit."""
    from maxwellfdtd_amd import Grid, Model, SoftSource
    d = 40e-9
    L = (n - 1) * d
    grid = Grid(xmin=0, xmax=L, Nx=n, ymin=0, ymax=L, Ny=n, zmin=0, zmax=L, Nz=n)
    src = SoftSource(geometry=(L / 2, L / 2, L / 2), waveform=lambda x, y, z, t: 1.0 + 0 * t,
                     component="Ex")
    dt = 1.5 / 3e8 / np.sqrt(3 / d ** 2)
    return Model(grid, src, tmax=(steps - 1) * dt, CN=1.5, pml=0)


def test_batch_and_server_isolate_a_failing_solve():
    """A solve that goes non-finite is reported as such and drops out; the others run to
    the end with the results they have alone.  The failing solve has a soft source, so
    while it is in the batch the TFSF faces of the others take the per-face kernels
    instead of the gathered pass: the fields stay bitwise equal, the Sa integral of the
    face cells is summed in another order (round-off)."""
    good = [m.describe() for m in small_3d(count=2, seed=31)]
    bad = unstable_3d().describe()
    b = Batch([good[0], bad, good[1]])
    with pytest.raises(_lib.EngineError, match="I found (NaN|Inf) field values"):
        b.run(0)
    assert b.active == 0
    assert b.handles[1].status in (_lib.MFD_ERR_NAN, _lib.MFD_ERR_INF)
    for d, hb in ((good[0], b.handles[0]), (good[1], b.handles[2])):
        assert hb.status == 0 and hb.it == d.Nt
        ha = alone(d)
        assert_same(hb, ha, 3, sa_tol=1e-6)
        ha.close()
    b.close()
    srv = Server()
    t0 = srv.submit(good[0])[0]
    tb = srv.submit(bad)[0]
    assert srv.wait(0, 120000) == _lib.MFD_REQ_DONE
    st, code, _ = srv.poll(tb)
    assert st == _lib.MFD_REQ_FAILED and code in (_lib.MFD_ERR_NAN, _lib.MFD_ERR_INF)
    st, code, _ = srv.poll(t0)
    assert st == _lib.MFD_REQ_DONE and code == 0
    assert srv.stats()["failed"] == 1
    srv.close()


def test_server_admits_least_work_first():
    """With room for one request at a time (fill of 1 Mi cells, requests of 81^3 cells),
    requests queued behind a running one are admitted least work first: the short
    requests finish while the long one, submitted before them, has not."""
    import cases
    mk = lambda steps: cases.spheres_3d(n=81, steps=steps, monitors=False)[0].describe()
    blocker, long_, short1, short2 = mk(6000), mk(400), mk(60), mk(120)
    srv = Server(options={_lib.MFD_OPT_FILL_MCELLS: 1})
    tb = srv.submit(blocker)[0]
    t0 = time.monotonic()
    while srv.poll(tb)[0] == _lib.MFD_REQ_QUEUED and time.monotonic() - t0 < 10:
        time.sleep(0.001)
    assert srv.poll(tb)[0] == _lib.MFD_REQ_RUNNING
    tl = srv.submit(long_)[0]
    t1 = srv.submit(short1)[0]
    t2 = srv.submit(short2)[0]
    assert srv.poll(tb)[0] == _lib.MFD_REQ_RUNNING     # the others queue behind it
    assert srv.wait(t1, 120000) == _lib.MFD_REQ_DONE
    assert srv.poll(tl)[0] != _lib.MFD_REQ_DONE
    assert srv.wait(t2, 120000) == _lib.MFD_REQ_DONE
    assert srv.poll(tl)[0] != _lib.MFD_REQ_DONE
    assert srv.wait(0, 120000) == _lib.MFD_REQ_DONE
    assert srv.stats()["done"] == 4
    srv.close()
