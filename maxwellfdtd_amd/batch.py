"""Batched runs and the serving loop over the engine's C ABI.

Batch   many built models of one dimensionality advanced together (mfd_batch_*): one H
        and one E launch per half-step over the packed tile list of every unfinished
        model, hipGraph replay, an autotuner over launch shapes.
Server  a worker thread in the engine that admits queued models against an HBM budget
        and advances the active set in rounds (mfd_server_*); requests that arrive
        while others run join between rounds.

Neither has a counterpart in the reference, which runs one model per solve! call
(src/models.jl:166-242).  Every model's results are those of stepping it alone.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from .models import Handle


class Batch:
    def __init__(self, descs, lib=None, options=None):
        self.lib = lib or _lib.engine()
        self.b = C.c_void_p()
        self.lib.check(self.lib.batch_create(C.byref(self.b)), "batch_create")
        self.handles = []
        try:
            for d in descs:
                h = Handle(d, self.lib)
                self.handles.append(h)
                self.lib.check(self.lib.batch_add(self.b, h.h), "batch_add")
            for k, v in (options or {}).items():
                self.lib.check(self.lib.batch_set_options(self.b, int(k), int(v)), "batch_set_options")
        except Exception:
            self.close()
            raise

    def autotune(self, trial_steps=16):
        zc, ty, pers = C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        self.lib.check(self.lib.batch_autotune(self.b, int(trial_steps), C.byref(zc), C.byref(ty),
                                               C.byref(pers), C.byref(ms)), "batch_autotune")
        return dict(zchunk=zc.value, rows=ty.value, persistent=pers.value, ms_per_step=ms.value)

    def run(self, nsteps=0):
        """Advance every unfinished model by nsteps (0: until all are done)."""
        self.lib.check(self.lib.batch_run(self.b, int(nsteps)), "batch_run")

    @property
    def active(self):
        n = C.c_int32()
        self.lib.check(self.lib.batch_active(self.b, C.byref(n)), "batch_active")
        return n.value

    def close(self):
        if self.b is not None and self.b.value:
            self.lib.batch_destroy(self.b)
        self.b = None
        for h in self.handles:
            h.close()
        self.handles = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The engine's request server.  submit() returns a ticket and the model's Handle;
    the handle can be read once poll() reports MFD_REQ_DONE."""

    def __init__(self, lib=None, hbm_budget=0, max_active=64, options=None):
        self.lib = lib or _lib.engine()
        self.s = C.c_void_p()
        self.lib.check(self.lib.server_create(int(hbm_budget), int(max_active), C.byref(self.s)),
                       "server_create")
        self.handles = {}
        for k, v in (options or {}).items():
            self.lib.check(self.lib.server_set_options(self.s, int(k), int(v)), "server_set_options")

    def submit(self, desc):
        h = Handle(desc, self.lib)
        t = C.c_int64()
        try:
            self.lib.check(self.lib.server_submit(self.s, h.h, C.byref(t)), "server_submit")
        except Exception:
            h.close()
            raise
        self.handles[t.value] = h
        return t.value, h

    def poll(self, ticket):
        st, code, lat = C.c_int32(), C.c_int32(), C.c_double()
        self.lib.check(self.lib.server_poll(self.s, int(ticket), C.byref(st), C.byref(code),
                                            C.byref(lat)), "server_poll")
        return st.value, code.value, lat.value

    def wait(self, ticket=0, timeout_ms=-1):
        st = C.c_int32()
        self.lib.check(self.lib.server_wait(self.s, int(ticket), int(timeout_ms), C.byref(st)),
                       "server_wait")
        return st.value

    def stats(self):
        s = _lib.mfd_server_stats()
        self.lib.check(self.lib.server_get_stats(self.s, C.byref(s)), "server_get_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def release(self, ticket):
        """Destroy a finished request's model, freeing its device memory."""
        h = self.handles.pop(ticket, None)
        if h is not None:
            h.close()

    def close(self):
        if self.s is not None and self.s.value:
            self.lib.server_destroy(self.s)
        self.s = None
        for h in self.handles.values():
            h.close()
        self.handles = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def replay(server, arrivals, describe, on_done=None, time_scale=1.0, poll_s=0.002):
    """Submit (t_arrival, model) pairs at their arrival times (scaled by time_scale) and
    wait for all of them.  `describe(model)` returns the ModelDescription to submit;
    `on_done(ticket, handle, latency_ms)` runs as each request finishes (the handle is
    released afterwards).  Returns the tickets in submission order."""
    t0 = time.perf_counter()
    tickets, pending = [], set()

    def reap():
        for t in list(pending):
            st, code, lat = server.poll(t)
            if st >= _lib.MFD_REQ_DONE:
                pending.discard(t)
                if on_done is not None:
                    on_done(t, server.handles.get(t), lat if st == _lib.MFD_REQ_DONE else None)
                server.release(t)

    for t_arr, model in arrivals:
        desc = describe(model)
        while time.perf_counter() - t0 < t_arr * time_scale:
            reap()
            time.sleep(poll_s)
        t, _ = server.submit(desc)
        tickets.append(t)
        pending.add(t)
        reap()
    while pending:
        server.wait(0, 50)
        reap()
    return tickets
