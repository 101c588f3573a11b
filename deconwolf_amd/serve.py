"""Batched and served SHB deconvolution (libdwhip's engine, include/dwhip.h).

    eng = Engine(hbm_budget_gb=64, max_active=32)
    eng.submit(tag, im, psf, DwOpts(...))     # torch GPU tensors, float32
    done = eng.step()                          # one SHB iteration of every active job
    out = eng.output(tag)

`serve()` drives an engine with a stream of timed arrivals (continuous
batching: requests that arrive while others run are admitted between
rounds as memory frees up) and returns per-job latencies.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import DwhipStats, EngineConfig, check, lib
from .method_shb import DwOpts, _params


class Engine:
    """order: _lib.ORDER_FIFO (arrival order) or ORDER_SHORTEST (least work
    first); active_mvox: cap on the transform voxels (millions) of the jobs
    iterating together, 0 for none; groups: 2 runs the active jobs as two
    groups on two streams whose passes overlap; rounds_per_step: rounds a
    step may run while every active job is FIXED (replayed from a captured
    HIP graph of the round), None for the library's default; aging_ms:
    with ORDER_SHORTEST, a job that has waited that long goes before every
    younger one (oldest first), 0 for least work first only."""

    def __init__(self, hbm_budget_gb=32.0, max_active=64, stream=None, order=0, active_mvox=0, groups=1,
                 rounds_per_step=None, aging_ms=0):
        import torch
        cfg = EngineConfig()
        lib().dwhip_engine_config_default(ctypes.byref(cfg))
        cfg.hbm_budget = int(hbm_budget_gb * (1 << 30))
        cfg.max_active = int(max_active)
        cfg.order = int(order)
        cfg.active_mvox = int(active_mvox)
        cfg.groups = int(groups)
        if rounds_per_step is not None:
            cfg.rounds_per_step = int(rounds_per_step)
        cfg.aging_ms = int(aging_ms)
        h = ctypes.c_void_p()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().dwhip_engine_create(ctypes.byref(cfg), s, ctypes.byref(h)), "dwhip_engine_create")
        self.h = h
        self.jobs = {}   # tag -> (im, psf, out, start, opts)

    def close(self):
        if self.h:
            lib().dwhip_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, tag, im, psf, opts: DwOpts, start_image=None):
        import torch
        im = im.contiguous()
        psf = psf.contiguous()
        out = torch.empty_like(im)
        P, N, M = im.shape
        pP, pN, pM = psf.shape
        p = _params(opts)
        if start_image is not None:
            start_image = start_image.contiguous()
            p.start_image = start_image.data_ptr()
        check(lib().dwhip_engine_submit(self.h, im.data_ptr(), M, N, P, psf.data_ptr(), pM, pN, pP,
                                        ctypes.byref(p), out.data_ptr(), int(tag)), "dwhip_engine_submit")
        self.jobs[int(tag)] = (im, psf, out, start_image, opts)

    def step(self, max_done=4096):
        tags = (ctypes.c_int64 * max_done)()
        n = ctypes.c_int()
        check(lib().dwhip_engine_step(self.h, tags, max_done, ctypes.byref(n)), "dwhip_engine_step")
        return [int(tags[i]) for i in range(n.value)]

    def status(self):
        q, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        check(lib().dwhip_engine_status(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(b)))
        return q.value, a.value, b.value

    def result(self, tag, release=True):
        """(output tensor, stats dict with the error trace) of a completed job."""
        im, psf, out, start, opts = self.jobs[int(tag)]
        n = max(int(opts.nIter), int(opts.maxiter), 1)
        err = np.zeros(n, np.float32)
        st = DwhipStats()
        st.err_trace = err.ctypes.data
        st.err_trace_len = n
        check(lib().dwhip_engine_result(self.h, int(tag), ctypes.byref(st), 1 if release else 0),
              "dwhip_engine_result")
        if release:
            del self.jobs[int(tag)]
        return out, {"bg": st.bg, "iterations": st.iterations, "work": tuple(st.work),
                     "ref_work": tuple(st.ref_work), "err_trace": err[:st.iterations].copy()}


def serve(engine, requests, arrivals):
    """requests: list of (tag, im, psf, opts); arrivals: seconds after start
    (sorted).  Submits each request when its time has come, steps the
    engine until all are done.  Returns {tag: (arrival_s, done_s)}."""
    import torch
    t0 = time.perf_counter()
    pending = list(zip(arrivals, requests))
    k = 0
    times = {}
    left = len(requests)
    while left:
        now = time.perf_counter() - t0
        while k < len(pending) and pending[k][0] <= now:
            arr, (tag, im, psf, opts) = pending[k]
            engine.submit(tag, im, psf, opts)
            times[tag] = [arr, None]
            k += 1
        q, a, _ = engine.status()
        if q == 0 and a == 0:
            # idle until the next arrival
            if k < len(pending):
                time.sleep(max(0.0, pending[k][0] - (time.perf_counter() - t0)))
            continue
        for tag in engine.step():
            times[tag][1] = time.perf_counter() - t0
            left -= 1
    torch.cuda.synchronize()
    return {t: tuple(v) for t, v in times.items()}
