"""Serving: a request loop with continuous batching over the engine's jobs.

Requests are (rng, target, algorithm, precision) tuples, i.e. the arguments of
``importance_sample(rng, target, algorithm)`` (src/methods/importance_sampling.jl:857-874).
The native loop (is_server_* in include/issampler.h) admits them in arrival
order as execution slots and device memory free up, keeps each admitted job's
proposal slots, mixture sums and samples resident on the device, and runs the
jobs on their own streams so their work overlaps.  Each result equals the same
request run alone.
"""

from __future__ import annotations

import ctypes

from . import _lib
from ._lib import I32P, VP, ServerConfig, ServerStats, ServerTiming
from .api import FP32, FP64, PreparedSampler, _check, _job_desc, context

STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    def __init__(self, memory_budget=200e9, max_active=8, graph=False):
        self._lib = _lib.load()
        cfg = ServerConfig(int(memory_budget), int(max_active), int(bool(graph)))
        h = VP()
        _check(self._lib.is_server_create(context(), ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._requests = {}

    def submit(self, rng, target, algorithm, precision="fp32"):
        """Queues one request; returns its ticket (MemoryError if it alone exceeds the budget)."""
        prec = {"fp64": FP64, "fp32": FP32, FP64: FP64, FP32: FP32}[precision]
        desc, keep, p, sched = _job_desc(rng, target, algorithm, prec)
        t = ctypes.c_uint64()
        rc = self._lib.is_server_submit(self._h, ctypes.byref(desc), ctypes.byref(t))
        if rc == _lib.IS_ENOMEM:
            raise MemoryError(_lib.last_error())
        _check(rc)
        self._requests[t.value] = (rng, target, algorithm, prec)
        return t.value

    def poll(self, ticket):
        st = ctypes.c_int32()
        _check(self._lib.is_server_poll(self._h, ticket, ctypes.byref(st)))
        return STATES[st.value]

    def wait(self, ticket, timeout=None):
        st = ctypes.c_int32()
        ms = -1 if timeout is None else int(timeout * 1000)
        _check(self._lib.is_server_wait(self._h, ticket, ms, ctypes.byref(st)))
        return STATES[st.value]

    def take(self, ticket):
        """Hands the finished request over: (PreparedSampler, timing dict, ok)."""
        job = VP()
        tm = ServerTiming()
        rc = self._lib.is_server_take(self._h, ticket, ctypes.byref(job), ctypes.byref(tm))
        rng, target, algorithm, prec = self._requests.pop(ticket)
        timing = {"submit_ns": tm.submit_ns, "admit_ns": tm.admit_ns, "done_ns": tm.done_ns,
                  "device_bytes": tm.device_bytes, "latency_s": (tm.done_ns - tm.submit_ns) * 1e-9,
                  "queue_s": (tm.admit_ns - tm.submit_ns) * 1e-9}
        if rc not in (_lib.IS_OK, _lib.IS_EFAILED) or not job.value:
            _check(rc)
        sampler = PreparedSampler.from_handle(job, rng, target, algorithm, prec)
        return sampler, timing, rc == _lib.IS_OK

    def result(self, ticket, timeout=None):
        """Waits for the request and returns (WeightedSamples, timing); raises the
        request's own error (AMISRoundError, SamplerExecutionError, ...) when it failed."""
        self.wait(ticket, timeout)
        sampler, timing, ok = self.take(ticket)
        if not ok:
            sampler._raise_failure()
        return sampler.collect(), timing

    def summary(self, ticket, timeout=None):
        """Waits for the request, returns (round ESS, round log Z, timing, error) and frees
        the job without copying its samples to the host; error is None, or the exception
        the request's own failure would raise."""
        import numpy as np

        from ._lib import F64P

        self.wait(ticket, timeout)
        sampler, timing, ok = self.take(ticket)
        T = len(sampler.schedule)
        ess, lz, aess = np.zeros(T), np.zeros(T), np.zeros(T)
        error = None
        if ok:
            _check(sampler._lib.is_job_get_round_summary(sampler._job, ess.ctypes.data_as(F64P),
                                                         lz.ctypes.data_as(F64P), aess.ctypes.data_as(F64P)))
        else:
            try:
                sampler._raise_failure()
            except Exception as exc:  # the request's failure, reported to the caller
                error = exc
        del sampler  # destroys the job and its device buffers
        return ess, lz, timing, error

    def stats(self):
        s = ServerStats()
        _check(self._lib.is_server_get_stats(self._h, ctypes.byref(s)))
        return {n: int(getattr(s, n)) for n, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.is_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
