"""Batched AMIS: many independent posteriors prepared together and run in one
persistent, graph-captured launch on the MI355X.

A batch is a vector of prepared samplers, each of which means
``importance_sample!(prepare_sampler(rng, target, AMIS(q0; rounds, round_size)))``
(src/methods/amis.jl:42-47, src/adaptive_gaussian_execution.jl:897-1112) for a
built-in regression target; every job keeps its own counter-RNG key, so its
results are those of the same sampler run alone.  The reference runs such jobs
one prepared sampler at a time; the batch is the engine's launch form for them.
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib
from ._lib import F64P, I32P, VP, JobDesc
from .api import (AMIS, AMISRoundError, AllZeroWeightsError, FACTOR, FP32, FP64, NativeProposal, PHASES,
                  SamplerExecutionError, WeightedSamples, _check, _job_desc, _p, context)


class BatchRoundErrors(RuntimeError):
    """Raised by importance_sample_batch_ when jobs failed: maps job index to the
    AMISRoundError the job would raise on its own; the other jobs' results stand."""

    def __init__(self, errors):
        super().__init__(f"{len(errors)} job(s) of the batch failed: {sorted(errors)[:8]}")
        self.errors = errors


class PreparedBatch:
    def __init__(self, rngs, targets, algorithms, precision=FP32):
        n = len(algorithms)
        if not (len(rngs) == len(targets) == n) or n == 0:
            raise ValueError("a batch needs one rng, target and algorithm per job")
        for a in algorithms:
            if not isinstance(a, AMIS) or type(a) is not AMIS:
                raise TypeError("the batch runs AMIS jobs")
        self.precision = precision
        self.algorithms = list(algorithms)
        self.targets = list(targets)
        lib = _lib.load()
        ctx = context()
        descs = (JobDesc * n)()
        self._keep = []
        self.dims, self.schedules, self.proposals0 = [], [], []
        for j, (rng, tgt, alg) in enumerate(zip(rngs, targets, algorithms)):
            desc, keep, p, sched = _job_desc(rng, tgt, alg, precision)
            descs[j] = desc
            self._keep.append(keep)
            self.dims.append(p.dim)
            self.schedules.append(list(sched))
            self.proposals0.append(p)
        h = VP()
        _check(lib.is_batch_create(ctx, descs, n, precision, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.njobs = n

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.is_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ---- launch shape
    def set_tuning(self, waves, blocks_per_cu):
        _check(self._lib.is_batch_set_tuning(self._h, int(waves), int(blocks_per_cu)))

    @property
    def tuning(self):
        w, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.is_batch_get_tuning(self._h, ctypes.byref(w), ctypes.byref(b)))
        return int(w.value), int(b.value)

    def set_order(self, order=None):
        if order is None:
            _check(self._lib.is_batch_set_order(self._h, None))
            return
        o = np.ascontiguousarray(np.asarray(order, dtype=np.int32))
        if o.size != self.njobs:
            raise ValueError("order must list every job once")
        _check(self._lib.is_batch_set_order(self._h, o.ctypes.data_as(I32P)))

    @property
    def order(self):
        o = np.zeros(self.njobs, dtype=np.int32)
        _check(self._lib.is_batch_get_order(self._h, o.ctypes.data_as(I32P)))
        return o

    def trial(self, waves, blocks_per_cu):
        """One event-timed run with this launch shape and no commit; returns the
        persistent kernel's milliseconds (raises for a shape that does not fit)."""
        ms = ctypes.c_double()
        rc = self._lib.is_batch_trial(self._h, int(waves), int(blocks_per_cu), ctypes.byref(ms))
        if rc not in (_lib.IS_OK, _lib.IS_EFAILED):
            _check(rc)
        return float(ms.value)

    def set_timing(self, enable=True):
        _check(self._lib.is_batch_set_timing(self._h, int(bool(enable))))

    @property
    def kernel_ms(self):
        ms = ctypes.c_double()
        _check(self._lib.is_batch_kernel_ms(self._h, ctypes.byref(ms)))
        return float(ms.value)

    # ---- execution
    def run(self):
        """Runs every job; returns the per-job status (0 ok, else the failing round)."""
        rc = self._lib.is_batch_run(self._h)
        if rc not in (_lib.IS_OK, _lib.IS_EFAILED):
            _check(rc)
        return self.status()

    def status(self):
        st = np.zeros(self.njobs, dtype=np.int32)
        _check(self._lib.is_batch_status(self._h, st.ctypes.data_as(I32P)))
        return st

    def failure(self, job):
        f = _lib.Failure()
        _check(self._lib.is_batch_job_failure(self._h, int(job), ctypes.byref(f)))
        return f

    def round_error(self, job):
        f = self.failure(job)
        if f.round == 0:
            return None
        phase = PHASES.get(f.phase)
        if f.phase in (5, 7):
            cause = AllZeroWeightsError()
        elif f.phase == 6:
            cause = np.linalg.LinAlgError("fitted covariance is not positive definite")
        else:
            index = f.draw_first_logical_index or f.first_logical_index
            cause = SamplerExecutionError(phase, index, f.draw_reason_bits or f.reason_bits)
        return AMISRoundError(f.round, phase, cause, {"failure_count": int(f.count), "job": int(job)})

    # ---- results
    def final_summary(self):
        """(ESS, log Z) of every job's final round (the AMIS estimator over all rounds)."""
        ess, lz = np.zeros(self.njobs), np.zeros(self.njobs)
        _check(self._lib.is_batch_final_summary(self._h, _p(ess), _p(lz)))
        return ess, lz

    def lognormalizers(self):
        return self.final_summary()[1]

    def round_summary(self, job):
        T = len(self.schedules[job])
        ess, lz = np.zeros(T), np.zeros(T)
        _check(self._lib.is_batch_job_round_summary(self._h, int(job), _p(ess), _p(lz)))
        return ess, lz

    def result(self, job):
        """The job's WeightedSamples (copied to the host)."""
        job = int(job)
        d, sched = self.dims[job], self.schedules[job]
        n = int(self._lib.is_batch_job_total(self._h, job))
        x = np.zeros(n * d)
        _check(self._lib.is_batch_job_samples(self._h, job, _p(x)))
        lw = np.zeros(n)
        _check(self._lib.is_batch_job_logweights(self._h, job, _p(lw)))
        ess, lz = self.round_summary(job)
        rounds = np.repeat(np.arange(1, len(sched) + 1), sched)
        diag = {"method": "amis", "execution": "mi355x-batch", "rounds": len(sched), "round_sizes": list(sched),
                "round_ess": ess, "round_lognormalizers": lz, "target_evaluations": n,
                "proposal_evaluations": len(sched) * n,
                "precision": "fp32" if self.precision == FP32 else "fp64", "job": job}
        res = WeightedSamples(x.reshape(n, d), lw, {"round": rounds}, diag)
        res._lognormalizer = float(lz[-1])
        return res

    def history(self, job):
        T, d = len(self.schedules[job]), self.dims[job]
        m, f, l = np.zeros((T + 1) * d), np.zeros((T + 1) * d * d), np.zeros(T + 1)
        _check(self._lib.is_batch_job_history(self._h, int(job), _p(m), _p(f), _p(l)))
        return m.reshape(T + 1, d), f.reshape(T + 1, d, d).transpose(0, 2, 1).copy(), l

    def current_proposal(self, job):
        """The job's committed proposal (src/adaptive_gaussian.jl:382-412)."""
        d = self.dims[job]
        loc, fac, ln = np.zeros(d), np.zeros(d * d), np.zeros(1)
        _check(self._lib.is_batch_job_proposal(self._h, int(job), _p(loc), _p(fac), _p(ln)))
        p0 = self.proposals0[job]
        return NativeProposal(p0.family, FACTOR, loc, fac.reshape(d, d).T.copy().ravel(), float(ln[0]), p0.dof)

    def export_randomness(self, job, round):
        """The counter-RNG values the next run consumes in `round` of `job`."""
        d, n, p = self.dims[job], self.schedules[job][round], self.proposals0[job]
        z = np.zeros(d * n)
        _check(self._lib.is_batch_export_normals(self._h, int(job), int(round), _p(z), z.size))
        out = {"normals": z}
        if p.family == 1:
            rs = 1 if p.dof == 1.0 else 8
            us = 0 if p.dof == 1.0 else 8 + (1 if p.dof < 2 else 0)
            rn, ru = np.zeros(rs * n), np.zeros(us * n)
            _check(self._lib.is_batch_export_radial(self._h, int(job), int(round), _p(rn), rn.size,
                                                    _p(ru) if us else None, ru.size))
            out["radial_normals"] = rn
            if us:
                out["radial_uniforms"] = ru
        return out


def prepare_batch(rngs, targets, algorithms, precision="fp32"):
    """Prepares one batch of AMIS samplers (one job per (rng, target, algorithm))."""
    prec = {"fp64": FP64, "fp32": FP32, FP64: FP64, FP32: FP32}[precision]
    return PreparedBatch(rngs, targets, algorithms, prec)


def importance_sample_batch_(batch: PreparedBatch):
    """`importance_sample!` over every job of the batch in one launch; raises
    BatchRoundErrors listing the failed jobs (the others' results stand)."""
    st = batch.run()
    bad = np.nonzero(st)[0]
    if bad.size:
        raise BatchRoundErrors({int(j): batch.round_error(j) for j in bad})
    return batch


# ---------------------------------------------------------------- autotuner
def _signature(batch):
    """Workload key of the tuning cache: device, precision, job count, and the
    histograms of padded dimension, data-row tiles and samples per job."""
    dims = np.bincount((np.asarray(batch.dims) + 3) // 4, minlength=9).tolist()
    rows = []
    for t in batch.targets:
        n = np.asarray(t.X).shape[0]
        rows.append(int(np.ceil(np.log2(max(16, n)))))
    rows = np.bincount(rows, minlength=16).tolist()
    work = sorted({(len(s), s[0]) for s in batch.schedules})
    return json.dumps({"arch": "gfx950", "precision": int(batch.precision), "jobs": batch.njobs, "dims": dims,
                       "rows": rows, "work": work}, sort_keys=True)


def autotune(batch: PreparedBatch, cache=None, candidates=None, repeats=2):
    """Picks the batch's launch shape by timing candidates on the device.

    Candidates are (waves per workgroup, workgroups per CU, job order) with the
    orders "lpt" (longest estimated job first) and "input"; each is run `repeats`
    times as a no-commit trial and scored by its fastest persistent-kernel time.
    The winner is applied to the batch and, with `cache` (a JSON file path),
    remembered under the workload signature so later batches of the same shape
    skip the search.  Returns {"waves", "blocks_per_cu", "order", "ms", "cached"}."""
    key = _signature(batch)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        best = table[key]
        _apply(batch, best)
        return dict(best, cached=True)
    if candidates is None:
        candidates = [(w, b, o) for o in ("lpt", "input") for w in (2, 4, 8) for b in (1, 2, 3, 4)]
    best = None
    for w, b, o in candidates:
        _apply(batch, {"order": o})
        try:
            ms = min(batch.trial(w, b) for _ in range(repeats))
        except ValueError:
            continue  # launch shape does not fit (LDS or waves per CU)
        if best is None or ms < best["ms"]:
            best = {"waves": w, "blocks_per_cu": b, "order": o, "ms": ms}
    if best is None:
        raise RuntimeError("no launch shape fits this batch")
    _apply(batch, best)
    if cache:
        table[key] = best
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return dict(best, cached=False)


def _apply(batch, choice):
    if "order" in choice:
        batch.set_order(None if choice["order"] == "lpt" else np.arange(batch.njobs))
    if "waves" in choice:
        batch.set_tuning(choice["waves"], choice["blocks_per_cu"])
