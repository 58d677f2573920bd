"""Uneven cells side by side on one GPU.

Several (k, q) workloads of different cells (FFT grids, plane-wave counts,
bands, modes, precisions) run at once, one context per cell: each context
owns its own non-blocking HIP stream, and a host thread per cell issues its
batch (`lelph_gkq_batch`; ctypes releases the interpreter lock for the
call), so small cells' kernels overlap instead of queueing one cell after
another.  `run_sequential` is the per-cell baseline.
"""
from __future__ import annotations

import threading
import time

import numpy as np


def run_sequential(jobs):
    """jobs: [(ctx, qpt, pairs)]; returns (seconds, [g])"""
    t0 = time.perf_counter()
    out = [ctx.gkq_batch(pairs, qpt=q) for ctx, q, pairs in jobs]
    return time.perf_counter() - t0, out


def run_concurrent(jobs):
    """the same jobs, one host thread (and so one stream) per cell"""
    out = [None] * len(jobs)
    err = [None] * len(jobs)
    start = threading.Barrier(len(jobs) + 1)

    def work(i, ctx, q, pairs):
        start.wait()
        try:
            out[i] = ctx.gkq_batch(pairs, qpt=q)
        except Exception as e:  # re-raised in the caller
            err[i] = e

    th = [threading.Thread(target=work, args=(i, *j)) for i, j in enumerate(jobs)]
    for t in th:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    for e in err:
        if e is not None:
            raise e
    return dt, out


def g_elements(jobs):
    return sum(len(p) * ctx.sys.nmodes * ctx.sys.nspin * ctx.sys.nbnds ** 2
               for ctx, _, p in jobs)
