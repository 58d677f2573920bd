"""Uneven batches of cells on one GPU.

A job is (Context, qpt, pairs): one cell (its own FFT grid, plane-wave,
band and mode counts, FP64 or FP32) and the (k, k+q) pairs of one q.

- `run_multi`: every job as one unit through lelph_multi - per-pair boxes,
  one persistent work-queue contraction launch per precision over the
  tiles of all pairs of all cells (largest first), the reductions and
  non-local parts - captured in a HIP graph once and replayed
  (letzelphc_amd.lelph.MultiBatch).  Each context first gets the launch
  shapes the autotuner picked for its grid (letzelphc_amd.autotune).
- `run_sequential`: the per-cell baseline, one lelph_gkq_batch per cell,
  one cell after another.
- `run_pairwise`: one lelph_gkq per pair, the launches the multi batch
  reproduces bit for bit.
"""
from __future__ import annotations

import time

import numpy as np

from . import autotune
from .lelph import MultiBatch


def tune(jobs, path=None, force=False):
    """Autotunes (or applies the cached tuning of) every job's context;
    returns the per-cell results."""
    cache = autotune.load_cache(path)
    out = [autotune.tune(ctx, pairs, q, cache=cache, force=force) for ctx, q, pairs in jobs]
    autotune.save_cache(cache, path)
    return out


def run_sequential(jobs):
    """jobs: [(ctx, qpt, pairs)]; returns (seconds, [g])"""
    t0 = time.perf_counter()
    out = [ctx.gkq_batch(pairs, qpt=q) for ctx, q, pairs in jobs]
    return time.perf_counter() - t0, out


def run_pairwise(jobs):
    """one lelph_gkq per pair; returns (seconds, [g])"""
    t0 = time.perf_counter()
    out = [np.stack([ctx.gkq(*p, qpt=q) for p in pairs]) for ctx, q, pairs in jobs]
    return time.perf_counter() - t0, out


def run_multi(jobs, reps=1, multi=None):
    """the jobs as one captured multi-cell batch (built unless given);
    returns (seconds per replay, [g], the MultiBatch)"""
    m = multi or MultiBatch(jobs)
    out = m.run()  # first replay
    t0 = time.perf_counter()
    for _ in range(reps):
        out = m.run()
    return (time.perf_counter() - t0) / reps, [o.copy() for o in out], m


def g_elements(jobs):
    return sum(len(p) * ctx.sys.nmodes * ctx.sys.nspin * ctx.sys.nbnds ** 2
               for ctx, _, p in jobs)
