"""Autotuner of a context's launch shapes (lelph_tuning), cached per grid.

The reference measures an FFTW plan for every (k, q) pair
(`wfc_plan` with FFTW_MEASURE, src/fft/planner.c:226-312, called from
src/dvloc/compute_Vbare_loc.c:214-215,250-251).  Here the shapes are chosen
once per cell shape: the FFT grid, the precision, bands, modes, spins,
spinors and potential components, and the plane-wave count (in steps of
256).  The knobs:

- split_target: blocks aimed at by the FP64/FP32 contraction's K splits
  (more splits fill the GPU on small cells, fewer keep the partials small);
  it moves the boundaries of the fixed-order FP64 sum, so g changes at the
  rounding level (held to 1e-12 relative in FP64, 1e-7 in FP32);
- fft_lines: lines per block of the compiled line FFTs of the box build
  (bit-identical g);
- fp32_form: FP32 contraction with dV and psi_k staged separately, or
  element-wise (bit-identical g);
- ctile: block tile of the FP64 / FP32 contraction (2 modes x 32 x 64, or
  4 modes x 32 x 32, 16 x 48, 16 x 16; 0 picks by the mode and band counts):
  the padding a tile wastes when the band count is not a multiple of it,
  against the fragment reuse of larger tiles (bit-identical g).

The search is coordinate descent from the defaults, one knob at a time,
each candidate timed on a few pairs of the actual workload with the
library's own HIP-event timings (lelph_last_timings, device time of the
batch), the median of `reps` runs, and checked against the g of the
shape it departs from.  Results go to a JSON cache
(LELPH_TUNE_CACHE, default letzelphc_amd/tune_cache.json).
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "tune_cache.json")

KNOBS = [
    ("split_target", (512, 1024, 2048, 4096)),
    ("fft_lines", (0, 8, 16, 32)),
    ("fp32_form", (0, 1)),
    ("ctile", (0, 1, 2, 3, 4)),
]


def cache_path() -> str:
    return os.environ.get("LELPH_TUNE_CACHE", DEFAULT_CACHE)


def shape_key(ctx) -> str:
    """The cell shape a tuning is valid for."""
    s = ctx.sys
    npw = max(w.npw for w in s.wfcs if w is not None)
    nx, ny, nz = (int(x) for x in s.fft_dims)
    return (f"gfx950|{nx}x{ny}x{nz}|{ctx.effective_precision()}|nb{s.nbnds}|nm{s.nmodes}"
            f"|ns{s.nspin}|nsp{s.nspinor}|nmag{s.nmag}|npw{npw // 256}")


def load_cache(path: str | None = None) -> dict:
    path = path or cache_path()
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_cache(cache: dict, path: str | None = None):
    path = path or cache_path()
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time(ctx, pairs, qpt, reps):
    out = ctx.gkq_batch(pairs, qpt=qpt)  # warm-up (allocations)
    ts = []
    for _ in range(reps):
        out = ctx.gkq_batch(pairs, qpt=qpt, out=out)
        ts.append(ctx.last_timings()["total_ms"])
    return float(np.median(ts)), out.copy()


def _same(g, ref, knob, fp32) -> bool:
    if knob != "split_target":
        return bool(np.array_equal(g, ref))
    den = float(np.max(np.abs(ref))) or 1.0
    return float(np.max(np.abs(g - ref))) / den <= (1e-7 if fp32 else 1e-12)


def tune(ctx, pairs, qpt, sample: int = 3, reps: int = 3, cache: dict | None = None,
         path: str | None = None, force: bool = False) -> dict:
    """Tunes ctx on up to `sample` of its pairs (or applies the cached
    tuning of its shape) and returns {"tuning", "ms", "default_ms",
    "tried", "cached"}."""
    own = cache is None
    if own:
        cache = load_cache(path)
    key = shape_key(ctx)
    if key in cache and not force:
        ctx.set_tuning(**cache[key]["tuning"])
        return dict(cache[key], cached=True)
    sub = list(pairs[:sample])
    best = {"split_target": 0, "fft_lines": 0, "fp32_form": 0, "ctile": 0}
    ctx.set_tuning(**best)
    base_ms, ref = _time(ctx, sub, qpt, reps)
    best_ms, best_g = base_ms, ref
    tried = 1
    fp32 = ctx.effective_precision() == "fp32"
    for name, values in KNOBS:
        if name == "fp32_form" and not fp32:
            continue
        for v in values:
            if v == best[name] or (name == "split_target" and best[name] == 0 and v == 2048):
                continue
            cand = dict(best, **{name: v})
            ctx.set_tuning(**cand)
            ms, g = _time(ctx, sub, qpt, reps)
            tried += 1
            # the split count against the defaults (to rounding); the other
            # knobs against the best shape so far, which they differ from in
            # that knob only (bit for bit)
            if not _same(g, ref if name == "split_target" else best_g, name, fp32):
                raise RuntimeError(f"tuning {cand} changed g on {key}")
            if ms < best_ms:
                best, best_ms, best_g = cand, ms, g
    ctx.set_tuning(**best)
    entry = {"tuning": best, "ms": best_ms, "default_ms": base_ms, "tried": tried,
             "pairs_timed": len(sub)}
    cache[key] = entry
    if own:
        save_cache(cache, path)
    return dict(entry, cached=False)
