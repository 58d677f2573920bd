"""The batching workload: eight uneven cells (FFT 24^3, 32^3, 36^3,
45x45x108, 48^3, 60^3, 64^3, 72^3; 8 to 64 bands, 6 to 36 modes, 16
(k, k+q) pairs each, cells alternately FP64 and FP32) as one multi-cell
batch: per-pair boxes, one persistent work-queue contraction launch per
precision over all tiles, the rest, captured in a HIP graph and replayed.

- every cell's g bit-identical to its own lelph_gkq_batch (same tiles, same
  fixed-order reductions, the same psi boxes taken over along k -> k+q
  chains), after the autotuner has set each grid's shapes, and within
  rounding (1e-12 FP64, 1e-6 FP32) of one lelph_gkq per pair, which
  transforms every box;
- graph replays, with other work and synchronisations between them, each
  equal to the eager pass;
- the first pair of every cell against the FP64 oracle (FP64 cells 1e-10,
  FP32 cells 1e-6);
- the autotuner's candidates reproduce the default g (bit for bit, or to
  rounding for the K-split count).
"""
import os

import numpy as np
import pytest

from letzelphc_amd import autotune, batch, lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def cells(tmp_path_factory):
    cs = synth.config_c3()
    jobs = []
    for i, s in enumerate(cs):
        ctx = lelph.Context(s, "fp64" if i % 2 == 0 else "fp32")
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        jobs.append((ctx, s.qpts[0], synth.kq_pairs(s, 0)))
    cache = str(tmp_path_factory.mktemp("tune") / "tune.json")
    res = batch.tune(jobs, path=cache)
    for (ctx, _, _), r in zip(jobs, res):
        say(f"  tuned {autotune.shape_key(ctx)}: {r['tuning']} {r['ms']:.3f} ms "
            f"(default {r['default_ms']:.3f} ms, {r['tried']} candidates)")
    m = lelph.MultiBatch(jobs)
    yield cs, jobs, m, cache
    m.close()
    for ctx, _, _ in jobs:
        ctx.close()


def test_multi_bit_identical_to_cell_batches(cells):
    cs, jobs, m, _ = cells
    say("case: multi-cell batch vs per-cell batches and one lelph_gkq per pair, "
        "8 cells x 16 pairs")
    gm = [g.copy() for g in m.run()]
    _, gs = batch.run_sequential(jobs)
    _, gp = batch.run_pairwise(jobs)
    say(f"  timings {m.timings()}")
    for i, (a, b, c) in enumerate(zip(gm, gs, gp)):
        e = rel_err(a, c)
        say(f"  cell {i} {cs[i].fft_dims}: vs per-cell batch max |d| = "
            f"{float(np.max(np.abs(a - b))):.3e}, vs per pair {e:.3e}")
        assert np.array_equal(a, b), i
        assert e < (1e-12 if i % 2 == 0 else 1e-6), i


def test_multi_replays(cells):
    cs, jobs, m, _ = cells
    say("case: multi-cell graph replays")
    first = [g.copy() for g in m.run()]
    for r in range(3):
        ctx, q, pairs = jobs[r]
        ctx.gkq(*pairs[0], qpt=q)  # other work on the device between replays
        again = m.run()
        for i, (a, b) in enumerate(zip(again, first)):
            assert np.array_equal(a, b), (r, i)


def test_multi_vs_oracle(cells):
    cs, jobs, m, _ = cells
    g = m.run()
    O.set_threads(1)
    O.set_inner_threads(max(1, min(16, os.cpu_count() or 1)))
    try:
        for i, (s, (ctx, q, pairs)) in enumerate(zip(cs, jobs)):
            o = O.OracleSystem(s)
            dv = o.dv_total(0)
            pr = pairs[0]
            ref = o.add_nonlocal(0, *pr, o.elph_local(0, *pr, dv))
            err = rel_err(g[i][0], ref)
            tol = 1e-10 if i % 2 == 0 else 1e-6
            say(f"  cell {i} {s.fft_dims} nb {s.nbnds} nm {s.nmodes} "
                f"{ctx.effective_precision()}: {err:.3e}")
            assert err < tol, i
    finally:
        O.set_inner_threads(1)


def test_autotune_candidates_bit_identical(cells):
    """a forced re-tune of two cells (one per precision): tune() raises if
    a candidate changes g beyond its bound; the cache records the choice"""
    cs, jobs, _, cache = cells
    c = autotune.load_cache(cache)
    for i in (2, 3):
        ctx, q, pairs = jobs[i]
        r = autotune.tune(ctx, pairs, q, cache=c, force=True)
        say(f"  cell {i}: {r}")
        assert autotune.shape_key(ctx) in c
