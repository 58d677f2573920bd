"""The multi-cell batch on mixed cells: an FP32 spinor cell with the 2x2
potential (element-wise staging) and an FP32 collinear cell (separate
staging) in the same work queue, an FP64 two-spin cell, a cell with no
pairs, and a one-pair FP32 cell with its own contraction tile (a second
FP32 queue).  Every cell's g is bit-identical to its own lelph_gkq_batch,
replay after replay."""
import numpy as np
import pytest

from letzelphc_amd import lelph, synth

from .util import say

pytestmark = pytest.mark.gpu

CELLS = [
    ("spinor_nmag4_fp32", dict(nbnds=6, nspinor=2, nmag=4), "fp32", None, 0),
    ("collinear_fp32", dict(nbnds=8), "fp32", None, 0),
    ("spin2_fp64", dict(nbnds=10, nspin=2, nmag=2), "fp64", None, 0),
    ("empty_fp64", dict(nbnds=4), "fp64", 0, 0),
    ("one_pair_fp32_tile32", dict(nbnds=24, ecut=14.0, fft=(16, 16, 16)), "fp32", 1, 2),
]


def test_multi_mixed_cells():
    jobs, ctxs = [], []
    try:
        for name, kw, prec, npair, ctile in CELLS:
            s = synth.tiny(**kw)
            ctx = lelph.Context(s, prec)
            ctxs.append(ctx)
            ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
            if ctile:
                ctx.set_tuning(ctile=ctile)
            pairs = synth.kq_pairs(s, 0)
            if npair is not None:
                pairs = pairs[:npair]
            jobs.append((ctx, s.qpts[0], pairs))
        say("case: multi-cell batch on mixed cells: " + ", ".join(
            f"{c[0]} ({len(j[2])} pairs)" for c, j in zip(CELLS, jobs)))
        m = lelph.MultiBatch(jobs)
        try:
            t = m.timings()
            say(f"  items fp64 {t['items_fp64']}, fp32 {t['items_fp32']}")
            refs = [ctx.gkq_batch(p, qpt=q) if len(p) else None for ctx, q, p in jobs]
            for rep in range(3):
                out = m.run()
                for (name, *_), g, r in zip(CELLS, out, refs):
                    if r is None:
                        assert g.shape[0] == 0, name
                        continue
                    assert np.array_equal(g, r), (rep, name)
        finally:
            m.close()
    finally:
        for ctx in ctxs:
            ctx.close()
