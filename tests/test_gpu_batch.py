"""Uneven cells side by side (letzelphc_amd/batch.py): the concurrent run,
one context and stream per cell, gives bit for bit the per-cell sequential
results, and the FP64 cells match the oracle."""
import numpy as np
import pytest

from letzelphc_amd import batch, lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu


def test_uneven_cells_concurrent():
    cells = synth.config_c3(cells=[0, 1, 2, 3])
    say(f"case: {len(cells)} uneven cells concurrently")
    ctxs, jobs = [], []
    try:
        for i, s in enumerate(cells):
            ctx = lelph.Context(s, "fp64" if i % 2 == 0 else "fp32")
            ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
            ctxs.append(ctx)
            jobs.append((ctx, s.qpts[0], synth.kq_pairs(s, 0)[:4]))
        _, seq = batch.run_sequential(jobs)
        _, con = batch.run_concurrent(jobs)
        for i in range(len(jobs)):
            assert np.array_equal(seq[i], con[i]), i
        s = cells[2]
        o = O.OracleSystem(s)
        dv = o.dv_total(0)
        for j, pr in enumerate(jobs[2][2][:2]):
            ref = o.add_nonlocal(0, *pr, o.elph_local(0, *pr, dv))
            assert rel_err(con[2][j], ref) < 1e-10, j
    finally:
        for c in ctxs:
            c.close()
