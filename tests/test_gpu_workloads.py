"""GPU parity at the full workload shapes, against the FP64 CPU oracle run
on the host cores in the same process.

- The 16-atom cubic cell (FFT 96^3, 48 modes, 128 bands, npw ~36,800): a
  chain of three pairs k -> k+q -> k+2q -> k+3q through one batch, so the
  second pair's k box is the first pair's k+q box and the third pair's k box
  is the second's k+q box shifted back by its umklapp vector.  This is the
  production tiling of the contraction (27,648 chunks of 32 points, the
  producer/consumer bf16x3 kernel with 8 K splits and 512-point FP32
  chains).  The oracle runs the last two pairs at all 48 modes
  (about 30 s each on 16 threads) on a worker thread while the GPU runs.
- The layered cell at full size (45x45x216, 40 bands, 12x12x1 k, 12
  operations with time reversal): pairs whose k and k+q need the identity,
  rotations and time-reversed operations, in both the k+q and k-q
  conventions, one pair by one and as a chained batch.

Bars: FP64 |dg|max/|g|max <= 1e-10, FP32 / bf16x6 / bf16x3 <= 1e-6.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say, tr_partner

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

TOL64 = 1e-10
TOL32 = 1e-6
PRECS = ("fp64", "fp32", "bf16x6", "bf16x3")


def _cores():
    return max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def c2():
    s = synth.config("c2", with_wfc=False)
    allp = synth.kq_pairs(s, 0)
    # full-BZ k 0 -> 20 -> 40 -> 60 along q = (1/4, 1/4, 0); 20 -> 40 crosses
    # the zone boundary (k+q = (1/2, 1/2, 0) is the grid point (-1/2, -1/2, 0))
    chain = [allp[0], allp[20], allp[40]]
    assert chain[1][1] == chain[0][0] and chain[2][1] == chain[1][0]
    need = sorted({p[0] for p in chain} | {p[1] for p in chain})
    synth.make_wfcs(s, workers=4, only=need)
    say(f"case: c2 chain {chain}, npw {[s.wfcs[k].npw for k in need]}")
    o = O.OracleSystem(s)
    O.set_threads(1)
    O.set_inner_threads(_cores())
    dv = o.dv_total(0)
    ex = ThreadPoolExecutor(1)
    futs = {i: ex.submit(lambda p=chain[i]: o.add_nonlocal(0, *p, o.elph_local(0, *p, dv)))
            for i in (1, 2)}
    ctx = lelph.Context(s, "fp64")
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
    res = {}
    try:
        dv_err = rel_err(ctx.get_dvscf(), dv)
        for prec in PRECS:
            ctx.set_precision(prec)
            res[prec] = (ctx.effective_precision(),
                         ctx.gkq_batch(chain, non_loc=True, qpt=s.qpts[0]))
            say(f"  c2 batch {prec} done")
    finally:
        ctx.close()
    yield s, chain, futs, res, dv_err
    ex.shutdown(wait=True)
    O.set_inner_threads(1)


def test_c2_dvscf(c2):
    _, _, _, _, dv_err = c2
    say(f"case: c2 dV_q (dvscf + bare local), |d|/|dv|max = {dv_err:.3e}")
    assert dv_err < TOL64


@pytest.mark.parametrize("i", [1, 2])
def test_c2_chain_vs_oracle(c2, i):
    s, chain, futs, res, _ = c2
    say(f"case: c2 pair {chain[i]} ({'umklapp k+q' if i == 1 else 'shifted box reuse'})")
    ref = futs[i].result()
    for prec in PRECS:
        eff, g = res[prec]
        err = rel_err(g[i], ref)
        say(f"  c2 {prec} (runs as {eff}): |dg|/|g|max = {err:.3e}")
        assert eff == prec
        assert err < (TOL64 if prec == "fp64" else TOL32), prec


def test_c2_first_pair_consistent(c2):
    """The first pair of the batch (both boxes transformed) against the
    FP64 result of the same batch: the reduced forms stay at 1e-6."""
    _, _, _, res, _ = c2
    g64 = res["fp64"][1]
    for prec in PRECS[1:]:
        err = rel_err(res[prec][1][0], g64[0])
        say(f"  c2 first pair {prec} vs FP64: {err:.3e}")
        assert err < TOL32


def _c1_pick(s, iq, kminusq):
    """One pair per (k+q rotated?, k rotated?) class of the pair list, plus
    the first non-trivial pair again with its k side, and then its k+q
    side, reached through a time-reversed operation."""
    pick, seen = [], set()
    for pr in synth.kq_pairs(s, iq, kminusq=kminusq):
        key = (pr[2] > 0, pr[3] > 0)
        if key not in seen:
            seen.add(key)
            pick.append(pr)
    ikq, ik, kqsym, ksym = next(p for p in pick if p[2] > 0 or p[3] > 0)
    jk, jkq = tr_partner(s, ik, ksym), tr_partner(s, ikq, kqsym)
    assert jk is not None and jkq is not None
    pick += [(ikq, ik, kqsym, jk), (ikq, ik, jkq, ksym)]
    return pick


@pytest.fixture(scope="module")
def c1():
    s = synth.config("c1")
    o = O.OracleSystem(s)
    O.set_threads(1)
    O.set_inner_threads(_cores())
    yield s, o
    O.set_inner_threads(1)


@pytest.mark.parametrize("iq,kminusq", [(0, False), (0, True), (1, False)])
def test_c1_full(c1, iq, kminusq):
    s, o = c1
    pick = _c1_pick(s, iq, kminusq)
    say(f"case: c1 full size, q {iq}, kminusq={kminusq}, pairs {pick}")
    assert len(pick) >= 5
    dv = o.dv_total(iq)
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[iq], s.eigvec(iq), s.dvscf(iq), kernel="dfpt")
        assert rel_err(ctx.get_dvscf(), dv) < TOL64
        refs = [o.add_nonlocal(iq, *pr, o.elph_local(iq, *pr, dv)) for pr in pick]
        for prec in PRECS:
            ctx.set_precision(prec)
            eff = ctx.effective_precision()
            for pr, ref in zip(pick, refs):
                g = ctx.gkq(*pr, qpt=s.qpts[iq])
                err = rel_err(g, ref)
                say(f"  {pr} {prec} (runs as {eff}): {err:.3e}")
                assert err < (TOL64 if prec == "fp64" else TOL32), (pr, prec)
    finally:
        ctx.close()


def test_c1_batch_chain(c1):
    """All 144 pairs of q 0 in one batch (the batch orders them along
    k -> k+q chains and reuses boxes across symmetry-rotated sides) against
    the oracle's k loop on the first 12 full-BZ k."""
    s, o = c1
    say("case: c1 full size, batch of all pairs")
    dv = o.dv_total(0)
    ref = o.compute_elphq(0, dv, non_loc=True, k_first=0, k_count=12)
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        for prec in ("fp64", "bf16x3"):
            ctx.set_precision(prec)
            g = ctx.gkq_batch(synth.kq_pairs(s, 0), non_loc=True, qpt=s.qpts[0])
            err = rel_err(g[:12], ref)
            say(f"  c1 batch {prec}: {err:.3e}")
            assert err < (TOL64 if prec == "fp64" else TOL32)
    finally:
        ctx.close()
