"""The Kleinman-Bylander part in the reduced precisions: the <beta|psi>
projector GEMMs in FP32 (FP32 MFMA; the default of the bf16 contexts too)
and BF16-split (tuning nl_form 1: bf16x3 / bf16x6 contexts on BF16 MFMA,
both with 3 pieces and 6 products), beta's phases and radial / Y_lm tables
and the K/K' sums and mode rotation in FP64.  The non-local part alone
against the oracle's, and the proof that the reduced path ran (its error
is above the FP64 path's); a bf16 context's default non-local part is the
FP32 context's bit for bit.  Cells: 80 bands (bf16x3 runs as bf16x3), 8
bands (bf16x3 runs as bf16x6), spinors."""
import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu

CASES = {
    "nb80": dict(nbnds=80, ecut=14.0, fft=(16, 16, 16)),
    "nb8": dict(nbnds=8),
    "spinor": dict(nbnds=4, nspinor=2, nmag=1),
    "cubic_syms": dict(nbnds=6, syms="cubic", q=(0.0, 0.5, 0.0)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_nonlocal_reduced(case):
    s = synth.tiny(**CASES[case])
    o = O.OracleSystem(s)
    pairs = synth.kq_pairs(s, 0)[:4]
    z = np.zeros((s.nmodes, s.nspin, s.nbnds, s.nbnds), np.complex128)
    refs = [o.add_nonlocal(0, *p, z) for p in pairs]
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        errs, gs = {}, {}
        for prec, form in (("fp64", 0), ("fp32", 0), ("bf16x6", 0), ("bf16x3", 0),
                           ("bf16x6", 1), ("bf16x3", 1)):
            ctx.set_precision(prec)
            ctx.set_tuning(nl_form=form)
            g = [ctx.add_nonlocal(*p, z) for p in pairs]
            key = f"{prec}/nl{form}"
            gs[key] = g
            errs[key] = (ctx.effective_precision(), max(rel_err(a, r) for a, r in zip(g, refs)))
        ctx.set_tuning()
        say(f"case: non-local part alone, {case}: " +
            ", ".join(f"{k} (as {v[0]}) {v[1]:.2e}" for k, v in errs.items()))
        assert errs["fp64/nl0"][1] < 1e-10
        for key, (_, e) in errs.items():
            if key.startswith("fp64"):
                continue
            assert e < 1e-6, key
            assert e > 1e-13, key  # the reduced path ran
        for prec in ("bf16x6", "bf16x3"):
            assert all(np.array_equal(a, b) for a, b in zip(gs[f"{prec}/nl0"], gs["fp32/nl0"]))
            assert not all(np.array_equal(a, b)
                           for a, b in zip(gs[f"{prec}/nl1"], gs["fp32/nl0"]))
    finally:
        ctx.close()
