"""Block tile shapes of the FP64 / FP32 contraction (lelph_tuning.ctile):
two modes x 32 x 64 (1), four modes x 32 x 32 (2), x 16 x 48 (3) and x 16 x
16 (4), and the automatic choice (0).  Every shape issues the same MFMAs on
the same operands in the same order for each g element and keeps the K
splits of the wide tile, so g is bit-identical across shapes; each against
the FP64 oracle (FP64 <= 1e-10, FP32 <= 1e-6).  Band counts that pad the
tiles differently (8, 24, 40, 80), a spinor cell with the 2x2 potential
(element-wise FP32 form) and a two-spin cell."""
import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu

CASES = {
    "nb8": dict(nbnds=8),
    "nb24": dict(nbnds=24, ecut=14.0, fft=(16, 16, 16)),
    "nb40": dict(nbnds=40, ecut=14.0, fft=(16, 16, 16)),
    "nb80": dict(nbnds=80, ecut=14.0, fft=(16, 16, 16)),
    "spinor_nmag4": dict(nbnds=6, nspinor=2, nmag=4),
    "spin2": dict(nbnds=10, nspin=2, nmag=2),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tile_shapes(case, prec):
    s = synth.tiny(**CASES[case])
    o = O.OracleSystem(s)
    dv = o.dv_total(0)
    pairs = synth.kq_pairs(s, 0)[:3]
    ref = [o.elph_local(0, *p, dv) for p in pairs]
    ctx = lelph.Context(s, prec)
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        gs = {}
        for ct in (0, 1, 2, 3, 4):
            ctx.set_tuning(ctile=ct)
            gs[ct] = np.stack([ctx.elph_local(*p) for p in pairs])
        err = max(rel_err(g, r) for g, r in zip(gs[0], ref))
        say(f"case: tile shapes {case} {prec}: vs oracle {err:.2e}; "
            + ", ".join(f"ctile {k} max|d| {float(np.max(np.abs(v - gs[1]))):.1e}"
                        for k, v in gs.items()))
        for k, v in gs.items():
            assert np.array_equal(v, gs[1]), k
        assert err < (1e-10 if prec == "fp64" else 1e-6)
    finally:
        ctx.close()
