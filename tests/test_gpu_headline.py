"""GPU parity of the producer/consumer bf16x3 contraction, the kernel the
benchmark measures, against the FP64 CPU oracle through the C-ABI.

That kernel runs on the aligned path (one spinor, diagonal potential, FFT
size a multiple of 32) with at least 64 bands; smaller band counts take the
one-wave-per-tile split kernel.  The cases cover what its tiling pads: a
mode count that is not a multiple of four (the last group of four modes
has two padding modes whose results are not stored), band counts that are
not multiples of 16 (rows) or 128 (columns, two column tiles at 136 bands),
and two spins with a spin-resolved potential.

Bars: bf16x3 |dg|max/|g|max <= 1e-6 against the oracle; the FP64 path of the
same context <= 1e-10.
"""
import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL_BF16X3 = 1e-6

CASES = {
    "nm6_nb80": dict(nbnds=80, ecut=14.0, fft=(16, 16, 16)),
    "spin2_nb72": dict(nbnds=72, nspin=2, nmag=2, ecut=14.0, fft=(16, 16, 16)),
    "nb136_two_column_tiles": dict(nbnds=136, ecut=20.0, fft=(24, 24, 24)),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kminusq", [False, True])
def test_headline_bf16x3(case, kminusq):
    s = synth.tiny(**CASES[case])
    n = int(np.prod(s.fft_dims))
    # the producer/consumer kernel's conditions
    assert s.nspinor == 1 and s.nmag in (1, 2) and n % 32 == 0 and s.nbnds >= 64
    if case == "nm6_nb80":
        assert s.nmodes % 4 != 0  # a padded mode group
    o = O.OracleSystem(s)
    O.set_threads(8)
    dv = o.dv_total(0)
    ref = o.compute_elphq(0, dv, non_loc=True, kminusq=kminusq)
    say(f"case: headline bf16x3 {case} kminusq={kminusq} nb={s.nbnds} nm={s.nmodes} "
        f"nspin={s.nspin} N={n}")
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        g64 = ctx.compute_elphq(s.qpts[0], non_loc=True, kminusq=kminusq)
        ctx.set_precision("bf16x3")
        assert ctx.effective_precision() == "bf16x3"
        g3 = ctx.compute_elphq(s.qpts[0], non_loc=True, kminusq=kminusq)
    finally:
        ctx.close()
    e64, e3 = rel_err(g64, ref), rel_err(g3, ref)
    say(f"  fp64 {e64:.3e}  bf16x3 {e3:.3e}")
    assert e64 < TOL64
    assert e3 < TOL_BF16X3
