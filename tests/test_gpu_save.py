"""The System read back from Yambo SAVE + QE ph_save files drives the HIP
path: g for every k of the q against the FP64 oracle on the same System."""
import numpy as np
import pytest

from letzelphc_amd import lelph, save, synth
from oracle import oracle as O

from .test_save import _write_both
from .util import rel_err, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["cubic", "soc"])
def test_compute_elphq_from_save(tmp_path, kind):
    kw = dict(syms="cubic", q=(0.0, 0.5, 0.0)) if kind == "cubic" else \
        dict(nspinor=2, nmag=1, soc=True)
    s = synth.tiny(**kw)
    _write_both(tmp_path, s, mag=None if kind == "cubic" else "noncolin")
    r, ph = save.read_save(tmp_path / "SAVE", tmp_path / "ph_save")
    r._eig = {i: s.eigvec(0) for i in range(len(r.qpts))}
    iq = int(np.argmax([np.abs(q).sum() for q in r.qpts]))
    r._dv = {iq: synth.System.dvscf(r, iq)}
    say(f"case: compute_elphq from SAVE, {kind}, q {r.qpts[iq]}")
    o = O.OracleSystem(r)
    dv = o.dv_total(iq)
    ctx = lelph.Context(r, "fp64")
    try:
        ctx.set_dvscf(r.qpts[iq], r.eigvec(iq), r.dvscf(iq), kernel="dfpt")
        assert rel_err(ctx.get_dvscf(), dv) < 1e-10
        g = ctx.compute_elphq(r.qpts[iq], non_loc=True)
    finally:
        ctx.close()
    ref = o.compute_elphq(iq, dv, non_loc=True)
    assert rel_err(g, ref) < 1e-10
