"""g computed on the GPU, written as ndb.elph through liblelph_nc4.so, read
back with h5dump and compared with the oracle's g."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from letzelphc_amd import lelph, ndb, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(ndb.LIB_PATH) or
                                 not (shutil.which("h5dump") or
                                      os.path.exists("/opt/conda/bin/h5dump")),
                                 reason="liblelph_nc4.so or h5dump missing")]


def test_ndb_elph_from_gpu(tmp_path):
    s = synth.tiny(nbnds=4)
    ctx = lelph.Context(s, "fp64")
    path = str(tmp_path / "ndb.elph")
    say("case: ndb.elph from the GPU path")
    try:
        ndb.write_elph_db(ctx, path)
    finally:
        ctx.close()
    h5 = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = str(tmp_path / "g.bin")
    subprocess.run([h5, "-d", "/elph_mat", "-b", "LE", "-o", out, path], check=True,
                   capture_output=True)
    nk = len(s.kpt_bz_crys)
    raw = np.fromfile(out, np.float64).reshape(1, nk, s.nmodes, s.nspin, s.nbnds, s.nbnds, 2)
    g = raw[0, ..., 0] + 1j * raw[0, ..., 1]
    o = O.OracleSystem(s)
    ref = o.compute_elphq(0, o.dv_total(0))
    assert rel_err(g, ref) < 1e-10
