"""The C-ABI from a C caller: tests/c_abi/boundary.c (compiled with gcc
against include/lelph_hip.h, linked to liblelph_hip.so) runs the sequence a
reference-side caller would - lelph_create, lelph_upload_wfc per iBZ k,
lelph_set_dvscf, lelph_gkq per pair, lelph_gkq_batch - on a seeded cell
whose pairs take rotated and time-reversed sides; g of both calls against
the FP64 oracle at 1e-10."""
import os
import struct
import subprocess

import numpy as np
import pytest

from letzelphc_amd import synth
from oracle import oracle as O

from .util import rel_err, say, tr_partner

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c_abi", "lelph_boundary")


def _rec(f, name, arr):
    b = np.ascontiguousarray(arr).tobytes()
    f.write(name.encode().ljust(16, b"\0")[:16])
    f.write(struct.pack("<q", len(b)))
    f.write(b)


def _write_inputs(path, s, pairs):
    ps = s.pseudo
    with open(path, "wb") as f:
        _rec(f, "lat_ints", np.array([s.natom, len(s.syms), s.timerev, s.nspin, s.nspinor,
                                      s.nbnds, s.nmodes, s.nmag, ord(s.dimension), 0], np.int64))
        _rec(f, "fft_dims", np.array(s.fft_dims, np.int64))
        _rec(f, "alat_vec", np.asarray(s.alat, np.float64))
        _rec(f, "blat_vec", np.asarray(s.blat, np.float64))
        _rec(f, "volume", np.array([s.volume], np.float64))
        _rec(f, "atomic_pos", np.asarray(s.atomic_pos, np.float64))
        _rec(f, "atom_type", np.asarray(s.atom_type, np.int32))
        _rec(f, "kpt_iredBZ", np.asarray(s.kpt_ibz_cart, np.float64))
        _rec(f, "kpt_fullBZ", np.asarray(s.kpt_bz_crys, np.float64))
        _rec(f, "kmap", np.asarray(s.kmap, np.int32))
        _rec(f, "syms", np.array([list(np.asarray(R, np.float64).reshape(9)) +
                                  list(np.asarray(t, np.float64)) + [float(tr)]
                                  for (R, t, tr) in s.syms], np.float64))
        _rec(f, "pp_ints", np.array([ps.ntype, ps.lmax, ps.nltimesj], np.int64))
        _rec(f, "PP_table", np.asarray(ps.PP_table, np.float64))
        _rec(f, "Fsign", np.asarray(ps.Fsign, np.float64))
        _rec(f, "qmax_abs", np.array([ps.qmax_abs], np.float64))
        for t in range(ps.ntype):
            v, r, rab, z = ps.loc[t]
            _rec(f, f"loc{t}_v", np.asarray(v, np.float64))
            _rec(f, f"loc{t}_r", np.asarray(r, np.float64))
            _rec(f, f"loc{t}_rab", np.asarray(rab, np.float64))
            _rec(f, f"loc{t}_z", np.array([z], np.float64))
        for ik, w in enumerate(s.wfcs):
            _rec(f, f"wfc{ik}", w.wfc)
            _rec(f, f"gvec{ik}", w.gvec)
            _rec(f, f"Fk{ik}", w.Fk)
        _rec(f, "qpt", np.asarray(s.qpts[0], np.float64))
        _rec(f, "eigvec", s.eigvec(0))
        _rec(f, "dvscf", s.dvscf(0))
        _rec(f, "pairs", np.asarray(pairs, np.int32))


def test_c_caller_sequence(tmp_path):
    assert os.path.exists(EXE), "tests/c_abi/lelph_boundary is built by __graft_entry__.build()"
    s = synth.tiny(nbnds=6, syms="cubic", q=(0.0, 0.5, 0.0))
    pairs = synth.kq_pairs(s, 0)
    ikq, ik, kqsym, ksym = pairs[4]
    pairs += [(ikq, ik, kqsym, tr_partner(s, ik, ksym)),
              (ikq, ik, tr_partner(s, ikq, kqsym), ksym)]  # time-reversed sides
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "g.bin")
    _write_inputs(inp, s, pairs)
    say(f"case: C caller, {len(pairs)} pairs {pairs}")
    r = subprocess.run([EXE, inp, out], capture_output=True, text=True, timeout=120)
    say(r.stdout.strip() + " " + r.stderr.strip())
    assert r.returncode == 0
    gsz = s.nmodes * s.nspin * s.nbnds ** 2
    g = np.fromfile(out, np.complex128).reshape(2, len(pairs), s.nmodes, s.nspin, s.nbnds,
                                                 s.nbnds)
    o = O.OracleSystem(s)
    dv = o.dv_total(0)
    ref = np.stack([o.add_nonlocal(0, *p, o.elph_local(0, *p, dv)) for p in pairs])
    e1, e2 = rel_err(g[0], ref), rel_err(g[1], ref)
    say(f"  lelph_gkq {e1:.3e}, lelph_gkq_batch {e2:.3e} ({gsz} g per pair)")
    assert e1 < 1e-10 and e2 < 1e-10
