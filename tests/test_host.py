"""Host-side checks that need no GPU: the C-ABI libraries load and export
every entry point declared in include/lelph_hip.h and include/lelph_nc4.h,
the host-only entry points
(status strings, k+q indices, q-star rotation) agree with the oracle and the
reference fixture, the seeded generators are deterministic, and the bench
refuses more than one GPU."""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lelph_hip.h")


def declared_functions(header=HEADER):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lelph_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    say("case: C-ABI exports")
    names = declared_functions()
    assert len(names) >= 15
    L = C.CDLL(lelph.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_nc4_library_exports_header():
    """include/lelph_nc4.h against liblelph_nc4.so (host netCDF-4 I/O)"""
    from letzelphc_amd import ndb
    if not os.path.exists(ndb.LIB_PATH):
        import pytest
        pytest.skip("liblelph_nc4.so not built (no HDF5)")
    names = declared_functions(os.path.join(ROOT, "include", "lelph_nc4.h"))
    assert len(names) == 7, names
    L = C.CDLL(ndb.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    L = lelph.lib()
    seen = set()
    for st in range(0, 10):
        s = lelph.strerror(st)
        assert s and s not in seen
        seen.add(s)
    assert lelph.strerror(12345)


def test_kplusq_idxs_host_vs_fixture():
    with np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")) as z:
        k, qs, idx = z["kq_k"], z["kq_q"], z["kq_idx"]
    for i, q in enumerate(qs):
        got = lelph.kplusq_idxs(k, q)
        np.testing.assert_array_equal(got, idx[i])


def test_elph_q_rotate_host_vs_oracle():
    rng = np.random.default_rng(21)
    nm, ns, nb = 4, 2, 5
    g = rng.standard_normal((nm, ns, nb, nb)) + 1j * rng.standard_normal((nm, ns, nb, nb))
    Dl = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
    Dr = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
    from letzelphc_amd.abi import dptr
    for tr in (0, 1):
        got = lelph.elph_q_rotate(Dl, g, Dr, bool(tr))
        ref = np.zeros_like(g)
        O.lib().orc_elph_q_rotate(dptr(Dl), dptr(g), dptr(Dr), nm, ns, nb, tr, dptr(ref))
        assert rel_err(got, ref) < 1e-14


def test_synth_deterministic_and_orthonormal():
    a = synth.tiny(nbnds=4, syms="cubic", nspin=2, nmag=2)
    b = synth.tiny(nbnds=4, syms="cubic", nspin=2, nmag=2)
    for wa, wb in zip(a.wfcs, b.wfcs):
        assert np.array_equal(wa.wfc, wb.wfc)
        assert np.array_equal(wa.gvec, wb.gvec)
        for s in range(a.nspin):
            M = wa.wfc[s].reshape(a.nbnds, -1)
            assert np.allclose(M @ M.conj().T, np.eye(a.nbnds), atol=1e-13)
    assert np.array_equal(a.dvscf(0), b.dvscf(0))
    assert np.array_equal(a.eigvec(0), b.eigvec(0))
    e = a.eigvec(0)
    assert np.allclose(e @ e.conj().T, np.eye(a.nmodes), atol=1e-13)


def test_kq_pairs_cover_grid():
    """Every full-BZ k maps to an iBZ point and symmetry that reproduce it,
    and k+q lands on the grid."""
    s = synth.config("c0", nbnds=2, with_wfc=False)
    pairs = synth.kq_pairs(s, 0)
    assert len(pairs) == len(s.kpt_bz_crys)
    for i, (ikq, ik, kqsym, ksym) in enumerate(pairs):
        Sc = synth._sym_crys(np.asarray(s.syms[ksym][0]), s.alat, s.blat)
        d = Sc @ s.kpt_ibz_crys[ik] - s.kpt_bz_crys[i]
        assert np.allclose(d, np.rint(d), atol=1e-6)


def test_bench_refuses_several_gpus():
    say("case: bench --gpus 2")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"] or "1 GPU" in line["error"]
