"""The real-input driver end to end (letzelphc_amd/driver.py): SAVE +
ph_save files (dyn, dvscf, patterns per iBZ q) -> ndb.elph and ndb.Dmats
written from the device, checked slab by slab against the FP64 oracle:
D matrices, iBZ rows, and every star member's rows rebuilt from oracle g
and oracle D with the host q-star rotation."""
import numpy as np
import pytest

from letzelphc_amd import driver, lelph, ndb, save, synth
from letzelphc_amd import savefiles as SF
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu


def _cplx(a):
    return a[..., 0] + 1j * a[..., 1]


def test_run_elph_end_to_end(tmp_path):
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), nbnds=4)
    masses = np.array([28.0855, 15.999])[np.asarray(s.atom_type)]
    qibz = SF.write_inputs(tmp_path, s, (2, 2, 2), masses)
    ph_dir = tmp_path / "ph_save"
    nm = s.nmodes
    sh = (nm, s.nmag) + tuple(s.fft_dims)
    say(f"case: run_elph tiny cubic, {len(qibz)} iBZ q")
    p_elph, p_dmat = driver.run_elph(tmp_path / "SAVE", ph_dir, tmp_path / "out")
    r, ph = save.read_save(tmp_path / "SAVE", ph_dir)
    nk = len(r.kpt_bz_crys)
    with ndb.NC4File(p_dmat, "r") as f:
        Dg = _cplx(f.get("Dmats"))
    with ndb.NC4File(p_elph, "r") as f:
        G = _cplx(f.get("elph_mat"))
        E = _cplx(f.get("POLARIZATION_VECTORS"))
        assert f.get("qpoints").shape == (len(ph.qpts_bz), 3)
    O.set_threads(8)
    o = O.OracleSystem(r)
    Do = np.empty_like(Dg)
    for isym, (R, tau, tr) in enumerate(ph.ph_syms):
        for k in range(nk):
            Do[isym, k] = o.electronic_reps(R, tau, tr, k)
    assert rel_err(Dg, Do) < 1e-10
    eye = np.eye(nm, dtype=np.complex128)
    for iq in range(len(qibz)):
        qpos = int(np.sum(ph.nqstar[:iq]))
        eig = np.ascontiguousarray(E[qpos].reshape(nm, nm))
        mode = np.empty(sh, np.complex128)
        O.check(O.lib().orc_read_dvscf_qe(str(ph_dir / f"dvscf{iq + 1}").encode(), o.lat,
                                          O.dptr(eig), O.dptr(eye), O.dptr(mode)))
        r._eig[iq] = eig
        r._dv[iq] = mode
        oq = O.OracleSystem(r)
        gq = oq.compute_elphq(iq, oq.dv_total(iq), non_loc=True)
        assert rel_err(G[qpos], gq) < 1e-10, iq
        q = ph.qpts_ibz[iq]
        for istar in range(1, int(ph.nqstar[iq])):
            qs = qpos + istar
            isym = int(ph.qmap[qs, 1])
            R, _, tr = ph.ph_syms[isym]
            for k in range(nk):
                kq = save.find_kidx_in_list(r.kpt_bz_crys, r.kpt_bz_crys[k] + q)
                kc = (r.blat / (2 * synth.RPI)) @ r.kpt_bz_crys[k]
                Sk = save.find_kidx_in_list(r.kpt_bz_crys, r.alat.T @ (R @ kc))
                want = lelph.elph_q_rotate(Do[isym, kq], gq[k], Do[isym, k], tr)
                assert rel_err(G[qs, Sk], want) < 1e-10, (iq, istar, k)


def test_cli_from_input_file(tmp_path):
    """`python -m letzelphc_amd.driver -F elph.in` (the reference's lelphc
    -F elph.in) with the yambo convention and a band window gives the same
    ndb.elph as run_elph called with those arguments"""
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), nbnds=4)
    masses = np.array([28.0855, 15.999])[np.asarray(s.atom_type)]
    SF.write_inputs(tmp_path, s, (2, 2, 2), masses)
    inp = tmp_path / "elph.in"
    inp.write_text(f"save_dir = '{tmp_path / 'SAVE'}'\nph_save_dir = {tmp_path / 'ph_save'}\n"
                   "kernel = dfpt\nconvention = yambo\nstart_bnd = 2\nend_bnd = 4\n")
    say("case: driver command line from elph.in")
    assert driver.main(["-F", str(inp), "--out", str(tmp_path / "cli")]) == 0
    p2, _ = driver.run_elph(tmp_path / "SAVE", tmp_path / "ph_save", tmp_path / "direct",
                            kminusq=True, start_band=2, end_band=4)
    with ndb.NC4File(tmp_path / "cli" / "ndb.elph", "r") as f:
        a = f.get("elph_mat")
    with ndb.NC4File(p2, "r") as f:
        b = f.get("elph_mat")
    assert a.shape == b.shape and a.shape[-3] == 3
    assert np.array_equal(a, b)
    say(f"  ndb.elph {a.shape} identical to run_elph's")
