"""Readers of real inputs (letzelphc_amd/save.py over the host C parsers):
QE ph_save (dyn0, data-file-schema.xml, UPF-1/2, tensors.xml,
quadrupole.fmt) and the Yambo SAVE netCDF-4 databases.  Each test writes
the files from a synthetic System in the reference's layouts
(letzelphc_amd/savefiles.py) and checks that the readers give the System back,
plus the conversions the reference applies on the way (phonon symmetries
to cartesian, time-reversal doubling, magnetic reordering, q/k star
expansion, band window, precision check)."""
import numpy as np
import pytest

from letzelphc_amd import lelph, save, synth
from oracle import oracle as O

from letzelphc_amd import savefiles as SF

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")


def _grid(n):
    return np.array([[i / n[0], j / n[1], k / n[2]] for i in range(n[0]) for j in range(n[1])
                     for k in range(n[2])])


def _ibz_qpts(s, qgrid, syms):
    ibz, _ = synth._reduce_kgrid(_grid(qgrid), syms, s.alat, s.blat)
    return ibz


def test_upf_v2_and_v1(tmp_path):
    r = np.linspace(0.001, 10.0, 37)
    rab = np.gradient(r) * (1 + 1e-3 * np.arange(37))
    v = -2.0 * 3.0 * np.tanh(r) / r
    SF.write_upf2(tmp_path / "a.upf", v, r, rab, 3.0)
    SF.write_upf1(tmp_path / "b.upf", v, r, rab, 5.0)
    for name, z in (("a.upf", 3.0), ("b.upf", 5.0)):
        vv, rr, rb, zz = lelph.parse_upf(tmp_path / name)
        assert zz == z
        assert np.array_equal(rr, r) and np.array_equal(rb, rab) and np.array_equal(vv, v)
    SF.write_upf2(tmp_path / "us.upf", v, r, rab, 3.0, pseudo_type="US")
    with pytest.raises(lelph.LelphError):
        lelph.parse_upf(tmp_path / "us.upf")
    with pytest.raises(lelph.LelphError):
        lelph.parse_upf(tmp_path / "missing.upf")


def test_qexml_fields_and_phonon_symmetries(tmp_path):
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), with_wfc=False)
    SF.write_qexml(tmp_path / "d.xml", s, ["x0.upf", "x1.upf"], ["Si", "O"])
    x = lelph.parse_qexml(tmp_path / "d.xml")
    assert (x["natom"], x["ntype"], x["nsym"]) == (s.natom, 2, 4)
    assert x["fft_dims"] == s.fft_dims and x["dimension"] == "3"
    assert np.array_equal(x["lat_vec"], s.alat) and x["alat"][0] == s.alat[0, 0]
    assert np.array_equal(x["atom_type"], s.atom_type)
    assert np.array_equal(x["atomic_pos"], s.atomic_pos)
    assert x["pseudo_files"] == ["x0.upf", "x1.upf"] and x["pseudo_dir"] == "./pseudo/"
    assert (x["nmag"], x["nspinor"], x["is_soc_present"]) == (1, 1, False)
    # no inversion among the proper operations: time reversal kept, the
    # group doubled with -S (read_qexml.c:355-380, get_data_from_qe.c:333-338)
    assert x["ph_tim_rev"] and not x["ph_mag_symm"]
    ph = save._phonon_syms(x, 4)
    for k, (R, v, tr) in enumerate(s.syms):
        Rp, vp, trp = ph[k]
        assert trp == bool(tr)
        assert np.allclose(Rp, R, atol=1e-12) and np.allclose(vp, v, atol=1e-12)


@pytest.mark.parametrize("case", ["lsda", "noncolin", "noncolin_mag", "noinv", "inversion",
                                  "isolated", "magnetic_group"])
def test_qexml_flags(tmp_path, case):
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), with_wfc=False)
    kw = {}
    if case in ("lsda", "noncolin", "noncolin_mag"):
        kw["mag"] = case
    elif case == "noinv":
        kw["noinv"] = True
    elif case == "isolated":
        kw["isolated"] = "2D"
    elif case == "inversion":
        I = np.eye(3)
        kw["extra_syms"] = [(I, np.zeros(3), None), (-I, np.zeros(3), None)]
    elif case == "magnetic_group":
        I = np.eye(3)
        C2 = np.diag([-1.0, -1.0, 1.0])
        kw["mag"] = "noncolin_mag"
        kw["extra_syms"] = [(I, np.zeros(3), False), (I, [0.5, 0, 0], True),
                            (C2, [0, 0.25, 0], True), (C2, np.zeros(3), False)]
    SF.write_qexml(tmp_path / "d.xml", s, ["x0.upf", "x1.upf"], ["Si", "O"], **kw)
    x = lelph.parse_qexml(tmp_path / "d.xml")
    if case == "lsda":
        assert (x["nmag"], x["nspinor"], x["ph_tim_rev"]) == (2, 1, True)
    elif case == "noncolin":
        assert (x["nmag"], x["nspinor"], x["ph_tim_rev"]) == (1, 2, True)
    elif case == "noncolin_mag":
        assert (x["nmag"], x["nspinor"], x["ph_tim_rev"]) == (4, 2, False)
    elif case == "noinv":
        assert not x["ph_tim_rev"]
    elif case == "inversion":
        assert not x["ph_tim_rev"] and x["nsym"] == 2
    elif case == "isolated":
        assert x["dimension"] == "2"
    else:
        # operations without time reversal first, each half in file order
        assert x["ph_mag_symm"] and x["ph_tim_rev"] and x["nmag"] == 4
        assert list(x["trevs"]) == [False, False, True, True]
        assert np.array_equal(x["sym_mats"][1], np.diag([-1.0, -1, 1]))
        assert np.array_equal(x["sym_tau"][2], [0.5, 0, 0])
        assert np.array_equal(x["sym_tau"][3], [0, 0.25, 0])


def test_tensors_and_quadrupoles(tmp_path):
    rng = np.random.default_rng(3)
    eps = rng.standard_normal((3, 3))
    zb = rng.standard_normal((2, 3, 3))
    Q = rng.standard_normal((2, 3, 3, 3))
    Q = 0.5 * (Q + Q.transpose(0, 2, 1, 3))  # Q[a][x][y][dir] symmetric in x, y
    SF.write_tensors(tmp_path / "tensors.xml", eps, zb)
    SF.write_quadrupole(tmp_path / "quadrupole.fmt", Q)
    e, z, q = lelph.read_ph_tensors_qe(tmp_path / "tensors.xml", tmp_path / "quadrupole.fmt", 2)
    assert np.array_equal(e, eps) and np.array_equal(z, zb) and np.array_equal(q, Q)
    e, z, q = lelph.read_ph_tensors_qe(tmp_path / "none.xml", tmp_path / "none.fmt", 2)
    assert e is None and z is None and q is None
    SF.write_tensors(tmp_path / "t2.xml", eps, None)
    e, z, q = lelph.read_ph_tensors_qe(tmp_path / "t2.xml", None, 2)
    assert np.array_equal(e, eps) and z is None and q is None


def test_dyn0_and_qstar(tmp_path):
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), with_wfc=False)
    qgrid = (2, 2, 2)
    x_syms = [(R, v, bool(tr)) for (R, v, tr) in s.syms]
    qibz = _ibz_qpts(s, qgrid, x_syms)
    SF.write_ph_save(tmp_path, s, qgrid, qibz)
    qs, nbz = lelph.read_qpts_qe(tmp_path / "dyn0")
    assert nbz == 8 and len(qs) == len(qibz)
    x, ph, loc = save.read_ph_save(tmp_path)
    assert np.allclose(ph.qpts_ibz, qibz, atol=1e-12)
    assert ph.qpts_bz.shape == (8, 3) and ph.nqstar.sum() == 8
    # every grid point once, and each reached from its iBZ point by its map
    for g in _grid(qgrid):
        assert save.find_kidx_in_list(ph.qpts_bz, g) >= 0
    A = s.alat
    for ib, (iq, isym) in enumerate(ph.qmap):
        R = ph.ph_syms[isym][0]
        qc = (synth.reciprocal_vecs(A) / (2 * synth.RPI)) @ ph.qpts_ibz[iq]
        d = A.T @ (R @ qc) - ph.qpts_bz[ib]
        assert np.allclose(d, np.rint(d), atol=1e-9)
    for (v, r, rab, z), (v0, r0, rab0, z0) in zip(loc, s.pseudo.loc):
        assert np.array_equal(v, v0) and np.array_equal(r, r0) and z == z0


def _write_both(tmp_path, s, dtype=np.float64, **kw):
    qgrid = (2, 2, 2)
    qibz = _ibz_qpts(s, qgrid, [(R, v, bool(tr)) for (R, v, tr) in s.syms])
    SF.write_save(tmp_path / "SAVE", s, dtype=dtype)
    SF.write_ph_save(tmp_path / "ph_save", s, qgrid, qibz, **kw)
    return qibz


def _check_system(r, s, b0=0, b1=None):
    b1 = s.nbnds if b1 is None else b1
    assert (r.natom, r.nspin, r.nspinor, r.nbnds, r.nmag) == (s.natom, s.nspin, s.nspinor,
                                                               b1 - b0, s.nmag)
    assert np.array_equal(r.alat, s.alat) and np.allclose(r.blat, s.blat, rtol=0, atol=1e-14)
    assert r.fft_dims == s.fft_dims and r.timerev == s.timerev
    assert len(r.syms) == len(s.syms)
    for (R, v, tr), (R0, v0, tr0) in zip(r.syms, s.syms):
        assert np.array_equal(R, R0) and bool(tr) == bool(tr0)
    assert np.allclose(r.kpt_ibz_cart, s.kpt_ibz_cart, rtol=0, atol=1e-14)
    for w, w0 in zip(r.wfcs, s.wfcs):
        assert np.array_equal(w.gint, w0.gint)
        assert np.allclose(w.gvec, w0.gvec, rtol=0, atol=1e-13)
        assert np.array_equal(w.wfc, np.asarray(w0.wfc)[:, b0:b1])
        assert np.array_equal(w.Fk, w0.Fk)
    ps, ps0 = r.pseudo, s.pseudo
    assert np.array_equal(ps.PP_table, ps0.PP_table) and np.array_equal(ps.Fsign, ps0.Fsign)
    assert (ps.nltimesj, ps.lmax, ps.ntype) == (ps0.nltimesj, ps0.lmax, ps0.ntype)
    # full BZ: the same points, each the image of its iBZ point
    assert len(r.kpt_bz_crys) == len(s.kpt_bz_crys)
    for k in s.kpt_bz_crys:
        assert save.find_kidx_in_list(r.kpt_bz_crys, k) >= 0
    for ib, (ik, isym) in enumerate(r.kmap):
        d = r.alat.T @ (np.asarray(r.syms[isym][0]) @ r.kpt_ibz_cart[ik]) - r.kpt_bz_crys[ib]
        assert np.allclose(d, np.rint(d), atol=1e-9)


@pytest.mark.parametrize("kind", ["cubic", "spin2", "soc"])
def test_read_save_roundtrip(tmp_path, kind):
    kw = dict(syms="cubic", q=(0.0, 0.5, 0.0)) if kind == "cubic" else {}
    if kind == "spin2":
        kw.update(nspin=2, nmag=2)
    if kind == "soc":
        kw.update(nspinor=2, nmag=1, soc=True)
    s = synth.tiny(**kw)
    mag = {"spin2": "lsda", "soc": "noncolin"}.get(kind)
    _write_both(tmp_path, s, mag=mag)
    r, ph = save.read_save(tmp_path / "SAVE", tmp_path / "ph_save")
    _check_system(r, s)
    if kind == "soc":
        assert r.pseudo.soc and len(r.pseudo.fcoeff) == len(s.pseudo.fcoeff)
        for f, f0 in zip(r.pseudo.fcoeff, s.pseudo.fcoeff):
            assert np.array_equal(f, f0)
    else:
        assert not r.pseudo.soc and r.pseudo.fcoeff is None
    # a band window (1-based, inclusive)
    r2, _ = save.read_save(tmp_path / "SAVE", tmp_path / "ph_save", start_band=2, end_band=3)
    _check_system(r2, s, 1, 3)


def test_read_save_precision_and_mismatch(tmp_path):
    s = synth.tiny()
    _write_both(tmp_path, s, dtype=np.float32)
    with pytest.raises(ValueError, match="precision"):
        save.read_save(tmp_path / "SAVE", tmp_path / "ph_save")
    r, _ = save.read_save(tmp_path / "SAVE", tmp_path / "ph_save", nc4_float=np.float32)
    for w, w0 in zip(r.wfcs, s.wfcs):
        assert np.array_equal(w.gint, w0.gint)
        assert np.max(np.abs(w.wfc - w0.wfc)) <= 2 ** -23 * np.max(np.abs(w0.wfc))
    # a SAVE of another cell
    s2 = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0))
    SF.write_save(tmp_path / "SAVE2", s2)
    with pytest.raises(ValueError, match="Lattice vectors"):
        save.read_save(tmp_path / "SAVE2", tmp_path / "ph_save")


def test_read_save_feeds_the_oracle(tmp_path):
    """The assembled System drives the FP64 restatement: every k whose
    (iBZ k, symmetry) the star expansion picked as the synthetic grid did
    gives the same g."""
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0), nbnds=4)
    _write_both(tmp_path, s)
    r, _ = save.read_save(tmp_path / "SAVE", tmp_path / "ph_save")
    iq = [i for i, q in enumerate(r.qpts) if np.allclose(q, s.qpts[0], atol=1e-12)]
    assert iq, "the system's q is one of the iBZ q points"
    # V_loc table range from the iBZ q of dyn0 (get_save_data.c:596-607)
    assert r.pseudo.qmax_abs == np.ceil(np.abs(np.asarray(r.qpts)).max()) + 1
    r.pseudo.qmax_abs = s.pseudo.qmax_abs
    r.qpts = [s.qpts[0]]
    r._eig = {0: s.eigvec(0)}
    r._dv = {0: s.dvscf(0)}
    O.set_threads(4)
    o, o0 = O.OracleSystem(r), O.OracleSystem(s)
    dv = o.dv_total(0)
    assert np.array_equal(dv, o0.dv_total(0))
    g = o.compute_elphq(0, dv, non_loc=True)
    g0 = o0.compute_elphq(0, dv, non_loc=True)
    same = 0
    for ib, k in enumerate(r.kpt_bz_crys):
        j = save.find_kidx_in_list(s.kpt_bz_crys, k)
        kq = k + np.asarray(s.qpts[0])
        if tuple(r.kmap[ib]) != tuple(s.kmap[j]):
            continue
        jr = save.find_kidx_in_list(r.kpt_bz_crys, kq)
        j0 = save.find_kidx_in_list(s.kpt_bz_crys, kq)
        if tuple(r.kmap[jr]) != tuple(s.kmap[j0]):
            continue
        # G and k are rebuilt from the stored (2 pi / a) units: equal to
        # rounding, not bit for bit
        err = np.abs(g[ib] - g0[j]).max() / np.abs(g0).max()
        assert err < 1e-12, (ib, err)
        same += 1
    assert same >= 2
