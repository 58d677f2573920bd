"""Readers of real inputs: a Yambo SAVE directory plus a QE ph_save
directory -> the System the library's context takes, and the phonon data
(symmetries, q grid, dielectric tensors) around it.

Mirrors, step by step:
  * get_data_from_qe   src/io/qe/get_data_from_qe.c:20-399
        dyn0 q points, data-file-schema.xml, tensors.xml / quadrupole.fmt,
        phonon symmetries to cartesian (x' = Sx + v, v = -tau_QE), q-star
        expansion, UPF local parts
  * read_and_alloc_save_data   src/io/get_save_data.c:51-614
        ndb.kindx (nXkbz), ns.db1 (DIMENSIONS, LATTICE_PARAMETER,
        LATTICE_VECTORS, SYMMETRY, K-POINTS, N_ATOMS, ATOM_MAP, WFC_NG,
        G-VECTORS, WFC_GRID), ns.kb_pp_pwscf (PARS, PP_TABLE, PP_KBS),
        ns.kb_pp_pwscf_fragment_<ik> (PP_KB_K<ik>),
        ns.wf_fragments_<is*nk+ik>_1 (WF_COMPONENTS_@_SP_POL<is>_K<ik>_BAND_GRP_1)
  * bz_expand   src/symmetries/BZ_expand.c:11-100
The file parsers are the library's host C code (lelph_parse_qexml,
lelph_parse_upf, lelph_read_ph_tensors_qe, lelph_read_qpts_qe, and the
netCDF-4 reads of liblelph_nc4.so); this module only assembles.  Errors
raise (the reference aborts).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import lelph, ndb, synth

ELPH_EPS = 1e-5  # src/common/constants.h:9


@dataclass
class PhononData:
    """struct Phonon (src/common/dtypes.h) as get_data_from_qe fills it."""
    qpts_ibz: np.ndarray   # (nq_iBZ, 3) crystal
    qpts_bz: np.ndarray    # (nq_BZ, 3) crystal
    qmap: np.ndarray       # (nq_BZ, 2) (iq_iBZ, isym)
    nqstar: np.ndarray     # (nq_iBZ,)
    ph_syms: list          # [(R cart, tau cart, time_rev)], nph_sym entries
    ph_tim_rev: bool
    ph_mag_symm: bool
    epsilon: np.ndarray | None
    Zborn: np.ndarray | None
    Qpole: np.ndarray | None


def find_kidx_in_list(klist, k) -> int:
    """find_kidx_in_list (src/common/numerical_func.c:471-497): first index
    whose crystal difference is an integer vector within ELPH_EPS."""
    if len(klist) == 0:
        return -1
    d = np.asarray(klist) - np.asarray(k)[None, :]
    d = d - np.rint(d)
    hit = np.nonzero(np.sqrt(np.sum(d * d, 1)) < ELPH_EPS)[0]
    return int(hit[0]) if len(hit) else -1


def bz_expand(kibz_cart, syms, alat):
    """bz_expand (src/symmetries/BZ_expand.c:11-100): full-BZ points
    (crystal), star sizes and (ik_iBZ, isym) map, the first symmetry that
    reaches each new point; a list that is already the full BZ maps every
    point to itself with the identity."""
    A = np.asarray(alat)
    kbz, kmap, kstar = [], [], []
    for ik, k in enumerate(np.asarray(kibz_cart)):
        n = 0
        for isym, (R, _, _) in enumerate(syms):
            kc = A.T @ (np.asarray(R) @ k)
            if find_kidx_in_list(kbz, kc) < 0:
                kbz.append(kc)
                kmap.append((ik, isym))
                n += 1
        kstar.append(n)
    nibz = len(kibz_cart)
    if len(syms) > 1 and len(kbz) == nibz:
        iden = -1
        for isym, (R, _, _) in enumerate(syms):
            if np.sqrt(np.sum((np.asarray(R) - np.eye(3)) ** 2)) / 3.0 < 1e-5:
                iden = isym
                break
        if iden < 0:
            raise ValueError("Identity matrix not found.")
        kmap = [(ik, iden) for ik in range(nibz)]
        kstar = [1] * nibz
        kbz = [A.T @ np.asarray(k) for k in kibz_cart]
    return (np.array(kbz).reshape(-1, 3), np.array(kstar, np.int64),
            np.array(kmap, np.int32).reshape(-1, 2))


def _phonon_syms(x, nph_sym):
    """get_data_from_qe.c:258-331: S_cart = A S_crys (B/2pi)^T, negated for
    time-reversed operations, tiny entries made +0; v = -A tau_crys (QE
    writes x' = Sx - v), negated again for time reversal; the second half
    holds -S with the complementary time-reversal flag."""
    A = x["lat_vec"]
    B = synth.reciprocal_vecs(A) / (2 * synth.RPI)
    first, second = [], []
    for k in range(nph_sym):
        S = A @ x["sym_mats"][k] @ B.T
        tr = bool(x["trevs"][k])
        if tr:
            S = -S
        S = np.where(np.abs(S) < ELPH_EPS, np.abs(S), S)
        St = -S
        vt = A @ x["sym_tau"][k]
        v = -vt
        if tr:
            v, vt = -v, -vt
        first.append((S, v, tr))
        second.append((St, vt, not tr))
    return first + second


def read_ph_save(ph_save_dir):
    """get_data_from_qe (src/io/qe/get_data_from_qe.c:20-399): returns the
    parsed data-file-schema.xml dict, the PhononData and the per-type local
    pseudopotentials [(Vloc, r, rab, Zval)]."""
    j = lambda name: os.path.join(ph_save_dir, name)
    qibz, nq_bz = lelph.read_qpts_qe(j("dyn0"))
    x = lelph.parse_qexml(j("data-file-schema.xml"))
    natom = x["natom"]
    eps, zb, qp = lelph.read_ph_tensors_qe(j("tensors.xml"), j("quadrupole.fmt"), natom)
    nph = x["nsym"]
    syms2 = _phonon_syms(x, nph)
    if x["ph_tim_rev"] and not x["ph_mag_symm"]:
        nph *= 2
    ph_syms = syms2[:nph]
    A = x["lat_vec"]
    qcart = qibz / x["alat"][None, :]
    qbz, nqstar, qmap = bz_expand(qcart, ph_syms, A)
    if len(qbz) != nq_bz:
        raise ValueError("Expansion of q-points in full BZ failed")
    qcrys = qcart @ A  # A^T q per row
    loc = [lelph.parse_upf(j(f)) for f in x["pseudo_files"]]
    ph = PhononData(qcrys, qbz, qmap, nqstar, ph_syms, x["ph_tim_rev"], x["ph_mag_symm"],
                    eps, zb, qp)
    return x, ph, loc


def read_save(save_dir, ph_save_dir, start_band=-1, end_band=-1, nc4_float=np.float64):
    """read_and_alloc_save_data (src/io/get_save_data.c:51-614): System
    (the iBZ wavefunctions of bands start_band..end_band, 1-based, G
    vectors, KB form factors, lattice, electronic symmetries, full-BZ k
    grid) and PhononData.  nc4_float is the floating type the databases
    must hold (the reference requires the precision it was built with,
    get_save_data.c:119-134)."""
    x, ph, loc = read_ph_save(ph_save_dir)
    j = lambda name: os.path.join(save_dir, name)
    want = {np.dtype(np.float64): ndb.NC_DOUBLE, np.dtype(np.float32): ndb.NC_FLOAT}[
        np.dtype(nc4_float)]
    with ndb.NC4File(j("ndb.kindx"), "r") as f:
        xt, _ = f.inq("HEAD_VERSION")
        if xt != want:
            raise ValueError("Yambo and LetzElPhC are compiled with different precision.")
        nkbz = int(f.get("nXkbz", dtype=np.int32).reshape(-1)[0])
    with ndb.NC4File(j("ns.db1"), "r") as f:
        dims = f.get("DIMENSIONS").reshape(-1)
        nspinor, nspin = int(np.rint(dims[11])), int(np.rint(dims[12]))
        timerev, nbtot = int(np.rint(dims[9])), int(np.rint(dims[5]))
        nsym, nibz = int(np.rint(dims[10])), int(np.rint(dims[6]))
        ng_total, ng_shell = int(np.rint(dims[7])), int(np.rint(dims[8]))
        if start_band < 1 or end_band < 1 or start_band > nbtot or end_band > nbtot \
                or start_band >= end_band:
            start_band, end_band = 1, nbtot
        nb = end_band - start_band + 1
        lat_param = f.get("LATTICE_PARAMETER").reshape(-1)
        latv = f.get("LATTICE_VECTORS").reshape(-1)
        A = x["lat_vec"]
        if np.any(np.abs(latv - A.reshape(-1)) > 1e-4):
            raise ValueError("Lattice vectors mismatch b/w yambo and dft code.")
        symm = f.get("SYMMETRY").reshape(nsym, 3, 3)
        kpts = f.get("K-POINTS").reshape(3, nibz)
        kibz = (kpts / lat_param[:, None]).T.copy()
        syms = []
        for i in range(nsym):
            syms.append((symm[i].T.copy(), np.zeros(3), i >= nsym // (timerev + 1)))
        ntype = int(np.rint(f.get("number_of_atom_species").reshape(-1)[0]))
        if ntype != x["ntype"]:
            raise ValueError("Number of types in save not same as dft code.")
        natom_t = np.rint(f.get("N_ATOMS").reshape(-1)).astype(int)
        if natom_t.sum() != x["natom"]:
            raise ValueError("Number of atoms different in yambo and qexml.")
        amap = f.get("ATOM_MAP").reshape(-1)
        nspec_max = int(natom_t.max())
        for it in range(ntype):
            for isp in range(natom_t[it]):
                ia = int(np.rint(amap[isp + it * nspec_max] - 1))
                if x["atom_type"][ia] != it:
                    raise ValueError("Atomic type mismatch in yambo and qexml")
        ngmax = np.rint(f.get("WFC_NG").reshape(-1)).astype(int)
        gtot = f.get("G-VECTORS").reshape(3, ng_total)
        gidx = f.get("WFC_GRID").reshape(nibz, ng_shell)
    with ndb.NC4File(j("ns.kb_pp_pwscf"), "r") as f:
        pars = f.get("PARS").reshape(-1)
        nlj, lmax = int(np.rint(pars[2])), int(np.rint(pars[3]))
        pp_table = f.get("PP_TABLE").reshape(nlj, ntype, 3)
        fsign = f.get("PP_KBS").reshape(nlj, ntype)
    kbz, _, kmap = bz_expand(kibz, syms, A)
    if len(kbz) != nkbz:
        raise ValueError("K-point expansion over full BZ failed.")
    wfcs = []
    for ik in range(nibz):
        ng = int(ngmax[ik])
        idx = np.rint(gidx[ik, :ng] - 1).astype(int)
        if np.any(idx < 0):
            raise ValueError("Wrong g indices")
        gvec = (gtot[:, idx] / lat_param[:, None]).T.copy()
        wfc = np.empty((nspin, nb, nspinor, ng), np.complex128)
        for s in range(nspin):
            name = j(f"ns.wf_fragments_{s * nibz + ik + 1}_1")
            with ndb.NC4File(name, "r") as f:
                raw = f.get(f"WF_COMPONENTS_@_SP_POL{s + 1}_K{ik + 1}_BAND_GRP_1",
                            (start_band - 1, 0, 0, 0), (nb, nspinor, ng, 2))
            wfc[s] = raw[..., 0] + 1j * raw[..., 1]
        with ndb.NC4File(j(f"ns.kb_pp_pwscf_fragment_{ik + 1}"), "r") as f:
            Fk = f.get(f"PP_KB_K{ik + 1}", (0, 0, 0), (nlj, ntype, ng))
        gint = np.rint(gvec @ A).astype(np.int64)  # crystal: A^T G per row
        wfcs.append(synth.Wfc(wfc, gvec, Fk, gint))
    soc = bool(x["is_soc_present"]) and nspinor == 2
    fco = None
    if soc:
        fco = []
        for lidx in range(nlj):
            for it in range(ntype):
                l = int(np.rint(pp_table[lidx, it, 0] - 1))
                tj = int(np.rint(pp_table[lidx, it, 1]))
                fco.append(None if (l < 0 or tj == 0) else synth.soc_fcoeff(l, tj))
    qmax_abs = float(np.ceil(np.abs(ph.qpts_ibz).max())) + 1.0
    ps = synth.PseudoData(ntype, lmax, nlj, pp_table, fsign, np.zeros((nlj, ntype)), 0.0,
                          loc, 0.0, 1.0, qmax_abs, soc, fco)
    blat = synth.reciprocal_vecs(A)
    sys_ = synth.System(
        name=os.path.basename(os.path.normpath(save_dir)), seed=0, alat=A, blat=blat,
        volume=float(abs(np.linalg.det(A))), fft_dims=tuple(int(v) for v in x["fft_dims"]),
        natom=x["natom"], atomic_pos=x["atomic_pos"], atom_type=x["atom_type"], nspin=nspin,
        nspinor=nspinor, nbnds=nb, nmodes=3 * x["natom"], nmag=x["nmag"],
        dimension=x["dimension"], syms=syms, timerev=timerev, kpt_ibz_crys=kibz @ A,
        kpt_ibz_cart=kibz, kpt_bz_crys=kbz, kmap=kmap, ecut=0.0, pseudo=ps,
        qpts=[q for q in ph.qpts_ibz], wfcs=wfcs)
    sys_.band_window = (start_band, end_band)  # 1-based, inclusive
    return sys_, ph
