"""ctypes front end of the FP64 CPU oracle (oracle/liblelph_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
bench's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from letzelphc_amd.abi import (Lattice, LongRange, LrParams, Pseudo, Symmetry, c_double_p,
                                dptr, i64ptr, pack_system)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblelph_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "liblelph_ref.so")


class OrcWfc(C.Structure):
    _fields_ = [("wfc", c_double_p), ("gvec", c_double_p), ("Fk", c_double_p),
                ("npw", C.c_int64)]


class OrcTable(C.Structure):
    _fields_ = [("npts", C.c_int64), ("dg", C.c_double), ("g_co", c_double_p),
                ("vlocg", c_double_p), ("vploc", c_double_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_legendre.restype = C.c_double
        L.orc_legendre.argtypes = [C.c_int, C.c_int, C.c_double]
        L.orc_Ylm.restype = C.c_double
        L.orc_Ylm.argtypes = [C.c_int, C.c_int, c_double_p]
        L.orc_simpson.restype = C.c_double
        L.orc_simpson.argtypes = [c_double_p, c_double_p, C.c_int64]
        L.orc_spline_interpolate.restype = C.c_double
        L.orc_spline_interpolate.argtypes = [C.c_double, C.c_int64, c_double_p,
                                             c_double_p, c_double_p]
        L.orc_prepare_spline.argtypes = [C.c_int64, c_double_p, c_double_p, c_double_p]
        L.orc_SU2mat.argtypes = [c_double_p, C.c_int, C.c_int, C.c_int, c_double_p]
        L.orc_fft3d.argtypes = [c_double_p, C.POINTER(C.c_int64), C.c_int]
        L.orc_fft1d.argtypes = [c_double_p, C.c_int64, C.c_int]
        L.orc_Vloc_Gspace.restype = C.c_double
        L.orc_Vloc_Gspace.argtypes = [c_double_p, C.c_char, C.c_double, C.c_int64,
                                      c_double_p, c_double_p, c_double_p, C.c_double,
                                      C.c_double, C.c_double]
        L.orc_create_vlocg_table.argtypes = [C.POINTER(Lattice), C.POINTER(Pseudo),
                                             C.POINTER(OrcTable)]
        L.orc_free_vloc_table.argtypes = [C.POINTER(OrcTable)]
        L.orc_dVlocq.argtypes = [c_double_p, C.POINTER(Lattice), C.POINTER(Pseudo),
                                 C.POINTER(OrcTable), c_double_p, c_double_p]
        L.orc_add_dvscf_qe.argtypes = [c_double_p, c_double_p, C.POINTER(Lattice)]
        L.orc_dvscf_pattern_to_mode.argtypes = [c_double_p, c_double_p, c_double_p,
                                                C.POINTER(Lattice)]
        L.orc_elphLocal.argtypes = [c_double_p, C.POINTER(OrcWfc), C.POINTER(Lattice),
                                    C.c_int, C.c_int, C.c_int, C.c_int, c_double_p,
                                    c_double_p]
        L.orc_add_elphNonLocal.argtypes = [C.POINTER(OrcWfc), C.POINTER(Lattice),
                                           C.POINTER(Pseudo), C.c_int, C.c_int, C.c_int,
                                           C.c_int, c_double_p, c_double_p]
        L.orc_elph_q_rotate.argtypes = [c_double_p, c_double_p, c_double_p, C.c_int,
                                        C.c_int, C.c_int, C.c_int, c_double_p]
        L.orc_get_KplusQ_idxs.argtypes = [C.c_int64, c_double_p, c_double_p,
                                          C.POINTER(C.c_int)]
        L.orc_compute_elphq.argtypes = [c_double_p, C.POINTER(OrcWfc), C.POINTER(Lattice),
                                        C.POINTER(Pseudo), c_double_p, c_double_p, C.c_int,
                                        C.c_int, C.c_int64, C.c_int64, c_double_p]
        L.orc_reciprocal_vecs.argtypes = [c_double_p, c_double_p]
        L.orc_su2rotate.argtypes = [C.c_int, C.c_int64, c_double_p, c_double_p]
        L.orc_apply_trans_wfc.argtypes = [c_double_p, c_double_p, C.c_int64, C.c_int64,
                                          c_double_p, c_double_p, C.c_int]
        L.orc_sum_VNL_KKp.argtypes = [c_double_p, c_double_p, c_double_p, C.c_int64,
                                      C.c_int64, C.c_int64, c_double_p]
        L.orc_find_gvecs_idxs.argtypes = [C.c_int64, c_double_p, C.c_int64, c_double_p,
                                          C.POINTER(C.c_int64)]
        L.orc_electronic_reps.argtypes = [C.POINTER(OrcWfc), C.POINTER(Lattice), c_double_p,
                                          c_double_p, C.c_int, C.c_int64, c_double_p]
        i64p = C.POINTER(C.c_int64)
        L.orc_get_fft_box.argtypes = [C.c_double, c_double_p, i64p]
        L.orc_ws_vectors.restype = C.c_int64
        L.orc_ws_vectors.argtypes = [i64p, c_double_p, C.c_double, c_double_p, C.c_int64,
                                     c_double_p, C.c_int64, C.c_int64, i64p, i64p]
        L.orc_dVlong_range_kernel.argtypes = [c_double_p, c_double_p, C.c_int64, c_double_p,
                                              c_double_p, c_double_p, c_double_p, C.c_int64,
                                              c_double_p, C.c_char, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_double, c_double_p]
        L.orc_dV_add_longrange.argtypes = [c_double_p, C.POINTER(Lattice), C.POINTER(LongRange),
                                           c_double_p, C.c_int, c_double_p]
        L.orc_multiply_eikr.argtypes = [c_double_p, c_double_p, i64p, C.c_int64, C.c_int]
        L.orc_dvscf_change_basis.argtypes = [c_double_p, c_double_p, C.c_int64, C.c_int64,
                                             C.c_int64, C.c_int64, C.c_char]
        L.orc_rotate_dvscf.argtypes = [c_double_p, c_double_p, C.POINTER(Symmetry),
                                       C.POINTER(Lattice), C.c_int, c_double_p]
        L.orc_fft_q2R.argtypes = [c_double_p, i64p, C.c_int64]
        L.orc_fft_R2q_dvscf.argtypes = [c_double_p, c_double_p, i64p, C.c_int64, C.c_int64,
                                        i64p, C.c_int64, i64p, c_double_p]
        L.orc_find_qpt_grid.argtypes = [C.c_int64, c_double_p, i64p]
        L.orc_sorted_qpts_idxs.argtypes = [C.c_int64, c_double_p, i64p]
        L.orc_pol_vecs_to_dyn.argtypes = [c_double_p, C.c_int64, c_double_p, c_double_p]
        L.orc_dyn_lr_asr.argtypes = [C.POINTER(Lattice), c_double_p, c_double_p, c_double_p, i64p,
                                     C.c_double, c_double_p]
        L.orc_add_ph_dyn_long_range.argtypes = [c_double_p, C.POINTER(Lattice), c_double_p,
                                                c_double_p, c_double_p, i64p, C.c_int,
                                                c_double_p, c_double_p, C.c_double, c_double_p]
        L.orc_fft_R2q_dyn.argtypes = [c_double_p, c_double_p, i64p, C.c_int64, i64p, C.c_int64,
                                      i64p, c_double_p]
        L.orc_mass_normalize_fc.argtypes = [c_double_p, C.c_int64, C.c_int64, C.c_double,
                                            c_double_p]
        L.orc_asr_simple_fc.argtypes = [i64p, C.c_int64, c_double_p]
        L.orc_read_dvscf_qe.argtypes = [C.c_char_p, C.POINTER(Lattice), c_double_p, c_double_p,
                                        c_double_p]
        L.orc_write_dvscf_qe.argtypes = [C.c_char_p, C.POINTER(Lattice), c_double_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_inner_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle call failed with status {rc}")


def fft3d(box: np.ndarray, sign: int) -> np.ndarray:
    b = np.ascontiguousarray(box, np.complex128).copy()
    dims = (C.c_int64 * 3)(*b.shape)
    check(lib().orc_fft3d(dptr(b), dims, sign))
    return b


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def set_inner_threads(n: int):
    lib().orc_set_inner_threads(int(n))


class OracleSystem:
    """Holds the packed structs and the per-k WFC table of a synth.System."""

    def __init__(self, sys):
        self.sys = sys
        self.packed = pack_system(sys)
        self.wtab = (OrcWfc * len(sys.wfcs))()
        for i, w in enumerate(sys.wfcs):
            if w is None:  # a k the caller did not generate: never touched
                continue
            self.wtab[i].wfc = dptr(w.wfc)
            self.wtab[i].gvec = dptr(w.gvec)
            self.wtab[i].Fk = dptr(w.Fk)
            self.wtab[i].npw = w.npw
        self.table = None

    @property
    def lat(self):
        return C.byref(self.packed.lattice)

    @property
    def pp(self):
        return C.byref(self.packed.pseudo)

    def vloc_table(self):
        if self.table is None:
            t = OrcTable()
            check(lib().orc_create_vlocg_table(self.lat, self.pp, C.byref(t)))
            self.table = t
        t = self.table
        n = t.npts
        nt = self.sys.pseudo.ntype
        g = np.ctypeslib.as_array(t.g_co, (n,)).copy()
        v = np.ctypeslib.as_array(t.vlocg, (nt * n,)).reshape(nt, n).copy()
        d = np.ctypeslib.as_array(t.vploc, (nt * n,)).reshape(nt, n).copy()
        return g, v, d

    def dvlocq(self, iq: int, qpt=None) -> np.ndarray:
        self.vloc_table()
        s = self.sys
        q = np.ascontiguousarray(s.qpts[iq] if qpt is None else qpt, np.float64)
        out = np.zeros((s.nmodes,) + tuple(s.fft_dims), np.complex128)
        check(lib().orc_dVlocq(dptr(q), self.lat, self.pp, C.byref(self.table),
                               dptr(s.eigvec(iq)), dptr(out)))
        return out

    def dv_total(self, iq: int, kernel: str = "dfpt") -> np.ndarray:
        """dvscf (mode basis) + bare local, as elph_driver builds it."""
        s = self.sys
        if kernel in ("dfpt", "dfpt_local"):
            dv = s.dvscf(iq).copy()
        else:
            dv = np.zeros((s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        vl = self.dvlocq(iq)
        lib().orc_add_dvscf_qe(dptr(dv), dptr(vl), self.lat)
        return dv

    def elph_local(self, iq, ikq, ik, kqsym, ksym, dv) -> np.ndarray:
        s = self.sys
        q = np.ascontiguousarray(s.qpts[iq], np.float64)
        out = np.zeros((s.nmodes, s.nspin, s.nbnds, s.nbnds), np.complex128)
        dv = np.ascontiguousarray(dv, np.complex128)
        check(lib().orc_elphLocal(dptr(q), self.wtab, self.lat, ikq, ik, kqsym,
                                  ksym, dptr(dv), dptr(out)))
        return out

    def add_nonlocal(self, iq, ikq, ik, kqsym, ksym, g) -> np.ndarray:
        g = np.ascontiguousarray(g, np.complex128).copy()
        check(lib().orc_add_elphNonLocal(self.wtab, self.lat, self.pp, ikq, ik, kqsym,
                                         ksym, dptr(self.sys.eigvec(iq)), dptr(g)))
        return g

    def compute_elphq(self, iq, dv, non_loc=True, kminusq=False, k_first=0,
                      k_count=-1) -> np.ndarray:
        s = self.sys
        nbz = len(s.kpt_bz_crys)
        if k_count < 0:
            k_count = nbz - k_first
        q = np.ascontiguousarray(s.qpts[iq], np.float64)
        out = np.zeros((k_count, s.nmodes, s.nspin, s.nbnds, s.nbnds), np.complex128)
        dv = np.ascontiguousarray(dv, np.complex128)
        check(lib().orc_compute_elphq(dptr(q), self.wtab, self.lat, self.pp,
                                      dptr(s.eigvec(iq)), dptr(dv), int(non_loc),
                                      int(kminusq), k_first, k_count, dptr(out)))
        return out

    def electronic_reps(self, Rsym, tauR, tim_rev, ikbz) -> np.ndarray:
        s = self.sys
        R = np.ascontiguousarray(Rsym, np.float64).reshape(9)
        t = np.ascontiguousarray(tauR, np.float64).reshape(3)
        out = np.zeros((s.nspin, s.nbnds, s.nbnds), np.complex128)
        check(lib().orc_electronic_reps(self.wtab, self.lat, dptr(R), dptr(t),
                                        int(bool(tim_rev)), int(ikbz), dptr(out)))
        return out

    def __del__(self):
        if self.table is not None:
            try:
                lib().orc_free_vloc_table(C.byref(self.table))
            except Exception:
                pass


# ---------------------------------------------------------------------------
# Fourier interpolation of dV_q (oracle/interp_oracle.c), composed in the
# order of src/interpolation/interpolation_driver.c:190-420, 790-835.
# ---------------------------------------------------------------------------

def make_sym(R, tau, tr) -> Symmetry:
    s = Symmetry()
    Rf = np.asarray(R, np.float64).reshape(9)
    for i in range(9):
        s.Rmat[i] = float(Rf[i])
    for i in range(3):
        s.tau[i] = float(np.asarray(tau, np.float64)[i])
    s.time_rev = int(bool(tr))
    return s


def ws_vectors(grid, lat, eps, rvec_m=None, rvec_n=None):
    g = np.ascontiguousarray(grid, np.int64)
    a = np.ascontiguousarray(lat, np.float64)
    rm = None if rvec_m is None else np.ascontiguousarray(rvec_m, np.float64)
    rn = None if rvec_n is None else np.ascontiguousarray(rvec_n, np.float64)
    nm = 1 if rm is None else len(rm)
    nn = 1 if rn is None else len(rn)
    degen = np.zeros(int(np.prod(g)) * nm * nn, np.int64)
    cap = 64 * len(degen)
    vecs = np.zeros((cap, 3), np.int64)
    n = lib().orc_ws_vectors(i64ptr(g), dptr(a), float(eps), None if rm is None else dptr(rm),
                             nm, None if rn is None else dptr(rn), nn, cap, i64ptr(vecs),
                             i64ptr(degen))
    if n < 0:
        raise RuntimeError(f"orc_ws_vectors failed ({n})")
    return vecs[:n].copy(), degen


def lr_kernel(qcart, gvecs, Zvals, eps, Zeu, Qpole, atom_pos, dimension, volume, zlat,
              EcutRy, eta_bare, eta_ind):
    """dVlong_range_kernel: (npw, natom, 3) complex."""
    g = np.ascontiguousarray(gvecs, np.float64)
    pos = np.ascontiguousarray(atom_pos, np.float64)
    na = len(pos)
    out = np.zeros((len(g), na, 3), np.complex128)
    opt = lambda a: None if a is None else dptr(np.ascontiguousarray(a, np.float64))  # noqa
    keep = [None if a is None else np.ascontiguousarray(a, np.float64)
            for a in (Zvals, eps, Zeu, Qpole)]
    lib().orc_dVlong_range_kernel(dptr(np.ascontiguousarray(qcart, np.float64)), dptr(g),
                                  len(g), *(None if a is None else dptr(a) for a in keep), na,
                                  dptr(pos), dimension.encode(), float(volume), float(zlat),
                                  float(EcutRy), float(eta_bare), float(eta_ind), dptr(out))
    del opt
    return out


class OracleInterp:
    """The oracle's Fourier interpolation for a synth.System-like object."""

    def __init__(self, sys, qgrid, masses=None, lr: LrParams | None = None):
        self.sys = sys
        self.packed = pack_system(sys)
        self.grid = np.ascontiguousarray(qgrid, np.int64)
        self.nq = int(np.prod(self.grid))
        self.masses = masses
        self.lr = lr
        self.plr = None if lr is None else lr.pack()
        self.ws, self.degen = ws_vectors(self.grid, sys.alat, 1e-6, sys.atomic_pos)
        nm, nmag = sys.nmodes, sys.nmag
        self.shape = (nm, nmag) + tuple(sys.fft_dims)
        self.data = np.zeros((self.nq,) + self.shape, np.complex128)
        self.have = np.zeros(self.nq, bool)
        self.last = None
        self.ready = False

    def slot(self, q):
        idx = []
        for i in range(3):
            v = (q[i] - np.floor(q[i])) * self.grid[i]
            idx.append(int(np.rint(v)) % int(self.grid[i]))
        return (idx[0] * self.grid[1] + idx[1]) * self.grid[2] + idx[2]

    def longrange(self, q, dv, eig, sign):
        dv = np.ascontiguousarray(dv, np.complex128).copy()
        e = None if eig is None else np.ascontiguousarray(eig, np.complex128)
        check(lib().orc_dV_add_longrange(dptr(np.ascontiguousarray(q, np.float64)),
                                         C.byref(self.packed.lattice), C.byref(self.plr),
                                         None if e is None else dptr(e), int(sign), dptr(dv)))
        return dv

    def _store(self, q, eig, dv):
        s = self.sys
        rot = np.ascontiguousarray(eig, np.complex128).copy()
        if self.masses is not None:
            rot = rot * np.repeat(np.asarray(self.masses, np.float64), 3)[None, :]
        d = np.ascontiguousarray(dv, np.complex128).copy()
        lib().orc_dvscf_change_basis(dptr(d), dptr(rot), 1, s.nmodes, s.nmag, s.nfft, b"C")
        dims = np.ascontiguousarray(s.fft_dims, np.int64)
        lib().orc_multiply_eikr(dptr(d), dptr(np.ascontiguousarray(q, np.float64)), i64ptr(dims),
                                s.nmodes * s.nmag, 1)
        i = self.slot(q)
        assert not self.have[i]
        self.data[i] = d
        self.have[i] = True

    def add_coarse(self, q, eig, dv):
        d = np.ascontiguousarray(dv, np.complex128).copy()
        if self.lr is not None:
            d = self.longrange(q, d, eig, -1)
        self.last = (np.asarray(q, np.float64), d)
        self._store(q, eig, d)

    def rotate(self, q, dv, sym, composite=False):
        out = np.empty_like(dv)
        check(lib().orc_rotate_dvscf(dptr(np.ascontiguousarray(q, np.float64)),
                                     dptr(np.ascontiguousarray(dv, np.complex128)),
                                     C.byref(make_sym(*sym)), C.byref(self.packed.lattice),
                                     int(bool(composite)), dptr(out)))
        return out

    def add_star(self, sym, qSq, eigSq):
        q0, d0 = self.last
        self._store(qSq, eigSq, self.rotate(q0, d0, sym, False))

    def finish(self):
        assert self.have.all()
        flat = np.ascontiguousarray(self.data.reshape(self.nq, -1))
        lib().orc_fft_q2R(dptr(flat), i64ptr(self.grid), flat.shape[1])
        self.dataR = flat
        self.ready = True

    def dvscf(self, q):
        s = self.sys
        out = np.zeros(self.shape, np.complex128)
        check(lib().orc_fft_R2q_dvscf(dptr(self.dataR), dptr(np.ascontiguousarray(q, np.float64)),
                                      i64ptr(self.grid), s.natom, s.nmag * s.nfft,
                                      i64ptr(self.ws), len(self.ws), i64ptr(self.degen),
                                      dptr(out)))
        dims = np.ascontiguousarray(s.fft_dims, np.int64)
        lib().orc_multiply_eikr(dptr(out), dptr(np.ascontiguousarray(q, np.float64)),
                                i64ptr(dims), s.nmodes * s.nmag, -1)
        if self.lr is not None:
            out = self.longrange(q, out, None, 1)
        return out


def _opt(a):
    return None if a is None else dptr(np.ascontiguousarray(a, np.float64))


_ref_lib = None


def ref_lib():
    """oracle/_ref/liblelph_ref.so: the reference's own helper TUs (built
    here by `make -C oracle ref`); raises if it is not built"""
    global _ref_lib
    if _ref_lib is None:
        if not os.path.exists(REF_LIB):
            raise RuntimeError(f"{REF_LIB} missing: run `make -C oracle ref`")
        L = C.CDLL(REF_LIB)
        ll = C.c_longlong
        L.apply_acoustic_sum_rule_fc.argtypes = [C.c_int, C.POINTER(ll), ll, c_double_p,
                                                 c_double_p, c_double_p, C.POINTER(ll), ll,
                                                 C.POINTER(ll)]
        L.apply_acoustic_sum_rule_fc.restype = None
        _ref_lib = L
    return _ref_lib


def ref_asr_fc(mode, grid, natom, fc, pos, lat, ws, degen):
    """apply_acoustic_sum_rule_fc (src/phonon/asr.c:269-512) of the
    reference, in place on fc (ngrid, nm * nm) complex"""
    L = ref_lib()
    ll = C.POINTER(C.c_longlong)
    g = np.ascontiguousarray(grid, np.int64)
    p = np.ascontiguousarray(pos, np.float64)
    a = np.ascontiguousarray(lat, np.float64)
    w = np.ascontiguousarray(ws, np.int64)
    d = np.ascontiguousarray(degen, np.int64)
    L.apply_acoustic_sum_rule_fc(int(mode), g.ctypes.data_as(ll), int(natom), dptr(fc), dptr(p),
                                 dptr(a), w.ctypes.data_as(ll), len(w), d.ctypes.data_as(ll))


class OracleDynInterp:
    """The oracle's dynamical-matrix interpolation, in the order of
    interpolation_driver.c:190-530 (coarse) and :855-992 (fine), with
    numpy's LAPACK eigh standing in for the reference's zheev."""

    def __init__(self, sys, qgrid, masses=None, lr: LrParams | None = None, eta_ph=1.0,
                 ecut=30.0, asr="simple"):
        self.sys = sys
        self.packed = pack_system(sys)
        self.grid = np.ascontiguousarray(qgrid, np.int64)
        self.nq = int(np.prod(self.grid))
        self.masses = None if masses is None else np.ascontiguousarray(masses, np.float64)
        self.lr = lr
        self.eps = None if lr is None else lr.epsilon
        self.zb = None if lr is None else lr.Zborn
        self.qp = None if lr is None else lr.Qpole
        self.eta = float(eta_ph)
        self.asr = asr
        nm = sys.nmodes
        self.Gg = np.zeros(3, np.int64)
        lib().orc_get_fft_box(float(ecut), dptr(np.ascontiguousarray(sys.blat)), i64ptr(self.Gg))
        self.ws, self.degen = ws_vectors(self.grid, sys.alat, 1e-6, sys.atomic_pos, sys.atomic_pos)
        self.asr_lr = np.zeros(9 * sys.natom, np.complex128)
        lib().orc_dyn_lr_asr(C.byref(self.packed.lattice), _opt(self.eps), _opt(self.zb),
                             _opt(self.qp), i64ptr(self.Gg), self.eta, dptr(self.asr_lr))
        self.dyns = np.zeros((self.nq, nm, nm), np.complex128)
        self.lrs = np.zeros((self.nq, nm, nm), np.complex128)
        self.have = np.zeros(self.nq, bool)

    def slot(self, q):
        idx = [int(np.rint((q[i] - np.floor(q[i])) * self.grid[i])) % int(self.grid[i])
               for i in range(3)]
        return (idx[0] * self.grid[1] + idx[1]) * self.grid[2] + idx[2]

    def lr_term(self, q, dyn, sign, masses=True, asr=True):
        d = np.ascontiguousarray(dyn, np.complex128).copy()
        check(lib().orc_add_ph_dyn_long_range(
            dptr(np.ascontiguousarray(q, np.float64)), C.byref(self.packed.lattice),
            _opt(self.eps), _opt(self.zb), _opt(self.qp), i64ptr(self.Gg), int(sign),
            _opt(self.masses) if masses else None, dptr(self.asr_lr) if asr else None,
            self.eta, dptr(d)))
        return d

    def add_coarse(self, q, omega, pol):
        d = np.ascontiguousarray(pol, np.complex128).copy()
        lib().orc_pol_vecs_to_dyn(dptr(np.ascontiguousarray(omega, np.float64)), self.sys.natom,
                                  _opt(self.masses), dptr(d))
        d = self.lr_term(q, d, -1)
        i = self.slot(q)
        assert not self.have[i]
        self.dyns[i] = d
        self.lrs[i] = self.lr_term(q, np.zeros_like(d), 1)
        self.have[i] = True

    def finish(self):
        assert self.have.all()
        nm = self.sys.nmodes
        fc = np.ascontiguousarray(self.dyns.reshape(self.nq, nm * nm)).copy()
        lib().orc_fft_q2R(dptr(fc), i64ptr(self.grid), nm * nm)
        m = _opt(self.masses)
        rot = self.asr in ("all", "all_huang") and self.eps is not None and (
            self.zb is not None or self.qp is not None)
        if rot:
            # the long-range force constants join before the rule and leave
            # after it (interpolation_driver.c:432-442,499-511,533-545)
            fl = np.ascontiguousarray(self.lrs.reshape(self.nq, nm * nm)).copy()
            lib().orc_fft_q2R(dptr(fl), i64ptr(self.grid), nm * nm)
            fc += fl
        lib().orc_mass_normalize_fc(m, self.nq, self.sys.natom, 0.5, dptr(fc))
        if self.asr == "simple":
            lib().orc_asr_simple_fc(i64ptr(self.grid), self.sys.natom, dptr(fc))
        elif self.asr in ("crystal", "all", "all_huang"):
            # the reference's own apply_acoustic_sum_rule_fc (LSMR), compiled
            # into oracle/_ref
            ref_asr_fc({"crystal": 2, "all": 3, "all_huang": 4}[self.asr], self.grid,
                       self.sys.natom, fc, self.sys.atomic_pos, self.sys.alat, self.ws,
                       self.degen)
        lib().orc_mass_normalize_fc(m, self.nq, self.sys.natom, -0.5, dptr(fc))
        if rot:
            fc -= fl
        self.fc = fc.reshape(self.nq, nm, nm)

    def dyn(self, q):
        nm = self.sys.nmodes
        out = np.zeros((nm, nm), np.complex128)
        check(lib().orc_fft_R2q_dyn(dptr(np.ascontiguousarray(self.fc)),
                                    dptr(np.ascontiguousarray(q, np.float64)), i64ptr(self.grid),
                                    self.sys.natom, i64ptr(self.ws), len(self.ws),
                                    i64ptr(self.degen), dptr(out)))
        return self.lr_term(q, out, 1)

    def modes(self, dyn):
        """symmetrised lower triangle, eigh, sign(w) sqrt|w|, rows / sqrt(M)"""
        lo = np.tril(0.5 * (dyn + dyn.conj().T))
        h = lo + np.tril(lo, -1).conj().T
        w, v = np.linalg.eigh(h)
        f = np.sign(w) * np.sqrt(np.abs(w))
        e = v.T.copy()
        if self.masses is not None:
            e = e / np.sqrt(np.repeat(self.masses, 3))[None, :]
        return f, e
