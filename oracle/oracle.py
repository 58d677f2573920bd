"""ctypes front end of the FP64 CPU oracle (oracle/liblelph_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
bench's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from letzelphc_amd.abi import Lattice, Pseudo, c_double_p, dptr, pack_system

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
