"""Python front end of liblelph_hip.so (ctypes), mirroring the reference's
operator calls for the e-ph hot path:

  Context.dvlocq          dVlocq            src/dvloc/dVbare_loc.c:16-231
  Context.set_dvscf       get_dvscf_dyn_qe (pattern->mode) + dVlocq +
                          add_dvscf_qe      src/elph/elph_driver.c:210-245
  Context.elph_local      elphLocal         src/dvloc/compute_Vbare_loc.c:17-353
  Context.add_nonlocal    add_elphNonLocal  src/nonloc/Vbare_nloc.c:35-577
  Context.compute_elphq   the k loop of compute_and_write_elphq
                          src/elph/compute_elph_matq.c:20-234

The library must be the compiled HIP build: there is no fallback.  Errors
come back as LelphError carrying the C status and its text.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (KERNELS, PRECISIONS, Lattice, Pair, Pseudo, c_double_p, dptr,
                  pack_system)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblelph_hip.so")

EXPORTS = [
    "lelph_create", "lelph_destroy", "lelph_strerror", "lelph_set_precision",
    "lelph_get_device_count", "lelph_upload_wfc", "lelph_create_vlocg_table",
    "lelph_dvlocq", "lelph_set_dvscf", "lelph_get_dvscf", "lelph_elph_local",
    "lelph_add_elph_nonlocal", "lelph_gkq", "lelph_gkq_batch", "lelph_compute_elphq",
    "lelph_elph_q_rotate", "lelph_get_kplusq_idxs", "lelph_last_timings",
    "lelph_electronic_reps", "lelph_elph_q_rotate_gpu",
]


class LelphError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: status {status} ({strerror(status)})")


_lib = None


def lib():
    """Loads liblelph_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run letzelphc_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        L.lelph_strerror.restype = C.c_char_p
        L.lelph_strerror.argtypes = [C.c_int]
        L.lelph_create.argtypes = [C.POINTER(Lattice), C.POINTER(Pseudo), C.c_int,
                                   C.POINTER(C.c_void_p)]
        L.lelph_destroy.argtypes = [C.c_void_p]
        L.lelph_destroy.restype = None
        L.lelph_set_precision.argtypes = [C.c_void_p, C.c_int]
        L.lelph_get_device_count.argtypes = [C.POINTER(C.c_int)]
        L.lelph_upload_wfc.argtypes = [C.c_void_p, C.c_int64, c_double_p, c_double_p,
                                       c_double_p, C.c_int64]
        L.lelph_create_vlocg_table.argtypes = [C.c_void_p, C.POINTER(C.c_int64), c_double_p,
                                               c_double_p, c_double_p]
        L.lelph_dvlocq.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p]
        L.lelph_set_dvscf.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p,
                                      c_double_p, C.c_int]
        L.lelph_get_dvscf.argtypes = [C.c_void_p, c_double_p]
        L.lelph_elph_local.argtypes = [C.c_void_p, c_double_p, C.c_int, C.c_int, C.c_int,
                                       C.c_int, c_double_p]
        L.lelph_add_elph_nonlocal.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                              c_double_p, c_double_p]
        L.lelph_gkq.argtypes = [C.c_void_p, c_double_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_int, c_double_p]
        L.lelph_gkq_batch.argtypes = [C.c_void_p, c_double_p, C.c_int64, C.POINTER(Pair),
                                      C.c_int, c_double_p]
        L.lelph_compute_elphq.argtypes = [C.c_void_p, c_double_p, C.c_int, C.c_int, c_double_p]
        L.lelph_elph_q_rotate.argtypes = [c_double_p, c_double_p, c_double_p, C.c_int,
                                          C.c_int, C.c_int, C.c_int, c_double_p]
        L.lelph_get_kplusq_idxs.argtypes = [C.c_int64, c_double_p, c_double_p,
                                            C.POINTER(C.c_int)]
        L.lelph_last_timings.argtypes = [C.c_void_p, c_double_p, C.c_int]
        L.lelph_electronic_reps.argtypes = [C.c_void_p, c_double_p, c_double_p, C.c_int,
                                            C.c_int64, c_double_p]
        L.lelph_elph_q_rotate_gpu.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p,
                                              C.c_int, C.c_int, C.c_int, C.c_int, c_double_p]
        _lib = L
    return _lib


def strerror(status: int) -> str:
    return lib().lelph_strerror(int(status)).decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise LelphError(rc, what)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().lelph_get_device_count(C.byref(n)), "lelph_get_device_count")
    return n.value


class Context:
    """One device context for a synth.System (or any object with the same
    fields).  Uploads every iBZ wavefunction at construction."""

    def __init__(self, system, precision: str = "fp64", upload: bool = True):
        self.sys = system
        self.packed = pack_system(system)
        h = C.c_void_p()
        _check(lib().lelph_create(C.byref(self.packed.lattice), C.byref(self.packed.pseudo),
                                  PRECISIONS[precision], C.byref(h)), "lelph_create")
        self.h = h
        self.precision = precision
        if upload:
            for ik, w in enumerate(system.wfcs):
                self.upload_wfc(ik, w.wfc, w.gvec, w.Fk)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().lelph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p: str):
        _check(lib().lelph_set_precision(self.h, PRECISIONS[p]), "lelph_set_precision")
        self.precision = p

    def upload_wfc(self, ik, wfc, gvec, Fk=None):
        wfc = np.ascontiguousarray(wfc, np.complex128)
        gvec = np.ascontiguousarray(gvec, np.float64)
        fk = None if Fk is None else dptr(np.ascontiguousarray(Fk, np.float64))
        _check(lib().lelph_upload_wfc(self.h, int(ik), dptr(wfc), dptr(gvec), fk,
                                      int(gvec.shape[0])), "lelph_upload_wfc")

    def vloc_table(self):
        n = C.c_int64(0)
        _check(lib().lelph_create_vlocg_table(self.h, C.byref(n), None, None, None),
               "lelph_create_vlocg_table")
        nt = self.sys.pseudo.ntype
        g = np.zeros(n.value)
        v = np.zeros((nt, n.value))
        d = np.zeros((nt, n.value))
        _check(lib().lelph_create_vlocg_table(self.h, C.byref(n), dptr(g), dptr(v), dptr(d)),
               "lelph_create_vlocg_table")
        return g, v, d

    def dvlocq(self, qpt, eig) -> np.ndarray:
        s = self.sys
        out = np.zeros((s.nmodes,) + tuple(s.fft_dims), np.complex128)
        q = np.ascontiguousarray(qpt, np.float64)
        _check(lib().lelph_dvlocq(self.h, dptr(q), dptr(np.ascontiguousarray(eig, np.complex128)),
                                  dptr(out)), "lelph_dvlocq")
        return out

    def set_dvscf(self, qpt, eig, dvscf=None, patterns=None, kernel: str = "dfpt"):
        q = np.ascontiguousarray(qpt, np.float64)
        e = np.ascontiguousarray(eig, np.complex128)
        dv = None if dvscf is None else dptr(np.ascontiguousarray(dvscf, np.complex128))
        pt = None if patterns is None else dptr(np.ascontiguousarray(patterns, np.complex128))
        _check(lib().lelph_set_dvscf(self.h, dptr(q), dptr(e), dv, pt, KERNELS[kernel]),
               "lelph_set_dvscf")
        self.q = q
        self.eig = e

    def get_dvscf(self) -> np.ndarray:
        s = self.sys
        out = np.zeros((s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        _check(lib().lelph_get_dvscf(self.h, dptr(out)), "lelph_get_dvscf")
        return out

    def _gshape(self):
        s = self.sys
        return (s.nmodes, s.nspin, s.nbnds, s.nbnds)

    def elph_local(self, ikq, ik, kqsym, ksym, qpt=None) -> np.ndarray:
        q = np.ascontiguousarray(self.q if qpt is None else qpt, np.float64)
        out = np.zeros(self._gshape(), np.complex128)
        _check(lib().lelph_elph_local(self.h, dptr(q), ikq, ik, kqsym, ksym, dptr(out)),
               "lelph_elph_local")
        return out

    def add_nonlocal(self, ikq, ik, kqsym, ksym, g, eig=None) -> np.ndarray:
        g = np.ascontiguousarray(g, np.complex128).copy()
        e = np.ascontiguousarray(self.eig if eig is None else eig, np.complex128)
        _check(lib().lelph_add_elph_nonlocal(self.h, ikq, ik, kqsym, ksym, dptr(e), dptr(g)),
               "lelph_add_elph_nonlocal")
        return g

    def gkq(self, ikq, ik, kqsym, ksym, non_loc=True, qpt=None) -> np.ndarray:
        q = np.ascontiguousarray(self.q if qpt is None else qpt, np.float64)
        out = np.zeros(self._gshape(), np.complex128)
        _check(lib().lelph_gkq(self.h, dptr(q), ikq, ik, kqsym, ksym, int(non_loc), dptr(out)),
               "lelph_gkq")
        return out

    def gkq_batch(self, pairs, non_loc=True, qpt=None, out=None) -> np.ndarray:
        q = np.ascontiguousarray(self.q if qpt is None else qpt, np.float64)
        n = len(pairs)
        arr = (Pair * max(n, 1))()
        for i, (ikq, ik, kqsym, ksym) in enumerate(pairs):
            arr[i].ikq, arr[i].ik, arr[i].kqsym, arr[i].ksym = ikq, ik, kqsym, ksym
        if out is None:
            out = np.zeros((n,) + self._gshape(), np.complex128)
        _check(lib().lelph_gkq_batch(self.h, dptr(q), n, arr, int(non_loc), dptr(out)),
               "lelph_gkq_batch")
        return out

    def compute_elphq(self, qpt=None, non_loc=True, kminusq=False) -> np.ndarray:
        q = np.ascontiguousarray(self.q if qpt is None else qpt, np.float64)
        nbz = len(self.sys.kpt_bz_crys)
        out = np.zeros((nbz,) + self._gshape(), np.complex128)
        _check(lib().lelph_compute_elphq(self.h, dptr(q), int(non_loc), int(kminusq), dptr(out)),
               "lelph_compute_elphq")
        return out

    def electronic_reps(self, Rsym, tauR, tim_rev, ikbz) -> np.ndarray:
        """D (nspin, nb, nb) = <R k, b | U(R) | k, a> at full-BZ point ikbz
        (electronic_reps, src/symmetries/elec_reps.c:16-413)."""
        s = self.sys
        R = np.ascontiguousarray(Rsym, np.float64).reshape(9)
        t = np.ascontiguousarray(tauR, np.float64).reshape(3)
        out = np.zeros((s.nspin, s.nbnds, s.nbnds), np.complex128)
        _check(lib().lelph_electronic_reps(self.h, dptr(R), dptr(t), int(bool(tim_rev)),
                                           int(ikbz), dptr(out)), "electronic_reps")
        return out

    def elph_q_rotate(self, Dl, g, Dr, tim_rev: bool) -> np.ndarray:
        """q-star rotation on the device (elph_q_rotate,
        src/symmetries/elphq_expand.c:16-73); g (nmodes, nspin, nb, nb)."""
        g = np.ascontiguousarray(g, np.complex128)
        Dl = np.ascontiguousarray(Dl, np.complex128)
        Dr = np.ascontiguousarray(Dr, np.complex128)
        nm, ns, nb, _ = g.shape
        out = np.zeros_like(g)
        _check(lib().lelph_elph_q_rotate_gpu(self.h, dptr(Dl), dptr(g), dptr(Dr), nm, ns, nb,
                                             int(bool(tim_rev)), dptr(out)),
               "elph_q_rotate_gpu")
        return out

    def last_timings(self) -> dict:
        """Device milliseconds of the last call, summed over its pairs."""
        t = np.zeros(6)
        _check(lib().lelph_last_timings(self.h, dptr(t), 6), "lelph_last_timings")
        return {"boxes_ms": t[0], "contract_ms": t[1], "reduce_ms": t[2],
                "nonlocal_ms": t[3], "total_ms": t[4], "pairs": int(t[5])}


def elph_q_rotate(Dl, g, Dr, tim_rev: bool) -> np.ndarray:
    """elph_q_rotate (src/symmetries/elphq_expand.c:16-73) through the C-ABI."""
    g = np.ascontiguousarray(g, np.complex128)
    nm, ns, nb, _ = g.shape
    out = np.zeros_like(g)
    _check(lib().lelph_elph_q_rotate(dptr(np.ascontiguousarray(Dl, np.complex128)), dptr(g),
                                     dptr(np.ascontiguousarray(Dr, np.complex128)), nm, ns, nb,
                                     int(tim_rev), dptr(out)), "lelph_elph_q_rotate")
    return out


def kplusq_idxs(kpts_crys, q) -> np.ndarray:
    k = np.ascontiguousarray(kpts_crys, np.float64)
    out = np.zeros(len(k), np.int32)
    _check(lib().lelph_get_kplusq_idxs(len(k), dptr(k), dptr(np.ascontiguousarray(q, np.float64)),
                                       out.ctypes.data_as(C.POINTER(C.c_int))),
           "lelph_get_kplusq_idxs")
    return out
