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

from .abi import (KERNELS, PRECISIONS, Lattice, LongRange, LrParams, Pair, Pseudo, Symmetry,
                  Tuning, c_double_p, c_int_p, dptr, i64ptr, pack_system)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblelph_hip.so")

EXPORTS = [
    "lelph_create", "lelph_destroy", "lelph_strerror", "lelph_set_precision",
    "lelph_get_precision", "lelph_elph_q_rotate_noalloc", "lelph_set_tuning", "lelph_get_tuning",
    "lelph_multi_create", "lelph_multi_run", "lelph_multi_timings", "lelph_multi_destroy",
    "lelph_get_device_count", "lelph_upload_wfc", "lelph_create_vlocg_table",
    "lelph_dvlocq", "lelph_set_dvscf", "lelph_get_dvscf", "lelph_elph_local",
    "lelph_add_elph_nonlocal", "lelph_gkq", "lelph_gkq_batch", "lelph_compute_elphq",
    "lelph_elph_q_rotate", "lelph_get_kplusq_idxs", "lelph_last_timings",
    "lelph_electronic_reps", "lelph_elph_q_rotate_gpu", "lelph_dv_longrange",
    "lelph_rotate_dvscf", "lelph_wigner_seitz", "lelph_interp_begin", "lelph_interp_add_coarse",
    "lelph_interp_add_star", "lelph_interp_finish", "lelph_interp_dvscf",
    "lelph_interp_set_dvscf", "lelph_read_dvscf_qe", "lelph_set_dvscf_file",
    "lelph_write_dvscf_qe", "lelph_interp_timings", "lelph_dyn_interp_begin",
    "lelph_dyn_interp_add_coarse", "lelph_dyn_interp_finish", "lelph_dyn_interp",
    "lelph_read_dyn_qe", "lelph_read_patterns_qe", "lelph_write_dyn_qe",
    "lelph_parse_qexml", "lelph_parse_upf", "lelph_read_ph_tensors_qe", "lelph_read_qpts_qe",
    "lelph_asr_fc", "lelph_asr_born", "lelph_quadrupole_3d_to_2d", "lelph_loto_qpoint",
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
        L.lelph_get_precision.argtypes = [C.c_void_p, c_int_p, c_int_p]
        L.lelph_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.lelph_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.lelph_multi_create.argtypes = [C.c_int, C.POINTER(C.c_void_p), c_double_p,
                                         C.POINTER(C.c_int64), C.POINTER(C.POINTER(Pair)),
                                         C.c_int, C.POINTER(C.c_void_p)]
        L.lelph_multi_run.argtypes = [C.c_void_p, C.POINTER(c_double_p)]
        L.lelph_multi_timings.argtypes = [C.c_void_p, c_double_p]
        L.lelph_multi_destroy.argtypes = [C.c_void_p]
        L.lelph_multi_destroy.restype = None
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
        i64p = C.POINTER(C.c_int64)
        L.lelph_dv_longrange.argtypes = [C.c_void_p, c_double_p, C.POINTER(LongRange),
                                         c_double_p, C.c_int, c_double_p]
        L.lelph_rotate_dvscf.argtypes = [C.c_void_p, c_double_p, c_double_p,
                                         C.POINTER(Symmetry), C.c_int, c_double_p]
        L.lelph_wigner_seitz.argtypes = [i64p, c_double_p, C.c_double, c_double_p, C.c_int64,
                                         c_double_p, C.c_int64, C.c_int64, i64p, i64p, i64p]
        L.lelph_interp_begin.argtypes = [C.c_void_p, i64p, c_double_p, C.POINTER(LongRange)]
        L.lelph_interp_add_coarse.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p]
        L.lelph_interp_add_star.argtypes = [C.c_void_p, C.POINTER(Symmetry), c_double_p,
                                            c_double_p]
        L.lelph_interp_finish.argtypes = [C.c_void_p]
        L.lelph_interp_dvscf.argtypes = [C.c_void_p, C.c_int64, c_double_p, c_double_p]
        L.lelph_interp_set_dvscf.argtypes = [C.c_void_p, c_double_p, c_double_p, C.c_int]
        L.lelph_read_dvscf_qe.argtypes = [C.c_void_p, C.c_char_p, c_double_p, c_double_p,
                                          c_double_p]
        L.lelph_set_dvscf_file.argtypes = [C.c_void_p, c_double_p, c_double_p, C.c_char_p,
                                           c_double_p, C.c_int]
        L.lelph_write_dvscf_qe.argtypes = [C.c_void_p, C.c_char_p, c_double_p]
        L.lelph_interp_timings.argtypes = [C.c_void_p, c_double_p, C.c_int]
        L.lelph_dyn_interp_begin.argtypes = [C.c_void_p, i64p, c_double_p, C.POINTER(LongRange),
                                             C.c_double, C.c_double, C.c_int]
        L.lelph_dyn_interp_add_coarse.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p]
        L.lelph_dyn_interp_finish.argtypes = [C.c_void_p]
        L.lelph_read_dyn_qe.argtypes = [C.c_void_p, C.c_char_p, c_double_p, c_double_p,
                                        c_double_p, c_double_p]
        L.lelph_read_patterns_qe.argtypes = [C.c_char_p, C.c_int, c_double_p]
        L.lelph_parse_qexml.argtypes = [C.c_char_p, C.POINTER(QexmlInfo), c_int_p, c_double_p,
                                        C.c_char_p, c_double_p, c_double_p, c_int_p]
        i64p = C.POINTER(C.c_int64)
        L.lelph_parse_upf.argtypes = [C.c_char_p, i64p, c_double_p, c_double_p, c_double_p,
                                      c_double_p]
        L.lelph_read_ph_tensors_qe.argtypes = [C.c_char_p, C.c_char_p, C.c_int, c_double_p,
                                               c_double_p, c_double_p, c_int_p]
        L.lelph_read_qpts_qe.argtypes = [C.c_char_p, i64p, i64p, c_double_p]
        L.lelph_asr_fc.argtypes = [C.c_int, i64p, C.c_int, c_double_p, c_double_p, c_double_p,
                                   i64p, C.c_int64, i64p]
        L.lelph_asr_born.argtypes = [C.c_int, c_double_p, C.c_int]
        L.lelph_quadrupole_3d_to_2d.argtypes = [C.c_int, c_double_p, c_double_p, c_double_p,
                                                c_double_p]
        L.lelph_loto_qpoint.argtypes = [c_double_p, c_double_p, c_double_p]
        L.lelph_write_dyn_qe.argtypes = [C.c_char_p, C.c_int, c_double_p, c_double_p,
                                         c_double_p]
        L.lelph_dyn_interp.argtypes = [C.c_void_p, C.c_int64, c_double_p, c_double_p, c_double_p,
                                       c_double_p]
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
    fields).  Uploads every iBZ wavefunction the system holds (entries left
    None are skipped) at construction."""

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
                if w is not None:
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

    def effective_precision(self) -> str:
        """The precision the contractions run in on this shape (bf16x3 runs
        as bf16x6 off the aligned producer/consumer path)."""
        req, eff = C.c_int(0), C.c_int(0)
        _check(lib().lelph_get_precision(self.h, C.byref(req), C.byref(eff)),
               "lelph_get_precision")
        return {v: k for k, v in PRECISIONS.items()}[eff.value]

    def set_tuning(self, split_target=0, fft_lines=0, fp32_form=0, ctile=0, nl_form=0):
        t = Tuning(int(split_target), int(fft_lines), int(fp32_form), int(ctile), int(nl_form))
        _check(lib().lelph_set_tuning(self.h, C.byref(t)), "lelph_set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().lelph_get_tuning(self.h, C.byref(t)), "lelph_get_tuning")
        return {"split_target": t.split_target, "fft_lines": t.fft_lines,
                "fp32_form": t.fp32_form, "ctile": t.ctile, "nl_form": t.nl_form}

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

    # ---- QE dvscf files ----
    def read_dvscf_qe(self, path, eig, patterns) -> np.ndarray:
        """read_dvscf_qe (src/io/qe/read_dvscf.c:17-172): mode-basis dV."""
        s = self.sys
        out = np.empty((s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        _check(lib().lelph_read_dvscf_qe(self.h, os.fsencode(path),
                                         dptr(np.ascontiguousarray(eig, np.complex128)),
                                         dptr(np.ascontiguousarray(patterns, np.complex128)),
                                         dptr(out)), "lelph_read_dvscf_qe")
        return out

    def set_dvscf_file(self, qpt, eig, path, patterns, kernel: str = "dfpt"):
        q = np.ascontiguousarray(qpt, np.float64)
        e = np.ascontiguousarray(eig, np.complex128)
        _check(lib().lelph_set_dvscf_file(self.h, dptr(q), dptr(e), os.fsencode(path),
                                          dptr(np.ascontiguousarray(patterns, np.complex128)),
                                          KERNELS[kernel]), "lelph_set_dvscf_file")
        self.q = q
        self.eig = e

    def write_dvscf_qe(self, path, dvscf):
        _check(lib().lelph_write_dvscf_qe(self.h, os.fsencode(path),
                                          dptr(np.ascontiguousarray(dvscf, np.complex128))),
               "lelph_write_dvscf_qe")

    # ---- Fourier interpolation of dV_q ----
    def dv_longrange(self, qpt, lr: LrParams, dvscf, eig=None, sign: int = 1) -> np.ndarray:
        """dV_add_longrange (src/dvloc/dVpot_add_long_range.c:18-241) on the
        device; returns dvscf + sign * long range."""
        out = np.ascontiguousarray(dvscf, np.complex128).copy()
        q = np.ascontiguousarray(qpt, np.float64)
        e = None if eig is None else np.ascontiguousarray(eig, np.complex128)
        plr = lr.pack()
        _check(lib().lelph_dv_longrange(self.h, dptr(q), C.byref(plr),
                                        None if e is None else dptr(e), int(sign), dptr(out)),
               "lelph_dv_longrange")
        return out

    def rotate_dvscf(self, qpt, dvscf, sym, composite: bool = False) -> np.ndarray:
        """rotate_dvscf (src/symmetries/rotate_dvscf.c:22-368) on the device;
        sym = (R, tau, time_rev)."""
        R, tau, tr = sym
        s = Symmetry()
        for i in range(9):
            s.Rmat[i] = float(np.asarray(R, np.float64).reshape(9)[i])
        for i in range(3):
            s.tau[i] = float(np.asarray(tau, np.float64)[i])
        s.time_rev = int(bool(tr))
        din = np.ascontiguousarray(dvscf, np.complex128)
        out = np.empty_like(din)
        q = np.ascontiguousarray(qpt, np.float64)
        _check(lib().lelph_rotate_dvscf(self.h, dptr(q), dptr(din), C.byref(s),
                                        int(bool(composite)), dptr(out)), "lelph_rotate_dvscf")
        return out

    def interp_begin(self, qgrid, masses=None, lr: LrParams | None = None):
        g = np.ascontiguousarray(qgrid, np.int64)
        self._ip_mass = None if masses is None else np.ascontiguousarray(masses, np.float64)
        self._ip_lr = lr
        plr = None if lr is None else lr.pack()
        self._ip_plr = plr
        _check(lib().lelph_interp_begin(self.h, i64ptr(g),
                                        None if self._ip_mass is None else dptr(self._ip_mass),
                                        None if plr is None else C.byref(plr)),
               "lelph_interp_begin")

    def interp_add_coarse(self, qpt, eig, dvscf):
        q = np.ascontiguousarray(qpt, np.float64)
        _check(lib().lelph_interp_add_coarse(self.h, dptr(q),
                                             dptr(np.ascontiguousarray(eig, np.complex128)),
                                             dptr(np.ascontiguousarray(dvscf, np.complex128))),
               "lelph_interp_add_coarse")

    def interp_add_star(self, sym, qpt_Sq, eig_Sq):
        R, tau, tr = sym
        s = Symmetry()
        for i in range(9):
            s.Rmat[i] = float(np.asarray(R, np.float64).reshape(9)[i])
        for i in range(3):
            s.tau[i] = float(np.asarray(tau, np.float64)[i])
        s.time_rev = int(bool(tr))
        q = np.ascontiguousarray(qpt_Sq, np.float64)
        _check(lib().lelph_interp_add_star(self.h, C.byref(s), dptr(q),
                                           dptr(np.ascontiguousarray(eig_Sq, np.complex128))),
               "lelph_interp_add_star")

    def interp_finish(self):
        _check(lib().lelph_interp_finish(self.h), "lelph_interp_finish")

    def interp_dvscf(self, qpts) -> np.ndarray:
        """(nq, nmodes, nmag, Nx, Ny, Nz) interpolated dV, cartesian basis."""
        q = np.ascontiguousarray(np.atleast_2d(qpts), np.float64)
        s = self.sys
        out = np.empty((len(q), s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        _check(lib().lelph_interp_dvscf(self.h, len(q), dptr(q), dptr(out)),
               "lelph_interp_dvscf")
        return out

    # ---- Fourier interpolation of dynamical matrices ----
    def dyn_interp_begin(self, qgrid, masses=None, lr: LrParams | None = None, eta_ph=1.0,
                         ecut=30.0, asr: str = "simple"):
        g = np.ascontiguousarray(qgrid, np.int64)
        self._dy_mass = None if masses is None else np.ascontiguousarray(masses, np.float64)
        self._dy_plr = None if lr is None else lr.pack()
        self._dy_lr = lr
        _check(lib().lelph_dyn_interp_begin(self.h, i64ptr(g),
                                            None if self._dy_mass is None else dptr(self._dy_mass),
                                            None if self._dy_plr is None else C.byref(self._dy_plr),
                                            float(eta_ph), float(ecut),
                                            0 if asr == "none" else asr_kind(asr)),
               "lelph_dyn_interp_begin")

    def dyn_interp_add_coarse(self, qpt, omega, pol_vecs):
        _check(lib().lelph_dyn_interp_add_coarse(
            self.h, dptr(np.ascontiguousarray(qpt, np.float64)),
            dptr(np.ascontiguousarray(omega, np.float64)),
            dptr(np.ascontiguousarray(pol_vecs, np.complex128))), "lelph_dyn_interp_add_coarse")

    def dyn_interp_finish(self):
        _check(lib().lelph_dyn_interp_finish(self.h), "lelph_dyn_interp_finish")

    def dyn_interp(self, qpts):
        """(dyn (nq,nm,nm), omega (nq,nm), eig (nq,nm,nm)) at the fine q."""
        q = np.ascontiguousarray(np.atleast_2d(qpts), np.float64)
        nm = self.sys.nmodes
        dyn = np.empty((len(q), nm, nm), np.complex128)
        om = np.empty((len(q), nm))
        eig = np.empty((len(q), nm, nm), np.complex128)
        _check(lib().lelph_dyn_interp(self.h, len(q), dptr(q), dptr(dyn), dptr(om), dptr(eig)),
               "lelph_dyn_interp")
        return dyn, om, eig

    def read_dyn_qe(self, path):
        """read_dyn_qe (src/io/qe/read_dyn.c:38-293): (q, omega, pol_vecs, masses)."""
        na, nm = self.sys.natom, self.sys.nmodes
        q = np.zeros(3)
        om = np.zeros(nm)
        pv = np.zeros((nm, nm), np.complex128)
        am = np.zeros(na)
        _check(lib().lelph_read_dyn_qe(self.h, os.fsencode(path), dptr(q), dptr(om), dptr(pv),
                                       dptr(am)), "lelph_read_dyn_qe")
        return q, om, pv, am

    def interp_timings(self) -> dict:
        t = np.zeros(5)
        _check(lib().lelph_interp_timings(self.h, dptr(t), 5), "lelph_interp_timings")
        return {"weights_ms": t[0], "r2q_ms": t[1], "longrange_ms": t[2], "total_ms": t[3],
                "nq": int(t[4])}

    def interp_set_dvscf(self, qpt, eig, kernel: str = "dfpt"):
        q = np.ascontiguousarray(qpt, np.float64)
        _check(lib().lelph_interp_set_dvscf(self.h, dptr(q),
                                            dptr(np.ascontiguousarray(eig, np.complex128)),
                                            KERNELS[kernel]), "lelph_interp_set_dvscf")
        self.q = q
        self.eig = np.ascontiguousarray(eig, np.complex128)

    def last_timings(self) -> dict:
        """Device milliseconds of the last call, summed over its pairs."""
        t = np.zeros(7)
        _check(lib().lelph_last_timings(self.h, dptr(t), 7), "lelph_last_timings")
        return {"boxes_ms": t[0], "contract_ms": t[1], "reduce_ms": t[2],
                "nonlocal_ms": t[3], "total_ms": t[4], "pairs": int(t[5]),
                "box_sides": int(t[6])}


class MultiBatch:
    """lelph_multi: the (k, k+q) pairs of several contexts (cells) as one
    unit - per-pair boxes, one persistent work-queue contraction launch per
    precision, reductions and non-local parts - captured in a HIP graph at
    construction and replayed by run().  jobs: [(Context, qpt, pairs)]."""

    def __init__(self, jobs, non_loc=True):
        self.jobs = jobs
        n = len(jobs)
        ctxs = (C.c_void_p * n)(*[j[0].h.value for j in jobs])
        q = np.ascontiguousarray(np.array([np.asarray(j[1], np.float64) for j in jobs]))
        npairs = (C.c_int64 * n)(*[len(j[2]) for j in jobs])
        self._arrs = []
        parr = (C.POINTER(Pair) * n)()
        for i, (_, _, pairs) in enumerate(jobs):
            a = (Pair * max(len(pairs), 1))()
            for k, (ikq, ik, kqsym, ksym) in enumerate(pairs):
                a[k].ikq, a[k].ik, a[k].kqsym, a[k].ksym = ikq, ik, kqsym, ksym
            self._arrs.append(a)
            parr[i] = C.cast(a, C.POINTER(Pair))
        self._q = q
        h = C.c_void_p()
        _check(lib().lelph_multi_create(n, ctxs, dptr(q), npairs, parr, int(non_loc),
                                        C.byref(h)), "lelph_multi_create")
        self.h = h
        self.out = [np.zeros((len(p),) + ctx._gshape(), np.complex128) for ctx, _, p in jobs]

    def run(self):
        ptrs = (c_double_p * len(self.out))(*[dptr(o) for o in self.out])
        _check(lib().lelph_multi_run(self.h, ptrs), "lelph_multi_run")
        return self.out

    def timings(self) -> dict:
        t = np.zeros(8)
        _check(lib().lelph_multi_timings(self.h, dptr(t)), "lelph_multi_timings")
        return {"boxes_ms": t[0], "contract_ms": t[1], "rest_ms": t[2], "replay_ms": t[3],
                "items_fp64": int(t[4]), "items_fp32": int(t[5]), "grid_fp64": int(t[6]),
                "grid_fp32": int(t[7])}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().lelph_multi_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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


def wigner_seitz(grid, lat, eps, rvec_m=None, rvec_n=None):
    """build_wigner_seitz_vectors (src/interpolation/wigner_seitz.c:27-202),
    host code of the library: (ws_vecs (nws,3) int64, ws_degen)."""
    g = np.ascontiguousarray(grid, np.int64)
    a = np.ascontiguousarray(lat, np.float64)
    rm = None if rvec_m is None else np.ascontiguousarray(rvec_m, np.float64)
    rn = None if rvec_n is None else np.ascontiguousarray(rvec_n, np.float64)
    nm = 1 if rm is None else len(rm)
    nn = 1 if rn is None else len(rn)
    degen = np.zeros(int(np.prod(g)) * nm * nn, np.int64)
    cap = 16 * len(degen)
    while True:
        vecs = np.zeros((cap, 3), np.int64)
        n = C.c_int64(0)
        rc = lib().lelph_wigner_seitz(i64ptr(g), dptr(a), float(eps),
                                      None if rm is None else dptr(rm), nm,
                                      None if rn is None else dptr(rn), nn, cap, i64ptr(vecs),
                                      i64ptr(degen), C.byref(n))
        if rc == 4:
            cap = int(n.value)
            continue
        _check(rc, "lelph_wigner_seitz")
        return vecs[: n.value].copy(), degen


def read_patterns_qe(path, natom) -> np.ndarray:
    """read_pattern_qe (src/io/qe/read_patterns.c:18-125), host code."""
    nm = 3 * natom
    out = np.zeros((nm, nm), np.complex128)
    _check(lib().lelph_read_patterns_qe(os.fsencode(path), int(natom), dptr(out)),
           "lelph_read_patterns_qe")
    return out


def write_dyn_qe(path, qpt, dyn, masses):
    """write_dyn_qe (src/io/qe/write_dyn_dvscf_qe.c:18-97), host code."""
    m = np.ascontiguousarray(masses, np.float64)
    _check(lib().lelph_write_dyn_qe(os.fsencode(path), len(m),
                                    dptr(np.ascontiguousarray(qpt, np.float64)),
                                    dptr(np.ascontiguousarray(dyn, np.complex128)), dptr(m)),
           "lelph_write_dyn_qe")


class QexmlInfo(C.Structure):
    """lelph_qexml_info (include/lelph_hip.h)"""
    _fields_ = [("natom", C.c_int64), ("ntype", C.c_int64), ("nsym", C.c_int64),
                ("fft_dims", C.c_int64 * 3), ("lat_vec", C.c_double * 9),
                ("alat", C.c_double * 3), ("dimension", C.c_char),
                ("is_soc_present", C.c_int), ("nmag", C.c_int), ("nspinor", C.c_int),
                ("ph_mag_symm", C.c_int), ("ph_tim_rev", C.c_int),
                ("pseudo_dir", C.c_char * 1024)]


def _path(p):
    return None if p is None else os.fsencode(p)


def parse_qexml(path) -> dict:
    """parse_qexml (src/io/qe/read_qexml.c:20-643), host code: the fields of
    data-file-schema.xml the reference reads, as a dict."""
    info = QexmlInfo()
    _check(lib().lelph_parse_qexml(_path(path), C.byref(info), None, None, None, None, None,
                                   None), "lelph_parse_qexml")
    na, nt, ns = info.natom, info.ntype, info.nsym
    atype = np.zeros(na, np.int32)
    pos = np.zeros((na, 3))
    files = C.create_string_buffer(256 * nt)
    R = np.zeros((ns, 3, 3))
    tau = np.zeros((ns, 3))
    trev = np.zeros(ns, np.int32)
    _check(lib().lelph_parse_qexml(_path(path), C.byref(info), atype.ctypes.data_as(c_int_p),
                                   dptr(pos), files, dptr(R), dptr(tau),
                                   trev.ctypes.data_as(c_int_p)), "lelph_parse_qexml")
    raw = files.raw
    names = [raw[256 * t:256 * (t + 1)].split(b"\0", 1)[0].decode().strip() for t in range(nt)]
    return {"natom": na, "ntype": nt, "nsym": ns, "fft_dims": tuple(info.fft_dims),
            "lat_vec": np.array(info.lat_vec).reshape(3, 3), "alat": np.array(info.alat),
            "dimension": info.dimension.decode(), "is_soc_present": bool(info.is_soc_present),
            "nmag": info.nmag, "nspinor": info.nspinor, "ph_mag_symm": bool(info.ph_mag_symm),
            "ph_tim_rev": bool(info.ph_tim_rev), "pseudo_dir": info.pseudo_dir.decode(),
            "atom_type": atype, "atomic_pos": pos, "pseudo_files": names,
            "sym_mats": R, "sym_tau": tau, "trevs": trev.astype(bool)}


def parse_upf(path):
    """parse_upf (src/io/qe/parse_upf.c:28-336), host code: (Vloc, r, rab,
    Zval) of a norm-conserving UPF-1 or UPF-2 file."""
    n = C.c_int64()
    z = C.c_double()
    _check(lib().lelph_parse_upf(_path(path), C.byref(n), C.byref(z), None, None, None),
           "lelph_parse_upf")
    r, rab, v = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
    _check(lib().lelph_parse_upf(_path(path), C.byref(n), C.byref(z), dptr(r), dptr(rab),
                                 dptr(v)), "lelph_parse_upf")
    return v, r, rab, z.value


def read_ph_tensors_qe(tensors_xml, quadrupole_fmt, natom):
    """read_ph_tensors_qe + read_quadrupole_fmt (src/io/qe/read_tensors.c),
    host code: (epsilon, Zborn, Qpole), None where absent."""
    eps, zb, qp = np.zeros((3, 3)), np.zeros((natom, 3, 3)), np.zeros((natom, 3, 3, 3))
    pres = C.c_int()
    _check(lib().lelph_read_ph_tensors_qe(_path(tensors_xml), _path(quadrupole_fmt), int(natom),
                                          dptr(eps), dptr(zb), dptr(qp), C.byref(pres)),
           "lelph_read_ph_tensors_qe")
    p = pres.value
    return (eps if p & 1 else None, zb if p & 2 else None, qp if p & 4 else None)


def read_qpts_qe(dyn0):
    """read_qpts_qe (src/io/qe/read_dyn.c:551-609), host code: (iBZ q points
    (nq, 3) cartesian in 2 pi / alat, number of q in the full grid)."""
    ni, nb = C.c_int64(), C.c_int64()
    _check(lib().lelph_read_qpts_qe(_path(dyn0), C.byref(ni), C.byref(nb), None),
           "lelph_read_qpts_qe")
    q = np.zeros((ni.value, 3))
    _check(lib().lelph_read_qpts_qe(_path(dyn0), C.byref(ni), C.byref(nb), dptr(q)),
           "lelph_read_qpts_qe")
    return q, nb.value


ASR_KINDS = {"no": 0, "simple": 1, "crystal": 2, "all": 3, "all_huang": 4}


def asr_kind(kind) -> int:
    """asr_kind_from_string (src/phonon/asr.c:73-116): a name (any case) or
    an int; unknown names mean no sum rule."""
    if isinstance(kind, str):
        return ASR_KINDS.get(kind.lower(), 0)
    return int(kind)


def asr_fc(kind, qgrid, frc, atomic_pos, lat_vecs=None, ws_vecs=None, ws_degen=None):
    """apply_acoustic_sum_rule_fc (src/phonon/asr.c:269-512), host code of
    the library: a copy of frc (ngrid, 3 natom, 3 natom) with the rule
    applied; ws_vecs / ws_degen from wigner_seitz(qgrid, lat, eps, pos, pos)."""
    out = np.ascontiguousarray(frc, np.complex128).copy()
    g = np.ascontiguousarray(qgrid, np.int64)
    pos = np.ascontiguousarray(atomic_pos, np.float64)
    lat = None if lat_vecs is None else np.ascontiguousarray(lat_vecs, np.float64)
    wv = None if ws_vecs is None else np.ascontiguousarray(ws_vecs, np.int64)
    wd = None if ws_degen is None else np.ascontiguousarray(ws_degen, np.int64)
    _check(lib().lelph_asr_fc(asr_kind(kind), i64ptr(g), len(pos), dptr(out), dptr(pos),
                              None if lat is None else dptr(lat),
                              None if wv is None else i64ptr(wv), 0 if wv is None else len(wv),
                              None if wd is None else i64ptr(wd)), "lelph_asr_fc")
    return out


def asr_born(kind, Zborn):
    """apply_acoustic_sum_rule_born_charges (src/phonon/asr.c:155-267)"""
    z = np.ascontiguousarray(Zborn, np.float64).copy()
    _check(lib().lelph_asr_born(asr_kind(kind), dptr(z), len(z)), "lelph_asr_born")
    return z


def quadrupole_3d_to_2d(epsilon, atomic_pos, Zborn, Qpole):
    """quadrupole_3d_to_2d (src/phonon/quadrupole_3Dto2D.c:18-117): the
    2-D quadrupoles (natom, 3, 3, 3) of a slab from its 3-D ones"""
    q = np.ascontiguousarray(Qpole, np.float64).copy()
    z = None if Zborn is None else np.ascontiguousarray(Zborn, np.float64)
    _check(lib().lelph_quadrupole_3d_to_2d(
        len(q), dptr(np.ascontiguousarray(epsilon, np.float64)),
        dptr(np.ascontiguousarray(atomic_pos, np.float64)), None if z is None else dptr(z),
        dptr(q)), "lelph_quadrupole_3d_to_2d")
    return q


def loto_qpoint(lat_vecs, direction):
    """interpolation_driver.c:734-786: the crystal q standing for Gamma
    approached along the cartesian direction"""
    out = np.zeros(3)
    _check(lib().lelph_loto_qpoint(dptr(np.ascontiguousarray(lat_vecs, np.float64)),
                                   dptr(np.ascontiguousarray(direction, np.float64)),
                                   dptr(out)), "lelph_loto_qpoint")
    return out
