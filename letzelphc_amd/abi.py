"""ctypes mirror of include/lelph_hip.h (the C-ABI boundary types).

The structs follow the reference's struct Lattice / Pseudo / WFC
(src/common/dtypes.h:60-266).  `pack_system` builds them from a
`letzelphc_amd.synth.System` and keeps every numpy buffer alive for as long
as the returned `Packed` object lives.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)

LELPH_OK = 0
STATUS_NAMES = {
    0: "LELPH_OK", 1: "LELPH_ERR_ARG", 2: "LELPH_ERR_ALLOC", 3: "LELPH_ERR_HIP",
    4: "LELPH_ERR_RANGE", 5: "LELPH_ERR_UNSUPPORTED", 6: "LELPH_ERR_STATE",
    7: "LELPH_ERR_NOKQ", 8: "LELPH_ERR_SYM", 9: "LELPH_ERR_SCRATCH",
}

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3}
KERNELS = {"dfpt": 0, "dfpt_local": 1, "bare": 2, "bare_local": 3}


class Symmetry(C.Structure):
    _fields_ = [("Rmat", C.c_double * 9), ("tau", C.c_double * 3),
                ("time_rev", C.c_int)]


class Lattice(C.Structure):
    _fields_ = [
        ("natom", C.c_int), ("nsym", C.c_int), ("timerev", C.c_int),
        ("nspin", C.c_int), ("nspinor", C.c_int), ("nbnds", C.c_int),
        ("nmodes", C.c_int), ("nmag", C.c_int), ("dimension", C.c_char),
        ("is_soc_present", C.c_int), ("fft_dims", C.c_int64 * 3),
        ("alat_vec", C.c_double * 9), ("blat_vec", C.c_double * 9),
        ("volume", C.c_double), ("atomic_pos", c_double_p),
        ("atom_type", c_int_p), ("nkpts_iBZ", C.c_int64),
        ("kpt_iredBZ", c_double_p), ("nkpts_BZ", C.c_int64),
        ("kpt_fullBZ_crys", c_double_p), ("kmap", c_int_p),
        ("syms", C.POINTER(Symmetry)),
    ]


class LocalPseudo(C.Structure):
    _fields_ = [("Vloc_atomic", c_double_p), ("r_grid", c_double_p),
                ("rab_grid", c_double_p), ("ngrid", C.c_int64),
                ("Zval", C.c_double)]


class Pseudo(C.Structure):
    _fields_ = [("ntype", C.c_int), ("lmax", C.c_int), ("nltimesj", C.c_int64),
                ("PP_table", c_double_p), ("Fsign", c_double_p),
                ("fCoeff", C.POINTER(c_double_p)),
                ("loc_pseudo", C.POINTER(LocalPseudo)), ("qmax_abs", C.c_double)]


class Tuning(C.Structure):
    """lelph_tuning: launch shapes of a context (fft_lines, fp32_form and
    ctile give bit-identical g, split_target changes it at the rounding
    level) and the form of the bf16 contexts' non-local GEMMs (nl_form)."""
    _fields_ = [("split_target", C.c_int), ("fft_lines", C.c_int), ("fp32_form", C.c_int),
                ("ctile", C.c_int), ("nl_form", C.c_int)]


class Pair(C.Structure):
    _fields_ = [("ikq", C.c_int), ("ik", C.c_int), ("kqsym", C.c_int),
                ("ksym", C.c_int)]


class LongRange(C.Structure):
    """lelph_longrange: parameters of dV_add_longrange
    (src/dvloc/dVpot_add_long_range.c:18-241)."""
    _fields_ = [("epsilon", c_double_p), ("Zborn", c_double_p), ("Qpole", c_double_p),
                ("Zvals", c_double_p), ("only_induced", C.c_int), ("eta_bare", C.c_double),
                ("eta_induced", C.c_double), ("EcutRy", C.c_double), ("nmags_add", C.c_int * 4)]


@dataclass
class LrParams:
    """Host-side long-range parameters; `pack()` returns a LongRange whose
    pointers stay valid while this object lives."""
    epsilon: np.ndarray | None = None
    Zborn: np.ndarray | None = None
    Qpole: np.ndarray | None = None
    Zvals: np.ndarray | None = None
    only_induced: bool = False
    eta_bare: float = 1.0
    eta_induced: float = 1.0
    EcutRy: float = 30.0
    nmags_add: tuple = (1, 0, 0, 0)
    _keep: list = field(default_factory=list)

    def pack(self) -> LongRange:
        def p(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            self._keep.append(a)
            return dptr(a)
        lr = LongRange()
        lr.epsilon = p(self.epsilon)
        lr.Zborn = p(self.Zborn)
        lr.Qpole = p(self.Qpole)
        lr.Zvals = p(self.Zvals)
        lr.only_induced = int(bool(self.only_induced))
        lr.eta_bare = float(self.eta_bare)
        lr.eta_induced = float(self.eta_induced)
        lr.EcutRy = float(self.EcutRy)
        for i in range(4):
            lr.nmags_add[i] = int(self.nmags_add[i]) if i < len(self.nmags_add) else 0
        return lr


def i64ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.int64
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def dptr(a: np.ndarray):
    """double* to a C-contiguous float64 / complex128 array."""
    assert a.flags["C_CONTIGUOUS"], "array must be C contiguous"
    assert a.dtype in (np.float64, np.complex128), a.dtype
    return a.ctypes.data_as(c_double_p)


def iptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.int32
    return a.ctypes.data_as(c_int_p)


@dataclass
class Packed:
    lattice: Lattice
    pseudo: Pseudo
    keep: list = field(default_factory=list)


def pack_system(sys) -> Packed:
    """Builds Lattice/Pseudo structs for a synth.System."""
    keep = []

    def hold(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    lat = Lattice()
    lat.natom = sys.natom
    lat.nsym = len(sys.syms)
    lat.timerev = sys.timerev
    lat.nspin = sys.nspin
    lat.nspinor = sys.nspinor
    lat.nbnds = sys.nbnds
    lat.nmodes = sys.nmodes
    lat.nmag = sys.nmag
    lat.dimension = sys.dimension.encode()
    lat.is_soc_present = int(bool(getattr(sys.pseudo, "soc", False)))
    for i in range(3):
        lat.fft_dims[i] = int(sys.fft_dims[i])
    for i in range(9):
        lat.alat_vec[i] = float(sys.alat.reshape(-1)[i])
        lat.blat_vec[i] = float(sys.blat.reshape(-1)[i])
    lat.volume = float(sys.volume)
    lat.atomic_pos = dptr(hold(sys.atomic_pos, np.float64))
    lat.atom_type = iptr(hold(sys.atom_type, np.int32))
    lat.nkpts_iBZ = len(sys.kpt_ibz_cart)
    lat.kpt_iredBZ = dptr(hold(sys.kpt_ibz_cart, np.float64))
    lat.nkpts_BZ = len(sys.kpt_bz_crys)
    lat.kpt_fullBZ_crys = dptr(hold(sys.kpt_bz_crys, np.float64))
    lat.kmap = iptr(hold(sys.kmap, np.int32))
    syms = (Symmetry * len(sys.syms))()
    for i, (R, tau, tr) in enumerate(sys.syms):
        for j in range(9):
            syms[i].Rmat[j] = float(np.asarray(R).reshape(-1)[j])
        for j in range(3):
            syms[i].tau[j] = float(tau[j])
        syms[i].time_rev = int(tr)
    keep.append(syms)
    lat.syms = C.cast(syms, C.POINTER(Symmetry))

    pp = Pseudo()
    ps = sys.pseudo
    pp.ntype = ps.ntype
    pp.lmax = ps.lmax
    pp.nltimesj = ps.nltimesj
    pp.PP_table = dptr(hold(ps.PP_table, np.float64))
    pp.Fsign = dptr(hold(ps.Fsign, np.float64))
    fco = getattr(ps, "fcoeff", None)
    if fco is not None:
        arr = (c_double_p * len(fco))()
        for i, f in enumerate(fco):
            arr[i] = None if f is None else dptr(hold(f, np.complex128))
        keep.append(arr)
        pp.fCoeff = C.cast(arr, C.POINTER(c_double_p))
    else:
        pp.fCoeff = None
    locs = (LocalPseudo * ps.ntype)()
    for t, (v, r, rab, z) in enumerate(ps.loc):
        locs[t].Vloc_atomic = dptr(hold(v, np.float64))
        locs[t].r_grid = dptr(hold(r, np.float64))
        locs[t].rab_grid = dptr(hold(rab, np.float64))
        locs[t].ngrid = len(r)
        locs[t].Zval = float(z)
    keep.append(locs)
    pp.loc_pseudo = C.cast(locs, C.POINTER(LocalPseudo))
    pp.qmax_abs = float(ps.qmax_abs)
    return Packed(lat, pp, keep)
