"""Seeded synthetic systems in the layouts the reference reads.

Every array is generated from a Philox counter-based stream keyed by
(system seed, role, index), so the CPU oracle and the GPU path consume
bit-identical complex128 inputs.  Layout conventions are those of the
reference (src/common/dtypes.h:80-266):

  alat[:, i]       lattice vector a_i (row-major 3x3, bohr)
  blat[:, i]       reciprocal vector b_i, includes 2*pi computed with the
                   reference's truncated pi (src/common/numerical_func.c:280-302)
  kpt_ibz_cart     iBZ k points, cartesian, no 2*pi
  kpt_bz_crys      full-BZ k points, crystal
  kmap[ik]         (ik_ibz, isym) with k_bz = Rmat[isym] @ k_ibz (mod G)
  wfc              (nspin, nbnds, nspinor, npw) complex128
  gvec             (npw, 3) cartesian, no 2*pi
  Fk               (nltimesj, ntype, npw) Kleinman-Bylander form factors
  dvscf            (nmodes, nmag, Nx, Ny, Nz) complex128, mode basis

The five presets follow the workload descriptions: "c0" 2-atom fcc 24^3,
"c1" 3-atom hexagonal monolayer 45x45x216 with 12 symmetry operations,
"c2" 16-atom cubic supercell 96^3, "c3" the uneven batch cells and "c4"
(the c0 cell on a coarse q grid, used by the interpolation path).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

RPI = 3.1415927  # ELPH_PI, src/common/constants.h:11
E2 = 2.0         # ELPH_e2 (Ry), src/common/constants.h:13


def reciprocal_vecs(alat: np.ndarray) -> np.ndarray:
    """blat with b_i = blat[:, i]; 2*pi with the reference's truncated pi."""
    a = alat.reshape(-1)
    det = np.linalg.det(alat)
    d = 2 * RPI / det
    b = np.array([
        (a[4] * a[8] - a[7] * a[5]), (a[5] * a[6] - a[3] * a[8]),
        (a[3] * a[7] - a[6] * a[4]), (a[2] * a[7] - a[1] * a[8]),
        (a[0] * a[8] - a[2] * a[6]), (a[1] * a[6] - a[0] * a[7]),
        (a[1] * a[5] - a[2] * a[4]), (a[2] * a[3] - a[0] * a[5]),
        (a[0] * a[4] - a[1] * a[3])]) * d
    return b.reshape(3, 3)


def rng_for(seed: int, *key: int) -> np.random.Generator:
    """Philox stream for (seed, key...)."""
    k = [seed & 0xFFFFFFFF] + [int(x) & 0xFFFFFFFF for x in key]
    while len(k) < 4:
        k.append(0)
    key128 = 0
    for i, x in enumerate(k[:4]):
        key128 |= x << (32 * i)
    return np.random.Generator(np.random.Philox(key=key128))


def cgauss(rng, shape):
    """complex normal, unit variance: (re, im) pairs from one draw."""
    n = int(np.prod(shape))
    z = rng.standard_normal(2 * n).view(np.complex128)
    z *= 1.0 / math.sqrt(2)
    return z.reshape(shape)


@dataclass
class PseudoData:
    ntype: int
    lmax: int
    nltimesj: int
    PP_table: np.ndarray     # (nltimesj, ntype, 3)
    Fsign: np.ndarray        # (nltimesj, ntype)
    Dl: np.ndarray           # (nltimesj, ntype) KB energies, Ry
    kb_alpha: float          # beta_l(r) = r^l exp(-r^2 / kb_alpha)
    loc: list                # per type (Vloc(r), r, rab, Zval)
    loc_A: float             # V_loc(r) = -e2 Z erf(r)/r + loc_A exp(-r^2/loc_b)
    loc_b: float
    qmax_abs: float
    soc: bool = False        # fully relativistic projectors (j = l +- 1/2)
    fcoeff: list = None      # [lidx * ntype + itype] -> (2l+1, 2l+1, 2, 2) or None


@dataclass
class Wfc:
    wfc: np.ndarray
    gvec: np.ndarray
    Fk: np.ndarray
    gint: np.ndarray  # (npw,3) integer crystal coordinates

    @property
    def npw(self):
        return self.gvec.shape[0]


@dataclass
class System:
    name: str
    seed: int
    alat: np.ndarray
    blat: np.ndarray
    volume: float
    fft_dims: tuple
    natom: int
    atomic_pos: np.ndarray
    atom_type: np.ndarray
    nspin: int
    nspinor: int
    nbnds: int
    nmodes: int
    nmag: int
    dimension: str
    syms: list
    timerev: int
    kpt_ibz_crys: np.ndarray
    kpt_ibz_cart: np.ndarray
    kpt_bz_crys: np.ndarray
    kmap: np.ndarray
    ecut: float
    pseudo: PseudoData
    qpts: list
    wfcs: list = field(default_factory=list)
    _eig: dict = field(default_factory=dict)
    _dv: dict = field(default_factory=dict)

    @property
    def nfft(self):
        return int(np.prod(self.fft_dims))

    def eigvec(self, iq: int) -> np.ndarray:
        """(nmodes, 3*natom) random unitary polarisation matrix of q iq."""
        if iq not in self._eig:
            rng = rng_for(self.seed, 3, iq)
            m = cgauss(rng, (self.nmodes, self.nmodes))
            q, r = np.linalg.qr(m)
            d = np.diag(r)
            self._eig[iq] = np.ascontiguousarray(q * (d / np.abs(d))[None, :])
        return self._eig[iq]

    def dvscf(self, iq: int) -> np.ndarray:
        """(nmodes, nmag, Nx, Ny, Nz) mode-basis screened dV of q iq:
        per set, Fourier components N(0,1) exp(-|q+G|^2/4) on the full box,
        then the unnormalised backward transform."""
        if iq not in self._dv:
            Nx, Ny, Nz = self.fft_dims
            q = np.asarray(self.qpts[iq], dtype=np.float64)
            mi = [miller_axis(n) for n in (Nx, Ny, Nz)]
            gx, gy, gz = np.meshgrid(mi[0] + q[0], mi[1] + q[1], mi[2] + q[2],
                                     indexing="ij")
            qg = np.stack([gx, gy, gz], -1) @ self.blat.T
            env = np.exp(-np.sum(qg * qg, -1) / 4.0)
            out = np.empty((self.nmodes, self.nmag, Nx, Ny, Nz), np.complex128)
            for v in range(self.nmodes):
                for m in range(self.nmag):
                    rng = rng_for(self.seed, 4, iq, v * 8 + m)
                    cg = cgauss(rng, (Nx, Ny, Nz)) * env
                    out[v, m] = np.fft.ifftn(cg) * (Nx * Ny * Nz)
            self._dv[iq] = out
        return self._dv[iq]


def miller_axis(n: int) -> np.ndarray:
    """get_miller_idx (src/common/numerical_func.c:436-449) for 0..n-1."""
    i = np.arange(n)
    mid = (n - 1) // 2 + 1
    return np.where(i < mid, i, (i - mid) - n // 2).astype(np.float64)


def _sym_crys(R: np.ndarray, alat: np.ndarray, blat: np.ndarray) -> np.ndarray:
    """k_crys' = A^T R (B/2pi) k_crys."""
    return alat.T @ R @ (blat / (2 * RPI))


def _mp_grid(n):
    pts = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                v = np.array([i / n[0], j / n[1], k / n[2]])
                pts.append(((v + 0.5) % 1.0) - 0.5)  # to [-0.5, 0.5)
    return np.array(pts)


def _reduce_kgrid(kbz, syms, alat, blat):
    if len(syms) == 1:  # the identity alone: every k is its own star
        kmap = np.zeros((len(kbz), 2), np.int32)
        kmap[:, 0] = np.arange(len(kbz))
        return np.array(kbz), kmap
    ibz, kmap = [], np.zeros((len(kbz), 2), np.int32)
    Sc = [_sym_crys(np.asarray(R), alat, blat) for (R, _, _) in syms]
    for ik, k in enumerate(kbz):
        found = False
        for ii, ki in enumerate(ibz):
            for isym, S in enumerate(Sc):
                d = S @ ki - k
                if np.linalg.norm(d - np.rint(d)) < 1e-5:
                    kmap[ik] = (ii, isym)
                    found = True
                    break
            if found:
                break
        if not found:
            kmap[ik] = (len(ibz), 0)
            ibz.append(k.copy())
    return np.array(ibz), kmap


def _radial_grid(Z: float, n: int = 1001, xmin: float = -7.0, dx: float = 0.0125):
    x = xmin + dx * np.arange(n)
    r = np.exp(x) / Z
    return r, r * dx


ELPH_SQRT2 = 1.4142136  # src/common/constants.h:12, enters the f coefficients


def _cmat(m1, m2):
    """real -> complex spherical-harmonic transformation element
    (Cmat, src/nonloc/fcoeff.c:203-248)"""
    a1, a2 = abs(m1), abs(m2)
    if a1 != a2:
        return 0.0
    if m1 == 0 and m2 == 0:
        return 1.0
    if m1 == a1 and m2 == a2:
        return (-1.0) ** a1 / ELPH_SQRT2
    if m1 == -a1 and m2 == a2:
        return 1.0 / ELPH_SQRT2
    if m1 == a1 and m2 == -a2:
        return 1j * (-1.0) ** a1 / ELPH_SQRT2
    return -1j / ELPH_SQRT2


def _cg(jp1, down, l, twomj):
    """Clebsch-Gordan alpha (CGCoeff, src/nonloc/fcoeff.c:250-283)"""
    mj = twomj / 2.0
    if jp1:
        m = mj - 0.5
        return math.sqrt((l - m) / (2 * l + 1.0)) if down else math.sqrt((l + m + 1) / (2 * l + 1.0))
    m = mj + 0.5
    return -math.sqrt((l + m) / (2 * l + 1.0)) if down else math.sqrt((l - m + 1) / (2 * l + 1.0))


def _umat(jp1, down, l, twomj, m2):
    """U (Umat, src/nonloc/fcoeff.c:166-201)"""
    if jp1:
        m = (twomj - 1) // 2
        return _cmat(m + 1 if down else m, m2)
    m = (twomj + 1) // 2
    return _cmat(m if down else m - 1, m2)


def soc_fcoeff(l: int, twoj: int) -> np.ndarray:
    """f^{l,j}_{m m' s s'} = sum_mj alpha U conj(alpha U), (2l+1, 2l+1, 2, 2)
    (alloc_and_Compute_f_Coeff, src/nonloc/fcoeff.c:24-143)"""
    jp1 = twoj == 2 * l + 1
    if not jp1 and twoj != 2 * l - 1:
        raise ValueError("inconsistent j")
    f = np.zeros((2 * l + 1, 2 * l + 1, 2, 2), np.complex128)
    for i1 in range(2 * l + 1):
        for i2 in range(2 * l + 1):
            for s1 in range(2):
                for s2 in range(2):
                    acc = 0.0
                    for mj in range(-twoj, twoj + 1, 2):
                        acc += (_cg(jp1, s1, l, mj) * _cg(jp1, s2, l, mj) *
                                _umat(jp1, s1, l, mj, i1 - l) *
                                np.conj(_umat(jp1, s2, l, mj, i2 - l)))
                    f[i1, i2, s1, s2] = acc
    return f


def _make_pseudo(ntype, Zvals, qmax_abs, lmax=1, Dl=(1.0, -0.5), kb_alpha=1.5,
                 loc_A=-2.0, loc_b=1.2, soc=False):
    from scipy.special import erf
    # scalar: one channel per l; fully relativistic: s 1/2, then p 1/2 and
    # p 3/2 with the same projector and D (so that sum_j f^{l,j} = 1 leaves
    # the non-local part of the scalar pseudopotential)
    chans = [(l, 0) for l in range(lmax + 1)] if not soc else \
        [(0, 1)] + [(l, tj) for l in range(1, lmax + 1) for tj in (2 * l - 1, 2 * l + 1)]
    nlj = len(chans)
    PP = np.zeros((nlj, ntype, 3))
    Fs = np.zeros((nlj, ntype))
    D = np.zeros((nlj, ntype))
    for li, (l, tj) in enumerate(chans):
        for t in range(ntype):
            PP[li, t] = (l + 1, tj, 1)
            D[li, t] = Dl[l]
            Fs[li, t] = np.sign(Dl[l])
    loc = []
    for t in range(ntype):
        Z = Zvals[t]
        r, rab = _radial_grid(Z if Z > 1 else 1.0)
        v = -E2 * Z * erf(r) / r + loc_A * np.exp(-r * r / loc_b)
        loc.append((v, r, rab, float(Z)))
    fco = None
    if soc:
        fco = [soc_fcoeff(l, tj) for (l, tj) in chans for _ in range(ntype)]
    return PseudoData(ntype, lmax, nlj, PP, Fs, D, kb_alpha, loc, loc_A, loc_b,
                      qmax_abs, soc, fco)


def kb_radial_ft(l: int, q: np.ndarray, a: float) -> np.ndarray:
    """int r^2 j_l(q r) r^l exp(-r^2/a) dr = sqrt(pi) q^l a^(l+3/2) / 2^(l+2)
    exp(-q^2 a / 4)  (closed form for the Gaussian KB projector)."""
    return math.sqrt(math.pi) * q ** l * a ** (l + 1.5) / 2 ** (l + 2) * np.exp(-q * q * a / 4)


def _gsphere(kcrys, blat, ecut):
    """Integer G (crystal) with |2pi (k+G)|^2 <= ecut, |.| with blat (2pi)."""
    bn = np.linalg.norm(blat, axis=0)
    nmax = [int(math.ceil(math.sqrt(ecut) / bn[i])) + 2 for i in range(3)]
    rng = [np.arange(-m, m + 1) for m in nmax]
    g = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    kg = (g + kcrys[None, :]) @ blat.T
    e = np.sum(kg * kg, -1)
    sel = e <= ecut
    g = g[sel]
    e = e[sel]
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.round(e, 10)))
    return g[order].astype(np.int64)


def _gram_schmidt_rows(M):
    """Orthonormalise the rows of M (nb, n) in order: the Gram-Schmidt result
    M = L U with L lower triangular, positive diagonal, computed as Cholesky
    QR applied twice (S = M M^H = L L^H, U = L^-1 M)."""
    U = M
    for _ in range(2):
        S = U @ U.conj().T
        L = np.linalg.cholesky(S)
        U = np.linalg.inv(L) @ U
    return np.ascontiguousarray(U)


def _make_wfc(sys: System, ik: int) -> Wfc:
    ps = sys.pseudo
    bscale = sys.blat / (2 * RPI)
    kc = sys.kpt_ibz_crys[ik]
    g = _gsphere(kc, sys.blat, sys.ecut)
    npw = len(g)
    gcart = np.ascontiguousarray(g.astype(np.float64) @ bscale.T)
    kg = (g + kc[None, :]) @ sys.blat.T
    kg2 = np.sum(kg * kg, -1)
    rng = rng_for(sys.seed, 1, ik)
    c = cgauss(rng, (sys.nspin, sys.nbnds, sys.nspinor, npw))
    c *= np.exp(-kg2 / 8.0)[None, None, None, :]
    for s in range(sys.nspin):
        M = c[s].reshape(sys.nbnds, sys.nspinor * npw)
        c[s] = _gram_schmidt_rows(M).reshape(sys.nbnds, sys.nspinor, npw)
    qn = np.sqrt(kg2)
    Fk = np.zeros((ps.nltimesj, ps.ntype, npw))
    pref = 4 * RPI / math.sqrt(sys.volume)
    for li in range(ps.nltimesj):
        l = int(round(ps.PP_table[li, 0, 0] - 1))
        for t in range(ps.ntype):
            Fk[li, t] = pref * math.sqrt(abs(ps.Dl[li, t])) * kb_radial_ft(l, qn, ps.kb_alpha)
    return Wfc(np.ascontiguousarray(c), gcart, np.ascontiguousarray(Fk), g)


def make_wfcs(sys: System, workers: int = 0, only=None):
    """The iBZ wavefunctions; independent per k, generated on a thread pool
    (every k has its own Philox stream, so the result does not depend on the
    number of workers).  `only`: generate just these iBZ k, the other
    entries stay None (a pair list that touches a few k of a large cell)."""
    from concurrent.futures import ThreadPoolExecutor
    nk = len(sys.kpt_ibz_crys)
    ks = list(range(nk)) if only is None else sorted(set(int(k) for k in only))
    workers = workers or 1
    if workers <= 1 or len(ks) <= 1:
        made = [_make_wfc(sys, ik) for ik in ks]
    else:
        with ThreadPoolExecutor(workers) as ex:
            made = list(ex.map(lambda ik: _make_wfc(sys, ik), ks))
    wf = [None] * nk
    for ik, w in zip(ks, made):
        wf[ik] = w
    sys.wfcs = wf
    return wf


def _rot_z(th):
    c, s = math.cos(th), math.sin(th)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def make_system(name, seed, alat, crys_pos, atom_type, Zvals, fft_dims, kgrid,
                ecut, nbnds, syms, timerev, qpts, nspin=1, nspinor=1, nmag=1,
                dimension="3", with_wfc=True, soc=False):
    alat = np.asarray(alat, np.float64)
    blat = reciprocal_vecs(alat)
    vol = float(np.linalg.det(alat))
    pos = np.asarray(crys_pos, np.float64) @ alat.T
    natom = len(pos)
    kbz = _mp_grid(kgrid)
    kibz, kmap = _reduce_kgrid(kbz, syms, alat, blat)
    kcart = kibz @ (blat / (2 * RPI)).T
    qmax = max([float(np.max(np.abs(q))) for q in qpts] + [0.0])
    if soc and nspinor != 2:
        raise ValueError("fully relativistic projectors need nspinor = 2")
    ps = _make_pseudo(int(max(atom_type)) + 1, Zvals, qmax, soc=soc)
    s = System(name, seed, alat, blat, vol, tuple(int(x) for x in fft_dims),
               natom, pos, np.asarray(atom_type, np.int32), nspin, nspinor,
               nbnds, 3 * natom, nmag, dimension, syms, timerev, kibz,
               np.ascontiguousarray(kcart), kbz, kmap, ecut, ps,
               [np.asarray(q, np.float64) for q in qpts])
    if with_wfc:
        make_wfcs(s)
    return s


def fcc_cell(a):
    return np.array([[-a / 2, 0, -a / 2], [0, a / 2, a / 2], [a / 2, a / 2, 0]])


def hex_cell(a, c):
    return np.array([[a, -a / 2, 0], [0, a * math.sqrt(3) / 2, 0], [0, 0, c]])


@lru_cache(maxsize=None)
def _syms_c1():
    rot = [(_rot_z(j * math.pi / 3), np.zeros(3), 0) for j in range(6)]
    tr = [(-_rot_z(j * math.pi / 3), np.zeros(3), 1) for j in range(6)]
    return tuple(rot + tr)


def config(cfg: str, nbnds=None, kgrid=None, with_wfc=True, **kw) -> System:
    """Preset systems.  `nbnds`/`kgrid` shrink a preset for fast tests."""
    I3 = np.eye(3)
    if cfg == "c0":
        a = 10.26
        syms = [(I3, np.zeros(3), 0), (-I3, np.zeros(3), 1)]
        return make_system("c0", 1000, fcc_cell(a), [[0, 0, 0], [0.25, 0.25, 0.25]],
                           [0, 0], [4.0], (24, 24, 24), kgrid or (4, 4, 4), 20.0,
                           nbnds or 8, syms, 1, [(0.0, 0.0, 0.25)],
                           with_wfc=with_wfc, **kw)
    if cfg == "c1":
        a, c = 5.9, 38.0
        z = 3.0 / c
        return make_system("c1", 1001, hex_cell(a, c),
                           [[1 / 3, 2 / 3, 0.5], [2 / 3, 1 / 3, 0.5 + z], [2 / 3, 1 / 3, 0.5 - z]],
                           [0, 1, 1], [14.0, 6.0], (45, 45, 216), kgrid or (12, 12, 1),
                           60.0, nbnds or 40, list(_syms_c1()), 1,
                           [(1 / 6, 0.0, 0.0), (1 / 3, 1 / 3, 0.0)], dimension="2",
                           with_wfc=with_wfc, **kw)
    if cfg == "c2":
        a = 20.5
        pos, typ = [], []
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    pos.append([i / 2, j / 2, k / 2])
                    typ.append(0)
                    pos.append([i / 2 + 0.25, j / 2 + 0.25, k / 2 + 0.25])
                    typ.append(1)
        return make_system("c2", 1002, a * I3, pos, typ, [4.0, 6.0], (96, 96, 96),
                           kgrid or (4, 4, 4), 40.0, nbnds or 128,
                           [(I3, np.zeros(3), 0)], 0, [(0.25, 0.25, 0.0)],
                           with_wfc=with_wfc, **kw)
    raise ValueError(cfg)


# c3: eight uneven cells (FFT grid, lattice constant or hex (a, c), atoms,
# bands, target plane waves); 16 (k, k+q) pairs each
C3_CELLS = [
    ((24, 24, 24), "fcc", 10.26, 2, 8, None),
    ((32, 32, 32), "sc", 9.0, 2, 16, 500),
    ((36, 36, 36), "sc", 10.5, 4, 16, 900),
    ((45, 45, 108), "hex", (5.9, 14.2), 3, 24, 1500),
    ((48, 48, 48), "sc", 13.0, 6, 32, 3000),
    ((60, 60, 60), "sc", 15.5, 8, 32, 5000),
    ((64, 64, 64), "sc", 16.5, 10, 48, 8000),
    ((72, 72, 72), "sc", 19.0, 12, 64, 12000),
]


def config_c3(with_wfc=True, cells=None):
    """The batching workload: eight cells with FFT grids 24^3 ... 72^3 and
    45x45x108, 180 ... 12,000 plane waves, 8 ... 64 bands, 6 ... 36 modes,
    16 (k, k+q) pairs each (identity symmetry, so all 16 k are iBZ).  The
    cut-off of each cell gives its target plane-wave count, capped so that
    4 G_max fits the grid (no aliasing of the products)."""
    out = []
    I3 = np.eye(3)
    for ic, (dims, kind, a, nat, nb, npw) in enumerate(C3_CELLS):
        if cells is not None and ic not in cells:
            continue
        rng = rng_for(3000 + ic, 9)
        if kind == "fcc":
            out.append(config("c0", nbnds=nb, kgrid=(2, 2, 4), with_wfc=with_wfc))
            continue
        if kind == "hex":
            alat = hex_cell(*a)
            kgrid, q = (4, 4, 1), (0.25, 0.0, 0.0)
            dim = "2"
        else:
            alat = a * I3
            kgrid, q = (2, 2, 4), (0.0, 0.0, 0.25)
            dim = "3"
        pos = rng.uniform(0.05, 0.95, (nat, 3))
        if kind == "hex":
            pos[:, 2] = 0.5 + 0.1 * (pos[:, 2] - 0.5)
        typ = [i % 2 for i in range(nat)]
        blat = reciprocal_vecs(np.asarray(alat, np.float64))
        bn = np.linalg.norm(blat, axis=0)
        cap = min(((dims[i] - 1) / 4.0 * bn[i]) ** 2 for i in range(3))
        vol = abs(np.linalg.det(alat))
        ecut = min(0.95 * cap, (npw * 6 * math.pi ** 2 / vol) ** (2.0 / 3.0))
        out.append(make_system(f"c3_{ic}", 3000 + ic, alat, pos, typ, [4.0, 6.0], dims, kgrid,
                               ecut, nb, [(I3, np.zeros(3), 0)], 0, [q], dimension=dim,
                               with_wfc=with_wfc))
    return out


def tiny(nspin=1, nspinor=1, nmag=1, nbnds=4, fft=(12, 12, 12), kgrid=(2, 2, 2),
         syms="trev", ecut=6.0, seed=7, q=(0.0, 0.0, 0.5), with_wfc=True,
         dimension="3", soc=False, crys_pos=None) -> System:
    """A small fcc cell for branch coverage (spin, spinors, symmetries);
    dimension "2": a hexagonal slab cell (a = 5, c = 20 bohr, 12x12x48)
    with the 2-D Coulomb cut-off.  crys_pos replaces the two atoms'
    crystal positions of the cubic cell."""
    a = 8.0
    I3 = np.eye(3)
    if dimension == "2":
        sl = [(I3, np.zeros(3), 0), (-I3, np.zeros(3), 1)]
        return make_system("tiny2d", seed, hex_cell(5.0, 20.0),
                           [[1 / 3, 2 / 3, 0.5], [2 / 3, 1 / 3, 0.56]], [0, 1], [3.0, 5.0],
                           (12, 12, 48), (kgrid[0], kgrid[1], 1), ecut, nbnds, sl, 1,
                           [(0.5, 0.0, 0.0)], nspin=nspin, nspinor=nspinor, nmag=nmag,
                           dimension="2", with_wfc=with_wfc, soc=soc)
    if syms == "id":
        sl, tr = [(I3, np.zeros(3), 0)], 0
    elif syms == "trev":
        sl, tr = [(I3, np.zeros(3), 0), (-I3, np.zeros(3), 1)], 1
    elif syms == "cubic":
        # C4 about z and its time-reversed partner
        r = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
        sl = [(I3, np.zeros(3), 0), (r, np.zeros(3), 0), (r @ r, np.zeros(3), 0),
              (r @ r @ r, np.zeros(3), 0), (-I3, np.zeros(3), 1), (-r, np.zeros(3), 1),
              (-(r @ r), np.zeros(3), 1), (-(r @ r @ r), np.zeros(3), 1)]
        tr = 1
        pos = [[0, 0, 0], [0.5, 0.5, 0.3]] if crys_pos is None else crys_pos
        return make_system("tiny", seed, a * I3, pos, [0, 1],
                           [3.0, 5.0], fft, kgrid, ecut, nbnds, sl, tr, [q],
                           nspin=nspin, nspinor=nspinor, nmag=nmag, with_wfc=with_wfc, soc=soc)
    else:
        raise ValueError(syms)
    return make_system("tiny", seed, fcc_cell(a), [[0, 0, 0], [0.25, 0.25, 0.25]], [0, 1],
                       [3.0, 5.0], fft, kgrid, ecut, nbnds, sl, tr, [q], nspin=nspin,
                       nspinor=nspinor, nmag=nmag, with_wfc=with_wfc, soc=soc)


def kq_pairs(sys: System, iq: int, kminusq: bool = False):
    """The (ikq, ik, kqsym, ksym) work items of compute_and_write_elphq
    (src/elph/compute_elph_matq.c:67-133) for q iq, in full-BZ k order."""
    q = np.asarray(sys.qpts[iq])
    qq = -q if kminusq else q
    kb = sys.kpt_bz_crys
    out = []
    for i in range(len(kb)):
        kq = kb[i] + qq
        d = kb - kq[None, :]
        d = d - np.rint(d)
        j = int(np.argmin(np.linalg.norm(d, axis=1)))
        if np.linalg.norm(d[j]) >= 1e-5:
            raise ValueError("k+q not on the grid")
        ik_, ikq_ = i, j
        if kminusq:
            ik_, ikq_ = ikq_, ik_
        ik, ksym = sys.kmap[ik_]
        ikq, kqsym = sys.kmap[ikq_]
        out.append((int(ikq), int(ik), int(kqsym), int(ksym)))
    return out
