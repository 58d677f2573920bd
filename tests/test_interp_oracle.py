"""The oracle's Fourier interpolation of dV_q (oracle/interp_oracle.c) and
the library's host Wigner-Seitz search, pinned on the CPU:

  * against tests/golden/interp_vectors.npz, written from the reference's own
    wigner_seitz.c (kd-tree), dVlong_range_kernel.c, multiply_eikr.c and
    rotate_dvscf.c compiled into oracle/_ref (and live on a second seed when
    that library is present);
  * fft_q2R against numpy's FFT;
  * the long-range G sum placed on the grid against a brute-force DFT;
  * the whole coarse -> R -> q chain reproduces the coarse potentials;
  * the dipole branch against the monopole branch at equivalent parameters;
  * the cartesian/mode basis change as a round trip.
"""
import ctypes as C
import os

import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from letzelphc_amd.abi import LrParams, dptr, i64ptr
from oracle import oracle as O

from .util import rel_err, say

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "interp_vectors.npz")


def _load():
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _ws_sets(vecs, degen):
    out, o = [], 0
    for d in degen:
        out.append(sorted(map(tuple, vecs[o:o + d])))
        o += d
    return out


def _check_ws(f, impl):
    for name in ("fcc", "fccdyn", "hex"):
        lat, g, pos = f[f"ws_{name}_lat"], f[f"ws_{name}_grid"], f[f"ws_{name}_pos"]
        pair = bool(f[f"ws_{name}_pair"])
        v, d = impl(g, lat, 1e-6, pos, pos if pair else None)
        np.testing.assert_array_equal(d, f[f"ws_{name}_degen"])
        assert _ws_sets(v, d) == _ws_sets(f[f"ws_{name}_vecs"], f[f"ws_{name}_degen"]), name


def check_against(f):
    _check_ws(f, O.ws_vectors)
    for dim in ("3", "2"):
        q, g, pos = f[f"lr{dim}_q"], f[f"lr{dim}_g"], f[f"lr{dim}_pos"]
        eps, zb, qp, zv = (f[f"lr{dim}_{k}"] for k in ("eps", "zb", "qp", "zv"))
        for case in range(4):
            args = {0: (zv, None, None, None), 1: (None, eps, zb, None),
                    2: (None, eps, zb, qp), 3: (zv, eps, zb, qp)}[case]
            got = O.lr_kernel(q, g, args[0], args[1], args[2], args[3], pos, dim,
                              float(f[f"lr{dim}_vol"]), float(f[f"lr{dim}_zlat"]), 30.0, 1.0, 1.3)
            ref = f[f"lr{dim}_out"][case]
            assert np.abs(ref).max() > 0
            assert rel_err(got, ref) < 1e-12, (dim, case)
    dims = np.array([4, 5, 6], np.int64)
    for i, sg in enumerate((1, -1)):
        t = f["ek_pot"].copy()
        O.lib().orc_multiply_eikr(dptr(t), dptr(f["ek_q"]), i64ptr(dims), 2, sg)
        assert rel_err(t, f["ek_out"][i]) < 1e-13
    lat = _rot_lattice(f)
    for nmag in (1, 4):
        lat.nmag = nmag
        din = f[f"rot{nmag}_in"]
        j = 0
        for R, tau, tr in zip(f["rot_R"], f["rot_tau"], f["rot_tr"]):
            for comp in ((0, 1) if nmag == 4 else (0,)):
                out = np.empty_like(din)
                O.check(O.lib().orc_rotate_dvscf(dptr(f["rot_q"]), dptr(din),
                                                 C.byref(O.make_sym(R, tau, tr)), C.byref(lat),
                                                 comp, dptr(out)))
                assert rel_err(out, f[f"rot{nmag}_out"][j]) < 1e-13, (nmag, j)
                j += 1


def _rot_lattice(f):
    from letzelphc_amd.abi import Lattice
    L = Lattice()
    L.natom = 0
    L.nmodes = 2
    L.nspinor = 2
    for i in range(3):
        L.fft_dims[i] = 6
    for i in range(9):
        L.alat_vec[i] = float(f["rot_alat"].reshape(9)[i])
        L.blat_vec[i] = float(f["rot_blat"].reshape(9)[i])
    return L


def test_interp_oracle_vs_reference_fixture():
    say("case: interpolation helpers vs reference fixture")
    check_against(_load())


@pytest.mark.skipif(not os.path.exists(O.REF_LIB), reason="oracle/_ref not built")
def test_interp_oracle_vs_reference_live():
    say("case: interpolation helpers vs reference library, second seed")
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_ref_vectors", os.path.join(HERE, "golden", "make_ref_vectors.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    L = mod.load_interp(mod.load_ref())
    check_against(mod.make_interp(L, seed=5))


def test_host_wigner_seitz_vs_fixture():
    """The library's own host search (lelph_wigner_seitz) against the
    reference's kd-tree results."""
    say("case: lelph_wigner_seitz vs reference fixture")
    _check_ws(_load(), lelph.wigner_seitz)


def test_fft_q2R_vs_numpy():
    rng = np.random.default_rng(3)
    g = np.array([3, 4, 2], np.int64)
    d = rng.standard_normal((3, 4, 2, 5)) + 1j * rng.standard_normal((3, 4, 2, 5))
    t = np.ascontiguousarray(d.reshape(24, 5)).copy()
    O.lib().orc_fft_q2R(dptr(t), i64ptr(g), 5)
    ref = np.fft.fftn(d, axes=(0, 1, 2)) / 24
    assert rel_err(t.reshape(3, 4, 2, 5), ref) < 1e-14


def _small_system(dimension="3"):
    if dimension == "2":
        s = synth.tiny(dimension="2", with_wfc=False)
    else:
        s = synth.tiny(with_wfc=False)
    return s


def _lr_params(s, quad=True):
    rng = np.random.default_rng(11)
    na = s.natom
    zb = rng.standard_normal((na, 3, 3))
    zb -= zb.mean(0)  # acoustic sum rule
    eps = 6.0 * np.eye(3) + 0.2 * np.diag(rng.uniform(-1, 1, 3))
    return LrParams(epsilon=eps, Zborn=zb, Qpole=rng.standard_normal((na, 3, 3, 3)) if quad
                    else None, Zvals=np.array([3.0, 5.0])[:na], only_induced=False,
                    eta_bare=1.0, eta_induced=1.0, EcutRy=12.0, nmags_add=(1, 1, 0, 0))


@pytest.mark.parametrize("dimension", ["3", "2"])
def test_longrange_grid_vs_bruteforce(dimension):
    """dV_add_longrange's G box, placement and backward FFT against a direct
    sum over the box of the kernel's coefficients."""
    s = _small_system(dimension)
    lr = _lr_params(s)
    say(f"case: long range on the grid vs brute force, dimension {dimension}")
    oi = O.OracleInterp(s, (2, 2, 1), lr=lr)
    q = np.array([0.31, -0.12, 0.0])
    nm, nmag = s.nmodes, s.nmag
    dv0 = np.zeros((nm, nmag) + tuple(s.fft_dims), np.complex128)
    got = oi.longrange(q, dv0, None, 1)
    # brute force: every G of the box, e^{2 pi i G.r} on the grid
    gb = np.zeros(3, np.int64)
    O.lib().orc_get_fft_box(lr.EcutRy, dptr(np.ascontiguousarray(s.blat)), i64ptr(gb))
    gb = np.minimum(gb, s.fft_dims)
    if dimension == "2":
        gb[2] = s.fft_dims[2]
    ms = [synth.miller_axis(int(n)) for n in gb]
    M = np.stack(np.meshgrid(*ms, indexing="ij"), -1).reshape(-1, 3)
    gv = (M @ s.blat.T) / (2 * synth.RPI)
    qc = s.blat @ q / (2 * synth.RPI)
    V = O.lr_kernel(qc, gv, lr.Zvals, lr.epsilon, lr.Zborn, lr.Qpole, s.atomic_pos, dimension,
                    s.volume, s.alat[2, 2], lr.EcutRy, lr.eta_bare, lr.eta_induced)
    V = V.reshape(len(M), nm)
    N = np.array(s.fft_dims)
    r = [np.arange(n) / n for n in N]
    ph = np.exp(2j * np.pi * (M[:, 0, None] * r[0][None, :]))[:, :, None, None] * \
        np.exp(2j * np.pi * (M[:, 1, None] * r[1][None, :]))[:, None, :, None] * \
        np.exp(2j * np.pi * (M[:, 2, None] * r[2][None, :]))[:, None, None, :]
    ref = np.einsum("gv,gxyz->vxyz", V, ph)
    assert np.abs(ref).max() > 0
    assert rel_err(got[:, 0], ref) < 1e-11
    # nmags_add (1, 1): the second magnetic component gets the same term
    if nmag > 1:
        assert rel_err(got[:, 1], ref) < 1e-11


def test_dipole_vs_monopole_branch():
    """With Z* = z I, eps = e I and eta_bare = eta_ind / e, the dipole term
    equals the monopole term of a charge -z/e (both are -d/dtau of one
    screened Coulomb potential)."""
    rng = np.random.default_rng(4)
    lat = synth.fcc_cell(10.26)
    blat = 2 * np.pi * np.linalg.inv(lat).T
    m = rng.integers(-3, 4, (30, 3)).astype(np.float64)
    gv = (m @ blat.T) / (2 * np.pi)
    qc = blat @ rng.uniform(-0.5, 0.5, 3) / (2 * np.pi)
    pos = rng.uniform(0, 1, (2, 3)) @ lat.T
    z, e, eta = 1.7, 5.0, 1.3
    vol = abs(np.linalg.det(lat))
    dip = O.lr_kernel(qc, gv, None, e * np.eye(3), np.stack([z * np.eye(3)] * 2), None, pos,
                      "3", vol, lat[2, 2], 40.0, eta / e, eta)
    mono = O.lr_kernel(qc, gv, np.full(2, -z / e), None, None, None, pos, "3", vol, lat[2, 2],
                       40.0, eta / e, eta)
    assert np.abs(dip).max() > 0
    assert rel_err(dip, mono) < 1e-12


def test_change_basis_round_trip():
    rng = np.random.default_rng(5)
    nm, nmag, nf = 6, 2, 30
    u, _ = np.linalg.qr(rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm)))
    u = np.ascontiguousarray(u)
    d = rng.standard_normal((nm, nmag, nf)) + 1j * rng.standard_normal((nm, nmag, nf))
    t = np.ascontiguousarray(d).copy()
    O.lib().orc_dvscf_change_basis(dptr(t), dptr(u), 1, nm, nmag, nf, b"N")
    assert rel_err(t, np.einsum("ab,bmf->amf", u, d)) < 1e-13
    O.lib().orc_dvscf_change_basis(dptr(t), dptr(u), 1, nm, nmag, nf, b"C")
    assert rel_err(t, d) < 1e-13


def coarse_grid(grid):
    g = np.stack(np.meshgrid(*[np.arange(n) / n for n in grid], indexing="ij"), -1)
    return g.reshape(-1, 3)


@pytest.mark.parametrize("with_lr", [False, True])
def test_interpolation_exact_on_coarse_grid(with_lr):
    """q -> R -> q returns every coarse potential in the cartesian basis
    (the long range, when present, is removed and added back).  The
    reference's q -> R uses FFTW's exact pi while R -> q and e^{iqr} use
    ELPH_PI = 3.1415927, so the round trip holds to ~1e-7, not to rounding."""
    s = synth.tiny(with_wfc=False)
    grid = (2, 2, 2)
    qs = coarse_grid(grid)
    rng = np.random.default_rng(8)
    masses = np.array([1.5, 2.5])
    lr = _lr_params(s, quad=False) if with_lr else None
    oi = O.OracleInterp(s, grid, masses=masses, lr=lr)
    dvs, eigs = [], []
    for i, q in enumerate(qs):
        dv = rng.standard_normal(oi.shape) + 1j * rng.standard_normal(oi.shape)
        e, _ = np.linalg.qr(rng.standard_normal((s.nmodes,) * 2)
                            + 1j * rng.standard_normal((s.nmodes,) * 2))
        # displacement patterns e / sqrt(M), the convention in which the
        # mass-weighted conjugate transpose is the inverse
        e = e / np.sqrt(np.repeat(masses, 3))[None, :]
        dvs.append(dv)
        eigs.append(e)
        oi.add_coarse(q, e, dv)
    oi.finish()
    say(f"case: interpolation exact on the coarse grid, long range {with_lr}")
    for i, q in enumerate(qs):
        got = oi.dvscf(q)
        rot = eigs[i] * np.repeat(masses, 3)[None, :]
        ref = np.einsum("va,vm...->am...", rot.conj(), dvs[i])
        assert rel_err(got, ref) < 2e-6, i


def test_dvscf_qe_files(tmp_path):
    """read_dvscf_qe / write_dvscf_qe of the oracle against numpy's view of
    the QE layout: sets of (Nz, Ny, Nx) complex128, x fastest."""
    s = synth.tiny(nmag=2, nspin=2, fft=(4, 5, 6), with_wfc=False)
    from letzelphc_amd.abi import pack_system
    pk = pack_system(s)
    rng = np.random.default_rng(6)
    sh = (s.nmodes, s.nmag) + tuple(s.fft_dims)
    dv = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    f1 = str(tmp_path / "dv1")
    np.ascontiguousarray(dv.transpose(0, 1, 4, 3, 2)).tofile(f1)
    eye = np.ascontiguousarray(np.eye(s.nmodes, dtype=np.complex128))
    out = np.empty(sh, np.complex128)
    O.check(O.lib().orc_read_dvscf_qe(f1.encode(), C.byref(pk.lattice), dptr(eye), dptr(eye),
                                      dptr(out)))
    assert np.array_equal(out, dv)
    f2 = str(tmp_path / "dv2")
    O.check(O.lib().orc_write_dvscf_qe(f2.encode(), C.byref(pk.lattice), dptr(dv)))
    assert open(f1, "rb").read() == open(f2, "rb").read()
    # patterns: eig @ pats^H @ dV
    e = rng.standard_normal((s.nmodes,) * 2) + 1j * rng.standard_normal((s.nmodes,) * 2)
    p, _ = np.linalg.qr(rng.standard_normal((s.nmodes,) * 2)
                        + 1j * rng.standard_normal((s.nmodes,) * 2))
    e, p = np.ascontiguousarray(e), np.ascontiguousarray(p)
    O.check(O.lib().orc_read_dvscf_qe(f1.encode(), C.byref(pk.lattice), dptr(e), dptr(p),
                                      dptr(out)))
    assert rel_err(out, np.einsum("ab,bm...->am...", e @ p.conj().T, dv)) < 1e-13


# ---------------------------------------------------------------------------
# dynamical matrices
# ---------------------------------------------------------------------------

def _dyn_lr(s, lr):
    return O.OracleDynInterp(s, (1, 1, 1), masses=None, lr=lr, eta_ph=1.0, ecut=30.0,
                             asr="none")


def test_pol_vecs_to_dyn_eigensystem():
    """pol_vecs_to_dyn builds the matrix whose eigenpairs are sign(w) w^2 and
    the mass-weighted polarisation vectors."""
    rng = np.random.default_rng(2)
    na, nm = 2, 6
    masses = np.array([1.7, 3.1])
    e, _ = np.linalg.qr(rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm)))
    om = np.array([-0.01, 0.02, 0.5, 0.7, 1.1, 1.3])
    pol = np.ascontiguousarray(e / np.sqrt(np.repeat(masses, 3))[None, :])
    d = pol.copy()
    O.lib().orc_pol_vecs_to_dyn(dptr(om), na, dptr(masses), dptr(d))
    f = np.sign(om) * om ** 2
    assert rel_err(np.linalg.eigvalsh(d), np.sort(f)) < 1e-13
    for i in range(nm):
        assert np.abs(d @ e[i] - f[i] * e[i]).max() < 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("dimension", ["3", "2"])
def test_dyn_longrange_invariants(dimension):
    """The long-range dynamical matrix is Hermitian, satisfies D(-q) =
    conj D(q), is unchanged by a common shift of the atoms, and with
    charge-neutral Born charges its q = 0 row sums are the ASR correction
    the reference subtracts."""
    s = _small_system(dimension)
    lr = _lr_params(s)
    od = _dyn_lr(s, lr)
    q = np.array([0.23, -0.11, 0.0 if dimension == "2" else 0.17])
    nm = s.nmodes
    z = np.zeros((nm, nm), np.complex128)
    D = od.lr_term(q, z, 1, masses=False, asr=False)
    assert np.abs(D).max() > 0
    assert rel_err(D, D.conj().T) < 1e-12
    assert rel_err(od.lr_term(-q, z, 1, masses=False, asr=False), D.conj()) < 1e-12
    s2 = _small_system(dimension)
    s2.atomic_pos = s2.atomic_pos + np.array([0.37, -0.21, 0.0 if dimension == "2" else 0.5])
    D2 = _dyn_lr(s2, lr).lr_term(q, z, 1, masses=False, asr=False)
    assert rel_err(D2, D) < 1e-11
    D0 = od.lr_term(np.zeros(3), z, 1, masses=False, asr=False)
    rows = D0.reshape(s.natom, 3, s.natom, 3).sum(2)
    assert np.abs(rows.reshape(-1) - od.asr_lr).max() <= 1e-12 * max(np.abs(D0).max(), 1e-300)
    Dc = od.lr_term(np.zeros(3), z, 1, masses=False, asr=True)
    assert np.abs(Dc.reshape(s.natom, 3, s.natom, 3).sum(2)).max() <= 1e-10 * np.abs(D0).max()


def test_dyn_longrange_nonanalytic_limit():
    """Near q = 0 only the G = 0 term depends on the direction of q:
    D(d1 dq) - D(d2 dq) -> (4 pi e^2 / Omega) [Z^T d (Z^T d)^T / (d eps d)]
    evaluated for d1 minus d2 (Gonze's non-analytic term)."""
    s = synth.tiny(with_wfc=False)
    lr = _lr_params(s, quad=False)
    od = _dyn_lr(s, lr)
    nm = s.nmodes
    z = np.zeros((nm, nm), np.complex128)
    dq = 2e-5  # |q| above the reference's q + G = 0 threshold (1e-5, cartesian)
    dirs = [np.array([1.0, 0, 0]), np.array([0.3, -0.5, 0.8])]
    Ds, NA = [], []
    for d in dirs:
        dcrys = np.linalg.solve(s.blat, d / np.linalg.norm(d))  # crystal of a unit cart vector
        Ds.append(od.lr_term(dq * dcrys, z, 1, masses=False, asr=False))
        qc = s.blat @ (dq * dcrys)
        zq = np.einsum("aij,i->aj", lr.Zborn, qc).reshape(-1)
        NA.append(4 * synth.RPI * 2.0 / s.volume * np.outer(zq, zq) / (qc @ lr.epsilon @ qc))
    assert rel_err(Ds[0] - Ds[1], NA[0] - NA[1]) < 2e-4  # O(|q|) from the analytic rest


def test_dyn_longrange_nonanalytic_limit_2d():
    """2-D cells: near q = 0 the in-plane G = 0 term is the published 2-D
    form (2 pi e^2 / A) (Z^T q)(Z^T q)^T / (|q| (1 + r_eff |q|)) (Sohier et
    al., PRB 94, 085415 (2016)), linear in |q| and direction dependent; A is
    the in-plane area |a1 x a2|.  The even part in q removes the odd
    analytic terms, the difference of two in-plane directions the q = 0
    limit and the direction-independent out-of-plane |q| term, and a
    Richardson step the O(|q|^2) rest, leaving (2 pi e^2 / A) [(Z^T d1)
    (Z^T d1)^T - (Z^T d2)(Z^T d2)^T] per unit |q|.  The 3-D kernel (an O(1)
    direction dependence), a transposed Z, or a prefactor missing the slab
    height's c / 2 fail it; the out-of-plane term and r_eff are not pinned."""
    s = _small_system("2")
    lr = _lr_params(s, quad=False)
    od = _dyn_lr(s, lr)
    nm = s.nmodes
    z = np.zeros((nm, nm), np.complex128)
    area = np.linalg.norm(np.cross(s.alat[0], s.alat[1]))
    dirs = [np.array([1.0, 0, 0]), np.array([0.6, -0.8, 0.0])]

    def even(h, d):
        dcrys = np.linalg.solve(s.blat, h * d)
        return 0.5 * (od.lr_term(dcrys, z, 1, masses=False, asr=False) +
                      od.lr_term(-dcrys, z, 1, masses=False, asr=False))

    h = 2e-5  # |q| above the reference's q + G = 0 threshold (1e-5, cartesian)
    slope = {}
    for hh in (h, 2 * h):
        slope[hh] = (even(hh, dirs[0]) - even(hh, dirs[1])) / hh
    got = 2 * slope[h] - slope[2 * h]
    want = np.zeros((nm, nm))
    for d, sg in zip(dirs, (1, -1)):
        zq = np.einsum("aij,i->aj", lr.Zborn, d).reshape(-1)
        want += sg * 2 * synth.RPI * 2.0 / area * np.outer(zq, zq)
    say(f"case: 2-D non-analytic slope, |got - want|/|want| = {rel_err(got, want):.2e}")
    assert rel_err(got.real, want) < 1e-5
    assert np.abs(got.imag).max() < 1e-5 * np.abs(want).max()
    # without the Richardson step the O(|q|) rest shows (a check on h)
    assert rel_err(slope[h].real, want) > 10 * rel_err(got.real, want)


@pytest.mark.parametrize("asr", ["none", "simple"])
def test_dyn_interpolation_coarse_and_asr(asr):
    """coarse -> R -> q returns each coarse dynamical matrix (asr none, to
    the reference's pi mismatch), and the simple sum rule zeroes the
    R-summed mass-unweighted force constants of every atom row."""
    s = synth.tiny(with_wfc=False)
    grid = (2, 2, 2)
    masses = np.array([1.5, 2.5])
    lr = _lr_params(s, quad=False)
    od = O.OracleDynInterp(s, grid, masses=masses, lr=lr, eta_ph=1.0, asr=asr)
    rng = np.random.default_rng(3)
    dyns = []
    for q in coarse_grid(grid):
        e, _ = np.linalg.qr(rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)))
        om = rng.uniform(0.1, 1.0, 6)
        pol = e / np.sqrt(np.repeat(masses, 3))[None, :]
        od.add_coarse(q, om, pol)
        d = pol.copy()
        O.lib().orc_pol_vecs_to_dyn(dptr(om), 2, dptr(masses), dptr(d))
        dyns.append(d)
    od.finish()
    say(f"case: dyn interpolation, asr {asr}")
    if asr == "none":
        for i, q in enumerate(coarse_grid(grid)):
            assert rel_err(od.dyn(q), dyns[i]) < 2e-6, i
    else:
        M = np.sqrt(np.outer(np.repeat(masses, 3), np.repeat(masses, 3)))
        fc = od.fc * M[None]
        s_ = fc.reshape(len(fc), 2, 3, 2, 3).sum(axis=(0, 3))
        assert np.abs(s_).max() < 1e-12 * np.abs(fc).max()


# ---------------------------------------------------------------------------
# spin-orbit f coefficients and the fully relativistic non-local part
# ---------------------------------------------------------------------------

def test_soc_fcoeff_vs_reference_and_completeness():
    """f^{l,j} (the synthetic inputs' restatement of fcoeff.c) against the
    reference's compiled alloc_and_Compute_f_Coeff, and sum_j f^{l,j} = 1
    (to the reference's truncated sqrt(2) = 1.4142136)"""
    f = _load()
    for l in range(4):
        tjs = (1,) if l == 0 else (2 * l - 1, 2 * l + 1)
        tot = 0
        for tj in tjs:
            got = synth.soc_fcoeff(l, tj)
            assert np.abs(got - f[f"fc_l{l}_j{tj}"]).max() < 1e-15, (l, tj)
            tot = tot + got
        I = np.einsum("ab,st->abst", np.eye(2 * l + 1), np.eye(2))
        assert np.abs(tot - I).max() < 1e-7, l


def test_soc_nonlocal_reduces_to_scalar():
    """With both j = l +- 1/2 channels carrying the same projector and D,
    the fully relativistic non-local term of the oracle equals the scalar
    spinor one (sum_j f^{l,j} = 1), so g agrees to the sqrt(2) truncation."""
    say("case: SOC non-local with identical j channels vs scalar spinors")
    a = synth.tiny(nspinor=2, soc=True, nbnds=4)
    b = synth.tiny(nspinor=2, soc=False, nbnds=4)
    assert np.array_equal(a.wfcs[0].wfc, b.wfcs[0].wfc)
    oa, ob = O.OracleSystem(a), O.OracleSystem(b)
    dv = oa.dv_total(0)
    for pr in synth.kq_pairs(a, 0)[:2]:
        z = np.zeros((a.nmodes, a.nspin, a.nbnds, a.nbnds), np.complex128)
        ga = oa.add_nonlocal(0, *pr, z)
        gb = ob.add_nonlocal(0, *pr, z)
        assert np.abs(gb).max() > 0
        assert rel_err(ga, gb) < 1e-6, pr


def _dd_short_and_self(s, lr, q_crys, eta, rmax=36.0):
    """Real-space part and self term of the Ewald sum of the multipole
    (Born charge + quadrupole) dynamical matrix (Gonze & Lee, PRB 55, 10355
    (1997), sec. IV; the quadrupole terms as in Smith's multipole Ewald sum),
    derived here from the split 1/(sqrt(det eps) D) = [erf(L D) + erfc(L D)]
    / (sqrt(det eps) D), D^2 = d.M.d, M = eps^-1, L^2 = eta, e^2 = 2
    (Rydberg), on the lattice whose reciprocal vectors are the system's blat
    (which carries the reference's truncated pi).  An atom's displacement
    charge in reciprocal space is X(K) = Z^T K - (i/2) K.Q.K, so the pair
    kernel conj(X_a)_i X_b_j is a polynomial in K of degree 2..4 and, with
    K -> -i grad, the real-space kernel is
      - Z_a,ki Z_b,lj d_kl phi + 1/2 Z_a,ki Q_b,lmj d_klm phi
      - 1/2 Q_a,lmi Z_b,kj d_klm phi + 1/4 Q_a,lmi Q_b,npj d_lmnp phi,
    with the derivatives of phi = g(u = D^2) / sqrt(det eps) from the B_n
    recursion (B_0 = erfc(L D) / D, B_n = [(2n-1) B_(n-1) + (2 L^2)^n
    e^(-L^2 D^2) / (L sqrt(pi))] / D^2, and 2 g1 = -B_1, 4 g2 = B_2,
    8 g3 = -B_3, 16 g4 = B_4 for the u-derivatives g1 .. g4) and w = M d.
    Returns sum'_R F_short(tau_a - tau_b + R) e^(-i q.R) minus the
    long-range part's own term at d = 0 for a = b (its second and fourth
    derivatives there: -(4 L^3 / (3 sqrt(pi))) M and (8 L^5 / (5 sqrt(pi)))
    times the three pairings M M), in the index order of the oracle's
    dynamical matrix [(a, i), (b, j)]."""
    from scipy.special import erfc
    e2 = 2.0
    B = np.asarray(s.blat, np.float64)
    A = 2 * np.pi * np.linalg.inv(B).T  # columns a_j with a_j . b_i = 2 pi delta_ij
    eps = np.asarray(lr.epsilon, np.float64)
    M = np.linalg.inv(eps)
    sq = np.sqrt(np.linalg.det(eps))
    L = np.sqrt(eta)
    Z = np.asarray(lr.Zborn, np.float64)  # Z[a, k, i]: (Z^T K)_i = sum_k Z[a, k, i] K_k
    # Q[a, i, j, k]: (K.Q.K)_k = sum_ij Q[a, i, j, k] K_i K_j
    Q = None if lr.Qpole is None else np.asarray(lr.Qpole, np.float64)
    tau = np.asarray(s.atomic_pos, np.float64)
    na = s.natom
    nmax = int(np.ceil(rmax / np.min(np.linalg.norm(A, axis=0)))) + 1
    rng_ = np.arange(-nmax, nmax + 1)
    n = np.stack(np.meshgrid(rng_, rng_, rng_, indexing="ij"), -1).reshape(-1, 3)
    R = n @ A.T
    phase = np.exp(-2j * np.pi * (n @ np.asarray(q_crys, np.float64)))

    def sym_m_ww(w):  # the 6 placements of M among 4 indices, w on the other two
        return (np.einsum("kn,xl,xm->xklmn", M, w, w) + np.einsum("ln,xk,xm->xklmn", M, w, w)
                + np.einsum("mn,xk,xl->xklmn", M, w, w) + np.einsum("km,xl,xn->xklmn", M, w, w)
                + np.einsum("lm,xk,xn->xklmn", M, w, w) + np.einsum("kl,xm,xn->xklmn", M, w, w))
    MM = (np.einsum("km,ln->klmn", M, M) + np.einsum("lm,kn->klmn", M, M)
          + np.einsum("kl,mn->klmn", M, M))
    out = np.zeros((na, 3, na, 3), np.complex128)
    for a in range(na):
        for b in range(na):
            d = tau[a] - tau[b] + R
            keep = np.linalg.norm(d, axis=1) > 1e-9
            d, ph = d[keep], phase[keep]
            w = d @ M
            D = np.sqrt(np.einsum("xi,xi->x", d, w))
            ex = np.exp(-(L * D) ** 2) / (L * np.sqrt(np.pi))
            Bn = [erfc(L * D) / D]
            for k in range(1, 5):
                Bn.append(((2 * k - 1) * Bn[-1] + (2 * L ** 2) ** k * ex) / D ** 2)
            d2 = (np.einsum("x,xk,xl->xkl", Bn[2], w, w) - Bn[1][:, None, None] * M[None]) / sq
            F = -np.einsum("ki,xkl,lj->xij", Z[a], d2, Z[b])
            if Q is not None:
                d3 = (-np.einsum("x,xk,xl,xm->xklm", Bn[3], w, w, w)
                      + Bn[2][:, None, None, None]
                      * (np.einsum("km,xl->xklm", M, w) + np.einsum("lm,xk->xklm", M, w)
                         + np.einsum("kl,xm->xklm", M, w))) / sq
                d4 = (np.einsum("x,xk,xl,xm,xn->xklmn", Bn[4], w, w, w, w)
                      - Bn[3][:, None, None, None, None] * sym_m_ww(w)
                      + Bn[2][:, None, None, None, None] * MM[None]) / sq
                F = F + 0.5 * np.einsum("ki,lmj,xklm->xij", Z[a], Q[b], d3)
                F = F - 0.5 * np.einsum("lmi,kj,xklm->xij", Q[a], Z[b], d3)
                F = F + 0.25 * np.einsum("lmi,npj,xlmnp->xij", Q[a], Q[b], d4)
            out[a, :, b, :] = np.einsum("xij,x->ij", F, ph)
        s2 = -(4 * L ** 3 / (3 * np.sqrt(np.pi) * sq)) * M
        self_ = -np.einsum("ki,kl,lj->ij", Z[a], s2, Z[a])
        if Q is not None:
            s4 = (8 * L ** 5 / (5 * np.sqrt(np.pi) * sq)) * MM
            self_ = self_ + 0.25 * np.einsum("lmi,npj,lmnp->ij", Q[a], Q[a], s4)
        out[a, :, a, :] -= self_
    return e2 * out.reshape(3 * na, 3 * na)


@pytest.mark.parametrize("quad", [False, True])
def test_dyn_longrange_ewald_invariance(quad):
    """The oracle's long-range dynamical matrix (the reciprocal sum over
    q + G != 0 with the Gaussian e^{-(q+G).eps.(q+G)/(4 eta)} of
    add_ph_dyn_long_range, src/phonon/dyn_long_range.c:125-404) is the
    reciprocal half of an Ewald sum: adding the real-space erfc half and the
    self term derived independently above gives the full dipole-dipole
    lattice sum, which cannot depend on eta.  Checked for three eta at a
    general q, with anisotropic eps and non-symmetric Born charges; the
    reciprocal parts alone differ by O(1) between the eta, and a reciprocal
    part off by 0.1 % breaks the invariance.  The reference's 4 pi and
    volume use ELPH_PI and the true cell while its reciprocal vectors carry
    ELPH_PI: the exact Ewald normalisation of that reciprocal lattice is
    (pi / ELPH_PI) (Omega / Omega'), applied to the oracle's sum."""
    s = synth.tiny(with_wfc=False)
    lr = _lr_params(s, quad=quad)
    lr.epsilon = np.array([[5.6, 0.3, -0.1], [0.3, 6.2, 0.2], [-0.1, 0.2, 5.9]])
    q = np.array([0.23, -0.11, 0.17])
    nm = s.nmodes
    z = np.zeros((nm, nm), np.complex128)
    B = np.asarray(s.blat, np.float64)
    vol_p = abs(np.linalg.det(2 * np.pi * np.linalg.inv(B).T))
    norm = (np.pi / synth.RPI) * (s.volume / vol_p)
    say(f"case: Ewald invariance of the long-range dynamical matrix, quadrupoles {quad}")
    tot, rec = [], []
    for eta in (0.6, 1.0, 1.8):
        od = O.OracleDynInterp(s, (1, 1, 1), masses=None, lr=lr, eta_ph=eta, ecut=320.0,
                               asr="none")
        r = od.lr_term(q, z, 1, masses=False, asr=False) * norm
        rec.append(r)
        tot.append(r + _dd_short_and_self(s, lr, q, eta))
    scale = np.abs(tot[0]).max()
    spread = max(np.abs(t - tot[0]).max() for t in tot[1:]) / scale
    rec_change = np.abs(rec[2] - rec[0]).max() / scale
    planted = np.abs((1.001 * rec[2] + tot[2] - rec[2]) - tot[0]).max() / scale
    say(f"  total spread over eta {spread:.2e}; reciprocal parts differ by {rec_change:.2e}; "
        f"reciprocal x 1.001: {planted:.2e}")
    assert rec_change > 1e-1
    assert spread < 1e-10
    assert planted > 1e-4
    if quad:
        # quadrupoles of the wrong sign in the reciprocal sum only (the
        # dipole-quadrupole cross terms flip) break the invariance
        import copy
        lr2 = copy.copy(lr)
        lr2.Qpole = -np.asarray(lr.Qpole)
        t2 = []
        for eta in (0.6, 1.8):
            od = O.OracleDynInterp(s, (1, 1, 1), masses=None, lr=lr2, eta_ph=eta, ecut=320.0,
                                   asr="none")
            t2.append(od.lr_term(q, z, 1, masses=False, asr=False) * norm
                      + _dd_short_and_self(s, lr, q, eta))
        wrong = np.abs(t2[1] - t2[0]).max() / scale
        say(f"  quadrupole sign flipped in the reciprocal sum: spread {wrong:.2e}")
        assert wrong > 1e-4
