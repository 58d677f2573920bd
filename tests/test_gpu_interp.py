"""GPU parity of the dV_q Fourier interpolation against the oracle
(oracle/interp_oracle.c), through the C-ABI: the long-range term
(dV_add_longrange), rotate_dvscf, the whole coarse -> R -> q chain with star
members, and g(k,q) computed from an interpolated potential.

Bar: FP64, |d|max / |ref|max <= 1e-10.
"""
import os

import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from letzelphc_amd.abi import LrParams
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def lr_params(s, quad=True, nmags=(1, 1, 0, 0), seed=11):
    rng = np.random.default_rng(seed)
    na = s.natom
    zb = rng.standard_normal((na, 3, 3))
    zb -= zb.mean(0)
    eps = 6.0 * np.eye(3) + 0.2 * np.diag(rng.uniform(-1, 1, 3))
    return LrParams(epsilon=eps, Zborn=zb, Qpole=rng.standard_normal((na, 3, 3, 3)) if quad
                    else None, Zvals=np.linspace(3.0, 5.0, na), only_induced=False,
                    eta_bare=1.0, eta_induced=1.0, EcutRy=12.0, nmags_add=nmags)


def masses_for(s):
    return np.linspace(1.5, 2.5, s.natom)


def patterns(s, rng, masses):
    """random displacement patterns e / sqrt(M) (rows: modes)"""
    e, _ = np.linalg.qr(rng.standard_normal((s.nmodes,) * 2)
                        + 1j * rng.standard_normal((s.nmodes,) * 2))
    return e / np.sqrt(np.repeat(masses, 3))[None, :]


@pytest.mark.parametrize("kind", ["3d", "3d_nmag2", "2d"])
def test_dv_longrange(kind):
    if kind == "2d":
        s = synth.tiny(dimension="2", with_wfc=False)
        q = np.array([0.21, -0.13, 0.0])
    elif kind == "3d_nmag2":
        s = synth.tiny(nspin=2, nmag=2, with_wfc=False)
        q = np.array([0.3, 0.1, -0.2])
    else:
        s = synth.config("c0", with_wfc=False)
        q = np.array([0.1, 0.25, -0.35])
    lr = lr_params(s)
    rng = np.random.default_rng(1)
    dv = rng.standard_normal((s.nmodes, s.nmag) + tuple(s.fft_dims)) * (1 + 1j)
    eig = patterns(s, rng, masses_for(s))
    oi = O.OracleInterp(s, (1, 1, 1), lr=lr)
    ctx = lelph.Context(s, "fp64", upload=False)
    say(f"case: dv_longrange {kind}")
    try:
        for e, sg in ((None, 1), (eig, -1)):
            got = ctx.dv_longrange(q, lr, dv, eig=e, sign=sg)
            ref = oi.longrange(q, dv, e, sg)
            assert rel_err(got - dv, ref - dv) < TOL64, (kind, sg)
    finally:
        ctx.close()


@pytest.mark.parametrize("nmag", [1, 4])
def test_rotate_dvscf(nmag):
    s = synth.tiny(nspinor=2 if nmag == 4 else 1, nmag=nmag, syms="cubic", with_wfc=False)
    a = s.alat[0, 0]
    C4 = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    syms = [(C4, np.zeros(3), 0), (C4, np.array([a / 2, 0, 0]), 0),
            (np.diag([1.0, -1, 1]), np.array([0, a / 2, 0]), 1), (-np.eye(3), np.zeros(3), 1)]
    rng = np.random.default_rng(2)
    dv = rng.standard_normal((s.nmodes, nmag) + tuple(s.fft_dims)) + \
        1j * rng.standard_normal((s.nmodes, nmag) + tuple(s.fft_dims))
    q = np.array([0.25, 0.5, 0.0])
    oi = O.OracleInterp(s, (1, 1, 1))
    ctx = lelph.Context(s, "fp64", upload=False)
    say(f"case: rotate_dvscf nmag={nmag}")
    try:
        for sym in syms:
            for comp in ((False, True) if nmag == 4 else (False,)):
                got = ctx.rotate_dvscf(q, dv, sym, comp)
                ref = oi.rotate(q, dv, sym, comp)
                assert rel_err(got, ref) < TOL64, (sym[2], comp)
    finally:
        ctx.close()


def coarse_grid(grid):
    g = np.stack(np.meshgrid(*[np.arange(n) / n for n in grid], indexing="ij"), -1)
    return g.reshape(-1, 3)


def _run_chain(s, grid, lr, fine, star_sym=None, seed=3):
    """the same coarse data into the device and the oracle; returns the
    device context (interpolation finished) and the oracle"""
    rng = np.random.default_rng(seed)
    masses = masses_for(s)
    oi = O.OracleInterp(s, grid, masses=masses, lr=lr)
    ctx = lelph.Context(s, "fp64", upload=bool(s.wfcs))
    ctx.interp_begin(grid, masses=masses, lr=lr)
    qs = coarse_grid(grid)
    done = set()
    for q in qs:
        sl = oi.slot(q)
        if sl in done:
            continue
        dv = rng.standard_normal(oi.shape) + 1j * rng.standard_normal(oi.shape)
        e = patterns(s, rng, masses)
        ctx.interp_add_coarse(q, e, dv)
        oi.add_coarse(q, e, dv)
        done.add(sl)
        if star_sym is not None:
            # S q (crystal) as a star member, if it is a new grid point
            R = np.asarray(star_sym[0])
            qc = s.blat @ q
            Sq = np.linalg.solve(s.blat, R @ qc)
            if star_sym[2]:
                Sq = -Sq
            ss = oi.slot(Sq)
            if ss not in done:
                eS = patterns(s, rng, masses)
                ctx.interp_add_star(star_sym, Sq, eS)
                oi.add_star(star_sym, Sq, eS)
                done.add(ss)
    ctx.interp_finish()
    oi.finish()
    return ctx, oi


@pytest.mark.parametrize("case", ["c0_lr", "c0_nolr", "slab_lr", "star"])
def test_interpolation_chain(case):
    """interp_dvscf at fine q (off the coarse grid) against the oracle chain"""
    rng = np.random.default_rng(9)
    if case.startswith("c0"):
        s = synth.config("c0", with_wfc=False)
        grid = (3, 3, 3)
        lr = lr_params(s, nmags=(1, 0, 0, 0)) if case == "c0_lr" else None
        fine = rng.uniform(-0.6, 0.6, (7, 3))
        star = None
    elif case == "slab_lr":
        s = synth.tiny(dimension="2", with_wfc=False)
        grid = (3, 3, 1)
        lr = lr_params(s, nmags=(1, 0, 0, 0))
        fine = np.c_[rng.uniform(-0.5, 0.5, (5, 2)), np.zeros(5)]
        star = None
    else:
        s = synth.tiny(syms="cubic", with_wfc=False)
        grid = (4, 4, 2)
        lr = lr_params(s, quad=False, nmags=(1, 0, 0, 0))
        fine = rng.uniform(-0.5, 0.5, (5, 3))
        star = (np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]]), np.zeros(3), 0)
    say(f"case: interpolation chain {case}, grid {grid}")
    ctx, oi = _run_chain(s, grid, lr, fine, star)
    try:
        got = ctx.interp_dvscf(fine)
        for i, q in enumerate(fine):
            ref = oi.dvscf(q)
            assert rel_err(got[i], ref) < TOL64, (case, i)
        # on the coarse grid the chain returns what went in (to the reference's
        # own pi mismatch between q -> R and R -> q)
        q0 = coarse_grid(grid)[1]
        assert rel_err(ctx.interp_dvscf(q0)[0], oi.dvscf(q0)) < TOL64
    finally:
        ctx.close()


def test_interp_set_dvscf_gkq():
    """g(k,q) from the interpolated potential (identity patterns, the mode
    rotation with eig and the bare local part added) against the oracle"""
    s = synth.tiny(nbnds=4, kgrid=(2, 2, 2))
    grid = (3, 3, 3)
    lr = lr_params(s, quad=False, nmags=(1, 0, 0, 0))
    ctx, oi = _run_chain(s, grid, lr, None)
    say("case: g(k,q) from an interpolated dV")
    try:
        q = s.qpts[0]  # (0, 0, 0.5): on the 2x2x2 k grid, off the 3x3x3 q grid
        eig = s.eigvec(0)
        ctx.interp_set_dvscf(q, eig, kernel="dfpt")
        o = O.OracleSystem(s)
        dv = np.einsum("vj,jm...->vm...", eig, oi.dvscf(q))
        dv = np.ascontiguousarray(dv)
        vl = o.dvlocq(0)
        O.lib().orc_add_dvscf_qe(O.dptr(dv), O.dptr(vl), o.lat)
        for pr in synth.kq_pairs(s, 0)[:3]:
            g = ctx.gkq(*pr, non_loc=True)
            ref = o.add_nonlocal(0, *pr, o.elph_local(0, *pr, dv))
            assert rel_err(g, ref) < TOL64, pr
    finally:
        ctx.close()


def test_interp_errors():
    """call order and grid checks return status codes"""
    s = synth.tiny(with_wfc=False)
    ctx = lelph.Context(s, "fp64", upload=False)
    try:
        with pytest.raises(lelph.LelphError):
            ctx.interp_finish()  # not begun
        ctx.interp_begin((2, 2, 2))
        rng = np.random.default_rng(0)
        dv = np.zeros((s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        e = np.eye(s.nmodes, dtype=np.complex128)
        with pytest.raises(lelph.LelphError):
            ctx.interp_add_coarse([0.3, 0.0, 0.0], e, dv)  # not a grid point
        ctx.interp_add_coarse([0.5, 0.0, 0.0], e, dv)
        with pytest.raises(lelph.LelphError):
            ctx.interp_add_coarse([-0.5, 0.0, 1.0], e, dv)  # same point twice
        with pytest.raises(lelph.LelphError):
            ctx.interp_finish()  # grid incomplete
        with pytest.raises(lelph.LelphError):
            ctx.interp_dvscf(rng.uniform(size=(1, 3)))  # not finished
    finally:
        ctx.close()


def test_dvscf_qe_files(tmp_path):
    """read_dvscf_qe / write_dvscf_qe / set_dvscf_file through the device
    transpose, against the oracle"""
    s = synth.tiny(nbnds=4, nmag=2, nspin=2)
    rng = np.random.default_rng(12)
    sh = (s.nmodes, s.nmag) + tuple(s.fft_dims)
    dv = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    f1 = str(tmp_path / "dvscf1")
    np.ascontiguousarray(dv.transpose(0, 1, 4, 3, 2)).tofile(f1)
    eig = s.eigvec(0)
    pats, _ = np.linalg.qr(rng.standard_normal((s.nmodes,) * 2)
                           + 1j * rng.standard_normal((s.nmodes,) * 2))
    pats = np.ascontiguousarray(pats)
    o = O.OracleSystem(s)
    ref = np.empty(sh, np.complex128)
    O.check(O.lib().orc_read_dvscf_qe(f1.encode(), o.lat, O.dptr(np.ascontiguousarray(eig)),
                                      O.dptr(pats), O.dptr(ref)))
    ctx = lelph.Context(s, "fp64")
    say("case: QE dvscf files")
    try:
        got = ctx.read_dvscf_qe(f1, eig, pats)
        assert rel_err(got, ref) < TOL64
        f2 = str(tmp_path / "dvscf2")
        ctx.write_dvscf_qe(f2, dv)
        assert open(f1, "rb").read() == open(f2, "rb").read()
        # the file-fed potential gives the same g as the host-fed one
        ctx.set_dvscf_file(s.qpts[0], eig, f1, pats, kernel="dfpt")
        pr = synth.kq_pairs(s, 0)[1]
        g_file = ctx.gkq(*pr, non_loc=True)
        ctx.set_dvscf(s.qpts[0], eig, dv, patterns=pats, kernel="dfpt")
        g_host = ctx.gkq(*pr, non_loc=True)
        assert rel_err(g_file, g_host) < 1e-13
    finally:
        ctx.close()


@pytest.mark.parametrize("case,asr", [("3d", "simple"), ("2d", "simple"), ("3d", "crystal"),
                                      ("3d", "all"), ("3d", "all_huang"), ("2d", "all")])
def test_dyn_interpolation(case, asr):
    """dynamical matrices interpolated on the device against the oracle
    (long range with Born charges and quadrupoles; the simple sum rule, and
    the crystal / all / all_huang projections against the reference's own
    LSMR code in oracle/_ref), and the device eigensolver against LAPACK's
    eigenvalues and a reconstruction of the Hermitian part from its
    eigenvectors"""
    if asr != "simple" and not os.path.exists(O.REF_LIB):
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(17)
    if case == "3d":
        s = synth.config("c0", with_wfc=False)
        grid = (3, 3, 3)
        fine = rng.uniform(-0.6, 0.6, (9, 3))
        fine[0] = 0.0
    else:
        s = synth.tiny(dimension="2", with_wfc=False)
        grid = (3, 3, 1)
        fine = np.c_[rng.uniform(-0.5, 0.5, (6, 2)), np.zeros(6)]
    nm = s.nmodes
    masses = masses_for(s)
    lr = lr_params(s, quad=True)
    od = O.OracleDynInterp(s, grid, masses=masses, lr=lr, eta_ph=1.0, ecut=30.0, asr=asr)
    # the projections: exact here, LSMR to 1e-12 in the reference
    tol = TOL64 if asr == "simple" else 1e-8
    ctx = lelph.Context(s, "fp64", upload=False)
    say(f"case: dynamical-matrix interpolation {case} asr {asr}")
    try:
        ctx.dyn_interp_begin(grid, masses=masses, lr=lr, eta_ph=1.0, ecut=30.0, asr=asr)
        for q in coarse_grid(grid):
            e, _ = np.linalg.qr(rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm)))
            om = np.sort(rng.uniform(0.05, 1.0, nm))
            pol = e / np.sqrt(np.repeat(masses, 3))[None, :]
            ctx.dyn_interp_add_coarse(q, om, pol)
            od.add_coarse(q, om, pol)
        ctx.dyn_interp_finish()
        od.finish()
        dyn, om, eig = ctx.dyn_interp(fine)
        sq = np.sqrt(np.repeat(masses, 3))
        for i, q in enumerate(fine):
            ref = od.dyn(q)
            assert rel_err(dyn[i], ref) < tol, (case, asr, i)
            f_ref, _ = od.modes(ref)
            w = np.sign(om[i]) * om[i] ** 2
            w_ref = np.sign(f_ref) * f_ref ** 2
            assert np.abs(w - w_ref).max() < 10 * tol * np.abs(w_ref).max(), (case, i)
            V = (eig[i] * sq[None, :]).T  # columns: unit eigenvectors
            assert np.abs(V.conj().T @ V - np.eye(nm)).max() < 1e-10
            lo = np.tril(0.5 * (ref + ref.conj().T))
            h = lo + np.tril(lo, -1).conj().T
            assert rel_err(V @ np.diag(w) @ V.conj().T, h) < 1e-10, (case, i)
    finally:
        ctx.close()
