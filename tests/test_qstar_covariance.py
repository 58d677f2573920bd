"""The q-star rotation of the oracle (elph_q_rotate, src/symmetries/
elphq_expand.c:16-73, as used at src/elph/compute_elph_matq.c:162-217)
pinned by covariance: for every operation S of the group and every
full-BZ k,

    g(Sk, Sq) computed directly   ==   elph_q_rotate(D(S, k+q), g(k, q), D(S, k))

where the direct side contracts the potential rotated by rotate_dvscf (the
oracle's restatement, pinned against the reference's compiled
rotate_dvscf.c) with the states of Sk and Sk+Sq, and its non-local part
takes the polarisation vectors rotated to Sq.  The D matrices are the
oracle's electronic_reps (pinned by explicit G matching).

The identity is exact only when each iBZ band set is closed under the
little group of its k (a symmetric Hamiltonian's eigenstates are), so the
synthetic states are replaced here by orthonormalised orbits of random
vectors under the little group; the D matrices are then unitary and far
from the identity at high-symmetry k.  The same check with D_l and D_r
swapped, or D in place of D^dagger, misses by O(1)."""
import dataclasses

import numpy as np
import pytest

from letzelphc_amd import driver, synth
from letzelphc_amd.abi import dptr
from oracle import oracle as O

from .util import rel_err, say


def _little_group_orbits(s):
    """Replace each iBZ k's bands by an orthonormal basis of orbits of
    random vectors under the little group of k (ops with S k = k mod G,
    time-reversed ones included): c'(S G + G_S) = c(G), conjugated for
    time reversal, G_S = S k - k."""
    Sc = [synth._sym_crys(np.asarray(R), s.alat, s.blat) for (R, _, _) in s.syms]
    rng = np.random.default_rng(77)
    for ik, w in enumerate(s.wfcs):
        k = s.kpt_ibz_crys[ik]
        pos = {tuple(g): i for i, g in enumerate(w.gint)}
        ops = []
        for j, S in enumerate(Sc):
            d = S @ k - k
            if np.linalg.norm(d - np.rint(d)) < 1e-6:
                G0 = np.rint(d).astype(np.int64)
                Gr = np.rint(w.gint @ S.T).astype(np.int64) + G0
                perm = np.array([pos[tuple(g)] for g in Gr])
                ops.append((perm, int(s.syms[j][2])))
        L = len(ops)
        assert s.nbnds % L == 0, (ik, L)
        for sp in range(s.nspin):
            rows = []
            for _ in range(s.nbnds // L):
                v = (rng.standard_normal(w.npw) + 1j * rng.standard_normal(w.npw)) * \
                    np.abs(w.wfc[sp, 0, 0])
                for perm, tr in ops:
                    u = np.zeros(w.npw, np.complex128)
                    u[perm] = np.conj(v) if tr else v
                    rows.append(u)
            Q, R = np.linalg.qr(np.array(rows).T)
            assert np.min(np.abs(np.diag(R))) > 1e-8 * np.max(np.abs(np.diag(R)))
            w.wfc[sp, :, 0, :] = Q.T
    return s


def _kidx(kbz, k):
    d = kbz - k[None, :]
    d = d - np.rint(d)
    j = int(np.argmin(np.linalg.norm(d, axis=1)))
    assert np.linalg.norm(d[j]) < 1e-5
    return j


@pytest.fixture(scope="module")
def sym_sys():
    # both atoms on the C4 axis: every operation maps each atom onto itself,
    # not onto a lattice image (an image would bring in the truncated pi of
    # the reference's projector phases, exp(-2 i ELPH_PI K.T) != 1, a 1e-8
    # break of the covariance that belongs to the reference itself)
    s = _little_group_orbits(synth.tiny(nbnds=8, syms="cubic", q=(0.0, 0.5, 0.0),
                                        crys_pos=[[0, 0, 0], [0, 0, 0.4]]))
    o = O.OracleSystem(s)
    dv = o.dv_total(0)
    g = o.compute_elphq(0, dv, non_loc=True)
    return s, o, dv, g


def _rotate(Dl, g, Dr, tr):
    out = np.zeros_like(g)
    nm, ns, nb = g.shape[0], g.shape[1], g.shape[2]
    O.check(O.lib().orc_elph_q_rotate(dptr(np.ascontiguousarray(Dl)), dptr(g),
                                      dptr(np.ascontiguousarray(Dr)), nm, ns, nb, int(tr),
                                      dptr(out)))
    return out


@pytest.mark.parametrize("isym", range(8))
def test_qstar_rotation_covariance(sym_sys, isym):
    s, o, dv, g = sym_sys
    R, tau, tr = s.syms[isym]
    R = np.asarray(R, np.float64)
    q = np.asarray(s.qpts[0], np.float64)
    kbz = s.kpt_bz_crys
    bscale = s.blat / (2 * synth.RPI)
    # S q in crystal coordinates (R is the cartesian Rmat, -R_spatial for
    # time reversal)
    Sq = s.alat.T @ (R @ (bscale @ q))
    dvS = np.empty_like(dv)
    O.check(O.lib().orc_rotate_dvscf(dptr(q), dptr(dv), C_sym(R, tau, tr), o.lat, 0, dptr(dvS)))
    eigS = driver.rotate_eig_vecs((R, tau, tr), s, q, s.eigvec(0))
    s2 = dataclasses.replace(s, qpts=[Sq], _eig={0: eigS}, _dv={})
    gS = O.OracleSystem(s2).compute_elphq(0, dvS, non_loc=True)
    say(f"case: q-star covariance, op {isym} (time reversal {tr}), Sq = {np.round(Sq, 3)}")
    worst, worst_swap, worst_dag = 0.0, np.inf, np.inf
    n_nontrivial = 0
    for ik in range(len(kbz)):
        ikq = _kidx(kbz, kbz[ik] + q)
        iSk = _kidx(kbz, s.alat.T @ (R @ (bscale @ kbz[ik])))
        Dl = o.electronic_reps(R, tau, tr, ikq)
        Dr = o.electronic_reps(R, tau, tr, ik)
        # D is unitary on these band sets
        for D in (Dl, Dr):
            assert np.max(np.abs(D[0] @ D[0].conj().T - np.eye(s.nbnds))) < 1e-12
        if np.max(np.abs(Dl[0] - np.eye(s.nbnds))) > 0.1:
            n_nontrivial += 1
        got = _rotate(Dl, g[ik], Dr, tr)
        worst = max(worst, rel_err(got, gS[iSk]))
        # planted convention errors, where they change the matrices at all
        if not np.allclose(Dl, Dr):
            worst_swap = min(worst_swap, rel_err(_rotate(Dr, g[ik], Dl, tr), gS[iSk]))
        dag = np.ascontiguousarray(np.conj(np.swapaxes(Dl, 1, 2)))
        dagr = np.ascontiguousarray(np.conj(np.swapaxes(Dr, 1, 2)))
        if not (np.allclose(dag, Dl) and np.allclose(dagr, Dr)):
            worst_dag = min(worst_dag, rel_err(_rotate(dag, g[ik], dagr, tr), gS[iSk]))
    say(f"  max |d|/|g|max = {worst:.3e}; swapped D {worst_swap:.3e}, D^+ {worst_dag:.3e}; "
        f"{n_nontrivial} k with a non-trivial D(k+q)")
    assert worst < 1e-10
    assert worst_swap > 1e-2 and worst_dag > 1e-2
    if isym > 0:
        assert n_nontrivial > 0 and np.isfinite(worst_swap)
        # D^+ = D^-1 = D for an operation of order two: only the others can
        # tell D from D^+
        if not np.allclose(R @ R, np.eye(3)):
            assert np.isfinite(worst_dag)


def C_sym(R, tau, tr):
    import ctypes
    return ctypes.byref(O.make_sym(R, tau, tr))


def test_covariance_needs_nontrivial_d(sym_sys):
    """The orbit bands give D matrices far from the identity (so the
    covariance above is not satisfied trivially)."""
    s, o, _, _ = sym_sys
    big = 0
    for R, tau, tr in s.syms[1:]:
        for ik in range(len(s.kpt_bz_crys)):
            D = o.electronic_reps(R, tau, tr, ik)
            big = max(big, float(np.max(np.abs(D[0] - np.eye(s.nbnds)))))
    assert big > 0.5
