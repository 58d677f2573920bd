"""Writes tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own helper routines, compiled from its sources into oracle/_ref/liblelph_ref.so
(`make -C oracle ref`, COMPILE_ELPH_DOUBLE):

  legendre, Ylm, simpson             src/common/numerical_func.c:28-192
  prepare_spline, spline_interpolate src/common/spline.c:13-139
  reciprocal_vecs                    src/common/numerical_func.c
  get_KplusQ_idxs                    src/common/numerical_func.c:471-524
  find_gvecs_idxs                    src/wfc/gsort.c:49-95
  SU2mat                             src/wfc/su2mat.c:16-117
  su2rotate                          src/wfc/su2rotate.c:13-41
  apply_trans_wfc                    src/wfc/trans_wfc.c:15-52
  Vloc_Gspace                        src/dvloc/VlocG_atom.c:10-57
  sum_VNL_KKp                        src/nonloc/sum_VNL_KKp.c:8-58

and tests/golden/interp_vectors.npz:

  build_wigner_seitz_vectors         src/interpolation/wigner_seitz.c:27-313
  dVlong_range_kernel                src/dvloc/dVlong_range_kernel.c:31-458
  multiply_eikr                      src/dvloc/multiply_eikr.c:14-64
  rotate_dvscf                       src/symmetries/rotate_dvscf.c:22-368
  alloc_and_Compute_f_Coeff          src/nonloc/fcoeff.c:24-143

and tests/golden/asr_vectors.npz:

  apply_acoustic_sum_rule_fc         src/phonon/asr.c:269-512 (with lsmr_solver.c)
  apply_acoustic_sum_rule_born_charges src/phonon/asr.c:155-267
  quadrupole_3d_to_2d                src/phonon/quadrupole_3Dto2D.c:18-117

Run:  python tests/golden/make_ref_vectors.py
The fixture is data (seeded inputs, the reference's outputs); the tests
compare the oracle's restatement against it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "liblelph_ref.so")
OUT = os.path.join(HERE, "ref_vectors.npz")

dp = C.POINTER(C.c_double)
ll = C.c_longlong


def P(a):
    return a.ctypes.data_as(dp)


def load_ref():
    L = C.CDLL(REF_LIB)
    L.legendre.restype = C.c_double
    L.legendre.argtypes = [C.c_int, C.c_int, C.c_double]
    L.Ylm.restype = C.c_double
    L.Ylm.argtypes = [C.c_int, C.c_int, dp]
    L.simpson.restype = C.c_double
    L.simpson.argtypes = [dp, dp, ll]
    L.prepare_spline.argtypes = [ll, dp, dp, dp]
    L.spline_interpolate.restype = C.c_double
    L.spline_interpolate.argtypes = [C.c_double, ll, dp, dp, dp]
    L.reciprocal_vecs.argtypes = [dp, dp]
    L.get_KplusQ_idxs.argtypes = [ll, dp, dp, C.POINTER(C.c_int)]
    L.SU2mat.argtypes = [dp, ll, C.c_bool, C.c_bool, dp]
    L.su2rotate.argtypes = [C.c_int, ll, ll, dp, dp]
    L.apply_trans_wfc.argtypes = [dp, dp, ll, ll, dp, dp, C.c_bool]
    L.Vloc_Gspace.restype = C.c_double
    L.Vloc_Gspace.argtypes = [dp, C.c_char, C.c_double, ll, dp, dp, dp, C.c_double,
                              C.c_double, C.c_double]
    L.sum_VNL_KKp.argtypes = [dp, dp, dp, ll, ll, ll, dp]
    L.find_gvecs_idxs.argtypes = [ll, dp, ll, dp, C.POINTER(ll)]
    return L


def rotations(rng, n):
    """n random proper rotations and their improper partners."""
    out = []
    for _ in range(n):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(q) < 0:
            q = -q
        out.append(q)
        out.append(-q)
    # special cases of the Euler-angle branches: identity, C2 about z, C4 about z
    out.append(np.eye(3))
    out.append(np.diag([-1.0, -1.0, 1.0]))
    out.append(np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]]))
    out.append(np.diag([1.0, -1.0, -1.0]))
    return np.array(out)


def make(L, seed=2024):
    rng = np.random.default_rng(seed)
    f = {}
    # legendre / Ylm
    lm = [(l, m) for l in range(5) for m in range(-l, l + 1)]
    xs = np.concatenate([rng.uniform(-1, 1, 9), [-1.0, 0.0, 1.0]])
    f["leg_lm"] = np.array(lm, np.int64)
    f["leg_x"] = xs
    f["leg_out"] = np.array([[L.legendre(l, m, x) for x in xs] for l, m in lm])
    vecs = np.concatenate([rng.standard_normal((10, 3)),
                           [[0, 0, 0], [0, 0, 1.5], [0, 0, -2.0], [1e-7, 0, 0],
                            [0.3, 0, 0], [0, -0.4, 0]]])
    ylm = [(l, m) for l in range(4) for m in range(-l, l + 1)]
    f["ylm_lm"] = np.array(ylm, np.int64)
    f["ylm_v"] = vecs
    out = np.zeros((len(ylm), len(vecs)))
    for i, (l, m) in enumerate(ylm):
        for j, v in enumerate(vecs):
            vv = np.ascontiguousarray(v, np.float64)
            out[i, j] = L.Ylm(l, m, P(vv))
    f["ylm_out"] = out
    # simpson on odd and even point counts
    for n in (5, 6, 101, 1000):
        fv = rng.standard_normal(n)
        dx = rng.uniform(0.01, 0.02, n)
        f[f"simp{n}_f"] = fv
        f[f"simp{n}_dx"] = dx
        f[f"simp{n}_out"] = np.array(L.simpson(P(fv), P(dx), n))
    # spline on a uniform table (as the V_loc(|G|) table)
    n = 40
    x = np.linspace(0.0, 3.9, n)
    y = np.sin(1.3 * x) * np.exp(-0.2 * x) + 0.1 * rng.standard_normal(n)
    dy = np.zeros(n)
    L.prepare_spline(n, P(x), P(y), P(dy))
    xi = np.sort(rng.uniform(0.0, 3.89, 25))
    yi = np.array([L.spline_interpolate(t, int(np.floor(t / (x[1] - x[0]))), P(x), P(y), P(dy))
                   for t in xi])
    f.update(spl_x=x, spl_y=y, spl_dy=dy, spl_xi=xi, spl_yi=yi)
    # reciprocal vectors (lattice vectors as columns, as struct Lattice)
    lat = rng.standard_normal((3, 3)) + 4 * np.eye(3)
    blat = np.zeros((3, 3))
    L.reciprocal_vecs(P(np.ascontiguousarray(lat)), P(blat))
    f.update(rec_lat=lat, rec_blat=blat)
    # k+q indices on a shuffled 4x4x3 grid
    g = np.stack(np.meshgrid(np.arange(4) / 4, np.arange(4) / 4, np.arange(3) / 3,
                             indexing="ij"), -1).reshape(-1, 3)
    g = g[rng.permutation(len(g))] + rng.integers(-1, 2, g.shape)
    g = np.ascontiguousarray(g)
    qs = np.array([[0.25, 0.0, 1 / 3], [-0.5, 0.75, 0.0], [0.0, 0.0, 0.0]])
    idx = np.zeros((len(qs), len(g)), np.int32)
    for i, q in enumerate(qs):
        qq = np.ascontiguousarray(q)
        L.get_KplusQ_idxs(len(g), P(g), P(qq), idx[i].ctypes.data_as(C.POINTER(C.c_int)))
    f.update(kq_k=g, kq_q=qs, kq_idx=idx)
    # SU(2) matrices
    R = rotations(rng, 4)
    su = np.zeros((len(R), 2, 2, 2, 2), np.complex128)
    for i, r in enumerate(R):
        rr = np.ascontiguousarray(r)
        for inv in (0, 1):
            for tr in (0, 1):
                m = np.zeros(4, np.complex128)
                L.SU2mat(P(rr), 2, bool(inv), bool(tr), m.ctypes.data_as(dp))
                su[i, inv, tr] = m.reshape(2, 2)
    f.update(su2_R=R, su2_out=su)
    # su2rotate
    npw, nsets = 17, 3
    w = (rng.standard_normal((nsets, 2, npw)) + 1j * rng.standard_normal((nsets, 2, npw)))
    w = np.ascontiguousarray(w)
    m = np.ascontiguousarray(su[0, 0, 0].reshape(4))
    wo = w.copy()
    L.su2rotate(2, npw, nsets, m.ctypes.data_as(dp), wo.ctypes.data_as(dp))
    f.update(su2rot_su=m, su2rot_in=w, su2rot_out=wo)
    # apply_trans_wfc
    tau = rng.standard_normal(3)
    kv = rng.standard_normal(3)
    gv = np.ascontiguousarray(rng.integers(-3, 4, (npw, 3)).astype(np.float64) * 0.37)
    ws = np.ascontiguousarray(rng.standard_normal((nsets, npw)) + 1j * rng.standard_normal((nsets, npw)))
    outs = []
    for cj in (0, 1):
        t = ws.copy()
        L.apply_trans_wfc(P(tau), P(kv), nsets, npw, P(gv), t.ctypes.data_as(dp), bool(cj))
        outs.append(t)
    f.update(tr_tau=tau, tr_k=kv, tr_g=gv, tr_in=ws, tr_out=np.array(outs))
    # Vloc_Gspace on the synthetic local pseudopotential
    from letzelphc_amd import synth
    ps = synth.tiny(with_wfc=False).pseudo
    v, r, rab, Z = ps.loc[1]
    vol = 123.4
    Gs = np.array([0.0, 5e-6, 0.05, 0.7, 2.3, 9.1])
    vg = np.zeros((2, len(Gs)))
    work = np.zeros(len(r) + 1)
    v = np.ascontiguousarray(v)
    r = np.ascontiguousarray(r)
    rab = np.ascontiguousarray(rab)
    for ic, cut in enumerate((b"3", b"2")):
        for j, G in enumerate(Gs):
            vg[ic, j] = L.Vloc_Gspace(P(work), cut, G, len(r), P(v), P(r), P(rab), Z, 1.0, vol)
    f.update(vl_v=v, vl_r=r, vl_rab=rab, vl_Z=np.array(Z), vl_vol=np.array(vol), vl_G=Gs,
             vl_out=vg)
    # sum_VNL_KKp
    for nspinor in (1, 2):
        nspin, nb = 2, 3
        sh = (4, nspin, nb, nspinor)
        K = np.ascontiguousarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
        Kp = np.ascontiguousarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
        fc = np.ascontiguousarray(rng.standard_normal(nspinor * nspinor)
                                  + 1j * rng.standard_normal(nspinor * nspinor))
        o0 = np.ascontiguousarray(rng.standard_normal((3, nspin, nb, nb))
                                  + 1j * rng.standard_normal((3, nspin, nb, nb)))
        o = o0.copy()
        L.sum_VNL_KKp(K.ctypes.data_as(dp), Kp.ctypes.data_as(dp), fc.ctypes.data_as(dp),
                      nspin, nb, nspinor, o.ctypes.data_as(dp))
        f.update({f"vnl{nspinor}_K": K, f"vnl{nspinor}_Kp": Kp, f"vnl{nspinor}_f": fc,
                  f"vnl{nspinor}_in": o0, f"vnl{nspinor}_out": o})
    # G-vector matching: a shuffled integer sphere, searched with a rotated
    # copy in which some vectors are missing and some are new
    g = np.stack(np.meshgrid(*[np.arange(-3, 4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    g = g[np.sum(g * g, 1) <= 9].astype(np.float64)
    gl = np.ascontiguousarray(g[rng.permutation(len(g))])
    sr = g[:, [1, 0, 2]] * np.array([1.0, -1.0, 1.0]) + rng.choice([0, 0, 0, 1], (len(g), 1))
    sr = np.ascontiguousarray(sr[rng.permutation(len(sr))] + 1e-9 * rng.standard_normal(sr.shape))
    out = np.zeros(len(sr), np.int64)
    L.find_gvecs_idxs(len(sr), P(sr), len(gl), P(gl), out.ctypes.data_as(C.POINTER(ll)))
    f.update(gs_list=gl, gs_search=sr, gs_idx=out)
    return f


def load_interp(L):
    """Signatures of the interpolation helpers (and the flat glue of
    oracle/ref_glue.c, which only fills the reference's structs)."""
    L.rg_init.restype = C.c_int
    L.rg_ws.restype = ll
    L.rg_ws.argtypes = [C.POINTER(ll), dp, C.c_double, dp, ll, dp, ll, ll, C.POINTER(ll),
                        C.POINTER(ll)]
    L.dVlong_range_kernel.argtypes = [dp, dp, ll, dp, dp, dp, dp, ll, dp, C.c_char, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, dp]
    L.rg_multiply_eikr.argtypes = [dp, dp, C.POINTER(ll), ll, C.c_int]
    L.rg_rotate_dvscf.argtypes = [dp, dp, dp, dp, C.c_int, C.POINTER(ll), ll, ll, C.c_int, dp,
                                  dp, C.c_int, dp]
    L.rg_fcoeff.argtypes = [C.c_int, C.c_int, dp]
    L.rg_fcoeff.restype = C.c_int
    if L.rg_init() != 0:
        raise RuntimeError("MPI initialisation of the reference library failed")
    return L


def _lp(a):
    return a.ctypes.data_as(C.POINTER(ll))


WS_CASES = {
    # name: (lattice a[:, i], grid, atoms (crystal), pair form)
    "fcc": (lambda: np.array([[-5.13, 0, -5.13], [0, 5.13, 5.13], [5.13, 5.13, 0]]),
            (4, 4, 4), [[0, 0, 0], [0.25, 0.25, 0.25]], False),
    "fccdyn": (lambda: np.array([[-5.13, 0, -5.13], [0, 5.13, 5.13], [5.13, 5.13, 0]]),
               (3, 3, 3), [[0, 0, 0], [0.25, 0.25, 0.25]], True),
    "hex": (lambda: np.array([[5.9, -2.95, 0], [0, 5.9 * np.sqrt(3) / 2, 0], [0, 0, 38.0]]),
            (6, 6, 1), [[1 / 3, 2 / 3, 0.5], [2 / 3, 1 / 3, 0.58], [2 / 3, 1 / 3, 0.42]], False),
}


def make_interp(L, seed=77):
    """Wigner-Seitz vectors (src/interpolation/wigner_seitz.c:27),
    dVlong_range_kernel (src/dvloc/dVlong_range_kernel.c:31), multiply_eikr
    (src/dvloc/multiply_eikr.c:14) and rotate_dvscf
    (src/symmetries/rotate_dvscf.c:22) on seeded inputs."""
    rng = np.random.default_rng(seed)
    f = {}
    for name, (latf, grid, crys, pair) in WS_CASES.items():
        lat = np.ascontiguousarray(latf())
        pos = np.ascontiguousarray(np.asarray(crys, np.float64) @ lat.T)
        g = np.ascontiguousarray(grid, np.int64)
        na = len(pos)
        nb = int(np.prod(g)) * na * (na if pair else 1)
        cap = 64 * nb
        vecs = np.zeros((cap, 3), np.int64)
        degen = np.zeros(nb, np.int64)
        n = L.rg_ws(_lp(g), P(lat), 1e-6, P(pos), na, P(pos) if pair else None,
                    na if pair else 0, cap, _lp(vecs), _lp(degen))
        assert n > 0
        f.update({f"ws_{name}_lat": lat, f"ws_{name}_grid": g, f"ws_{name}_pos": pos,
                  f"ws_{name}_pair": np.array(int(pair)), f"ws_{name}_vecs": vecs[:n].copy(),
                  f"ws_{name}_degen": degen})
    # long-range kernel: 3-D and 2-D, each with four sets of terms
    lat3 = np.array([[-5.13, 0, -5.13], [0, 5.13, 5.13], [5.13, 5.13, 0]])
    lat2 = np.array([[5.9, -2.95, 0], [0, 5.9 * np.sqrt(3) / 2, 0], [0, 0, 38.0]])
    for dim, lat in (("3", lat3), ("2", lat2)):
        blat = 2 * np.pi * np.linalg.inv(lat).T
        m = rng.integers(-3, 4, (40, 3)).astype(np.float64)
        q = rng.uniform(-0.5, 0.5, 3)
        if dim == "2":
            # the induced 2-D terms decay as 1 - tanh(|q+G|_par L/2): keep
            # most in-plane q+G small so that they are not cut off
            m[:, 2] = rng.integers(-12, 13, 40)
            m[:30, :2] = 0
            q = np.array([rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), 0.0])
        m[0] = -np.round(q)  # q + G small but not zero
        gv = np.ascontiguousarray((m @ blat.T) / (2 * np.pi))
        qc = np.ascontiguousarray(blat @ q / (2 * np.pi))
        na = 2
        pos = np.ascontiguousarray(rng.uniform(0, 1, (na, 3)) @ lat.T)
        A = rng.standard_normal((3, 3))
        eps = np.ascontiguousarray(4.0 * np.eye(3) + 0.3 * (A + A.T))
        if dim == "2":
            eps[0, 2] = eps[2, 0] = eps[1, 2] = eps[2, 1] = 0.0
        zb = np.ascontiguousarray(rng.standard_normal((na, 3, 3)))
        qp = np.ascontiguousarray(rng.standard_normal((na, 3, 3, 3)))
        zv = np.ascontiguousarray(rng.uniform(1, 6, na))
        vol = abs(np.linalg.det(lat))
        outs = []
        for case in range(4):
            args = {0: (zv, None, None, None), 1: (None, eps, zb, None),
                    2: (None, eps, zb, qp), 3: (zv, eps, zb, qp)}[case]
            out = np.zeros((len(gv), na, 3), np.complex128)
            L.dVlong_range_kernel(P(qc), P(gv), len(gv),
                                  *(None if a is None else P(a) for a in args), na, P(pos),
                                  dim.encode(), vol, lat[2, 2], 30.0, 1.0, 1.3,
                                  out.ctypes.data_as(dp))
            outs.append(out)
        f.update({f"lr{dim}_q": qc, f"lr{dim}_g": gv, f"lr{dim}_pos": pos, f"lr{dim}_eps": eps,
                  f"lr{dim}_zb": zb, f"lr{dim}_qp": qp, f"lr{dim}_zv": zv,
                  f"lr{dim}_vol": np.array(vol), f"lr{dim}_zlat": np.array(lat[2, 2]),
                  f"lr{dim}_out": np.array(outs)})
    # multiply_eikr
    dims = np.array([4, 5, 6], np.int64)
    pot = np.ascontiguousarray(rng.standard_normal((2, 4, 5, 6)) + 1j * rng.standard_normal((2, 4, 5, 6)))
    q = np.ascontiguousarray(rng.uniform(-1, 1, 3))
    outs = []
    for sg in (1, -1):
        t = pot.copy()
        L.rg_multiply_eikr(t.ctypes.data_as(dp), P(q), _lp(dims), 2, sg)
        outs.append(t)
    f.update(ek_pot=pot, ek_q=q, ek_out=np.array(outs))
    # rotate_dvscf on a 6^3 cubic grid: C4z with a half translation, and a
    # time-reversed mirror; nmag 1 and 4 (composite and not)
    a = 8.0
    alat = np.ascontiguousarray(a * np.eye(3))
    blat = np.ascontiguousarray(2 * 3.1415927 * np.linalg.inv(alat).T)
    dims = np.array([6, 6, 6], np.int64)
    C4 = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    syms = [(C4, np.array([a / 2, 0, 0]), 0), (np.diag([1.0, -1, 1]), np.zeros(3), 1)]
    qr = np.ascontiguousarray([0.25, 0.5, 0.0])
    for nmag in (1, 4):
        din = np.ascontiguousarray(rng.standard_normal((2, nmag, 6, 6, 6))
                                   + 1j * rng.standard_normal((2, nmag, 6, 6, 6)))
        outs = []
        for R, tau, tr in syms:
            for comp in ((0, 1) if nmag == 4 else (0,)):
                o = np.zeros_like(din)
                L.rg_rotate_dvscf(P(qr), din.ctypes.data_as(dp), P(np.ascontiguousarray(R)),
                                  P(np.ascontiguousarray(tau)), tr, _lp(dims), 2, nmag, 2, P(alat),
                                  P(blat), comp, o.ctypes.data_as(dp))
                outs.append(o)
        f.update({f"rot{nmag}_in": din, f"rot{nmag}_out": np.array(outs)})
    # spin-orbit f coefficients of every channel up to l = 3
    for l in range(4):
        for tj in ((1,) if l == 0 else (2 * l - 1, 2 * l + 1)):
            fo = np.zeros((2 * l + 1, 2 * l + 1, 2, 2), np.complex128)
            assert L.rg_fcoeff(l, tj, fo.ctypes.data_as(dp)) == 0
            f[f"fc_l{l}_j{tj}"] = fo
    f.update(rot_q=qr, rot_alat=alat, rot_blat=blat,
             rot_R=np.array([s[0] for s in syms]), rot_tau=np.array([s[1] for s in syms]),
             rot_tr=np.array([s[2] for s in syms]))
    return f


OUT_INTERP = os.path.join(HERE, "interp_vectors.npz")

ASR_CASES = {
    # name: (lattice a[:, i], grid, atoms (crystal))
    "fcc2": (np.array([[-5.13, 0, -5.13], [0, 5.13, 5.13], [5.13, 5.13, 0]]), (2, 2, 2),
             [[0, 0, 0], [0.25, 0.25, 0.25]]),
    "hex3": (np.array([[5.9, -2.95, 0], [0, 5.9 * np.sqrt(3) / 2, 0], [0, 0, 30.0]]), (3, 3, 1),
             [[1 / 3, 2 / 3, 0.5], [2 / 3, 1 / 3, 0.58], [2 / 3, 1 / 3, 0.42]]),
    "tri2": (np.array([[7.0, 0.9, 0.3], [0.4, 6.1, -0.7], [-0.2, 0.5, 8.3]]), (3, 2, 2),
             [[0.1, 0.2, 0.05], [0.55, 0.4, 0.6]]),
}


def load_asr(L):
    """apply_acoustic_sum_rule_fc / _born_charges (src/phonon/asr.c), called
    directly: their arguments are plain pointers"""
    L.apply_acoustic_sum_rule_fc.argtypes = [C.c_int, C.POINTER(ll), ll, dp, dp, dp,
                                             C.POINTER(ll), ll, C.POINTER(ll)]
    L.apply_acoustic_sum_rule_fc.restype = None
    L.apply_acoustic_sum_rule_born_charges.argtypes = [C.c_int, dp, ll]
    L.apply_acoustic_sum_rule_born_charges.restype = None
    L.rg_quad2d.argtypes = [ll, dp, dp, dp, dp]
    L.rg_quad2d.restype = None
    return L


def make_asr(L, seed=31):
    """The force-constant sum rules simple (1), crystal (2), all (3) and
    all_huang (4) on seeded complex force constants, with the reference's
    own Wigner-Seitz vectors of the atom pairs; the Born-charge rule for
    modes 1 and 2."""
    rng = np.random.default_rng(seed)
    f = {}
    for name, (lat, grid, crys) in ASR_CASES.items():
        lat = np.ascontiguousarray(lat)
        pos = np.ascontiguousarray(np.asarray(crys, np.float64) @ lat.T)
        g = np.ascontiguousarray(grid, np.int64)
        na = len(pos)
        nm = 3 * na
        ng = int(np.prod(g))
        nb = ng * na * na
        cap = 64 * nb
        vecs = np.zeros((cap, 3), np.int64)
        degen = np.zeros(nb, np.int64)
        n = L.rg_ws(_lp(g), P(lat), 1e-6, P(pos), na, P(pos), na, cap, _lp(vecs), _lp(degen))
        assert n > 0
        vecs = np.ascontiguousarray(vecs[:n])
        frc = (rng.standard_normal((ng, nm, nm)) + 0.1j * rng.standard_normal((ng, nm, nm)))
        frc = np.ascontiguousarray(frc)
        f.update({f"asr_{name}_lat": lat, f"asr_{name}_grid": g, f"asr_{name}_pos": pos,
                  f"asr_{name}_ws": vecs, f"asr_{name}_degen": degen, f"asr_{name}_frc": frc})
        for mode in (1, 2, 3, 4):
            out = frc.copy()
            L.apply_acoustic_sum_rule_fc(mode, _lp(g), na, out.ctypes.data_as(dp), P(pos),
                                         P(lat), _lp(vecs), n, _lp(degen))
            f[f"asr_{name}_out{mode}"] = out
        zb = np.ascontiguousarray(rng.standard_normal((na, 3, 3)))
        f[f"asr_{name}_zb"] = zb
        for mode in (1, 2):
            z = zb.copy()
            L.apply_acoustic_sum_rule_born_charges(mode, P(z), na)
            f[f"asr_{name}_zb{mode}"] = z
        # quadrupole_3d_to_2d (src/phonon/quadrupole_3Dto2D.c:18-117)
        A = rng.standard_normal((3, 3))
        eps = np.ascontiguousarray(5.0 * np.eye(3) + 0.4 * (A + A.T))
        Q = np.ascontiguousarray(rng.standard_normal((na, 3, 3, 3)))
        q2 = Q.copy()
        L.rg_quad2d(na, P(eps), P(pos), P(zb), P(q2))
        f.update({f"q2d_{name}_eps": eps, f"q2d_{name}_Q": Q, f"q2d_{name}_out": q2})
    return f


OUT_ASR = os.path.join(HERE, "asr_vectors.npz")


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit(f"{REF_LIB} missing: run `make -C oracle ref` first")
    L = load_ref()
    f = make(L)
    np.savez_compressed(OUT, **f)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(f)} arrays)")
    fi = make_interp(load_interp(L))
    np.savez_compressed(OUT_INTERP, **fi)
    print(f"wrote {OUT_INTERP} ({os.path.getsize(OUT_INTERP)} bytes, {len(fi)} arrays)")
    fa = make_asr(load_asr(L))
    np.savez_compressed(OUT_ASR, **fa)
    print(f"wrote {OUT_ASR} ({os.path.getsize(OUT_ASR)} bytes, {len(fa)} arrays)")


if __name__ == "__main__":
    main()
