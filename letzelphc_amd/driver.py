"""The electron-phonon driver on real inputs: Yambo SAVE + QE ph_save ->
ndb.elph and ndb.Dmats, every matrix element on the device.

Follows elph_driver (src/elph/elph_driver.c:33-342) and
compute_and_write_elphq (src/elph/compute_elph_matq.c:20-234):
  1. read the SAVE and ph_save (read_and_alloc_save_data, get_data_from_qe)
  2. D matrices of every phonon symmetry at every full-BZ k
     (compute_and_write_dmats, src/elph/compute_dmats.c:20-131)
  3. per iBZ q: dynamical matrix dyn<iq>, patterns.<iq>.xml and dvscf<iq>
     (get_dvscf_dyn_qe, src/io/qe/get_dvscf_dyn_qe.c:15-107), the bare
     local part added on the device; polarisation vectors and frequencies of
     every star member (rotate_eig_vecs, src/symmetries/rotate_eig_vecs.c);
     g for all k in one batch; each star member's g from the D matrices
     (elph_q_rotate) written at (S q, S k)
  4. kpoints, qpoints, maps, symmetries, convention, kernel and the
     dielectric tensors (write_basic_data, src/io/write_basic_data.c)
"""
from __future__ import annotations

import os

import numpy as np

from . import lelph, ndb, save, synth


def rotate_eig_vecs(sym, sys_, qpt_crys, eig_q):
    """rotate_eig_vecs (src/symmetries/rotate_eig_vecs.c:14-160): phonon
    polarisation vectors of S q from those of q.  Atom a goes to the atom
    at S r_a + v (negated for time reversal); each mode's displacement is
    rotated, its e^{-i q.r_a} e^{-i Sq.v} phase removed, conjugated with a
    sign for time reversal, and e^{i Sq.r_a'} put back."""
    R, v, trev = sym
    R = np.asarray(R, np.float64)
    v = np.asarray(v, np.float64)
    blat = sys_.blat
    pos = np.asarray(sys_.atomic_pos, np.float64)
    na = sys_.natom
    crys = pos @ blat / (2 * synth.RPI)  # (natom, 3): row a = blat^T r_a / 2pi
    rot_map = np.empty(na, np.int64)
    for a in range(na):
        ra = R @ pos[a] + v
        if trev:
            ra = -ra
        rc = (blat.T @ ra) / (2 * synth.RPI)
        rot_map[a] = save.find_kidx_in_list(crys, rc)
        if rot_map[a] < 0:
            raise ValueError("Unable to find the equalivalent atom for the symmetry operation.")
    qc = blat @ np.asarray(qpt_crys, np.float64)
    Sq = R @ qc
    qphase = np.exp(-1j * (Sq @ v))
    e_Sqr = np.exp(1j * (pos @ Sq))
    e_qr = np.exp(-1j * (pos @ qc))
    nm = 3 * na
    out = np.zeros((nm, nm), np.complex128)
    eq = np.asarray(eig_q, np.complex128).reshape(nm, na, 3)
    rot = eq @ R.T  # (nm, na, 3): R e_a per mode
    rot = rot * (e_qr[None, :, None] * qphase)
    if trev:
        rot = -np.conj(rot)
    o = out.reshape(nm, na, 3)
    for a in range(na):
        o[:, rot_map[a], :] = e_Sqr[rot_map[a]] * rot[:, a, :]
    return out


def compute_dmats(ctx, sys_, ph_syms):
    """compute_and_write_dmats (src/elph/compute_dmats.c:20-131): D for
    every (phonon symmetry, full-BZ k), (nsym, nk, nspin, nb, nb)."""
    nk = len(sys_.kpt_bz_crys)
    D = np.empty((len(ph_syms), nk, sys_.nspin, sys_.nbnds, sys_.nbnds), np.complex128)
    for isym, (R, tau, trev) in enumerate(ph_syms):
        for ik in range(nk):
            D[isym, ik] = ctx.electronic_reps(R, tau, trev, ik)
    return D


def _kplusq(kbz, q):
    idx = np.empty(len(kbz), np.int64)
    for i, k in enumerate(kbz):
        idx[i] = save.find_kidx_in_list(kbz, k + q)
        if idx[i] < 0:
            raise ValueError("K+Q point cannot be found in kgrid.")
    return idx


def unfold_star(g, sys_, ph, iq, D, qpos, rotate, kminusq=False):
    """compute_elph_matq.c:152-217: for every star member S q of iBZ q iq
    and every k, g(S q, S k) = elph_q_rotate(D[S, k+q], g(q, k), D[S, k]);
    yields (qpos of S q, index of S k, g).  rotate(Dl, g, Dr, trev) is the
    rotation (device or host)."""
    q = np.asarray(ph.qpts_ibz[iq], np.float64)
    kbz = sys_.kpt_bz_crys
    kq = _kplusq(kbz, -q if kminusq else q)
    kcart = kbz @ (sys_.blat.T / (2 * synth.RPI))  # (nk, 3) cartesian, no 2 pi
    A = sys_.alat
    for istar in range(1, int(ph.nqstar[iq])):
        qs = qpos + istar
        isym = int(ph.qmap[qs, 1])
        R, _, trev = ph.ph_syms[isym]
        for i in range(len(kbz)):
            idx_k, idx_kq = i, int(kq[i])
            if kminusq:
                idx_k, idx_kq = idx_kq, idx_k
            gs = rotate(D[isym, idx_kq], g[i], D[isym, idx_k], trev)
            Sk = A.T @ (np.asarray(R) @ kcart[i])
            iSk = save.find_kidx_in_list(kbz, Sk)
            if iSk < 0:
                raise ValueError("Unable to to find S*k index")
            yield qs, iSk, gs


def run_elph(save_dir, ph_save_dir, out_dir=".", kernel="dfpt", kminusq=False,
             start_band=-1, end_band=-1, precision="fp64", non_loc=True,
             nc4_float=np.float64):
    """elph_driver: writes out_dir/ndb.elph and out_dir/ndb.Dmats; returns
    their paths.  kernel: dfpt, dfpt_local, bare or bare_local."""
    sys_, ph = save.read_save(save_dir, ph_save_dir, start_band, end_band, nc4_float)
    nq_bz, nk = len(ph.qpts_bz), len(sys_.kpt_bz_crys)
    nm = sys_.nmodes
    os.makedirs(out_dir, exist_ok=True)
    p_elph = os.path.join(out_dir, "ndb.elph")
    p_dmat = os.path.join(out_dir, "ndb.Dmats")
    ctx = lelph.Context(sys_, precision)
    try:
        D = compute_dmats(ctx, sys_, ph.ph_syms)
        ndb.write_dmats(p_dmat, D)
        w = ndb.ElphWriter(p_elph, nq_bz, nk, nm, sys_.nspin, sys_.nbnds)
        try:
            for iq in range(len(ph.qpts_ibz)):
                q = ph.qpts_ibz[iq]
                j = lambda name: os.path.join(ph_save_dir, name)
                _, omega, eig, _ = ctx.read_dyn_qe(j(f"dyn{iq + 1}"))
                if kernel.startswith("dfpt"):
                    pats = lelph.read_patterns_qe(j(f"patterns.{iq + 1}.xml"), sys_.natom)
                    ctx.set_dvscf_file(q, eig, j(f"dvscf{iq + 1}"), pats, kernel=kernel)
                else:
                    ctx.set_dvscf(q, eig, None, kernel=kernel)
                qpos = int(np.sum(ph.nqstar[:iq]))
                w.put_phonon(qpos, eig, omega)
                for istar in range(1, int(ph.nqstar[iq])):
                    sym = ph.ph_syms[int(ph.qmap[qpos + istar, 1])]
                    w.put_phonon(qpos + istar, rotate_eig_vecs(sym, sys_, q, eig), omega)
                g = ctx.compute_elphq(q, non_loc=non_loc, kminusq=kminusq)
                w.put_elph(qpos, 0, g)
                for qs, iSk, gs in unfold_star(g, sys_, ph, iq, D, qpos, ctx.elph_q_rotate,
                                               kminusq):
                    w.put_elph(qs, iSk, gs[None])
            w.write_basic_data(sys_.kpt_bz_crys, ph.qpts_bz, ph.qpts_ibz,
                               np.asarray(ph.qmap, np.int32), np.asarray(sys_.kmap, np.int32),
                               list(sys_.band_window),
                               ph.ph_syms, kernel=kernel,
                               convention="yambo" if kminusq else "standard",
                               epsilon=ph.epsilon, Zborn=ph.Zborn, Qpole=ph.Qpole,
                               natom=sys_.natom)
        finally:
            w.close()
    finally:
        ctx.close()
    return p_elph, p_dmat


# ---------------------------------------------------------------------------
# the input file (elph.in) and the command line
# ---------------------------------------------------------------------------
KERNELS = ("dfpt", "dfpt_local", "bare", "bare_local")


def _strip_quotes(v: str) -> str:
    """strip_quotes_in_string (src/common/string_func.c:329): a value in
    matching single or double quotes loses them and the blanks inside"""
    if len(v) > 1 and v[0] in "\"'" and v[-1] == v[0]:
        return v[1:-1].strip()
    return v


def read_elph_input(path) -> dict:
    """read_elph_input_file (src/parser/elph_input_parser.c:22-161): INI
    lines `name = value` (';' and '#' comments, sections ignored), the keys
    and defaults of the reference: nkpool 1, nqpool 1, start_bnd 0,
    end_bnd 0 (0: every band), save_dir SAVE, ph_save_dir ph_save, kernel
    dfpt (lower-cased), convention standard (k+q) or yambo (k-q).  An empty
    value, an unknown key or another convention is an error, as there."""
    out = {"nkpool": 1, "nqpool": 1, "start_bnd": 0, "end_bnd": 0, "save_dir": "SAVE",
           "ph_save_dir": "ph_save", "kernel": "dfpt", "kminusq": False}
    with open(path) as f:
        for ln, raw in enumerate(f, 1):
            line = raw.strip()
            if not line or line[0] in ";#" or (line[0] == "[" and line.endswith("]")):
                continue
            for c in (" ;", " #"):  # inline comments
                if c in line:
                    line = line[:line.index(c)].rstrip()
            if "=" not in line and ":" not in line:
                raise ValueError(f"{path}:{ln}: not a `name = value` line")
            sep = "=" if "=" in line else ":"
            name, value = (x.strip() for x in line.split(sep, 1))
            if not value.replace(" ", ""):
                raise ValueError(f"Invalid input for {name}")
            if name in ("nkpool", "nqpool", "start_bnd", "end_bnd"):
                out[name] = int(value)
            elif name in ("save_dir", "ph_save_dir"):
                out[name] = _strip_quotes(value)
            elif name == "kernel":
                out["kernel"] = _strip_quotes(value).lower()
            elif name == "convention":
                if "yambo" in value:
                    out["kminusq"] = True
                elif "standard" in value:
                    out["kminusq"] = False
                else:
                    raise ValueError("Invalid value for convention in the input file. "
                                     "Only yambo or standard are accepted.")
            else:
                raise ValueError("Invalid variable in the elph input file.")
    if out["kernel"] not in KERNELS:
        raise ValueError(f"unknown kernel {out['kernel']!r}; one of {KERNELS}")
    return out


def run_from_input(path, out_dir=".", precision="fp64"):
    """elph_driver from an elph.in: paths relative to the input file's
    directory are taken as given (the reference reads them relative to the
    working directory); nkpool / nqpool are pools of MPI ranks there and map
    to GPUs here (one GPU per process), so a single process ignores them."""
    inp = read_elph_input(path)
    sb = inp["start_bnd"] if inp["start_bnd"] > 0 else -1
    eb = inp["end_bnd"] if inp["end_bnd"] > 0 else -1
    return run_elph(inp["save_dir"], inp["ph_save_dir"], out_dir, kernel=inp["kernel"],
                    kminusq=inp["kminusq"], start_band=sb, end_band=eb, precision=precision)


def main(argv=None):
    """`python -m letzelphc_amd.driver -F elph.in` (the reference's `lelphc
    -F elph.in`, src/preprocessor/cli.c:14-80); --precision picks the
    contraction precision of this library."""
    import argparse
    ap = argparse.ArgumentParser(prog="letzelphc_amd.driver")
    ap.add_argument("-F", dest="input", required=True, help="elph input file")
    ap.add_argument("--out", default=".", help="directory of ndb.elph and ndb.Dmats")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32", "bf16x3", "bf16x6"))
    args = ap.parse_args(argv)
    p_elph, p_dmat = run_from_input(args.input, args.out, args.precision)
    print(p_elph)
    print(p_dmat)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
