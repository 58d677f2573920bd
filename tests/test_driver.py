"""Host parts of the real-input driver (letzelphc_amd/driver.py): the
phonon eigenvector rotation q -> S q (rotate_eig_vecs) against the group
structure, and the star-unfolding loop's indices against the k grid."""
import numpy as np
import pytest

from letzelphc_amd import driver, save, synth

from .test_save import _write_both
from .util import say


@pytest.fixture(scope="module")
def cubic(tmp_path_factory):
    d = tmp_path_factory.mktemp("cubic")
    s = synth.tiny(syms="cubic", q=(0.0, 0.5, 0.0))
    _write_both(d, s)
    return save.read_save(d / "SAVE", d / "ph_save")


def _sq(sys_, R, q):
    B = sys_.blat / (2 * synth.RPI)
    return sys_.alat.T @ (np.asarray(R) @ (B @ q))


def test_rotate_eig_vecs_identity_and_inverse(cubic):
    r, ph = cubic
    rng = np.random.default_rng(5)
    nm = r.nmodes
    e, _ = np.linalg.qr(rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm)))
    q = np.array([0.25, -0.5, 0.125])
    # the identity leaves the vectors alone
    I = [k for k, (R, v, tr) in enumerate(ph.ph_syms)
         if np.allclose(R, np.eye(3)) and not tr][0]
    assert np.allclose(driver.rotate_eig_vecs(ph.ph_syms[I], r, q, e), e, atol=1e-14)
    for (R, v, tr) in ph.ph_syms:
        es = driver.rotate_eig_vecs((R, v, tr), r, q, e)
        # unitary: rows stay orthonormal
        assert np.allclose(es @ es.conj().T, np.eye(nm), atol=1e-12)
        # S^-1 (R^T, -R^T v, same time reversal) brings S q's vectors back
        Ri = np.asarray(R).T
        back = driver.rotate_eig_vecs((Ri, -Ri @ np.asarray(v), tr), r, _sq(r, R, q), es)
        assert np.allclose(back, e, atol=1e-12)


def test_unfold_star_indices(cubic):
    """every (S q, S k) slot of the star is written exactly once, from the
    D matrices of the right symmetry and k indices"""
    r, ph = cubic
    nk = len(r.kpt_bz_crys)
    nm, ns, nb = r.nmodes, r.nspin, r.nbnds
    g = np.zeros((nk, nm, ns, nb, nb), np.complex128)
    g[:, 0, 0, 0, 0] = np.arange(nk)  # tag k in g
    D = np.zeros((len(ph.ph_syms), nk, ns, nb, nb), np.complex128)
    seen = set()
    for iq in range(len(ph.qpts_ibz)):
        qpos = int(np.sum(ph.nqstar[:iq]))
        calls = []

        def rot(Dl, gk, Dr, trev):
            calls.append(1)
            return gk

        for qs, iSk, gs in driver.unfold_star(g, r, ph, iq, D, qpos, rot):
            assert (qs, iSk) not in seen
            seen.add((qs, iSk))
            k = int(gs[0, 0, 0, 0].real)
            R = ph.ph_syms[int(ph.qmap[qs, 1])][0]
            kc = r.kpt_bz_crys[k] @ (r.blat.T / (2 * synth.RPI))
            assert save.find_kidx_in_list(r.kpt_bz_crys, r.alat.T @ (R @ kc)) == iSk
            # S q of the map is the star member's q
            d = _sq(r, R, ph.qpts_ibz[iq]) - ph.qpts_bz[qs]
            assert np.allclose(d, np.rint(d), atol=1e-9)
        assert len(calls) == (int(ph.nqstar[iq]) - 1) * nk
    assert len(seen) == (len(ph.qpts_bz) - len(ph.qpts_ibz)) * nk


def test_read_elph_input(tmp_path):
    """the reference's elph.in keys, defaults, quotes and errors
    (src/parser/elph_input_parser.c:22-161)"""
    from letzelphc_amd import driver
    p = tmp_path / "elph.in"
    p.write_text("# comment\n[elph]\nnkpool = 2\nstart_bnd = 3\nend_bnd=10\n"
                 "save_dir = \" ../my SAVE \"\nph_save_dir='phs'\nkernel = DFPT_Local\n"
                 "convention = yambo\n")
    d = driver.read_elph_input(p)
    say("case: elph.in parsing")
    assert d == {"nkpool": 2, "nqpool": 1, "start_bnd": 3, "end_bnd": 10,
                 "save_dir": "../my SAVE", "ph_save_dir": "phs", "kernel": "dfpt_local",
                 "kminusq": True}
    p.write_text("")
    assert driver.read_elph_input(p)["save_dir"] == "SAVE"
    for bad in ("convention = other\n", "unknown = 1\n", "kernel =   \n"):
        p.write_text(bad)
        with pytest.raises(ValueError):
            driver.read_elph_input(p)
