"""QE phonon files: patterns.N.xml (read_pattern_qe) and dyn files
(write_dyn_qe / read_dyn_qe, old text format), host parsing on the CPU and
the device diagonalisation of read_dyn_qe against LAPACK on the GPU."""
import os
import re

import numpy as np
import pytest

from letzelphc_amd import lelph, synth

from .util import rel_err, say


def _patterns_xml(pats, nirr_split=(2, 4)):
    """a patterns file in QE's layout: IRREPS_INFO / REPRESENTION.i /
    PERTURBATION.j / DISPLACEMENT_PATTERN with (re, im) pairs"""
    nm = len(pats)
    out = ["<?xml version=\"1.0\"?>", "<Root>", "<IRREPS_INFO>",
           f"<NUMBER_IRR_REP type=\"integer\" size=\"1\">{len(nirr_split)}</NUMBER_IRR_REP>"]
    m = 0
    for r, npert in enumerate(nirr_split):
        out.append(f"<REPRESENTION.{r + 1}>")
        out.append(f"<NUMBER_OF_PERTURBATIONS type=\"integer\" size=\"1\">{npert}"
                   "</NUMBER_OF_PERTURBATIONS>")
        for p in range(npert):
            vals = "\n".join(f"{z.real:.16E},{z.imag:.16E}" for z in pats[m])
            out.append(f"<PERTURBATION.{p + 1}>")
            out.append(f"<DISPLACEMENT_PATTERN type=\"complex\" size=\"{nm}\">\n{vals}\n"
                       "</DISPLACEMENT_PATTERN>")
            out.append(f"</PERTURBATION.{p + 1}>")
            m += 1
        out.append(f"</REPRESENTION.{r + 1}>")
    out += ["</IRREPS_INFO>", "</Root>"]
    assert m == nm
    return "\n".join(out)


def test_read_patterns(tmp_path):
    say("case: patterns.xml parsing")
    rng = np.random.default_rng(0)
    pats, _ = np.linalg.qr(rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)))
    f = tmp_path / "patterns.1.xml"
    f.write_text(_patterns_xml(pats))
    got = lelph.read_patterns_qe(str(f), 2)
    assert np.abs(got - pats).max() < 1e-15
    g = tmp_path / "patterns.2.xml"
    g.write_text("<?xml version=\"1.0\"?>\n<DVSCF_IS_IN_CARTESIAN_BASIS/>\n")
    assert np.array_equal(lelph.read_patterns_qe(str(g), 2), np.eye(6))
    bad = tmp_path / "patterns.3.xml"
    bad.write_text(_patterns_xml(pats).replace("<NUMBER_IRR_REP type=\"integer\" size=\"1\">2",
                                               "<NUMBER_IRR_REP type=\"integer\" size=\"1\">1"))
    with pytest.raises(lelph.LelphError):
        lelph.read_patterns_qe(str(bad), 2)


def parse_dyn(path, natom):
    """an independent reader of the text written by write_dyn_qe"""
    lines = open(path).read().splitlines()
    masses = [float(l.split()[-1]) for l in lines[3:3 + natom]]
    qi = next(i for i, l in enumerate(lines) if l.strip().startswith("q = ("))
    q = [float(x) for x in re.findall(r"[-+]?\d+\.\d+", lines[qi])]
    nm = 3 * natom
    D = np.zeros((nm, nm), np.complex128)
    i = qi + 2
    for a in range(natom):
        for b in range(natom):
            assert [int(x) for x in lines[i].split()] == [a + 1, b + 1]
            for x in range(3):
                v = [float(t) for t in lines[i + 1 + x].split()]
                for y in range(3):
                    D[3 * a + x, 3 * b + y] = v[2 * y] + 1j * v[2 * y + 1]
            i += 4
    return np.array(q), np.array(masses), D


def test_write_dyn_format(tmp_path):
    say("case: dyn file writer")
    rng = np.random.default_rng(1)
    masses = np.array([12.0107, 15.999])
    dyn = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    q = np.array([0.125, -0.25, 0.5])
    f = str(tmp_path / "dyn1")
    lelph.write_dyn_qe(f, q, dyn, masses)
    q2, m2, D = parse_dyn(f, 2)
    assert np.allclose(q2, q, atol=1e-9) and np.allclose(m2, masses, atol=1e-6)
    ms = np.sqrt(np.outer(np.repeat(masses, 3), np.repeat(masses, 3)))
    assert np.abs(D - dyn * ms).max() <= 5e-9 * max(1.0, np.abs(dyn * ms).max())


@pytest.mark.gpu
def test_read_dyn_gpu(tmp_path):
    """read_dyn_qe: the file's matrix / sqrt(Ma Mb), its Hermitian part
    diagonalised on the device, against LAPACK (eigenvalues, and the
    reconstruction from the returned eigenvectors)"""
    say("case: read_dyn_qe with the device eigensolver")
    s = synth.tiny(with_wfc=False)
    rng = np.random.default_rng(2)
    masses = np.array([12.0107, 15.999])
    A = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    dyn = (A + A.conj().T) * 1e-3
    f = str(tmp_path / "dyn1")
    lelph.write_dyn_qe(f, np.array([0.0, 0.5, 0.0]), dyn, masses)
    _, mf, D = parse_dyn(f, 2)
    Dn = D / np.sqrt(np.outer(np.repeat(mf, 3), np.repeat(mf, 3)))
    H = 0.5 * (Dn + Dn.conj().T)
    w_ref = np.linalg.eigvalsh(H)
    ctx = lelph.Context(s, "fp64", upload=False)
    try:
        q, om, pv, am = ctx.read_dyn_qe(f)
    finally:
        ctx.close()
    assert np.allclose(q, [0.0, 0.5, 0.0]) and np.allclose(am, mf)
    w = np.sign(om) * om ** 2
    assert np.abs(w - w_ref).max() < 1e-12 * np.abs(w_ref).max()
    V = (pv * np.sqrt(np.repeat(am, 3))[None, :]).T
    assert np.abs(V.conj().T @ V - np.eye(6)).max() < 1e-12
    assert rel_err(V @ np.diag(w) @ V.conj().T, H) < 1e-12


def write_dyn_xml(path, q, D, type_masses_amu, atom_types):
    """QE's XML dyn layout as read_dyn_xml reads it (src/io/qe/read_dyn.c:
    306-548): PHI.a.b with the first cartesian index fastest"""
    na = len(atom_types)
    L = ['<?xml version="1.0"?>', "<Root>", "  <GEOMETRY_INFO>",
         f"    <NUMBER_OF_TYPES>{len(type_masses_amu)}</NUMBER_OF_TYPES>",
         f"    <NUMBER_OF_ATOMS>{na}</NUMBER_OF_ATOMS>"]
    for t, m in enumerate(type_masses_amu):
        L.append(f"    <MASS.{t + 1}>{m:.17e}</MASS.{t + 1}>")
    for a, t in enumerate(atom_types):
        L.append(f'    <ATOM.{a + 1} SPECIES="X{t}" INDEX="{t + 1}" TAU="0 0 0"/>')
    L += ["  </GEOMETRY_INFO>", "  <DYNAMICAL_MAT_.1>",
          f"    <Q_POINT>{q[0]:.17e} {q[1]:.17e} {q[2]:.17e}</Q_POINT>"]
    for a in range(na):
        for b in range(na):
            blk = D[3 * a:3 * a + 3, 3 * b:3 * b + 3]
            vals = []
            for y in range(3):
                for x in range(3):
                    vals += [blk[x, y].real, blk[x, y].imag]
            L.append(f"    <PHI.{a + 1}.{b + 1}>" + " ".join(f"{v:.17e}" for v in vals)
                     + f"</PHI.{a + 1}.{b + 1}>")
    L += ["  </DYNAMICAL_MAT_.1>", "  <DYNAMICAL_MAT_.2>", "    <Q_POINT>9 9 9</Q_POINT>",
          "  </DYNAMICAL_MAT_.2>", "</Root>"]
    open(path, "w").write("\n".join(L) + "\n")


def test_parse_dyn_xml(tmp_path):
    """the XML dyn format gives the file's matrix / sqrt(Ma Mb) bit for bit,
    masses in amu times 911.444243096, the first dynamical matrix only;
    the text format of the same matrix agrees to its printed digits"""
    say("case: XML dyn file parsing")
    import ctypes as C
    rng = np.random.default_rng(4)
    A = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    D = (A + A.conj().T) * 1e-3
    amu = [12.0107, 15.999]
    f = tmp_path / "dyn1.xml"
    write_dyn_xml(f, [0.0, 0.5, 0.25], D, amu, [1, 0])
    L = C.CDLL(lelph.LIB_PATH)
    q = np.zeros(3)
    m = np.zeros(2)
    d = np.zeros((6, 6), np.complex128)
    dp = C.POINTER(C.c_double)
    assert L.lelph_qe_parse_dyn(os.fsencode(f), 2, q.ctypes.data_as(dp), m.ctypes.data_as(dp),
                                d.ctypes.data_as(dp)) == 0
    mass = np.array([amu[1], amu[0]]) * 911.444243096
    assert np.array_equal(q, [0.0, 0.5, 0.25]) and np.array_equal(m, mass)
    ms = np.sqrt(np.outer(np.repeat(mass, 3), np.repeat(mass, 3)))
    assert np.allclose(d, D / ms, rtol=0, atol=1e-18)
    g = tmp_path / "dyn1"
    lelph.write_dyn_qe(str(g), np.array([0.0, 0.5, 0.25]), D / ms, mass)
    d2 = np.zeros_like(d)
    assert L.lelph_qe_parse_dyn(os.fsencode(g), 2, q.ctypes.data_as(dp), m.ctypes.data_as(dp),
                                d2.ctypes.data_as(dp)) == 0
    # the text form prints D sqrt(Ma Mb) with 8 decimals
    assert np.abs(d2 - d).max() <= 1e-8 / ms.min()
