"""ndb.elph / ndb.Dmats written through liblelph_nc4.so (netCDF-4 layout on
HDF5), read back with HDF5's own h5dump: variable and dimension names,
shapes, types, chunking, the dimension-scale attachments a netCDF reader
resolves, and the data bit for bit."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from letzelphc_amd import ndb

from .util import say

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP) or not os.path.exists(ndb.LIB_PATH),
                                reason="h5dump or liblelph_nc4.so missing")


def h5(*args):
    return subprocess.run([H5DUMP, *args], check=True, capture_output=True, text=True).stdout


def header(path):
    """{dataset: (dims, type, layout)} from h5dump -H -p"""
    out = {}
    for m in re.finditer(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+).*?DATASPACE\s+SIMPLE \{ \( '
                         r'([0-9, ]+) \).*?STORAGE_LAYOUT \{\s*(\w+)( \( ([0-9, ]+) \))?',
                         h5("-H", "-p", path), re.S):
        dims = tuple(int(x) for x in m.group(3).split(","))
        chunk = tuple(int(x) for x in m.group(6).split(",")) if m.group(6) else None
        out[m.group(1)] = (dims, m.group(2), m.group(4), chunk)
    return out


def dim_list(path, name):
    txt = h5("-A", "-d", f"/{name}", path)
    m = re.search(r'ATTRIBUTE "DIMENSION_LIST".*?DATA \{(.*?)\n\s*\}', txt, re.S)
    return re.findall(r"DATASET \d+ /(\w+)", m.group(1))


def raw(path, name, dtype, shape, tmp):
    out = str(tmp / f"{name}.bin")
    h5("-d", f"/{name}", "-b", "LE", "-o", out, path)
    return np.fromfile(out, dtype).reshape(shape)


def test_ndb_elph_layout_and_data(tmp_path):
    say("case: ndb.elph written and read back with h5dump")
    rng = np.random.default_rng(0)
    nq, nk, nm, ns, nb = 3, 8, 6, 1, 4
    path = str(tmp_path / "ndb.elph")
    w = ndb.ElphWriter(path, nq, nk, nm, ns, nb)
    g = rng.standard_normal((nq, nk, nm, ns, nb, nb)) + 1j * rng.standard_normal((nq, nk, nm, ns, nb, nb))
    eig = rng.standard_normal((nq, nm, nm)) + 1j * rng.standard_normal((nq, nm, nm))
    fr = rng.uniform(0, 1, (nq, nm))
    for iq in range(nq):
        w.put_phonon(iq, eig[iq], fr[iq])
        w.put_elph(iq, 0, g[iq, :5])
        w.put_elph(iq, 5, g[iq, 5:])
    kp = rng.uniform(size=(nk, 3))
    qp = rng.uniform(size=(nq, 3))
    syms = [(np.eye(3), np.zeros(3), 0), (-np.eye(3), np.zeros(3), 1)]
    w.write_basic_data(kp, qp, qp[:1], np.zeros((nq, 2), np.int32), np.zeros((nk, 2), np.int32),
                       [1, nb], syms, kernel="dfpt", convention="standard")
    w.close()
    hd = header(path)
    exp = {"elph_mat": (nq, nk, nm, ns, nb, nb, 2), "POLARIZATION_VECTORS": (nq, nm, nm // 3, 3, 2),
           "FREQ": (nq, nm), "kpoints": (nk, 3), "qpoints": (nq, 3), "qpoints_iBZ": (1, 3),
           "qmap": (nq, 2), "kmap": (nk, 2), "bands": (2,), "number_of_phonon_symmetries": (1,),
           "time_reversal_phonon": (1,), "symmetry_matrices": (2, 3, 3),
           "fractional_translation": (2, 3), "kernel": (5,), "convention": (9,),
           "epsilon": (3, 3), "Born_charges": (2, 3, 3), "Quadrupole_tensor": (2, 3, 3, 3)}
    for k, sh in exp.items():
        assert hd[k][0] == sh, k
    assert hd["elph_mat"][1] == "H5T_IEEE_F64LE" and hd["qmap"][1] == "H5T_STD_I32LE"
    assert hd["elph_mat"][2] == "CHUNKED"
    assert hd["elph_mat"][3] == (1, ndb.elph_chunk_nk(nk, nm, ns, nb), nm, ns, nb, nb, 2)
    assert hd["kpoints"][2] == "CONTIGUOUS"
    # dimensions as netCDF sees them: scale datasets of the reference's names
    for d, n in (("nq", nq), ("nk", nk), ("nmodes", nm), ("re_im", 2), ("pol", 3)):
        assert hd[d][0] == (n,), d
    assert dim_list(path, "elph_mat") == ["nq", "nk", "nmodes", "nspin", "initial_band",
                                          "final_band_PH_abs", "re_im"]
    assert dim_list(path, "symmetry_matrices") == ["nsym_ph", "pol", "pol"]
    got = raw(path, "elph_mat", np.float64, (nq, nk, nm, ns, nb, nb, 2), tmp_path)
    assert np.array_equal(got[..., 0] + 1j * got[..., 1], g)
    pv = raw(path, "POLARIZATION_VECTORS", np.float64, (nq, nm, nm // 3, 3, 2), tmp_path)
    assert np.array_equal((pv[..., 0] + 1j * pv[..., 1]).reshape(nq, nm, nm), eig)
    txt = h5("-d", "/convention", path)
    assert re.search(r'"s", "t", "a", "n", "d", "a", "r", "d", ""', txt)
    assert "H5T_C_S1" in txt and "STRSIZE 1" in txt
    txt = h5("-A", "-d", "/nq", path)
    assert "This is a netCDF dimension but not a netCDF variable." in txt
    assert "_Netcdf4Dimid" in txt


def test_ndb_dmats_and_errors(tmp_path):
    say("case: ndb.Dmats and writer argument checks")
    rng = np.random.default_rng(1)
    D = rng.standard_normal((2, 8, 1, 4, 4)) + 1j * rng.standard_normal((2, 8, 1, 4, 4))
    path = str(tmp_path / "ndb.Dmats")
    ndb.write_dmats(path, D)
    hd = header(path)
    assert hd["Dmats"][0] == (2, 8, 1, 4, 4, 2) and hd["Dmats"][2] == "CHUNKED"
    assert dim_list(path, "Dmats") == ["nsym_ph", "nkpts", "nspin", "Rk_band", "k_band", "re_im"]
    got = raw(path, "Dmats", np.float64, (2, 8, 1, 4, 4, 2), tmp_path)
    assert np.array_equal(got[..., 0] + 1j * got[..., 1], D)
    f = ndb.NC4File(str(tmp_path / "x.nc"))
    f.def_var("a", ndb.NC_DOUBLE, (3, 4), ("n3", "n4"))
    with pytest.raises(RuntimeError):
        f.def_var("b", ndb.NC_DOUBLE, (5,), ("n3",))  # dimension length mismatch
    with pytest.raises(RuntimeError):
        f.def_var("a", ndb.NC_DOUBLE, (3,), ("n3",))  # name taken
    with pytest.raises(RuntimeError):
        f.put("a", np.zeros((2, 4)), (2, 0), (2, 4))  # slab past the end
    f.close()
