"""The ctypes mirror of the C-ABI structs (letzelphc_amd/abi.py and the
parser's QexmlInfo) against what a C compiler makes of include/lelph_hip.h:
every struct's sizeof and every field's offsetof, from tests/c_abi/layout.c
compiled with gcc on the spot (host only)."""
import ctypes as C
import os
import subprocess

import pytest

from letzelphc_amd import abi, lelph

from .util import say

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "lelph_symmetry": abi.Symmetry,
    "lelph_lattice": abi.Lattice,
    "lelph_local_pseudo": abi.LocalPseudo,
    "lelph_pseudo": abi.Pseudo,
    "lelph_pair": abi.Pair,
    "lelph_longrange": abi.LongRange,
    "lelph_qexml_info": lelph.QexmlInfo,
    "lelph_tuning": abi.Tuning,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("layout") / "layout")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_abi", "layout.c"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lay = {}
    for line in out.splitlines():
        key, val = line.rsplit(" ", 1)
        lay[key] = int(val)
    return lay


@pytest.mark.parametrize("name", list(MIRRORS))
def test_struct_layout(c_layout, name):
    M = MIRRORS[name]
    say(f"case: layout of {name}")
    assert c_layout[f"{name} sizeof"] == C.sizeof(M), name
    cfields = {k.split(".", 1)[1]: v for k, v in c_layout.items()
               if k.startswith(name + ".")}
    pyfields = [f[0] for f in M._fields_]
    assert sorted(cfields) == sorted(pyfields), (name, sorted(cfields), sorted(pyfields))
    for f in pyfields:
        assert getattr(M, f).offset == cfields[f], (name, f)


def test_status_names_cover_header():
    """every status of the header's enum has a name in the mirror"""
    import re
    txt = open(os.path.join(ROOT, "include", "lelph_hip.h")).read()
    codes = {int(v) for v in re.findall(r"LELPH_ERR_[A-Z]+ = (\d+)", txt)}
    assert codes <= set(abi.STATUS_NAMES)
