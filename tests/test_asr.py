"""Acoustic sum rules (host code of the library, lelph_asr_fc /
lelph_asr_born) and the slab quadrupole conversion against the reference's own apply_acoustic_sum_rule_fc /
_born_charges compiled from src/phonon/asr.c and src/common/lsmr_solver.c
into oracle/_ref: the committed fixture tests/golden/asr_vectors.npz and,
where the library is built, a second seed live.  The library computes the
projection exactly (normal equations, pivoted Cholesky) where the
reference iterates LSMR to 1e-12, so they agree to the solver's
tolerance.  Plus what fixes the result independently of either code: the
translational rule holds, the R / -R blocks are transposes and R = 0 is
symmetric, and the projection is idempotent."""
import os

import numpy as np
import pytest

from letzelphc_amd import lelph
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "asr_vectors.npz")
CASES = ("fcc2", "hex3", "tri2")


def _run(f, name, mode):
    g = f[f"asr_{name}_grid"]
    return lelph.asr_fc(mode, g, f[f"asr_{name}_frc"], f[f"asr_{name}_pos"],
                        f[f"asr_{name}_lat"], f[f"asr_{name}_ws"], f[f"asr_{name}_degen"])


def _check(f, tol=1e-8):
    for name in CASES:
        for mode in (1, 2, 3, 4):
            ref = f[f"asr_{name}_out{mode}"]
            out = _run(f, name, mode)
            err = np.abs(out - ref).max() / np.abs(ref).max()
            assert err < (1e-15 if mode == 1 else tol), (name, mode, err)
        for mode in (1, 2):
            z = lelph.asr_born(mode, f[f"asr_{name}_zb"])
            ref = f[f"asr_{name}_zb{mode}"]
            if mode == 2 and not np.isfinite(ref).all():
                # the reference's LSMR path for Born charges breaks down on
                # some inputs (NaN, e.g. seed 8 hex3); the projection it
                # aims at is the mean subtraction of its simple mode
                ref = f[f"asr_{name}_zb1"]
            assert np.abs(z - ref).max() < 1e-14


def test_asr_vs_reference_fixture():
    with np.load(FIX) as f:
        _check(dict(f))


@pytest.mark.skipif(not os.path.exists(O.REF_LIB), reason="oracle/_ref not built")
def test_asr_vs_reference_live():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_ref_vectors", os.path.join(HERE, "golden", "make_ref_vectors.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    L = mod.load_asr(mod.load_interp(mod.load_ref()))
    _check(mod.make_asr(L, seed=8))


def _mR(g, grid):
    gyz = grid[1] * grid[2]
    rx, ry, rz = g // gyz, (g % gyz) // grid[2], g % grid[2]
    return ((grid[2] - rz) % grid[2]) + ((grid[1] - ry) % grid[1]) * grid[2] + \
        ((grid[0] - rx) % grid[0]) * gyz


@pytest.mark.parametrize("mode", ["crystal", "all", "all_huang"])
def test_asr_invariants(mode):
    with np.load(FIX) as f:
        f = dict(f)
    for name in CASES:
        grid = f[f"asr_{name}_grid"]
        out = _run(f, name, mode)
        ng, nm, _ = out.shape
        na = nm // 3
        # real, -R the transpose of R, R = 0 symmetric (a zone-boundary R
        # that is its own -R modulo the grid is left as it is, asr.c:325-354)
        assert np.abs(out.imag).max() == 0
        assert np.array_equal(out[0], out[0].T)
        for g in range(1, ng):
            if _mR(g, grid) != g:
                assert np.array_equal(out[_mR(g, grid)], out[g].T)
        # translational: sum over R and b of Phi(a alpha, b beta) vanishes
        # for the blocks the Wigner-Seitz search keeps (degeneracy > 0)
        dg = f[f"asr_{name}_degen"].reshape(ng, na, na)
        tot = np.zeros((na, 3, 3))
        blk = out.real.reshape(ng, na, 3, na, 3)
        for g in range(ng):
            for a in range(na):
                for b in range(na):
                    if dg[g, a, b] > 0:
                        tot[a] += blk[g, a, :, b, :]
        assert np.abs(tot).max() < 1e-12 * np.abs(out).max(), (name, mode)
        # a projection: applying it again changes nothing
        again = lelph.asr_fc(mode, grid, out, f[f"asr_{name}_pos"], f[f"asr_{name}_lat"],
                             f[f"asr_{name}_ws"], f[f"asr_{name}_degen"])
        assert np.abs(again - out).max() < 1e-12 * np.abs(out).max(), (name, mode)


def test_quadrupole_3d_to_2d_vs_reference():
    with np.load(FIX) as f:
        for name in CASES:
            out = lelph.quadrupole_3d_to_2d(f[f"q2d_{name}_eps"], f[f"asr_{name}_pos"],
                                            f[f"asr_{name}_zb"], f[f"q2d_{name}_Q"])
            assert np.abs(out - f[f"q2d_{name}_out"]).max() < 1e-14, name


def test_loto_qpoint():
    lat = np.array([[-5.13, 0, -5.13], [0, 5.13, 5.13], [5.13, 5.13, 0]])
    q = lelph.loto_qpoint(lat, [0.0, 0.0, 2.0])
    # back to cartesian: (B / 2 pi) q = 1e-5 z
    assert np.allclose(np.linalg.inv(lat).T @ q, [0, 0, 1e-5], atol=1e-18)
    with pytest.raises(lelph.LelphError):
        lelph.loto_qpoint(lat, [0, 0, 1e-9])


def test_asr_kinds_and_born():
    assert [lelph.asr_kind(k) for k in ("no", "SIMPLE", "Crystal", "all", "all_huang", "x")] \
        == [0, 1, 2, 3, 4, 0]
    z = np.random.default_rng(1).standard_normal((3, 3, 3))
    out = lelph.asr_born("crystal", z)
    assert np.abs(out.sum(0)).max() < 1e-15
    assert np.array_equal(lelph.asr_born("no", z), z)
