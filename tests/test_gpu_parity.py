"""GPU parity of the HIP path against the FP64 CPU oracle, through the C-ABI.

Bars: FP64 path |dg|max/|g|max <= 1e-10; FP32 path <= 1e-6.
"""
import bisect

import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from oracle import oracle as O

from .util import rel_err, say

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-6


@pytest.fixture(scope="module")
def c0():
    s = synth.config("c0")
    o = O.OracleSystem(s)
    O.set_threads(8)
    dv = o.dv_total(0)
    return s, o, dv


@pytest.fixture(scope="module")
def c0ctx(c0):
    s, o, dv = c0
    ctx = lelph.Context(s, "fp64")
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
    yield ctx
    ctx.close()


def test_vloc_table(c0, c0ctx):
    s, o, _ = c0
    say("case: V_loc(G) table c0")
    g1, v1, d1 = c0ctx.vloc_table()
    g0, v0, d0 = o.vloc_table()
    assert g1.shape == g0.shape
    assert rel_err(g1, g0) == 0.0
    assert rel_err(v1, v0) < 1e-12
    assert rel_err(d1, d0) < 1e-9


def test_dvlocq(c0, c0ctx):
    s, o, _ = c0
    say("case: dVlocq c0 q0")
    vg = c0ctx.dvlocq(s.qpts[0], s.eigvec(0))
    vc = o.dvlocq(0)
    assert rel_err(vg, vc) < TOL64


def test_dvscf_total(c0, c0ctx):
    s, o, dv = c0
    say("case: dvscf + bare local c0")
    c0ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
    assert rel_err(c0ctx.get_dvscf(), dv) < TOL64


@pytest.mark.parametrize("ip", [0, 1, 3, 17])
def test_local_fp64(c0, c0ctx, ip):
    s, o, dv = c0
    pr = synth.kq_pairs(s, 0)[ip]
    say(f"case: elph_local fp64 c0 pair {pr}")
    g = c0ctx.elph_local(*pr)
    ref = o.elph_local(0, *pr, dv)
    assert rel_err(g, ref) < TOL64


@pytest.mark.parametrize("ip", [0, 3, 17])
def test_nonlocal_fp64(c0, c0ctx, ip):
    s, o, dv = c0
    pr = synth.kq_pairs(s, 0)[ip]
    say(f"case: add_elph_nonlocal fp64 c0 pair {pr}")
    base = o.elph_local(0, *pr, dv)
    g = c0ctx.add_nonlocal(*pr, base)
    ref = o.add_nonlocal(0, *pr, base)
    assert rel_err(g - base, ref - base) < TOL64


def test_compute_elphq_fp64(c0, c0ctx):
    s, o, dv = c0
    say("case: compute_elphq fp64 c0 all k")
    g = c0ctx.compute_elphq(s.qpts[0], non_loc=True, kminusq=False)
    ref = o.compute_elphq(0, dv, non_loc=True, kminusq=False)
    assert rel_err(g, ref) < TOL64


def test_compute_elphq_kminusq(c0, c0ctx):
    s, o, dv = c0
    say("case: compute_elphq fp64 c0 yambo convention")
    g = c0ctx.compute_elphq(s.qpts[0], non_loc=True, kminusq=True)
    ref = o.compute_elphq(0, dv, non_loc=True, kminusq=True, k_first=0, k_count=16)
    assert rel_err(g[:16], ref) < TOL64


@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3"])
def test_compute_elphq_reduced(c0, c0ctx, prec):
    """FP32 and split-product forms at 1e-6.  c0's 8 bands are off the
    producer/consumer path, so a bf16x3 request runs as bf16x6 there."""
    s, o, dv = c0
    say(f"case: compute_elphq {prec} c0 all k")
    c0ctx.set_precision(prec)
    try:
        eff = c0ctx.effective_precision()
        assert eff == ("bf16x6" if prec == "bf16x3" else prec)
        g = c0ctx.compute_elphq(s.qpts[0], non_loc=True)
    finally:
        c0ctx.set_precision("fp64")
    ref = o.compute_elphq(0, dv, non_loc=True)
    err = rel_err(g, ref)
    say(f"  {prec} (runs as {eff}): |dg|/|g|max = {err:.3e}")
    assert err < TOL32


@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3"])
@pytest.mark.parametrize("kind", ["spin2", "spinor_nmag1", "spinor_nmag4", "soc_nmag4"])
def test_tiny_variants_reduced(kind, prec):
    """FP32 and split-product contractions on the spin / spinor / 2x2
    potential branches (the split kernel's element-wise staging path for
    spinors and nmag 4, spin-resolved dV for nmag 2), against the oracle."""
    kw = dict(nbnds=4)
    if kind == "spin2":
        kw.update(nspin=2, nmag=2)
    elif kind == "spinor_nmag1":
        kw.update(nspinor=2, nmag=1)
    elif kind == "spinor_nmag4":
        kw.update(nspinor=2, nmag=4)
    elif kind == "soc_nmag4":
        kw.update(nspinor=2, nmag=4, soc=True, syms="cubic", q=(0.0, 0.5, 0.0))
    s = synth.tiny(**kw)
    o = O.OracleSystem(s)
    say(f"case: tiny {kind} {prec}")
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        ctx.set_precision(prec)
        g = ctx.compute_elphq(s.qpts[0], non_loc=True)
    finally:
        ctx.close()
    ref = o.compute_elphq(0, o.dv_total(0), non_loc=True)
    err = rel_err(g, ref)
    say(f"  {prec}: |dg|/|g|max = {err:.3e}")
    assert err < TOL32


@pytest.mark.parametrize("kind", ["spin2", "spinor_nmag1", "spinor_nmag4", "cubic_syms",
                                  "bare", "dfpt_local", "slab_2d", "soc", "soc_nmag4"])
def test_tiny_variants(kind):
    kw = dict(nbnds=4)
    kernel = "dfpt"
    if kind == "spin2":
        kw.update(nspin=2, nmag=2)
    elif kind == "spinor_nmag1":
        kw.update(nspinor=2, nmag=1)
    elif kind == "spinor_nmag4":
        kw.update(nspinor=2, nmag=4)
    elif kind == "cubic_syms":
        kw.update(syms="cubic", q=(0.0, 0.5, 0.0))
    elif kind in ("bare", "dfpt_local"):
        kernel = kind
    elif kind == "slab_2d":
        kw.update(dimension="2")
    elif kind == "soc":
        kw.update(nspinor=2, nmag=1, soc=True)
    elif kind == "soc_nmag4":
        kw.update(nspinor=2, nmag=4, soc=True, syms="cubic", q=(0.0, 0.5, 0.0))
    s = synth.tiny(**kw)
    o = O.OracleSystem(s)
    say(f"case: tiny {kind}")
    ctx = lelph.Context(s, "fp64")
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel=kernel)
    dv = o.dv_total(0, kernel=kernel)
    assert rel_err(ctx.get_dvscf(), dv) < TOL64
    non_loc = kernel in ("dfpt", "bare")
    g = ctx.compute_elphq(s.qpts[0], non_loc=non_loc)
    ref = o.compute_elphq(0, dv, non_loc=non_loc)
    ctx.close()
    assert rel_err(g, ref) < TOL64


def test_patterns_rotation():
    s = synth.tiny(nbnds=4)
    o = O.OracleSystem(s)
    say("case: tiny dvscf in pattern basis")
    rng = np.random.default_rng(5)
    m = rng.standard_normal((s.nmodes, s.nmodes)) + 1j * rng.standard_normal((s.nmodes, s.nmodes))
    pats, _ = np.linalg.qr(m)
    pats = np.ascontiguousarray(pats)
    dvp = s.dvscf(0).copy()
    ctx = lelph.Context(s, "fp64")
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), dvp, patterns=pats, kernel="dfpt")
    ref = dvp.copy()
    import ctypes
    from letzelphc_amd.abi import dptr
    O.lib().orc_dvscf_pattern_to_mode(dptr(ref), dptr(s.eigvec(0)), dptr(pats), o.lat)
    O.lib().orc_add_dvscf_qe(dptr(ref), dptr(o.dvlocq(0)), o.lat)
    got = ctx.get_dvscf()
    ctx.close()
    assert rel_err(got, ref) < TOL64


def test_error_paths():
    s = synth.tiny(nbnds=2)
    say("case: error paths")
    ctx = lelph.Context(s, "fp64")
    with pytest.raises(lelph.LelphError) as e:
        ctx.elph_local(0, 0, 0, 0, qpt=np.zeros(3))
    assert e.value.status == 6  # no dV set yet
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0))
    with pytest.raises(lelph.LelphError) as e:
        ctx.elph_local(0, 999, 0, 0)
    assert e.value.status == 1
    with pytest.raises(lelph.LelphError) as e:
        ctx.compute_elphq(np.array([0.1234, 0.0, 0.0]))
    assert e.value.status == 7
    ctx.close()


def test_c1_symmetry_pairs():
    """The layered c1 cell (45x45x216, 12 operations incl. time reversal),
    16 bands: pairs whose k and k+q need non-trivial and time-reversed
    operations, FP64 to 1e-10 and the reduced variants (generic, unaligned
    grid path; bf16x3 runs as bf16x6 there) to 1e-6."""
    s = synth.config("c1", nbnds=16, kgrid=(6, 6, 1))
    o = O.OracleSystem(s)
    O.set_inner_threads(8)
    pairs = synth.kq_pairs(s, 0)
    pick, seen = [], set()
    for pr in pairs:
        key = (pr[2] > 0, pr[3] > 0, pr[2] >= 6 or pr[3] >= 6)
        if key not in seen:
            seen.add(key)
            pick.append(pr)
    say(f"case: c1 pairs {pick}")
    dv = o.dv_total(0)
    ctx = lelph.Context(s, "fp64")
    ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
    try:
        for pr in pick:
            ref = o.add_nonlocal(0, *pr, o.elph_local(0, *pr, dv))
            for prec, tol in (("fp64", TOL64), ("fp32", TOL32), ("bf16x3", TOL32)):
                ctx.set_precision(prec)
                g = ctx.gkq(*pr)
                err = rel_err(g, ref)
                say(f"  {pr} {prec}: {err:.3e}")
                assert err < tol, (pr, prec)
    finally:
        O.set_inner_threads(1)
        ctx.close()


@pytest.mark.parametrize("kind", ["cubic", "spinor", "c1"])
def test_electronic_reps(kind):
    """Representation matrices D = <Rk,b|U(R)|k,a> against the oracle for
    every symmetry of the system as R (time-reversed ones included)."""
    if kind == "cubic":
        s = synth.tiny(nbnds=4, syms="cubic", q=(0.0, 0.5, 0.0))
        ks = range(len(s.kpt_bz_crys))
    elif kind == "spinor":
        s = synth.tiny(nbnds=4, nspinor=2, nmag=1, syms="cubic", q=(0.0, 0.5, 0.0))
        ks = range(len(s.kpt_bz_crys))
    else:
        s = synth.config("c1", nbnds=8, kgrid=(6, 6, 1))
        ks = range(0, len(s.kpt_bz_crys), 5)
    o = O.OracleSystem(s)
    ctx = lelph.Context(s, "fp64")
    say(f"case: electronic_reps {kind}, {len(s.syms)} symmetries")
    try:
        for R, tau, tr in s.syms:
            for ik in ks:
                got = ctx.electronic_reps(R, tau, tr, ik)
                ref = o.electronic_reps(R, tau, tr, ik)
                assert rel_err(got, ref) < TOL64, (kind, ik)
    finally:
        ctx.close()


@pytest.mark.parametrize("nb", [5, 40, 70])
def test_elph_q_rotate_gpu(nb):
    """q-star rotation on the device against the oracle's elph_q_rotate
    (both time-reversal branches; nb below, at and across the 32-wide tile)."""
    from letzelphc_amd.abi import dptr
    rng = np.random.default_rng(nb)
    nm, ns = 6, 2
    g = rng.standard_normal((nm, ns, nb, nb)) + 1j * rng.standard_normal((nm, ns, nb, nb))
    Dl = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
    Dr = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
    s = synth.tiny(nbnds=4)
    ctx = lelph.Context(s, "fp64")
    say(f"case: elph_q_rotate_gpu nb={nb}")
    try:
        for tr in (0, 1):
            got = ctx.elph_q_rotate(Dl, g, Dr, bool(tr))
            ref = np.zeros_like(g)
            O.lib().orc_elph_q_rotate(dptr(Dl), dptr(g), dptr(Dr), nm, ns, nb, tr, dptr(ref))
            assert rel_err(got, ref) < TOL64, (nb, tr)
    finally:
        ctx.close()


def _chain_links(pairs):
    """pairs whose k side (iBZ k, symmetry) is another pair's k+q side: the
    psi box the batch takes over instead of transforming again"""
    ks = {(ik, ksym) for (_, ik, _, ksym) in pairs}
    return sum((ikq, kqsym) in ks and (ikq, kqsym) != (ik, ksym)
               for (ikq, ik, kqsym, ksym) in pairs)


def _chain_sides(pairs):
    """psi sides a batch transforms: the chain order of the library (a pair
    follows the first pair, in (iBZ k, symmetry) order, whose k side is its
    k+q side and that has no predecessor yet; open chains first, then what
    is left of cycles), one transform for a chain head's k side and one for
    every k+q side except where it is the head's k side (a closed cycle)."""
    n = len(pairs)
    keys = sorted(((ik, ks), i) for i, (_, ik, _, ks) in enumerate(pairs))
    nxt, has_prev = [-1] * n, [False] * n
    for i, (ikq, _, kqs, _) in enumerate(pairs):
        j = bisect.bisect_left(keys, ((ikq, kqs), -1))
        if j < n and keys[j][0] == (ikq, kqs) and keys[j][1] != i and not has_prev[keys[j][1]]:
            nxt[i] = keys[j][1]
            has_prev[keys[j][1]] = True
    order, used = [], [False] * n
    for first_pass in (True, False):
        for i in range(n):
            if used[i] or (first_pass and has_prev[i]):
                continue
            j = i
            while j >= 0 and not used[j]:
                used[j] = True
                order.append(j)
                j = nxt[j]
    sides, prev, head = 0, None, None
    for i in order:
        ikq, ik, kqs, ks = pairs[i]
        if prev != (ik, ks):
            sides, head = sides + 1, (ik, ks)
        if (ikq, kqs) != head:
            sides += 1
        prev = (ikq, kqs)
    return sides


def test_chain_sides_counts():
    """A chain a -> b -> c -> a (a closed cycle) needs three transforms, an
    open chain a -> b -> c four, a self-loop one, two disjoint pairs four."""
    cyc = [(1, 0, 0, 0), (2, 1, 0, 0), (0, 2, 0, 0)]
    assert _chain_sides(cyc) == 3
    assert _chain_sides(cyc[::-1]) == 3
    assert _chain_sides([(1, 0, 0, 0), (2, 1, 0, 0), (3, 2, 0, 0)]) == 4
    assert _chain_sides([(0, 0, 0, 0)]) == 1
    assert _chain_sides([(1, 0, 0, 0), (3, 2, 0, 0)]) == 4


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3", "bf16x6"])
def test_batch_gamma_self_loops(prec):
    """At q = 0 every pair's k+q side is its own k side (a cycle of one
    pair): a batch transforms one psi side per pair, takes the k+q box over
    from the pair's own psi_k box (a copy in FP64 / FP32, read as it is by
    the bf16 presplit), and gives the g of pair-by-pair calls bit for bit
    and of the oracle to the precision's bar."""
    s = synth.tiny(nbnds=4, q=(0.0, 0.0, 0.0))
    o = O.OracleSystem(s)
    pairs = synth.kq_pairs(s, 0)
    assert all(ikq == ik and kqs == ks for (ikq, ik, kqs, ks) in pairs)
    say(f"case: q = 0 batch {prec}, {len(pairs)} self-loop pairs")
    ctx = lelph.Context(s, "fp64")
    try:
        ctx.set_dvscf(s.qpts[0], s.eigvec(0), s.dvscf(0), kernel="dfpt")
        ctx.set_precision(prec)
        gb = ctx.gkq_batch(pairs, non_loc=True, qpt=s.qpts[0])
        sides = ctx.last_timings()["box_sides"]
        one = np.stack([ctx.gkq(*p, non_loc=True, qpt=s.qpts[0]) for p in pairs])
    finally:
        ctx.close()
    assert sides == len(pairs)
    assert np.array_equal(gb, one)
    dv = o.dv_total(0)
    ref = np.stack([o.add_nonlocal(0, *p, o.elph_local(0, *p, dv)) for p in pairs[:3]])
    err = rel_err(gb[:3], ref)
    say(f"  vs oracle: {err:.3e}")
    assert err < (TOL64 if prec == "fp64" else TOL32)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("bf16x3", 0.0)])
def test_batch_shuffled_with_repeats(c0, c0ctx, prec, tol):
    """A batch in random order with repeated pairs: repeats compete for the
    same chain links, so heads, cycle closes and mirrored non-local sides
    fall in irregular places; every pair's g equals its pair-by-pair g (bit
    for bit in the split precisions, which build every box in k-side form)
    and the batch transforms exactly the sides its chain walk leaves."""
    s, o, dv = c0
    rng = np.random.default_rng(5)
    base = synth.kq_pairs(s, 0)
    pairs = base + [base[i] for i in rng.choice(len(base), 12, replace=False)]
    pairs = [pairs[i] for i in rng.permutation(len(pairs))]
    want_sides = _chain_sides(pairs)
    say(f"case: shuffled batch with repeats {prec}, {len(pairs)} pairs, sides {want_sides}")
    c0ctx.set_precision(prec)
    try:
        gb = c0ctx.gkq_batch(pairs, non_loc=True, qpt=s.qpts[0])
        sides = c0ctx.last_timings()["box_sides"]
        one = {p: c0ctx.gkq(*p, non_loc=True, qpt=s.qpts[0]) for p in set(pairs)}
    finally:
        c0ctx.set_precision("fp64")
    assert sides == want_sides
    ref = np.stack([one[p] for p in pairs])
    if tol == 0.0:
        assert np.array_equal(gb, ref)
    else:
        assert rel_err(gb, ref) < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", TOL32), ("bf16x3", TOL32)])
@pytest.mark.parametrize("kminusq", [False, True])
def test_batch_box_reuse(c0, c0ctx, prec, tol, kminusq):
    """A batch walks k -> k+q chains and takes each pair's k box from the
    previous pair's k+q box, shifted back by its umklapp vector, and the
    last pair of a closed cycle its k+q box from the head's k box, shifted
    by its own; pair by pair (every box transformed) gives the same g (the
    bf16 split precisions bit for bit).  The pairs go in reversed order so
    the batch has to rebuild the chains itself; it transforms exactly the
    sides the chains leave."""
    s, o, dv = c0
    pairs = synth.kq_pairs(s, 0, kminusq=kminusq)[::-1]
    links = _chain_links(pairs)
    want_sides = _chain_sides(pairs)
    say(f"case: box reuse {prec} kminusq={kminusq} pairs={len(pairs)} links={links} "
        f"sides={want_sides}")
    assert links >= len(pairs) // 2
    c0ctx.set_precision(prec)
    try:
        gb = c0ctx.gkq_batch(pairs, non_loc=True, qpt=s.qpts[0])
        sides = c0ctx.last_timings()["box_sides"]
        one = np.stack([c0ctx.gkq(*p, non_loc=True, qpt=s.qpts[0]) for p in pairs])
        assert c0ctx.last_timings()["box_sides"] == 2
    finally:
        c0ctx.set_precision("fp64")
    assert sides == want_sides
    err = rel_err(gb, one)
    say(f"  batch vs pair by pair: {err:.3e}")
    assert err < tol
    if prec.startswith("bf16"):
        # split precisions build every box in k-side form (the umklapp phase
        # goes into the presplit), so a batch takes over exactly the boxes
        # pair-by-pair calls transform: bit-identical
        assert np.array_equal(gb, one)
    if prec == "fp64":
        ref = o.compute_elphq(0, dv, non_loc=True, kminusq=kminusq, k_first=0, k_count=4)
        n = len(pairs)
        assert rel_err(gb[::-1][:4], ref) < TOL64
        assert n == len(s.kpt_bz_crys)
