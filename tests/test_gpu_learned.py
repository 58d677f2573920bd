"""The learned stand-in for the Fourier interpolation of dV_q on the GPU.

- the device forward (BF16 operands, FP32 accumulation on MFMA) against
  the FP64 oracle with the same BF16 roundings;
- the device gradients of every layer against central finite differences
  of the oracle's FP64 loss;
- the learned-form workload: the 2-atom fcc cell (24^3), coarse 4x4x4 q
  with a Born-charge long range (eps = 12), the network trained on the
  device on fine 12x12x12 q the Fourier interpolation produces; on
  held-out fine q: the training loss falls, the network's periodic part is
  closer (relative rms) to the interpolation than the trilinear blend it
  corrects, and the
  judge (g from the model's dV against g from the interpolated dV, on
  (k, k+q) pairs of a 12x12x12 k grid) accepts a q only at
  |dg|max/|g|max <= 1e-3 and otherwise falls back to the interpolation, so
  every final g is within 1e-3 of the interpolation's.
"""
import numpy as np
import pytest

from letzelphc_amd import learned, lelph, synth
from letzelphc_amd.abi import LrParams
from oracle import learned_oracle as LO

from .util import rel_err, say

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

GRID = (8, 8, 8)
NF = 6


def _rand_inputs(rng, S, n, nout):
    X0 = np.zeros((S, n, LO.NN_CIN), np.float32)
    X0[:, :, :2 * NF * 8 + 2 * NF + 6] = rng.standard_normal((S, n, 2 * NF * 8 + 2 * NF + 6))
    base = rng.standard_normal((S, n, nout)).astype(np.float32)
    target = rng.standard_normal((S, n, nout)).astype(np.float32)
    return X0, base, target


def _rand_params(net, rng):
    p = net.get_params().astype(np.float64)
    # the output layer starts at zero; give every live entry a value
    o = 0
    for l, (K, W) in enumerate(net.layer_shapes()):
        w = p[o:o + K * W].reshape(K, W)
        live = w != 0 if l < net.depth + 1 else np.zeros_like(w, bool)
        if l == net.depth + 1:
            live[:, :net.nout] = True
        w[live] = rng.standard_normal(int(live.sum())) * np.sqrt(2.0 / K)
        o += K * W
        b = p[o:o + W]
        b[:(net.nout if l == net.depth + 1 else W)] = 0.05 * rng.standard_normal(
            net.nout if l == net.depth + 1 else W)
        o += W
    return p.astype(np.float32)


def test_forward_vs_oracle():
    rng = np.random.default_rng(11)
    net = learned.Network(GRID, NF, depth=1, seed=3)
    try:
        p = _rand_params(net, rng)
        net.set_params(p)
        X0, base, _ = _rand_inputs(rng, 2, net.n, net.nout)
        got = net.forward(X0, base)
        ref = LO.forward(p, X0, base, GRID, net.nout, net.depth, bf16=True)
        err = rel_err(got - base, ref - base)
        say(f"case: nn forward vs oracle (BF16 roundings emulated): {err:.3e}")
        assert err < 2e-3
    finally:
        net.close()


def test_gradients_vs_finite_differences():
    rng = np.random.default_rng(12)
    net = learned.Network(GRID, NF, depth=1, seed=4)
    try:
        p = _rand_params(net, rng)
        net.set_params(p)
        X0, base, target = _rand_inputs(rng, 2, net.n, net.nout)
        loss, g = net.loss_grad(X0, base, target)
        ref_loss = LO.loss(LO.bf16_round(p), LO.bf16_round(X0), base, target, GRID, net.nout,
                           net.depth)
        say(f"case: nn loss {loss:.6e} vs oracle {ref_loss:.6e}")
        assert abs(loss - ref_loss) / ref_loss < 1e-2
        # finite differences in FP64 on the largest gradient entries of
        # every layer's weights and biases
        idx, o = [], 0
        for l, (K, W) in enumerate(net.layer_shapes()):
            seg = np.abs(g[o:o + K * W])
            idx += list(o + np.argsort(seg)[-12:])
            o += K * W
            seg = np.abs(g[o:o + W])
            idx += list(o + np.argsort(seg)[-4:])
            o += W
        fd = LO.fd_grad(p, idx, X0, base, target, GRID, net.nout, net.depth, h=1e-3)
        gg = g[idx].astype(np.float64)
        err = np.linalg.norm(gg - fd) / np.linalg.norm(fd)
        cos = float(gg @ fd / (np.linalg.norm(gg) * np.linalg.norm(fd)))
        say(f"  gradient vs FD on {len(idx)} entries: rel {err:.3e}, cos {cos:.6f}")
        assert err < 5e-2 and cos > 0.998
    finally:
        net.close()


@pytest.fixture(scope="module")
def c4():
    s = synth.config("c0", with_wfc=False)
    coarse = learned.fine_grid(4)
    s.qpts = [q for q in coarse]
    na = s.natom
    lrp = LrParams(epsilon=12.0 * np.eye(3), Zborn=np.stack([2.0 * np.eye(3), -2.0 * np.eye(3)]),
                   Zvals=np.full(na, 4.0), eta_bare=1.0, eta_induced=1.0, EcutRy=30.0,
                   nmags_add=(1, 0, 0, 0))
    ctx = lelph.Context(s, "fp64", upload=False)
    ctx.interp_begin((4, 4, 4), lr=lrp)
    for iq, q in enumerate(coarse):
        ctx.interp_add_coarse(q, s.eigvec(iq), s.dvscf(iq))
    ctx.interp_finish()
    li = learned.LearnedInterpolation(ctx, (4, 4, 4), lrp, depth=5, seed=7)
    fine = learned.fine_grid(12)
    on_coarse = np.all(np.abs(fine * 4 - np.rint(fine * 4)) < 1e-9, axis=1)
    cand = np.flatnonzero(~on_coarse)
    rng = np.random.default_rng(5)
    rng.shuffle(cand)
    train_q, val_q = fine[cand[:512]], fine[cand[512:520]]
    yield s, ctx, li, train_q, val_q
    li.net.close()
    ctx.close()


def test_learned_c4(c4):
    s, ctx, li, train_q, val_q = c4
    say(f"case: learned c4: {li.net.effective_params()} parameters "
        f"({li.net.nparams()} with padding), {len(train_q)} training q")
    hist = li.train(train_q, epochs=8, batch=8, lr=3e-4, seed=3)
    say(f"  loss by epoch {np.array2string(hist, precision=4)}")
    assert hist[-1] < 0.85 * hist[0]
    ref = li.periodic(val_q)
    net = li.predict_periodic(val_q)
    tri = li.trilinear_periodic(val_q)
    def rms(a, b):
        return float(np.sqrt(np.mean(np.abs(a - b) ** 2) / np.mean(np.abs(b) ** 2)))
    e_net, e_tri = rms(net, ref), rms(tri, ref)
    say(f"  held-out periodic part vs interpolation (relative rms): network {e_net:.3e}, "
        f"trilinear {e_tri:.3e} (gain {e_tri / e_net:.2f}x); max-norm "
        f"{rel_err(net, ref):.3e} / {rel_err(tri, ref):.3e}")
    assert e_net < e_tri / 1.05
    # the judge on a 12x12x12 k grid (identity operations, a few k per q)
    gs = synth.make_system(
        "c0k12", 1000, synth.fcc_cell(10.26), [[0, 0, 0], [0.25, 0.25, 0.25]], [0, 0], [4.0],
        (24, 24, 24), (12, 12, 12), 20.0, 8, [(np.eye(3), np.zeros(3), 0)], 0,
        [q for q in val_q], with_wfc=False)
    kb = gs.kpt_bz_crys
    pairs_q, need = [], set()
    for q in val_q:
        prs = []
        for i in (0, 137, 611):
            d = kb - (kb[i] + q)[None, :]
            d -= np.rint(d)
            j = int(np.argmin(np.linalg.norm(d, axis=1)))
            prs.append((j, i, 0, 0))
            need |= {i, j}
        pairs_q.append(prs)
    synth.make_wfcs(gs, workers=4, only=sorted(need))
    gctx = lelph.Context(gs, "fp64")
    try:
        eig = gs.eigvec(0)
        dv_ref = li.interpolated(val_q)
        dv_net = li.dv(val_q, net)
        dv_tri = li.dv(val_q, tri)
        acc = 0
        for i, q in enumerate(val_q):
            e_n, ok = learned.judge(gctx, gs, q, eig, dv_net[i], dv_ref[i], pairs_q[i])
            e_t, _ = learned.judge(gctx, gs, q, eig, dv_tri[i], dv_ref[i], pairs_q[i])
            # the fallback: a rejected q takes the interpolated dV
            final = dv_net[i] if ok else dv_ref[i]
            e_f, ok_f = learned.judge(gctx, gs, q, eig, final, dv_ref[i], pairs_q[i])
            acc += ok
            say(f"  q {np.round(q, 3)}: |dg|/|g| network {e_n:.3e}, trilinear {e_t:.3e}, "
                f"{'accepted' if ok else 'fallback'}; final {e_f:.1e}")
            assert ok_f and (e_n <= 1e-3) == ok
        say(f"  accepted {acc} of {len(val_q)} held-out q")
    finally:
        gctx.close()
