"""Pins of the learned network's FP64 oracle (oracle/learned_oracle.py) on
the CPU: BF16 rounding against torch's bfloat16 conversion, the periodic
3-D convolution against scipy.ndimage.correlate with wrap-around, and the
finite-difference gradient against the closed form of a linear model."""
import numpy as np
import pytest

from oracle import learned_oracle as LO

from .util import say


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.standard_normal(10000) * 10 ** rng.uniform(-6, 6, 10000),
                        [0.0, 1.0, -1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9]])
    ref = torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    say("case: bf16 rounding vs torch")
    assert np.array_equal(LO.bf16_round(a), ref)


def test_conv_matches_scipy_correlate():
    from scipy.ndimage import correlate
    rng = np.random.default_rng(4)
    grid = (4, 5, 6)
    cin = 3
    A = rng.standard_normal((1, int(np.prod(grid)), cin))
    W = rng.standard_normal((27 * cin, LO.NN_W))
    b = rng.standard_normal(LO.NN_W)
    Z = LO._conv(A, W, b, grid)
    say("case: learned-oracle convolution vs scipy correlate (wrap)")
    A5 = A.reshape(grid + (cin,))
    for c in (0, 17, 63):
        ref = np.full(grid, b[c])
        for ci in range(cin):
            k = np.zeros((3, 3, 3))
            for t, (dx, dy, dz) in enumerate(LO.TAPS):
                k[dx + 1, dy + 1, dz + 1] = W[t * cin + ci, c]
            ref += correlate(A5[..., ci], k, mode="wrap")
        assert np.max(np.abs(Z[0, :, c].reshape(grid) - ref)) < 1e-12


def test_fd_grad_linear_closed_form():
    """depth 0 (two layers) with the first layer's bias only: for a
    single-sample model the gradient of the mean square loss with respect
    to the output bias is 2 mean(residual) per channel"""
    rng = np.random.default_rng(5)
    grid = (4, 4, 4)
    nout = 4
    n = int(np.prod(grid))
    X0 = np.zeros((1, n, LO.NN_CIN))
    base = rng.standard_normal((1, n, nout))
    target = rng.standard_normal((1, n, nout))
    sizes = [27 * LO.NN_CIN * LO.NN_W + LO.NN_W, 27 * LO.NN_W * LO.NN_W + LO.NN_W]
    p = np.zeros(sum(sizes))
    p[sizes[0] + 27 * LO.NN_W * LO.NN_W:] = rng.standard_normal(LO.NN_W) * 0.1
    idx = [sizes[0] + 27 * LO.NN_W * LO.NN_W + c for c in range(nout)]
    g = LO.fd_grad(p, idx, X0, base, target, grid, nout, depth=0)
    out = LO.forward(p, X0, base, grid, nout, depth=0)
    want = 2 * np.mean(out - target, axis=(0, 1)) / nout
    say("case: learned-oracle finite differences vs closed form")
    assert np.max(np.abs(g - want)) < 1e-9
