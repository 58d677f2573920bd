"""FP64 CPU restatement of the learned dV_q network's loss (numpy).

TEST INFRASTRUCTURE: imported only by tests/ as the checker of the device
network in letzelphc_amd/csrc/learned.hip; the product never imports it.

The network: layers l = 0 .. L-1, each a 3-D convolution on the periodic
FFT grid with 27 taps t = (dx+1)*9 + (dy+1)*3 + (dz+1), dx, dy, dz in
{-1, 0, 1}, the tap reading the input at (x+dx, y+dy, z+dz):

    Z_l[p][c] = b_l[c] + sum_{t, ci} A_{l-1}[p + off_t][ci] W_l[t * cin + ci][c]

with A_{-1} the input features (128 channels), A_l = SiLU(Z_l) for hidden
layers and the output out = base + Z_{L-1}[:, :nout].  Loss: the mean of
(out - target)^2 over points, samples and the nout channels.  Parameters
flat as [W_0 (27 cin_0, 64), b_0 (64), W_1, b_1, ...].

`forward(..., bf16=True)` rounds weights, inputs and hidden activations to
BF16 (round to nearest even) where the device stores them in BF16, so that
the device forward can be compared at the FP32-accumulation level;
`loss` (FP64 throughout) is the function whose finite differences pin the
device gradients.
"""
from __future__ import annotations

import numpy as np

NN_CIN = 128
NN_W = 64
TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16_round(a: np.ndarray) -> np.ndarray:
    """round float values to the nearest BF16 (ties to even)"""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, depth: int):
    out, o = [], 0
    for l in range(depth + 2):
        cin = NN_CIN if l == 0 else NN_W
        K = 27 * cin
        W = params[o:o + K * NN_W].reshape(K, NN_W).astype(np.float64)
        o += K * NN_W
        b = params[o:o + NN_W].astype(np.float64)
        o += NN_W
        out.append((W, b))
    assert o == params.size
    return out


def _conv(A, W, b, grid):
    """A (S, N, cin) -> Z (S, N, 64)"""
    S, N, cin = A.shape
    nx, ny, nz = grid
    A5 = A.reshape(S, nx, ny, nz, cin)
    Z = np.zeros((S, nx, ny, nz, NN_W))
    for t, (dx, dy, dz) in enumerate(TAPS):
        sh = np.roll(A5, shift=(-dx, -dy, -dz), axis=(1, 2, 3))  # sh[x] = A[x + d]
        Z += sh @ W[t * cin:(t + 1) * cin]
    return Z.reshape(S, N, NN_W) + b


def _silu(z):
    return z / (1.0 + np.exp(-z))


def forward(params, X0, base, grid, nout: int, depth: int, bf16: bool = False):
    layers = unpack(np.asarray(params), depth)
    A = bf16_round(X0) if bf16 else np.asarray(X0, np.float64)
    Z = None
    for l, (W, b) in enumerate(layers):
        if bf16:
            W = bf16_round(W)
        Z = _conv(A, W, b, grid)
        if l + 1 < len(layers):
            A = _silu(Z)
            if bf16:
                A = bf16_round(A)
    return np.asarray(base, np.float64) + Z[:, :, :nout]


def loss(params, X0, base, target, grid, nout: int, depth: int) -> float:
    out = forward(params, X0, base, grid, nout, depth)
    return float(np.mean((out - np.asarray(target, np.float64)) ** 2))


def fd_grad(params, idx, X0, base, target, grid, nout, depth, h=1e-3):
    """central finite differences of `loss` in FP64 at the entries idx"""
    p = np.asarray(params, np.float64).copy()
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        v = p[i]
        p[i] = v + h
        lp = loss(p, X0, base, target, grid, nout, depth)
        p[i] = v - h
        lm = loss(p, X0, base, target, grid, nout, depth)
        p[i] = v
        g[j] = (lp - lm) / (2 * h)
    return g
