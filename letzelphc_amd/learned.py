"""Learned stand-in for the Fourier interpolation of dV_q (host side).

The reference interpolates the screened potential from a coarse q grid onto
fine q points (src/interpolation/interpolation_driver.c:790-856: long range
removed, q -> R, then for each fine q the Wigner-Seitz weighted R -> q sum,
e^{-iqr}, long range added back; fft_R2q_dvscf,
src/interpolation/interpolation_utilities.c:165-238).  Here a small 3-D
convolutional network (lelph_nn, letzelphc_amd/csrc/learned.hip) predicts
the short-range periodic part at a fine q,

    u_q(r) = e^{-2 pi i q.x} (dV_q(r) - dV_q^LR(r))      (cartesian basis),

as a correction to the trilinear blend of the 8 coarse q around it; the long
range comes back analytically (lelph_dv_longrange).  It is trained on the
device, by its own forward and backward kernels, on potentials the Fourier
interpolation itself produces (lelph_interp_dvscf); the interpolation stays
its judge and fallback: for each fine q the network's dV and the
interpolated dV give g on the same (k, k+q) pairs, and a q whose
|dg|max/|g|max exceeds the tolerance (1e-3) takes the interpolated dV.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lelph
from .abi import c_double_p, dptr

_f32p = C.POINTER(C.c_float)


def _fptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float32
    return a.ctypes.data_as(_f32p)


_bound = False


def _lib():
    global _bound
    L = lelph.lib()
    if not _bound:
        i64p = C.POINTER(C.c_int64)
        L.lelph_nn_create.argtypes = [i64p, C.c_int, C.c_int, C.c_uint64,
                                      C.POINTER(C.c_void_p)]
        L.lelph_nn_destroy.argtypes = [C.c_void_p]
        L.lelph_nn_destroy.restype = None
        L.lelph_nn_nparams.argtypes = [C.c_void_p, i64p]
        L.lelph_nn_get_params.argtypes = [C.c_void_p, _f32p]
        L.lelph_nn_set_params.argtypes = [C.c_void_p, _f32p]
        L.lelph_nn_set_coarse.argtypes = [C.c_void_p, i64p, c_double_p, C.c_double]
        L.lelph_nn_features.argtypes = [C.c_void_p, C.c_int, c_double_p, _f32p, _f32p]
        L.lelph_nn_forward.argtypes = [C.c_void_p, C.c_int, _f32p, _f32p, _f32p]
        L.lelph_nn_loss_grad.argtypes = [C.c_void_p, C.c_int, _f32p, _f32p, _f32p,
                                         C.POINTER(C.c_double), _f32p]
        L.lelph_nn_set_targets.argtypes = [C.c_void_p, C.c_int64, _f32p]
        L.lelph_nn_train.argtypes = [C.c_void_p, C.c_int64, c_double_p, C.c_int, C.c_int,
                                     C.c_double, C.c_uint64, c_double_p]
        L.lelph_nn_predict.argtypes = [C.c_void_p, C.c_int64, c_double_p, _f32p]
        _bound = True
    return L


NN_CIN = 128
NN_W = 64
NN_TAPS = 27


class Network:
    """lelph_nn: the convolutional network and its training state."""

    def __init__(self, grid, nfields: int, depth: int = 5, seed: int = 0):
        self.grid = tuple(int(x) for x in grid)
        self.nfields = int(nfields)
        self.nout = 2 * self.nfields
        self.depth = int(depth)
        g = np.ascontiguousarray(self.grid, np.int64)
        h = C.c_void_p()
        lelph._check(_lib().lelph_nn_create(g.ctypes.data_as(C.POINTER(C.c_int64)),
                                            self.nfields, self.depth, int(seed), C.byref(h)),
                     "lelph_nn_create")
        self.h = h
        self.n = int(np.prod(self.grid))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib().lelph_nn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layer_shapes(self):
        """(K, 64) weight shape of every layer, K = 27 * cin"""
        return [(NN_TAPS * (NN_CIN if l == 0 else NN_W), NN_W) for l in range(self.depth + 2)]

    def nparams(self) -> int:
        n = C.c_int64(0)
        lelph._check(_lib().lelph_nn_nparams(self.h, C.byref(n)), "lelph_nn_nparams")
        return n.value

    def effective_params(self) -> int:
        """parameters outside the zero padding (first layer's unused input
        channels, the output layer's unused columns)"""
        cin0 = 2 * self.nfields * 8 + 2 * self.nfields + 6
        n = NN_TAPS * cin0 * NN_W + NN_W
        n += self.depth * (NN_TAPS * NN_W * NN_W + NN_W)
        n += NN_TAPS * NN_W * self.nout + self.nout
        return n

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparams(), np.float32)
        lelph._check(_lib().lelph_nn_get_params(self.h, _fptr(out)), "lelph_nn_get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams()
        lelph._check(_lib().lelph_nn_set_params(self.h, _fptr(p)), "lelph_nn_set_params")

    def set_coarse(self, qgrid, fields: np.ndarray, scale: float):
        """fields (nqc, nfields, N) complex: short-range cartesian potentials
        of the coarse grid in its crystal order (i, j, k), k fastest"""
        g = np.ascontiguousarray(qgrid, np.int64)
        f = np.ascontiguousarray(fields, np.complex128)
        assert f.shape == (int(np.prod(qgrid)), self.nfields, self.n), f.shape
        self._coarse = f
        lelph._check(_lib().lelph_nn_set_coarse(self.h, g.ctypes.data_as(C.POINTER(C.c_int64)),
                                                dptr(f), float(scale)), "lelph_nn_set_coarse")

    def features(self, qpts):
        q = np.ascontiguousarray(np.atleast_2d(qpts), np.float64)
        S = len(q)
        X0 = np.zeros((S, self.n, NN_CIN), np.float32)
        base = np.zeros((S, self.n, self.nout), np.float32)
        lelph._check(_lib().lelph_nn_features(self.h, S, dptr(q), _fptr(X0), _fptr(base)),
                     "lelph_nn_features")
        return X0, base

    def forward(self, X0, base) -> np.ndarray:
        X0 = np.ascontiguousarray(X0, np.float32)
        base = np.ascontiguousarray(base, np.float32)
        out = np.zeros_like(base)
        lelph._check(_lib().lelph_nn_forward(self.h, X0.shape[0], _fptr(X0), _fptr(base),
                                             _fptr(out)), "lelph_nn_forward")
        return out

    def loss_grad(self, X0, base, target):
        X0 = np.ascontiguousarray(X0, np.float32)
        base = np.ascontiguousarray(base, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        loss = C.c_double(0)
        g = np.zeros(self.nparams(), np.float32)
        lelph._check(_lib().lelph_nn_loss_grad(self.h, X0.shape[0], _fptr(X0), _fptr(base),
                                               _fptr(target), C.byref(loss), _fptr(g)),
                     "lelph_nn_loss_grad")
        return loss.value, g

    def set_targets(self, targets: np.ndarray):
        t = np.ascontiguousarray(targets, np.float32)
        assert t.shape[1:] == (self.n, self.nout)
        lelph._check(_lib().lelph_nn_set_targets(self.h, t.shape[0], _fptr(t)),
                     "lelph_nn_set_targets")

    def train(self, qpts, epochs: int, batch: int = 8, lr: float = 2e-3, seed: int = 1):
        q = np.ascontiguousarray(qpts, np.float64)
        hist = np.zeros(int(epochs))
        lelph._check(_lib().lelph_nn_train(self.h, len(q), dptr(q), int(epochs), int(batch),
                                           float(lr), int(seed), dptr(hist)), "lelph_nn_train")
        return hist

    def predict(self, qpts) -> np.ndarray:
        q = np.ascontiguousarray(np.atleast_2d(qpts), np.float64)
        out = np.zeros((len(q), self.n, self.nout), np.float32)
        lelph._check(_lib().lelph_nn_predict(self.h, len(q), dptr(q), _fptr(out)),
                     "lelph_nn_predict")
        return out


def fine_grid(n: int) -> np.ndarray:
    """the n x n x n q grid in crystal units, [0, 1), k fastest"""
    a = np.arange(n) / n
    return np.stack(np.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3)


def _phase(sys_, qpts, sign):
    """e^{sign 2 pi i q.x} on the FFT grid, (nq, N)"""
    dims = sys_.fft_dims
    ax = [np.arange(n) / n for n in dims]
    x = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    return np.exp(sign * 2j * math.pi * (np.atleast_2d(qpts) @ x.T))


def to_channels(u: np.ndarray) -> np.ndarray:
    """(nq, nf, N) complex -> (nq, N, 2 nf) float32, re/im interleaved"""
    nq, nf, n = u.shape
    out = np.empty((nq, n, 2 * nf), np.float32)
    out[:, :, 0::2] = u.real.transpose(0, 2, 1)
    out[:, :, 1::2] = u.imag.transpose(0, 2, 1)
    return out


def from_channels(c: np.ndarray) -> np.ndarray:
    """(nq, N, 2 nf) -> (nq, nf, N) complex128"""
    return (c[:, :, 0::2] + 1j * c[:, :, 1::2]).transpose(0, 2, 1).astype(np.complex128)


class LearnedInterpolation:
    """The c4 pipeline: a context whose Fourier interpolation is set up
    (lelph_interp_begin / add_coarse / finish done), its long-range
    parameters, and a Network trained on the interpolation's own fine-q
    potentials."""

    def __init__(self, ctx, qgrid, lr_params, depth: int = 5, seed: int = 0):
        self.ctx = ctx
        self.sys = ctx.sys
        self.qgrid = tuple(int(x) for x in qgrid)
        self.lrp = lr_params
        s = self.sys
        assert s.nmag == 1, "one potential component"
        self.nf = s.nmodes
        self.net = Network(s.fft_dims, self.nf, depth=depth, seed=seed)
        coarse = fine_grid(self.qgrid[0]) if len(set(self.qgrid)) == 1 else None
        assert coarse is not None, "cubic coarse grids"
        sr = self.short_range(coarse)
        self.scale = float(np.sqrt(np.mean(np.abs(sr) ** 2)))
        self.net.set_coarse(self.qgrid, sr, self.scale)

    def interpolated(self, qpts) -> np.ndarray:
        """the Fourier interpolation's dV (nq, nf, N), cartesian basis"""
        d = self.ctx.interp_dvscf(qpts)
        return d.reshape(len(np.atleast_2d(qpts)), self.nf, -1)

    def long_range(self, q) -> np.ndarray:
        s = self.sys
        z = np.zeros((s.nmodes, s.nmag) + tuple(s.fft_dims), np.complex128)
        return self.ctx.dv_longrange(q, self.lrp, z, eig=None, sign=1).reshape(self.nf, -1)

    def short_range(self, qpts) -> np.ndarray:
        """interpolated dV minus its long range, (nq, nf, N) (not periodic)"""
        q = np.atleast_2d(qpts)
        d = self.interpolated(q)
        for i in range(len(q)):
            d[i] -= self.long_range(q[i])
        return d

    def periodic(self, qpts) -> np.ndarray:
        """u_q = e^{-2 pi i q.x} (dV_q - dV_q^LR), (nq, nf, N)"""
        q = np.atleast_2d(qpts)
        return self.short_range(q) * _phase(self.sys, q, -1)[:, None, :]

    def train(self, qpts, epochs: int, batch: int = 8, lr: float = 2e-3, seed: int = 1):
        self.net.set_targets(to_channels(self.periodic(qpts)))
        return self.net.train(qpts, epochs, batch=batch, lr=lr, seed=seed)

    def predict_periodic(self, qpts) -> np.ndarray:
        return from_channels(self.net.predict(qpts))

    def trilinear_periodic(self, qpts) -> np.ndarray:
        """the untrained model's answer (the trilinear blend of the 8 coarse
        periodic parts around q)"""
        _, base = self.net.features(qpts)
        return from_channels(base) * self.scale

    def dv(self, qpts, periodic: np.ndarray) -> np.ndarray:
        """dV (nq, nf, N) from periodic parts: e^{2 pi i q.x} u + long range"""
        q = np.atleast_2d(qpts)
        d = periodic * _phase(self.sys, q, 1)[:, None, :]
        for i in range(len(q)):
            d[i] += self.long_range(q[i])
        return d


def judge(gctx, gsys, qpt, eig, dv_model, dv_ref, pairs, tol=1e-3):
    """g on `pairs` from the model's dV and from the interpolation's dV
    (both cartesian, (nf, N)) on a context of the fine-k system; returns
    (|dg|max/|g|max, accepted)."""
    eye = np.eye(gsys.nmodes, dtype=np.complex128)
    shape = (gsys.nmodes, gsys.nmag) + tuple(gsys.fft_dims)
    out = []
    for dv in (dv_model, dv_ref):
        gctx.set_dvscf(qpt, eig, np.ascontiguousarray(dv.reshape(shape)), patterns=eye,
                       kernel="dfpt")
        out.append(np.stack([gctx.gkq(*p, qpt=qpt) for p in pairs]))
    err = float(np.max(np.abs(out[0] - out[1])) / np.max(np.abs(out[1])))
    return err, err <= tol
