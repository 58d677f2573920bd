"""The no-allocation q-star rotation (lelph_elph_q_rotate_noalloc) through
the C-ABI on device pointers, against lelph_elph_q_rotate_gpu (the entry
point it is a variant of) on the same inputs: two calls in a row, the second
smaller, both time-reversal branches; and the size limit of its static
scratch.  Bar: FP64, |d|max/|ref|max <= 1e-10 (the kernels are the same, so
the results are expected to be bit-identical)."""
import ctypes as C

import numpy as np
import pytest

from letzelphc_amd import lelph, synth
from letzelphc_amd.abi import dptr

from .util import rel_err, say

pytestmark = pytest.mark.gpu


def _hip():
    lelph.lib()  # the runtime the library links, already loaded
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class _Dev:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, np.complex128)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def _fn():
    f = lelph.lib().lelph_elph_q_rotate_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                  C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def test_q_rotate_noalloc_matches():
    h = _hip()
    f = _fn()
    s = synth.tiny(nbnds=4)
    ctx = lelph.Context(s, "fp64")
    big = 6 * 2 * 70 * 70 * 16
    bufs = [_Dev(h, big) for _ in range(4)]  # Dl, g, Dr, out (sized for the first call)
    try:
        for (nm, ns, nb) in [(6, 2, 70), (4, 1, 40)]:
            rng = np.random.default_rng(100 + nb)
            g = rng.standard_normal((nm, ns, nb, nb)) + 1j * rng.standard_normal((nm, ns, nb, nb))
            Dl = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
            Dr = rng.standard_normal((ns, nb, nb)) + 1j * rng.standard_normal((ns, nb, nb))
            for tr in (0, 1):
                say(f"case: q-rotate no-alloc nm={nm} nspin={ns} nb={nb} tim_rev={tr}")
                ref = ctx.elph_q_rotate(Dl, g, Dr, bool(tr))
                bufs[0].put(Dl)
                bufs[1].put(g)
                bufs[2].put(Dr)
                rc = f(bufs[0].p, bufs[1].p, bufs[2].p, nm, ns, nb, tr, bufs[3].p, None)
                assert rc == 0, lelph.strerror(rc)
                assert h.hipDeviceSynchronize() == 0
                got = bufs[3].get(g.shape)
                err = rel_err(got, ref)
                say(f"  vs lelph_elph_q_rotate_gpu: {err:.3e} (bit-identical: "
                    f"{np.array_equal(got, ref)})")
                assert err <= 1e-10
        # beyond the scratch: a clear status, nothing launched
        rc = f(bufs[0].p, bufs[1].p, bufs[2].p, 1, 1, 2049, 0, bufs[3].p, None)
        assert rc == 9
        assert "scratch" in lelph.strerror(rc)
    finally:
        for b in bufs:
            b.free()
        ctx.close()
