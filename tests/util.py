"""Shared helpers of the test suite."""
import numpy as np


def rel_err(a, b) -> float:
    """|a - b|_max / |b|_max, the parity metric of the e-ph path."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def say(msg: str):
    print(msg, flush=True)


def tr_partner(s, ik, isym):
    """A time-reversed operation j (index >= nsym/2) with S_j k = S_isym k
    (mod G) for iBZ k ik, or None.  A k-grid map that reaches every star by
    a rotation first never picks one, so pairs through the
    time-reversed branch are built with it: the same full-BZ k, its
    wavefunction obtained as the conjugate of a rotated one."""
    half = len(s.syms) // 2
    from letzelphc_amd import synth
    Sc = [synth._sym_crys(np.asarray(R), s.alat, s.blat) for (R, _, _) in s.syms]
    k = s.kpt_ibz_crys[ik]
    want = Sc[isym] @ k
    for j in range(half, len(s.syms)):
        d = Sc[j] @ k - want
        if np.linalg.norm(d - np.rint(d)) < 1e-5:
            return j
    return None
