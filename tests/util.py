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
