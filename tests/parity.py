"""Comparison helpers with the tolerances of the parity bar.

Forces: max_i |dF_i| / RMS|F| (the north-star force metric).
Scalars and 3x3 tensors: |d| / max(|ref|, scale) with scale the magnitude of
the quantity's own terms, so values that cancel to ~0 are judged against
the size of what was summed.
"""
import numpy as np

FP64_TOL = 1e-12


def force_error(f, fref):
    rms = np.sqrt((fref ** 2).sum(axis=1).mean())
    if rms == 0.0:
        return float(np.abs(f - fref).max())
    return float(np.sqrt(((f - fref) ** 2).sum(axis=1)).max() / rms)


def rel(a, b, scale=0.0):
    d = abs(a - b)
    s = max(abs(b), scale)
    return d / s if s > 0 else d


def tensor_error(p, pref):
    s = np.abs(pref).max()
    return float(np.abs(p - pref).max() / s) if s > 0 else float(np.abs(p - pref).max())
