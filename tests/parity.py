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


def half_pair_keys(orc_lib, system, cap_per_atom=128):
    """Sorted keys i * N + j (i < j) of the half lists held in a system's
    AoS (any library with the seppart layout), read by the oracle's
    orc_half_pairs so that large systems need no Python loop."""
    n = system.n
    cap = n * cap_per_atom
    buf = np.empty(2 * cap, dtype=np.int32)
    c = orc_lib.orc_half_pairs(system.atoms, n, buf.ctypes.data, cap)
    assert c >= 0, "more than %d pairs per atom" % cap_per_atom
    a = buf[:2 * c].reshape(-1, 2).astype(np.int64)
    key = a[:, 0] * n + a[:, 1]
    key.sort()
    return key


def pair_keys(pairs, n):
    """Sorted keys of an (m, 2) array of (i, j) pairs with i < j."""
    key = pairs[:, 0].astype(np.int64) * n + pairs[:, 1].astype(np.int64)
    key.sort()
    return key
