"""Shared helpers for the parity tests (CPU oracle vs HIP path)."""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps


def rel_fro(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def rank_tie(sigma_ref, r_ref, r_got, acc, truncated=False, slack=1e3):
    """True when a rank mismatch is explained by a singular value within rounding of the
    threshold: |sigma_i - thr| <= slack * eps * sigma_0 for some i between the two ranks."""
    if r_ref == r_got:
        return True
    s = np.asarray(sigma_ref, dtype=np.float64)
    if len(s) == 0:
        return False
    thr = acc * s[0] if truncated else acc
    lo, hi = min(r_ref, r_got), max(r_ref, r_got)
    win = slack * EPS * s[0]
    for i in range(max(lo - 1, 0), min(hi + 1, len(s))):
        if abs(s[i] - thr) <= win:
            return True
    return False
