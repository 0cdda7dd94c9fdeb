"""Shared helpers of the test suite."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def max_ulp(a, b):
    """Largest distance in units of the last place between two float64 arrays."""
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    ia = a.view(np.int64).astype(object) if a.size < 0 else a.view(np.int64)
    ib = b.view(np.int64)
    # map sign-magnitude to a monotone integer line
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    d = np.abs(ia - ib)
    return int(d.max()) if d.size else 0


def compare_fields(got: dict, ref: dict, names=None, exact=True, tol=0.0):
    """Returns {name: (equal, rel_l2, max_abs)} and a list of failing names."""
    out, bad = {}, []
    for n in names or ref.keys():
        a, b = got[n], ref[n]
        eq = np.array_equal(a, b)
        r = rel_l2(a, b)
        m = float(np.max(np.abs(a - b))) if a.size else 0.0
        out[n] = (eq, r, m)
        if (exact and not eq) or (not exact and not (r <= tol)):
            bad.append(n)
    return out, bad
