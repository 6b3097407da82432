"""numpy restatement of the social-network classifier (Common.cuh:100-250,
CSC.cuh:176,187-217).  TEST INFRASTRUCTURE ONLY: the checker for blest_csc_is_social."""
from __future__ import annotations

import math

import numpy as np


def social_degree_distribution(deg) -> bool:
    deg = np.asarray(deg, dtype=np.int64)
    n = deg.size
    if n == 0 or deg.max(initial=0) == 0 or deg.sum() == 0:
        return False
    total = float(deg.sum())
    s = np.sort(deg)[::-1]

    def share(frac):
        k = int(math.ceil(frac * n))
        k = min(max(k, 1), n)
        return float(s[:k].sum()) / total

    shares = share(0.01) >= 0.05 and share(0.10) >= 0.40
    freq = np.bincount(deg)
    ks = np.nonzero(freq)[0]
    ks = ks[ks >= 5]
    power = False
    if ks.size >= 3:
        x = np.log(ks.astype(np.float64))
        y = np.log(freq[ks].astype(np.float64))
        m = float(ks.size)
        sx, sy, sxy, sxx = x.sum(), y.sum(), (x * y).sum(), (x * x).sum()
        den = m * sxx - sx * sx
        if den > 0:
            slope = (m * sxy - sx * sy) / den
            icpt = (sy - slope * sx) / m
            tot = ((y - sy / m) ** 2).sum()
            res = ((y - (icpt + slope * x)) ** 2).sum()
            if tot > 0:
                r2, gamma = 1.0 - res / tot, -slope
                power = r2 >= 0.70 and 1.0 <= gamma <= 5.0
    return shares or power


def is_social(n: int, col_ptrs, rows) -> bool:
    cp = np.asarray(col_ptrs, dtype=np.int64)
    out_deg = np.diff(cp)
    in_deg = np.bincount(np.asarray(rows, dtype=np.int64), minlength=n)
    avg = float(cp[-1]) / n if n else 0.0
    return social_degree_distribution(in_deg) or social_degree_distribution(out_deg) or avg > 18.0
