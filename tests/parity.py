"""Shared parity helpers for the tests.

Tolerance of the plain FP32 path against the reference (float64 arithmetic):
  * identical NaN pattern (the rule |P|, |S| > DBL_EPSILON, cpp:1437-1443);
  * max |dHU| <= 0.5 on the fused value before rounding, mapped to HU with
    HU = 4095*v/255 - 1024 (the final output is quantised in 16.06 HU steps,
    so parity is taken before rounding);
  * max |dv| / |v_ref| <= 1e-4 on the fused value in label units.
"""
from __future__ import annotations

import numpy as np

MAX_DHU = 0.5
MAX_REL = 1e-4


def hu_raw(v):
    return 4095.0 * (np.asarray(v, np.float64) / 255.0) - 1024.0


def compare(got, ref, max_dhu=MAX_DHU, max_rel=MAX_REL, nan_mismatch_allowed=0):
    """Returns a dict of statistics; raises AssertionError outside tolerance."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    gn, rn = np.isnan(got), np.isnan(ref)
    mism = int(np.count_nonzero(gn != rn))
    ok = ~(gn | rn)
    d = np.abs(hu_raw(got[ok]) - hu_raw(ref[ok]))
    rv = np.abs(np.asarray(ref[ok], np.float64))
    rel = np.abs(np.asarray(got[ok], np.float64) - np.asarray(ref[ok], np.float64)) / \
        np.maximum(rv, 1e-30)
    rel = np.where(rv > 0, rel, np.abs(np.asarray(got[ok], np.float64)))
    stats = {
        "n": int(got.size), "nan_ref": int(rn.sum()), "nan_mismatch": mism,
        "max_dhu": float(d.max()) if d.size else 0.0,
        "mean_dhu": float(d.mean()) if d.size else 0.0,
        "max_rel": float(rel.max()) if rel.size else 0.0,
    }
    assert mism <= nan_mismatch_allowed, f"NaN pattern differs: {stats}"
    assert stats["max_dhu"] <= max_dhu, f"|dHU| too large: {stats}"
    assert stats["max_rel"] <= max_rel, f"relative error too large: {stats}"
    return stats
