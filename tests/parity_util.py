"""TEST INFRASTRUCTURE: comparison helpers for HIP-path vs oracle parity tests."""
from __future__ import annotations

import numpy as np

STATE_FIELDS = ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "totsed", "maxvel", "maxdepth", "velx", "vely")


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b|| / ||b|| over the finite entries of b (NaN positions must agree)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb, na = np.isnan(b), np.isnan(a)
    if not np.array_equal(na, nb):
        return float("inf")
    m = ~nb
    den = np.linalg.norm(b[m])
    num = np.linalg.norm(a[m] - b[m])
    if den == 0:
        return 0.0 if num == 0 else float("inf")
    return float(num / den)


def compare_states(gpu: dict, ora: dict, nclass: int):
    """Relative L2 error per field (per class for ch and m)."""
    out = {}
    for k in STATE_FIELDS:
        out[k] = rel_l2(gpu[k], ora[k])
    for c in range(nclass):
        out["ch%d" % (c + 1)] = rel_l2(gpu["ch"][c], ora["ch"][c])
        out["m%d" % (c + 1)] = rel_l2(gpu["m"][c], ora["m"][c])
    return out


def worst(errs: dict):
    k = max(errs, key=lambda x: errs[x])
    return k, errs[k]
