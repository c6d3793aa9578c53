"""Shared helpers of the parity tests."""
import numpy as np


def eq_rel_err(a, b, nb, neq):
    """Equation-norm relative error: max_e,j |a - b| / max_e,j |b| per equation m, the max over m."""
    a = np.asarray(a).reshape(-1, neq, nb)
    b = np.asarray(b).reshape(-1, neq, nb)
    errs = []
    for m in range(neq):
        den = np.abs(b[:, m, :]).max()
        errs.append(np.abs(a[:, m, :] - b[:, m, :]).max() / (den if den > 0 else 1.0))
    return max(errs)


def block_rel_err(a, b):
    """max |a - b| / max |b| over all block entries."""
    return float(np.abs(a - b).max() / np.abs(b).max())
