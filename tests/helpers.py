"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_err(a, b, floor_frac=1e-13):
    """max |a-b| / (|b| + floor_frac * max|b|): relative error with a floor for entries
    that are cancellations near zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    return float(np.max(np.abs(a - b) / (np.abs(b) + floor_frac * scale + 1e-300))) if b.size else 0.0


def bcrs_to_dense(rowptr, colidx, vals, N):
    A = np.zeros((2 * N, 2 * N))
    for i in range(N):
        for p in range(rowptr[i], rowptr[i + 1]):
            j = colidx[p]
            A[2 * i:2 * i + 2, 2 * j:2 * j + 2] = vals[4 * p:4 * p + 4].reshape(2, 2)
    return A
