"""Error metrics of the parity tests (shared helpers, no test functions).

normwise(x, ref)  per matrix ||x - ref||_2 / ||ref||_2 over its entries (Frobenius); NaN positions must agree
                  (both NaN counts as agreement, as cuda/tests/test_cuda_api.cu:18); a mismatch gives inf
maxrel(x, ref)    the reference tests' batch metric max|x - ref| / max|ref| (c/tests/test_c.c:18-25, NaN pairs skipped)
"""
import numpy as np


def normwise(x, ref):
    x = np.asarray(x, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    nx, nr = np.isnan(x), np.isnan(ref)
    mism = (nx != nr).any(axis=1)
    d = np.where(nr, 0.0, x - ref)
    r = np.where(nr, 0.0, ref)
    num = np.sqrt((d * d).sum(axis=1))
    den = np.sqrt((r * r).sum(axis=1))
    err = np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
    err[mism] = np.inf
    return err


def maxrel(x, ref):
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    both = np.isnan(x) & np.isnan(ref)
    d = np.abs(x - ref)[~both]
    if np.isnan(d).any():
        return np.inf
    s = np.abs(ref[~both]).max() if (~both).any() else 0.0
    return d.max() / s if d.size else 0.0
