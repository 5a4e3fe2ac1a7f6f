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


def spread_check(name, got, ora, refv, tol=1e-12, factor=32.0):
    """per-matrix normwise error of `got` against the oracle, bounded per matrix by max(tol, factor s) with s the
    difference between two correct builds of the reference on that matrix -- its SIMD build (libmvec; refv) against
    its scalar build (= the oracle).  Without the reference build (refv None): p99.9 <= 2 tol and max <= 1e-10.
    Prints the distributions (max, p99.9, share of matrices above tol) of both as a line starting with "DIST"."""
    e = normwise(got, ora)
    line = (f"DIST {name}: n {len(e)} max {e.max():.2e} p99.9 {np.quantile(e, 0.999):.2e} "
            f"above {tol:g} {np.mean(e > tol):.3%}")
    if refv is None:
        print(line + " | reference SIMD build not available", flush=True)
        assert np.quantile(e, 0.999) <= 2 * tol and e.max() <= 1e-10, (name, e.max(), np.quantile(e, 0.999))
        return e
    s = normwise(refv, ora)
    print(line + f" | reference SIMD vs scalar: max {s.max():.2e} p99.9 {np.quantile(s, 0.999):.2e} "
                 f"above {tol:g} {np.mean(s > tol):.3%}", flush=True)
    bound = np.maximum(tol, factor * s)
    bad = np.nonzero(e > bound)[0]
    assert bad.size == 0, (f"{name}: {bad.size} matrices over tolerance, first {bad[0]}: {e[bad[0]]:.2e} "
                           f"(spread {s[bad[0]]:.2e})")
    return e
