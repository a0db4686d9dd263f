"""Parity metric shared by the tests.

Per entry: err_i = |got_i - ref_i| / A_i, where A_i is the oracle's sum of
absolute contributions to bin i (entries with A_i = 0 must be exactly 0).
Entries whose contributions are formed through products below the FP64
normal range (N ~ 1e-300 next to rates ~1e11) depend on the order of the
multiplications at the underflow level, so the comparison carries an absolute
floor of 2^-1000 (~1e-301) times the largest A_j of the problem -- about 290
orders of magnitude below the largest entry.
Volume: |sum_i x_i got_i - D| <= tol * sum_i x_i A_i, with D the exact first-
moment defect the oracle reports (boundary clipping, daughter moments).
"""
import numpy as np

RHS_TOL = 1e-12      # per-entry bound on |gpu - oracle| / A_i
M1_TOL = 1e-13       # first-moment bound relative to sum x_i A_i


UNDERFLOW_FLOOR = 2.0 ** -1000


def entry_error(got, ref, A):
    got, ref, A = (np.asarray(v, dtype=np.float64) for v in (got, ref, A))
    nz = A > 0
    err = np.zeros_like(ref)
    floor = UNDERFLOW_FLOOR * (float(A.max()) if A.size else 0.0)
    diff = np.maximum(np.abs(got - ref) - floor, 0.0)
    err[nz] = diff[nz] / A[nz]
    zero_ok = bool(np.all(got[~nz] == 0.0))
    return float(err.max()) if err.size else 0.0, zero_ok


def assert_parity(got, ref, A, tol=RHS_TOL, what=""):
    worst, zero_ok = entry_error(got, ref, A)
    assert zero_ok, f"{what}: nonzero rhs where the oracle has no contribution"
    if worst > tol:
        got, ref, A = (np.asarray(v, dtype=np.float64) for v in (got, ref, A))
        e = np.where(A > 0, np.abs(got - ref) / np.where(A > 0, A, 1.0), 0.0)
        i = int(np.argmax(e))
        detail = f" (bin {i}: got {got[i]:.17g} ref {ref[i]:.17g} A {A[i]:.3e})"
    else:
        detail = ""
    assert worst <= tol, f"{what}: max |gpu-oracle|/A = {worst:.3e} > {tol:.1e}{detail}"
    return worst


def assert_m1(got, x, A, defect, tol=M1_TOL, what=""):
    x = np.asarray(x, dtype=np.float64)
    scale = float(np.sum(x * A))
    if scale == 0.0:
        return 0.0
    m1 = float(np.sum(x * np.asarray(got)))
    rel = abs(m1 - defect) / scale
    assert rel <= tol, f"{what}: |sum x rhs - defect|/sum x A = {rel:.3e} > {tol:.1e}"
    return rel
