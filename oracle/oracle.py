"""ctypes front end of the CPU oracle (liboracle.so, built from this folder).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by bench.py's cpu_baseline leg, always as the checker or the CPU baseline, never
as the product.  The C sources cite the reference lines they restate.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

BRACKET_REFERENCE = 0
BRACKET_EXACT = 1

_dp = ctypes.POINTER(ctypes.c_double)


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i, d = ctypes.c_int, ctypes.c_double
        L.oracle_beta.restype = d
        L.oracle_beta.argtypes = [i, d, d, d, d, d]
        L.oracle_selection.restype = d
        L.oracle_selection.argtypes = [i, d, d, d, d, d]
        L.oracle_aggregation_rhs.restype = i
        L.oracle_aggregation_rhs.argtypes = [i, _dp, _dp, i, d, d, d, d, d, i,
                                             _dp, _dp, _dp]
        L.oracle_aggregation_rhs_threads.restype = i
        L.oracle_aggregation_rhs_threads.argtypes = [i, _dp, _dp, i, d, d, d, d, d, i,
                                                     _dp, _dp, _dp, i]
        L.oracle_breakage_rhs.restype = i
        L.oracle_breakage_rhs.argtypes = [i, i, _dp, _dp, _dp, _dp, i, d, d, d,
                                          d, _dp, _dp, _dp]
        L.oracle_rk4.restype = i
        L.oracle_rk4.argtypes = [ctypes.c_void_p, _dp, d, i]
        L.port_aggregation_rhs.restype = i
        L.port_aggregation_rhs.argtypes = [i, _dp, _dp, i, d, d, d, d, d, _dp, i]
        L.port_breakage_rhs.restype = i
        L.port_breakage_rhs.argtypes = [i, i, _dp, _dp, _dp, _dp, i, d, d, d, d,
                                        _dp, i]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle returned error {code}")
        self.code = code


def aggregation_rhs(N, x, kernel, beta0=1.0, beta_br=0.0, beta_sh=0.0,
                    log_x0=0.0, inv_log_r=0.0, bracket=BRACKET_EXACT, n=None, threads=1):
    """Return (rhs, A, m1_defect) of one aggregation evaluation from zero rhs.

    A[i] is the sum of |contributions| landing in bin i; m1_defect is the exact
    sum_i x_i rhs_i produced by boundary clipping.  threads > 1 splits the rows
    over OpenMP threads (same sums to double-double precision).
    """
    N, x = _f64(N), _f64(x)
    n = len(N) if n is None else n
    rhs = np.zeros(max(n, 1))
    A = np.zeros(max(n, 1))
    defect = ctypes.c_double(0.0)
    e = lib().oracle_aggregation_rhs_threads(n, _p(N), _p(x), int(kernel), beta0,
                                             beta_br, beta_sh, log_x0, inv_log_r,
                                             int(bracket), _p(rhs), _p(A),
                                             ctypes.byref(defect), int(threads))
    if e:
        raise OracleError(e)
    return rhs, A, defect.value


def breakage_rhs(N, x, t_q, bw_q, selection, S0=1.0, v_ref=1.0, alpha=1.0,
                 v_min=0.0, n=None, n_quad=None):
    """Return (rhs, A, m1_defect) of one breakage evaluation from zero rhs.

    m1_defect is the exact sum_i x_i rhs_i (daughter-moment and clip defects).
    """
    N, x, t_q, bw_q = _f64(N), _f64(x), _f64(t_q), _f64(bw_q)
    n = len(N) if n is None else n
    nq = len(t_q) if n_quad is None else n_quad
    rhs = np.zeros(max(n, 1))
    A = np.zeros(max(n, 1))
    defect = ctypes.c_double(0.0)
    e = lib().oracle_breakage_rhs(n, nq, _p(N), _p(x), _p(t_q), _p(bw_q),
                                  int(selection), S0, v_ref, alpha, v_min,
                                  _p(rhs), _p(A), ctypes.byref(defect))
    if e:
        raise OracleError(e)
    return rhs, A, defect.value


def beta(kernel, xj, xk, b0=1.0, b_br=0.0, b_sh=0.0):
    return lib().oracle_beta(int(kernel), xj, xk, b0, b_br, b_sh)


def selection(sel, v, S0=1.0, v_ref=1.0, alpha=1.0, v_min=0.0):
    return lib().oracle_selection(int(sel), v, S0, v_ref, alpha, v_min)


class _Problem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("x", _dp), ("use_agg", ctypes.c_int),
        ("kernel", ctypes.c_int), ("beta0", ctypes.c_double),
        ("beta_br", ctypes.c_double), ("beta_sh", ctypes.c_double),
        ("log_x0", ctypes.c_double), ("inv_log_r", ctypes.c_double),
        ("bracket_mode", ctypes.c_int), ("use_brk", ctypes.c_int),
        ("n_quad", ctypes.c_int), ("t_q", _dp), ("bw_q", _dp),
        ("selection", ctypes.c_int), ("S0", ctypes.c_double),
        ("v_ref", ctypes.c_double), ("alpha", ctypes.c_double),
        ("v_min", ctypes.c_double), ("threads", ctypes.c_int),
    ]


def rk4(N0, x, dt, steps, agg=None, brk=None, t_q=None, bw_q=None,
        bracket=BRACKET_EXACT, threads=1):
    """Classic RK4 of the reference harness; agg/brk are dicts of parameters
    (keys as in aggregation_rhs / breakage_rhs); threads > 1 splits each
    aggregation evaluation over OpenMP threads."""
    N = _f64(N0).copy()
    x = _f64(x)
    keep = [x]
    pr = _Problem()
    pr.n = len(N)
    pr.x = _p(x)
    pr.bracket_mode = bracket
    pr.threads = int(threads)
    if agg:
        pr.use_agg = 1
        pr.kernel = int(agg["kernel"])
        pr.beta0 = agg.get("beta0", 1.0)
        pr.beta_br = agg.get("beta_br", 0.0)
        pr.beta_sh = agg.get("beta_sh", 0.0)
        pr.log_x0 = agg.get("log_x0", 0.0)
        pr.inv_log_r = agg.get("inv_log_r", 0.0)
    if brk:
        tq, bq = _f64(t_q), _f64(bw_q)
        keep += [tq, bq]
        pr.use_brk = 1
        pr.n_quad = len(tq)
        pr.t_q, pr.bw_q = _p(tq), _p(bq)
        pr.selection = int(brk["selection"])
        pr.S0 = brk.get("S0", 1.0)
        pr.v_ref = brk.get("v_ref", 1.0)
        pr.alpha = brk.get("alpha", 1.0)
        pr.v_min = brk.get("v_min", 0.0)
    e = lib().oracle_rk4(ctypes.byref(pr), _p(N), dt, int(steps))
    if e:
        raise OracleError(e)
    return N


def port_aggregation_rhs(N, x, kernel, beta0=1.0, beta_br=0.0, beta_sh=0.0,
                         log_x0=0.0, inv_log_r=0.0, threads=1):
    N, x = _f64(N), _f64(x)
    rhs = np.zeros(len(N))
    e = lib().port_aggregation_rhs(len(N), _p(N), _p(x), int(kernel), beta0,
                                   beta_br, beta_sh, log_x0, inv_log_r,
                                   _p(rhs), int(threads))
    if e:
        raise OracleError(e)
    return rhs


def port_breakage_rhs(N, x, t_q, bw_q, selection, S0=1.0, v_ref=1.0,
                      alpha=1.0, v_min=0.0, threads=1):
    N, x, t_q, bw_q = _f64(N), _f64(x), _f64(t_q), _f64(bw_q)
    rhs = np.zeros(len(N))
    e = lib().port_breakage_rhs(len(N), len(t_q), _p(N), _p(x), _p(t_q),
                                _p(bw_q), int(selection), S0, v_ref, alpha,
                                v_min, _p(rhs), int(threads))
    if e:
        raise OracleError(e)
    return rhs


# ------------------------------------------------------------ complex step --
class _CsProblem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("x", _dp), ("use_agg", ctypes.c_int),
                ("kernel", ctypes.c_int), ("use_brk", ctypes.c_int), ("nq", ctypes.c_int),
                ("t_q", _dp), ("bw_q", _dp), ("sel", ctypes.c_int),
                ("v_ref", ctypes.c_double), ("v_min", ctypes.c_double),
                ("threads", ctypes.c_int)]


def _cs_setup(x, agg, brk, t_q, bw_q, threads):
    x = _f64(x)
    keep = [x]
    p = _CsProblem()
    p.n = len(x)
    p.x = _p(x)
    p.threads = int(threads)
    th = np.zeros(5)
    if agg:
        p.use_agg, p.kernel = 1, int(agg["kernel"])
        th[0], th[1], th[2] = agg.get("beta0", 0.0), agg.get("beta_br", 0.0), agg.get("beta_sh", 0.0)
    if brk:
        tq, bq = _f64(t_q), _f64(bw_q)
        keep += [tq, bq]
        p.use_brk, p.nq, p.t_q, p.bw_q = 1, len(tq), _p(tq), _p(bq)
        p.sel = int(brk["selection"])
        p.v_ref, p.v_min = brk.get("v_ref", 1.0), brk.get("v_min", 0.0)
        th[3], th[4] = brk.get("S0", 1.0), brk.get("alpha", 1.0)
    return p, th, keep


def _cs_fns():
    L = lib()
    if not hasattr(L, "_cs_ready"):
        L.cs_rhs.restype = ctypes.c_int
        L.cs_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.cs_rk4_loss.restype = ctypes.c_int
        L.cs_rk4_loss.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_double,
                                  ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
        L._cs_ready = True
    return L


THETA_NAMES = ("beta0", "beta_br", "beta_sh", "S0", "alpha")


def cs_rhs(N, x, agg=None, brk=None, t_q=None, bw_q=None, dN=None, dtheta=None,
           h=1e-30, threads=1):
    """Complex-step directional derivative of R at (N, theta) along
    (dN, dtheta): returns (R, dR) with dR = Im R(N + i h dN, theta + i h dtheta)/h."""
    p, th, keep = _cs_setup(x, agg, brk, t_q, bw_q, threads)
    n = len(N)
    Nim = np.zeros(n) if dN is None else h * _f64(dN)
    thim = np.zeros(5) if dtheta is None else h * _f64(dtheta)
    ore, oim = np.zeros(n), np.zeros(n)
    e = _cs_fns().cs_rhs(ctypes.byref(p), _p(th), _p(thim), _p(_f64(N)), _p(Nim),
                         _p(ore), _p(oim))
    if e:
        raise OracleError(e)
    return ore, oim / h


def cs_loss(N0, x, dt, steps, logT, mask, eps, agg=None, brk=None, t_q=None, bw_q=None,
            dN0=None, dtheta=None, h=1e-30, threads=1):
    """Complex-step derivative of the RK4 log-misfit loss along (dN0, dtheta):
    returns (L, dL)."""
    p, th, keep = _cs_setup(x, agg, brk, t_q, bw_q, threads)
    n = len(N0)
    Nim = np.zeros(n) if dN0 is None else h * _f64(dN0)
    thim = np.zeros(5) if dtheta is None else h * _f64(dtheta)
    Lr, Li = ctypes.c_double(), ctypes.c_double()
    e = _cs_fns().cs_rk4_loss(ctypes.byref(p), _p(th), _p(thim), _p(_f64(N0)), _p(Nim),
                              dt, int(steps), _p(_f64(logT)), _p(_f64(mask)), eps,
                              ctypes.byref(Lr), ctypes.byref(Li), None, None)
    if e:
        raise OracleError(e)
    return Lr.value, Li.value / h


def cs_rk4_final(N0, x, dt, steps, agg=None, brk=None, t_q=None, bw_q=None, threads=1):
    """Real RK4 final state through the complex-step code path (imaginary
    parts zero): the same trajectory cs_loss differentiates."""
    p, th, keep = _cs_setup(x, agg, brk, t_q, bw_q, threads)
    n = len(N0)
    zero = np.zeros(n)
    Lr, Li = ctypes.c_double(), ctypes.c_double()
    fre, fim = np.zeros(n), np.zeros(n)
    e = _cs_fns().cs_rk4_loss(ctypes.byref(p), _p(th), _p(np.zeros(5)), _p(_f64(N0)), _p(zero),
                              dt, int(steps), _p(zero), _p(zero), 1.0,
                              ctypes.byref(Lr), ctypes.byref(Li), _p(fre), _p(fim))
    if e:
        raise OracleError(e)
    return fre
