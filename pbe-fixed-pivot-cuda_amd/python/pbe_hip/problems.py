"""Host-side builders for pivot grids, initial conditions and daughter tables.

The reference keeps these in its example harness, not in the library:
grids ``examples/common/pbe_examples_utils.hpp:83-99``, initial conditions
``:104-158``, daughter quadrature conventions ``include/pbe_cuda/breakage.cuh:43-50``
and ``docs/api.md`` (symmetric ``{0.5}/{2}``, erosion ``{1-eps, eps}/{1, 1}``,
uniform ``GL nodes / 2*GL weights``, power law ``2 x GL order``).

Everything here is plain numpy on the host and produces float64 arrays that
are handed unchanged to both the HIP library and the CPU oracle.
"""
import math

import numpy as np

# ---------------------------------------------------------------- grids ----


def geometric_grid(n, x0, r):
    """x[0] = x0, x[i] = x[i-1]*r by repeated multiplication, exactly as
    ``make_geometric_grid`` (pbe_examples_utils.hpp:83-90)."""
    x = np.empty(n)
    x[0] = x0
    for i in range(1, n):
        x[i] = x[i - 1] * r
    return x


def geometric_grid_range(n, v_min, v_max):
    """``make_geometric_grid_range`` (pbe_examples_utils.hpp:93-99)."""
    r = (v_max / v_min) ** (1.0 / (n - 1))
    return geometric_grid(n, v_min, r)


def grid_lookup(x):
    """(log_x0, inv_log_r) as the reference's callers compute them
    (include/pbe_cuda/aggregation.cuh:105-108)."""
    return math.log(x[0]), 1.0 / math.log(x[1] / x[0])


def jittered_grid(n, b0, r, jitter, seed):
    """Non-geometric grid: section boundaries b_i = b0*r^(i+d_i) with
    d_i ~ U[-jitter, jitter] (jitter < 0.5 keeps them increasing) and pivots
    x_i = sqrt(b_i b_{i+1}).  Callers pass inv_log_r = 0 for such a grid,
    which selects the binary-search bracket path."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-jitter, jitter, n + 1)
    b = b0 * np.power(r, np.arange(n + 1) + d)
    x = np.sqrt(b[:-1] * b[1:])
    assert np.all(np.diff(x) > 0)
    return x

# ------------------------------------------------------ initial conditions --


def exponential_ic_cells(x, N0, v0):
    """N_i = N0*exp(-x_i/v0)*dx_i/v0 with geometric-midpoint cell width
    dx_i = x_i(sqrt(r) - 1/sqrt(r)) (the cell convention of
    pbe_examples_utils.hpp:125-133)."""
    r = x[1] / x[0]
    sr = math.sqrt(r)
    dx = x * (sr - 1.0 / sr)
    return N0 * np.exp(-x / v0) * dx / v0


def exponential_ic_exact(x, N0, vc):
    """``make_exponential_ic`` (pbe_examples_utils.hpp:120-136): exact cell
    integral of (N0/vc) exp(-v/vc)."""
    r = x[1] / x[0]
    sr = math.sqrt(r)
    return N0 * (np.exp(-(x / sr) / vc) - np.exp(-(x * sr) / vc))


def lognormal_ic(x, x_med, sigma_ln, total, noise=0.0, seed=None):
    """Lognormal number density in volume, following the pattern of
    ``make_lognormal_ic_nd`` (pbe_examples_utils.hpp:142-158) with an explicit
    median: N_i ~ exp(-0.5((ln x_i - ln x_med)/sigma)^2) x_i ln r, optional
    seeded multiplicative noise U[1-noise, 1+noise], normalised to sum = total.
    Bins whose density underflows are exactly zero."""
    log_r = math.log(x[1] / x[0])
    z = (np.log(x) - math.log(x_med)) / sigma_ln
    N = np.exp(-0.5 * z * z) * x * log_r
    if noise:
        rng = np.random.default_rng(seed)
        N = N * rng.uniform(1.0 - noise, 1.0 + noise, len(x))
    return N * (total / N.sum())

# ------------------------------------------------------- daughter tables ---


def gauss_legendre(m, a=0.0, b=1.0):
    """m-point Gauss-Legendre nodes and weights on (a, b)."""
    t, w = np.polynomial.legendre.leggauss(m)
    return 0.5 * (b - a) * t + 0.5 * (b + a), 0.5 * (b - a) * w


def daughter_symmetric():
    """b(v|v') = 2 delta(v - v'/2): t = {0.5}, bw = {2} (breakage.cuh:47)."""
    return np.array([0.5]), np.array([2.0])


def daughter_erosion(eps):
    """b(v|v') = delta(v - eps v') + delta(v - (1-eps) v'):
    t = {1-eps, eps}, bw = {1, 1} (breakage.cuh:48; flocculation example)."""
    return np.array([1.0 - eps, eps]), np.array([1.0, 1.0])


def daughter_uniform(n_quad):
    """b(v|v') = 2/v': t = GL nodes on (0,1), bw = 2*GL weights
    (examples/breakage_linear_uniform/main.cpp:76-80)."""
    t, w = gauss_legendre(n_quad)
    return t, 2.0 * w


def daughter_power_law(q, n_quad):
    """b(v|v') = C(v') v^q normalised by volume conservation,
    C(v') = (q+2)/v'^(q+1), i.e. b = (q+2)/v' (v/v')^q, so that
    bw = (q+2) t^q w.  Integrated with two Gauss-Legendre panels (0, 1/2) and
    (1/2, 1) of n_quad/2 points each (breakage.cuh:50: n_quad = 2 x GL order).
    The reference documents this table but ships no code for it."""
    if n_quad % 2:
        raise ValueError("power-law daughter needs an even n_quad (2 panels)")
    m = n_quad // 2
    t1, w1 = gauss_legendre(m, 0.0, 0.5)
    t2, w2 = gauss_legendre(m, 0.5, 1.0)
    t = np.concatenate([t1, t2])
    w = np.concatenate([w1, w2])
    return t, (q + 2.0) * np.power(t, q) * w


def daughter_table(kind, n_quad=None, q=2.0, eps=0.05):
    kind = kind.lower()
    if kind in ("symmetric", "binary", "symmetric_binary"):
        return daughter_symmetric()
    if kind == "erosion":
        return daughter_erosion(eps)
    if kind == "uniform":
        return daughter_uniform(n_quad)
    if kind in ("powerlaw", "power_law"):
        return daughter_power_law(q, n_quad)
    raise ValueError(f"unknown daughter distribution {kind!r}")

# ------------------------------------------------------------ workloads ---


def config_smoluchowski(n=256):
    """Constant-kernel aggregation on x_i = 1e-18 2^(i/4), beta0 = 1e-12,
    N_i = N0 exp(-x_i/v0) dx_i/v0 with N0 = 1e12, v0 = 10 x_0."""
    r = 2.0 ** 0.25
    x = geometric_grid(n, 1e-18, r)
    N = exponential_ic_cells(x, 1e12, 10.0 * x[0])
    log_x0, inv_log_r = grid_lookup(x)
    agg = dict(kernel=0, beta0=1e-12, beta_br=0.0, beta_sh=0.0,
               log_x0=log_x0, inv_log_r=inv_log_r)
    return x, N, agg


def config_flocculation(n, seed=20261203, daughter="symmetric", n_quad=16,
                        G=100.0, r=1.02):
    """Single-instance aggregation + breakage: geometric grid r = 1.02 (or
    the ratio given; n ln r must stay below about 700) from x_0 = 1e-18, Brownian+shear kernel beta_br = 6.73e-18, beta_sh = 4/3 G,
    threshold breakage S0 = 1e-3 with v_min = x_{n/2}, symmetric-binary
    daughters (or a GL uniform variant), lognormal N (median x_{n/4},
    sigma_ln = 0.8, total 1e12, +-5 % seeded noise)."""
    x = geometric_grid(n, 1e-18, r)
    N = lognormal_ic(x, x[n // 4], 0.8, 1e12, noise=0.05, seed=seed)
    log_x0, inv_log_r = grid_lookup(x)
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=4.0 / 3.0 * G,
               log_x0=log_x0, inv_log_r=inv_log_r)
    brk = dict(selection=3, S0=1e-3, v_ref=1.0, alpha=1.0, v_min=x[n // 2])
    if daughter == "symmetric":
        t_q, bw_q = daughter_symmetric()
    else:
        t_q, bw_q = daughter_table(daughter, n_quad)
    return x, N, agg, brk, t_q, bw_q


def pair_units(n, n_quad):
    """Work units of one aggregation + breakage evaluation: n(n+1)/2 pairs
    plus n*n_quad breakage (parent, quadrature point) updates."""
    return n * (n + 1) // 2 + n * n_quad


# Algorithmic FP64 operations per pair for each aggregation kernel (FMA = 2),
# and per breakage (parent, quad) update / per parent.
PAIR_FLOPS = {0: 11, 1: 12, 2: 13, 3: 16, 4: 15, 5: 21}
BREAKAGE_UPDATE_FLOPS = 8
BREAKAGE_PARENT_FLOPS = 2


def birth_pairs(x):
    """Pairs j <= k of the grid x (increasing) whose birth weight is not
    exactly zero: fl(x_j + x_k) != x_k.  A pair with fl(x_j + x_k) == x_k
    (x_j below half an ulp of x_k) lands wholly in bin k, where its birth
    cancels the column's loss, so it only takes its rate from bin j: a
    separable death, O(n) work for all such pairs together.  For each k the
    pairs form a suffix j >= j0(k) (the sum grows with x_j); j0 by a
    vectorised bisection over all columns at once."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    k = np.arange(n)
    lo = np.zeros(n, dtype=np.int64)    # rows [0, lo) give x_j + x_k == x_k
    hi = k.copy()                       # row k itself: 2 x_k != x_k
    while np.any(lo < hi):
        mid = (lo + hi) // 2
        same = (x[mid] + x) == x
        lo = np.where((lo < hi) & same, mid + 1, lo)
        hi = np.where((lo < hi) & ~same, mid, hi)
    return int(np.sum(k - lo + 1))


def config_batched(count=4096, seed=4096, n_min=64, n_max=2048):
    """Batched workload: `count` independent instances, n log-uniform in
    [n_min, n_max] on geometric grids r = 1.03 from x_0 = 1e-18, each with its
    own shear rate G log-uniform in [10, 500] 1/s (beta_sh = 4/3 G,
    beta_br = 6.73e-18), power-law breakage S0 = 1e-4 G^1.5 with exponent 1/3
    and v_ref = x_{n/2}, symmetric-binary daughters, and its own lognormal N
    (median pivot index uniform in [n/8, n/2], sigma_ln uniform in [0.5, 1.2],
    total 1e12, +-5 % noise).  Returns (instances, N, x, t_q, bw_q) with N and
    x packed back to back in instance order."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(n_min), math.log(n_max + 1), count)).astype(int)
    ns = np.clip(ns, n_min, n_max)
    Gs = np.exp(rng.uniform(math.log(10.0), math.log(500.0), count))
    sig = rng.uniform(0.5, 1.2, count)
    xs, Ns, inst = [], [], []
    off = 0
    for i in range(count):
        n = int(ns[i])
        x = geometric_grid(n, 1e-18, 1.03)
        med = int(rng.integers(max(1, n // 8), max(2, n // 2)))
        N = lognormal_ic(x, x[med], float(sig[i]), 1e12, noise=0.05,
                         seed=int(rng.integers(1 << 30)))
        log_x0, inv_log_r = grid_lookup(x)
        G = float(Gs[i])
        inst.append(dict(n=n, offset=off, log_x0=log_x0, inv_log_r=inv_log_r, beta0=0.0,
                         beta_br=6.73e-18, beta_sh=4.0 / 3.0 * G, S0=1e-4 * G ** 1.5,
                         v_ref=float(x[n // 2]), alpha=1.0 / 3.0, v_min=0.0, G=G))
        xs.append(x)
        Ns.append(N)
        off += n
    t_q, bw_q = daughter_symmetric()
    return inst, np.concatenate(Ns), np.concatenate(xs), t_q, bw_q


def instance_params(d):
    """(agg, brk) parameter dicts of one batched instance for the plain calls."""
    agg = dict(kernel=5, beta0=d["beta0"], beta_br=d["beta_br"], beta_sh=d["beta_sh"],
               log_x0=d["log_x0"], inv_log_r=d["inv_log_r"])
    brk = dict(selection=2, S0=d["S0"], v_ref=d["v_ref"], alpha=d["alpha"], v_min=d["v_min"])
    return agg, brk


def config_adjoint(n=1024, seed=5, G=1.0, total=1e4, beta_br=1e-8, median_frac=0.6):
    """Adjoint / parameter-fit workload: geometric grid r = 1.03 from
    x_0 = 1e-18, Brownian + shear aggregation (beta_sh = 4/3 G), power-law
    breakage S0 = 1e-4 G^1.5 with exponent 1/3 and v_ref = x_{n/2},
    symmetric-binary daughters, lognormal N (median x_{0.6 n}, sigma_ln = 0.8).

    Scales are chosen so that explicit RK4 with dt = 0.5 s is stable on the
    whole 13-decade grid: the fastest decay rate, the death rate of the top
    pivot sum_k beta(x_max, x_k) N_k ~ beta_sh x_max N_tot, stays near 0.2/s
    (limit 2.78/dt), while Brownian, shear and breakage rates at the median
    are all ~3e-4/s, so each of the four fitted parameters moves the 1000-s
    solution.  beta_br is scaled up from the physical 6.73e-18 m^3/s to keep
    the Brownian term identifiable at N_tot = 1e4 m^-3."""
    x = geometric_grid(n, 1e-18, 1.03)
    N = lognormal_ic(x, x[int(median_frac * n)], 0.8, total, noise=0.05, seed=seed)
    log_x0, inv_log_r = grid_lookup(x)
    agg = dict(kernel=5, beta0=0.0, beta_br=beta_br, beta_sh=4.0 / 3.0 * G,
               log_x0=log_x0, inv_log_r=inv_log_r)
    brk = dict(selection=2, S0=1e-4 * G ** 1.5, v_ref=float(x[n // 2]), alpha=1.0 / 3.0,
               v_min=0.0)
    t_q, bw_q = daughter_symmetric()
    return x, N, agg, brk, t_q, bw_q


def config_batched_fit(count=256, seed=256, n_min=128, n_max=1024):
    """Batched fit workload: `count` instances with n log-uniform in
    [n_min, n_max], each as config_adjoint with its own shear rate G
    log-uniform in [0.5, 2] (beta_sh = 4/3 G, S0 = 1e-4 G^1.5) and its own
    lognormal N (median pivot index uniform in [0.5 n, 0.65 n], sigma_ln
    uniform in [0.5, 1.2], total 1e4) -- stable under RK4 with dt = 0.5 s.
    Returns (instances, N, x, t_q, bw_q) packed like config_batched."""
    rng = np.random.default_rng(seed)
    ns = np.clip(np.exp(rng.uniform(math.log(n_min), math.log(n_max + 1), count)).astype(int),
                 n_min, n_max)
    Gs = np.exp(rng.uniform(math.log(0.5), math.log(2.0), count))
    xs, Ns, inst = [], [], []
    off = 0
    for i in range(count):
        n = int(ns[i])
        x = geometric_grid(n, 1e-18, 1.03)
        med = int(rng.uniform(0.5, 0.65) * n)
        N = lognormal_ic(x, x[med], float(rng.uniform(0.5, 1.2)), 1e4, noise=0.05,
                         seed=int(rng.integers(1 << 30)))
        log_x0, inv_log_r = grid_lookup(x)
        G = float(Gs[i])
        inst.append(dict(n=n, offset=off, log_x0=log_x0, inv_log_r=inv_log_r, beta0=0.0,
                         beta_br=1e-8, beta_sh=4.0 / 3.0 * G, S0=1e-4 * G ** 1.5,
                         v_ref=float(x[n // 2]), alpha=1.0 / 3.0, v_min=0.0, G=G))
        xs.append(x)
        Ns.append(N)
        off += n
    t_q, bw_q = daughter_symmetric()
    return inst, np.concatenate(Ns), np.concatenate(xs), t_q, bw_q


FIT_PARAMS = ("beta_br", "beta_sh", "S0", "alpha")


def with_params(agg, brk, theta):
    """(agg, brk) with (beta_br, beta_sh, S0, alpha) replaced by theta."""
    a, b = dict(agg), dict(brk)
    a["beta_br"], a["beta_sh"] = float(theta[0]), float(theta[1])
    b["S0"], b["alpha"] = float(theta[2]), float(theta[3])
    return a, b


def log_misfit(N, logT, mask, eps):
    """L = sum_i mask_i (log max(N_i, eps) - logT_i)^2 and dL/dN."""
    Nc = np.maximum(N, eps)
    d = np.log(Nc) - logT
    L = float(np.sum(mask * d * d))
    g = np.where(N > eps, 2.0 * mask * d / Nc, 0.0)
    return L, g


# ------------------------------------------- two-stage flocculation example --

FLOC_CONSTANTS = dict(kB=1.380649e-23, d0=0.5e-6, phi=5.0e-5, T=298.0, mu=1.0e-3,
                      d_min=0.05e-6, d_max=200.0e-6, G_high=100.0, G_low=20.0, G_ref=1.0,
                      dt_s=0.5, d_crit=5.0e-6, k_br=1.0e-3, alpha_br=2.0, eps_ero=0.05)


def config_flocculation_two_stage(n=960, t_end_s=1800.0, t_mid_s=600.0):
    """The reference's two-stage flocculation case study in its dimensionless
    form (examples/flocculation_two_stage/main.cpp:97-205, 226-331): grid of
    n geometric sections over d in [0.05, 200] um (x = v / v0), log-normal
    start sigma_ln = 0.3 centred on x = 1 (normalised by a running sum, in
    index order), Brownian + shear aggregation (beta_br = 0.25,
    beta_sh = G v0 / (pi beta0)), power-law selection S0 = k_br G t_c,
    v_ref = (d_crit / d0)^3, alpha = 2, erosion daughter (0.95, 0.05),
    RK4 with dt = 0.5 s / t_c.  Returns a dict with x, N0, the parameter
    dicts for each shear rate, the daughter table, dt, the step count, the
    number of high-shear steps (a step is high-shear while the accumulated
    dimensionless time times t_c is below t_mid, as in the reference's loop)
    and the step of the state at t_mid."""
    c = FLOC_CONSTANTS
    v0 = (math.pi / 6.0) * c["d0"] * c["d0"] * c["d0"]
    N0s = c["phi"] / v0
    beta0 = 4.0 * (2.0 * c["kB"] * c["T"] / (3.0 * c["mu"]))
    t_c = 1.0 / (beta0 * N0s)
    x = geometric_grid_range(n, (c["d_min"] / c["d0"]) ** 3.0, (c["d_max"] / c["d0"]) ** 3.0)
    lr = math.log(x[1] / x[0])
    N = np.empty(n)
    s = 0.0
    for i in range(n):
        z = math.log(x[i]) / 0.3
        N[i] = math.exp(-0.5 * z * z) * (x[i] * lr)
        s += N[i]
    N = N / s
    log_x0, inv_log_r = grid_lookup(x)

    def agg(G):
        return dict(kernel=5, beta0=0.0, beta_br=0.25, beta_sh=G * v0 / (math.pi * beta0),
                    log_x0=log_x0, inv_log_r=inv_log_r)

    def brk(G):
        return dict(selection=2, S0=c["k_br"] * G / c["G_ref"] * t_c,
                    v_ref=(c["d_crit"] / c["d0"]) ** 3.0, alpha=c["alpha_br"], v_min=0.0)

    dt_nd = c["dt_s"] / t_c
    steps = int(t_end_s / c["dt_s"])
    n_high, t_nd = 0, 0.0
    for k in range(steps):
        if t_nd * t_c < t_mid_s:
            n_high = k + 1
        t_nd += dt_nd
    return dict(x=x, N0=N, agg_high=agg(c["G_high"]), agg_low=agg(c["G_low"]),
                brk_high=brk(c["G_high"]), brk_low=brk(c["G_low"]),
                t_q=np.array([1.0 - c["eps_ero"], c["eps_ero"]]), bw_q=np.array([1.0, 1.0]),
                dt=dt_nd, steps=steps, n_high=n_high,
                mid_step=int(t_mid_s / c["dt_s"] + 0.5), t_c=t_c, d0=c["d0"])


def d43(N, x, d_ref):
    """Volume-weighted mean diameter sum N d^4 / sum N d^3 of pivots x (volumes
    in units of a reference volume of diameter d_ref), summed in index order."""
    num = den = 0.0
    for Ni, xi in zip(N, x):
        num += Ni * xi ** (4.0 / 3.0)
        den += Ni * xi
    return d_ref * num / den if den > 0.0 else 0.0
