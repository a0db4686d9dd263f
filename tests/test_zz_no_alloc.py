"""The no-allocation entry point pbe_launch_aggregation_rhs_deterministic_noalloc
(its workspace is the library's static per-device array) against the
workspace form pbe_launch_aggregation_rhs_deterministic: the same kernels on
the same inputs, so the results must be identical bit for bit.  Two calls in a
row, the second on a smaller problem (the array keeps the first call's
contents), then a size past the array's limit, which is refused before any
launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n, kernel, seed):
    from pbe_hip import problems as P
    x = P.geometric_grid(n, 1e-18, 1.02)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=seed)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
               beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0, inv_log_r=inv_log_r)
    return x, N, agg


def _both(pbe, x, N, agg):
    """(no-allocation result, workspace-form result) through the C ABI."""
    from pbe_hip import hip
    L = pbe.library()
    n = len(N)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    p = pbe.agg_params(n, agg)
    r_static = hip.DeviceArray(n)
    r_static.zero()
    hip.check(L.pbe_launch_aggregation_rhs_deterministic_noalloc(dN.ptr, dx.ptr, r_static.ptr,
                                                                 ctypes.byref(p), None),
              "noalloc")
    hip.synchronize()
    got = r_static.download()
    ws_bytes = pbe.aggregation_deterministic_workspace_bytes(n)
    ws = hip.DeviceArray(ws_bytes // 8 + 1)
    r_ws = hip.DeviceArray(n)
    r_ws.zero()
    hip.check(pbe.launch_aggregation_rhs_deterministic(dN, dx, r_ws, p, ws), "workspace")
    hip.synchronize()
    return got, r_ws.download()


@pytest.mark.parametrize("kernel", [5, 0])
def test_noalloc_matches_workspace_form(pbe, oracle_mod, kernel):
    from parity import assert_parity
    for n, seed in ((12000, 1), (3000, 2)):   # the second call smaller than the first
        x, N, agg = _problem(n, kernel, seed)
        got, ref = _both(pbe, x, N, agg)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (kernel, n)
        r, A, _ = oracle_mod.aggregation_rhs(N, x, **agg)
        assert_parity(got, r, A, what=f"noalloc kernel {kernel} n={n}")


def test_noalloc_refuses_sizes_past_its_array(pbe):
    from pbe_hip import hip
    L = pbe.library()
    n = 65537   # PBE_DETERMINISTIC_NOALLOC_MAX_N + 1
    x, N, agg = _problem(64, 5, 3)
    dN, dx, dr = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x), hip.DeviceArray(64)
    p = pbe.agg_params(n, agg)   # n beyond the buffers too: refused before any launch
    assert L.pbe_launch_aggregation_rhs_deterministic_noalloc(dN.ptr, dx.ptr, dr.ptr,
                                                              ctypes.byref(p), None) == 1
