"""Python front end of the gfx950 fixed-pivot PBE library (libpbe_hip.so).

Mirrors the reference's C++ interface (include/pbe_cuda/aggregation.cuh and
include/pbe_cuda/breakage.cuh): the same enum values, parameter structs with
the same fields and defaults, and ``launch_aggregation_rhs`` /
``launch_breakage_rhs`` that accumulate into a caller-zeroed ``rhs`` and return
the HIP status code.  Calls go through the C ABI declared in
include/pbe_hip/pbe_hip.h; there is no other implementation behind them --
if the shared library is missing, importing this package raises.
"""
import ctypes
import enum
import os

from . import hip

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libpbe_hip.so")


class AggregationKernel(enum.IntEnum):
    """include/pbe_cuda/aggregation.cuh:32-39"""
    Constant = 0
    Sum = 1
    Product = 2
    Brownian = 3
    Shear = 4
    BrownianShear = 5


class BreakageSelection(enum.IntEnum):
    """include/pbe_cuda/breakage.cuh:30-35"""
    Constant = 0
    Linear = 1
    PowerLaw = 2
    Threshold = 3


class AggregationParams(ctypes.Structure):
    """Layout of AggregationParams (include/pbe_cuda/aggregation.cuh:59-73)."""
    _fields_ = [("n", ctypes.c_int), ("log_x0", ctypes.c_double),
                ("inv_log_r", ctypes.c_double), ("kernel_type", ctypes.c_int),
                ("beta0", ctypes.c_double), ("beta_br", ctypes.c_double),
                ("beta_sh", ctypes.c_double), ("block_size", ctypes.c_int)]

    def __init__(self, n=0, log_x0=0.0, inv_log_r=0.0, kernel_type=0, beta0=1.0,
                 beta_br=0.0, beta_sh=0.0, block_size=256, kernel=None):
        if kernel is not None:
            kernel_type = kernel
        super().__init__(int(n), float(log_x0), float(inv_log_r), int(kernel_type),
                         float(beta0), float(beta_br), float(beta_sh), int(block_size))


class AggOptions(ctypes.Structure):
    """pbe_agg_options: block width (columns per lane), row slices per
    diagonal-band tile and launch form (1: band and strips apart, 2: one
    launch of both, 3: one launch of every tile, 4: as 2 at three
    workgroups per CU); 0 selects the autotuned value."""
    _fields_ = [("tile_u", ctypes.c_int), ("band_split", ctypes.c_int),
                ("form", ctypes.c_int)]


class BreakageParams(ctypes.Structure):
    """Layout of BreakageParams (include/pbe_cuda/breakage.cuh:63-77)."""
    _fields_ = [("n", ctypes.c_int), ("n_quad", ctypes.c_int),
                ("selection", ctypes.c_int), ("S0", ctypes.c_double),
                ("v_ref", ctypes.c_double), ("alpha", ctypes.c_double),
                ("v_min", ctypes.c_double), ("block_size", ctypes.c_int)]

    def __init__(self, n=0, n_quad=0, selection=0, S0=1.0, v_ref=1.0, alpha=1.0,
                 v_min=0.0, block_size=256):
        super().__init__(int(n), int(n_quad), int(selection), float(S0), float(v_ref),
                         float(alpha), float(v_min), int(block_size))


_lib = None


def library():
    """Load libpbe_hip.so (after the HIP runtime, so it binds to that one)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; "
                "g.build()'` from the repository root")
        hip.runtime()
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        vp = ctypes.c_void_p
        L.pbe_launch_aggregation_rhs.restype = ctypes.c_int
        L.pbe_launch_aggregation_rhs.argtypes = [vp, vp, vp, ctypes.POINTER(AggregationParams), vp]
        L.pbe_launch_breakage_rhs.restype = ctypes.c_int
        L.pbe_launch_breakage_rhs.argtypes = [vp, vp, vp, vp, vp,
                                              ctypes.POINTER(BreakageParams), vp]
        L.pbe_version.restype = ctypes.c_char_p
        L.pbe_version.argtypes = []
        from . import _abi_ext
        _abi_ext.declare(L)
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, hip.DeviceArray):
        return a.ptr
    if isinstance(a, ctypes.c_void_p):
        return a
    return ctypes.c_void_p(int(a))


def _stream(s):
    if s is None:
        return None
    return s.handle if isinstance(s, hip.Stream) else ctypes.c_void_p(int(s))


def launch_aggregation_rhs(N, x, rhs, params, stream=None):
    """rhs += aggregation source (pbe_launch_aggregation_rhs).  Returns the
    HIP status code, as the reference's C++ function does."""
    return library().pbe_launch_aggregation_rhs(_ptr(N), _ptr(x), _ptr(rhs),
                                                ctypes.byref(params), _stream(stream))


def launch_aggregation_rhs_ex(N, x, rhs, params, tile_u=0, band_split=0, form=0, stream=None):
    """pbe_launch_aggregation_rhs_ex with explicit launch options."""
    opt = AggOptions(int(tile_u), int(band_split), int(form))
    return library().pbe_launch_aggregation_rhs_ex(_ptr(N), _ptr(x), _ptr(rhs),
                                                   ctypes.byref(params), ctypes.byref(opt),
                                                   _stream(stream))


def aggregation_default_options(params):
    """(tile_u, band_split, form) that options of 0 resolve to for these
    params (pbe_aggregation_default_options)."""
    opt = AggOptions(0, 0, 0)
    hip.check(library().pbe_aggregation_default_options(ctypes.byref(params),
                                                        ctypes.byref(opt)),
              "pbe_aggregation_default_options")
    return opt.tile_u, opt.band_split, opt.form


def launch_aggregation_rhs_deterministic(N, x, rhs, params, workspace, stream=None):
    """rhs += aggregation source, atomic-free and bitwise repeatable
    (pbe_launch_aggregation_rhs_deterministic).  `workspace` is a device array
    of at least aggregation_deterministic_workspace_bytes(n) bytes."""
    return library().pbe_launch_aggregation_rhs_deterministic(
        _ptr(N), _ptr(x), _ptr(rhs), ctypes.byref(params), _ptr(workspace),
        workspace.nbytes if workspace is not None else 0, _stream(stream))


def launch_breakage_rhs_deterministic(N, x, t_q, bw_q, rhs, params, workspace, stream=None):
    """rhs += breakage source, atomic-free and bitwise repeatable."""
    return library().pbe_launch_breakage_rhs_deterministic(
        _ptr(N), _ptr(x), _ptr(t_q), _ptr(bw_q), _ptr(rhs), ctypes.byref(params),
        _ptr(workspace), workspace.nbytes if workspace is not None else 0, _stream(stream))


def aggregation_deterministic_workspace_bytes(n):
    return library().pbe_aggregation_deterministic_workspace_bytes(int(n))


def breakage_deterministic_workspace_bytes(n):
    return library().pbe_breakage_deterministic_workspace_bytes(int(n))


def launch_breakage_rhs(N, x, t_q, bw_q, rhs, params, stream=None):
    """rhs += breakage source (pbe_launch_breakage_rhs).  Returns the HIP
    status code."""
    return library().pbe_launch_breakage_rhs(_ptr(N), _ptr(x), _ptr(t_q), _ptr(bw_q),
                                             _ptr(rhs), ctypes.byref(params),
                                             _stream(stream))


def version():
    return library().pbe_version().decode()


def agg_params(n, agg):
    """AggregationParams from a problems.* parameter dict."""
    return AggregationParams(n=n, log_x0=agg.get("log_x0", 0.0),
                             inv_log_r=agg.get("inv_log_r", 0.0),
                             kernel_type=int(agg["kernel"]), beta0=agg.get("beta0", 1.0),
                             beta_br=agg.get("beta_br", 0.0), beta_sh=agg.get("beta_sh", 0.0))


def brk_params(n, n_quad, brk):
    """BreakageParams from a problems.* parameter dict."""
    return BreakageParams(n=n, n_quad=n_quad, selection=int(brk["selection"]),
                          S0=brk.get("S0", 1.0), v_ref=brk.get("v_ref", 1.0),
                          alpha=brk.get("alpha", 1.0), v_min=brk.get("v_min", 0.0))


def rhs_on_device(N, x, agg=None, brk=None, t_q=None, bw_q=None):
    """Convenience: upload host arrays, evaluate aggregation and/or breakage
    into a zeroed rhs on the GPU, synchronise and return rhs on the host."""
    import numpy as np
    n = len(N)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    if agg is not None:
        hip.check(launch_aggregation_rhs(dN, dx, drhs, agg_params(n, agg)),
                  "launch_aggregation_rhs")
    if brk is not None:
        dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
        hip.check(launch_breakage_rhs(dN, dx, dt, db, drhs,
                                      brk_params(n, len(np.atleast_1d(t_q)), brk)),
                  "launch_breakage_rhs")
    hip.synchronize()
    return drhs.download()


# ------------------------------------------------------------ integration --
METHOD_RK4 = 0
METHOD_EULER = 1


class Integrator:
    """Device RK4 / Euler over R_agg + R_brk (pbe_integrate and the hipGraph
    entry points).  Holds device copies of x, t_q, bw_q and the workspace;
    N is a DeviceArray advanced in place."""

    def __init__(self, x, agg=None, brk=None, t_q=None, bw_q=None, method=METHOD_RK4):
        import numpy as np
        L = library()
        self.n = len(x)
        self.x = hip.DeviceArray.from_host(x)
        self.agg = agg_params(self.n, agg) if agg is not None else None
        self.brk = None
        self.t_q = self.bw_q = None
        if brk is not None:
            t_q, bw_q = np.atleast_1d(t_q), np.atleast_1d(bw_q)
            self.t_q = hip.DeviceArray.from_host(t_q)
            self.bw_q = hip.DeviceArray.from_host(bw_q)
            self.brk = brk_params(self.n, len(t_q), brk)
        self.method = method
        self.ws_bytes = L.pbe_integrator_workspace_bytes(self.n, method)
        self.ws = hip.DeviceArray(self.ws_bytes // 8)
        self.graphs = []

    def _args(self):
        return (_ptr(self.x), _ptr(self.t_q), _ptr(self.bw_q),
                ctypes.byref(self.agg) if self.agg is not None else None,
                ctypes.byref(self.brk) if self.brk is not None else None)

    def step(self, N, dt, steps=1, stream=None):
        x, tq, bw, a, b = self._args()
        return library().pbe_integrate(_ptr(N), x, tq, bw, a, b, float(dt), int(steps),
                                       self.method, self.ws.ptr, self.ws_bytes,
                                       _stream(stream))

    def graph(self, N, dt, steps_per_launch):
        """Capture `steps_per_launch` steps into a hipGraph; returns a Graph."""
        x, tq, bw, a, b = self._args()
        g = ctypes.c_void_p()
        hip.check(library().pbe_integrator_graph_create(
            ctypes.byref(g), _ptr(N), x, tq, bw, a, b, float(dt), int(steps_per_launch),
            self.method, self.ws.ptr, self.ws_bytes), "pbe_integrator_graph_create")
        gr = Graph(g)
        self.graphs.append(gr)
        return gr


class Graph:
    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream=None):
        return library().pbe_graph_launch(self.handle, _stream(stream))

    def destroy(self):
        if self.handle is not None and self.handle.value:
            library().pbe_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # noqa: BLE001
            pass


# ---------------------------------------------------------------- batched --
class BatchInstance(ctypes.Structure):
    """pbe_batch_instance (include/pbe_hip/pbe_hip.h)."""
    _fields_ = [("n", ctypes.c_int), ("offset", ctypes.c_int), ("log_x0", ctypes.c_double),
                ("inv_log_r", ctypes.c_double), ("beta0", ctypes.c_double),
                ("beta_br", ctypes.c_double), ("beta_sh", ctypes.c_double),
                ("S0", ctypes.c_double), ("v_ref", ctypes.c_double),
                ("alpha", ctypes.c_double), ("v_min", ctypes.c_double)]


class BatchPlan:
    """A pbe_batch_plan over instances given as dicts with keys n, offset and
    the coefficients (log_x0, inv_log_r, beta0, beta_br, beta_sh, S0, v_ref,
    alpha, v_min).  kernel_type / selection = -1 switch a term off."""

    def __init__(self, instances, kernel_type, selection, n_quad, tile_u=0):
        arr = (BatchInstance * len(instances))()
        for a, d in zip(arr, instances):
            a.n, a.offset = int(d["n"]), int(d["offset"])
            for f in ("log_x0", "inv_log_r", "beta0", "beta_br", "beta_sh", "S0", "v_ref",
                      "alpha", "v_min"):
                setattr(a, f, float(d.get(f, 0.0)))
        self.handle = ctypes.c_void_p()
        if tile_u:
            hip.check(library().pbe_batch_plan_create_ex(ctypes.byref(self.handle), arr,
                                                         len(instances), int(kernel_type),
                                                         int(selection), int(n_quad),
                                                         int(tile_u)),
                      "pbe_batch_plan_create_ex")
        else:
            hip.check(library().pbe_batch_plan_create(ctypes.byref(self.handle), arr,
                                                      len(instances), int(kernel_type),
                                                      int(selection), int(n_quad)),
                      "pbe_batch_plan_create")
        self.size = library().pbe_batch_plan_size(self.handle)
        self.instances = instances

    def launch(self, N, x, t_q, bw_q, rhs, stream=None):
        return library().pbe_launch_batched_rhs(self.handle, _ptr(N), _ptr(x), _ptr(t_q),
                                                _ptr(bw_q), _ptr(rhs), _stream(stream))

    def integrate(self, N, x, t_q, bw_q, dt, steps, ws, ws_bytes, method=METHOD_RK4,
                  stream=None):
        return library().pbe_batch_integrate(self.handle, _ptr(N), _ptr(x), _ptr(t_q),
                                             _ptr(bw_q), float(dt), int(steps), int(method),
                                             _ptr(ws), int(ws_bytes), _stream(stream))

    def graph(self, N, x, t_q, bw_q, dt, steps_per_launch, ws, ws_bytes, method=METHOD_RK4):
        g = ctypes.c_void_p()
        hip.check(library().pbe_batch_integrator_graph_create(
            ctypes.byref(g), self.handle, _ptr(N), _ptr(x), _ptr(t_q), _ptr(bw_q), float(dt),
            int(steps_per_launch), int(method), _ptr(ws), int(ws_bytes)),
            "pbe_batch_integrator_graph_create")
        return Graph(g)

    def destroy(self):
        if self.handle is not None and self.handle.value:
            library().pbe_batch_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # noqa: BLE001
            pass


# ---------------------------------------------------------- differentiation --
GRAD_SLOTS = ("beta0", "beta_br", "beta_sh", "S0", "alpha")
STORE_ALL_STAGES = 0   # checkpoint interval that keeps every RK4 stage state


def rhs_vjp(N, x, lam, agg=None, brk=None, t_q=None, bw_q=None):
    """Host convenience over pbe_rhs_vjp: returns (gN, gtheta[5]) with
    gN = lam . dR/dN and gtheta = lam . dR/d(beta0, beta_br, beta_sh, S0, alpha)."""
    import numpy as np
    L = library()
    n = len(N)
    it = Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    dN, dl = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(lam)
    gN, gth = hip.DeviceArray(n), hip.DeviceArray(5)
    wsb = L.pbe_vjp_workspace_bytes(n)
    ws = hip.DeviceArray(wsb // 8)
    xp, tq, bw, a, b = it._args()
    hip.check(L.pbe_rhs_vjp(_ptr(dN), xp, tq, bw, _ptr(dl), a, b, _ptr(gN), _ptr(gth),
                            ws.ptr, wsb, None), "pbe_rhs_vjp")
    hip.synchronize()
    return gN.download(), gth.download()


class Adjoint:
    """Checkpointed discrete RK4 adjoint (pbe_adjoint_rk4_forward/backward)
    of `steps` steps of size dt, checkpoint every `every` steps (every = 0:
    keep every stage state, no recompute in the reverse pass)."""

    def __init__(self, x, dt, steps, every=50, agg=None, brk=None, t_q=None, bw_q=None):
        self.it = Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
        self.n, self.dt, self.steps, self.every = len(x), float(dt), int(steps), int(every)
        self.ws_bytes = library().pbe_adjoint_workspace_bytes(self.n, self.steps, self.every)
        self.ws = hip.DeviceArray(self.ws_bytes // 8)

    def forward(self, N, stream=None):
        """Advance the DeviceArray / device pointer N in place, storing checkpoints."""
        x, tq, bw, a, b = self.it._args()
        return library().pbe_adjoint_rk4_forward(_ptr(N), x, tq, bw, a, b, self.dt,
                                                 self.steps, self.every, self.ws.ptr,
                                                 self.ws_bytes, _stream(stream))

    def backward(self, gN_final, gN0, gtheta, stream=None):
        x, tq, bw, a, b = self.it._args()
        return library().pbe_adjoint_rk4_backward(_ptr(gN_final), _ptr(gN0), _ptr(gtheta), x,
                                                  tq, bw, a, b, self.dt, self.steps,
                                                  self.every, self.ws.ptr, self.ws_bytes,
                                                  _stream(stream))


class BatchAdjoint:
    """Checkpointed RK4 adjoint over every instance of a BatchPlan
    (pbe_batch_adjoint_rk4_forward/backward); gtheta slots 0-3 are
    d/d ln(common scale) of beta0, beta_br, beta_sh, S0, slot 4 d/d alpha."""

    def __init__(self, plan, x, t_q, bw_q, dt, steps, every=50):
        self.plan = plan
        self.x = hip.DeviceArray.from_host(x)
        self.t_q = hip.DeviceArray.from_host(t_q) if t_q is not None else None
        self.bw_q = hip.DeviceArray.from_host(bw_q) if bw_q is not None else None
        self.dt, self.steps, self.every = float(dt), int(steps), int(every)
        self.ws_bytes = library().pbe_adjoint_workspace_bytes(plan.size, self.steps, self.every)
        self.ws = hip.DeviceArray(self.ws_bytes // 8)

    def forward(self, N, stream=None):
        return library().pbe_batch_adjoint_rk4_forward(
            self.plan.handle, _ptr(N), self.x.ptr, _ptr(self.t_q), _ptr(self.bw_q), self.dt,
            self.steps, self.every, self.ws.ptr, self.ws_bytes, _stream(stream))

    def backward(self, gN_final, gN0, gtheta, stream=None):
        return library().pbe_batch_adjoint_rk4_backward(
            self.plan.handle, _ptr(gN_final), _ptr(gN0), _ptr(gtheta), self.x.ptr,
            _ptr(self.t_q), _ptr(self.bw_q), self.dt, self.steps, self.every, self.ws.ptr,
            self.ws_bytes, _stream(stream))


def batch_rhs_vjp(plan, N, x, t_q, bw_q, lam):
    """Host convenience over pbe_batch_rhs_vjp: (gN, gtheta[5])."""
    L = library()
    n = plan.size
    dN, dx, dl = (hip.DeviceArray.from_host(a) for a in (N, x, lam))
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    gN, gth = hip.DeviceArray(n), hip.DeviceArray(5)
    wsb = L.pbe_vjp_workspace_bytes(n)
    ws = hip.DeviceArray(wsb // 8)
    hip.check(L.pbe_batch_rhs_vjp(plan.handle, dN.ptr, dx.ptr, dt.ptr, db.ptr, dl.ptr, gN.ptr,
                                  gth.ptr, ws.ptr, wsb, None), "pbe_batch_rhs_vjp")
    hip.synchronize()
    return gN.download(), gth.download()
