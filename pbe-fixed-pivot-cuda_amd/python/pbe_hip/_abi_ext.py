"""ctypes signatures of the C-ABI entry points beyond the two reference
replacements (include/pbe_hip/pbe_hip.h)."""
import ctypes

vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t


def declare(lib):
    sig = {
        "pbe_launch_aggregation_rhs_ex": (i, [vp, vp, vp, vp, vp, vp]),
        "pbe_aggregation_default_options": (i, [vp, vp]),
        "pbe_aggregation_deterministic_workspace_bytes": (sz, [i]),
        "pbe_breakage_deterministic_workspace_bytes": (sz, [i]),
        "pbe_launch_aggregation_rhs_deterministic": (i, [vp, vp, vp, vp, vp, sz, vp]),
        "pbe_launch_breakage_rhs_deterministic": (i, [vp, vp, vp, vp, vp, vp, vp, sz, vp]),
        "pbe_launch_aggregation_rhs_deterministic_noalloc": (i, [vp, vp, vp, vp, vp]),
        "pbe_integrator_workspace_bytes": (sz, [i, i]),
        "pbe_integrate": (i, [vp, vp, vp, vp, vp, vp, d, i, i, vp, sz, vp]),
        "pbe_integrator_graph_create": (i, [ctypes.POINTER(vp), vp, vp, vp, vp, vp, vp, d,
                                            i, i, vp, sz]),
        "pbe_graph_launch": (i, [vp, vp]),
        "pbe_graph_destroy": (i, [vp]),
        "pbe_batch_plan_create": (i, [ctypes.POINTER(vp), vp, i, i, i, i]),
        "pbe_batch_plan_create_ex": (i, [ctypes.POINTER(vp), vp, i, i, i, i, i]),
        "pbe_batch_plan_destroy": (i, [vp]),
        "pbe_batch_plan_size": (i, [vp]),
        "pbe_launch_batched_rhs": (i, [vp, vp, vp, vp, vp, vp, vp]),
        "pbe_batch_integrate": (i, [vp, vp, vp, vp, vp, d, i, i, vp, sz, vp]),
        "pbe_batch_integrator_graph_create": (i, [ctypes.POINTER(vp), vp, vp, vp, vp, vp, d, i,
                                                  i, vp, sz]),
        "pbe_vjp_workspace_bytes": (sz, [i]),
        "pbe_rhs_vjp": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
        "pbe_adjoint_workspace_bytes": (sz, [i, i, i]),
        "pbe_adjoint_rk4_forward": (i, [vp, vp, vp, vp, vp, vp, d, i, i, vp, sz, vp]),
        "pbe_adjoint_rk4_backward": (i, [vp, vp, vp, vp, vp, vp, vp, vp, d, i, i, vp, sz, vp]),
        "pbe_batch_rhs_vjp": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
        "pbe_batch_adjoint_rk4_forward": (i, [vp, vp, vp, vp, vp, d, i, i, vp, sz, vp]),
        "pbe_batch_adjoint_rk4_backward": (i, [vp, vp, vp, vp, vp, vp, vp, d, i, i, vp, sz,
                                               vp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib
