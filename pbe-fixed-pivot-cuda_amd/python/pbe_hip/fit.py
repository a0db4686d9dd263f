"""Joint parameter fit over a batch of independent PBE instances.

Every instance of the batch carries its own coefficients (its own shear rate
G, hence beta_sh and S0); the fit moves them all by common factors,

    beta_br_i = s_br * beta_br_i(0),  beta_sh_i = s_sh * beta_sh_i(0),
    S0_i = s_S0 * S0_i(0),             alpha_i = alpha,

so the fitted vector is theta = (ln s_br, ln s_sh, ln s_S0, alpha).  The loss
is the summed L2 misfit of the log final PSD of every instance against a
target, L = sum_i mask_i (log max(N_i(T), eps) - logT_i)^2.  Its gradient
comes from one batched discrete RK4 adjoint (pbe_batch_adjoint_rk4_forward /
_backward): the gtheta slots 1-3 of that call are d/d ln(scale) of beta_br,
beta_sh and S0 over all instances, slot 4 is d/d alpha -- exactly dL/dtheta.

The reference has no parameter fit (it evaluates source terms only); its RK4
arithmetic, which the forward solve and the adjoint follow, is that of
examples/common/pbe_examples_utils.hpp:417-457.
"""
import math

import numpy as np

from . import BatchAdjoint, BatchPlan, hip
from . import problems as P

_SLOTS = (1, 2, 3, 4)   # gtheta slots of (ln s_br, ln s_sh, ln s_S0, alpha)


class BatchedFit:
    """A batch of instances (dicts as problems.config_batched_fit returns
    them, N0 and x packed in instance order) whose common parameter factors
    are fitted.  `kernel_type` / `selection` are the batch's aggregation
    kernel and breakage selection; `every` the adjoint's checkpoint interval
    (0: keep every RK4 stage state)."""

    def __init__(self, instances, N0, x, t_q, bw_q, dt, steps, every=50, kernel_type=5,
                 selection=2, eps=1e-26):
        self.base = [dict(d) for d in instances]
        self.N0 = np.ascontiguousarray(N0, dtype=np.float64)
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.t_q, self.bw_q = t_q, bw_q
        self.dt, self.steps, self.every = float(dt), int(steps), int(every)
        self.kernel_type, self.selection = int(kernel_type), int(selection)
        self.eps = float(eps)
        self.theta0 = np.array([0.0, 0.0, 0.0, self.base[0]["alpha"]])
        if any(d["alpha"] != self.theta0[3] for d in self.base):
            raise ValueError("BatchedFit: the instances must share one breakage exponent")
        self.logT = None
        self.mask = None
        self.evaluations = 0

    def instances(self, theta):
        """The instance dicts with theta's factors applied."""
        s_br, s_sh, s_s0 = (math.exp(float(t)) for t in theta[:3])
        out = []
        for d in self.base:
            e = dict(d)
            e["beta_br"] = d["beta_br"] * s_br
            e["beta_sh"] = d["beta_sh"] * s_sh
            e["S0"] = d["S0"] * s_s0
            e["alpha"] = float(theta[3])
            out.append(e)
        return out

    def _adjoint(self, theta):
        plan = BatchPlan(self.instances(theta), self.kernel_type, self.selection, len(self.t_q))
        return BatchAdjoint(plan, self.x, self.t_q, self.bw_q, self.dt, self.steps, self.every)

    def solve(self, theta, stream=None):
        """N(T) of every instance at theta (packed, on the host)."""
        adj = self._adjoint(theta)
        dN = hip.DeviceArray.from_host(self.N0)
        hip.check(adj.forward(dN, stream), "pbe_batch_adjoint_rk4_forward")
        hip.synchronize()
        return dN.download()

    def set_target(self, theta_true, rel_floor=1e-8):
        """Target log PSD from a solve at theta_true; bins below rel_floor of
        their instance's largest entry are left out of the loss."""
        T = self.solve(theta_true)
        self.logT = np.log(np.maximum(T, self.eps))
        self.mask = np.zeros_like(T)
        for d in self.base:
            lo, hi = d["offset"], d["offset"] + d["n"]
            self.mask[lo:hi] = T[lo:hi] > rel_floor * T[lo:hi].max()
        return T

    def loss_and_grad(self, theta, mask=None, stream=None):
        """(L, dL/dtheta) at theta; `mask` overrides the target's mask (for
        example to keep a few instances only)."""
        if self.logT is None:
            raise RuntimeError("BatchedFit: set_target first")
        m = self.mask if mask is None else mask
        adj = self._adjoint(theta)
        dN = hip.DeviceArray.from_host(self.N0)
        hip.check(adj.forward(dN, stream), "pbe_batch_adjoint_rk4_forward")
        hip.synchronize()
        L, g = P.log_misfit(dN.download(), self.logT, m, self.eps)
        dg = hip.DeviceArray.from_host(g)
        gN0, gth = hip.DeviceArray(len(self.N0)), hip.DeviceArray(5)
        hip.check(adj.backward(dg, gN0, gth, stream), "pbe_batch_adjoint_rk4_backward")
        hip.synchronize()
        self.evaluations += 1
        return L, gth.download()[list(_SLOTS)]

    def fit(self, theta_start=None, max_evals=60, gtol=1e-10, bounds=None):
        """L-BFGS-B from theta_start (default: all factors 1); returns
        (theta, history of losses, scipy result)."""
        from scipy.optimize import minimize
        th0 = self.theta0 if theta_start is None else np.asarray(theta_start, dtype=float)
        hist = []

        def f(th):
            L, g = self.loss_and_grad(th)
            hist.append(L)
            return L, g

        if bounds is None:
            b = math.log(4.0)
            bounds = [(-b, b)] * 3 + [(0.05, 1.0)]
        res = minimize(f, th0, jac=True, method="L-BFGS-B", bounds=bounds,
                       options={"maxfun": int(max_evals), "gtol": gtol, "ftol": 1e-15})
        return res.x, hist, res
