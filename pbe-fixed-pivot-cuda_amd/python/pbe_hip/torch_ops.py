"""Thin torch.autograd.Function over the checkpointed RK4 adjoint, for
parameter fitting with torch optimisers.

Import this module before anything else loads the HIP runtime: torch brings
its own copy of libamdhip64.so.7, and the library must bind to that same copy
(pbe_hip.hip loads the runtime by soname, which then resolves to torch's).

    import torch
    from pbe_hip.torch_ops import PBESolve
    N_end = PBESolve.apply(N0, theta, problem)   # theta = (beta_br, beta_sh, S0, alpha)

Tensors are float64 on the current HIP device; work is enqueued on torch's
current stream.  The forward keeps a checkpoint every `problem.every` steps;
the backward recomputes each segment and runs the hand-written VJP kernels.
"""
import torch  # noqa: F401  (must precede the HIP runtime load)

import pbe_hip as pbe
from pbe_hip import hip, problems

_SLOTS = (1, 2, 3, 4)   # gtheta slots of beta_br, beta_sh, S0, alpha


class Problem:
    """The fixed part of a fit: grid, daughter table, step size and count, and
    the kernel/selection templates whose fit parameters theta replaces."""

    def __init__(self, x, agg, brk, t_q, bw_q, dt, steps, every=50):
        self.x, self.agg, self.brk = x, agg, brk
        self.t_q, self.bw_q = t_q, bw_q
        self.dt, self.steps, self.every = float(dt), int(steps), int(every)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check_inputs(N0, theta, prob):
    """Refuse anything the device kernels would read out of bounds or from
    host memory: N0 must be a float64 tensor on the HIP device with one entry
    per pivot, theta must hold the 4 fitted parameters."""
    if not isinstance(N0, torch.Tensor) or not N0.is_cuda:
        raise ValueError("PBESolve: N0 must be a tensor on the HIP device")
    if N0.dtype != torch.float64:
        raise ValueError(f"PBESolve: N0 must be float64, got {N0.dtype}")
    if N0.numel() != len(prob.x):
        raise ValueError(f"PBESolve: N0 has {N0.numel()} entries, the grid {len(prob.x)}")
    if not isinstance(theta, torch.Tensor) or theta.numel() != len(_SLOTS):
        raise ValueError("PBESolve: theta must be a tensor of 4 parameters "
                         "(beta_br, beta_sh, S0, alpha)")


class PBESolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, N0, theta, prob):
        _check_inputs(N0, theta, prob)
        th = theta.detach().to("cpu", torch.float64).numpy()
        agg, brk = problems.with_params(prob.agg, prob.brk, th)
        adj = pbe.Adjoint(prob.x, prob.dt, prob.steps, prob.every, agg=agg, brk=brk,
                          t_q=prob.t_q, bw_q=prob.bw_q)
        N = N0.detach().to(torch.float64).contiguous().clone()
        hip.check(adj.forward(N.data_ptr(), _stream()), "pbe_adjoint_rk4_forward")
        ctx.adj = adj
        ctx.theta_device = theta.device
        return N

    @staticmethod
    def backward(ctx, gN):
        if not gN.is_cuda or gN.numel() != ctx.adj.n:
            raise ValueError("PBESolve: the output gradient must be a device tensor of n entries")
        gN = gN.detach().to(torch.float64).contiguous()
        gN0 = torch.empty_like(gN)
        gth = torch.empty(pbe_grad_count(), dtype=torch.float64, device=gN.device)
        hip.check(ctx.adj.backward(gN.data_ptr(), gN0.data_ptr(), gth.data_ptr(), _stream()),
                  "pbe_adjoint_rk4_backward")
        return gN0, gth[list(_SLOTS)].to(ctx.theta_device), None


def pbe_grad_count():
    return len(pbe.GRAD_SLOTS)
