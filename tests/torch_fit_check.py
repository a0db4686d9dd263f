"""Run in its own process by test_gpu_torch.py (torch must load the HIP
runtime first): gradient of a log-PSD misfit through PBESolve against the
C-ABI adjoint and central differences, then a short Adam fit."""
import json
import os
import sys

import torch

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "python"))

import numpy as np  # noqa: E402

from pbe_hip import problems as P  # noqa: E402
from pbe_hip.torch_ops import PBESolve, Problem  # noqa: E402


def main():
    dev = torch.device("cuda:0")
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(256)
    dt, steps = 0.5, 200
    prob = Problem(x, agg, brk, t_q, bw_q, dt, steps, every=25)
    th0 = np.array([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]])
    rng = np.random.default_rng(7)
    th_true = th0 * rng.uniform(0.5, 2.0, 4)
    N0t = torch.tensor(N0, dtype=torch.float64, device=dev)
    with torch.no_grad():
        T = PBESolve.apply(N0t, torch.tensor(th_true, dtype=torch.float64, device=dev), prob)
    eps = 1e-30 * float(N0.sum())
    mask = (T > 1e-8 * T.max()).to(torch.float64)
    logT = torch.log(torch.clamp(T, min=eps))

    def loss_of(log_th):
        th = torch.exp(log_th)
        N = PBESolve.apply(N0t, th, prob)
        d = torch.log(torch.clamp(N, min=eps)) - logT
        return torch.sum(mask * d * d)

    log_th = torch.tensor(np.log(th0), dtype=torch.float64, device=dev, requires_grad=True)
    L = loss_of(log_th)
    L.backward()
    g = log_th.grad.detach().cpu().numpy()
    fd = []
    for p in range(4):
        h = 1e-6
        e = torch.zeros(4, dtype=torch.float64, device=dev)
        e[p] = h
        with torch.no_grad():
            fd.append(float((loss_of(log_th + e) - loss_of(log_th - e)) / (2 * h)))
    opt = torch.optim.Adam([log_th], lr=0.05)
    hist = []
    for it in range(15):
        opt.zero_grad()
        Lk = loss_of(log_th)
        Lk.backward()
        opt.step()
        hist.append(float(Lk))
    print(json.dumps({"loss": float(L), "grad": g.tolist(), "fd": fd, "hist": hist}))


if __name__ == "__main__":
    main()
