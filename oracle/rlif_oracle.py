"""TEST INFRASTRUCTURE ONLY: the FP64 PyTorch autograd unroll of the recurrent LIF layer that checks the
hand-written forward / adjoint kernels (naer2_amd/rlif.py, naer2_amd/csrc/rlif.hip).

Only tests/ and bench.py's gradient check use it.  Same equations as the kernels:
    I_t = alpha I_{t-1} + W_in x_t + W_rec s_{t-1};  V_t = (beta V_{t-1} + I_t) - theta s_{t-1}
    s_t = H(V_t - theta) with surrogate derivative gamma / (1 + slope |V_t - theta|)^2  (mode 0), or
    s_t = f(V_t - theta), f(x) = x / (1 + slope |x|)                                      (mode 1, smooth)
    J_t = alpha J_{t-1} + W_out s_t;  R_t = beta R_{t-1} + J_t
`forced` (mode 0) replaces the spike values of the forward by given ones (those of the fp32 run), so that
an fp32/fp64 comparison is not thrown off by potentials within rounding of the threshold; the surrogate
derivative is still taken at the oracle's own float64 potentials.
"""
from __future__ import annotations

import torch


class _Spike(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, slope, gamma, forced):
        ctx.save_for_backward(x)
        ctx.slope, ctx.gamma = slope, gamma
        return forced.to(x.dtype) if forced is not None else (x > 0).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        return g * ctx.gamma / (1.0 + ctx.slope * x.abs()) ** 2, None, None, None


def smooth_spike(x: torch.Tensor, slope: float) -> torch.Tensor:
    return x / (1.0 + slope * x.abs())


def unroll(x, w_in, w_rec, w_out, alpha, beta, theta, slope, gamma=1.0, mode=0, forced=None):
    """(R [T,B,O], spikes [T,B,H], V [T,B,H]) in the dtype of the weights (use float64)."""
    T, B, _ = x.shape
    H, O = w_rec.shape[0], w_out.shape[0]
    dt = w_in.dtype
    x = x.to(dt)
    I = torch.zeros(B, H, dtype=dt, device=x.device)
    V = torch.zeros(B, H, dtype=dt, device=x.device)
    s = torch.zeros(B, H, dtype=dt, device=x.device)
    J = torch.zeros(B, O, dtype=dt, device=x.device)
    R = torch.zeros(B, O, dtype=dt, device=x.device)
    rs, ss, vs = [], [], []
    for t in range(T):
        I = alpha * I + (x[t] @ w_in.t() + s @ w_rec.t())
        V = (beta * V + I) - theta * s
        if mode == 0:
            s = _Spike.apply(V - theta, slope, gamma, None if forced is None else forced[t])
        else:
            s = smooth_spike(V - theta, slope)
        J = alpha * J + s @ w_out.t()
        R = beta * R + J
        rs.append(R)
        ss.append(s)
        vs.append(V)
    return torch.stack(rs), torch.stack(ss), torch.stack(vs)


def loss_fn(r: torch.Tensor, labels: torch.Tensor) -> torch.Tensor:
    return torch.nn.functional.cross_entropy(r.max(dim=0).values, labels)


def grads(x, w_in, w_rec, w_out, labels, **kw):
    """(loss, [g_w_in, g_w_rec, g_w_out], spikes) by autograd through the unroll."""
    ws = [w.detach().clone().requires_grad_(True) for w in (w_in, w_rec, w_out)]
    r, s, _ = unroll(x, *ws, **kw)
    loss = loss_fn(r, labels)
    gs = torch.autograd.grad(loss, ws)
    return loss.detach(), [g.detach() for g in gs], s.detach()


def rel_err(a: torch.Tensor, b: torch.Tensor) -> float:
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm().clamp_min(1e-300))
