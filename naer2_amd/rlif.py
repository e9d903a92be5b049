"""Surrogate-gradient BPTT through a recurrent LIF network: a torch.autograd.Function over the hand-written
forward and adjoint kernels of naer2_amd/csrc/rlif.hip (C-ABI n2h_rlif_*).

    net = RecurrentLIF(n_in=700, n_hidden=1024, n_out=20, generator=g)   # W ~ N(0, 1/fan_in)
    r, s = net(x)                       # x: [T, B, n_in] input spikes (0/1); r: readout [T, B, n_out]
    loss = max_over_time_ce(r, labels)  # cross-entropy of the max-over-time readout
    loss.backward()                     # adjoint kernels, checkpoints every `checkpoint` steps

Dynamics (dt = 1 ms; alpha = exp(-dt/tau_syn), beta = exp(-dt/tau_m)):
    I_t = alpha I_{t-1} + W_in x_t + W_rec s_{t-1};  V_t = beta V_{t-1} + I_t - theta s_{t-1}
    s_t = H(V_t - theta), surrogate ds/dV = 1 / (1 + slope |V_t - theta|)^2 (fast sigmoid, slope 25)
    J_t = alpha J_{t-1} + W_out s_t;  R_t = beta R_{t-1} + J_t (leaky readout integrators)
The reference has no autograd (naer2 is forward-only): a surrogate-gradient adjoint added beside its plain path.
Plain library GEMMs (the input projection, the readout projection and the weight gradients) run through
rocBLAS / hipBLASLt; delivery and the LIF update and their adjoints are the hand-written kernels: the
event-driven recurrent delivery, the compensated-fp32 LIF update, and per reverse step the delivery adjoint
W_rec^T gI on the matrix cores fused with the LIF adjoint.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


class RLIFLayer(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int32), ("n_hidden", ctypes.c_int32), ("n_steps", ctypes.c_int32),
                ("checkpoint", ctypes.c_int32), ("mode", ctypes.c_int32), ("alpha", ctypes.c_float),
                ("beta", ctypes.c_float), ("theta", ctypes.c_float), ("slope", ctypes.c_float),
                ("gamma", ctypes.c_float), ("alpha_lo", ctypes.c_float), ("beta_lo", ctypes.c_float)]


def layer(B: int, H: int, T: int, cfg: "RLIFConfig") -> RLIFLayer:
    """The C layer description; the decays as float + float remainder (for the compensated state)."""
    a_hi, b_hi = float(np.float32(cfg.alpha)), float(np.float32(cfg.beta))
    return RLIFLayer(B, H, T, cfg.checkpoint, cfg.mode, a_hi, b_hi, cfg.theta, cfg.slope, cfg.gamma,
                     float(np.float32(cfg.alpha - a_hi)), float(np.float32(cfg.beta - b_hi)))


def _bind(lib):
    if getattr(lib, "_rlif_bound", False):
        return lib
    vp, i32, i64, f = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    P = ctypes.POINTER(RLIFLayer)
    lib.n2h_rlif_last_error.argtypes, lib.n2h_rlif_last_error.restype = [ctypes.c_char_p, ctypes.c_size_t], i32
    lib.n2h_rlif_forward.argtypes, lib.n2h_rlif_forward.restype = [P, vp, vp, vp, vp, vp, vp], i32
    lib.n2h_rlif_readout.argtypes, lib.n2h_rlif_readout.restype = [i32, i64, f, f, vp, vp, vp], i32
    lib.n2h_rlif_readout_adjoint.argtypes, lib.n2h_rlif_readout_adjoint.restype = [i32, i64, f, f, vp, vp, vp], i32
    lib.n2h_rlif_backward.argtypes, lib.n2h_rlif_backward.restype = [P, vp, vp, vp, vp, vp, vp, vp, vp], i32
    lib.n2h_rlif_delivery_adjoint.argtypes, lib.n2h_rlif_delivery_adjoint.restype = [P, vp, vp, vp, vp], i32
    lib.n2h_rlif_work_floats.argtypes, lib.n2h_rlif_work_floats.restype = [P], i64
    lib._rlif_bound = True
    return lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        _lib.load().n2h_rlif_last_error(buf, len(buf))
        raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")


@dataclass
class RLIFConfig:
    tau_m: float = 20.0
    tau_syn: float = 5.0
    theta: float = 1.0
    slope: float = 25.0
    dt: float = 1.0
    checkpoint: int = 25
    mode: int = 0         # 0: spikes; 1: the smooth model s = f(V - theta) (gradient checks)
    gamma: float = 1.0    # surrogate scale

    @property
    def alpha(self) -> float:
        return math.exp(-self.dt / self.tau_syn)

    @property
    def beta(self) -> float:
        return math.exp(-self.dt / self.tau_m)


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class RecurrentLIFFunction(torch.autograd.Function):
    """(x [T,B,C], w_in [H,C], w_rec [H,H], w_out [O,H], cfg) -> (readout R [T,B,O], spikes [T,B,H])."""

    @staticmethod
    def forward(ctx, x, w_in, w_rec, w_out, cfg: RLIFConfig):
        lib = _bind(_lib.load())
        T, B, C = x.shape
        H, O = w_rec.shape[0], w_out.shape[0]
        dev = x.device
        x = x.contiguous().float()
        u = torch.matmul(x.reshape(T * B, C), w_in.t()).reshape(T, B, H).contiguous()
        w_rec_t = w_rec.t().contiguous()
        spikes = torch.empty(T, B, H, device=dev)
        n_seg = (T + cfg.checkpoint - 1) // cfg.checkpoint
        ckpt = torch.empty(n_seg, 4, B, H, device=dev)
        state = torch.empty(4, B, H, device=dev)
        lay = layer(B, H, T, cfg)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _check(lib.n2h_rlif_forward(ctypes.byref(lay), _ptr(u), _ptr(w_rec_t), _ptr(state), _ptr(spikes),
                                    _ptr(ckpt), stream), "n2h_rlif_forward")
        z = torch.matmul(spikes.reshape(T * B, H), w_out.t()).reshape(T, B, O).contiguous()
        r = torch.empty(T, B, O, device=dev)
        _check(lib.n2h_rlif_readout(T, B * O, cfg.alpha, cfg.beta, _ptr(z), _ptr(r), stream), "n2h_rlif_readout")
        ctx.save_for_backward(x, w_rec, w_rec_t, w_out, u, spikes, ckpt)
        ctx.cfg = cfg
        ctx.layer = lay
        ctx.mark_non_differentiable(spikes)
        return r, spikes

    @staticmethod
    def backward(ctx, g_r, g_spikes):
        lib = _bind(_lib.load())
        x, w_rec, w_rec_t, w_out, u, spikes, ckpt = ctx.saved_tensors
        cfg, lay = ctx.cfg, ctx.layer
        T, B, C = x.shape
        H, O = w_rec.shape[0], w_out.shape[0]
        dev = x.device
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        g_r = g_r.contiguous().float()
        g_j = torch.empty(T, B, O, device=dev)
        _check(lib.n2h_rlif_readout_adjoint(T, B * O, cfg.alpha, cfg.beta, _ptr(g_r), _ptr(g_j), stream),
               "n2h_rlif_readout_adjoint")
        gj2 = g_j.reshape(T * B, O)
        s2 = spikes.reshape(T * B, H)
        g_w_out = gj2.t() @ s2
        gso = (gj2 @ w_out).reshape(T, B, H).contiguous()
        g_i = torch.empty(T, B, H, device=dev)
        work = torch.empty(int(lib.n2h_rlif_work_floats(ctypes.byref(lay))), device=dev)
        _check(lib.n2h_rlif_backward(ctypes.byref(lay), _ptr(u), _ptr(w_rec_t), _ptr(spikes), _ptr(ckpt),
                                     _ptr(gso), _ptr(g_i), _ptr(work), stream),
               "n2h_rlif_backward")
        gi2 = g_i.reshape(T * B, H)
        g_w_rec = g_i[1:].reshape(-1, H).t() @ spikes[:-1].reshape(-1, H)
        g_w_in = gi2.t() @ x.reshape(T * B, C)
        return None, g_w_in, g_w_rec, g_w_out, None


class RecurrentLIF(torch.nn.Module):
    """700 -> 1024 (recurrent LIF) -> 20 (leaky readout) by default; weights N(0, 1/fan_in)."""

    def __init__(self, n_in: int = 700, n_hidden: int = 1024, n_out: int = 20, cfg: RLIFConfig | None = None,
                 generator: torch.Generator | None = None, device=None):
        super().__init__()
        self.cfg = cfg or RLIFConfig()
        dev = device if device is not None else (generator.device if generator is not None else None)
        kw = dict(generator=generator, device=dev)
        self.w_in = torch.nn.Parameter(torch.randn(n_hidden, n_in, **kw) / math.sqrt(n_in))
        self.w_rec = torch.nn.Parameter(torch.randn(n_hidden, n_hidden, **kw) / math.sqrt(n_hidden))
        self.w_out = torch.nn.Parameter(torch.randn(n_out, n_hidden, **kw) / math.sqrt(n_hidden))

    def forward(self, x: torch.Tensor):
        return RecurrentLIFFunction.apply(x, self.w_in, self.w_rec, self.w_out, self.cfg)


def max_over_time_ce(r: torch.Tensor, labels: torch.Tensor) -> torch.Tensor:
    """Cross-entropy of the max-over-time readout ([T, B, O] -> [B, O])."""
    return torch.nn.functional.cross_entropy(r.max(dim=0).values, labels)


def poisson_input(batch: int, n_in: int, n_steps: int, dt_ms: float, generator: torch.Generator,
                  max_rate_hz: float = 100.0) -> torch.Tensor:
    """[T, B, n_in] 0/1 Poisson spike trains, rates uniform in [0, max_rate_hz] per (sample, channel)."""
    dev = generator.device
    rates = torch.rand(batch, n_in, generator=generator, device=dev) * max_rate_hz
    p = (rates * dt_ms * 1e-3).clamp(max=1.0)
    u = torch.rand(n_steps, batch, n_in, generator=generator, device=dev)
    return (u < p).float()
