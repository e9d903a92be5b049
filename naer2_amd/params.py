"""Per-population constants of the fixed-dt step: fp32 propagators and the Poisson inversion table.

With a fixed dt the exact propagators of the reference's update hooks are constants, computed here once in
float64 and rounded to fp32, so the kernel needs no exp:
  - delta LIF (tools/brunel_reference.py:27): decay = exp(-dt / tau_m)
  - exp-synapse LIF (test/networks/PD14/neuron.py:37-42): P_V = exp(-dt/tau_m), P_I = exp(-dt/tau_syn),
    P_VI = R tau_syn / (tau_m - tau_syn) (P_V - P_I)
Refractory periods become whole steps:
  - delta LIF: active iff t > ref_until (brunel_reference.py:38), ref_until = t_spike + t_ref, so a neuron
    stays inactive for the round(t_ref/dt) steps after its spike;
  - exp-synapse LIF: refractory iff t_end < t_ref_end (test/fused_reference.py:94), t_ref_end = t_end of the
    spiking step + t_ref, so round(t_ref/dt) - 1 steps are clamped to V_reset.
The background draw (brunel_reference.py:35, torch.poisson per step) is Poisson inversion of one 32-bit
Philox word against thr[j] = min(floor(2^32 CDF(j)), 2^32 - 1).
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import POISSON_K, PopParams

F32 = np.float32


def poisson_thresholds(lam: float, k: int = POISSON_K) -> np.ndarray:
    """uint32 [k]: thr[j] = min(floor(2^32 * P(X <= j)), 2^32 - 1) for X ~ Poisson(lam); all 2^32 - 1
    for lam == 0 (no drive)."""
    if lam < 0 or not math.isfinite(lam):
        raise ValueError(f"Poisson mean {lam} must be finite and >= 0")
    out = np.empty(k, dtype=np.uint64)
    p = math.exp(-lam)
    cdf = p
    for j in range(k):
        out[j] = min(int(math.floor(cdf * 4294967296.0)), 4294967295)
        p *= lam / (j + 1)
        cdf += p
    out = np.maximum.accumulate(out)
    if lam == 0:
        out[:] = 4294967295
    return out.astype(np.uint32)


def _f32(x: float) -> float:
    return float(F32(x))


def lif_delta_params(dt: float, tau_m: float, theta: float, v_reset: float, t_ref: float, j_ext: float,
                     lam: float) -> PopParams:
    """Delta-synapse LIF (tools/brunel_reference.py:14-48)."""
    p = PopParams()
    p.ref_steps = int(round(t_ref / dt))
    p.decay = _f32(math.exp(-dt / tau_m))
    p.v_th = _f32(theta)
    p.v_reset = _f32(v_reset)
    p.j_ext = _f32(j_ext) if lam > 0 else 0.0
    p.poisson_thr[:] = [int(x) for x in poisson_thresholds(lam)]
    return p


def lif_exp_params(dt: float, t_mem: float, t_syn: float, r_mem: float, t_ref: float, v_rest: float,
                   v_thresh: float, v_reset: float, j_ext: float, lam: float) -> PopParams:
    """Exponential-current-synapse LIF with the exact propagator (test/networks/PD14/neuron.py:6-60)."""
    p = PopParams()
    pv = math.exp(-dt / t_mem)
    pi = math.exp(-dt / t_syn)
    p.ref_steps = max(0, int(round(t_ref / dt)) - 1)
    p.decay = _f32(pv)
    p.p_i = _f32(pi)
    p.p_vi = _f32(r_mem * t_syn / (t_mem - t_syn) * (pv - pi))
    p.v_rest = _f32(v_rest)
    p.v_th = _f32(v_thresh)
    p.v_reset = _f32(v_reset)
    p.j_ext = _f32(j_ext) if lam > 0 else 0.0
    p.poisson_thr[:] = [int(x) for x in poisson_thresholds(lam)]
    return p


def params_bytes(pops: list[PopParams]) -> bytes:
    """The parameter blocks as raw bytes (the oracle takes the same layout)."""
    return b"".join(bytes(p) for p in pops)
