"""TEST INFRASTRUCTURE ONLY: the FP64 PyTorch autograd unroll that checks the rate-loss adjoint of the
exponential-synapse LIF step (naer2_amd/adjoint.py, naer2_amd/csrc/adjoint.hip).

Only tests/ and bench.py's gradient check use it.  The forward is the step of the LIF_EXP kernel
(naer2_amd/csrc/naer2_hip.hip; PD14 neuron: test/networks/PD14/neuron.py:37-57, fused_reference.py:86-107)
with per-link delays, integer refractoriness and the Philox/Poisson background, in float64:
    x_t[k] = sum over links j -> k of alpha[pop k][pop j] w s_{t-d}[j]
    I' = I + x_t + j_ext count_t;  V^pre = V_rest + P_V (V - V_rest) + P_VI I' (or V_reset while refractory)
    I = P_I I';  s = spike(V^pre);  V = V^pre (1 - s) + V_reset s
loss = sum over populations of (rate_p - target_p)^2, rate_p in Hz.
mode "spike": s = H(V^pre - theta) (or the given `forced` spikes of an fp32 run), backward r h(V^pre) with
h = gamma (kappa/2) / dv / (1 + kappa |u|)^2 for |u| <= cutoff (else 0), u = (V^pre - theta) / dv,
dv = theta - V_reset;
mode "smooth": s = 1/2 + gamma (kappa/2) u' / (1 + kappa |u'|), u' = u clamped to [-cutoff, cutoff], no
refractoriness -- a differentiable model whose exact gradient central finite differences check, and whose
derivative is the surrogate h of the spike mode.
"""
from __future__ import annotations

import numpy as np
import torch

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85


def philox_word0(c0, c1, c2, c3, k0, k1) -> np.ndarray:
    """Word 0 of Philox4x32-10 (Salmon et al., SC'11), vectorised over numpy uint arrays."""
    m = np.uint64(0xFFFFFFFF)
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & m for c in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint64) & m
    k1 = np.asarray(k1, dtype=np.uint64) & m
    for r in range(10):
        p0 = np.uint64(M0) * c0
        p1 = np.uint64(M1) * c2
        h0, l0 = p0 >> np.uint64(32), p0 & m
        h1, l1 = p1 >> np.uint64(32), p1 & m
        c0, c1, c2, c3 = (h1 ^ c1 ^ k0) & m, l1, (h0 ^ c3 ^ k1) & m, l0
        if r < 9:
            k0 = (k0 + np.uint64(W0)) & m
            k1 = (k1 + np.uint64(W1)) & m
    return c0.astype(np.uint32)


def poisson_counts(thr: np.ndarray, seed: int, instance: int, step0: int, n_steps: int) -> np.ndarray:
    """int [n_steps, n]: the background counts of the step kernel (counter {neuron, step lo, step hi,
    instance}, key {seed lo, seed hi}; count = #thresholds <= the uniform word).  thr: uint32 [n, 32]."""
    n = thr.shape[0]
    ids = np.arange(n, dtype=np.uint64)
    out = np.zeros((n_steps, n), dtype=np.int64)
    for t in range(n_steps):
        s = step0 + t
        u = philox_word0(ids, np.uint64(s & 0xFFFFFFFF), np.uint64(s >> 32), np.uint64(instance),
                         np.uint64(seed & 0xFFFFFFFF), np.uint64((seed >> 32) & 0xFFFFFFFF))
        out[t] = (u[:, None] >= thr).sum(axis=1)
    return out


def surrogate(vpre, theta, dv, kappa: float, cutoff: float, gamma: float):
    """h(V) = gamma (kappa/2) / dv / (1 + kappa |u|)^2 for |u| <= cutoff, else 0; u = (V - theta) / dv."""
    u = (vpre - theta).abs() / dv
    h = (gamma * 0.5 * kappa / dv) / (1.0 + kappa * u) ** 2
    return torch.where(u <= cutoff, h, torch.zeros_like(h))


def smooth_spike(vpre, theta, dv, kappa: float, cutoff: float, gamma: float):
    """1/2 + gamma (kappa/2) u' / (1 + kappa |u'|), u' = u clamped to [-cutoff, cutoff]: its derivative is the
    surrogate."""
    u = ((vpre - theta) / dv).clamp(-cutoff, cutoff)
    return 0.5 + gamma * 0.5 * kappa * u / (1.0 + kappa * u.abs())


class _Spike(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vpre, r, theta, dv, sur, forced):
        ctx.save_for_backward(vpre, r, theta, dv)
        ctx.sur = sur
        if forced is not None:
            return forced.to(vpre.dtype)
        return ((vpre > theta) & (r > 0)).to(vpre.dtype)

    @staticmethod
    def backward(ctx, g):
        vpre, r, theta, dv = ctx.saved_tensors
        return g * r * surrogate(vpre, theta, dv, *ctx.sur), None, None, None, None, None


def unroll(net: dict, alpha: torch.Tensor, n_steps: int, counts: np.ndarray, mode: str = "spike",
           forced: np.ndarray | None = None, kappa: float = 25.0, frozen_reset: np.ndarray | None = None,
           cutoff: float = float("inf"), gamma: float = 1.0):
    """(spikes [T, n], rates [P] in Hz, vpre [T, n], r [T, n] the non-refractory gate) in float64.
    net: n, pop_of [n], per-population
    arrays P_V, P_I, P_VI, v_rest, v_th, v_reset, j_ext, ref_steps, the links src, tgt, w (base weights), d
    (steps), and the initial v0, i0; alpha [P, P] (target, source).
    The reset carries no gradient: in mode "spike" V = V^pre (1 - s) + V_reset s with s detached there; in
    mode "smooth" the reset uses frozen_reset[t] (e.g. the model's own spikes of an unperturbed run), a
    fixed input, which makes the smooth model's exact gradient the detached-reset one; frozen_reset None:
    the reset uses the live s (gradient through the reset)."""
    f64 = torch.float64
    n = int(net["n"])
    pop = torch.as_tensor(net["pop_of"], dtype=torch.long)
    col = lambda k: torch.as_tensor(np.asarray(net[k], dtype=np.float64))[pop]
    P_V, P_I, P_VI, v_rest, v_th, v_reset, j_ext = (col(k) for k in
                                                    ("P_V", "P_I", "P_VI", "v_rest", "v_th", "v_reset", "j_ext"))
    ref_steps = torch.as_tensor(np.asarray(net["ref_steps"]), dtype=torch.long)[pop]
    dv = v_th - v_reset
    src = torch.as_tensor(net["src"], dtype=torch.long)
    tgt = torch.as_tensor(net["tgt"], dtype=torch.long)
    d = torch.as_tensor(net["d"], dtype=torch.long)
    w_link = alpha[pop[tgt], pop[src]] * torch.as_tensor(net["w"], dtype=f64)
    D = int(d.max()) if d.numel() else 1
    groups = [(k, (d == k).nonzero().flatten()) for k in range(1, D + 1)]
    groups = [(k, idx) for k, idx in groups if idx.numel()]
    ring = [torch.zeros(n, dtype=f64) for _ in range(D + 1)]
    V = torch.as_tensor(net["v0"], dtype=f64)
    I = torch.as_tensor(net["i0"], dtype=f64)
    R = torch.zeros(n, dtype=torch.long)
    cnt = torch.as_tensor(counts, dtype=f64)
    ss, vs, rs = [], [], []
    for t in range(n_steps):
        slot = t % (D + 1)
        x = ring[slot]
        ring[slot] = torch.zeros(n, dtype=f64)
        ip = I + (x + j_ext * cnt[t])
        r = (R == 0).to(f64)
        vpre = r * (v_rest + P_V * (V - v_rest) + P_VI * ip) + (1.0 - r) * v_reset
        I = P_I * ip
        if mode == "spike":
            fz = None if forced is None else torch.as_tensor(forced[t])
            s = _Spike.apply(vpre, r, v_th, dv, (kappa, cutoff, gamma), fz)
            sd = s.detach() > 0
            R = torch.where(R > 0, R - 1, torch.where(sd, ref_steps, torch.zeros_like(R)))
        else:
            s = smooth_spike(vpre, v_th, dv, kappa, cutoff, gamma)
        if mode == "spike":
            sr = s.detach()
        elif frozen_reset is not None:
            sr = torch.as_tensor(frozen_reset[t], dtype=f64)
        else:
            sr = s
        V = vpre * (1.0 - sr) + v_reset * sr
        c = w_link * s[src]
        for k, idx in groups:
            q = (t + k) % (D + 1)
            ring[q] = ring[q].index_add(0, tgt[idx], c[idx])
        ss.append(s)
        vs.append(vpre)
        rs.append(r)
    S = torch.stack(ss)
    P = int(pop.max()) + 1
    dt_s = float(net["dt_ms"]) * 1e-3
    sizes = torch.bincount(pop, minlength=P).to(f64)
    per_pop = torch.zeros(P, dtype=f64).index_add(0, pop, S.sum(dim=0))
    rates = per_pop / (sizes * n_steps * dt_s)
    return S, rates, torch.stack(vs), torch.stack(rs)


def loss(rates: torch.Tensor, targets) -> torch.Tensor:
    return ((rates - torch.as_tensor(targets, dtype=rates.dtype)) ** 2).sum()


def net_from(pop_first, params, canon, v0, i0, dt_ms: float) -> dict:
    """The oracle's description of a built network: pop_first, its populations' n2h_pop_params, the
    canonical synapse table (on the CPU) and the initial state."""
    n = int(pop_first[-1])
    pop_of = np.repeat(np.arange(len(params)), np.diff(np.asarray(pop_first)))
    src_ptr = canon.src_ptr.cpu().numpy()
    return {
        "n": n, "pop_of": pop_of, "dt_ms": dt_ms,
        "P_V": [p.decay for p in params], "P_I": [p.p_i for p in params], "P_VI": [p.p_vi for p in params],
        "v_rest": [p.v_rest for p in params], "v_th": [p.v_th for p in params],
        "v_reset": [p.v_reset for p in params], "j_ext": [p.j_ext for p in params],
        "ref_steps": [p.ref_steps for p in params],
        "thr": np.stack([np.array(list(p.poisson_thr), dtype=np.uint32) for p in params])[pop_of],
        "src": np.repeat(np.arange(n), np.diff(src_ptr)), "tgt": canon.tgt.cpu().numpy().astype(np.int64),
        "w": canon.w.cpu().numpy().astype(np.float64), "d": canon.d.cpu().numpy().astype(np.int64),
        "v0": np.asarray(v0, dtype=np.float64), "i0": np.asarray(i0, dtype=np.float64),
    }


def loss_grad_per_spike(net: dict, rates, targets, n_steps: int) -> np.ndarray:
    """dL/ds of one spike of each population: 2 (rate_p - target_p) / (n_p T dt)."""
    pop = np.asarray(net["pop_of"])
    P = int(pop.max()) + 1
    sizes = np.bincount(pop, minlength=P)
    dt_s = float(net["dt_ms"]) * 1e-3
    return 2.0 * (np.asarray(rates, dtype=np.float64) - np.asarray(targets)) / (sizes * n_steps * dt_s)


def adjoint(net: dict, alpha, S, VPRE, RG, c_pop, kappa: float = 25.0, drop: str | None = None,
            with_reset: bool = False, cutoff: float = float("inf"), gamma: float = 1.0) -> np.ndarray:
    """dL/dalpha [P, P] by the reverse recursion the adjoint kernel implements, in float64 numpy, from the
    forward's spikes S, potentials before the threshold VPRE and non-refractory gates RG ([T, n] each):
        gs = c + sum over out-links of alpha w gx_{t+d}     (+ (V_reset - V^pre) gV with with_reset)
        gVpre = (1 - s) gV + r h(V^pre) gs;  gI' = r P_VI gVpre + P_I gI;  gV = r P_V gVpre;  gx_t = gI'
        dL/dalpha[a][pop j] += s_t[j] sum over j's links into a of w gx_{t+d}
    drop: leave one term out ("reset", "coupling", "delay"), to show the check notices."""
    S, VPRE, RG = (np.asarray(x, dtype=np.float64) for x in (S, VPRE, RG))
    T, n = S.shape
    pop = np.asarray(net["pop_of"])
    P = int(pop.max()) + 1
    col = lambda k: np.asarray(net[k], dtype=np.float64)[pop]
    P_V, P_I, P_VI, v_th, v_reset = (col(k) for k in ("P_V", "P_I", "P_VI", "v_th", "v_reset"))
    dv = v_th - v_reset
    src, tgt, w, d = (np.asarray(net[k]) for k in ("src", "tgt", "w", "d"))
    w = w.astype(np.float64)
    D = int(d.max()) if d.size else 1
    A = np.asarray(alpha, dtype=np.float64)
    c = np.asarray(c_pop, dtype=np.float64)[pop]
    gV = np.zeros(n)
    gI = np.zeros(n)
    gx = np.zeros((T + D + 2, n))
    gsum = np.zeros((n, P))
    shift = 1 if drop == "delay" else 0
    for t in range(T - 1, -1, -1):
        bins = np.zeros((n, P))
        np.add.at(bins, (src, pop[tgt]), w * gx[t + d - shift, tgt])
        coup = (A[:, pop] * bins.T).sum(axis=0)
        if drop == "coupling":
            coup = 0.0 * coup
        s, vp, r = S[t], VPRE[t], RG[t]
        gsum += s[:, None] * bins
        reset = 0.0 if drop == "reset" else (v_reset - vp) * gV
        gs = c + coup + (reset if with_reset else 0.0)
        u = np.abs(vp - v_th) / dv
        h = np.where(u <= cutoff, (gamma * 0.5 * kappa / dv) / (1.0 + kappa * u) ** 2, 0.0)
        gvpre = (1.0 - s) * gV + r * h * gs
        gip = r * P_VI * gvpre + P_I * gI
        gV = r * P_V * gvpre
        gI = gip
        gx[t] = gip
    G = np.zeros((P, P))
    for b in range(P):
        G[:, b] = gsum[pop == b].sum(axis=0)
    return G
