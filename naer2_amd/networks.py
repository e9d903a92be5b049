"""The two plain workloads, built with the description API.

brunel(): Brunel (2000) model A, time-driven (tools/brunel_reference.py:14-48), with the default
  parameters N_E=2000, N_I=500, C_E=200, C_I=50, J=0.2 mV, g=5, eta=2, D=1.5 ms, tau=20 ms, theta=20 mV,
  V_r=10 mV, t_ref=2 ms, dt=0.1 ms.  External drive: Poisson counts per step of mean
  C_E eta nu_thr dt with nu_thr = theta / (J C_E tau) (test/networks/Brunel/net.py:83-91), weight J.
  Initial potentials uniform in [0, theta).
brunel_exact(): Brunel (2000) model A as the reference's own event-driven network (test/networks/Brunel/
  net.py:1-118): exact-time delta LIF neurons (E, I) and N_E Poisson sources P at eta nu_thr, fixed in-degree
  C_E from E, C_I = gamma C_E from I, C_E from P, one delay D everywhere; initial potentials uniform in
  [0, theta).
pd14(): the Potjans & Diesmann (2014) microcircuit of test/networks/PD14/net.py:25-108 without the thalamic
  population, with a Poisson background of K_ext inputs at 8 Hz (J_ext = 87.81 pA) instead of the DC shift
  of V_rest, 8 populations, fixed-total-number connectivity (connection.py:9-31), weights normal +-10% with
  the sign kept (clamped), L4E -> L2/3E doubled, inhibitory -4 J; delays normal (E 1.5 +- 0.75 ms,
  I 0.75 +- 0.375 ms) clipped at 0.1 ms; initial potentials normal per population.
"""
from __future__ import annotations

import torch

from .network import Network
from .nodes import FixedInDegree, FixedTotalNumber, LIFDelta, LIFDeltaExact, LIFExp, PoissonExact

_J_E = 0.08781
_J_I = -0.35124

PD14 = {
    "name": ["E_23", "I_23", "E_4", "I_4", "E_5", "I_5", "E_6", "I_6"],
    "K_ext": [1600, 1500, 2100, 1900, 2000, 1900, 2900, 2100],
    "size": [20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948],
    "V_mean": [-68.28, -63.16, -63.33, -63.45, -63.11, -61.66, -66.72, -61.45],
    "V_std": [5.36, 4.57, 4.74, 4.94, 4.94, 4.55, 5.46, 4.48],
    # rows = source population, columns = target population (test/networks/PD14/net.py:45-56)
    "probabilities": [
        [0.1009, 0.1346, 0.0077, 0.0691, 0.1004, 0.0548, 0.0156, 0.0364],
        [0.1689, 0.1371, 0.0059, 0.0029, 0.0622, 0.0269, 0.0066, 0.0010],
        [0.0437, 0.0316, 0.0497, 0.0794, 0.0505, 0.0257, 0.0211, 0.0034],
        [0.0818, 0.0515, 0.1350, 0.1597, 0.0057, 0.0022, 0.0166, 0.0005],
        [0.0323, 0.0755, 0.0067, 0.0033, 0.0831, 0.0600, 0.0572, 0.0277],
        [0.0, 0.0, 0.0003, 0.0, 0.3726, 0.3158, 0.0197, 0.0080],
        [0.0076, 0.0042, 0.0453, 0.1057, 0.0204, 0.0086, 0.0396, 0.0658],
        [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2252, 0.1443],
    ],
    "weight_mean": [_J_E, _J_I, _J_E, _J_I, _J_E, _J_I, _J_E, _J_I],
    "delay_mean": [1.5, 0.75, 1.5, 0.75, 1.5, 0.75, 1.5, 0.75],
}


def brunel(N_E: int = 2000, N_I: int = 500, C_E: int = 200, C_I: int = 50, J: float = 0.2, g: float = 5.0,
           eta: float = 2.0, D: float = 1.5, tau: float = 20.0, theta: float = 20.0, V_r: float = 10.0,
           t_ref: float = 2.0, dt: float = 0.1, seed: int = 0, device=None, instance: int = 0):
    """(network, {"E": E, "I": I})."""
    nu_thr = theta / (J * C_E * tau)              # per ms
    rate_ext = C_E * eta * nu_thr                  # background events per ms per neuron
    node = LIFDelta(tau=tau, theta=theta, V_r=V_r, t_ref=t_ref, J_ext=J, rate_ext=rate_ext)
    E = node * N_E
    I = LIFDelta(tau=tau, theta=theta, V_r=V_r, t_ref=t_ref, J_ext=J, rate_ext=rate_ext) * N_I
    net = E.network
    net.merge(I.network)
    _configure(net, dt, seed, device, instance)
    gen = net.generator
    E.V = torch.rand(N_E, generator=gen, device=gen.device) * theta
    I.V = torch.rand(N_I, generator=gen, device=gen.device) * theta
    for target in (E, I):
        ce = E >> FixedInDegree(C_E) >> target
        ce.weight, ce.delay = J, D
        ci = I >> FixedInDegree(C_I) >> target
        ci.weight, ci.delay = -g * J, D
    return net, {"E": E, "I": I}


BRUNEL_EXACT = {
    "N": 10000,  # N_E + N_I
    "f_E": 0.8,
    "epsilon": 0.1,  # connection probability (strong scaling: C_E = epsilon N_E)
    "gamma": 0.25,   # C_I / C_E
    "J": 0.1,        # mV
    "g": 5.0,
    "eta": 2.0,
    "D": 1.5,        # ms, every connection
    "neuron": {"tau": 20.0, "theta": 20.0, "V_r": 10.0, "t_ref": 2.0},
}


def brunel_exact(N: int | None = None, K: int | None = None, eta: float | None = None, g: float | None = None,
                 desc: dict = BRUNEL_EXACT, seed: int = 0, device=None, instance: int = 0):
    """(network, {"E", "I", "P": populations}): test/networks/Brunel/net.py:47-80 (build).  N: N_E + N_I;
    K: a fixed in-degree C_E (None: epsilon N_E)."""
    N = N or desc["N"]
    eta = desc["eta"] if eta is None else eta
    g = desc["g"] if g is None else g
    J, D, nrn = desc["J"], desc["D"], desc["neuron"]
    N_E = round(desc["f_E"] * N)
    N_I = N - N_E
    C_E = K if K is not None else round(desc["epsilon"] * N_E)
    C_I = max(round(desc["gamma"] * C_E), 1)
    rate = eta * nrn["theta"] / (J * C_E * nrn["tau"])  # spikes per ms
    E = LIFDeltaExact(**nrn) * N_E
    I = LIFDeltaExact(**nrn) * N_I
    P = PoissonExact(rate) * N_E
    net = E.network
    net.merge(I.network)
    net.merge(P.network)
    _configure(net, 0.1, seed, device, instance)
    gen = net.generator
    E.V = torch.rand(N_E, generator=gen, device=gen.device) * nrn["theta"]
    I.V = torch.rand(N_I, generator=gen, device=gen.device) * nrn["theta"]
    for target in (E, I):
        for src, C, w in ((E, C_E, J), (I, C_I, -g * J), (P, C_E, J)):
            conn = src >> FixedInDegree(C) >> target
            conn.weight, conn.delay = w, D
    return net, {"E": E, "I": I, "P": P}


def brunel_exact_theory(N: int | None = None, K: int | None = None, eta: float | None = None,
                        desc: dict = BRUNEL_EXACT) -> dict:
    """test/networks/Brunel/net.py:83-91: C_E, C_I, nu_thr and the external rate per source (Hz)."""
    N = N or desc["N"]
    eta = desc["eta"] if eta is None else eta
    N_E = round(desc["f_E"] * N)
    C_E = K if K is not None else round(desc["epsilon"] * N_E)
    nu_thr = desc["neuron"]["theta"] / (desc["J"] * C_E * desc["neuron"]["tau"])
    return {"C_E": C_E, "C_I": max(round(desc["gamma"] * C_E), 1), "nu_thr_Hz": 1000 * nu_thr,
            "rate_ext_Hz": 1000 * eta * nu_thr}


def pd14(scale: float = 1.0, seed: int = 0, dt: float = 0.1, device=None, instance: int = 0,
         k_ext_scale: float = 1.0, weight_scale=None, sizes=None, prob_scale: float = 1.0):
    """(network, {name: population}); scale shrinks every population (connectivity follows the rule).
    weight_scale: [8][8] factors alpha[target][source] on the connections' weights (default all 1);
    sizes: explicit population sizes instead of scale; prob_scale multiplies the connection probabilities
    (capped at 0.9), for small subsampled circuits."""
    pops = {}
    net = None
    size_list = sizes if sizes is not None else [max(1, int(round(sz * scale))) for sz in PD14["size"]]
    for name, k_ext, n in zip(PD14["name"], PD14["K_ext"], size_list):
        n = int(n)
        pops[name] = LIFExp(K_ext=int(round(k_ext * k_ext_scale))) * n
        if net is None:
            net = pops[name].network
        else:
            net.merge(pops[name].network)
    _configure(net, dt, seed, device, instance)
    gen = net.generator
    dev = gen.device
    for name, vm, vs in zip(PD14["name"], PD14["V_mean"], PD14["V_std"]):
        p = pops[name]
        p.potential = vm + vs * torch.randn(p.n, generator=gen, device=dev)
    for si, src in enumerate(PD14["name"]):
        for ti, tgt in enumerate(PD14["name"]):
            prob = min(0.9, PD14["probabilities"][si][ti] * prob_scale)
            if prob == 0.0:
                continue
            w_mean = 2 * _J_E if (src == "E_4" and tgt == "E_23") else PD14["weight_mean"][si]
            conn = pops[src] >> FixedTotalNumber(prob) >> pops[tgt]
            n_links = conn.shape[0]
            d_mean = PD14["delay_mean"][si]
            conn.delay = (d_mean + 0.5 * d_mean * torch.randn(n_links, generator=gen, device=dev)).clamp_(min=0.1)
            w = w_mean + 0.1 * abs(w_mean) * torch.randn(n_links, generator=gen, device=dev)
            w = w.clamp_(min=0.0) if w_mean >= 0 else w.clamp_(max=0.0)
            if weight_scale is not None:
                w = w * float(weight_scale[ti][si])
            conn.weight = w
    return net, pops


def sweep_params(n_instances: int = 256, seed: int = 0, n_min: int = 1000, n_max: int = 50000) -> list[dict]:
    """Parameters of a batched Brunel sweep, all drawn from seed: N log-uniform in [n_min, n_max] (E:I =
    4:1), fixed in-degree 10 % of N_E and N_I, g uniform in [3, 8], external drive ratio eta uniform in
    [0.9, 4.0]; the rest as brunel()'s defaults."""
    import numpy as np
    rng = np.random.default_rng(np.random.PCG64(seed))
    out = []
    for i in range(n_instances):
        n = int(round(float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))))
        n_e = int(round(0.8 * n))
        n_i = n - n_e
        out.append({"N_E": n_e, "N_I": n_i, "C_E": max(1, int(round(0.1 * n_e))), "C_I": max(1, int(round(0.1 * n_i))),
                    "g": float(rng.uniform(3.0, 8.0)), "eta": float(rng.uniform(0.9, 4.0)), "seed": seed * 100003 + i,
                    "instance": i})
    return out


def brunel_sweep(params: list[dict], device=None, keep_canonical: bool = False, record: bool = False):
    """One built network per parameter set (sweep_params), ready to batch (FusedRound)."""
    nets = []
    for p in params:
        net, _ = brunel(N_E=p["N_E"], N_I=p["N_I"], C_E=p["C_E"], C_I=p["C_I"], g=p["g"], eta=p["eta"],
                        seed=p["seed"], instance=p["instance"], device=device)
        net.keep_canonical = keep_canonical
        net.keep_description = keep_canonical  # hundreds of networks: only the runtime's tables stay on the GPU
        net.record = record
        net.build()
        nets.append(net)
    return nets


def _configure(net: Network, dt: float, seed: int, device, instance: int) -> None:
    net.resolution = dt
    net.seed = seed
    net.instance = instance
    if device is not None:
        net.device = torch.device(device)
    net._generator = None
