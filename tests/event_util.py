"""Test helpers for the event-driven path: a DSL network in the oracle's terms."""
from __future__ import annotations

import numpy as np
import torch

from oracle import event as ev


def oracle_net(net, canon=None):
    """(EvNet, v0) of an event-driven DSL network, from its description (before or after the build)."""
    from naer2_amd.topology import canonical

    pops = sorted(net.populations, key=lambda p: p._order)
    index = {p: i for i, p in enumerate(pops)}
    sizes = [p.n for p in pops]
    pf = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    first = {p: int(f) for p, f in zip(pops, pf[:-1])}
    n = int(pf[-1])
    conns = list(net.connections)
    if canon is None:
        src = [c.src.cpu().to(torch.int64) + first[c.source] for c in conns]
        tgt = [c.tgt.cpu().to(torch.int64) + first[c.target] for c in conns]
        w = [c.weight.cpu() for c in conns]
        cat = (lambda xs, t: torch.cat(xs) if xs else torch.zeros(0, dtype=t))
        s = cat(src, torch.int64)
        canon = canonical(n, s, cat(tgt, torch.int64), cat(w, torch.float32), torch.ones(s.numel(), dtype=torch.int32))
    delays = np.array([float(np.float32(float(c.delay.reshape(-1)[0]))) for c in conns], np.float32)
    window = float(delays.min()) if conns else 1.0
    eps = []
    for p in pops:
        e = ev.EvPop()
        if p.family == "event_lif":
            e.kind = ev.NEURON
            inst = p.instance
            e.tau_inv = float(np.float32(1.0 / inst.tau))
            e.theta, e.v_reset, e.t_ref = inst.theta, inst.V_r, inst.t_ref
        elif p.family == "event_poisson":
            e.kind = ev.POISSON
            e.rate_inv = float(np.float32(1.0 / p.instance.rate)) if p.instance.rate > 0 else float("inf")
        else:
            e.kind = ev.TRAIN
        eps.append(e)
    train_ptr = train_t = None
    if any(p.family == "event_train" for p in pops):
        counts = np.zeros(n, np.int64)
        for p in pops:
            if p.family == "event_train":
                for i, ts in enumerate(p.instance.times):
                    counts[first[p] + i] = len(ts)
        train_ptr = np.zeros(n + 1, np.int64)
        train_ptr[1:] = np.cumsum(counts)
        train_t = np.zeros(int(train_ptr[-1]), np.float32)
        for p in pops:
            if p.family == "event_train":
                for i, ts in enumerate(p.instance.times):
                    j = first[p] + i
                    train_t[train_ptr[j]:train_ptr[j + 1]] = np.asarray(ts, np.float32)
    v0 = np.zeros(n, np.float32)
    for p in pops:
        if "V" in p._init:
            v0[first[p]:first[p] + p.n] = p._init["V"].cpu().numpy()
        elif p.family == "event_lif":
            v0[first[p]:first[p] + p.n] = float(p.instance.V)
    evnet = ev.EvNet(pf, eps, np.array([index[c.source] for c in conns], np.int32),
                     np.array([index[c.target] for c in conns], np.int32), window,
                     canon.src_ptr.cpu().numpy().astype(np.int64), canon.tgt.cpu().numpy().astype(np.int32),
                     canon.w.cpu().numpy().astype(np.float32), seed=int(net.seed), instance=int(net.instance),
                     train_ptr=train_ptr, train_t=train_t, conn_delay=delays)
    return evnet, v0


def train_net(times_per_source: list, targets: int, links: list, V0, window: float = 1.5, device="cpu",
              neuron: dict | None = None):
    """A network of spike-train sources (one population per entry of times_per_source, each a list of
    per-source time lists) into one population of `targets` exact-time neurons.  links: per source
    population, a list of (source index, target index, weight) in the connection's link order."""
    import naer2_amd as n2
    from naer2_amd.dsl import Points

    nrn = neuron or {"tau": 20.0, "theta": 20.0, "V_r": 10.0, "t_ref": 2.0}
    T = n2.LIFDeltaExact(**nrn) * targets
    net = T.network
    net.device = torch.device(device)
    srcs = []
    for times, lk in zip(times_per_source, links):
        S = n2.SpikeTrains(times) * len(times)
        net.merge(S.network)
        srcs.append((S, lk))
    for S, lk in srcs:
        pairs = sorted(range(len(lk)), key=lambda j: lk[j][0])  # source-major, stable
        s_idx = torch.tensor([lk[j][0] for j in pairs], dtype=torch.int64)
        t_idx = torch.tensor([lk[j][1] for j in pairs], dtype=torch.int64)
        w = torch.tensor([lk[j][2] for j in pairs], dtype=torch.float32)

        class _Given:
            def links(self, n_src, n_tgt, generator, s_idx=s_idx, t_idx=t_idx):
                return s_idx.to(generator.device), t_idx.to(generator.device)

            def __call__(self, *args):
                raise NotImplementedError

        conn = S >> n2.pattern(_Given()) >> T
        conn.weight = w
        conn.delay = window
    T.V = torch.as_tensor(V0, dtype=torch.float32).expand(targets).clone()
    return net, T
