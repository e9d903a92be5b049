"""Helpers of the GPU parity tests: networks from explicit topologies, and the oracle run that mirrors a
built runtime."""
from __future__ import annotations

import numpy as np
import torch

from naer2_amd import FixedInDegree, GainSynapse, LIFDelta, LIFExp, PlasticSynapse, StaticSynapse, node, pattern
from naer2_amd.params import params_bytes
from oracle import oracle
from tests.helpers import probe_params


@node
class ProbeNode(LIFDelta):
    """Delta LIF whose potential after a step is that step's ring input (decay 0, no threshold)."""

    def params(self, dt):
        return probe_params(dt)


@pattern
class Given:
    """A pattern returning a fixed (src, tgt) link list (source-major)."""

    def __init__(self, src, tgt):
        self.src = torch.as_tensor(src, dtype=torch.int64)
        self.tgt = torch.as_tensor(tgt, dtype=torch.int64)

    def links(self, n_src, n_tgt, generator):
        return self.src.to(generator.device), self.tgt.to(generator.device)


def network_from_topology(node_instance, n, topo, dt=0.1, seed=0, device="cuda"):
    """One population of n neurons with one connection carrying the canonical topology topo."""
    src_ptr, tgt, w, d = topo
    pop = node_instance * n
    net = pop.network
    net.resolution, net.seed, net.device = dt, seed, torch.device(device)
    net._generator = None
    src = np.repeat(np.arange(n), np.diff(src_ptr))
    conn = pop >> Given(src, tgt) >> pop
    conn.weight = torch.as_tensor(w)
    conn.delay = torch.as_tensor(d.astype(np.float64) * dt, dtype=torch.float32)
    return net, pop


def mixed_edges(seed=1, kinds=("static", "gain", "plastic", "static"), ports=("current",) * 4, sizes=(120, 60),
                in_degree=(12, 6), device="cpu"):
    """Two populations E, I of drive-fired exp LIF neurons (all start at V_reset: the first spikes come in
    a synchronous burst); connections E->E, E->I, I->E, I->I of the given edge kinds and ports, random
    weights, gains (0.5..2), factors (+-1e-3) and delays (1..8 steps)."""
    rng = np.random.default_rng(seed)
    E, I = LIFExp(K_ext=2000) * sizes[0], LIFExp(K_ext=1800) * sizes[1]
    cls = {"static": lambda: StaticSynapse(0.0), "gain": lambda: GainSynapse(0.0, 1.0),
           "plastic": lambda: PlasticSynapse(0.0, 0.0)}
    conns = []
    pairs = ((E, E, in_degree[0]), (E, I, in_degree[0]), (I, E, in_degree[1]), (I, I, in_degree[1]))
    for (s, t, K), kind, port in zip(pairs, kinds, ports):
        c = s >> cls[kind]() @ port @ FixedInDegree(K) >> t
        sign = 1.0 if s is E else -3.0
        scale = 0.4 if port == "potential" else 0.1
        c.weight = torch.as_tensor(sign * scale * rng.random(c.shape[0]), dtype=torch.float32)
        c.delay = torch.as_tensor(rng.integers(1, 9, c.shape[0]) * 0.1, dtype=torch.float32)
        if kind == "gain":
            c.gain = torch.as_tensor(rng.uniform(0.5, 2.0, c.shape[0]), dtype=torch.float32)
        if kind == "plastic":
            c.factor = torch.as_tensor(rng.uniform(-1e-3, 1e-3, c.shape[0]), dtype=torch.float32)
        conns.append(c)
    net = E.network
    net.resolution, net.seed, net.device = 0.1, seed, torch.device(device)
    net._generator = None
    return net, conns


def oracle_state(rt):
    return oracle.State(rt.n, rt.d_slots, rt.v.cpu().numpy(), None if rt.isyn is None else rt.isyn.cpu().numpy(),
                        rt.ref.cpu().numpy(), rt.ring.cpu().numpy())


def oracle_run(rt, st, step0, n_steps, record=True):
    c = rt.canon.to("cpu")
    model = oracle.LIF_DELTA if rt.model == 1 else oracle.LIF_EXP
    return oracle.run(model, rt.pop_first, params_bytes(rt.params), rt.d_slots, rt.net.seed, rt.net.instance,
                      c.src_ptr.numpy(), c.tgt.numpy(), c.w.numpy(), c.d.numpy(), st, step0, n_steps, record)


def oracle_run_edges(rt, st, ring_v, w, step0, n_steps, record=True):
    """The oracle's edge step (or_run_edges) from the runtime's canonical table and its payload (the runtime
    keeps both: Network.keep_canonical); updates st, ring_v and the canonical weights w in place."""
    from naer2_amd.network import PAYLOAD_ATTR, PAYLOAD_FLAGS
    c = rt.canon.to("cpu")
    fl = c.extra[PAYLOAD_FLAGS].numpy()
    kind, port = ((fl >> 8) & 3).astype(np.uint8), ((fl >> 10) & 1).astype(np.uint8)
    attr = c.extra[PAYLOAD_ATTR].contiguous().view(torch.float32).numpy()
    return oracle.run_edges(rt.pop_first, params_bytes(rt.params), rt.d_slots, rt.net.seed, rt.net.instance,
                            c.src_ptr.numpy(), c.tgt.numpy(), w, c.d.numpy(), kind, port, attr, st, ring_v, step0,
                            n_steps, record)


def canonical_weights(rt) -> np.ndarray:
    """The runtime's current link weights in canonical order."""
    from naer2_amd.network import PAYLOAD_INPUT
    return rt.link_weights().cpu().numpy()[rt.canon.extra[PAYLOAD_INPUT].cpu().numpy()]


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


def assert_state_equal(rt, st):
    assert bits_equal(rt.v.cpu().numpy(), st.v), "membrane potentials differ"
    assert np.array_equal(rt.ref.cpu().numpy(), st.ref), "refractory counters differ"
    if rt.isyn is not None:
        assert bits_equal(rt.isyn.cpu().numpy(), st.isyn), "synaptic currents differ"
    assert bits_equal(rt.ring.cpu().numpy(), st.ring), "ring contents differ"


__all__ = ["ProbeNode", "Given", "network_from_topology", "mixed_edges", "oracle_state", "oracle_run",
           "oracle_run_edges", "canonical_weights", "bits_equal", "assert_state_equal", "LIFExp"]
