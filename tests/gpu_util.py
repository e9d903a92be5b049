"""Helpers of the GPU parity tests: networks from explicit topologies, and the oracle run that mirrors a
built runtime."""
from __future__ import annotations

import numpy as np
import torch

from naer2_amd import LIFDelta, LIFExp, node, pattern
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


def oracle_state(rt):
    return oracle.State(rt.n, rt.d_slots, rt.v.cpu().numpy(), None if rt.isyn is None else rt.isyn.cpu().numpy(),
                        rt.ref.cpu().numpy(), rt.ring.cpu().numpy())


def oracle_run(rt, st, step0, n_steps, record=True):
    c = rt.canon.to("cpu")
    model = oracle.LIF_DELTA if rt.model == 1 else oracle.LIF_EXP
    return oracle.run(model, rt.pop_first, params_bytes(rt.params), rt.d_slots, rt.net.seed, rt.net.instance,
                      c.src_ptr.numpy(), c.tgt.numpy(), c.w.numpy(), c.d.numpy(), st, step0, n_steps, record)


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


def assert_state_equal(rt, st):
    assert bits_equal(rt.v.cpu().numpy(), st.v), "membrane potentials differ"
    assert np.array_equal(rt.ref.cpu().numpy(), st.ref), "refractory counters differ"
    if rt.isyn is not None:
        assert bits_equal(rt.isyn.cpu().numpy(), st.isyn), "synaptic currents differ"
    assert bits_equal(rt.ring.cpu().numpy(), st.ring), "ring contents differ"


__all__ = ["ProbeNode", "Given", "network_from_topology", "oracle_state", "oracle_run", "bits_equal",
           "assert_state_equal", "LIFExp"]
