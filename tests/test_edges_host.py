"""Edge kinds and ports on the host: the description API (`edge @ "port" @ pattern`, kinds, per-link second
attributes; naer2/_partial.py:1-44, test/edges/*.py), the per-link payload assemble() builds for them, and
the oracle's edge step (oracle/naer2_oracle.c, or_run_edges) pinned by reductions to the plain step and by
the closed-form response of a neuron driven through its potential port."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from naer2_amd import FixedInDegree, GainSynapse, LIFDelta, LIFExp, PlasticSynapse, StaticSynapse, edge
from naer2_amd.network import (EDGE_KIND_BITS, PAYLOAD_ATTR, PAYLOAD_CONN, PAYLOAD_FLAGS, PAYLOAD_INPUT,
                               PORT_POTENTIAL_BIT, assemble)
from naer2_amd.params import params_bytes
from oracle import oracle
from tests.gpu_util import Given, mixed_edges


def _net_of(pop, seed=0):
    net = pop.network
    net.resolution, net.seed, net.device = 0.1, seed, torch.device("cpu")
    net._generator = None
    return net


def test_ports_kinds_and_attributes_in_any_order():
    A, B = LIFExp() * 20, LIFExp() * 30
    c1 = A >> GainSynapse(0.5, 2.0) @ "potential" @ FixedInDegree(3) >> B
    c2 = B >> PlasticSynapse(0.25, 0.125) @ FixedInDegree(2) @ "current" >> A
    c3 = A >> FixedInDegree(4) >> A
    assert (c1.kind, c1.port, c2.kind, c2.port, c3.kind, c3.port) == ("gain", "potential", "plastic", "current",
                                                                      "static", None)
    assert torch.equal(c1.gain, torch.full(c1.shape, 2.0)) and torch.equal(c1.weight, torch.full(c1.shape, 0.5))
    assert torch.equal(c2.factor, torch.full(c2.shape, 0.125))
    assert not c1.plain and not c2.plain and c3.plain
    with pytest.raises(AttributeError, match="gain"):
        c2.gain = 1.0
    with pytest.raises(AttributeError, match="factor"):
        _ = c1.factor
    with pytest.raises(ValueError, match="port"):
        c3.port = "dendrite"
    with pytest.raises(ValueError, match="already set"):
        StaticSynapse(1.0) @ "current" @ "potential"
    with pytest.raises(ValueError, match="mappings"):
        StaticSynapse(1.0) @ {"weight": "w"}
    with pytest.raises(ValueError, match="edge kind"):
        @edge(kind="stdp")
        class _Bad:
            pass
    c1.gain = torch.arange(c1.shape[0], dtype=torch.float32)
    assert c1.gain[-1] == c1.shape[0] - 1
    c3.port = "potential"
    assert not c3.plain
    # a connection's kind can be changed afterwards; the new kind's attribute is then unset
    c3.kind = "plastic"
    with pytest.raises(ValueError, match="factor not set"):
        _ = c3.factor
    c3.factor = 0.5
    assert torch.equal(c3.factor, torch.full(c3.shape, 0.5)) and c3.second_attr is c3.factor
    with pytest.raises(ValueError, match="edge kind"):
        c3.kind = "stdp"


def test_payload_carries_kind_port_attribute_and_input_place():
    A, B = LIFExp() * 40, LIFExp() * 25
    c1 = A >> GainSynapse(0.5, 2.0) @ "potential" @ FixedInDegree(5) >> B
    c2 = B >> PlasticSynapse(0.25, 0.125) @ FixedInDegree(4) >> A
    c3 = A >> FixedInDegree(3) >> A
    for c in (c1, c2, c3):
        c.delay = 0.2
    c3.weight = torch.randn(c3.shape)
    c1.gain = torch.rand(c1.shape)
    a = assemble(_net_of(A), "cpu")
    assert a.edges
    x = a.canon.extra
    assert x.shape == (4, a.canon.n_links)
    # each canonical link's input place names the link it came from
    src = torch.cat([c.src.long() + (0 if c.source is A else 40) for c in (c1, c2, c3)])
    tgt = torch.cat([c.tgt.long() + (0 if c.target is A else 40) for c in (c1, c2, c3)])
    w = torch.cat([c.weight for c in (c1, c2, c3)])
    inp = x[PAYLOAD_INPUT].long()
    assert torch.equal(torch.sort(inp).values, torch.arange(a.canon.n_links))
    c_src = torch.repeat_interleave(torch.arange(a.n), a.canon.out_degree())
    assert torch.equal(src[inp], c_src) and torch.equal(tgt[inp], a.canon.tgt.long()) and torch.equal(w[inp], a.canon.w)
    flags = {0: EDGE_KIND_BITS["gain"] | PORT_POTENTIAL_BIT, 1: EDGE_KIND_BITS["plastic"], 2: 0}
    attr = torch.cat([c1.gain, c2.factor, torch.zeros(c3.shape)])
    for ci in range(3):
        m = x[PAYLOAD_CONN] == ci
        assert int(m.sum()) == (c1, c2, c3)[ci].shape[0]
        assert (x[PAYLOAD_FLAGS][m] == flags[ci]).all()
    assert torch.equal(x[PAYLOAD_ATTR].contiguous().view(torch.float32), attr[inp])
    # a network of static current edges carries no payload
    P = LIFExp() * 10
    (P >> FixedInDegree(2) >> P).delay = 0.1
    (P.network.connections[0]).weight = 1.0
    assert assemble(_net_of(P), "cpu").canon.extra is None


def test_edges_need_the_exponential_synapse_model():
    A = LIFDelta() * 10
    c = A >> StaticSynapse(0.1) @ "potential" @ FixedInDegree(2) >> A
    c.delay = 0.1
    with pytest.raises(NotImplementedError, match="exponential-synapse"):
        assemble(_net_of(A), "cpu")


# ---------------------------------------------------------------- the oracle's edge step

def _mixed(seed=1, kinds=("static", "gain", "plastic", "static"), ports=("current",) * 4):
    return mixed_edges(seed, kinds, ports, device="cpu")


def _edge_arrays(canon):
    x = canon.extra
    fl = x[PAYLOAD_FLAGS].numpy()
    return (((fl >> 8) & 3).astype(np.uint8), ((fl >> 10) & 1).astype(np.uint8),
            x[PAYLOAD_ATTR].contiguous().view(torch.float32).numpy())


def _state(a):
    return oracle.State(a.n, a.d_slots, a.v.numpy(), a.isyn.numpy(), a.ref.numpy())


def _run_edges(a, n_steps, w=None, kind=None, port=None, attr=None):
    c = a.canon
    k, p, at = _edge_arrays(c) if c.extra is not None else (np.zeros(c.n_links, np.uint8),) * 2 + (
        np.zeros(c.n_links, np.float32),)
    k, p, at = (k if kind is None else kind), (p if port is None else port), (at if attr is None else attr)
    w = c.w.numpy().copy() if w is None else w
    st = _state(a)
    ring_v = np.zeros((a.d_slots, a.n), np.float32)
    r, ev = oracle.run_edges(a.pop_first, params_bytes(a.params), a.d_slots, a.net_seed, 0, c.src_ptr.numpy(),
                             c.tgt.numpy(), w, c.d.numpy(), k, p, at, st, ring_v, 0, n_steps)
    return r, st, w, ring_v


def _run_plain(a, n_steps, w=None):
    c = a.canon
    st = _state(a)
    r, _ = oracle.run(oracle.LIF_EXP, a.pop_first, params_bytes(a.params), a.d_slots, a.net_seed, 0,
                      c.src_ptr.numpy(), c.tgt.numpy(), c.w.numpy() if w is None else w, c.d.numpy(), st, 0, n_steps)
    return r, st


def _assembled(net):
    a = assemble(net, "cpu")
    a.net_seed = net.seed
    return a


def _same(x, y):
    return np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))


def _assert_states(s1, s2):
    assert _same(s1.v, s2.v) and _same(s1.isyn, s2.isyn) and _same(s1.ring, s2.ring)
    assert np.array_equal(s1.ref, s2.ref)


def test_oracle_static_current_edges_are_the_plain_step():
    net, _ = _mixed(kinds=("static",) * 4)
    a = _assembled(net)
    a.canon.extra = None
    r1, s1, _, ring_v = _run_edges(a, 300)
    r2, s2 = _run_plain(a, 300)
    assert r1.sum() > 100 and np.array_equal(r1, r2)
    _assert_states(s1, s2)
    assert not ring_v.any()


def test_oracle_gain_edges_deliver_the_product():
    """A gain link delivers w * g, rounded once: the plain step with those weights, bit for bit."""
    net, _ = _mixed(seed=2, kinds=("gain",) * 4)
    a = _assembled(net)
    k, p, g = _edge_arrays(a.canon)
    assert (k == 1).all() and not p.any() and (g != 1.0).any()
    r1, s1, w_after, _ = _run_edges(a, 300)
    wg = (a.canon.w.numpy() * g).astype(np.float32)
    r2, s2 = _run_plain(a, 300, wg)
    assert r1.sum() > 100 and np.array_equal(r1, r2)
    _assert_states(s1, s2)
    assert _same(w_after, a.canon.w.numpy())  # gains leave the weights alone
    # a plausible slip (w + g) would change the run
    r3, _ = _run_plain(a, 300, (a.canon.w.numpy() + g).astype(np.float32))
    assert not np.array_equal(r1, r3)


def test_oracle_plastic_weights_grow_by_factor_per_delivery():
    """With dyadic weights and factors every update is exact: after a run each plastic link's weight is
    w0 + f * (the spikes of its source that were delivered); factor -0.0 is the plain step."""
    net, conns = _mixed(seed=3, kinds=("plastic",) * 4)
    for c in conns:
        c.weight = torch.round(c.weight * 64) / 64
        c.factor = torch.round(c.factor * 2 ** 14) / 2 ** 14
    a = _assembled(net)
    T = 250
    r, _, w_after, _ = _run_edges(a, T)
    c = a.canon
    # step s delivers the spikes of step s - 1: steps 0..T-1 deliver those of steps 0..T-2 (none before step 0)
    counts = r[:T - 1].sum(axis=0).astype(np.float64)
    per_link = np.repeat(counts, np.diff(c.src_ptr.numpy()))
    _, _, f = _edge_arrays(c)
    want = c.w.numpy().astype(np.float64) + f.astype(np.float64) * per_link
    assert (per_link > 0).sum() > 1000 and np.array_equal(w_after.astype(np.float64), want)
    # factor -0.0, the exact identity of fp32 addition (+0.0 would turn a weight of -0.0 into +0.0)
    r0, s0, w0, _ = _run_edges(a, T, attr=np.full(c.n_links, -0.0, np.float32))
    r1, s1 = _run_plain(a, T)
    assert np.array_equal(r0, r1) and _same(w0, c.w.numpy())
    _assert_states(s0, s1)


def test_oracle_potential_port_closed_form():
    """A neuron with no drive and no current input, reached only through its potential port, relaxes as
    V(t) = E_L + sum_k x_k exp(-(t - t_k) dt / tau_m) (t_k: the arrival steps; the jump is applied in the
    arrival step's update, after the propagation).  One plastic link (w0 = 0, factor 0.3: its k-th delivery
    brings 0.3 k, the weight after the update) and one static link (0.2), from a driven source, delays 3 and
    5 steps."""
    S, T = LIFExp(K_ext=2000) * 1, LIFExp() * 1
    c1 = S >> PlasticSynapse(0.0, 0.3) @ "potential" @ Given([0], [0]) >> T
    c2 = S >> StaticSynapse(0.2) @ "potential" @ Given([0], [0]) >> T
    c1.delay, c2.delay = 0.3, 0.5
    net = _net_of(S, seed=5)
    a = _assembled(net)
    n_steps = 1500
    r, st, w_after, _ = _run_edges(a, n_steps)
    spikes = np.nonzero(r[:, 0])[0]
    assert 4 <= len(spikes) and not r[:, 1].any()
    tau, dt, e_l = T.instance.t_mem, 0.1, T.instance.E_leak
    # the target's potential after the last step
    v = e_l
    for k, s in enumerate(spikes, 1):
        for x, d in ((0.3 * k, 3), (0.2, 5)):
            t_arr = s + d
            if t_arr <= n_steps - 1:
                v += x * math.exp(-(n_steps - 1 - t_arr) * dt / tau)
    assert abs(float(st.v[1]) - v) <= 2e-4, (float(st.v[1]), v)
    assert float(st.isyn[1]) == 0.0
    delivered = int((spikes <= n_steps - 2).sum())
    assert w_after[list(a.canon.extra[PAYLOAD_CONN]).index(0)] == np.float32(0.3 * delivered)
    # an arrival one step late (or early) would miss by far more than the tolerance
    late = sum(0.3 * k * math.exp(-(n_steps - 1 - s - 4) * dt / tau) for k, s in enumerate(spikes, 1) if s + 4 <= n_steps - 1)
    on_time = sum(0.3 * k * math.exp(-(n_steps - 1 - s - 3) * dt / tau) for k, s in enumerate(spikes, 1) if s + 3 <= n_steps - 1)
    assert abs(late - on_time) > 1e-2
