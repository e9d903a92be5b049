"""Host logic of the event-driven path on the CPU (naer2_amd.event.describe): what it runs, what it refuses, the
window and lags, shared input streams and region capacities; and the C-ABI's event entry points exported."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import naer2_amd as n2
from naer2_amd import _lib
from naer2_amd.event import describe, region_cap
from naer2_amd.networks import brunel_exact


def _net(delays=(1.5, 1.5, 1.5, 1.5, 1.5, 1.5)):
    net, pops = brunel_exact(400, K=20, device="cpu")
    for c, d in zip(net.connections, delays):
        c.delay = d
    return net, pops


def test_brunel_plan():
    net, pops = _net()
    plan = describe(net)
    assert plan.window == np.float32(1.5)
    assert plan.conn_lag == [1] * 6
    # E and I receive from (E, I, P) with one delay: one shared stream
    assert len(plan.streams) == 1
    assert plan.stream_of[pops["E"]] == plan.stream_of[pops["I"]] == 0
    assert pops["P"] not in plan.stream_of
    assert plan.pop_first.tolist() == [0, 320, 400, 720]
    # a neuron fires at most once per refractory period (2 ms > 1.5 ms): one spike per neuron and window
    assert plan.caps[0] == 64 and plan.caps[1] == 64
    assert plan.caps[2] % 64 == 0 and plan.caps[2] >= 64


def test_several_delays_plan():
    net, pops = _net((0.75, 1.5, 3.0, 0.75, 1.5, 0.75))
    plan = describe(net)
    assert plan.window == np.float32(0.75)
    assert plan.conn_lag == [1, 2, 4, 1, 2, 1]
    assert len(plan.streams) == 2  # E's and I's sources arrive with different delays
    assert plan.streams[0] == ((0, 0.75, 1), (1, 1.5, 2), (2, 3.0, 4))


@pytest.mark.parametrize("delays", [(1.5, 2.0, 1.5, 1.5, 1.5, 1.5), (0.1, 7.0, 1.5, 1.5, 1.5, 1.5)])
def test_delays_not_multiples_refused(delays):
    net, _ = _net(delays)
    with pytest.raises(NotImplementedError, match="whole number"):
        describe(net)


def test_refusals():
    net, pops = _net()
    c = pops["E"] >> n2.FixedInDegree(2) >> pops["I"]  # a second E -> I connection
    c.weight, c.delay = 0.1, 1.5
    with pytest.raises(NotImplementedError, match="one connection per"):
        describe(net)
    net2, pops2 = _net()
    c = pops2["E"] >> n2.FixedInDegree(2) >> pops2["P"]  # a Poisson source cannot receive
    c.weight, c.delay = 0.1, 1.5
    with pytest.raises(ValueError, match="only event_lif"):
        describe(net2)
    E = n2.LIFDeltaExact() * 10
    T = n2.LIFDelta() * 10
    net3 = E.network
    net3.merge(T.network)
    with pytest.raises(NotImplementedError, match="time-driven"):
        describe(net3)


def test_delay_per_link_must_be_uniform():
    net, _ = _net()
    c = net.connections[0]
    d = torch.full(c.shape, 1.5)
    d[0] = 3.0
    c.delay = d
    with pytest.raises(NotImplementedError, match="same delay"):
        describe(net)


def test_region_caps():
    P = n2.PoissonExact(2.0) * 64
    cap = region_cap(P, 1.5)
    mean = 64 * 2.0 * 1.5
    assert cap % 64 == 0 and cap >= mean + 10 * mean ** 0.5
    E = n2.LIFDeltaExact(t_ref=0.5) * 64
    assert region_cap(E, 1.5) == 64 * 4  # up to 1.5 // 0.5 + 1 spikes per neuron and window
    E.history_cap = 2
    assert region_cap(E, 1.5) == 128


def test_event_symbols_exported():
    lib = _lib.load()
    for name in _lib.EXPORTS:
        if name.startswith("n2h_ev_"):
            assert hasattr(lib, name), name
    hdr = open(_lib.LIB_PATH.parent.parent / "include" / "naer2_hip.h").read()
    for name in [x for x in _lib.EXPORTS if x.startswith("n2h_ev_")]:
        assert f"int {name}(" in hdr, name
