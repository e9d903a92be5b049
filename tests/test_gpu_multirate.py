"""Multi-rate networks on the GPU: populations with a timestep of their own (a whole multiple of the
resolution) beside populations updating every step, bit-exact against the oracle -- rasters, potentials,
currents, refractory counters and rings -- through eager steps, graph windows, both intake paths and a
batch (test/test_fusion.py:61-85 is the reference's case: I at 0.2 ms beside E at 0.1 ms)."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd import FusedRound
from naer2_amd.networks import brunel, pd14
from tests.gpu_util import (Given, ProbeNode, assert_state_equal, bits_equal, oracle_run, oracle_state)
from tests.helpers import tiny_topology

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _brunel_mixed(seed, ts_e=0.2, ts_i=0.3, **kw):
    net, pops = brunel(device="cuda", seed=seed, **kw)
    pops["E"].timestep, pops["I"].timestep = ts_e, ts_i
    return net


@pytest.mark.parametrize("ts_e,ts_i", [(None, 0.3), (0.2, 0.3), (0.4, None)])
def test_brunel_mixed_timesteps_bit_exact(ts_e, ts_i):
    _say(f"case: brunel N=2500 E timestep {ts_e} I timestep {ts_i}, 900 steps")
    net = _brunel_mixed(2, ts_e, ts_i)
    rt = net.build()
    assert rt.d_slots == 15 + max(int(round((t or 0.1) / 0.1)) for t in (ts_e, ts_i)) - 1
    st = oracle_state(rt)
    net.run(900)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, 900)
    assert raster.sum() > 500
    for name, ts, lo, hi in (("E", ts_e, 0, 2000), ("I", ts_i, 2000, 2500)):
        k = int(round((ts or 0.1) / 0.1))
        off = np.arange(900) % k != k - 1
        assert not raster[off, lo:hi].any(), name  # spikes at update steps only
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


def test_pd14_mixed_timesteps_bit_exact():
    _say("case: pd14 scale=0.05 seed=3, four populations at 0.2-0.5 ms, 800 steps")
    net, pops = pd14(scale=0.05, seed=3, device="cuda")
    for name, ts in (("E_23", 0.2), ("I_4", 0.5), ("E_5", 0.3), ("I_6", 0.4)):
        pops[name].timestep = ts
    rt = net.build()
    st = oracle_state(rt)
    net.run(800)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, 800)
    assert raster.sum() > 100
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


@pytest.mark.parametrize("dense_min", [0, 10**9])
def test_forced_spikes_window_sums_both_intakes(dense_min):
    """Probe populations (V = the window's input) at strides 1, 3 and 4 under host-forced spikes, with
    delay-1 links among the rest: every step's V and the ring equal the oracle's, through the row intake
    (dense_min 0) and the segment intake."""
    sizes, strides, D = [150, 130, 170], [1, 3, 4], 9
    n = sum(sizes)
    _say(f"case: forced spikes, strides {strides}, dense_min={dense_min}")
    src_ptr, tgt, w, d = tiny_topology(n, 40000, D, seed=31)
    src = np.repeat(np.arange(n), np.diff(src_ptr))
    first = np.concatenate([[0], np.cumsum(sizes)])
    pops = []
    for sz, k in zip(sizes, strides):
        p = ProbeNode() * sz
        p.timestep = 0.1 * k if k > 1 else None
        pops.append(p)
    net = pops[0].network
    for p in pops[1:]:
        net = net + p.network
    net.resolution, net.seed, net.device = 0.1, 0, torch.device("cuda")
    net._generator = None
    pop_of = np.searchsorted(first, np.arange(n), side="right") - 1
    for a in range(3):
        for b in range(3):
            m = (pop_of[src] == a) & (pop_of[tgt] == b)
            c = pops[a] >> Given(src[m] - first[a], tgt[m] - first[b]) >> pops[b]
            c.weight = torch.as_tensor(w[m])
            c.delay = torch.as_tensor(d[m].astype(np.float64) * 0.1, dtype=torch.float32)
    rt = net.build()
    assert rt.d_slots == D + 4 - 1
    rt.set_dense_min(dense_min)
    st = oracle_state(rt)
    rng = np.random.default_rng(32)
    forced = rng.random((40, n)) < np.linspace(0.02, 0.6, 40)[:, None]
    for s in range(forced.shape[0]):
        if s > 0:
            rt.force_spikes(s - 1, forced[s - 1])
            st.prev_spk[:] = forced[s - 1]
        rt.run(1)
        oracle_run(rt, st, s, 1, record=False)
        torch.cuda.synchronize()
        assert bits_equal(rt.v.cpu().numpy(), st.v), f"step {s}: window input differs"
        assert bits_equal(rt.ring.cpu().numpy(), st.ring), f"step {s}: ring differs"
    assert_state_equal(rt, st)


def test_graph_windows_and_continuation_equal_eager_steps():
    _say("case: multi-rate brunel, graph windows vs eager, off-phase starts")
    a, b = _brunel_mixed(6), _brunel_mixed(6)
    ra, rb = a.build(), b.build()
    a.run(5, graph_window=0)
    b.run(5, graph_window=0)
    for _ in range(5):
        a.run(61, graph_window=48)
        b.run(61, graph_window=0)
        torch.cuda.synchronize()
        _ = (ra.v * 2).sum().item()
        assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert np.array_equal(ra.raster(), rb.raster())
    assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_batch_with_multirate_plans_equals_plain_runs():
    _say("case: batch of a multi-rate and a single-rate brunel network")
    batched = [_brunel_mixed(8), brunel(device="cuda", seed=9)[0]]
    plain = [_brunel_mixed(8), brunel(device="cuda", seed=9)[0]]
    fr = FusedRound([n.runtime for n in batched], {"order": 0, "xcd": 8}, reserve_steps=200)
    fr.launch(200, graph_window=48)
    fr.close()
    for n in plain:
        n.run(200, graph_window=0)
    torch.cuda.synchronize()
    for x, y in zip(batched, plain):
        rx, ry = x.runtime, y.runtime
        assert np.array_equal(rx.raster(), ry.raster())
        assert bits_equal(rx.v.cpu().numpy(), ry.v.cpu().numpy())
        assert bits_equal(rx.ring.cpu().numpy(), ry.ring.cpu().numpy())
    assert batched[0].runtime.raster().sum() > 0


def test_potential_traces_refuse_multirate():
    net = _brunel_mixed(1)
    rt = net.build()
    with pytest.raises(RuntimeError, match="stride"):
        rt.trace_potentials(torch.zeros(10, rt.n, device="cuda"))


def test_run_planned_and_run_graph_equal_run():
    _say("case: multi-rate brunel, run_planned + run_graph vs run")
    a, b = _brunel_mixed(4), _brunel_mixed(4)
    a.build(), b.build()
    assert a.run_planned(37) == 37
    assert a.run_graph(until=20.0) == 200 - 37
    b.run(200, graph_window=0)
    torch.cuda.synchronize()
    ra, rb = a.runtime, b.runtime
    assert ra.step == rb.step == 200 and abs(a.t - 20.0) < 1e-9
    assert np.array_equal(ra.raster(), rb.raster())
    assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
