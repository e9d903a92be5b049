"""Parity of the HIP step (through the C-ABI) with the CPU oracle: spike rasters, potentials, currents,
refractory counters and ring contents bit for bit, on the same seeded inputs."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd.networks import brunel, pd14
from tests.gpu_util import (ProbeNode, assert_state_equal, bits_equal, network_from_topology, oracle_run,
                            oracle_state)
from tests.helpers import tiny_topology

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_brunel_config0_bit_exact_2000_steps():
    _say("case: brunel N=2500 seed=0 2000 steps")
    net, _ = brunel(device="cuda", seed=0)
    rt = net.build()
    st = oracle_state(rt)
    net.run(2000)
    torch.cuda.synchronize()
    raster, events = oracle_run(rt, st, 0, 2000)
    assert raster.sum() > 1000 and events > 0
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


def test_pd14_small_bit_exact_1000_steps():
    _say("case: pd14 scale=0.05 seed=1 1000 steps")
    net, _ = pd14(scale=0.05, seed=1, device="cuda")
    rt = net.build()
    st = oracle_state(rt)
    net.run(1000)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, 1000)
    assert raster.sum() > 100
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


@pytest.mark.parametrize("n,n_links,max_delay,p_spike", [
    (200, 6000, 7, 0.1),      # a few spikes a step: the sorted-list path
    (500, 20000, 9, 0.125),   # around 64 spikes a step: both paths
    (3000, 120000, 40, 0.6),  # bursts: > 512 spiking sources and > 1024 links a group per step
    (130, 4000, 128, 0.3),    # delays up to the ring's 128 slots, two 64-bit touched masks
])
def test_forced_spikes_deliver_like_the_oracle(n, n_links, max_delay, p_spike):
    """Host-forced spike words (the commanded-node fixture of test/runtimes.py:24-44): the ring input
    read at every step, and the ring, equal the oracle's."""
    _say(f"case: forced spikes n={n} links={n_links} D={max_delay} p={p_spike}")
    topo = tiny_topology(n, n_links, max_delay, seed=n)
    net, pop = network_from_topology(ProbeNode(), n, topo)
    rt = net.build()
    st = oracle_state(rt)
    rng = np.random.default_rng(n + 1)
    forced = rng.random((25, n)) < p_spike
    for s in range(forced.shape[0]):
        if s > 0:
            rt.force_spikes(s - 1, forced[s - 1])
            st.prev_spk[:] = forced[s - 1]
        rt.run(1)
        oracle_run(rt, st, s, 1, record=False)
        torch.cuda.synchronize()
        assert bits_equal(rt.v.cpu().numpy(), st.v), f"step {s}: ring input differs"
    assert_state_equal(rt, st)


@pytest.mark.parametrize("dense_min,fmt,uniform", [(0, "wide", False), (10**9, "wide", False), (0, "compact", False),
                                                   (0, "wide", True), (0, "compact", True), (0, "slot16", True),
                                                   (10**9, "slot16", True)])
def test_row_intake_and_segment_intake_agree_with_the_oracle(dense_min, fmt, uniform):
    """The same forced spike patterns through the row intake (every step; wide, compact and slot16 row
    entries, mixed delays and one delay for all links) and the segment intake (never)."""
    n, n_links, max_delay = 700, 30000, 12
    _say(f"case: row intake dense_min={dense_min} fmt={fmt} uniform={uniform} n={n}")
    src_ptr, tgt, w, d = tiny_topology(n, n_links, max_delay, seed=17)
    if fmt == "compact":  # 200 distinct weights of varied magnitude and sign, -0.0 among them
        w = np.concatenate([w[:199], np.array([-0.0], np.float32)])[np.random.default_rng(5).integers(0, 200, n_links)]
    if fmt == "slot16":  # the weight set by the source's side of slot 384 (one population: slot = neuron)
        w = np.where(np.repeat(np.arange(n), np.diff(src_ptr)) < 384, np.float32(0.2), np.float32(-1.0))
    if uniform:
        d = np.full_like(d, 7)
    net, pop = network_from_topology(ProbeNode(), n, (src_ptr, tgt, w, d))
    if fmt != "slot16":
        net.row_format = fmt
    rt = net.build()
    assert rt.rows is not None and rt.rows.fmt == {"wide": 0, "compact": 1, "slot16": 2}[fmt]
    assert rt.rows.uniform_delay == (7 if uniform else 0)
    rt.set_dense_min(dense_min)
    st = oracle_state(rt)
    rng = np.random.default_rng(18)
    forced = rng.random((20, n)) < np.linspace(0.02, 0.8, 20)[:, None]
    for s in range(forced.shape[0]):
        if s > 0:
            rt.force_spikes(s - 1, forced[s - 1])
            st.prev_spk[:] = forced[s - 1]
        rt.run(1)
        oracle_run(rt, st, s, 1, record=False)
        torch.cuda.synchronize()
        assert bits_equal(rt.v.cpu().numpy(), st.v), f"step {s}: ring input differs"
    assert_state_equal(rt, st)


def test_brunel_bursts_row_intake_bit_exact():
    """A synchronous Brunel regime (strong excitation) whose bursts take the row intake."""
    _say("case: brunel g=3 bursts, 600 steps")
    net, _ = brunel(N_E=4000, N_I=1000, C_E=400, C_I=100, g=3.0, eta=3.0, device="cuda", seed=4)
    net.dense_min_spikes = 200
    rt = net.build()
    assert rt.rows.fmt == 2  # Brunel: one delay, weights set by the source population (slot16 rows)
    st = oracle_state(rt)
    net.run(600)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, 600)
    assert raster.sum(axis=1).max() >= 200  # bursts did take the row intake
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


def test_graph_replays_equal_eager_steps():
    _say("case: graph windows vs eager, brunel N=2500")
    a, _ = brunel(device="cuda", seed=5)
    b, _ = brunel(device="cuda", seed=5)
    ra, rb = a.build(), b.build()
    a.run(3, graph_window=0)  # off the captured phase: run() steps eagerly up to it
    b.run(3, graph_window=0)
    for _ in range(6):  # replays with eager work between them
        a.run(52, graph_window=48)
        b.run(52, graph_window=0)
        torch.cuda.synchronize()
        _ = (ra.v * 2).sum().item()
        assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert np.array_equal(ra.raster(), rb.raster())
    assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_continuation_equals_one_run():
    _say("case: continuation 300+200 vs 500, pd14 scale=0.03")
    a, _ = pd14(scale=0.03, seed=2, device="cuda")
    b, _ = pd14(scale=0.03, seed=2, device="cuda")
    a.run(300)
    a.run(200)
    b.run(500)
    torch.cuda.synchronize()
    ra, rb = a.runtime, b.runtime
    assert np.array_equal(ra.raster(), rb.raster())
    assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert bits_equal(ra.isyn.cpu().numpy(), rb.isyn.cpu().numpy())


def test_no_recording_keeps_last_step():
    _say("case: record=False")
    a, _ = brunel(device="cuda", seed=7)
    a.record = False
    b, _ = brunel(device="cuda", seed=7)
    a.run(300)
    b.run(300)
    torch.cuda.synchronize()
    assert a.runtime.bit_slots == 2
    assert np.array_equal(a.runtime.raster(299, 300), b.runtime.raster(299, 300))
    assert bits_equal(a.runtime.v.cpu().numpy(), b.runtime.v.cpu().numpy())


def test_dsl_state_views_write_through():
    _say("case: population state views")
    net, pops = brunel(N_E=300, N_I=100, C_E=30, C_I=10, device="cuda", seed=9)
    rt = net.build()
    E = pops["E"]
    assert E.V.shape == (300,)
    E.V = torch.full((300,), 5.0)
    assert float(rt.v[:300].min()) == 5.0 and float(rt.v[:300].max()) == 5.0
    net.step(1.0)
    assert net.step_index == 10 and abs(net.t - 1.0) < 1e-9


@pytest.mark.parametrize("xcd,run", [(0, 0), (3, 0), (8, 0), (8, 4), (5, 7)])
def test_launch_mappings_equal_the_oracle(xcd, run):
    """Every block-to-group mapping of the single-plan step (n2h_set_tuning) gives the oracle's results,
    eager and through captured graphs (groups not a multiple of the XCD count or of its runs: padding
    blocks)."""
    _say(f"case: pd14 scale=0.05 seed=4, launch mapping xcd={xcd} run={run}, 400 steps")
    net, _ = pd14(scale=0.05, seed=4, device="cuda")
    rt = net.build()
    rt.set_launch_mapping(xcd, run)
    assert rt.n_groups % 8 != 0
    st = oracle_state(rt)
    net.run(200, graph_window=0)
    net.run(200, graph_window=100)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, 400)
    assert raster.sum() > 50
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)
