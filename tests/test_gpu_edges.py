"""Edge kinds and ports on the GPU (n2h_bind_edges; the reference's D-6 per-connection receive variants, D-7
per-link attribute blocks and edge forward, TODO.md:179-204, test/test_fused_generated_gpu.py:361-398)
against the oracle's edge step: rasters, potentials, currents, refractory counters, both rings and the
plastic weights bit for bit, eager and in captured windows, through both intakes (sorted spike list and
bitmap scan); the delivery log of an edge plan; snapshots carrying the plastic weights; what edge plans
refuse."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd import FusedRound
from naer2_amd.adjoint import RateLossGrad
from tests.gpu_util import (assert_state_equal, bits_equal, canonical_weights, mixed_edges, oracle_run_edges,
                            oracle_state)

pytestmark = pytest.mark.gpu

CASES = {
    "mixed": (("static", "gain", "plastic", "static"), ("current", "potential", "current", "potential")),
    "plastic": (("plastic",) * 4, ("potential", "current", "potential", "current")),
    "gain-potential": (("gain", "static", "gain", "plastic"), ("potential",) * 4),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _build(case, seed, device="cuda", sizes=(1600, 400), in_degree=(80, 20)):
    kinds, ports = CASES[case]
    net, conns = mixed_edges(seed, kinds, ports, sizes=sizes, in_degree=in_degree, device=device)
    net.keep_canonical = True
    net.record = True
    return net, conns


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("graph_window", [0, 100])
def test_edge_plans_match_the_oracle(case, graph_window):
    T = 300
    _say(f"case: edges {case}, 2000 neurons, {T} steps, graph window {graph_window}")
    net, _ = _build(case, seed=11)
    rt = net.build()
    assert rt.edges and rt.rows is None and rt.uniform_delay == 0
    st = oracle_state(rt)
    ring_v = rt.ring_v.cpu().numpy().copy()
    w = canonical_weights(rt).copy()
    w0 = w.copy()
    net.run(T, graph_window=graph_window)
    torch.cuda.synchronize()
    raster_o, _ = oracle_run_edges(rt, st, ring_v, w, 0, T)
    raster = rt.raster(0, T)
    per_step = raster.sum(axis=1)
    _say(f"spikes {int(per_step.sum())}, largest step {int(per_step.max())}")
    assert per_step.max() > 64 and ((per_step > 0) & (per_step <= 64)).any()  # both intakes
    assert np.array_equal(raster, raster_o)
    assert_state_equal(rt, st)
    assert bits_equal(rt.ring_v.cpu().numpy(), ring_v)
    assert bits_equal(canonical_weights(rt), w)
    if "plastic" in CASES[case][0]:
        assert not np.array_equal(w, w0)


def test_edge_plan_delivery_log_and_status_rows():
    _say("case: delivery log and status rows of an edge plan, 60 steps after 100")
    net, _ = _build("mixed", seed=12, sizes=(800, 200), in_degree=(40, 10))
    twin, _ = _build("mixed", seed=12, sizes=(800, 200), in_degree=(40, 10))
    rt, rt2 = net.build(), twin.build()
    net.run(100)
    twin.run(100)
    rt.set_delivery_log(2_000_000)
    rt.set_status_rows(True)
    net.run(60, graph_window=0)
    twin.run(60, graph_window=0)
    torch.cuda.synchronize()
    log = rt.deliveries()
    raster = rt.raster(99, 159)
    c = rt.canon.to("cpu")
    ptr, tgt, d = c.src_ptr.numpy(), c.tgt.numpy(), c.d.numpy()
    conn = c.extra[:, 0].numpy()
    exp = []
    for i, s in enumerate(range(99, 159)):
        for j in np.nonzero(raster[i])[0]:
            for e in range(ptr[j], ptr[j + 1]):
                exp.append((s + int(d[e]), s, int(conn[e]), int(j), int(tgt[e])))
    exp.sort()
    assert len(exp) > 1000 and sorted(log) == exp
    rows = rt.status_rows()
    assert (rows[:, 0] == 159).all() and (rows[:, 3] == raster[-1].sum()).all()
    for a, b in ((rt.v, rt2.v), (rt.ring, rt2.ring), (rt.ring_v, rt2.ring_v), (rt.links, rt2.links)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_snapshot_restores_plastic_weights():
    _say("case: snapshot and restore of a plastic edge plan, 100 + 60 steps twice")
    net, _ = _build("plastic", seed=13, sizes=(800, 200), in_degree=(40, 10))
    rt = net.build()
    net.run(100)
    snap = rt.snapshot()
    net.run(60, graph_window=0)
    first = (rt.raster(100, 160), rt.link_weights().clone(), rt.v.clone(), rt.ring_v.clone())
    rt.restore(snap)
    net.run(60, graph_window=20)
    assert np.array_equal(rt.raster(100, 160), first[0])
    for a, b in zip((rt.link_weights(), rt.v, rt.ring_v), first[1:]):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_edge_plans_step_alone():
    _say("case: what edge plans refuse")
    net, _ = _build("mixed", seed=14, sizes=(200, 100), in_degree=(20, 5))
    rt = net.build()
    with pytest.raises(RuntimeError, match="edge"):
        FusedRound([rt])
    with pytest.raises(ValueError, match="edge"):
        RateLossGrad([rt], np.ones((2, 2), np.float32))
    with pytest.raises(RuntimeError, match="compensated"):
        rt.trace_potentials(torch.empty(4, rt.n, device="cuda"), 0, torch.zeros(2, rt.n, device="cuda"))
