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

# kinds and ports of E->E, E->I, I->E, I->I, and the largest delay (steps) of potential-port links (None: 8, as
# every other link)
CASES = {
    "mixed": (("static", "gain", "plastic", "static"), ("current", "potential", "current", "potential"), None),
    "plastic": (("plastic",) * 4, ("potential", "current", "potential", "current"), None),
    "gain-potential": (("gain", "static", "gain", "plastic"), ("potential",) * 4, None),
    "kinds-only": (("plastic", "gain", "static", "plastic"), ("current",) * 4, None),
    "short-potential": (("static", "plastic", "gain", "static"), ("current", "potential", "current", "potential"), 3),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _build(case, seed, device="cuda", sizes=(1600, 400), in_degree=(80, 20)):
    kinds, ports, pot_max = CASES[case]
    net, conns = mixed_edges(seed, kinds, ports, sizes=sizes, in_degree=in_degree, device=device)
    if pot_max is not None:
        rng = np.random.default_rng(seed + 100)
        for c in conns:
            if c.port == "potential":
                c.delay = torch.as_tensor(rng.integers(1, pot_max + 1, c.shape[0]) * 0.1, dtype=torch.float32)
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
    kinds, ports, pot_max = CASES[case]
    want_v = 0 if "potential" not in ports else (pot_max or 8)
    assert rt.v_slots == want_v and (rt.ring_v is None) == (want_v == 0)
    if pot_max is not None:
        assert rt.v_slots < rt.d_slots
    st = oracle_state(rt)
    ring_v = (rt.ring_v.cpu().numpy().copy() if rt.ring_v is not None
              else np.zeros((rt.d_slots, rt.n), np.float32))
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
    if rt.ring_v is not None:
        assert bits_equal(rt.ring_v.cpu().numpy(), ring_v)
    else:
        assert not ring_v.any()
    assert bits_equal(canonical_weights(rt), w)
    if "plastic" in kinds:
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
    conn = c.extra[0].numpy()
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


def test_full_scale_pd14_edges_deliver_the_plain_weights():
    """Full-scale PD14 (298.8 M links) with every connection a gain edge of gain 1.0, then a plastic edge of
    factor -0.0: both deliver exactly the plain weights, so 300 steps give the plain run's raster bit for
    bit, and the weights read back in connection order are the connections' own (the payload's input places
    at this size)."""
    from naer2_amd.networks import pd14
    T = 300
    _say(f"case: full-scale pd14, plain vs gain 1.0 vs plastic -0.0 edges, {T} steps")
    net, _ = pd14(scale=1.0, seed=0, device="cuda")
    net.record = True
    rt = net.build()
    net.run(T, graph_window=100)
    plain = rt.raster_bits(0, T).clone()
    del rt, net
    torch.cuda.empty_cache()
    for kind, attr in (("gain", 1.0), ("plastic", -0.0)):
        net, _ = pd14(scale=1.0, seed=0, device="cuda")
        net.record = True
        net.keep_canonical = False
        for c in net.connections:
            c.kind = kind
            setattr(c, "gain" if kind == "gain" else "factor", attr)
        rt = net.build()
        assert rt.edges and rt.v_slots == 0 and rt.n_links == 298822986
        w0 = torch.cat([c.weight for c in net.connections])
        assert torch.equal(rt.link_weights().view(torch.int32), w0.view(torch.int32))
        net.run(T, graph_window=100)
        torch.cuda.synchronize()
        assert torch.equal(rt.raster_bits(0, T), plain), kind
        assert torch.equal(rt.link_weights().view(torch.int32), w0.view(torch.int32)), kind
        _say(f"{kind}: raster and weights as the plain run's")
        del rt, net, w0
        torch.cuda.empty_cache()
