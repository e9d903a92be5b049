"""Readback formats of the time-driven step, through the C-ABI: the delivery log (n2h_bind_log; the reference's
PopulationRuntime.deliveries, naer2/runtime.py:747-752, checked as test/test_fused_gpu.py:105-142 checks it:
the delivered set against the links the spikes must deliver, each target's links in order) and the
per-group status rows (n2h_set_status)."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd.networks import brunel
from tests.gpu_util import Given, ProbeNode, assert_state_equal, oracle_run, oracle_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _three_connections(seed=3, device="cuda"):
    """Probe populations A (150) and B (100) with connections A->A, A->B, B->A of random links (repeats
    kept), delays 1..12 steps; the links of each, as given."""
    rng = np.random.default_rng(seed)
    A, B = ProbeNode() * 150, ProbeNode() * 100
    net = A.network
    net.merge(B.network)
    net.resolution, net.seed, net.device = 0.1, seed, torch.device(device)
    net._generator = None
    given = []
    for ci, (src, tgt, ns, nt, m) in enumerate(((A, A, 150, 150, 2500), (A, B, 150, 100, 1800), (B, A, 100, 150, 1500))):
        s = np.sort(rng.integers(0, ns, m))  # source-major, as a connection keeps its links
        t = rng.integers(0, nt, m)
        d = rng.integers(1, 13, m)
        w = rng.standard_normal(m).astype(np.float32)
        conn = src >> Given(s, t) >> tgt
        conn.weight = torch.as_tensor(w)
        conn.delay = torch.as_tensor(d.astype(np.float64) * 0.1, dtype=torch.float32)
        off_s = 0 if src is A else 150
        off_t = 0 if tgt is A else 150
        given.append((ci, s + off_s, t + off_t, d))
    return net, given


def _expected(given, spikes_by_step):
    """Every (arrival step, emission step, connection, source, target) the spikes must deliver, and per
    (emission step, target) the canonical order: by source, then connection (creation order), then the
    link's place in its connection."""
    out = []
    for s, ids in spikes_by_step.items():
        fire = set(int(i) for i in ids)
        for ci, src, tgt, d in given:
            for k in range(len(src)):
                if int(src[k]) in fire:
                    out.append(((int(s) + int(d[k]), int(s), ci, int(src[k]), int(tgt[k])), (int(src[k]), ci, k)))
    return out


def test_delivery_log_of_forced_spikes_through_three_connections():
    _say("case: delivery log, forced spikes, connections A->A, A->B, B->A, delays 1..12")
    net, given = _three_connections()
    rt = net.build()
    rt.set_delivery_log(200000)
    rng = np.random.default_rng(4)
    forced = rng.random((15, rt.n)) < np.linspace(0.02, 0.5, 15)[:, None]  # up to ~125 spikes: both intakes
    spikes = {}
    for s in range(16):
        if s > 0:
            rt.force_spikes(s - 1, forced[s - 1])
            spikes[s - 1] = np.nonzero(forced[s - 1])[0]
        rt.run(1)
    log = rt.deliveries()
    exp = _expected(given, spikes)
    assert len(log) == len(exp) > 10000
    assert sorted(log) == sorted(r for r, _ in exp)
    # per (emission step, target): the log's order is the contract order
    order = {r: key for r, key in exp}
    by = {}
    for r in log:
        by.setdefault((r[1], r[4]), []).append(order[r])
    for v in by.values():
        assert v == sorted(v)


def test_delivery_log_of_a_brunel_run_and_unchanged_results():
    _say("case: delivery log of 60 Brunel steps after 200, vs the raster; results as without the log")
    net, _ = brunel(N_E=1600, N_I=400, C_E=160, C_I=40, seed=8, device="cuda")
    twin, _ = brunel(N_E=1600, N_I=400, C_E=160, C_I=40, seed=8, device="cuda")
    rt, rt2 = net.build(), twin.build()
    st = oracle_state(rt)
    net.run(200)
    twin.run(200)
    rt.set_delivery_log(3_000_000)
    net.run(60, graph_window=0)
    twin.run(60, graph_window=0)
    torch.cuda.synchronize()
    log = rt.deliveries()
    raster = rt.raster(199, 259)  # the spikes of steps 199..258 are delivered by steps 200..259
    c = rt.canon.to("cpu")
    ptr, tgt, d = c.src_ptr.numpy(), c.tgt.numpy(), c.d.numpy()
    # connection of a link: brunel() creates E->E, I->E, E->I, I->I in that order
    conn = lambda j, t: (0 if j < 1600 else 1) + (0 if t < 1600 else 2)
    exp = []
    for i, s in enumerate(range(199, 259)):
        for j in np.nonzero(raster[i])[0]:
            for e in range(ptr[j], ptr[j + 1]):
                exp.append((s + int(d[e]), s, conn(j, int(tgt[e])), int(j), int(tgt[e])))
    exp.sort()
    assert len(exp) > 1000 and sorted(log) == exp
    for a, b in ((rt.v, rt2.v), (rt.ring, rt2.ring)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    assert torch.equal(rt.ref, rt2.ref) and np.array_equal(rt.raster(200, 260), rt2.raster(200, 260))
    raster_o, _ = oracle_run(rt, st, 0, 260)
    assert np.array_equal(rt.raster(0, 260), raster_o)
    assert_state_equal(rt, st)


def test_delivery_log_overflow_keeps_counting():
    _say("case: delivery log past its capacity")
    net, given = _three_connections(5)
    rt = net.build()
    rt.set_delivery_log(100)
    rt.run(1)
    rt.force_spikes(0, np.ones(rt.n, bool))  # every source spikes at step 0: step 1 delivers every link
    rt.run(1)
    assert rt.delivery_count() == sum(len(s) for _, s, _, _ in given)
    with pytest.raises(RuntimeError, match="overflowed"):
        rt.deliveries()
    rec = rt._log[1][:100].cpu().numpy()
    assert (rec[:, 1] == 0).all() and (rec[:, 0] >= 1).all()
    rt.set_delivery_log(None)
    rt.run(1)


@pytest.mark.parametrize("dense_min", [None, 30])
def test_status_rows_account_for_every_step(dense_min):
    _say(f"case: status rows, bursting brunel 5000, dense_min {dense_min}")
    net, _ = brunel(N_E=4000, N_I=1000, C_E=400, C_I=100, g=3.0, eta=3.0, seed=9, device="cuda")
    if dense_min is not None:
        net.dense_min_spikes = dense_min
    rt = net.build()
    rt.set_status_rows(True)
    layout = rt.layout
    gf = layout.group_first.cpu().numpy()
    cnt = (layout.group_info.cpu().numpy() & 0xFFFF)
    paths = set()
    for s in range(200):
        rt.run(1)
        rows = rt.status_rows()
        r = rt.raster(s, s + 1)[0]
        prev = int(rt.raster(s - 1, s)[0].sum()) if s > 0 else 0
        assert (rows[:, 0] == s).all()
        assert all(rows[g, 1] == r[gf[g]:gf[g] + cnt[g]].sum() for g in range(rt.n_groups))
        assert (rows[:, 3] == prev).all()
        want = 0 if prev == 0 else (3 if rt.rows is not None and prev >= rt.dense_min else (1 if prev <= 64 else 2))
        assert (rows[:, 2] == want).all(), (s, prev, rows[0])
        paths.add(want)
    assert 1 in paths and (paths & {2, 3})
