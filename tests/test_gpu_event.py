"""The event-driven, exact-time path on the GPU (n2h_ev_*) against the oracle, bit for bit: spikes (neuron and
fp32 time, per window), final state; the fp32 exp/log; captured windows; bursts past the LDS sort; the
reference's own Brunel statistics (test/test_brunel_gpu.py:74-105); lost-data flags."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import event as ev
from tests.event_util import oracle_net, train_net

pytestmark = pytest.mark.gpu
F32 = np.float32


def _say(*a):
    print(*a, flush=True)


def _oracle(evn, v0, n_windows):
    st = ev.EvState.initial(evn, v0)
    nodes, tms, cnt, _ = ev.run(evn, st, n_windows)
    return st, nodes, tms, cnt


def _assert_parity(rt, st, nodes, tms, cnt):
    torch.cuda.synchronize()
    rt.check()
    assert np.array_equal(rt.window_counts(), cnt)
    gt, gn = rt.spikes()
    assert np.array_equal(gn, nodes)
    assert np.array_equal(gt.view(np.uint32), tms.view(np.uint32))
    for name, ref in (("v", st.v), ("last", st.last), ("tre", st.tre), ("nxt", st.nxt)):
        got = getattr(rt, name).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), name


def test_math_bit_exact():
    from naer2_amd import _lib

    lib = _lib.load()
    olib = ev._lib()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-90, 5, 100000), -rng.exponential(0.1, 50000)]).astype(F32)
    u = np.concatenate([rng.uniform(0, 1, 100000), 10.0 ** rng.uniform(-37, 37, 50000)]).astype(F32)
    xd, ud = torch.from_numpy(x).cuda(), torch.from_numpy(u).cuda()
    ex, lg = torch.empty_like(xd), torch.empty_like(ud)
    _say("n2h_ev_math", x.size)
    _lib.ev_check(lib.n2h_ev_math(xd.data_ptr(), ex.data_ptr(), ud.data_ptr(), lg.data_ptr(), x.size, None),
                  "n2h_ev_math")
    torch.cuda.synchronize()
    ge, gl = ex.cpu().numpy(), lg.cpu().numpy()
    oe = np.array([olib.or_ev_exp(float(v)) for v in x], F32)
    ol = np.array([olib.or_ev_log(float(v)) for v in u], F32)
    assert np.array_equal(ge.view(np.uint32), oe.view(np.uint32))
    assert np.array_equal(gl.view(np.uint32), ol.view(np.uint32))


@pytest.mark.parametrize("N,K,windows,graph,xcd", [(500, 50, 100, 0, 0), (2000, 200, 100, 20, 0),
                                                   (3000, 300, 60, 20, 8)])
def test_brunel_exact_parity(N, K, windows, graph, xcd):
    from naer2_amd.networks import brunel_exact

    _say("brunel_exact parity", N, K, windows, graph, xcd)
    net, _ = brunel_exact(N, K=K, device="cuda")
    evn, v0 = oracle_net(net)
    st, nodes, tms, cnt = _oracle(evn, v0, windows)
    rt = net.runtime
    rt.set_xcd(xcd)
    rt.run(windows, graph)
    assert cnt.sum() > 50
    _assert_parity(rt, st, nodes, tms, cnt)


def test_chunked_and_graph_windows_equal_eager():
    from naer2_amd.networks import brunel_exact

    outs = []
    for runs in ([(60, 0)], [(13, 0), (40, 10), (7, 0)], [(60, 30)]):
        _say("brunel_exact chunks", runs)
        net, _ = brunel_exact(1000, K=100, device="cuda", seed=3)
        rt = net.runtime
        for n, g in runs:
            rt.run(n, g)
        torch.cuda.synchronize()
        rt.check()
        outs.append((rt.spikes(), rt.v.cpu().numpy()))
    (t0, n0), v0 = outs[0]
    for (t, n), v in outs[1:]:
        assert np.array_equal(n, n0)
        assert np.array_equal(t.view(np.uint32), t0.view(np.uint32))
        assert np.array_equal(v.view(np.uint32), v0.view(np.uint32))


@pytest.mark.parametrize("seed", [1, 2])
def test_ties_and_trains_parity(seed):
    """Synchronous spike trains (equal times across sources and connections) with Poisson and recurrent
    inputs: arrival ties broken by connection, then source neuron."""
    from tests.test_event_oracle import _tiny_random

    import naer2_amd  # noqa: F401

    net = _tiny_random(seed, True)
    net.device = torch.device("cuda")
    _say("ties parity", seed)
    evn, v0 = oracle_net(net)
    st, nodes, tms, cnt = _oracle(evn, v0, 40)
    rt = net.runtime
    rt.run(40)
    _assert_parity(rt, st, nodes, tms, cnt)


@pytest.mark.parametrize("seed,delays", [(5, (1.5, 3.0, 4.5)), (6, (0.5, 1.0, 1.5, 0.5))])
def test_several_delays_parity(seed, delays):
    """Connections with delays of several windows (the smallest delay is the window): a window's input stream
    merges the sources' spikes of several past windows, from the sort kernel's history of windows."""
    from tests.test_event_oracle import _tiny_random

    net = _tiny_random(seed, True, delays)
    net.device = torch.device("cuda")
    evn, v0 = oracle_net(net)
    n_win = int(60 / evn.window)
    _say("several delays", seed, delays, n_win)
    st, nodes, tms, cnt = _oracle(evn, v0, n_win)
    rt = net.runtime
    rt.run(n_win, 20)
    _assert_parity(rt, st, nodes, tms, cnt)


def test_brunel_several_delays_parity():
    """Brunel with E delays of 1.5 ms, I delays of 3.0 ms and source delays of 0.75 ms (window 0.75 ms, lags
    2, 4 and 1), 2,000 neurons, eager and captured windows, chunked: bit for bit with the oracle."""
    from naer2_amd.networks import brunel_exact

    net, pops = brunel_exact(2000, K=200, device="cuda", seed=4)
    for c in net.connections:
        c.delay = {"E": 1.5, "I": 3.0, "P": 0.75}[[k for k, p in pops.items() if p is c.source][0]]
    evn, v0 = oracle_net(net)
    assert evn.window == F32(0.75)
    n_win = 160
    _say("brunel several delays", n_win)
    st, nodes, tms, cnt = _oracle(evn, v0, n_win)
    rt = net.runtime
    assert rt.max_lag == 4
    rt.run(37, 0)
    rt.run(100, 25)
    rt.run(23, 0)
    _assert_parity(rt, st, nodes, tms, cnt)


def test_equal_arrivals_from_distinct_emissions():
    """Emission times 1022.5 and the next float both arrive at 1024.0: the GPU orders them by source neuron as
    the oracle does (the stream's bins hold equal arrivals together)."""
    e1 = F32(1022.5)
    e2 = np.nextafter(e1, F32(2000.0))
    rng = np.random.default_rng(9)
    times = [[float(e2)], [float(e1)]] + [[float(t)] for t in rng.uniform(1020.0, 1025.0, 30).astype(F32)]
    links = [(s, t, float(w)) for s in range(len(times)) for t in range(8)
             for w in [rng.choice([7.0, 1e-3, -2.0, 0.3])]]
    net, T = train_net([times], 8, [links], 3.0, device="cuda")
    n_win = int(1026 / 1.5) + 2
    _say("equal arrivals", n_win)
    evn, v0 = oracle_net(net)
    st, nodes, tms, cnt = _oracle(evn, v0, n_win)
    rt = net.runtime
    rt.run(n_win, 50)
    _assert_parity(rt, st, nodes, tms, cnt)


def test_burst_beyond_lds_sort():
    """12,000 train sources firing in one window (some at equal times): the next window's input stream
    (12,000 spikes) is sorted in global memory, and each group takes several rounds of entries."""
    rng = np.random.default_rng(5)
    n_src, n_tgt = 12000, 256
    base = rng.uniform(0.0, 1.5, 3000).astype(F32)
    times = [[float(rng.choice(base))] for _ in range(n_src)]
    links = []
    for t in range(n_tgt):
        for s in rng.integers(0, n_src, 400):
            links.append((int(s), t, float(rng.choice([0.05, -0.2, 0.1]))))
    net, T = train_net([times], n_tgt, [links], rng.uniform(0, 20, n_tgt).astype(F32), device="cuda")
    _say("burst", n_src, n_tgt)
    evn, v0 = oracle_net(net)
    st, nodes, tms, cnt = _oracle(evn, v0, 4)
    rt = net.runtime
    rt.run(4)
    assert cnt[0] == n_src
    _assert_parity(rt, st, nodes, tms, cnt)


@pytest.mark.timeout(600)
def test_reference_brunel_gpu():
    """test/test_brunel_gpu.py:74-105 on the GPU -- N = 5000, K = 1000, 200 ms -- bit for bit with the oracle,
    and the reference's bounds: rates 30-46 Hz, sources within 5 % of theory, off-grid times, CV(ISI) > 0.3."""
    from naer2_amd.networks import brunel_exact, brunel_exact_theory

    N, K, T0 = 5000, 1000, 50.0
    net, pops = brunel_exact(N, K=K, device="cuda")
    evn, v0 = oracle_net(net)
    windows = 133
    _say("reference brunel", N, K, windows)
    st, nodes, tms, cnt = _oracle(evn, v0, windows)
    net.graph_window = 19
    net.step(windows * 1.5)
    rt = net.runtime
    _assert_parity(rt, st, nodes, tms, cnt)
    ref = brunel_exact_theory(N, K=K)
    t_end = windows * 1.5
    for name in ("E", "I", "P"):
        times, ids = net.spikes(pops[name])
        n = pops[name].n
        rate = float((times >= T0).sum()) / n / ((t_end - T0) / 1000.0)
        if name == "P":
            assert abs(rate - ref["rate_ext_Hz"]) < 0.05 * ref["rate_ext_Hz"], rate
            continue
        assert 30.0 < rate < 46.0, (name, rate)
        assert np.isclose(times * 10, np.round(times * 10), atol=1e-3).mean() < 0.05
        cvs = []
        for j in range(min(n, 500)):
            tj = np.sort(times[(ids == j) & (times >= T0)])
            if len(tj) >= 4:
                isi = np.diff(tj)
                cvs.append(float(isi.std(ddof=1) / isi.mean()))
        assert len(cvs) > 200 and sum(cvs) / len(cvs) > 0.3


def test_delivery_log():
    """The GPU's delivery log (window, arrival, connection, source, target) equals the oracle's: per target the
    same inputs in the same order, arrival times bit for bit (naer2/runtime.py:747-752's deliveries())."""
    from naer2_amd.networks import brunel_exact

    net, _ = brunel_exact(1000, K=100, device="cuda", seed=2)
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    olog: list = []
    ev.run(evn, st, 30, log=olog, log_cap=5_000_000)
    rt = net.runtime
    rt.log_deliveries(5_000_000)
    _say("delivery log", len(olog))
    rt.run(30, 10)
    glog = rt.deliveries()
    rt.log_deliveries(None)
    assert len(glog) == len(olog) > 1000
    tpop = {c: int(t) for c, t in enumerate(evn.conn_tgt)}
    by_o, by_g = {}, {}
    for r in olog:
        by_o.setdefault((tpop[r[2]], r[4]), []).append((r[0], F32(r[1]).view(np.uint32), r[2], r[3]))
    for r in glog:
        by_g.setdefault((tpop[r[2]], r[4]), []).append((r[0], F32(r[1]).view(np.uint32), r[2], r[3]))
    assert by_o == by_g


def test_lost_spikes_raise():
    """A region too small for a window's spikes (history_cap) is flagged, never overrun: check() raises."""
    import naer2_amd as n2

    E = n2.LIFDeltaExact() * 64
    P = n2.PoissonExact(2.0) * 64  # ~190 spikes per window into a region of 64
    net = E.network
    net.device = torch.device("cuda")
    net.merge(P.network)
    c = P >> n2.FixedInDegree(4) >> E
    c.weight, c.delay = 0.1, 1.5
    P.history_cap = 1
    _say("lost spikes flag")
    rt = net.runtime
    rt.run(3)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="region"):
        rt.check()


def test_long_segment_raises():
    """A source with more than 2048 links into one group: the excess is not delivered, and check() raises."""
    net, T = train_net([[[0.5]]], 64, [[(0, t % 64, 0.001) for t in range(3000)]], 0.0, device="cuda")
    _say("long segment flag")
    rt = net.runtime
    rt.run(3)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="segment"):
        rt.check()
