"""The event-driven oracle (oracle/event_oracle.c) pinned on the CPU: its fp32 exp/log against float64, closed
forms of the neuron, the one-threshold-per-time rule, refractoriness, the Poisson source's statistics, an
independent global event-queue formulation (bit for bit), and the reference's own Brunel statistics
(test/test_brunel_gpu.py:74-105) at its size."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import event as ev
from tests.event_util import oracle_net, train_net

F32 = np.float32
NRN = {"tau": 20.0, "theta": 20.0, "v_reset": 10.0, "t_ref": 2.0}


def _lib():
    return ev._lib()


def test_exp_accuracy():
    lib = _lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-87.0, 0.0, 20001), -rng.exponential(1.0, 5000), -rng.uniform(0, 0.1, 5000),
                         np.linspace(0.0, 10.0, 1001)]).astype(F32)
    y = np.array([lib.or_ev_exp(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    rel = np.abs(y - ref) / ref
    assert rel.max() < 2.5e-7, rel.max()
    assert lib.or_ev_exp(0.0) == 1.0
    assert lib.or_ev_exp(-100.0) == 0.0


def test_log_accuracy():
    lib = _lib()
    rng = np.random.default_rng(1)
    us = np.concatenate([rng.uniform(0, 1, 20000), np.exp(-rng.uniform(0, 17, 5000)),
                         np.linspace(0.5, 2.0, 4001), 10.0 ** rng.uniform(-30, 30, 2000)]).astype(F32)
    us = us[us > 0]
    y = np.array([lib.or_ev_log(float(u)) for u in us], np.float64)
    ref = np.log(us.astype(np.float64))
    err = np.abs(y - ref)
    big = np.abs(ref) > 1e-3
    assert (err[big] / np.abs(ref[big])).max() < 4e-7
    assert err[~big].max() < 1e-10
    assert lib.or_ev_log(1.0) == 0.0


def test_numpy_restatement_bit_identical():
    """The NumPy copies of exp/log (used by the event-queue check) are the C functions bit for bit."""
    lib = _lib()
    rng = np.random.default_rng(2)
    for x in np.concatenate([-rng.exponential(2.0, 300), rng.uniform(-87, 5, 300)]).astype(F32):
        assert F32(lib.or_ev_exp(float(x))).view(np.uint32) == ev.ev_exp(x).view(np.uint32)
    for u in np.concatenate([rng.uniform(0, 1, 300), np.exp(-rng.uniform(0, 17, 300))]).astype(F32):
        if u > 0:
            assert F32(lib.or_ev_log(float(u))).view(np.uint32) == ev.ev_log(u).view(np.uint32)


def _single(times_w, V0, window=1.5, n_windows=None, neuron=None):
    """One neuron fed by one spike-train source per input (time, weight): (its final state, its spikes)."""
    times = [[t] for t, _ in times_w]
    links = [[(0, 0, w)] for _, w in times_w]
    net, _ = train_net([[ts] for ts in times], 1, links, V0, window=window, neuron=neuron)
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    n_windows = n_windows or int(max(t for t, _ in times_w) / window) + 3
    nodes, tms, _, _ = ev.run(evn, st, n_windows)
    return st, nodes, tms


def test_closed_form_decay():
    """V after inputs at t_i + D: the exact solution sum w_i exp(-(t - a_i)/tau) (float64), no spike."""
    ins = [(0.3, 2.0), (1.1, 3.0), (4.7, -1.0), (9.2, 2.5)]
    V0, D, tau = 5.0, 1.5, 20.0
    st, nodes, _ = _single(ins, V0)
    assert not np.any(nodes == 0)  # node 0 is the neuron (its population was made first)
    a_last = ins[-1][0] + D
    v = V0 * math.exp(-(ins[0][0] + D) / tau)
    t = ins[0][0] + D
    for k, (ti, w) in enumerate(ins):
        a = ti + D
        v = v * math.exp(-(a - t) / tau) + w if k else v + w
        t = a
    assert abs(float(st.v[0]) - v) < 1e-5 * abs(v)
    assert st.last[0] == F32(F32(ins[-1][0]) + F32(D)) and abs(a_last - float(st.last[0])) < 1e-6


def test_threshold_after_all_inputs_at_one_time():
    """Two inputs at one exact time: +8 then -6 (or -6 then +8) from 15 mV: 15 e^{-a/20} + 2 < 20, no spike --
    a threshold between them would fire on the +8 first (fused.py:1511-1522: the send once per time)."""
    for order in ([(0.5, 8.0), (0.5, -6.0)], [(0.5, -6.0), (0.5, 8.0)]):
        st, nodes, _ = _single(order, 15.0)
        assert not np.any(nodes == 0)
        v = 15.0 * math.exp(-2.0 / 20.0) + 2.0
        assert abs(float(st.v[0]) - v) < 1e-5


def test_spike_time_reset_and_refractory_drop():
    """+12 at 2.0 from 15 fires at exactly 2.0 (resets to V_r = 10); an input 1 ms later falls in the 2 ms
    refractory period and is dropped; one 3 ms later decays from V_r over (a - (t + t_ref))."""
    st, nodes, tms = _single([(0.5, 12.0), (1.5, 5.0), (3.5, 1.0)], 15.0)
    fired = tms[nodes == 0]
    assert fired.tolist() == [F32(F32(0.5) + F32(1.5))]
    v = 10.0 * math.exp(-((3.5 + 1.5) - (2.0 + 2.0)) / 20.0) + 1.0
    assert abs(float(st.v[0]) - v) < 1e-5
    assert st.tre[0] == F32(4.0)


def test_connection_order_breaks_ties():
    """Equal arrivals from two connections: the lower connection's input first (_merge_segments' c2 < c).
    The weights are chosen so that the fp32 sum differs with the order: the result must be the c0-first one."""
    w0, w1 = 1e-3, 7.0
    st, _, _ = _single([(0.5, w1), (0.5, w0)], 3.0)  # connection 0 carries w1, connection 1 carries w0
    a = F32(F32(0.5) + F32(1.5))
    v = F32(F32(F32(3.0) * ev.ev_exp(F32(F32(F32(0.0) - a) * F32(1.0 / 20.0)))) + F32(w1))
    v = F32(v + F32(w0))
    assert st.v[0].view(np.uint32) == v.view(np.uint32)


def test_equal_arrivals_from_distinct_emissions_by_source():
    """Two emission times that round to one fp32 arrival (1022.5 and the next float, + 1.5 = 1024.0 both):
    inputs of one connection with equal arrival in source-neuron order (the source emitting later is neuron 0
    here, so it goes first), then the threshold once."""
    e1 = F32(1022.5)
    e2 = np.nextafter(e1, F32(2000.0))
    assert F32(e1 + F32(1.5)) == F32(e2 + F32(1.5))
    w0, w1 = 7.0, 1e-3  # source 0 fires at e2 (later), source 1 at e1
    net, _ = train_net([[[float(e2)], [float(e1)]]], 1, [[(0, 0, w0), (1, 0, w1)]], 3.0)
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    ev.run(evn, st, int(1024 / 1.5) + 3)
    a = F32(1024.0)
    v = F32(F32(F32(3.0) * ev.ev_exp(F32(F32(F32(0.0) - a) * F32(1.0 / 20.0)))) + F32(w0))
    v = F32(v + F32(w1))
    assert st.v[0].view(np.uint32) == v.view(np.uint32)
    hn, ht, hc, hv = ev.heap_run(evn, v0, int(1024 / 1.5) + 3)
    assert hv[0].view(np.uint32) == v.view(np.uint32)


def _tiny_random(seed: int, with_trains: bool, delays=None):
    import naer2_amd as n2
    import torch

    rng = np.random.default_rng(seed)
    E = n2.LIFDeltaExact(tau=20.0, theta=20.0, V_r=10.0, t_ref=2.0) * 24
    I = n2.LIFDeltaExact(tau=10.0, theta=18.0, V_r=5.0, t_ref=1.0) * 6
    P = n2.PoissonExact(0.5) * 20
    net = E.network
    net.device = torch.device("cpu")
    net.seed = seed
    net.merge(I.network)
    net.merge(P.network)
    pops = [E, I, P]
    if with_trains:
        # synchronous bursts: equal spike times across sources (ties in arrival, broken by connection and node)
        base = rng.uniform(0, 30, 6).astype(F32)
        times = [sorted(set(float(x) for x in rng.choice(base, 3))) for _ in range(8)]
        S = n2.SpikeTrains(times) * 8
        net.merge(S.network)
        pops.append(S)
    E.V = torch.as_tensor(rng.uniform(0, 20, 24), dtype=torch.float32)
    I.V = torch.as_tensor(rng.uniform(0, 18, 6), dtype=torch.float32)
    j = 0
    for tgt in (E, I):
        for src, K, w in ((E, 6, 1.2), (I, 2, -4.0), (P, 5, 1.5)) + (((pops[3], 2, 3.0),) if with_trains else ()):
            c = src >> n2.FixedInDegree(K) >> tgt
            c.weight = w
            c.delay = 1.5 if delays is None else delays[j % len(delays)]
            j += 1
    return net


@pytest.mark.parametrize("seed,trains", [(0, False), (1, True), (2, True)])
def test_windows_equal_one_event_queue(seed, trains):
    """The windowed oracle equals a global time-ordered event queue (oracle/event.py: heap_run) bit for bit:
    the spikes (neuron, time), per window, and the final potentials."""
    net = _tiny_random(seed, trains)
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    n_windows = 40
    nodes, tms, cnt, _ = ev.run(evn, st, n_windows)
    hn, ht, hc, hv = ev.heap_run(evn, v0, n_windows)
    assert cnt.sum() > 100
    assert np.array_equal(cnt, hc)
    assert np.array_equal(nodes, hn)
    assert np.array_equal(tms.view(np.uint32), ht.view(np.uint32))
    assert np.array_equal(st.v.view(np.uint32), hv.view(np.uint32))


def test_delivery_log_matches_event_queue():
    """The oracle's delivery log (window, arrival, connection, source, target) -- per target the inputs in
    delivery order -- equals the event queue's, and holds every input once."""
    net = _tiny_random(4, True)
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    log: list = []
    _, _, _, events = ev.run(evn, st, 30, log=log, log_cap=200000)
    hlog: list = []
    ev.heap_run(evn, v0, 30, log=hlog)
    assert len(log) == events == len(hlog) > 500
    # targets are indices within their population: key by (the connection's target population, target)
    tpop = {c: int(t) for c, t in enumerate(evn.conn_tgt)}
    by_o, by_h = {}, {}
    for r in log:
        by_o.setdefault((tpop[r[2]], r[4]), []).append(r)
    for r in hlog:
        by_h.setdefault((tpop[r[2]], r[4]), []).append(r)
    assert by_o == by_h


@pytest.mark.parametrize("seed,delays", [(5, (1.5, 3.0, 4.5)), (6, (0.5, 1.0, 1.5, 0.5)), (7, (2.0, 1.0))])
def test_several_delays_equal_one_event_queue(seed, delays):
    """Connections with different delays (whole multiples of the smallest, the window; the reference's delay
    classes, TODO.md:148-157): a spike reaches window k + delay / window, still in (arrival, connection,
    source) order per target -- windows equal the global event queue bit for bit."""
    net = _tiny_random(seed, True, delays)
    evn, v0 = oracle_net(net)
    assert evn.window == min(delays)
    st = ev.EvState.initial(evn, v0)
    n_windows = int(60 / evn.window)
    nodes, tms, cnt, _ = ev.run(evn, st, n_windows)
    hn, ht, hc, hv = ev.heap_run(evn, v0, n_windows)
    assert cnt.sum() > 100
    assert np.array_equal(cnt, hc)
    assert np.array_equal(nodes, hn)
    assert np.array_equal(tms.view(np.uint32), ht.view(np.uint32))
    assert np.array_equal(st.v.view(np.uint32), hv.view(np.uint32))
    # chunked runs carry the last windows' spikes
    b = ev.EvState.initial(evn, v0)
    parts = [ev.run(evn, b, k) for k in (n_windows // 3, 1, n_windows - n_windows // 3 - 1)]
    assert np.array_equal(np.concatenate([p[0] for p in parts]), nodes)
    assert np.array_equal(b.v.view(np.uint32), st.v.view(np.uint32))


def test_chunked_runs_equal_one_run():
    net = _tiny_random(3, True)
    evn, v0 = oracle_net(net)
    a = ev.EvState.initial(evn, v0)
    na, ta, ca, _ = ev.run(evn, a, 30)
    b = ev.EvState.initial(evn, v0)
    parts = [ev.run(evn, b, k) for k in (7, 1, 22)]
    assert np.array_equal(np.concatenate([p[0] for p in parts]), na)
    assert np.array_equal(np.concatenate([p[1] for p in parts]).view(np.uint32), ta.view(np.uint32))
    assert np.array_equal(a.v.view(np.uint32), b.v.view(np.uint32))


def test_poisson_statistics():
    """Counts per source over 200 ms at 20 Hz (test/networks/Brunel/poisson.py): mean rate T within 4 sigma;
    intervals exponential (Kolmogorov-Smirnov); the first spike an interval after 0."""
    from scipy import stats

    rate = 0.02
    pops = [ev.pops_from([(ev.POISSON, {"rate": rate})])[0]]
    n = 4000
    evn = ev.EvNet(np.array([0, n], np.int32), pops, np.zeros(0, np.int32), np.zeros(0, np.int32), 1.5,
                   np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32), seed=7)
    st = ev.EvState.initial(evn, np.zeros(n, np.float32))
    W = 134
    nodes, tms, cnt, _ = ev.run(evn, st, W)
    T = W * 1.5
    expect = n * rate * T
    assert abs(len(nodes) - expect) < 4 * math.sqrt(expect)
    # a Poisson process on [0, T]: given the counts, the times are uniform; the counts' variance equals their
    # mean; the first spike (an interval after 0) is exponential, truncated at T
    assert stats.kstest(tms.astype(np.float64) / T, "uniform").pvalue > 1e-3
    counts = np.bincount(nodes, minlength=n)
    disp = counts.var(ddof=1) / counts.mean()
    assert abs(disp - 1.0) < 6 * math.sqrt(2.0 / (n - 1)), disp
    order = np.lexsort((tms, nodes))
    nd, tt = nodes[order], tms[order].astype(np.float64)
    first = tt[np.r_[True, nd[1:] != nd[:-1]]]
    cdf = lambda x: (1 - np.exp(-rate * x)) / (1 - math.exp(-rate * T))
    assert stats.kstest(first, cdf).pvalue > 1e-3
    # every spike inside its window, times increasing per source
    k = np.repeat(np.arange(W), cnt)
    assert np.all(tms >= np.float32(k * 1.5)) and np.all(tms < np.float32((k + 1) * 1.5 + 1e-6))


@pytest.mark.timeout(600)
def test_reference_brunel_statistics():
    """test/test_brunel_gpu.py:74-105 on the oracle: N = 5000, K = 1000, 200 ms; after 50 ms the E and I rates
    in (30, 46) Hz, the sources' within 5 % of eta nu_thr, exact times off the 0.1 ms grid, CV(ISI) > 0.3."""
    import torch

    from naer2_amd.networks import brunel_exact, brunel_exact_theory

    N, K, T0, T = 5000, 1000, 50.0, 200.0
    torch.manual_seed(0)
    net, pops = brunel_exact(N, K=K, device="cpu")
    evn, v0 = oracle_net(net)
    st = ev.EvState.initial(evn, v0)
    n_windows = int(round(T / 1.5))
    nodes, tms, _, _ = ev.run(evn, st, n_windows)
    t_end = n_windows * 1.5
    ref = brunel_exact_theory(N, K=K)
    pf = evn.pop_first
    for p, name in enumerate(("E", "I", "P")):
        m = (nodes >= pf[p]) & (nodes < pf[p + 1])
        n = pf[p + 1] - pf[p]
        times = tms[m]
        rate = float((times >= T0).sum()) / n / ((t_end - T0) / 1000.0)
        if name == "P":
            assert abs(rate - ref["rate_ext_Hz"]) < 0.05 * ref["rate_ext_Hz"], rate
            continue
        assert 30.0 < rate < 46.0, (name, rate)
        on_grid = np.isclose(times * 10, np.round(times * 10), atol=1e-3).mean()
        assert on_grid < 0.05, on_grid
        ids = nodes[m] - pf[p]
        cvs = []
        for j in range(min(n, 500)):
            tj = np.sort(times[(ids == j) & (times >= T0)])
            if len(tj) >= 4:
                isi = np.diff(tj)
                cvs.append(float(isi.std(ddof=1) / isi.mean()))
        assert len(cvs) > 200 and sum(cvs) / len(cvs) > 0.3
