"""The oracle's multi-rate step: populations with a timestep of their own, a whole multiple (stride) of the
network's resolution (naer2/population.py:191-204, 382-410; test/test_scheduling.py and
test/test_fusion.py:61-85 run populations at 0.2 / 0.3 ms beside others at 0.1 ms).

Pinned without the oracle's own formulas:
  - each update's input is the sum of its window's arrivals: integer weights, so the dense brute force
    (float64) gives exact sums to compare with, and non-update steps change nothing;
  - a network whose populations all run at stride k is the same network at resolution k * dt, spike for
    spike and bit for bit (no background drive, delays multiples of k);
  - an isolated exp-synapse neuron at stride k follows the closed-form solution at t = n k dt;
  - host plumbing: parameters built for the population's interval, ring slots, refusals.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from naer2_amd import FixedInDegree, LIFDelta, LIFDeltaExact, LIFExp
from naer2_amd.network import assemble, pop_stride
from naer2_amd.params import lif_delta_params, lif_exp_params, params_bytes
from oracle import oracle
from tests.helpers import probe_params, tiny_topology


def _with_stride(p, k):
    p.stride = k
    return p


def test_window_input_is_the_sum_of_its_steps_arrivals():
    """Three populations of probe neurons (V = window input at an update, unchanged otherwise) at strides
    1, 3 and 4 under host-forced spikes; every update's V equals the brute-force sum of its window."""
    sizes, strides, D = [20, 20, 20], [1, 3, 4], 5
    n = sum(sizes)
    pop_first = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    src_ptr, tgt, _, d = tiny_topology(n, 1500, D, seed=11)
    w = np.random.default_rng(12).integers(-8, 9, len(tgt)).astype(np.float32)
    d_slots = D + max(strides) - 1
    pb = params_bytes([_with_stride(probe_params(), k) for k in strides])
    rng = np.random.default_rng(13)
    n_steps = 61
    forced = rng.random((n_steps, n)) < 0.25
    st = oracle.State(n, d_slots, np.zeros(n, np.float32))
    vs = []
    for s in range(n_steps):
        st.prev_spk[:] = forced[s - 1] if s > 0 else 0
        raster, _ = oracle.run(oracle.LIF_DELTA, pop_first, pb, d_slots, 0, 0, src_ptr, tgt, w, d, st, s, 1)
        assert not raster.any()
        vs.append(st.v.copy())
    vs = np.array(vs, dtype=np.float64)
    dense = oracle.ring_inputs_dense(n, src_ptr, tgt, w, d, forced)
    assert np.abs(dense).sum() > 0
    for p, k in enumerate(strides):
        lo, hi = pop_first[p], pop_first[p + 1]
        before = np.zeros(hi - lo)
        for s in range(n_steps):
            if (s + 1) % k == 0:
                want = dense[s - k + 1:s + 1, lo:hi].sum(axis=0)
                assert np.array_equal(vs[s, lo:hi], want), (k, s)
            else:
                assert np.array_equal(vs[s, lo:hi], before), (k, s)
            before = vs[s, lo:hi]
    # what arrived after a population's last update, up to the latest arrival of the delivered spikes
    # (those of steps 0 .. n_steps - 2), is still in its ring columns: nothing lost, nothing read twice
    ext = np.zeros((n_steps + D + 1, n), bool)
    ext[:n_steps - 1] = forced[:n_steps - 1]
    later = oracle.ring_inputs_dense(n, src_ptr, tgt, w, d, ext)
    for p, k in enumerate(strides):
        lo, hi = pop_first[p], pop_first[p + 1]
        last = max(s for s in range(n_steps) if (s + 1) % k == 0)
        held = st.ring[:, lo:hi].astype(np.float64).sum(axis=0)
        assert np.array_equal(held, later[last + 1:, lo:hi].sum(axis=0)), k


@pytest.mark.parametrize("model,k", [(oracle.LIF_DELTA, 2), (oracle.LIF_DELTA, 5), (oracle.LIF_EXP, 3)])
def test_one_stride_network_equals_the_coarser_resolution(model, k):
    """Every population at stride k on dt = 0.1 against the same network at dt = 0.1 k, stride 1: spikes at
    the update steps only, on the same coarse steps, and the same final state bit for bit."""
    dt, n, S = 0.1, 240, 60
    rng = np.random.default_rng(21 + k)
    src_ptr, tgt, _, dc = tiny_topology(n, 4800, 3, seed=22 + k)
    w = rng.uniform(-1.0, 3.0, len(tgt)).astype(np.float32)
    if model == oracle.LIF_EXP:
        w *= np.float32(0.4)
    pop_first = np.array([0, 120, n], np.int32)
    if model == oracle.LIF_DELTA:
        mk = lambda: lif_delta_params(k * dt, 20.0, 20.0, 10.0, 2.0, 0.0, 0.0)
        v0 = rng.uniform(0.0, 24.0, n).astype(np.float32)
        i0 = None
    else:
        mk = lambda: lif_exp_params(k * dt, 10.0, 0.5, 40.0, 2.0, -65.0, -50.0, -65.0, 0.0, 0.0)
        v0 = rng.uniform(-66.0, -48.0, n).astype(np.float32)
        i0 = rng.uniform(0.0, 0.6, n).astype(np.float32)
    coarse_p = params_bytes([mk(), mk()])
    fine_p = params_bytes([_with_stride(mk(), k), _with_stride(mk(), k)])
    dmax = int(dc.max())
    cs = oracle.State(n, dmax, v0, i0)
    fs = oracle.State(n, dmax * k + k - 1, v0, i0)
    craster, _ = oracle.run(model, pop_first, coarse_p, dmax, 7, 0, src_ptr, tgt, w, dc, cs, 0, S)
    fraster, _ = oracle.run(model, pop_first, fine_p, dmax * k + k - 1, 7, 0, src_ptr, tgt, w, dc * k, fs, 0,
                            S * k)
    assert craster.sum() > craster[0].sum() > 0, "spikes beyond the first update"
    upd = np.arange(S * k) % k == k - 1
    assert not fraster[~upd].any()
    assert np.array_equal(fraster[upd], craster)
    for a, b in [(cs.v, fs.v), (cs.ref, fs.ref)] + ([(cs.isyn, fs.isyn)] if i0 is not None else []):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("k", [1, 4, 7])
def test_exp_lif_at_a_stride_follows_the_closed_form(k):
    """No input, no threshold: V and I after n updates are the exact solution at t = n k dt, and between
    updates they stay put."""
    dt = 0.1
    p = _with_stride(lif_exp_params(k * dt, 10.0, 0.5, 40.0, 2.0, -65.0, 1e9, -65.0, 0.0, 0.0), k)
    st = oracle.State(1, k, np.array([-70.0], np.float32), np.array([0.6], np.float32))
    for s in range(12 * k):
        v_before = st.v[0]
        oracle.run(oracle.LIF_EXP, [0, 1], params_bytes([p]), k, 0, 0, np.zeros(2, np.int64),
                   np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros(0, np.int32), st, s, 1)
        n_upd = (s + 1) // k
        if (s + 1) % k:
            assert st.v[0] == v_before
            continue
        v, i = oracle.exact_lif_exp(-70.0, 0.6, n_upd * k * dt, 10.0, 0.5, 40.0, -65.0)
        assert abs(st.v[0] - v) <= 2e-5 * abs(v), (s, st.v[0], v)
        assert abs(st.isyn[0] - i) <= 2e-5 * max(abs(i), 1e-3), (s, st.isyn[0], i)


def test_strides_outside_the_ring_or_the_limit_are_refused():
    pb = params_bytes([_with_stride(probe_params(), 4)])
    args = (np.zeros(3, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros(0, np.int32))
    with pytest.raises(RuntimeError):
        oracle.run(oracle.LIF_DELTA, [0, 2], pb, 3, 0, 0, *args, oracle.State(2, 3, np.zeros(2)), 0, 1)
    with pytest.raises(RuntimeError):  # the order-hash receiver takes stride-1 networks only
        oracle.run(oracle.LIF_DELTA, [0, 2], pb, 4, 0, 0, *args, oracle.State(2, 4, np.zeros(2)), 0, 1,
                   accum_mode=1)


# ---------------------------------------------------------------- host side (no GPU)

def _two_pops(ts_e=None, ts_i=None):
    E, I = LIFDelta(rate_ext=5.0) * 50, LIFDelta(rate_ext=5.0) * 20
    E.timestep, I.timestep = ts_e, ts_i
    net = E.network + I.network
    net.resolution = 0.1
    c = E >> FixedInDegree(5) >> I
    c.weight, c.delay = 0.1, 1.5
    c = I >> FixedInDegree(4) >> E
    c.weight, c.delay = -0.2, 0.8
    return net, E, I


def test_timestep_sets_stride_parameters_and_ring_slots():
    net, E, I = _two_pops(ts_i=0.3)
    a = assemble(net, device="cpu")
    pe, pi = a.params
    assert (pe.stride, pi.stride) == (1, 3)
    assert pi.decay == np.float32(math.exp(-0.3 / I.instance.tau))
    assert pe.decay == np.float32(math.exp(-0.1 / E.instance.tau))
    assert pi.ref_steps == round(I.instance.t_ref / 0.3)
    want = lif_delta_params(0.3, I.instance.tau, I.instance.theta, I.instance.V_r, I.instance.t_ref,
                            I.instance.J_ext, I.instance.rate_ext * 0.3)
    assert list(pi.poisson_thr) == list(want.poisson_thr)
    assert a.d_slots == 15 + 3 - 1


def test_timestep_must_be_a_whole_multiple_of_the_resolution():
    assert pop_stride(_two_pops(ts_e=0.1)[1], 0.1) == 1
    with pytest.raises(ValueError, match="whole multiple"):
        assemble(_two_pops(ts_e=0.25)[0], device="cpu")
    with pytest.raises(ValueError, match="whole multiple"):
        assemble(_two_pops(ts_e=0.05)[0], device="cpu")
    with pytest.raises(ValueError, match="more than"):
        assemble(_two_pops(ts_e=6.5)[0], device="cpu")


def test_event_driven_populations_take_no_timestep():
    from naer2_amd.event import describe
    pop = LIFDeltaExact() * 10
    pop.timestep = 0.5
    with pytest.raises(NotImplementedError, match="timestep"):
        describe(pop.network)


def test_exp_lif_stride_parameters():
    E = LIFExp(K_ext=100) * 30
    E.timestep = 0.4
    net = E.network
    net.resolution = 0.1
    c = E >> FixedInDegree(3) >> E
    c.weight, c.delay = 0.01, 0.1
    a = assemble(net, device="cpu")
    (p,) = a.params
    assert p.stride == 4 and p.ref_steps == max(0, round(2.0 / 0.4) - 1)
    assert p.p_i == np.float32(math.exp(-0.4 / 0.5))
    assert a.d_slots == 1 + 4 - 1


def test_plan_lists_each_populations_windows():
    """Network.plan (naer2/network.py:154-204 -- test/test_scheduling.py:141-170 checks step() runs as
    planned): E every step, I (0.3 ms) in every third launch with its window; `until` bounds the plan."""
    net, E, I = _two_pops(ts_i=0.3)
    launches = net.plan(max_launches=7)
    assert len(launches) == 7
    for s, launch in enumerate(launches):
        got = {p: (round(t0, 9), round(t1, 9), g) for p, t0, t1, g in launch}
        assert got[E] == (round(s * 0.1, 9), round((s + 1) * 0.1, 9), False)
        if (s + 1) % 3 == 0:
            assert got[I] == (round((s - 2) * 0.1, 9), round((s + 1) * 0.1, 9), False)
        else:
            assert I not in got
    assert len(net.plan(until=0.5)) == 5
    assert net.plan(max_launches=0) == []


def test_plan_of_an_event_driven_network_gates_receivers():
    from naer2_amd import PoissonExact
    src, tgt = PoissonExact(rate=10.0) * 20, LIFDeltaExact() * 30
    net = src.network + tgt.network
    c = src >> FixedInDegree(5) >> tgt
    c.weight, c.delay = 0.1, 1.5
    launches = net.plan(max_launches=3)
    assert [[(p, round(a, 9), round(b, 9), g) for p, a, b, g in L] for L in launches] == [
        [(src, round(k * 1.5, 9), round((k + 1) * 1.5, 9), False), (tgt, round(k * 1.5, 9), round((k + 1) * 1.5, 9), True)]
        for k in range(3)]
