"""Pins of the CPU oracle (oracle/naer2_oracle.c) against published vectors, a library implementation,
closed forms, exact-arithmetic timing and brute force -- each chosen so that a plausible mistake (a wrong
constant, a dropped term, a wrong index, a changed order) fails one of them.  CPU only."""
from __future__ import annotations

import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

from naer2_amd.params import lif_delta_params, lif_exp_params, params_bytes, poisson_thresholds
from oracle import oracle
from tests.helpers import pd14_params, probe_params, tiny_topology

GOLDEN = Path(__file__).resolve().parent / "golden"

# Random123 known-answer vectors for philox4x32-10 (ctr[4], key[2] -> out[4]), the test file distributed
# with Salmon et al., SC'11.
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out
    assert oracle.philox_py(ctr, key) == out


def test_philox_matches_rocrand():
    vectors = json.loads((GOLDEN / "philox_rocrand.json").read_text())
    assert len(vectors) == 18
    for v in vectors:
        key = [v["seed"] & 0xFFFFFFFF, v["seed"] >> 32]
        assert list(oracle.philox(v["ctr"], key)) == v["out"]


@pytest.mark.parametrize("lam", [0.25, 1.0, 1.28, 2.32, 6.0])
def test_poisson_table_is_the_cdf(lam):
    thr = poisson_thresholds(lam)
    assert thr.dtype == np.uint32 and thr.shape == (32,)
    assert np.all(np.diff(thr.astype(np.int64)) >= 0)
    cdf = stats.poisson.cdf(np.arange(32), lam)
    expect = np.minimum(np.floor(cdf * 2.0**32), 2.0**32 - 1)
    assert np.all(np.abs(thr.astype(np.float64) - expect) <= 2.0)


def test_poisson_count_boundaries():
    thr = poisson_thresholds(1.0)
    assert oracle.poisson_count(0, thr) == 0
    assert oracle.poisson_count(int(thr[0]) - 1, thr) == 0
    assert oracle.poisson_count(int(thr[0]), thr) == 1
    assert oracle.poisson_count(int(thr[3]), thr) == 4
    none = poisson_thresholds(0.0)
    assert oracle.poisson_count(0xFFFFFFFE, none) == 0


@pytest.mark.parametrize("lam", [1.0, 2.32])
def test_poisson_draws_have_poisson_moments(lam):
    thr = poisson_thresholds(lam)
    n = 40000
    ks = np.array([oracle.poisson_count(int(oracle.philox([i, 7, 0, 0], [5, 0])[0]), thr) for i in range(n)])
    assert abs(ks.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(ks.var() - lam) < 5 * lam * math.sqrt(2.0 / n) + 0.02
    # the whole distribution: chi-square against the pmf over the bins with enough mass
    kmax = int(lam + 6 * math.sqrt(lam))
    obs = np.bincount(np.minimum(ks, kmax), minlength=kmax + 1)
    exp = stats.poisson.pmf(np.arange(kmax + 1), lam) * n
    exp[-1] = stats.poisson.sf(kmax - 1, lam) * n
    assert stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue > 1e-4


def _single(model, params, v0, isyn0=None, n_steps=1, ref0=0):
    st = oracle.State(1, 1, np.array([v0], np.float32), None if isyn0 is None else np.array([isyn0], np.float32),
                      np.array([ref0], np.int32))
    trace = []
    for s in range(n_steps):
        r, _ = oracle.run(model, [0, 1], params_bytes([params]), 1, 0, 0, np.zeros(2, np.int64), np.zeros(0, np.int32),
                          np.zeros(0, np.float32), np.zeros(0, np.int32), st, s, 1)
        trace.append((float(st.v[0]), None if st.isyn is None else float(st.isyn[0]), bool(r[0, 0])))
    return trace


def test_delta_lif_free_decay_is_exponential():
    p = lif_delta_params(0.1, 20.0, 1e9, 10.0, 2.0, 0.0, 0.0)
    tr = _single(oracle.LIF_DELTA, p, 15.0, n_steps=200)
    for n, (v, _, _) in enumerate(tr, start=1):
        # fp32 rounding of the decay constant and of each product: ~n * 2^-24 relative
        assert v == pytest.approx(15.0 * math.exp(-n * 0.1 / 20.0), rel=1e-6 + 1.2e-7 * n)


def test_delta_lif_refractory_window_matches_float_time_rule():
    """tools/brunel_reference.py:38-41: active iff t > ref_until, t = (n+1) dt, ref_until = t_spike + t_ref,
    evaluated here in exact rational arithmetic."""
    dt, t_ref = Fraction(1, 10), Fraction(2)
    p = lif_delta_params(0.1, 20.0, 20.0, 10.0, 2.0, 0.0, 0.0)
    tr = _single(oracle.LIF_DELTA, p, 100.0, n_steps=30)
    assert tr[0][2] and not any(f for _, _, f in tr[1:])
    ref_until = dt * 1 + t_ref
    for n in range(1, 30):
        active = dt * (n + 1) > ref_until
        v = tr[n][0]
        if not active:
            assert v == 10.0, n
        else:
            assert v < 10.0, n
    assert [dt * (n + 1) > ref_until for n in range(1, 30)].index(True) == 20


def test_exp_lif_matches_closed_form():
    """The exact propagator (Rotter & Diesmann 1999; test/networks/PD14/neuron.py:37-50): after n fixed
    steps V and I equal the float64 solution of the ODE at t = n dt."""
    p = lif_exp_params(0.1, 10.0, 0.5, 40.0, 2.0, -65.0, 1e9, -65.0, 0.0, 0.0)
    tr = _single(oracle.LIF_EXP, p, -70.0, isyn0=0.6, n_steps=80)
    for n, (v, i, _) in enumerate(tr, start=1):
        ve, ie = oracle.exact_lif_exp(-70.0, 0.6, n * 0.1, 10.0, 0.5, 40.0, -65.0)
        assert v == pytest.approx(ve, rel=1e-6 + 1.2e-7 * n, abs=1e-5)
        assert i == pytest.approx(ie, rel=1e-6 + 1.2e-7 * n, abs=1e-30)


def test_exp_lif_refractory_window_matches_float_time_rule():
    """test/fused_reference.py:94-103: refractory iff t_end < t_ref_end, t_ref_end = t_end(spike) + t_ref."""
    dt, t_ref = Fraction(1, 10), Fraction(2)
    p = lif_exp_params(0.1, 10.0, 0.5, 40.0, 2.0, -65.0, -50.0, -65.0, 0.0, 0.0)
    tr = _single(oracle.LIF_EXP, p, -40.0, isyn0=1.0, n_steps=30)
    assert tr[0][2]
    t_ref_end = dt * 1 + t_ref
    for n in range(1, 30):
        refractory = dt * (n + 1) < t_ref_end
        if refractory:
            assert tr[n][0] == -65.0, n
        else:
            assert tr[n][0] > -65.0, n
    assert sum(dt * (n + 1) < t_ref_end for n in range(1, 30)) == 19


def _forced_inputs(n, topo, forced, d_slots, accum_mode=0):
    """Run the probe model one step at a time with host-forced spikes (forced[s] = the spikes of step s,
    delivered from step s+1); returns the ring input read at every step."""
    src_ptr, tgt, w, d = topo
    st = oracle.State(n, d_slots, np.zeros(n, np.float32))
    xs = []
    for s in range(forced.shape[0]):
        st.prev_spk[:] = forced[s - 1] if s > 0 else 0
        oracle.run(oracle.LIF_DELTA, [0, n], params_bytes([probe_params()]), d_slots, 0, 0, src_ptr, tgt, w, d, st,
                   s, 1, accum_mode=accum_mode)
        xs.append(st.v.copy())
    return np.array(xs), st


def test_delivery_totals_match_dense_brute_force():
    n, D = 48, 6
    topo = tiny_topology(n, 900, D, seed=1)
    rng = np.random.default_rng(2)
    forced = rng.random((40, n)) < 0.2
    xs, _ = _forced_inputs(n, topo, forced, D)
    expect = oracle.ring_inputs_dense(n, *topo, forced)
    assert np.abs(xs - expect).max() <= 1e-4 * max(1.0, np.abs(expect).max())
    assert np.abs(expect).sum() > 0


def test_delivery_order_hash_matches_contract():
    """The order-hash receiver (test/test_fused_generated_gpu.py:535-551, acc = acc*0.5 + w): any change of
    the summation order changes the result.  The brute force walks per target and arrival step, the oracle
    per source; both must give the contract's order bit for bit."""
    n, D = 24, 4
    src_ptr, tgt, w, d = tiny_topology(n, 300, D, seed=4)
    rng = np.random.default_rng(5)
    forced = rng.random((30, n)) < 0.3
    xs, _ = _forced_inputs(n, (src_ptr, tgt, w, d), forced, D, accum_mode=1)
    src = np.repeat(np.arange(n), np.diff(src_ptr))
    ring = np.zeros((D, n), np.float32)
    half = np.float32(0.5)
    for s in range(forced.shape[0]):
        if s > 0:
            spk = forced[s - 1]
            for i in range(n):
                for dd in range(1, D + 1):
                    p = None
                    for l in range(len(tgt)):  # ascending canonical link index
                        if tgt[l] == i and d[l] == dd and spk[src[l]]:
                            p = np.float32(np.float32(0.0) + w[l]) if p is None else np.float32(p * half + w[l])
                    if p is not None:
                        c = (s - 1 + dd) % D
                        ring[c, i] = np.float32(ring[c, i] * half + p)
        x = ring[s % D].copy()
        ring[s % D] = 0
        assert np.array_equal(x.view(np.uint32), (np.float32(0.0) + xs[s]).view(np.uint32)), s


def test_arrival_is_exactly_spike_step_plus_delay():
    n = 10
    src_ptr = np.array([0, 3] + [3] * (n - 1), np.int64)
    tgt = np.array([4, 4, 7], np.int32)
    w = np.array([1.0, 2.0, 4.0], np.float32)
    d = np.array([3, 5, 5], np.int32)
    forced = np.zeros((12, n), bool)
    forced[2, 0] = True  # spike of step 2
    xs, _ = _forced_inputs(n, (src_ptr, tgt, w, d), forced, 5)
    got = {(s, i): float(xs[s, i]) for s in range(12) for i in range(n) if xs[s, i] != 0}
    assert got == {(5, 4): 1.0, (7, 4): 2.0, (7, 7): 4.0}


def test_ring_slot_is_clear_after_its_read():
    n, D = 30, 5
    topo = tiny_topology(n, 400, D, seed=8)
    forced = np.random.default_rng(9).random((20, n)) < 0.3
    _, st = _forced_inputs(n, topo, forced, D)
    assert np.all(st.ring[19 % D] == 0)


def test_exp_lif_state_stays_finite_in_a_small_network():
    n = 200
    topo = tiny_topology(n, 8000, 20, seed=11, wscale=0.05)
    st = oracle.State(n, 20, np.full(n, -60.0, np.float32), np.zeros(n, np.float32))
    r, ev = oracle.run(oracle.LIF_EXP, [0, n], params_bytes([pd14_params()]), 20, 3, 0, *topo, st, 0, 500)
    assert np.isfinite(st.v).all() and np.isfinite(st.isyn).all()
    assert r.sum() > 0 and ev > 0


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_oracle_equals_the_serial_one(threads):
    """The OpenMP oracle (target ranges per thread, the CPU baseline at every core) gives exactly the
    serial oracle's raster and state."""
    n = 300
    topo = tiny_topology(n, 12000, 20, seed=12, wscale=0.05)
    mk = lambda: oracle.State(n, 20, np.full(n, -60.0, np.float32), np.zeros(n, np.float32))
    a, b = mk(), mk()
    pb = params_bytes([pd14_params(), pd14_params()])
    ra, ea = oracle.run(oracle.LIF_EXP, [0, 200, n], pb, 20, 3, 0, *topo, a, 0, 400)
    par = oracle.Parallel(n, *topo, threads=threads)
    rb, eb = par.run(oracle.LIF_EXP, [0, 200, n], pb, 20, 3, 0, b, 0, 400)
    par.close()
    assert ra.sum() > 0 and ea == eb
    assert np.array_equal(ra, rb)
    for x, y in ((a.v, b.v), (a.isyn, b.isyn), (a.ring, b.ring)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(a.ref, b.ref)
