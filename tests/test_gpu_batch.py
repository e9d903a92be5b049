"""Batched sweeps (FusedRound over n2h_batch_*): every network in a batch computes exactly what it computes
alone -- rasters, potentials, refractory counters and rings bit for bit -- for every launch mapping the
autotuner can choose."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd import FusedRound
from naer2_amd.autotune import CANDIDATES, measure_work, tune, tune_intake
from naer2_amd.networks import brunel_sweep, pd14, sweep_params
from tests.gpu_util import assert_state_equal, bits_equal, oracle_run, oracle_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params():
    return sweep_params(n_instances=5, seed=3, n_min=300, n_max=3000)


def _same(a, b):
    ra, rb = a.runtime, b.runtime
    assert ra.step == rb.step
    assert np.array_equal(ra.raster(), rb.raster())
    assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert np.array_equal(ra.ref.cpu().numpy(), rb.ref.cpu().numpy())
    assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_batch_equals_plain_for_every_mapping():
    params = _params()
    _say(f"case: batch of {len(params)} brunel instances, sizes {[p['N_E'] + p['N_I'] for p in params]}")
    batched = brunel_sweep(params, device="cuda", record=True)
    plain = brunel_sweep(params, device="cuda", record=True)
    total = 0
    weights = [3.0, 1.0, 4.0, 1.5, 9.0][:len(params)]  # order 2: a permutation of the plans
    for cand in CANDIDATES:
        _say(f"case: mapping {cand}")
        fr = FusedRound([n.runtime for n in batched], cand, reserve_steps=110, weights=weights)
        fr.launch(110, graph_window=48)
        fr.close()
        for n in plain:
            n.run(110, graph_window=0)
        total += 110
        torch.cuda.synchronize()
        for a, b in zip(batched, plain):
            _same(a, b)
    assert sum(int(n.runtime.raster().sum()) for n in plain) > 0
    assert total == batched[0].runtime.step


def test_batch_refuses_a_plan_stepped_alone():
    _say("case: batch refuses a diverged plan")
    nets = brunel_sweep(_params()[:3], device="cuda")
    fr = FusedRound([n.runtime for n in nets], {"order": 1, "xcd": 8})
    fr.launch(8)
    nets[1].run(1, graph_window=0)
    with pytest.raises(RuntimeError, match="step"):
        fr.launch(4)
    fr.close()


def test_autotune_returns_a_candidate_and_keeps_results():
    _say("case: autotune on a small batch")
    params = _params()
    batched = brunel_sweep(params, device="cuda", record=True)
    plain = brunel_sweep(params, device="cuda", record=True)
    rts = [n.runtime for n in batched]
    for n in batched + plain:  # off step 0, with spikes in the history
        n.run(40, graph_window=0)
    weights = measure_work(rts, steps=48, graph_window=48)
    assert all(w > 0 for w in weights) and all(rt.step == 40 for rt in rts)
    best, times = tune(rts, steps=48, graph_window=48, reserve_steps=400, use_cache=False, weights=weights)
    assert best in CANDIDATES and len(times) == len(CANDIDATES)
    assert all(rt.step == 40 for rt in rts)  # tuning and measuring restore the state
    fr = FusedRound(rts, best, reserve_steps=100, weights=weights)
    fr.launch(100, graph_window=48)
    fr.close()
    for n in plain:
        n.run(100, graph_window=0)
    torch.cuda.synchronize()
    for a, b in zip(batched, plain):
        ra, rb = a.runtime, b.runtime
        assert ra.step == rb.step == 140
        assert np.array_equal(ra.raster(40, 140), rb.raster(40, 140))
        assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
        assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_intake_tuning_restores_the_state_and_changes_no_result():
    """tune_intake times several row-intake thresholds from one snapshot and restores it: the run that
    follows equals an untuned run bit for bit (the intakes give identical results)."""
    p = dict(N_E=3200, N_I=800, C_E=320, C_I=80, g=3.2, eta=3.0, seed=8, instance=0)
    _say("case: intake tuning, bursting brunel N=4000")
    a, b = brunel_sweep([p, p], device="cuda", keep_canonical=False, record=True)
    a.run(40)
    b.run(40)
    best, times = tune_intake(a.runtime, steps=60, graph_window=20)
    assert best in times and len(times) >= 2
    assert a.runtime.step == b.runtime.step == 40
    a.run(200)
    b.run(200)
    torch.cuda.synchronize()
    ra, rb = a.runtime, b.runtime
    assert np.array_equal(ra.raster(40, 240), rb.raster(40, 240))
    assert ra.raster(40, 240).sum(axis=1).max() >= 200  # bursts: the thresholds mattered
    assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert np.array_equal(ra.ref.cpu().numpy(), rb.ref.cpu().numpy())
    assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_plan_mapping_tuning_restores_the_state_and_changes_no_result():
    """tune_plan times the single-network launch mappings from one snapshot and restores it: the run that
    follows equals an untuned run bit for bit (every mapping gives identical results)."""
    from naer2_amd.autotune import PLAN_MAPPINGS, tune_plan
    _say("case: launch-mapping tuning, pd14 scale=0.05")
    a, _ = pd14(scale=0.05, seed=6, device="cuda")
    b, _ = pd14(scale=0.05, seed=6, device="cuda")
    a.run(40)
    b.run(40)
    best, times = tune_plan(a.runtime, steps=60, graph_window=20)
    assert (best["xcd"], best["run"]) in PLAN_MAPPINGS and len(times) == len(PLAN_MAPPINGS)
    assert a.runtime.step == b.runtime.step == 40
    a.run(200)
    b.run(200)
    torch.cuda.synchronize()
    ra, rb = a.runtime, b.runtime
    assert np.array_equal(ra.raster(40, 240), rb.raster(40, 240)) and ra.raster(40, 240).sum() > 0
    assert bits_equal(ra.v.cpu().numpy(), rb.v.cpu().numpy())
    assert bits_equal(ra.isyn.cpu().numpy(), rb.isyn.cpu().numpy())
    assert np.array_equal(ra.ref.cpu().numpy(), rb.ref.cpu().numpy())
    assert bits_equal(ra.ring.cpu().numpy(), rb.ring.cpu().numpy())


def test_batch_of_large_networks_equals_plain_runs():
    """Networks of 513-832 groups take the batched kernel variant with a 13-word bitmap per lane (and the
    batch's smaller staging passes): the same results as each network alone, through sparse, burst and
    row-intake steps."""
    params = sweep_params(n_instances=3, seed=11, n_min=34000, n_max=52000)
    params[1]["g"], params[1]["eta"] = 3.0, 3.5  # a bursting one: bursts take the row intake
    _say(f"case: batch of large brunel instances, sizes {[p['N_E'] + p['N_I'] for p in params]}")
    batched = brunel_sweep(params, device="cuda", record=True)
    plain = brunel_sweep(params, device="cuda", record=True)
    assert all(513 <= n.runtime.n_groups <= 832 for n in batched)
    fr = FusedRound([n.runtime for n in batched], {"order": 1, "xcd": 0}, reserve_steps=144)
    fr.launch(144, graph_window=48)
    fr.close()
    for n in plain:
        n.run(144, graph_window=0)
    torch.cuda.synchronize()
    for a, b in zip(batched, plain):
        _same(a, b)
    counts = np.concatenate([n.runtime.raster().sum(axis=1) for n in plain])
    assert (counts <= 64).any() and (counts > 64).any()  # the sparse path and a bitmap or row intake


def test_batch_of_both_models_equals_plain_runs():
    """A heterogeneous round (naer2/fusion.py:61-116 over runtimes of different node types): delta-synapse
    Brunel networks and exponential-synapse PD14 microcircuits interleaved in one launch, each group running
    each network alone, under two launch mappings, and the first 96 batched steps of every network equal
    the oracle's (delta-synapse and exponential-synapse steps alike)."""
    from naer2_amd._lib import MODEL_LIF_EXP
    from naer2_amd.networks import pd14

    params = sweep_params(n_instances=3, seed=5, n_min=500, n_max=4000)

    def build():
        rts = [n.runtime for n in brunel_sweep(params, device="cuda", record=True, keep_canonical=True)]
        for i, scale in enumerate((0.02, 0.03)):
            net, _ = pd14(scale=scale, seed=7 + i, device="cuda", instance=i)
            net.record = True
            rts.append(net.build())
        return [rts[i] for i in (0, 3, 1, 4, 2)]  # the models interleaved in batch order

    batched, plain = build(), build()
    assert {rt.model for rt in batched} == {MODEL_LIF_EXP, batched[0].model} and batched[0].model != MODEL_LIF_EXP
    _say(f"case: mixed batch, sizes {[rt.n for rt in batched]}")
    states = [oracle_state(rt) for rt in batched]
    for cand in ({"order": 0, "xcd": 0}, {"order": 1, "xcd": 8}):
        _say(f"case: mapping {cand}")
        fr = FusedRound(batched, cand, reserve_steps=96)
        fr.launch(96, graph_window=48)
        fr.close()
        for rt in plain:
            rt.run(96, 0)
        torch.cuda.synchronize()
        for a, b in zip(batched, plain):
            assert a.step == b.step
            assert np.array_equal(a.raster(), b.raster())
            assert bits_equal(a.v.cpu().numpy(), b.v.cpu().numpy())
            assert np.array_equal(a.ref.cpu().numpy(), b.ref.cpu().numpy())
            assert bits_equal(a.ring.cpu().numpy(), b.ring.cpu().numpy())
            if a.model == MODEL_LIF_EXP:
                assert bits_equal(a.isyn.cpu().numpy(), b.isyn.cpu().numpy())
        if states is not None:  # after the first mapping's 96 steps: every network against the oracle
            for rt, st in zip(batched, states):
                raster, _ = oracle_run(rt, st, 0, 96)
                assert np.array_equal(rt.raster(0, 96), raster), f"network of {rt.n} neurons"
                assert_state_equal(rt, st)
            states = None
    assert all(int(rt.raster().sum()) > 0 for rt in plain)


def test_batch_sync_follows_restored_runtimes():
    """n2h_batch_sync: after the runtimes are restored to an earlier step, the same batch (synced, no new
    allocation) replays those steps bit for bit; a plan bound to a delivery log since is refused."""
    params = sweep_params(n_instances=3, seed=9, n_min=800, n_max=3000)
    rts = [n.runtime for n in brunel_sweep(params, device="cuda", record=True, keep_canonical=True)]
    _say(f"case: batch sync after restore, sizes {[rt.n for rt in rts]}")
    fr = FusedRound(rts, {"order": 0, "xcd": 0}, reserve_steps=200)
    fr.launch(40, graph_window=0)
    snaps = [rt.snapshot() for rt in rts]
    fr.launch(60, graph_window=0)
    torch.cuda.synchronize()
    first = [(rt.raster(40, 100), rt.v.clone(), rt.ring.clone()) for rt in rts]
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
    fr.sync()
    fr.launch(60, graph_window=0)
    torch.cuda.synchronize()
    for rt, (r, v, ring) in zip(rts, first):
        assert rt.step == 100 and np.array_equal(rt.raster(40, 100), r)
        assert torch.equal(rt.v.view(torch.int32), v.view(torch.int32))
        assert torch.equal(rt.ring.view(torch.int32), ring.view(torch.int32))
    rts[1].set_delivery_log(1000)
    with pytest.raises(RuntimeError, match="log"):
        fr.sync()
    rts[1].set_delivery_log(None)
    fr.close()
