"""Degenerate networks against the oracle, bit for bit (raster, V, I, refractory counters, ring), through the
segment intake and the row intake taken at every step: a network without a single link, a one-neuron
network, and ragged populations (1, 63, 65 and 130 neurons: groups of one lane, a full group, a group and
one lane) joined by connections of which one holds no links; and the three in one batched launch under
every launch mapping."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd import FixedInDegree, LIFExp
from tests.gpu_util import Given, assert_state_equal, oracle_run, oracle_state

pytestmark = pytest.mark.gpu

STEPS = 500


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _finish(net, seed):
    net.resolution, net.seed, net.device = 0.1, seed, torch.device("cuda")
    net._generator = None
    net.keep_canonical = True
    return net


def _unconnected():
    return _finish((LIFExp(K_ext=2000) * 300).network, 3)


def _one_neuron():
    return _finish((LIFExp(K_ext=2400) * 1).network, 4)


def _ragged():
    rng = np.random.default_rng(7)
    a, b, c, d = LIFExp(K_ext=2200) * 1, LIFExp(K_ext=2000) * 63, LIFExp(K_ext=2100) * 65, LIFExp(K_ext=1900) * 130
    conns = [c >> FixedInDegree(5) >> d, d >> FixedInDegree(3) >> a, b >> FixedInDegree(4) >> c,
             a >> FixedInDegree(1) >> b, d >> FixedInDegree(6) >> d]
    for k, conn in enumerate(conns):
        n = conn.shape[0]
        sign = -1.0 if k == 1 else 1.0
        conn.weight = torch.as_tensor(sign * 60.0 * rng.random(n), dtype=torch.float32)
        conn.delay = torch.as_tensor(rng.integers(1, 12, n) * 0.1, dtype=torch.float32)
    empty = b >> Given(np.zeros(0, np.int64), np.zeros(0, np.int64)) >> b  # a connection without links
    empty.weight = torch.zeros(0, dtype=torch.float32)
    empty.delay = torch.zeros(0, dtype=torch.float32)
    return _finish(a.network, 5)


@pytest.mark.parametrize("make", [_unconnected, _one_neuron, _ragged], ids=["unconnected", "one_neuron", "ragged"])
@pytest.mark.parametrize("rows", [False, True], ids=["segments", "rows_every_step"])
def test_degenerate_networks_equal_the_oracle(make, rows):
    net = make()
    net.row_intake = rows
    rt = net.build()
    _say(f"case: {make.__name__}, n={rt.n}, groups={rt.n_groups}, links={rt.n_links}, "
         f"row intake every step={rows}, {STEPS} steps")
    if rows:
        assert rt.rows is not None
        rt.set_dense_min(0)
    st = oracle_state(rt)
    net.run(STEPS)
    torch.cuda.synchronize()
    raster, _ = oracle_run(rt, st, 0, STEPS)
    assert raster.sum() > 0
    assert np.array_equal(rt.raster(), raster)
    assert_state_equal(rt, st)


def test_degenerate_networks_batched_equal_the_oracle():
    """The three networks in one batched launch (FusedRound), 40 steps under every launch mapping the
    autotuner can choose, then each network against the oracle over all the steps."""
    from naer2_amd import FusedRound
    from naer2_amd.autotune import CANDIDATES
    nets = [_unconnected(), _one_neuron(), _ragged()]
    rts = [n.build() for n in nets]
    _say(f"case: batch of degenerate networks, sizes {[rt.n for rt in rts]}, links {[rt.n_links for rt in rts]}, "
         f"{len(CANDIDATES)} mappings x 40 steps")
    states = [oracle_state(rt) for rt in rts]
    for cand in CANDIDATES:
        fr = FusedRound(rts, cand, reserve_steps=40, weights=[1.0, 2.0, 3.0])
        fr.launch(40, graph_window=0)
        fr.close()
    torch.cuda.synchronize()
    steps = 40 * len(CANDIDATES)
    for rt, st in zip(rts, states):
        assert rt.step == steps
        raster, _ = oracle_run(rt, st, 0, steps)
        assert np.array_equal(rt.raster(), raster)
        assert_state_equal(rt, st)
    assert sum(int(rt.raster().sum()) for rt in rts) > 0
