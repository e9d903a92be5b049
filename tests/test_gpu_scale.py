"""Parity at the configurations' full sizes, through the C-ABI, bit for bit against the oracle (the OpenMP
build, pinned equal to the serial one in the CPU tests): rasters, potentials, currents, refractory counters
and the whole ring.

- the full-scale PD14 microcircuit (77,169 neurons, 1,210 groups: the 32-word spike-bitmap step variant the
  benchmark times) over 1,000 steps;
- a 40,000-neuron Brunel network (625 groups: the 16-word variant) and a 150,000-neuron one (2,344 groups:
  the variant without an LDS bitmap, segment intake only);
- a batch of sweep networks spanning the sweep's whole size range, every network checked against the oracle;
- the 256-network sweep itself, batched against its networks run one after another.

Mirrors test/test_fused_generated_gpu.py:56-91 (the generated kernel against the hand-written reference,
round by round) and test/test_fusion.py:40-85 (a fused round equals the separate runs)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import torch

from naer2_amd import FusedRound
from naer2_amd.networks import brunel, brunel_sweep, pd14, sweep_params
from naer2_amd.params import params_bytes
from oracle import oracle
from tests.gpu_util import assert_state_equal, bits_equal, oracle_state

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _parallel_oracle(rt, st, step0, n_steps):
    """The OpenMP oracle over the runtime's own canonical table, in place on st; returns the raster."""
    c = rt.canon.to("cpu")
    par = oracle.Parallel(rt.n, c.src_ptr.numpy(), c.tgt.numpy(), c.w.numpy(), c.d.numpy(), threads=THREADS)
    model = oracle.LIF_DELTA if rt.model == 1 else oracle.LIF_EXP
    try:
        raster, _ = par.run(model, rt.pop_first, params_bytes(rt.params), rt.d_slots, rt.net.seed, rt.net.instance,
                            st, step0, n_steps, True)
    finally:
        par.close()
    return raster


def _check_against_oracle(net, n_steps, graph_window=100):
    rt = net.runtime
    st = oracle_state(rt)
    s0 = rt.step
    net.run(n_steps, graph_window=graph_window)
    torch.cuda.synchronize()
    raster = _parallel_oracle(rt, st, s0, n_steps)
    gpu = rt.raster(s0, s0 + n_steps)
    assert raster.sum() > 0, "no spikes: the comparison would be vacuous"
    if not np.array_equal(gpu, raster):
        bad = np.nonzero((gpu != raster).any(axis=1))[0]
        raise AssertionError(f"raster differs from step {s0 + int(bad[0])} on ({len(bad)} steps)")
    assert_state_equal(rt, st)
    return int(raster.sum())


def test_pd14_full_scale_bit_exact_1000_steps():
    _say("case: pd14 scale=1.0 seed=0, 1000 steps vs OpenMP oracle")
    net, _ = pd14(scale=1.0, seed=0, device="cuda")
    rt = net.build()
    assert rt.n == 77169 and rt.n_groups == 1210 and rt.d_slots > 32
    n_spk = _check_against_oracle(net, 1000)
    _say(f"pd14 full scale: {n_spk} spikes in 1000 steps, bit-exact")


@pytest.mark.parametrize("n_e,n_i,steps", [(32000, 8000, 1000), (120000, 30000, 400)])
def test_brunel_large_variants_bit_exact(n_e, n_i, steps):
    """625 groups (the 16-word bitmap variant, row intake built) and 2,344 groups (no LDS bitmap: the
    segment intake at every activity)."""
    _say(f"case: brunel N_E={n_e} N_I={n_i} C_E=200 C_I=50, {steps} steps vs OpenMP oracle")
    net, _ = brunel(N_E=n_e, N_I=n_i, seed=7, device="cuda")
    rt = net.build()
    groups = (n_e + 63) // 64 + (n_i + 63) // 64
    assert rt.n_groups == groups
    assert (rt.rows is not None) == (groups <= 1535)
    n_spk = _check_against_oracle(net, steps)
    _say(f"brunel {n_e + n_i}: {n_spk} spikes, bit-exact")


def test_sweep_size_range_batched_bit_exact_vs_oracle():
    """Networks of the 256-network sweep from its smallest to its largest (including g < 4 bursting ones),
    batched in one launch a step, every network equal to the oracle over 300 steps."""
    params = sweep_params(256, seed=0)
    sizes = np.array([p["N_E"] + p["N_I"] for p in params])
    gs = np.array([p["g"] for p in params])
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmin(gs))}
    for q in (0.25, 0.5, 0.75, 0.9):
        pick.add(int(np.argsort(sizes)[int(q * (len(sizes) - 1))]))
    chosen = [params[i] for i in sorted(pick)]
    _say(f"case: sweep batch vs oracle, sizes {[p['N_E'] + p['N_I'] for p in chosen]}, "
         f"g {[round(p['g'], 2) for p in chosen]}")
    nets = brunel_sweep(chosen, device="cuda", keep_canonical=True, record=True)
    rts = [n.runtime for n in nets]
    sts = [oracle_state(rt) for rt in rts]
    fr = FusedRound(rts, {"order": 1, "xcd": 0}, reserve_steps=300)
    fr.launch(300, graph_window=100)
    fr.close()
    torch.cuda.synchronize()
    for p, rt, st in zip(chosen, rts, sts):
        raster = _parallel_oracle(rt, st, 0, 300)
        assert np.array_equal(rt.raster(0, 300), raster), f"network N={p['N_E'] + p['N_I']} g={p['g']:.2f}"
        assert_state_equal(rt, st)
    assert max(sizes[i] for i in pick) > 45000 and min(sizes[i] for i in pick) < 1100


def test_sweep_256_batched_equals_sequential():
    """The whole 256-network sweep (9.1 G synapses): 200 batched steps equal the same networks run one
    after another from the same state, bit for bit (potentials, refractory counters, rings)."""
    params = sweep_params(256, seed=0)
    _say("case: 256-network sweep, 200 steps batched vs sequential")
    nets = []
    for i in range(0, 256, 32):
        nets += brunel_sweep(params[i:i + 32], device="cuda", keep_canonical=False, record=False)
        _say(f"built {len(nets)} networks")
    rts = [n.runtime for n in nets]
    snaps = [rt.snapshot() for rt in rts]
    fr = FusedRound(rts, {"order": 1, "xcd": 0})
    fr.launch(200, graph_window=100)
    fr.close()
    torch.cuda.synchronize()
    batched = [(rt.v.clone(), rt.ref.clone(), rt.ring.clone()) for rt in rts]
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
        rt.run(200, graph_window=100)
    torch.cuda.synchronize()
    for i, (rt, (v, r, ring)) in enumerate(zip(rts, batched)):
        assert bits_equal(rt.v.cpu().numpy(), v.cpu().numpy()), f"network {i}: V"
        assert torch.equal(rt.ref, r), f"network {i}: refractory counters"
        assert bits_equal(rt.ring.cpu().numpy(), ring.cpu().numpy()), f"network {i}: ring"
