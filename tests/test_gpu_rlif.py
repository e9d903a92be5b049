"""Surrogate-gradient BPTT kernels (naer2_amd.rlif) against the FP64 autograd unroll (oracle/rlif_oracle.py)."""
from __future__ import annotations

import ctypes
import math
import sys

import pytest
import torch

from naer2_amd import rlif as rl
from naer2_amd.rlif import RecurrentLIF, RLIFConfig, max_over_time_ce, poisson_input
from oracle import rlif_oracle as ro

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(net, x, labels):
    net.zero_grad(set_to_none=True)
    r, s = net(x)
    loss = max_over_time_ce(r, labels)
    loss.backward()
    return loss.detach(), r.detach(), s.detach(), [p.grad.detach().clone() for p in (net.w_in, net.w_rec, net.w_out)]


def _oracle(net, x, labels, cfg, forced=None):
    w = [p.detach().double() for p in (net.w_in, net.w_rec, net.w_out)]
    return ro.grads(x.double(), *w, labels, alpha=cfg.alpha, beta=cfg.beta, theta=cfg.theta, slope=cfg.slope,
                    gamma=cfg.gamma, mode=cfg.mode, forced=forced)


def test_smooth_model_gradients_match_fp64_autograd():
    _say("case: rlif mode 1 T=30 B=4 C=20 H=96 O=5")
    cfg = RLIFConfig(mode=1, checkpoint=7)
    g = torch.Generator(device="cuda").manual_seed(1)
    net = RecurrentLIF(20, 96, 5, cfg, generator=g)
    x = poisson_input(4, 20, 30, 1.0, g)
    labels = torch.randint(0, 5, (4,), generator=g, device="cuda")
    loss, r, _, grads = _run(net, x, labels)
    loss64, g64, _ = _oracle(net, x, labels, cfg)
    assert float(loss) == pytest.approx(float(loss64), rel=1e-5)
    for a, b in zip(grads, g64):
        assert ro.rel_err(a, b) <= 1e-4


def test_spiking_config_gradients_match_fp64_autograd():
    """The benchmark's shape: B=128, 700 inputs, H=1024, 20 readouts, T=100, checkpoints every 25 steps;
    the oracle replays the fp32 spike trains and takes the surrogate at its own float64 potentials."""
    _say("case: rlif mode 0 T=100 B=128 C=700 H=1024 O=20")
    cfg = RLIFConfig()
    g = torch.Generator(device="cuda").manual_seed(2)
    net = RecurrentLIF(700, 1024, 20, cfg, generator=g)
    x = poisson_input(128, 700, 100, 1.0, g)
    labels = torch.randint(0, 20, (128,), generator=g, device="cuda")
    loss, r, s, grads = _run(net, x, labels)
    assert 0.001 < float(s.mean()) < 0.5
    loss64, g64, s64 = _oracle(net, x, labels, cfg, forced=s.double())
    assert float(loss) == pytest.approx(float(loss64), rel=1e-5)
    for a, b in zip(grads, g64):
        assert ro.rel_err(a, b) <= 1e-4
    # the unforced fp64 forward fires the same spikes until rounding near the threshold lets the two runs
    # drift apart: identical over the first 10 steps
    _, s_free, _ = ro.unroll(x.double(), *[p.detach().double() for p in (net.w_in, net.w_rec, net.w_out)],
                             alpha=cfg.alpha, beta=cfg.beta, theta=cfg.theta, slope=cfg.slope)
    assert float((s_free[:10] != s[:10].double()).float().mean()) < 1e-4


def test_checkpoint_interval_does_not_change_gradients():
    """Spikes, readout and gradients are bit-identical: the recompute restarts the forward's own compensated
    state from each checkpoint, the same trajectory whatever the interval."""
    _say("case: rlif checkpoints 25 vs 7 vs 100")
    g = torch.Generator(device="cuda").manual_seed(3)
    x = poisson_input(16, 50, 100, 1.0, g)
    labels = torch.randint(0, 10, (16,), generator=g, device="cuda")
    out = []
    for k in (25, 7, 100):
        g2 = torch.Generator(device="cuda").manual_seed(4)
        net = RecurrentLIF(50, 256, 10, RLIFConfig(checkpoint=k), generator=g2)
        out.append(_run(net, x, labels))
    for o in out[1:]:
        assert torch.equal(o[2], out[0][2]) and torch.equal(o[1], out[0][1])
        for a, b in zip(o[3], out[0][3]):
            assert torch.equal(a, b)


@pytest.mark.parametrize("B,H", [(128, 1024), (5, 96), (33, 48)])
def test_delivery_adjoint_matches_a_float64_product(B, H):
    """The hand-written matrix-core product m = gI W_rec (through W_rec^T) against float64 torch, on ragged
    batch sizes and hidden sizes of 16 k."""
    _say(f"case: rlif delivery adjoint B={B} H={H}")
    lib = rl._bind(rl._lib.load())
    g = torch.Generator(device="cuda").manual_seed(B * 1000 + H)
    gi = torch.randn(B, H, generator=g, device="cuda")
    w = torch.randn(H, H, generator=g, device="cuda") / math.sqrt(H)
    w_t = w.t().contiguous()
    m = torch.full((B, H), float("nan"), device="cuda")
    lay = rl.layer(B, H, 1, RLIFConfig())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rl._check(lib.n2h_rlif_delivery_adjoint(ctypes.byref(lay), rl._ptr(gi), rl._ptr(w_t), rl._ptr(m), stream),
              "n2h_rlif_delivery_adjoint")
    torch.cuda.synchronize()
    ref = (gi.double().cpu() @ w.double().cpu())
    err = float((m.double().cpu() - ref).abs().max() / ref.abs().max())
    assert err < 1e-5, err
    lay_bad = rl.layer(B, H + 1, 1, RLIFConfig())
    assert lib.n2h_rlif_delivery_adjoint(ctypes.byref(lay_bad), rl._ptr(gi), rl._ptr(w_t), rl._ptr(m), stream) != 0
