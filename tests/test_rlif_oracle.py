"""Pins of the FP64 BPTT oracle (oracle/rlif_oracle.py): closed-form sub-threshold response, the
surrogate's shape, spike/reset timing, and central finite differences of the smooth model.  CPU only."""
from __future__ import annotations

import math

import pytest
import torch

from oracle import rlif_oracle as ro

ALPHA, BETA = math.exp(-1 / 5), math.exp(-1 / 20)


def _tiny(seed=0, T=12, B=3, C=5, H=7, O=4, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    x = (torch.rand(T, B, C, generator=g) < 0.3).double()
    w_in = torch.randn(H, C, generator=g, dtype=torch.float64) * scale
    w_rec = torch.randn(H, H, generator=g, dtype=torch.float64) * scale / math.sqrt(H)
    w_out = torch.randn(O, H, generator=g, dtype=torch.float64) / math.sqrt(H)
    labels = torch.randint(0, O, (B,), generator=g)
    return x, w_in, w_rec, w_out, labels


def test_subthreshold_response_is_the_double_exponential():
    """One input pulse at t = 0: V_t = u (beta^{t+1} - alpha^{t+1}) / (beta - alpha), no spike."""
    x = torch.zeros(40, 1, 1, dtype=torch.float64)
    x[0, 0, 0] = 1.0
    u = 0.2  # peak potential ~0.72, below threshold
    w_in = torch.full((1, 1), u, dtype=torch.float64)
    _, s, v = ro.unroll(x, w_in, torch.zeros(1, 1, dtype=torch.float64), torch.zeros(1, 1, dtype=torch.float64),
                        ALPHA, BETA, 1.0, 25.0)
    assert float(s.abs().sum()) == 0.0
    for t in range(40):
        expect = u * (BETA ** (t + 1) - ALPHA ** (t + 1)) / (BETA - ALPHA)
        assert float(v[t, 0, 0]) == pytest.approx(expect, rel=1e-12)


def test_spike_and_subtractive_reset_one_step_later():
    x = torch.zeros(3, 1, 1, dtype=torch.float64)
    x[0, 0, 0] = 1.0
    w_in = torch.full((1, 1), 1.5, dtype=torch.float64)
    _, s, v = ro.unroll(x, w_in, torch.zeros(1, 1, dtype=torch.float64), torch.zeros(1, 1, dtype=torch.float64),
                        ALPHA, BETA, 1.0, 25.0)
    # the threshold subtracted one step after each spike
    v1 = (BETA * 1.5 + ALPHA * 1.5) - 1.0
    v2 = (BETA * v1 + ALPHA ** 2 * 1.5) - 1.0
    assert float(v[1, 0, 0]) == pytest.approx(v1, rel=1e-12)
    assert float(v[2, 0, 0]) == pytest.approx(v2, rel=1e-12)
    assert s[:, 0, 0].tolist() == [1.0, float(v1 > 1.0), float(v2 > 1.0)]


def test_surrogate_is_the_derivative_of_the_smooth_spike():
    xs = torch.linspace(-0.5, 0.5, 101, dtype=torch.float64).requires_grad_(True)
    (g_sm,) = torch.autograd.grad(ro.smooth_spike(xs, 25.0).sum(), xs)
    x2 = xs.detach().clone().requires_grad_(True)
    (g_su,) = torch.autograd.grad(ro._Spike.apply(x2, 25.0, 1.0, None).sum(), x2)
    assert torch.allclose(g_sm, g_su, rtol=1e-12, atol=0)
    assert float(g_su[50]) == pytest.approx(1.0, rel=1e-12)  # 1 / (1 + 0)^2 at the threshold


def test_smooth_model_gradient_matches_central_differences():
    x, w_in, w_rec, w_out, labels = _tiny(scale=0.8)
    kw = dict(alpha=ALPHA, beta=BETA, theta=1.0, slope=25.0, mode=1)
    _, gs, _ = ro.grads(x, w_in, w_rec, w_out, labels, **kw)
    eps = 1e-6
    ws = [w_in, w_rec, w_out]
    g = torch.Generator().manual_seed(7)
    for k, w in enumerate(ws):
        for _ in range(6):
            idx = tuple(int(torch.randint(0, n, (1,), generator=g)) for n in w.shape)
            wp = [v.clone() for v in ws]
            wm = [v.clone() for v in ws]
            wp[k][idx] += eps
            wm[k][idx] -= eps
            lp = ro.loss_fn(ro.unroll(x, *wp, **kw)[0], labels)
            lm = ro.loss_fn(ro.unroll(x, *wm, **kw)[0], labels)
            fd = float((lp - lm) / (2 * eps))
            assert float(gs[k][idx]) == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_spiking_model_gradient_uses_the_surrogate_chain():
    """Mode 0 with gamma = 0 switches the surrogate off: only the readout weights get a gradient, equal to
    the exact gradient of the loss as a function of W_out (spike trains fixed), checked by differences."""
    x, w_in, w_rec, w_out, labels = _tiny(seed=3, scale=3.0)
    kw = dict(alpha=ALPHA, beta=BETA, theta=1.0, slope=25.0, mode=0, gamma=0.0)
    _, gs, s = ro.grads(x, w_in, w_rec, w_out, labels, **kw)
    assert float(s.sum()) > 0
    assert float(gs[0].abs().max()) == 0.0 and float(gs[1].abs().max()) == 0.0
    eps = 1e-6
    for idx in [(0, 0), (1, 3), (3, 6)]:
        wp, wm = w_out.clone(), w_out.clone()
        wp[idx] += eps
        wm[idx] -= eps
        fd = float((ro.loss_fn(ro.unroll(x, w_in, w_rec, wp, **kw)[0], labels)
                    - ro.loss_fn(ro.unroll(x, w_in, w_rec, wm, **kw)[0], labels)) / (2 * eps))
        assert float(gs[2][idx]) == pytest.approx(fd, rel=1e-5, abs=1e-9)
