"""The FP64 oracle of the rate-loss adjoint (oracle/adjoint_oracle.py), pinned on the CPU: its vectorised
Philox against the C oracle, its background counts against the C Poisson draw, and its autograd gradient
of the smooth model against FP64 central finite differences on a 4-neuron-per-population PD14 subsample."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from naer2_amd.network import assemble
from naer2_amd.networks import pd14
from oracle import adjoint_oracle as ao
from oracle import oracle


def _subsample(seed=3):
    net, _ = pd14(sizes=[4] * 8, prob_scale=10.0, seed=seed, device="cpu")
    a = assemble(net, "cpu")
    return ao.net_from(a.pop_first, a.params, a.canon, a.v.numpy(), np.zeros(a.n), a.dt), net


def test_vectorised_philox_matches_the_oracle():
    rng = np.random.default_rng(7)
    for _ in range(20):
        c = rng.integers(0, 2 ** 32, 4, dtype=np.uint64)
        k = rng.integers(0, 2 ** 32, 2, dtype=np.uint64)
        assert int(ao.philox_word0(*c, *k)) == int(oracle.philox(c.astype(np.uint32), k.astype(np.uint32))[0])


def test_background_counts_match_the_oracle_draw():
    d, net = _subsample()
    cnt = ao.poisson_counts(d["thr"], seed=net.seed, instance=0, step0=5, n_steps=3)
    for t in range(3):
        for i in (0, 9, 31):
            u = oracle.philox(np.array([i, 5 + t, 0, 0], np.uint32), np.array([net.seed, 0], np.uint32))[0]
            assert cnt[t, i] == oracle.poisson_count(int(u), d["thr"][i])
    assert cnt.mean() > 0.5  # K_ext * 8 Hz * 0.1 ms: ~1.2-2.3 inputs a step


SUR = {"cutoff": 0.8, "gamma": 0.5}  # truncated (cuts ~30 % of the smooth model's potentials), scaled


def _loss(d, alpha, counts, T, targets, frozen=None, sur=None):
    _, rates, _, _ = ao.unroll(d, alpha, T, counts, mode="smooth", frozen_reset=frozen, **(sur or {}))
    return ao.loss(rates, targets)


@pytest.mark.parametrize("reset,sur", [("live", None), ("frozen", None), ("frozen", SUR)])
def test_smooth_model_gradient_matches_central_differences(reset, sur):
    """Autograd through the smooth model equals FP64 central differences: with the reset live, and with
    the reset driven by the model's own unperturbed spikes held fixed (the detached-reset gradient the
    adjoint kernel computes)."""
    d, net = _subsample()
    T = 120
    counts = ao.poisson_counts(d["thr"], seed=net.seed, instance=0, step0=0, n_steps=T)
    rng = np.random.default_rng(11)
    alpha0 = torch.as_tensor(1.0 + 0.2 * rng.standard_normal((8, 8)), dtype=torch.float64)
    # targets near the model's own rates: a small loss, so that a small step leaves little roundoff (and a
    # small step seldom straddles the truncated surrogate's kinks)
    targets = ao.unroll(d, alpha0, T, counts, mode="smooth", **(sur or {}))[1].numpy() + rng.uniform(-5, 5, 8)
    frozen = None
    if reset == "frozen":
        S0, _, _, _ = ao.unroll(d, alpha0, T, counts, mode="smooth", **(sur or {}))
        frozen = S0.numpy()
    alpha = alpha0.clone().requires_grad_(True)
    L = _loss(d, alpha, counts, T, targets, frozen, sur)
    L.backward()
    g = alpha.grad.clone()
    assert torch.isfinite(g).all() and g.abs().max() > 0
    h = 1e-6
    # the truncated surrogate's clamp has kinks: a step that straddles one leaves an O(h) error
    tol = 1e-6 if sur is None else 1e-4
    entries = [(0, 0), (0, 2), (2, 3), (4, 5), (6, 7), (7, 6), (0, 5)]  # (0, 5): I_5 -> E_23 has no links
    for a, b in entries:
        e = torch.zeros(8, 8, dtype=torch.float64)
        e[a, b] = h
        fd = (_loss(d, alpha0 + e, counts, T, targets, frozen, sur) - _loss(d, alpha0 - e, counts, T, targets, frozen, sur)) / (2 * h)
        assert abs(float(fd) - float(g[a, b])) <= tol * max(1.0, abs(float(g[a, b]))), (a, b, float(fd), float(g[a, b]))
    assert float(g[0, 5]) == 0.0
    for _ in range(2):  # directional derivatives along random directions
        v = torch.as_tensor(rng.standard_normal((8, 8)), dtype=torch.float64)
        fd = (_loss(d, alpha0 + h * v, counts, T, targets, frozen, sur) - _loss(d, alpha0 - h * v, counts, T, targets, frozen, sur)) / (2 * h)
        ad = float((g * v).sum())
        assert abs(float(fd) - ad) <= tol * max(1.0, abs(ad)), (float(fd), ad)


@pytest.mark.parametrize("mode,sur", [("smooth-live", None), ("smooth-frozen", None), ("spike", None),
                                      ("smooth-frozen", SUR), ("spike", SUR)])
def test_reverse_recursion_matches_autograd(mode, sur):
    """The recursion the adjoint kernel implements (ao.adjoint, reset detached) equals autograd through
    the unroll: the smooth model with a frozen reset and the spiking model with the surrogate (the gradient
    the kernel must produce); with the reset term added it equals the smooth model with a live reset."""
    d, net = _subsample()
    T = 150
    counts = ao.poisson_counts(d["thr"], seed=net.seed, instance=0, step0=0, n_steps=T)
    rng = np.random.default_rng(5)
    targets = rng.uniform(0.5, 10.0, 8)
    alpha0 = torch.as_tensor(1.0 + 0.2 * rng.standard_normal((8, 8)), dtype=torch.float64)
    kind = "spike" if mode == "spike" else "smooth"
    frozen = None
    if mode == "smooth-frozen":
        frozen = ao.unroll(d, alpha0, T, counts, mode="smooth", **(sur or {}))[0].numpy()
    alpha = alpha0.clone().requires_grad_(True)
    S, rates, VP, RG = ao.unroll(d, alpha, T, counts, mode=kind, frozen_reset=frozen, **(sur or {}))
    ao.loss(rates, targets).backward()
    if mode == "spike":
        assert float(S.detach().sum()) > 20  # spikes did happen, and refractory steps with them
        assert float((RG == 0).sum()) > 0
    c = ao.loss_grad_per_spike(d, rates.detach().numpy(), targets, T)
    args = (d, alpha0.numpy(), S.detach().numpy(), VP.detach().numpy(), RG.numpy(), c)
    live = mode == "smooth-live"
    G = ao.adjoint(*args, with_reset=live, **(sur or {}))
    ref = alpha.grad.numpy()
    assert np.abs(G - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
    for drop in ("coupling", "delay"):  # each term of the recursion matters
        Gd = ao.adjoint(*args, drop=drop, with_reset=live, **(sur or {}))
        assert np.abs(Gd - ref).max() > 1e-7 * np.abs(ref).max(), drop  # far above the 1e-9 match
    if live:
        assert np.abs(ao.adjoint(*args, with_reset=False, **(sur or {})) - ref).max() > 1e-4 * np.abs(ref).max()
