"""The rate-loss adjoint on the GPU (naer2_amd/adjoint.py, n2h_adj_step) against the FP64 autograd unroll
of oracle/adjoint_oracle.py on the same spike trains (the fp32 run's, forced), batched over a
4-neuron-per-population PD14 subsample and a 2 % PD14; checkpoint-interval and batching invariance; the
compensated trace of the step kernel (its V, I and spikes unchanged, its distances to the threshold those of
the FP64 unroll)."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd.adjoint import RateLossGrad
from naer2_amd.networks import pd14
from oracle import adjoint_oracle as ao

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _nets(alpha):
    a, _ = pd14(sizes=[4] * 8, prob_scale=10.0, seed=3, device="cuda", weight_scale=alpha, instance=0)
    b, _ = pd14(scale=0.02, seed=1, device="cuda", weight_scale=alpha, instance=1)
    for n in (a, b):
        n.record = True
        n.keep_canonical = True
    return [a, b]


def _alpha(seed=2):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.1 * rng.standard_normal((8, 8))).astype(np.float32)


@pytest.mark.parametrize("sur", [{}, {"cutoff": 0.5, "gamma": 0.5}])
def test_adjoint_matches_fp64_autograd_on_the_same_spikes(sur):
    T = 200
    alpha = _alpha()
    nets = _nets(alpha)
    rts = [n.build() for n in nets]
    _say(f"case: adjoint vs fp64 autograd, networks of {[rt.n for rt in rts]} neurons, T={T}, checkpoint 64, "
         f"surrogate {sur}")
    init = [(rt.v.cpu().numpy().copy(), rt.isyn.cpu().numpy().copy()) for rt in rts]
    # the trajectory the adjoint differentiates: its compensated distances to the threshold, traced
    traces = []
    for rt in rts:
        sn = rt.snapshot()
        buf = torch.empty(T, rt.n, device="cuda")
        comp = torch.zeros(2, rt.n, device="cuda")
        rt.trace_potentials(buf, 0, comp)
        rt.run(T, 0)
        rt.trace_potentials(None)
        torch.cuda.synchronize()
        v_th = np.repeat([p.v_th for p in rt.params], np.diff(rt.pop_first)).astype(np.float64)
        traces.append(buf.cpu().numpy().astype(np.float64) + v_th[None, :])  # potentials (+inf refractory)
        rt.restore(sn)
    targets = np.random.default_rng(9).uniform(0.5, 10.0, (2, 8))
    grad = RateLossGrad(rts, alpha, checkpoint=64, **sur)
    loss, g = grad(T, targets)
    a64 = torch.as_tensor(alpha.astype(np.float64), dtype=torch.float64).requires_grad_(True)
    total = 0.0
    g_rec = np.zeros((8, 8))
    for i, rt in enumerate(rts):
        d = ao.net_from(rt.pop_first, rt.params, rt.canon.to("cpu"), init[i][0], init[i][1], rt.dt)
        pop = d["pop_of"]
        d["w"] = d["w"] / alpha.astype(np.float64)[pop[d["tgt"]], pop[d["src"]]]  # base weights
        counts = ao.poisson_counts(d["thr"], seed=rt.net.seed, instance=rt.net.instance, step0=0, n_steps=T)
        forced = rt.raster(0, T)
        assert forced.sum() > 0
        _, rates, _, _ = ao.unroll(d, a64, T, counts, mode="spike", forced=forced, **sur)
        total = total + ao.loss(rates, targets[i])
        # the kernel's recursion in float64 along the fp32 trajectory (its traced potentials)
        vp = traces[i]
        rg = np.isfinite(vp).astype(np.float64)
        vp = np.where(rg > 0, vp, 0.0)
        c = ao.loss_grad_per_spike(d, rates.detach().numpy(), targets[i], T)
        g_rec += ao.adjoint(d, alpha.astype(np.float64), forced.astype(np.float64), vp, rg, c, **sur)
    total.backward()
    ref = a64.grad.numpy()
    assert abs(loss - float(total.detach())) <= 1e-9 * max(1.0, abs(float(total.detach())))
    err_rec = np.linalg.norm(g - g_rec) / np.linalg.norm(g_rec)
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    _say(f"loss {loss:.6g}, |grad| {np.linalg.norm(ref):.4g}; rel err vs the fp64 recursion on the fp32 "
         f"trajectory {err_rec:.3g}, vs the fp64 autograd unroll {err:.3g}")
    assert err_rec <= 1e-5   # the kernel's arithmetic
    assert err <= 1e-4       # against the FP64 unroll's own potentials
    assert np.abs(ref).max() > 0


def test_gradient_is_independent_of_checkpoints_and_batching():
    T = 160
    alpha = _alpha(4)
    targets = np.random.default_rng(1).uniform(0.5, 10.0, (2, 8))
    _say(f"case: checkpoint 32 vs 160, recompute overlapped or serial, batched vs alone, T={T}")
    res = []
    for K, overlap in ((32, True), (160, True), (32, False)):
        rts = [n.build() for n in _nets(alpha)]
        res.append(RateLossGrad(rts, alpha, checkpoint=K, overlap=overlap)(T, targets))
    # checkpoint intervals, and the recompute overlapped on its own stream or not, change nothing
    for r in res[1:]:
        assert r[0] == res[0][0] and np.array_equal(r[1], res[0][1])
    alone = []
    for i in range(2):
        rt = _nets(alpha)[i].build()
        alone.append(RateLossGrad([rt], alpha, checkpoint=64)(T, targets[i:i + 1]))
    g_alone = np.zeros((8, 8), dtype=np.float32)  # networks added in fp32 in batch order, as the batch does
    for _, g in alone:
        g_alone = g_alone + g.astype(np.float32)
    assert np.array_equal(g_alone.astype(np.float64), res[0][1])


def test_gradient_agrees_over_window_lengths():
    """The long-link windows and target ranges only regroup each source's sum (short links + long links, the
    long ones range by range): every window length, range count and range layout (equal ranges across
    populations, walked with 4 population bins a lane at every window length from 2 to 8, or
    population-pure ones walked eight tasks a wavefront), with checkpoints that are not a multiple of the
    window, gives the single-step gradient to fp32 rounding."""
    T = 120
    alpha = _alpha(5)
    targets = np.random.default_rng(2).uniform(0.5, 10.0, (2, 8))
    sur = {"cutoff": 0.5, "gamma": 0.5}
    res = {}
    cases = ((1, 1, False), (3, 5, False), (4, 8, False), (8, 16, True), (4, 1, False), (4, 24, False), (4, 24, True),
             (4, 32, True), (2, 8, False), (3, 8, False), (8, 16, False))
    for W, P, pure in cases:
        _say(f"case: long_delay {W}, long_parts {P}, pure ranges {pure}, checkpoint 50, T={T}")
        rts = [n.build() for n in _nets(alpha)]
        res[W, P, pure] = RateLossGrad(rts, alpha, checkpoint=50, long_delay=W, long_parts=P, pure_ranges=pure,
                                       **sur)(T, targets)
    g1 = res[1, 1, False][1]
    assert np.abs(g1).max() > 0
    for key in cases[1:]:
        assert res[key][0] == res[1, 1, False][0]
        err = np.linalg.norm(res[key][1] - g1) / np.linalg.norm(g1)
        _say(f"long_delay, long_parts {key}: rel diff {err:.3g}")
        assert err <= 1e-5


def test_ranges_spanning_many_populations_agree():
    """A target range whose targets span more populations than the long kernel keeps bins for (4) is walked
    in several sweeps, each summing 4 of its populations: with six two-neuron populations between
    populations of 90 and 110, range 3 of 8 holds targets of all 8 populations (two sweeps) and range 7 of 16
    of populations 2-7.  The gradient, the sums of the second sweeps included, is the single-step one to
    fp32 rounding."""
    T = 120
    alpha = _alpha(6)
    sizes = [90, 2, 2, 2, 2, 2, 2, 110]
    first = np.concatenate([[0], np.cumsum(sizes)])
    n = int(first[-1])
    span = [int(np.searchsorted(first, t, side="right")) - 1
            for t in (-(-3 * n // 8), -(-4 * n // 8) - 1)]  # range 3: its first and last target's population
    assert span == [0, 7]
    targets = np.random.default_rng(3).uniform(0.5, 10.0, (1, 8))
    sur = {"cutoff": 0.5, "gamma": 0.5}
    res = {}
    for W, P in ((1, 1), (4, 8), (4, 16)):
        _say(f"case: populations {sizes}, long_delay {W}, long_parts {P}, checkpoint 50, T={T}")
        net, _ = pd14(sizes=sizes, prob_scale=10.0, seed=5, device="cuda", weight_scale=alpha, instance=0)
        net.record = True
        res[W, P] = RateLossGrad([net.build()], alpha, checkpoint=50, long_delay=W, long_parts=P, **sur)(T, targets)
    g1 = res[1, 1][1]
    assert np.abs(g1[4:]).max() > 0  # links into populations 4-7 carry gradient
    for key in ((4, 8), (4, 16)):
        assert res[key][0] == res[1, 1][0]
        err = np.linalg.norm(res[key][1] - g1) / np.linalg.norm(g1)
        _say(f"long_delay, long_parts {key}: rel diff {err:.3g}")
        assert err <= 1e-5


def test_compensated_trace_changes_nothing_and_tracks_fp64_potentials():
    """With the compensation state bound, the step's potentials, currents, refractory counters, ring and
    spikes stay bit-identical to the plain run, and the traced distances to the threshold match the FP64
    unroll's (same spikes) far closer than the fp32 potentials do (their remaining difference is the fp32
    rounding of the inputs, which the unroll sums in FP64).  At kappa = 25 the surrogate's relative error is at
    most 50 times the error over V_th - V_reset, so 3e-7 keeps it under 1.5e-5."""
    T = 300
    alpha = _alpha(7)
    net_a, net_b = _nets(alpha)[1], _nets(alpha)[1]
    ra, rb = net_a.build(), net_b.build()
    _say(f"case: compensated trace vs plain run, {ra.n} neurons, T={T}")
    init = (ra.v.cpu().numpy().copy(), ra.isyn.cpu().numpy().copy())
    comp = torch.zeros(2, ra.n, device="cuda")
    buf = torch.empty(T, ra.n, device="cuda")
    plain = torch.empty(T, rb.n, device="cuda")
    ra.trace_potentials(buf, 0, comp)
    rb.trace_potentials(plain, 0)
    ra.run(T, 0)
    rb.run(T, 0)
    torch.cuda.synchronize()
    assert np.array_equal(ra.raster(), rb.raster()) and ra.raster().sum() > 0
    for x, y in ((ra.v, rb.v), (ra.isyn, rb.isyn), (ra.ring, rb.ring)):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))
    assert torch.equal(ra.ref, rb.ref)
    d = ao.net_from(ra.pop_first, ra.params, ra.canon.to("cpu"), init[0], init[1], ra.dt)
    pop = d["pop_of"]
    d["w"] = d["w"] / alpha.astype(np.float64)[pop[d["tgt"]], pop[d["src"]]]
    counts = ao.poisson_counts(d["thr"], seed=ra.net.seed, instance=ra.net.instance, step0=0, n_steps=T)
    a64 = torch.as_tensor(alpha.astype(np.float64))
    _, _, vpre, rg = ao.unroll(d, a64, T, counts, mode="spike", forced=ra.raster(0, T))
    v_th = np.asarray(d["v_th"])[pop]
    ref = vpre.numpy() - v_th[None, :]
    live = rg.numpy() > 0
    comp_t, plain_t = buf.cpu().numpy().astype(np.float64), plain.cpu().numpy().astype(np.float64)
    assert np.array_equal(np.isfinite(comp_t), live) and np.array_equal(np.isfinite(plain_t), live)
    e_comp = np.abs(comp_t - ref)[live].max() / 15.0
    e_plain = np.abs(plain_t - ref)[live].max() / 15.0
    _say(f"max |error| / (V_th - V_reset): compensated {e_comp:.3g}, fp32 {e_plain:.3g}")
    assert e_comp <= 3e-7 and e_plain > 3 * e_comp


@pytest.mark.parametrize("kind", ["uneven3", "ten"])
def test_batches_of_many_and_uneven_networks_equal_their_networks_alone(kind):
    """The reverse step deals each network's gatherers to XCDs of its own (a contiguous share of the 8 with at
    most 8 networks, XCD q mod 8 beyond) and sizes its work lists for the XCD dealt the most: three networks
    of uneven sizes (2 % PD14, the 32-neuron subsample, 5 % PD14: XCD shares of 2, 3 and 3) and ten
    subsamples (two networks on XCDs 0 and 1).  The batch's gradient equals the networks' own, added in
    batch order, bit for bit."""
    T = 120
    alpha = _alpha(7)

    def nets():
        if kind == "uneven3":
            made = [pd14(scale=0.02, seed=1, device="cuda", weight_scale=alpha, instance=0)[0],
                    pd14(sizes=[4] * 8, prob_scale=10.0, seed=3, device="cuda", weight_scale=alpha, instance=1)[0],
                    pd14(scale=0.05, seed=2, device="cuda", weight_scale=alpha, instance=2)[0]]
        else:
            made = [pd14(sizes=[4] * 8, prob_scale=10.0, seed=10 + i, device="cuda", weight_scale=alpha, instance=i)[0]
                    for i in range(10)]
        for n in made:
            n.record = True
            n.keep_canonical = True
        return made

    first = nets()
    targets = np.random.default_rng(4).uniform(0.5, 10.0, (len(first), 8))
    rts = [n.build() for n in first]
    _say(f"case: {kind}, batch of {len(rts)} networks, sizes {[rt.n for rt in rts]}, T={T}")
    loss, g = RateLossGrad(rts, alpha, checkpoint=40)(T, targets)
    g_alone = np.zeros((8, 8), dtype=np.float32)
    loss_alone = 0.0
    for i, n in enumerate(nets()):
        li, gi = RateLossGrad([n.build()], alpha, checkpoint=40)(T, targets[i:i + 1])
        g_alone = g_alone + gi.astype(np.float32)
        loss_alone += li
    assert np.abs(g).max() > 0
    assert np.array_equal(g_alone.astype(np.float64), g)
    assert abs(loss - loss_alone) <= 1e-9 * max(1.0, abs(loss))
