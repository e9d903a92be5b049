"""PD14 real-time factor and synaptic events/s on one MI355X.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 1000] [--no-cpu]

Workload: the full-scale Potjans & Diesmann (2014) microcircuit of naer2's test/networks/PD14 (77,169 LIF
neurons with exponential current synapses, ~298.8 M synapses by the reference's fixed-total-number rule,
Poisson background, dt = 0.1 ms), built with seed 0 from synthetic random draws.  A step is one fused
kernel launch advancing the whole network by dt; `steps` of them are timed (default 10,000 = 1 s of
biological time) after `warmup` untimed ones, replayed from captured hipGraphs with every input resident
in HBM.

value: real-time factor = simulated (bio) seconds / wall seconds.  Also reported: synaptic events/s
(links delivered per wall second), the roofline of the step kernel (algorithmic bytes per launch over its
average duration: HIP events recorded on the launch stream around the timed region, divided by the
launches in it), and the CPU oracle's rate on this host's
cores over a bounded sample of the same workload, from the same state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def _events_and_spikes(rt, step_lo: int, step_hi: int, out_deg: torch.Tensor) -> tuple[int, int]:
    """(links delivered, spikes) for the spikes of steps [step_lo, step_hi): each spike of source s is
    delivered over its out-degree."""
    words = rt.raster_bits(step_lo, step_hi)  # int64 [steps, G]
    G = rt.n_groups
    gf = rt.layout.group_first.to(torch.int64)
    cnt = (rt.layout.group_info & 0xFFFF).to(torch.int64)
    lane = torch.arange(64, device=words.device)
    valid = lane[None, :] < cnt[:, None]
    nid = (gf[:, None] + lane[None, :]).clamp(max=rt.n - 1)
    deg = torch.where(valid, out_deg[nid], torch.zeros_like(nid))  # [G, 64]
    ev = 0
    sp = 0
    for a in range(0, words.shape[0], 500):
        w = words[a:a + 500]
        bits = ((w[:, :, None] >> lane[None, None, :]) & 1).to(torch.int64)  # [k, G, 64]
        ev += int((bits * deg[None]).sum())
        sp += int(bits.sum())
    return ev, sp


def _lambdas(rt) -> list:
    """Poisson mean of each population's background per step."""
    out = []
    for p in rt.pops:
        inst = p.instance
        out.append(inst.rate_ext * rt.dt if hasattr(inst, "rate_ext") else inst.K_ext * inst.nu_ext * rt.dt)
    return out


def cpu_baselines(net, rt, args, kind: str) -> dict:
    """CPU baselines on the host's cores, from the GPU's state after the timed run:
    the C oracle on one thread and on `threads` (OpenMP over target ranges; both checked bit for bit
    against the GPU's next steps), and the reference's own algorithm (tools/brunel_reference.py's
    plain-torch loop, oracle/cpu_reference.py) on `threads`."""
    from naer2_amd.params import params_bytes
    from oracle import cpu_reference, oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    model = oracle.LIF_EXP if kind == "pd14" else oracle.LIF_DELTA
    c = rt.canon.to("cpu")
    prev = rt.raster(rt.step - 1, rt.step)[0].astype(np.uint8)
    snap = (rt.v.cpu().numpy(), None if rt.isyn is None else rt.isyn.cpu().numpy(), rt.ref.cpu().numpy(),
            rt.ring.cpu().numpy(), prev)
    # bounded samples of a few seconds of CPU work each (scaled by --cpu-steps, 2000 by default)
    f = args.cpu_steps / 2000.0
    c0 = rt.step
    n1 = int(2000 * f) if kind == "pd14" else int(50000 * f)        # one thread
    n_mt = int(20000 * f) if kind == "pd14" else int(100000 * f)    # every core
    net.run(max(n1, n_mt), graph_window=0)
    torch.cuda.synchronize()
    gpu_raster = rt.raster(c0, c0 + max(n1, n_mt))
    gpu_v = rt.v.cpu().numpy()
    pb = params_bytes(rt.params)
    arrays = (c.src_ptr.numpy(), c.tgt.numpy(), c.w.numpy(), c.d.numpy())
    bio_step = rt.dt / 1000.0

    def state():
        v, i, r, ring, pv = snap
        return oracle.State(rt.n, rt.d_slots, v.copy(), None if i is None else i.copy(), r.copy(), ring.copy(),
                            pv.copy())

    st = state()
    t0 = time.perf_counter()
    r1, ev1 = oracle.run(model, rt.pop_first, pb, rt.d_slots, rt.net.seed, rt.net.instance, *arrays, st, c0, n1, True)
    w1 = time.perf_counter() - t0
    same1 = bool(np.array_equal(gpu_raster[:n1], r1))
    name = "PD14" if kind == "pd14" else "Brunel"
    if any(p.stride > 1 for p in rt.params):  # the threaded oracle and the torch loop are single-rate only
        print(f"[bench] cpu: oracle 1 thread {n1} steps {w1:.1f} s (bit-exact {same1})", file=sys.stderr, flush=True)
        return {"value": round(n1 * bio_step / w1, 6), "unit": "bio_s/wall_s", "cores": 1, "kind": "port",
                "sample": f"{n1} steps ({n1 * bio_step * 1000:.0f} ms bio) of the same multi-rate {name} network from "
                          f"the GPU state after step {c0}: the C oracle (-O2) on one host thread; raster bit-exact "
                          f"vs GPU: {same1}",
                "events_per_s": round(ev1 / w1, 1), "bit_exact_vs_gpu": same1}
    par = oracle.Parallel(rt.n, *arrays, threads=threads)
    st = state()
    t0 = time.perf_counter()
    rm, evm = par.run(model, rt.pop_first, pb, rt.d_slots, rt.net.seed, rt.net.instance, st, c0, n_mt, True)
    wm = time.perf_counter() - t0
    par.close()
    samem = bool(np.array_equal(gpu_raster[:n_mt], rm)) and (
        n_mt != max(n1, n_mt) or bool(np.array_equal(gpu_v.view(np.uint32), st.v.view(np.uint32))))
    n_ref = int(100 * f) if kind == "pd14" else int(20000 * f)       # the reference's torch loop
    v, i, _, _, _ = snap
    ref = cpu_reference.run(model, rt.pop_first, rt.params, _lambdas(rt), c, rt.d_slots, v, i if i is not None else v,
                            n_ref, threads)
    print(f"[bench] cpu: oracle 1 thread {n1} steps {w1:.1f} s (bit-exact {same1}); oracle {threads} threads "
          f"{n_mt} steps {wm:.1f} s (bit-exact {samem}); reference algorithm (torch) {threads} threads {n_ref} steps "
          f"{ref['seconds']:.1f} s", file=sys.stderr, flush=True)
    return {
        "value": round(n_mt * bio_step / wm, 6), "unit": "bio_s/wall_s", "cores": threads, "kind": "port",
        "sample": f"{n_mt} steps ({n_mt * bio_step * 1000:.0f} ms bio) of the same {name} network from the GPU state "
                  f"after step {c0}: the C oracle (-O2) on {threads} host threads (OpenMP, target ranges); "
                  f"raster bit-exact vs GPU: {samem}",
        "events_per_s": round(evm / wm, 1), "bit_exact_vs_gpu": samem and same1,
        "single_thread": {"value": round(n1 * bio_step / w1, 6), "cores": 1, "steps": n1,
                          "events_per_s": round(ev1 / w1, 1), "bit_exact_vs_gpu": same1},
        "reference_algorithm": {"value": round(n_ref * bio_step / ref["seconds"], 6), "cores": threads,
                                "steps": n_ref, "events_per_s": round(ref["events"] / ref["seconds"], 1),
                                "what": "tools/brunel_reference.py's plain-torch time-driven loop (a sparse product "
                                        "per delay class and step), restated for this network on the CPU"},
    }


def bench_sweep(args) -> int:
    """Batched Brunel sweep: --instances networks drawn from --seed (N log-uniform in [1k, 50k], in-degree
    10 %, g in [3, 8], eta in [0.9, 4]), --warmup batched steps, then one batched launch per step (mapping
    from the autotuner: each candidate timed on --warmup steps from the same state) against the same
    networks run one after another from the same state.  The final
    states of both must be bit-identical."""
    from naer2_amd import FusedRound
    from naer2_amd.autotune import measure_work, tune, tune_intake
    from naer2_amd.networks import brunel_sweep, sweep_params

    params = sweep_params(args.instances or 256, seed=args.seed)
    t0 = time.perf_counter()
    nets = []
    for i, p in enumerate(params):
        nets += brunel_sweep([p], device="cuda", keep_canonical=False, record=False)
        if i % 16 == 15:
            print(f"[bench] built {i + 1}/{len(params)} networks ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    rts = [n.runtime for n in nets]
    n_syn = sum(rt.n_links for rt in rts)
    n_neurons = sum(rt.n for rt in rts)
    torch.cuda.synchronize()
    print(f"[bench] sweep: {len(rts)} networks, {n_neurons} neurons, {n_syn} synapses, "
          f"{sum(rt.n_groups for rt in rts)} groups, built in {time.perf_counter() - t0:.0f} s",
          file=sys.stderr, flush=True)
    if args.tune_intake:
        t0 = time.perf_counter()
        for rt in rts:
            tune_intake(rt, steps=200, graph_window=args.window)
        print(f"[bench] intake thresholds tuned in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    fr = FusedRound(rts, {"order": 1, "xcd": 0})  # warmup: past the onset transient, untimed
    fr.launch(args.warmup, args.window)
    fr.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    weights = measure_work(rts, steps=200, graph_window=args.window)
    print(f"[bench] per-network work measured in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    best, times = tune(rts, steps=args.warmup, graph_window=args.window, use_cache=False, weights=weights)
    print(f"[bench] autotune: {best} from {times}", file=sys.stderr, flush=True)
    snaps = [rt.snapshot() for rt in rts]
    fr = FusedRound(rts, best, weights=weights)
    fr.launch(args.window, args.window)  # capture, untimed (the state is restored below)
    torch.cuda.synchronize()
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
    fr.close()
    fr = FusedRound(rts, best, weights=weights)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fr.launch(args.steps, graph_window=args.window)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    fr.close()
    print(f"[bench] batched {args.steps} steps: {wall_b:.3f} s", file=sys.stderr, flush=True)
    final_b = [(rt.v.clone(), rt.ref.clone(), rt.ring.clone()) for rt in rts]
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
        rt.run(args.window, args.window)  # capture each network's own graph, untimed
    torch.cuda.synchronize()
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
    torch.cuda.synchronize()
    wall_s = 0.0
    for i, rt in enumerate(rts):
        t0 = time.perf_counter()
        rt.run(args.steps, args.window)
        torch.cuda.synchronize()
        wall_s += time.perf_counter() - t0
    print(f"[bench] sequential {args.steps} steps x {len(rts)}: {wall_s:.3f} s", file=sys.stderr, flush=True)
    same = all(torch.equal(a[0].view(torch.int32), rt.v.view(torch.int32)) and torch.equal(a[1], rt.ref)
               and torch.equal(a[2].view(torch.int32), rt.ring.view(torch.int32)) for a, rt in zip(final_b, rts))
    bio = args.steps * rts[0].dt / 1000.0
    line = {
        "metric": "batched Brunel sweep: aggregate instance-bio-s per wall-s, batched vs sequential plain runs",
        "value": round(len(rts) * bio / wall_b, 4), "unit": "instance_bio_s/wall_s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(rts)} Brunel networks (N log-uniform 1k-50k, in-degree 10 %, g U[3,8], "
                               f"eta U[0.9,4]), seed {args.seed}", "n_neurons": n_neurons, "n_synapses": n_syn},
        "sequential_value": round(len(rts) * bio / wall_s, 4),
        "batched_speedup": round(wall_s / wall_b, 3),
        "mapping": best, "autotune_s": times,
        "bit_exact_batched_vs_sequential": bool(same),
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_pd14grad(args) -> int:
    """Batched, differentiable PD14: --instances seed-distinct full-scale microcircuits in one launch per
    step, --steps steps (500 ms bio at 5000), loss = squared error of the per-population rates against
    seed-drawn targets in [0.5, 10] Hz, gradient with respect to the 8x8 weight scales from the surrogate
    adjoint (truncated at --cutoff, scaled by --gamma) with checkpoints every --checkpoint steps.  Times the
    forward (checkpointed, graph windows) and
    the backward (recompute + adjoint) with HIP events, and checks the same adjoint on a
    4-neuron-per-population PD14 subsample against the FP64 autograd unroll (on the same spike trains),
    whose own gradient FP64 central differences pin in the CPU tests."""
    import numpy as np

    from naer2_amd.adjoint import RateLossGrad
    from naer2_amd.networks import pd14

    n_inst, T, K = args.instances or 4, args.steps, args.checkpoint
    rng = np.random.default_rng(args.seed)
    alpha = np.ones((8, 8), dtype=np.float32)
    targets = rng.uniform(0.5, 10.0, (n_inst, 8))
    t0 = time.perf_counter()
    rts = []
    for i in range(n_inst):
        net, _ = pd14(seed=args.seed + i, instance=i, device="cuda", weight_scale=alpha)
        net.record = True
        net.keep_canonical = True
        rts.append(net.build())
    torch.cuda.synchronize()
    n_syn = sum(rt.n_links for rt in rts)
    print(f"[bench] built {n_inst} PD14 instances ({n_syn} synapses) in {time.perf_counter() - t0:.0f} s",
          file=sys.stderr, flush=True)
    sur = {"cutoff": args.cutoff, "gamma": args.gamma}
    grad = RateLossGrad(rts, alpha, checkpoint=K, long_delay=args.long_delay, long_parts=args.long_parts, **sur)
    t1 = time.perf_counter()
    loss, g = grad(T, targets)
    wall = time.perf_counter() - t1
    fw, bw = grad.last["forward_ms"], grad.last["backward_ms"]
    print(f"[bench] loss {loss:.4g}, forward {fw:.0f} ms, backward {bw:.0f} ms, wall {wall:.1f} s",
          file=sys.stderr, flush=True)
    # the check: the same adjoint on a 4-neuron-per-population subsample against the FP64 unroll
    from oracle import adjoint_oracle as ao
    Tc = 500
    sub, _ = pd14(sizes=[4] * 8, prob_scale=10.0, seed=args.seed, device="cuda", weight_scale=alpha)
    sub.record, sub.keep_canonical = True, True
    srt = sub.build()
    v0, i0 = srt.v.cpu().numpy().copy(), srt.isyn.cpu().numpy().copy()
    ts = rng.uniform(0.5, 10.0, (1, 8))
    _, gs = RateLossGrad([srt], alpha, checkpoint=100, **sur)(Tc, ts)
    d = ao.net_from(srt.pop_first, srt.params, srt.canon.to("cpu"), v0, i0, srt.dt)
    counts = ao.poisson_counts(d["thr"], seed=sub.seed, instance=sub.instance, step0=0, n_steps=Tc)
    a64 = torch.ones(8, 8, dtype=torch.float64, requires_grad=True)
    _, rates, _, _ = ao.unroll(d, a64, Tc, counts, mode="spike", forced=srt.raster(0, Tc), **sur)
    ao.loss(rates, ts[0]).backward()
    ref = a64.grad.numpy()
    err = float(np.linalg.norm(gs - ref) / max(np.linalg.norm(ref), 1e-300))
    bio = T * rts[0].dt / 1000.0
    line = {
        "metric": "batched differentiable PD14: backward/forward cost ratio (rate-loss surrogate adjoint)",
        "value": round(bw / fw, 3), "unit": "ratio", "n_gpus": 1, "steps": T, "warmup": 0,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n_inst} seed-distinct full-scale PD14 microcircuits (77169 neurons each) in one launch a step, {T} steps "
                               f"of 0.1 ms, loss sum (rate - target)^2 over populations, targets U[0.5, 10] Hz, "
                               f"grad wrt the 8x8 weight scales, checkpoints every {K} steps, seed {args.seed}",
                   "n_synapses": n_syn},
        "forward_ms": round(fw, 1), "backward_ms": round(bw, 1),
        "forward_rtf_aggregate": round(n_inst * bio / (fw / 1000.0), 3),
        "loss": loss, "grad_norm": float(np.linalg.norm(g)), "grad_finite": bool(np.isfinite(g).all()),
        "surrogate": {"kappa": 25.0, **sur}, "long_delay": args.long_delay, "long_parts": args.long_parts,
        "grad_rel_err_vs_fp64_subsample": err,
        "subsample": f"4 neurons per population (32), PD14 probabilities x10, {Tc} steps, same spike trains",
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_plain(args) -> int:
    """The plain path: the full-scale PD14 microcircuit (the headline) or the 2500-neuron Brunel network.  Warmup, then exactly
    --steps steps in graph windows, timed with HIP events on the launch stream; the CPU baselines on the
    same network from the GPU's state."""
    from naer2_amd.networks import brunel, pd14

    kind = args.workload
    t_build = time.perf_counter()
    if kind == "pd14":
        net, pops = pd14(scale=1.0, seed=args.seed, device="cuda")
    else:
        net, pops = brunel(seed=args.seed, device="cuda")
    timesteps = {}
    for spec in args.timestep or []:
        name, _, ms = spec.partition("=")
        if name not in pops:
            raise SystemExit(f"--timestep: no population {name!r} (have {sorted(pops)})")
        pops[name].timestep = timesteps[name] = float(ms)
    rt = net.build()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    print(f"[bench] built {kind}: {rt.n} neurons, {rt.canon.n_links} synapses, {rt.n_groups} groups, "
          f"{rt.d_slots} ring slots in {t_build:.1f} s", file=sys.stderr, flush=True)
    net.record = True
    rt._ensure_slots(args.warmup + args.steps + 400)

    net.run(args.warmup, graph_window=args.window)
    torch.cuda.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    s0 = rt.step
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    net.run(args.steps, graph_window=args.window)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    s1 = rt.step
    print(f"[bench] timed {args.steps} steps: {wall:.3f} s wall, {gpu_ms:.1f} ms on the stream",
          file=sys.stderr, flush=True)

    out_deg = rt.canon.out_degree()
    events, spikes_timed = _events_and_spikes(rt, s0 - 1, s1 - 1, out_deg)
    bio_s = args.steps * rt.dt / 1000.0
    rtf = bio_s / wall
    ms_per_step = wall * 1000.0 / args.steps

    # roofline of the step kernel: algorithmic bytes per launch (sum over populations of n (8 n_state / k + 8)
    # + 8 E + 8 S; n_state 3 for the exponential synapse: V, I, refractory counter; 2 for the delta synapse;
    # k the population's steps per update, whose state is read and written once per k steps while its ring
    # slots are read and cleared once each)
    n_state = 3 if kind == "pd14" else 2
    strides = [p.stride if p.stride > 1 else 1 for p in rt.params]
    state_bytes = sum(pop.n * (8 * n_state / k + 8) for pop, k in zip(rt.pops, strides))
    bytes_per_step = state_bytes + 8 * events / args.steps + 8 * spikes_timed / args.steps
    k_ms = gpu_ms / args.steps  # HIP events on the launch stream over the timed region, per launch
    achieved = bytes_per_step / (k_ms * 1e-3) / 1e9
    traffic, occupancy = None, None
    tfile = ROOT / "profiles" / "pd14_step_traffic.json"
    if kind == "pd14" and tfile.exists() and not timesteps:
        prof = json.loads(tfile.read_text())
        traffic, occupancy = prof.get("hbm_bytes_per_launch"), prof.get("occupancy")

    rates = net.rates(t_skip_ms=(s0 * rt.dt))
    rates_hz = {name: round(rates[p], 3) for name, p in pops.items()}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(net, rt, args, kind)
    line = {
        "metric": "PD14 real-time factor + synaptic events/s" if kind == "pd14" else
                  "Brunel real-time factor + synaptic events/s",
        "value": round(rtf, 4),
        "unit": "bio_s/wall_s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("pd14 full scale: 77169 LIF-exp neurons, 8 populations, "
                                f"{rt.canon.n_links} synapses, Poisson background" if kind == "pd14" else
                                "brunel: N_E 2000, N_I 500 delta-synapse LIF, C_E 200, C_I 50, J 0.2 mV, "
                                f"g 5, D 1.5 ms, {rt.canon.n_links} synapses, Poisson drive (eta 2)") +
                               f", dt 0.1 ms, {args.steps} steps timed, seed {args.seed}",
                   "n_neurons": rt.n, "n_synapses": rt.canon.n_links, "dt_ms": rt.dt,
                   **({"timesteps_ms": timesteps} if timesteps else {}),
                   "graph_window": args.window, "parallelism": "single GPU"},
        "events_per_s": round(events / wall, 1),
        "spikes_per_step": round(spikes_timed / args.steps, 3),
        "events_per_step": round(events / args.steps, 1),
        "rates_hz": rates_hz,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel_avg_us": round(k_ms * 1000.0, 3), "occupancy": occupancy,
                     "algorithmic_bytes_per_launch": round(bytes_per_step, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _event_oracle_net(net, rt):
    """The oracle's description of a built event-driven network (host copies of its canonical table)."""
    from oracle import event as ev

    index = {p: i for i, p in enumerate(rt.pops)}
    pops = []
    for p in rt.pops:
        e = ev.EvPop()
        if p.family == "event_lif":
            e.kind = ev.NEURON
            e.tau_inv = float(np.float32(1.0 / p.instance.tau))
            e.theta, e.v_reset, e.t_ref = p.instance.theta, p.instance.V_r, p.instance.t_ref
        else:
            e.kind = ev.POISSON
            e.rate_inv = float(np.float32(1.0 / p.instance.rate))
        pops.append(e)
    conns = net.connections
    return ev.EvNet(rt.pop_first.astype(np.int32), pops, np.array([index[c.source] for c in conns], np.int32),
                    np.array([index[c.target] for c in conns], np.int32), rt.window,
                    rt.canon.src_ptr.cpu().numpy().astype(np.int64), rt.canon.tgt.cpu().numpy(),
                    rt.canon.w.cpu().numpy(), seed=int(net.seed), instance=int(net.instance))


def bench_brunel_exact(args) -> int:
    """The event-driven, exact-time Brunel model A (the reference's own Brunel network, test/networks/Brunel):
    N = N_E + N_I exact-time delta LIF neurons and N_E Poisson sources, fixed in-degree C_E = 0.1 N_E, one delay
    D = 1.5 ms.  A step is one window of D (a deliver and a sort launch); --ev-warmup windows, then
    --ev-windows timed in captured windows with recording off; a split run with HIP events around each kernel
    for the deliver kernel's roofline; the CPU oracle from the GPU's state after warmup."""
    from naer2_amd.networks import brunel_exact, brunel_exact_theory

    N = args.neurons or 50000
    t_build = time.perf_counter()
    net, pops = brunel_exact(N, seed=args.seed, device="cuda")
    net.record = False
    net.keep_canonical = not args.no_cpu
    rt = net.build()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    W = rt.window
    print(f"[bench] built brunel_exact: {rt.n} nodes ({N} neurons), {rt.n_links} synapses, {rt.n_groups} groups "
          f"in {t_build:.1f} s", file=sys.stderr, flush=True)
    gw = args.ev_graph
    warm, n_win = args.ev_warmup, args.ev_windows
    rt.run(warm - 1, gw)
    net.record = True
    rt.run(1, 0)  # the last warmup window recorded: the CPU oracle starts from it
    net.record = False
    torch.cuda.synchronize()
    prev_t, prev_n = rt.spikes()
    snap = (rt.v.clone(), rt.last.clone(), rt.tre.clone(), rt.nxt.clone())
    print(f"[bench] warmup {warm} windows done", file=sys.stderr, flush=True)
    c0 = rt.counters()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    rt.run(n_win, gw)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    c1 = rt.counters()
    rt.check()
    events, spikes = c1[0] - c0[0], c1[1] - c0[1]
    bio_s = n_win * W / 1000.0
    rtf = bio_s / wall
    print(f"[bench] timed {n_win} windows: {wall:.3f} s wall, {gpu_ms:.1f} ms on the stream", file=sys.stderr,
          flush=True)
    # the deliver kernel alone: HIP events around each kernel of a split run
    n_split = min(100, n_win)
    c2 = rt.counters()
    d_ms, s_ms = rt.run_timed(n_split)
    c3 = rt.counters()
    ev_split, sp_split = c3[0] - c2[0], c3[1] - c2[1]
    n_nrn = sum(p.n for p in pops.values() if p.family == "event_lif")
    n_src = sum(p.n for p in pops.values() if p.family == "event_poisson")
    g_nrn = sum((p.n + 63) // 64 for p in pops.values() if p.family == "event_lif")
    # algorithmic bytes of one window's deliver launch: 8 B per link walked, 24 B per (input spike, neuron group)
    # (the stream entry and the two segment offsets), 24 B of neuron state read and written, 8 B per source's
    # next spike, 8 B per spike emitted into a region
    bytes_win = (8.0 * ev_split + 24.0 * g_nrn * sp_split + 8.0 * sp_split) / n_split + 24.0 * n_nrn + 8.0 * n_src
    k_ms = d_ms / n_split
    achieved = bytes_win / (k_ms * 1e-3) / 1e9
    traffic = occupancy = None
    tfile = ROOT / "profiles" / "brunel_exact_traffic.json"
    if tfile.exists() and N == 50000:
        prof = json.loads(tfile.read_text())
        traffic, occupancy = prof.get("hbm_bytes_per_launch"), prof.get("occupancy")
    cpu = None
    if not args.no_cpu:
        from oracle import event as ev

        evn = _event_oracle_net(net, rt)
        st = ev.EvState(snap[0].cpu().numpy(), snap[1].cpu().numpy(), snap[2].cpu().numpy(), snap[3].cpu().numpy(),
                        prev_n.astype(np.int32), prev_t.astype(np.float32), warm)
        k = args.ev_cpu_windows
        tc = time.perf_counter()
        _, _, cnt, ev_cpu = ev.run(evn, st, k, record=False)
        tc = time.perf_counter() - tc
        cpu = {"value": round(k * W / 1000.0 / tc, 6), "unit": "bio_s/wall_s", "cores": 1, "kind": "port",
               "sample": f"{k} windows ({k * W:.1f} ms bio) of the same network from the GPU's state after warmup, "
                         f"serial C oracle (oracle/event_oracle.c), {ev_cpu} inputs, {tc:.2f} s",
               "events_per_s": round(ev_cpu / tc, 1), "threads_available": os.cpu_count()}
    # steady-state rates: 100 ms more, recorded
    t_rec = rt.window_index * W
    net.record = True
    rt.run(67, 0)
    net.record = False
    r = rt.rates(t_skip_ms=t_rec)
    rates = {name: r[p] for name, p in pops.items()}
    ref_rtf = 1.0 / 2.94
    line = {
        "metric": "event-driven Brunel real-time factor + synaptic events/s",
        "value": round(rtf, 4),
        "unit": "bio_s/wall_s",
        "n_gpus": 1,
        "steps": n_win,
        "warmup": warm,
        "ms_per_step": round(wall * 1000.0 / n_win, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(rtf / ref_rtf, 2) if N == 50000 else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"brunel_exact: event-driven Brunel model A with exact spike times, N {N} "
                               f"(N_E {pops['E'].n}, N_I {pops['I'].n}), {pops['P'].n} Poisson sources at "
                               f"{brunel_exact_theory(N)['rate_ext_Hz']:.1f} Hz, C_E {brunel_exact_theory(N)['C_E']}, "
                               f"J 0.1 mV, g 5, D 1.5 ms, {rt.n_links} synapses; a step = one 1.5 ms window; "
                               f"{n_win} windows timed, seed {args.seed}",
                   "n_neurons": N, "n_synapses": rt.n_links, "window_ms": W, "graph_window": gw,
                   "parallelism": "single GPU"},
        "vs_baseline_note": "reference: N = 50,000 at 2.94 ms wall per simulated ms on an RTX 4090, captured "
                            "(test/networks/Brunel/net.py:77-78)",
        "events_per_s": round(events / wall, 1),
        "spikes_per_window": round(spikes / n_win, 1),
        "events_per_window": round(events / n_win, 1),
        "rates_hz": {name: round(float(v), 3) for name, v in rates.items()},
        "kernel_split_ms_per_window": {"deliver": round(d_ms / n_split, 4), "sort": round(s_ms / n_split, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": "ev_deliver_kernel", "kernel_avg_us": round(k_ms * 1000.0, 3),
                     "occupancy": occupancy, "algorithmic_bytes_per_launch": round(bytes_win, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_bptt(args) -> int:
    """Surrogate-gradient BPTT: B=128 Poisson inputs on 700 channels (rates U[0, 100] Hz), T=100 steps of
    1 ms, 1024 recurrent LIF units, 20 leaky readouts, cross-entropy of the max-over-time readout,
    checkpoints every 25 steps.  Times forward and backward with HIP events (median of --steps iterations
    after --warmup), and checks the gradients against the FP64 autograd unroll (same spikes)."""
    from naer2_amd.rlif import RecurrentLIF, RLIFConfig, max_over_time_ce, poisson_input
    B, C, H, O, T = 128, 700, 1024, 20, 100
    g = torch.Generator(device="cuda").manual_seed(args.seed)
    cfg = RLIFConfig(checkpoint=25)
    net = RecurrentLIF(C, H, O, cfg, generator=g)
    x = poisson_input(B, C, T, 1.0, g)
    labels = torch.randint(0, O, (B,), generator=g, device="cuda")
    fw, bw = [], []
    iters = max(1, args.steps)
    for it in range(args.warmup + iters):
        net.zero_grad(set_to_none=True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        r, s = net(x)
        loss = max_over_time_ce(r, labels)
        e1.record()
        loss.backward()
        e2.record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            fw.append(e0.elapsed_time(e1))
            bw.append(e1.elapsed_time(e2))
    fwd_ms, bwd_ms = float(np.median(fw)), float(np.median(bw))
    from oracle import rlif_oracle as ro
    grads = [p.grad.detach() for p in (net.w_in, net.w_rec, net.w_out)]
    w64 = [p.detach().double() for p in (net.w_in, net.w_rec, net.w_out)]
    loss64, g64, _ = ro.grads(x.double(), *w64, labels, alpha=cfg.alpha, beta=cfg.beta, theta=cfg.theta,
                              slope=cfg.slope, gamma=cfg.gamma, mode=0, forced=s.detach().double())
    errs = {n: ro.rel_err(a, b) for n, a, b in zip(("w_in", "w_rec", "w_out"), grads, g64)}
    flops = 2.0 * B * T * (C * H + H * H + H * O)
    line = {
        "metric": "BPTT backward/forward cost ratio (surrogate-gradient recurrent LIF)",
        "value": round(bwd_ms / fwd_ms, 4), "unit": "ratio", "n_gpus": 1, "steps": iters, "warmup": args.warmup,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "B=128, 700 Poisson inputs (0-100 Hz), 1024 recurrent LIF, 20 leaky readouts, "
                               "T=100 x 1 ms, fast-sigmoid slope 25, checkpoints every 25 steps"},
        "forward_ms": round(fwd_ms, 4), "backward_ms": round(bwd_ms, 4),
        "forward_dense_equiv_tflops": round(flops / (fwd_ms * 1e-3) / 1e12, 3),
        "spike_fraction": round(float(s.mean()), 5),
        "grad_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
        "loss": float(loss), "loss_fp64": float(loss64),
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 10000; pd14grad: 5000, i.e. 500 ms of bio time)")
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--window", type=int, default=100, help="steps per captured hipGraph")
    ap.add_argument("--cpu-steps", type=int, default=2000,
                    help="scales the CPU samples (PD14: 2000 steps on one thread, 10x that on every core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timestep", action="append", metavar="POP=MS",
                    help="pd14/brunel: a population's own timestep, a whole multiple of dt (repeatable)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", choices=["pd14", "brunel", "sweep", "bptt", "pd14grad", "brunel_exact"],
                    default="pd14",
                    help="pd14: the headline PD14 real-time factor; brunel: the 2500-neuron Brunel network; sweep: the batched "
                         "Brunel sweep; bptt: the recurrent-LIF BPTT; pd14grad: the batched differentiable PD14; "
                         "brunel_exact: the event-driven Brunel model with exact spike times")
    ap.add_argument("--neurons", type=int, default=None, help="brunel_exact: N = N_E + N_I (default 50,000)")
    ap.add_argument("--ev-windows", type=int, default=667, help="brunel_exact: timed windows of 1.5 ms")
    ap.add_argument("--ev-warmup", type=int, default=67, help="brunel_exact: untimed windows first")
    ap.add_argument("--ev-graph", type=int, default=23, help="brunel_exact: windows per captured hipGraph")
    ap.add_argument("--ev-cpu-windows", type=int, default=10, help="brunel_exact: windows the CPU oracle runs")
    ap.add_argument("--instances", type=int, default=None, help="networks (sweep: 256, pd14grad: 4)")
    ap.add_argument("--checkpoint", type=int, default=500, help="pd14grad: steps between saved states")
    ap.add_argument("--long-delay", type=int, default=4,
                    help="pd14grad: out-links with at least this delay (steps) are gathered once per window of "
                         "that many reverse steps")
    ap.add_argument("--long-parts", type=int, default=24,
                    help="pd14grad: target ranges the long out-links are walked in (a multiple of 8: one launch, "
                         "the XCDs taking ranges in turn; otherwise one launch a range)")
    ap.add_argument("--cutoff", type=float, default=0.2,
                    help="pd14grad: surrogate support |V - theta| / (theta - V_reset) (sparse gradient)")
    ap.add_argument("--gamma", type=float, default=0.01,
                    help="pd14grad: surrogate scale (0.1 and above diverge over 500 ms at full scale)")
    ap.add_argument("--tune-intake", action="store_true",
                    help="sweep: time each network's row-intake thresholds first (naer2_amd.autotune.tune_intake)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run 5 plain device-to-device copies of MiB each (the reference kernel that scales "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes), then exit")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 5000 if args.workload == "pd14grad" else 10000
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the spiking-network step has no multi-GPU path in this build (it would "
              "shard by target groups, with one all-gather of the step's spike lists)", file=sys.stderr)
        return 2
    if args.workload == "bptt":
        return bench_bptt(args)
    if args.workload == "pd14grad":
        return bench_pd14grad(args)
    if args.workload == "brunel_exact":
        return bench_brunel_exact(args)
    if args.window <= 0 or args.window % 4 or args.steps % args.window or args.warmup % args.window:
        print("--window must be a multiple of 4, --steps and --warmup multiples of --window", file=sys.stderr)
        return 2
    if args.workload == "sweep":
        return bench_sweep(args)

    return bench_plain(args)


if __name__ == "__main__":
    sys.exit(main())
