"""PD14 real-time factor and synaptic events/s on one MI355X, with the batched, BPTT and differentiable
configurations measured beside it on the same GPU.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 1000] [--no-cpu] [--no-secondary]

Headline workload: the full-scale Potjans & Diesmann (2014) microcircuit of naer2's test/networks/PD14
(77,169 LIF neurons with exponential current synapses, ~298.8 M synapses by the reference's
fixed-total-number rule, Poisson background, dt = 0.1 ms), built with seed 0 from synthetic random draws.
A step is one fused kernel launch advancing the whole network by dt.  `warmup` untimed steps run first;
then every hipGraph the timed region replays is captured (untimed), and exactly `steps` steps are timed
between two device synchronizations, with every input resident in HBM.

value: real-time factor = simulated (bio) seconds / wall seconds over the timed steps.  Beside it:
- synaptic events/s (links delivered per wall second) and the step kernel's roofline (algorithmic bytes per
  launch over its average duration: HIP events on the launch stream around the timed region, divided by
  the launches in it);
- `full_config`: the configuration's own run, 1 s of bio time (10,000 steps) after 1,000 settling steps;
- `cpu_baseline`: the CPU oracle on this host's usable cores over a bounded sample of the same network from
  the GPU's state, checked bit for bit against the GPU's next steps;
- `configs`: the other configurations at their full sizes -- the plain Brunel network, the batched Brunel
  sweep against 256 sequential plain runs, the BPTT cost ratio and gradient error, the batched
  differentiable PD14's cost ratio and gradient error -- each skipped, with the reason, when the run's time
  budget (--budget-s) would not hold it.
`--workload` runs one of them alone and prints its own line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def _process_start() -> float:
    """perf_counter() value at this process's start (the interpreter's and torch's imports included)."""
    try:
        import psutil
        return time.perf_counter() - max(0.0, time.time() - psutil.Process().create_time())
    except Exception:
        return time.perf_counter()


T_START = _process_start()


def log(msg: str) -> None:
    print(f"[bench {time.perf_counter() - T_START:6.1f} s] {msg}", file=sys.stderr, flush=True)


# ----------------------------------------------------------------------------------------------- host facts

def host_cpu() -> dict:
    """The host's CPU model, logical CPUs, the CPUs this process may run on and its cgroup CPU quota."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
        except (OSError, StopIteration):
            model = "unknown"
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            q, p = open(path).read().split()[:2]
            if q != "max":
                quota = float(q) / float(p)
        except (OSError, ValueError):
            pass
    usable = affinity if quota is None else max(1, min(affinity, int(math.floor(quota + 1e-9))))
    return {"model_name": model, "nproc": nproc, "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
            "usable_cpus": usable}


# ----------------------------------------------------------------------------------------------- plain path

def _events_and_spikes(rt, step_lo: int, step_hi: int, out_deg: torch.Tensor) -> tuple[int, int]:
    """(links delivered, spikes) for the spikes of steps [step_lo, step_hi): each spike of source s is
    delivered over its out-degree."""
    words = rt.raster_bits(step_lo, step_hi)  # int64 [steps, G]
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


def timed_run(net, rt, n_steps: int, window: int, out_deg: torch.Tensor, kind: str) -> dict:
    """Exactly n_steps steps from the current state, every graph captured beforehand: wall time between two
    device synchronizations, HIP events on the launch stream, delivered links and spikes, the roofline."""
    # the runtime's own schedule (Network.run): graph windows of `window` steps, eager steps before the first
    # window's phase and after the last; a run shorter than a window is all eager (a graph launch costs more
    # host time than the few launches it replaces); window 0: eager launches only
    gw = window
    rt.prepare(n_steps, gw)
    s0 = rt.step
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    net.run(n_steps, graph_window=gw)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    events, spikes = _events_and_spikes(rt, s0 - 1, rt.step - 1, out_deg)
    bio_s = n_steps * rt.dt / 1000.0
    # algorithmic bytes of one launch: per population n (8 n_state / k + 8) + 8 E + 8 S; n_state 3 for the
    # exponential synapse (V, I, refractory counter), 2 for the delta synapse; k the population's steps per
    # update (its state is read and written once per k steps, its ring slots read and cleared every step)
    n_state = 3 if kind == "pd14" else 2
    strides = [p.stride if p.stride > 1 else 1 for p in rt.params]
    state_bytes = sum(pop.n * (8 * n_state / k + 8) for pop, k in zip(rt.pops, strides))
    bytes_per_step = state_bytes + 8 * events / n_steps + 8 * spikes / n_steps
    k_ms = gpu_ms / n_steps
    achieved = bytes_per_step / (k_ms * 1e-3) / 1e9
    return {"steps": n_steps, "from_step": s0, "graph_window": gw, "wall_s": wall, "gpu_ms": gpu_ms,
            "rtf": bio_s / wall, "ms_per_step": wall * 1000.0 / n_steps, "events": events, "spikes": spikes,
            "events_per_s": events / wall, "kernel_avg_us": k_ms * 1000.0, "bytes_per_step": bytes_per_step,
            "achieved_gbs": achieved}


def _roofline(m: dict, traffic, occupancy, kernel: str) -> dict:
    return {"bound": "hbm", "achieved": round(m["achieved_gbs"], 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(m["achieved_gbs"] / HBM_PEAK_GBS, 6), "traffic": traffic, "kernel": kernel,
            "kernel_avg_us": round(m["kernel_avg_us"], 3), "occupancy": occupancy,
            "algorithmic_bytes_per_launch": round(m["bytes_per_step"], 1),
            "region": f"steps {m['from_step']}..{m['from_step'] + m['steps'] - 1}"}


def cpu_baselines(net, rt, args, kind: str, cpu: dict) -> dict:
    """CPU baselines on the host, from the GPU's state: the C oracle (built here with -O3 -march=native) on
    one thread and on every usable CPU (OpenMP over target ranges; both checked bit for bit against the
    GPU's next steps), and the reference's own algorithm (tools/brunel_reference.py's plain-torch loop,
    oracle/cpu_reference.py) on every usable CPU."""
    from naer2_amd.params import params_bytes
    from oracle import cpu_reference, oracle
    native = True
    try:
        oracle.use_native()
    except (OSError, subprocess.CalledProcessError, RuntimeError) as e:  # no compiler: the portable -O2 build
        native = False
        log(f"cpu: native oracle build failed ({e}); using the portable build")
    build = "-O3 -march=native" if native else "-O2 (portable)"
    threads = cpu["usable_cpus"]
    model = oracle.LIF_EXP if kind == "pd14" else oracle.LIF_DELTA
    c = rt.canon.to("cpu")
    prev = rt.raster(rt.step - 1, rt.step)[0].astype(np.uint8)
    snap = (rt.v.cpu().numpy(), None if rt.isyn is None else rt.isyn.cpu().numpy(), rt.ref.cpu().numpy(),
            rt.ring.cpu().numpy(), prev)
    # bounded samples of a few seconds of CPU work each (scaled by --cpu-steps)
    f = args.cpu_steps / 2000.0
    c0 = rt.step
    n1 = max(4, int(2000 * f)) if kind == "pd14" else max(4, int(50000 * f))                  # one thread
    n_mt = max(4, int(1000 * f * threads)) if kind == "pd14" else max(4, int(100000 * f))       # every CPU
    n_gpu = max(n1, n_mt)
    rt._ensure_slots(n_gpu)
    net.run(n_gpu, graph_window=0)
    torch.cuda.synchronize()
    gpu_raster = rt.raster(c0, c0 + n_gpu)
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
    host = {k: cpu[k] for k in ("model_name", "nproc", "affinity_cpus", "cgroup_cpu_quota")}
    if any(p.stride > 1 for p in rt.params):  # the threaded oracle and the torch loop are single-rate only
        log(f"cpu: oracle 1 thread {n1} steps {w1:.1f} s (bit-exact {same1})")
        return {"value": round(n1 * bio_step / w1, 6), "unit": "bio_s/wall_s", "cores": 1, "kind": "port",
                "sample": f"{n1} steps ({n1 * bio_step * 1000:.0f} ms bio) of the same multi-rate {name} network from "
                          f"the GPU state after step {c0}: the C oracle ({build}) on one host thread; raster "
                          f"bit-exact vs GPU: {same1}",
                "events_per_s": round(ev1 / w1, 1), "bit_exact_vs_gpu": same1, "host": host}
    par = oracle.Parallel(rt.n, *arrays, threads=threads)
    st = state()
    t0 = time.perf_counter()
    rm, evm = par.run(model, rt.pop_first, pb, rt.d_slots, rt.net.seed, rt.net.instance, st, c0, n_mt, True)
    wm = time.perf_counter() - t0
    par.close()
    samem = bool(np.array_equal(gpu_raster[:n_mt], rm)) and (
        n_mt != n_gpu or bool(np.array_equal(gpu_v.view(np.uint32), st.v.view(np.uint32))))
    n_ref = max(4, int(100 * f)) if kind == "pd14" else max(4, int(20000 * f))  # the reference's torch loop
    v, i, _, _, _ = snap
    ref = cpu_reference.run(model, rt.pop_first, rt.params, _lambdas(rt), c, rt.d_slots, v, i if i is not None else v,
                            n_ref, threads)
    log(f"cpu: oracle 1 thread {n1} steps {w1:.1f} s (bit-exact {same1}); oracle {threads} threads {n_mt} steps "
        f"{wm:.1f} s (bit-exact {samem}); reference algorithm (torch) {threads} threads {n_ref} steps "
        f"{ref['seconds']:.1f} s")
    return {
        "value": round(n_mt * bio_step / wm, 6), "unit": "bio_s/wall_s", "cores": threads, "kind": "port",
        "sample": f"{n_mt} steps ({n_mt * bio_step * 1000:.0f} ms bio) of the same {name} network from the GPU state "
                  f"after step {c0}: the C oracle ({build}, built on this host) on {threads} host threads (OpenMP, "
                  f"target ranges; every CPU this process may use); raster bit-exact vs GPU: {samem}",
        "events_per_s": round(evm / wm, 1), "bit_exact_vs_gpu": samem and same1, "host": host,
        "single_thread": {"value": round(n1 * bio_step / w1, 6), "cores": 1, "steps": n1,
                          "events_per_s": round(ev1 / w1, 1), "bit_exact_vs_gpu": same1},
        "reference_algorithm": {"value": round(n_ref * bio_step / ref["seconds"], 6), "cores": threads,
                                "steps": n_ref, "events_per_s": round(ref["events"] / ref["seconds"], 1),
                                "what": "tools/brunel_reference.py's plain-torch time-driven loop (a sparse product "
                                        "per delay class and step), restated for this network on the CPU"},
    }


def measure_plain(args, kind: str, steps: int, warmup: int, cpu_leg: bool, cpu: dict) -> dict:
    """The plain path: the full-scale PD14 microcircuit (the headline) or the 2500-neuron Brunel network.
    `warmup` steps, then exactly `steps` timed steps; then the configuration's own 1 s of bio time after
    1,000 settling steps (when the timed run is not already that); the CPU baselines from the GPU's state."""
    from naer2_amd.networks import brunel, pd14

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
    if args.xcd is None:  # the launch mapping from the autotuner (untimed, the state restored afterwards)
        from naer2_amd.autotune import tune_plan
        mapping, map_times = tune_plan(rt, steps=200, graph_window=args.window)
        log(f"launch mapping (autotuned): {mapping} from {map_times}")
    else:
        rt.set_launch_mapping(args.xcd, args.xcd_run)
        mapping, map_times = {"xcd": args.xcd, "run": args.xcd_run}, None
    log(f"built {kind}: {rt.n} neurons, {rt.canon.n_links} synapses, {rt.n_groups} groups, {rt.d_slots} ring "
        f"slots in {t_build:.1f} s")
    net.record = True
    out_deg = rt.canon.out_degree()

    if warmup:
        rt.prepare(warmup, args.window)
        net.run(warmup, graph_window=args.window)
    torch.cuda.synchronize()
    log(f"warmup {warmup} steps done")
    m = timed_run(net, rt, steps, args.window, out_deg, kind)
    log(f"timed {steps} steps: {m['wall_s'] * 1e3:.3f} ms wall, {m['gpu_ms']:.3f} ms on the stream, RTF {m['rtf']:.3f}")
    rates = net.rates(t_skip_ms=(m["from_step"] * rt.dt))
    timed_rates = {name: round(rates[p], 3) for name, p in pops.items()}

    # the configuration's own run: 1 s of bio time after 1,000 settling steps
    settle, full_n = 1000, int(round(1000.0 / rt.dt))
    if m["from_step"] >= settle and steps >= full_n:
        full, full_rates = m, timed_rates
    elif args.no_full:
        full = None
    else:
        if rt.step < settle:
            net.run(settle - rt.step, graph_window=0)
        full = timed_run(net, rt, full_n, args.window, out_deg, kind)
        rr = net.rates(t_skip_ms=(full["from_step"] * rt.dt))
        full_rates = {name: round(rr[p], 3) for name, p in pops.items()}
        log(f"full config: {full_n} steps from step {full['from_step']}: RTF {full['rtf']:.3f}, "
            f"{full['kernel_avg_us']:.2f} us per launch")

    traffic, occupancy = None, None
    tfile = ROOT / "profiles" / "pd14_step_traffic.json"
    if kind == "pd14" and tfile.exists() and not timesteps:
        prof = json.loads(tfile.read_text())
        traffic, occupancy = prof.get("hbm_bytes_per_launch"), prof.get("occupancy")
    kernel = "step_kernel<2, 32> (1210 groups)" if kind == "pd14" else "step_kernel<1, 8>"

    cpu_line = cpu_baselines(net, rt, args, kind, cpu) if cpu_leg else None
    workload = ("pd14 full scale: 77169 LIF-exp neurons, 8 populations, "
                f"{rt.canon.n_links} synapses, Poisson background" if kind == "pd14" else
                "brunel: N_E 2000, N_I 500 delta-synapse LIF, C_E 200, C_I 50, J 0.2 mV, "
                f"g 5, D 1.5 ms, {rt.canon.n_links} synapses, Poisson drive (eta 2)")
    line = {
        "metric": "PD14 real-time factor + synaptic events/s" if kind == "pd14" else
                  "Brunel real-time factor + synaptic events/s",
        "value": round(m["rtf"], 4),
        "unit": "bio_s/wall_s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(m["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload + f", dt 0.1 ms, seed {args.seed}; value over {steps} steps after "
                               f"{warmup} warmup steps",
                   "n_neurons": rt.n, "n_synapses": rt.canon.n_links, "dt_ms": rt.dt,
                   **({"timesteps_ms": timesteps} if timesteps else {}),
                   "graph_window": m["graph_window"],
                   "launch_mapping": {**mapping, **({"autotune_s": map_times} if map_times else {})},
                   "parallelism": "single GPU"},
        "events_per_s": round(m["events_per_s"], 1),
        "spikes_per_step": round(m["spikes"] / steps, 3),
        "events_per_step": round(m["events"] / steps, 1),
        "rates_hz": timed_rates,
        "roofline": _roofline(m, traffic, occupancy, kernel),
        "full_config": None if full is None else {
            "what": "1 s of bio time (10,000 steps) after 1,000 settling steps, graph windows of "
                    f"{full['graph_window']} steps",
            "rtf": round(full["rtf"], 4), "ms_per_step": round(full["ms_per_step"], 6),
            "events_per_s": round(full["events_per_s"], 1), "spikes_per_step": round(full["spikes"] / full["steps"], 3),
            "events_per_step": round(full["events"] / full["steps"], 1), "rates_hz": full_rates,
            "roofline": _roofline(full, traffic, occupancy, kernel),
            "reference_published": "PD14 ~25 us per 0.1 ms round (RTF ~4) on an RTX 4090, captured windows, DC "
                                   "background (TODO.md:47)" if kind == "pd14" else None,
        },
        "cpu_baseline": cpu_line,
    }
    del rt, net
    return line


# ----------------------------------------------------------------------------------------------- configs 2-4

def measure_sweep(args, steps: int, warmup: int, window: int, instances: int) -> dict:
    """Batched Brunel sweep: `instances` networks drawn from --seed (N log-uniform in [1k, 50k], in-degree
    10 %, g in [3, 8], eta in [0.9, 4]), `warmup` batched steps, then one batched launch per step (mapping
    from the autotuner: each candidate timed from the same state) against the same networks run one after
    another from the same state.  The final states of both must be bit-identical."""
    from naer2_amd import FusedRound
    from naer2_amd.autotune import measure_work, tune, tune_intake
    from naer2_amd.networks import brunel_sweep, sweep_params

    params = sweep_params(instances, seed=args.seed)
    t0 = time.perf_counter()
    nets = []
    for i, p in enumerate(params):
        nets += brunel_sweep([p], device="cuda", keep_canonical=False, record=False)
        if i % 32 == 31:
            log(f"sweep: built {i + 1}/{len(params)} networks")
    rts = [n.runtime for n in nets]
    n_syn = sum(rt.n_links for rt in rts)
    n_neurons = sum(rt.n for rt in rts)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    log(f"sweep: {len(rts)} networks, {n_neurons} neurons, {n_syn} synapses, {sum(rt.n_groups for rt in rts)} "
        f"groups, built in {t_build:.0f} s")
    if args.tune_intake:
        for rt in rts:
            tune_intake(rt, steps=200, graph_window=window)
    fr = FusedRound(rts, {"order": 1, "xcd": 0})  # warmup: past the onset transient, untimed
    fr.launch(warmup, window)
    fr.close()
    torch.cuda.synchronize()
    weights = measure_work(rts, steps=200, graph_window=window)
    best, times = tune(rts, steps=min(warmup, 1000), graph_window=window, use_cache=False, weights=weights)
    log(f"sweep: autotune {best} from {times}")
    snaps = [rt.snapshot() for rt in rts]
    fr = FusedRound(rts, best, weights=weights)
    fr.prepare(steps, window)  # every graph the timed launch replays, captured untimed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fr.launch(steps, graph_window=window)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    fr.close()
    log(f"sweep: batched {steps} steps: {wall_b:.3f} s")
    final_b = [(rt.v.clone(), rt.ref.clone(), rt.ring.clone()) for rt in rts]
    for rt, sn in zip(rts, snaps):
        rt.restore(sn)
        rt.prepare(steps, window)  # every graph of the sequential runs, captured untimed
    torch.cuda.synchronize()
    wall_s = 0.0
    for rt in rts:
        t0 = time.perf_counter()
        rt.run(steps, window)
        torch.cuda.synchronize()
        wall_s += time.perf_counter() - t0
    log(f"sweep: sequential {steps} steps x {len(rts)}: {wall_s:.3f} s")
    same = all(torch.equal(a[0].view(torch.int32), rt.v.view(torch.int32)) and torch.equal(a[1], rt.ref)
               and torch.equal(a[2].view(torch.int32), rt.ring.view(torch.int32)) for a, rt in zip(final_b, rts))
    bio = steps * rts[0].dt / 1000.0
    out = {
        "metric": "batched Brunel sweep: aggregate instance-bio-s per wall-s, batched vs sequential plain runs",
        "value": round(len(rts) * bio / wall_b, 4), "unit": "instance_bio_s/wall_s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(rts)} Brunel networks (N log-uniform 1k-50k, in-degree 10 %, g U[3,8], "
                               f"eta U[0.9,4]), seed {args.seed}, {steps} steps = {bio:g} s bio each",
                   "n_neurons": n_neurons, "n_synapses": n_syn},
        "sequential_value": round(len(rts) * bio / wall_s, 4),
        "batched_wall_s": round(wall_b, 4), "sequential_wall_s": round(wall_s, 4),
        "batched_speedup": round(wall_s / wall_b, 3),
        "mapping": best, "autotune_s": times,
        "bit_exact_batched_vs_sequential": bool(same),
        "build_s": round(t_build, 1),
    }
    del fr, rts, nets, snaps, final_b
    torch.cuda.empty_cache()
    return out


def measure_pd14grad(args, steps: int, instances: int) -> dict:
    """Batched, differentiable PD14: `instances` seed-distinct full-scale microcircuits in one launch per
    step, `steps` steps (500 ms bio at 5000), loss = squared error of the per-population rates against
    seed-drawn targets in [0.5, 10] Hz, gradient with respect to the 8x8 weight scales from the surrogate
    adjoint (truncated at --cutoff, scaled by --gamma) with checkpoints every --checkpoint steps.  Times the
    forward (checkpointed, graph windows) and the backward (recompute + adjoint) with HIP events, and checks
    the same adjoint on a 4-neuron-per-population PD14 subsample against the FP64 autograd unroll (on the
    same spike trains), whose own gradient FP64 central differences pin in the CPU tests."""
    from naer2_amd.adjoint import RateLossGrad
    from naer2_amd.networks import pd14

    n_inst, T, K = instances, steps, args.checkpoint
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
    log(f"pd14grad: built {n_inst} PD14 instances ({n_syn} synapses) in {time.perf_counter() - t0:.0f} s")
    sur = {"cutoff": args.cutoff, "gamma": args.gamma}
    grad = RateLossGrad(rts, alpha, checkpoint=K, long_delay=args.long_delay, long_parts=args.long_parts,
                        pure_ranges=args.pure_ranges, overlap=not args.serial_backward, **sur)
    t1 = time.perf_counter()
    loss, g = grad(T, targets)
    wall = time.perf_counter() - t1
    fw, bw, phases = grad.last["forward_ms"], grad.last["backward_ms"], grad.last["backward_phases_ms"]
    log(f"pd14grad: loss {loss:.4g}, forward {fw:.0f} ms, backward {bw:.0f} ms ({phases}), wall {wall:.1f} s")
    del grad, rts
    torch.cuda.empty_cache()
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
    bio = T * 0.1 / 1000.0
    return {
        "metric": "batched differentiable PD14: backward/forward cost ratio (rate-loss surrogate adjoint)",
        "value": round(bw / fw, 3), "unit": "ratio", "n_gpus": 1, "steps": T, "warmup": 0,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n_inst} seed-distinct full-scale PD14 microcircuits (77169 neurons each) in one "
                               f"launch a step, {T} steps of 0.1 ms, loss sum (rate - target)^2 over populations, "
                               f"targets U[0.5, 10] Hz, grad wrt the 8x8 weight scales, checkpoints every {K} steps, "
                               f"seed {args.seed}",
                   "n_synapses": n_syn},
        "forward_ms": round(fw, 1), "backward_ms": round(bw, 1),
        "backward_phases_ms": {k: round(v, 1) for k, v in phases.items()},
        "forward_rtf_aggregate": round(n_inst * bio / (fw / 1000.0), 3),
        "loss": loss, "grad_norm": float(np.linalg.norm(g)), "grad_finite": bool(np.isfinite(g).all()),
        "surrogate": {"kappa": 25.0, **sur}, "long_delay": args.long_delay, "long_parts": args.long_parts,
        "pure_ranges": bool(args.pure_ranges),
        "recompute_overlapped": not args.serial_backward,
        "grad_rel_err_vs_fp64_subsample": err,
        "subsample": f"4 neurons per population (32), PD14 probabilities x10, {Tc} steps, same spike trains",
    }


def measure_bptt(args, iters: int, warmup: int) -> dict:
    """Surrogate-gradient BPTT: B=128 Poisson inputs on 700 channels (rates U[0, 100] Hz), T=100 steps of
    1 ms, 1024 recurrent LIF units, 20 leaky readouts, cross-entropy of the max-over-time readout,
    checkpoints every 25 steps.  Times forward and backward with HIP events (median of `iters` iterations
    after `warmup`), and checks the gradients against the FP64 autograd unroll (same spikes)."""
    from naer2_amd.rlif import RecurrentLIF, RLIFConfig, max_over_time_ce, poisson_input
    B, C, H, O, T = 128, 700, 1024, 20, 100
    g = torch.Generator(device="cuda").manual_seed(args.seed)
    cfg = RLIFConfig(checkpoint=25)
    net = RecurrentLIF(C, H, O, cfg, generator=g)
    x = poisson_input(B, C, T, 1.0, g)
    labels = torch.randint(0, O, (B,), generator=g, device="cuda")
    fw, bw = [], []
    for it in range(warmup + iters):
        net.zero_grad(set_to_none=True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        r, s = net(x)
        loss = max_over_time_ce(r, labels)
        e1.record()
        loss.backward()
        e2.record()
        torch.cuda.synchronize()
        if it >= warmup:
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
    return {
        "metric": "BPTT backward/forward cost ratio (surrogate-gradient recurrent LIF)",
        "value": round(bwd_ms / fwd_ms, 4), "unit": "ratio", "n_gpus": 1, "steps": iters, "warmup": warmup,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "B=128, 700 Poisson inputs (0-100 Hz), 1024 recurrent LIF, 20 leaky readouts, "
                               "T=100 x 1 ms, fast-sigmoid slope 25, checkpoints every 25 steps"},
        "forward_ms": round(fwd_ms, 4), "backward_ms": round(bwd_ms, 4),
        "forward_dense_equiv_tflops": round(flops / (fwd_ms * 1e-3) / 1e12, 3),
        "spike_fraction": round(float(s.mean()), 5),
        "grad_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
        "grad_rel_err_max": float(f"{max(errs.values()):.3e}"),
        "loss": float(loss.detach()), "loss_fp64": float(loss64),
    }


# ----------------------------------------------------------------------------------------------- event path

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


def measure_brunel_exact(args) -> dict:
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
    log(f"built brunel_exact: {rt.n} nodes ({N} neurons), {rt.n_links} synapses, {rt.n_groups} groups in "
        f"{t_build:.1f} s")
    gw = args.ev_graph
    warm, n_win = args.ev_warmup, args.ev_windows
    rt.run(warm - 1, gw)
    net.record = True
    rt.run(1, 0)  # the last warmup window recorded: the CPU oracle starts from it
    net.record = False
    torch.cuda.synchronize()
    prev_t, prev_n = rt.spikes()
    snap = (rt.v.clone(), rt.last.clone(), rt.tre.clone(), rt.nxt.clone())
    c0 = rt.counters()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    rt.run(n_win, gw)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c1 = rt.counters()
    rt.check()
    events, spikes = c1[0] - c0[0], c1[1] - c0[1]
    bio_s = n_win * W / 1000.0
    rtf = bio_s / wall
    log(f"brunel_exact: timed {n_win} windows: {wall:.3f} s wall")
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
               "events_per_s": round(ev_cpu / tc, 1)}
    t_rec = rt.window_index * W
    net.record = True
    rt.run(67, 0)
    net.record = False
    r = rt.rates(t_skip_ms=t_rec)
    rates = {name: r[p] for name, p in pops.items()}
    ref_rtf = 1.0 / 2.94
    return {
        "metric": "event-driven Brunel real-time factor + synaptic events/s",
        "value": round(rtf, 4), "unit": "bio_s/wall_s", "n_gpus": 1, "steps": n_win, "warmup": warm,
        "ms_per_step": round(wall * 1000.0 / n_win, 6), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(rtf / ref_rtf, 2) if N == 50000 else None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"brunel_exact: event-driven Brunel model A with exact spike times, N {N} "
                               f"(N_E {pops['E'].n}, N_I {pops['I'].n}), {pops['P'].n} Poisson sources at "
                               f"{brunel_exact_theory(N)['rate_ext_Hz']:.1f} Hz, C_E {brunel_exact_theory(N)['C_E']}, "
                               f"J 0.1 mV, g 5, D 1.5 ms, {rt.n_links} synapses; a step = one 1.5 ms window; "
                               f"{n_win} windows timed, seed {args.seed}",
                   "n_neurons": N, "n_synapses": rt.n_links, "window_ms": W, "graph_window": gw,
                   "parallelism": "single GPU"},
        "vs_baseline_note": "reference: N = 50,000 at 2.94 ms wall per simulated ms on an RTX 4090, captured "
                            "(test/networks/Brunel/net.py:77-78)",
        "events_per_s": round(events / wall, 1), "spikes_per_window": round(spikes / n_win, 1),
        "events_per_window": round(events / n_win, 1),
        "rates_hz": {name: round(float(v), 3) for name, v in rates.items()},
        "kernel_split_ms_per_window": {"deliver": round(d_ms / n_split, 4), "sort": round(s_ms / n_split, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": "ev_deliver_kernel", "kernel_avg_us": round(k_ms * 1000.0, 3),
                     "occupancy": occupancy, "algorithmic_bytes_per_launch": round(bytes_win, 1)},
        "cpu_baseline": cpu,
    }


# ----------------------------------------------------------------------------------------------- main

# rough seconds each configuration takes at its full size (builds included), for the time budget
COST_S = {"brunel": 40.0, "bptt": 20.0, "pd14grad": 90.0, "sweep": 240.0}


def _edge_variant(net, name: str) -> None:
    """Turn PD14's connections into edges of the step's edge variant: "gain" (gain 1.0: delivers exactly w),
    "plastic" (factor -0.0: w + -0.0 = w, stored back every delivery), or "mixed" (excitatory connections
    plastic with factors U(-1e-4, 1e-4) nA, inhibitory ones into the potential port with their weight scaled
    by R_mem tau_syn / tau_m = 2 mV per nA: the jump that the charge of their current pulse would give)."""
    g = torch.Generator(device="cpu").manual_seed(7)
    for c in net.connections:
        if name == "gain":
            c.kind = "gain"
            c.gain = 1.0
        elif name == "plastic":
            c.kind = "plastic"
            c.factor = -0.0
        elif float(c.weight[0]) >= 0:
            c.kind = "plastic"
            c.factor = (torch.rand(c.shape, generator=g) * 2e-4 - 1e-4)
        else:
            c.port = "potential"
            c.weight = c.weight * 2.0


def measure_edges(args, steps: int, warmup: int) -> dict:
    """Full-scale PD14 through the step's edge variant (edge kinds and ports), timed like the headline: the
    plain step, then every connection a gain edge of gain 1.0 and a plastic edge of factor -0.0 -- both deliver
    exactly the plain weights, so their rasters must equal the plain run's bit for bit -- and a mixed network
    (plastic excitatory connections, inhibitory ones into the potential port)."""
    from naer2_amd.networks import pd14

    out, plain_bits = {}, None
    for name in ("plain", "gain", "plastic", "mixed"):
        t0 = time.perf_counter()
        net, pops = pd14(scale=1.0, seed=args.seed, device="cuda")
        if name != "plain":
            _edge_variant(net, name)
        rt = net.build()
        net.record = True
        out_deg = rt.canon.out_degree()
        build_s = time.perf_counter() - t0
        if warmup:
            rt.prepare(warmup, args.window)
            net.run(warmup, graph_window=args.window)
        m = timed_run(net, rt, steps, args.window, out_deg, "pd14")
        rates = net.rates(t_skip_ms=m["from_step"] * rt.dt)
        bits = rt.raster_bits(0, rt.step)
        row = {"rtf": round(m["rtf"], 4), "kernel_avg_us": round(m["kernel_avg_us"], 3),
               "events_per_s": round(m["events_per_s"], 1), "spikes_per_step": round(m["spikes"] / steps, 3),
               "rates_hz": {k: round(rates[p], 3) for k, p in pops.items()}, "edges": bool(rt.edges),
               "build_s": round(build_s, 1)}
        if name == "plain":
            plain_bits = bits.clone()
        elif name in ("gain", "plastic"):
            row["raster_equals_plain"] = bool(torch.equal(bits, plain_bits))
            if name == "plastic":
                w = rt.link_weights()
                w0 = torch.cat([c.weight.to(w.device) for c in net.connections])
                row["weights_unchanged"] = bool(torch.equal(w.view(torch.int32), w0.view(torch.int32)))
        out[name] = row
        log(f"edges {name}: RTF {row['rtf']}, {row['kernel_avg_us']} us per launch")
        del rt, net, bits
        torch.cuda.empty_cache()
    return {
        "metric": "PD14 real-time factor through the edge variant (plastic edges)",
        "value": out["plastic"]["rtf"], "unit": "bio_s/wall_s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"pd14 full scale, seed {args.seed}, {steps} timed steps after {warmup}: plain, every "
                               "connection gain 1.0, every connection plastic with factor -0.0, mixed (plastic "
                               "excitatory, inhibitory into the potential port)"},
        "variants": out,
    }


def configs(args, cpu: dict) -> dict:
    """The other configurations at their full sizes, in order of cost, each within the time budget."""
    out = {}
    plan = [
        ("config0_brunel", "brunel", lambda: measure_plain(args, "brunel", 10000, 1000, not args.no_cpu, cpu)),
        ("config3_bptt", "bptt", lambda: measure_bptt(args, 5, 2)),
        ("config4_pd14grad", "pd14grad", lambda: measure_pd14grad(args, 5000, 4)),
        ("config2_sweep", "sweep", lambda: measure_sweep(args, 10000, 1000, 100, 256)),
    ]
    for name, key, fn in plan:
        left = args.budget_s - (time.perf_counter() - T_START)
        if left < COST_S[key]:
            out[name] = {"skipped": f"time budget: {left:.0f} s left of --budget-s {args.budget_s:g}, "
                                    f"about {COST_S[key]:g} s needed"}
            log(f"{name}: skipped ({out[name]['skipped']})")
            continue
        t0 = time.perf_counter()
        try:
            out[name] = fn()
        except Exception as e:  # one configuration failing must not hide the headline line
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name}: {type(e).__name__}: {e}")
        out[name]["seconds"] = round(time.perf_counter() - t0, 1)
        torch.cuda.empty_cache()
        log(f"{name}: done in {out[name]['seconds']} s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 10000 = 1 s of bio time; pd14grad: 5000)")
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--window", type=int, default=100, help="steps per captured hipGraph of long runs")
    ap.add_argument("--cpu-steps", type=int, default=2000,
                    help="scales the CPU samples (PD14: 2000 steps on one thread, 1000 per thread on every CPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="pd14: leave out the other configurations (`configs` in the line)")
    ap.add_argument("--no-full", action="store_true",
                    help="pd14/brunel: leave out the configuration's own 10,000-step run when the timed run is shorter "
                         "(`full_config` null; for counter-collecting profiler runs, which slow every launch)")
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="pd14: wall seconds from start within which the other configurations are run")
    ap.add_argument("--timestep", action="append", metavar="POP=MS",
                    help="pd14/brunel: a population's own timestep, a whole multiple of dt (repeatable)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", choices=["pd14", "brunel", "sweep", "bptt", "pd14grad", "brunel_exact",
                                           "pd14_edges"],
                    default="pd14",
                    help="pd14: the headline PD14 real-time factor (with the other configurations beside it); "
                         "brunel: the 2500-neuron Brunel network; sweep: the batched Brunel sweep; bptt: the "
                         "recurrent-LIF BPTT; pd14grad: the batched differentiable PD14; brunel_exact: the "
                         "event-driven Brunel model with exact spike times; pd14_edges: PD14 through the step's "
                         "edge variant (gain, plastic and potential-port edges)")
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
    ap.add_argument("--pure-ranges", action="store_true",
                    help="pd14grad: population-pure target ranges (each population's targets split apart) instead of "
                         "equal ones across populations")
    ap.add_argument("--serial-backward", action="store_true",
                    help="pd14grad: recompute each checkpoint segment before its reverse steps on one stream, instead "
                         "of recomputing the next segment on a second stream while a segment is reversed")
    ap.add_argument("--cutoff", type=float, default=0.2,
                    help="pd14grad: surrogate support |V - theta| / (theta - V_reset) (sparse gradient)")
    ap.add_argument("--gamma", type=float, default=0.01,
                    help="pd14grad: surrogate scale (0.1 and above diverge over 500 ms at full scale)")
    ap.add_argument("--xcd", type=int, default=None,
                    help="pd14/brunel: launch mapping of the step (n2h_set_tuning): 0 block b runs group b; k > 1 "
                         "consecutive groups on each of k XCDs; default: the autotuner's choice (autotune.tune_plan)")
    ap.add_argument("--xcd-run", type=int, default=0,
                    help="pd14/brunel: with --xcd k > 1, runs of this many consecutive groups dealt to the XCDs in "
                         "turn (0: one contiguous chunk per XCD)")
    ap.add_argument("--tune-intake", action="store_true",
                    help="sweep: time each network's row-intake thresholds first (naer2_amd.autotune.tune_intake)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="only run 5 plain device-to-device copies of MiB each (the reference kernel that scales "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes), then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: this build measures one MI355X (the spiking-network step would shard by "
              "target groups, with one all-gather of each step's spike lists; not built)", file=sys.stderr)
        return 2
    if args.steps is None:
        args.steps = 5000 if args.workload == "pd14grad" else 10000
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be positive and --warmup non-negative", file=sys.stderr)
        return 2
    if args.window < 0 or args.window % 4:
        print("--window must be 0 (eager launches) or a positive multiple of 4 (the spike lists' phase)",
              file=sys.stderr)
        return 2
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n}))
        return 0
    cpu = host_cpu()
    log(f"host: {cpu}")
    if args.workload == "bptt":
        line = measure_bptt(args, max(1, args.steps if args.steps < 1000 else 5), min(args.warmup, 2))
    elif args.workload == "pd14grad":
        line = measure_pd14grad(args, args.steps, args.instances or 4)
    elif args.workload == "brunel_exact":
        line = measure_brunel_exact(args)
    elif args.workload == "sweep":
        line = measure_sweep(args, args.steps, max(args.warmup, args.window), args.window, args.instances or 256)
    elif args.workload == "pd14_edges":
        line = measure_edges(args, args.steps, args.warmup)
    else:
        line = measure_plain(args, args.workload, args.steps, args.warmup, not args.no_cpu, cpu)
        torch.cuda.empty_cache()
        if args.workload == "pd14" and not args.no_secondary:
            line["configs"] = configs(args, cpu)
        line["bench_wall_s"] = round(time.perf_counter() - T_START, 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
