"""Launch-mapping autotuner of batched sweeps (FusedRound): plan order, XCD-aware block mapping and stage size.

Every candidate mapping computes exactly the same results (each target group's work does not depend on
where or when it runs), so the tuner times the candidates on the batch's own first steps -- the warmup --
and keeps the fastest.  Choices are cached per device and batch signature in a JSON file
($NAER2_AUTOTUNE_CACHE, default ~/.cache/naer2_amd/autotune.json).

Candidates: order 0 (plans as given), 1 (most links first, so the heaviest groups start first and light
ones fill the tail) or 2 (most delivered links per group first, measured on each network alone:
measure_work); XCD mapping xcd 0 (blocks in tile order, a plan's groups dealt round-robin over the 8 XCDs),
xcd 8 run 0 (each XCD one contiguous eighth of the tiles: a plan's consecutive groups share an L2 -- spike
lists, offset rows, link segments -- but whole networks' uneven work lands on one XCD) or xcd 8 run 16
(runs of 16 neighbouring groups dealt to the XCDs in turn: the sharing within a run, the XCDs' shares
interleaved over the whole batch); stage 0 (the
lean batched stage: 128 sources and 384 links a pass, 4 row cells in flight, so more groups fit a CU) or 1
(the single-plan stage: 256, 512 and 8 -- fewer passes and more row cells in flight a group).  Every
candidate runs from the same snapshot of the state, which is restored afterwards.

tune_plan picks a single network's step launch mapping (tile order, one chunk per XCD or runs of 16 groups
per XCD) by timing each on the network's own steps from one snapshot of the state.

tune_intake picks, per network, the spike count from which a step takes the row intake (every in-link
walked against the spike bitmap) instead of the segment intake (the spiking sources' segments): both give
bit-identical results, and which is faster depends on the in-degree and the activity regime, so each
candidate is timed from the same snapshot of the state, which is restored afterwards.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

import torch

from ._lib import Tuning
from .fusion import FusedRound

# XCD mappings (xcd, run): tile order; one chunk of the batch per XCD; runs of 16 neighbouring groups dealt to
# the XCDs in turn
MAPPINGS = ((0, 0), (8, 0), (8, 16))
CANDIDATES = [{"order": o, "xcd": x, "run": r, "stage": g} for o in (0, 1, 2) for x, r in MAPPINGS for g in (0, 1)]


def cache_path() -> Path:
    return Path(os.environ.get("NAER2_AUTOTUNE_CACHE", Path.home() / ".cache" / "naer2_amd" / "autotune.json"))


def signature(runtimes) -> str:
    h = hashlib.sha1()
    for rt in runtimes:
        h.update(f"{rt.n}:{rt.n_groups}:{rt.n_links}:{rt.d_slots};".encode())
    # "3": the candidate set with stage sizes and XCD runs (entries cached before it lack those dimensions)
    return f"3|{torch.cuda.get_device_name(runtimes[0].device)}|{len(runtimes)}|{h.hexdigest()[:16]}"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _save(table: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(table, indent=1, sort_keys=True))
    except OSError:
        pass


def measure_work(runtimes, steps: int = 200, graph_window: int = 100) -> list[float]:
    """Delivered links per step and target group of each runtime, run alone for `steps` steps from its
    current state (which is restored afterwards): the order-2 weight."""
    out = []
    for rt in runtimes:
        rec = rt.net.record
        rt.net.record = True
        rt._ensure_slots(steps + graph_window + 4)
        snap = rt.snapshot()
        s0 = rt.step
        rt.run(steps, graph_window)
        raster = rt.raster(s0, rt.step)
        od = rt.out_degree.cpu().numpy()
        ev = float(od[raster.nonzero()[1]].sum()) / steps
        rt.restore(snap)
        rt.net.record = rec
        out.append(ev / rt.n_groups)
    return out


def tune(runtimes, steps: int = 200, graph_window: int = 100, reserve_steps: int = 0, use_cache: bool = True,
         weights: list | None = None):
    """(best mapping as a dict, {candidate: seconds}).  Order 2 needs weights (measure_work); without them
    it is left out.  The runtimes end in the state they started in."""
    key = signature(runtimes)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        return table[key]["best"], table[key]["times"]
    snaps = [rt.snapshot() for rt in runtimes]
    times = {}
    for cand in CANDIDATES:
        if cand["order"] == 2 and weights is None:
            continue
        for rt, sn in zip(runtimes, snaps):
            rt.restore(sn)
        tuning = Tuning(order=cand["order"], xcd=cand["xcd"], stage=cand.get("stage", 0), run=cand.get("run", 0))
        fr = FusedRound(runtimes, tuning, reserve_steps=reserve_steps, weights=weights)
        fr.launch(graph_window, graph_window)  # capture + first replay, untimed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fr.launch(steps, graph_window)
        torch.cuda.synchronize()
        times[json.dumps(cand, sort_keys=True)] = time.perf_counter() - t0
        fr.close()
    for rt, sn in zip(runtimes, snaps):
        rt.restore(sn)
    best = json.loads(min(times, key=times.get))
    if use_cache:
        table[key] = {"best": best, "times": times}
        _save(table)
    return best, times


# launch mappings (xcd, run) of a single network's step kernels (n2h_set_tuning): blocks in group order;
# one contiguous chunk of the groups per XCD (the runtime's default); runs of 16 groups dealt to the XCDs
PLAN_MAPPINGS = ((0, 0), (8, 0), (8, 16))


def tune_plan(rt, steps: int = 200, graph_window: int = 100, mappings=PLAN_MAPPINGS):
    """Set rt's step launch mapping (Runtime.set_launch_mapping) to the fastest of `mappings`, each timed
    over `steps` steps from the same snapshot of the state, which is restored afterwards (recording
    restarts there); every mapping gives the same results.  Returns ({"xcd", "run"}, {mapping: seconds})."""
    rt._ensure_slots(max(steps, graph_window) + 4)  # the recorded history must not grow under the snapshot
    snap = rt.snapshot()
    times = {}
    for xcd, run in mappings:
        rt.set_launch_mapping(xcd, run)
        rt.restore(snap)
        rt.prepare(steps, graph_window)  # graphs captured untimed
        torch.cuda.synchronize(rt.device)
        t0 = time.perf_counter()
        rt.run(steps, graph_window)
        torch.cuda.synchronize(rt.device)
        times[json.dumps({"xcd": xcd, "run": run}, sort_keys=True)] = time.perf_counter() - t0
    best = json.loads(min(times, key=times.get))
    rt.set_launch_mapping(best["xcd"], best["run"])
    rt.restore(snap)
    return best, times


INTAKE_FRACTIONS = (0.002, 0.005, 0.01, 0.02, 0.05)


def tune_intake(rt, steps: int = 200, graph_window: int = 100, fractions=INTAKE_FRACTIONS):
    """Set rt's row-intake threshold to the fastest of max(65, f * n) for f in fractions, or never; the
    state is restored afterwards (recording restarts there).  Returns (threshold, {threshold: seconds});
    a no-op without a row table."""
    if rt.rows is None:
        return None, {}
    never = 1 << 30
    cands = sorted({max(65, int(f * rt.n)) for f in fractions}) + [never]
    rt._ensure_slots(max(steps, graph_window) + 4)  # the recorded history must not grow under the snapshot
    snap = rt.snapshot()
    times = {}
    for c in cands:
        rt.set_dense_min(c)
        rt.restore(snap)
        rt.run(graph_window, graph_window)  # capture, untimed
        rt.restore(snap)
        torch.cuda.synchronize(rt.device)
        t0 = time.perf_counter()
        rt.run(steps, graph_window)
        torch.cuda.synchronize(rt.device)
        times[c] = time.perf_counter() - t0
    best = min(times, key=times.get)
    rt.set_dense_min(best)
    rt.restore(snap)
    return best, times
