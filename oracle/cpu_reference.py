"""CPU-BASELINE TIMING ONLY: the reference's own CPU-capable algorithm, timed on the host's cores.

tools/brunel_reference.py:14-48 is a plain-torch time-driven loop:
- take the due slot of a delayed-input buffer plus J x Poisson(lambda) background counts;
- apply the LIF update with a refractory period;
- add the spikes' input D steps ahead with a sparse matrix product W @ s.

This module restates that loop for a network built here: one sparse matrix per delay class (the networks
here have per-link delays), and the exponential synapse of test/fused_reference.py:86-107 for PD14. It runs
it on the CPU with torch.set_num_threads(threads). Only bench.py's cpu_baseline leg uses it. Its results
match the GPU's statistically, not bit for bit: torch's Poisson stream and the sparse product's summation
order differ from the build's fixed contract.
"""
from __future__ import annotations

import time
import warnings

import numpy as np
import torch


def run(model: int, pop_first, params, lam, canon, d_slots: int, v0, i0, n_steps: int, threads: int,
        record: bool = False) -> dict:
    """Time n_steps steps. model 1: delta LIF, 2: exponential synapse. params: the populations'
    n2h_pop_params; lam: Poisson mean per population and step; canon: the canonical CSR (CPU).  record: also
    return every spike as (step, neuron) in "spike_steps" / "spike_ids" (statistics such as CV-ISI)."""
    torch.set_num_threads(int(threads))
    f32 = torch.float32
    pf = np.asarray(pop_first, dtype=np.int64)
    n = int(pf[-1])
    pop = torch.as_tensor(np.repeat(np.arange(len(params)), np.diff(pf)))
    col = lambda vals: torch.as_tensor(np.asarray(vals, dtype=np.float32))[pop]
    decay = col([p.decay for p in params])
    p_i, p_vi = col([p.p_i for p in params]), col([p.p_vi for p in params])
    v_rest, v_th, v_reset = col([p.v_rest for p in params]), col([p.v_th for p in params]), col([p.v_reset for p in params])
    j_ext = col([p.j_ext for p in params])
    ref_steps = torch.as_tensor(np.asarray([p.ref_steps for p in params], dtype=np.int32))[pop]
    lam_n = col(lam)
    # one CSR matrix (rows = targets) per delay class, from one sort by (delay, target) -- on the canonical
    # table's device, then moved to the host
    dev = canon.tgt.device
    src = torch.repeat_interleave(torch.arange(n, device=dev), canon.out_degree())
    tgt = canon.tgt.to(torch.int64)
    d = canon.d.to(torch.int64)
    order = torch.argsort(d * n + tgt, stable=True)
    src, tgt, d, w = src[order].cpu(), tgt[order].cpu(), d[order].cpu(), canon.w[order].cpu().to(f32)
    del order
    warnings.filterwarnings("ignore", message="Sparse CSR tensor support is in beta")
    mats = []
    ks, counts = torch.unique_consecutive(d, return_counts=True)
    lo = 0
    for k, c in zip(ks.tolist(), counts.tolist()):
        rows = torch.bincount(tgt[lo:lo + c], minlength=n)
        crow = torch.zeros(n + 1, dtype=torch.int64)
        crow[1:] = torch.cumsum(rows, 0)
        mats.append((int(k), torch.sparse_csr_tensor(crow, src[lo:lo + c].clone(), w[lo:lo + c].clone(), (n, n))))
        lo += c
    del src, tgt, d, w
    out_deg = canon.out_degree().cpu()
    D = int(d_slots)
    buf = torch.zeros(D, n, dtype=f32)
    V = torch.as_tensor(np.asarray(v0, dtype=np.float32)).clone()
    I = torch.as_tensor(np.asarray(i0, dtype=np.float32)).clone() if model == 2 else None
    R = torch.zeros(n, dtype=torch.int32)
    per_neuron = torch.zeros(n, dtype=torch.int64)
    rec_steps, rec_ids = [], []
    n_spk = 0
    n_ev = 0
    t0 = time.perf_counter()
    for s in range(n_steps):
        slot = s % D
        inp = buf[slot] + j_ext * torch.poisson(lam_n)
        buf[slot] = 0
        active = R == 0
        if model == 1:
            V = torch.where(active, V * decay + inp, V)
            fired = active & (V >= v_th)
        else:
            I = I + inp
            V = torch.where(active, v_rest + decay * (V - v_rest) + p_vi * I, v_reset)
            I = p_i * I
            fired = active & (V > v_th)
        V = torch.where(fired, v_reset, V)
        R = torch.where(fired, ref_steps, torch.where(active, R, R - 1))
        idx = fired.nonzero().flatten()
        if idx.numel():
            sv = fired.to(f32)
            for k, W in mats:
                buf[(s + k) % D] += W @ sv
            n_spk += int(idx.numel())
            n_ev += int(out_deg[idx].sum())
            per_neuron[idx] += 1
            if record:
                rec_steps.append(np.full(idx.numel(), s, dtype=np.int64))
                rec_ids.append(idx.numpy().astype(np.int64))
    sec = time.perf_counter() - t0
    out = {"seconds": sec, "steps": n_steps, "spikes": n_spk, "events": n_ev, "threads": int(threads),
           "per_neuron": per_neuron.numpy()}
    if record:
        out["spike_steps"] = np.concatenate(rec_steps) if rec_steps else np.zeros(0, np.int64)
        out["spike_ids"] = np.concatenate(rec_ids) if rec_ids else np.zeros(0, np.int64)
    return out
