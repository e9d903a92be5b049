"""Surrogate-gradient adjoint of a batch of exponential-synapse LIF networks (the LIF_EXP step, e.g. PD14)
for a per-population rate loss, with respect to the weight scales alpha[target pop][source pop]:

    grad = RateLossGrad(runtimes, alpha, checkpoint=500)   # built with weights alpha * base (pd14(weight_scale=...))
    loss, g = grad(n_steps, targets)                       # targets [n_networks][n_pops] in Hz; g [P][P]

loss = sum over networks and populations of (rate - target)^2 (rates in Hz over the n_steps).  The forward is
the plain step (bit-exact, one batched launch per step, FusedRound), with the state saved every `checkpoint`
steps; the forward also tracks what fp32 rounding drops from each neuron's V and I (n2h_set_vtrace's
compensation state, checkpointed with the state; V, I and the spikes are unchanged), and traces the
compensated distance to the threshold through its last segment.  The backward runs the segments in reverse:
each before the last is recomputed from its checkpoint with that distance traced (beside the reverse of the
segment after it, on a stream of its own), then reversed in windows of `long_delay` steps:
n2h_adj_window gathers the input adjoints of every out-link with a delay of at least long_delay once per
window (they all lie past it), then n2h_adj_step reverses each step of the window -- one launch for all
networks, a wavefront per source gathering its short links' input adjoints and adding the window's long
sums (naer2_amd/csrc/adjoint.hip states the recursion).  The surrogate derivative is
gamma (kappa/2) / dv / (1 + kappa |u|)^2 for |u| = |V - theta| / dv <= cutoff and 0 beyond (dv = theta -
V_reset); the cutoff makes the gradient sparse (only neurons near the threshold, or spiking, gather).  The
reset and refractoriness carry no gradient (through the hard reset the adjoint diverges).

The reference is forward-only: this is a hand-written adjoint of delivery plus the LIF update, checkpointed
every K steps, added beside the plain PD14 path.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import MODEL_LIF_EXP, Tuning
from .fusion import FusedRound
from .topology import _stable_order_by_key


class AdjNet(ctypes.Structure):
    _fields_ = [("n_neurons", ctypes.c_int32), ("n_pops", ctypes.c_int32), ("ring_slots", ctypes.c_int32),
                ("long_delay", ctypes.c_int32), ("long_parts", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
                    (k, ctypes.c_void_p) for k in (
                    "pop_first", "out_ptr", "out_links", "long_ptr", "long_links", "alpha", "pop_coef", "vtrace",
                    "spikes", "gx_ring", "gx_recent", "long_bins", "g_v", "g_i", "g_sum", "range_pop")]


class Surrogate(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_float), ("cutoff", ctypes.c_float), ("gamma", ctypes.c_float),
                ("reserved", ctypes.c_float)]


def _bind(lib):
    if getattr(lib, "_adj_bound", False):
        return lib
    lib.n2h_adj_last_error.argtypes, lib.n2h_adj_last_error.restype = [ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int32
    lib.n2h_adj_step.argtypes = [ctypes.POINTER(AdjNet), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                 ctypes.POINTER(Surrogate), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.n2h_adj_step.restype = ctypes.c_int32
    lib.n2h_adj_window.argtypes = [ctypes.POINTER(AdjNet), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.POINTER(Surrogate), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.n2h_adj_window.restype = ctypes.c_int32
    lib.n2h_adj_work_words.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.n2h_adj_work_words.restype = ctypes.c_int64
    lib._adj_bound = True
    return lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        _lib.load().n2h_adj_last_error(buf, len(buf))
        raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")


def pure_ranges(pop_first, long_parts: int):
    """Population-pure target ranges: each non-empty population's targets split into near-equal contiguous
    ranges, their number proportional to its size (at least one; largest remainders), long_parts in all.
    (range_first int64 [long_parts + 1], range_pop int32 [long_parts]), or None when long_parts is not a
    multiple of 8 or is below the number of non-empty populations."""
    pf = np.asarray(pop_first, dtype=np.int64)
    sizes = np.diff(pf)
    live = np.nonzero(sizes > 0)[0]
    if long_parts % 8 or long_parts < len(live):
        return None
    share = sizes[live] / sizes[live].sum() * long_parts
    cnt = np.maximum(1, np.floor(share).astype(np.int64))
    while cnt.sum() > long_parts:  # the minimum of one each overshot: take from counts above one
        cnt[np.argmax(np.where(cnt > 1, cnt - share, -np.inf))] -= 1
    rem = share - cnt
    for k in np.argsort(-rem, kind="stable")[:long_parts - int(cnt.sum())]:
        cnt[k] += 1
    first, pops = [], []
    for p, c in zip(live, cnt):
        edges = pf[p] + (np.arange(c) * sizes[p]) // c
        first += edges.tolist()
        pops += [int(p)] * int(c)
    return np.asarray(first + [int(pf[-1])], dtype=np.int64), np.asarray(pops, dtype=np.int32)


def out_links(rt, alpha: torch.Tensor, long_delay: int = 1, long_parts: int = 1, pure: bool = False):
    """(short_ptr, short_links, long_ptr, long_links, range_pop): rt's canonical CSR by source packed for the
    adjoint, {target | (delay - 1) << 17 | target pop << 24, base weight bits} with base = weight /
    alpha[pt][ps], split into the links with delay < long_delay (ptr int64 [n + 1]) and the rest, these by
    (source, target range) (ptr int64 [n * long_parts + 1]); links int32 [E, 2], each run in canonical order.
    The target ranges are population-pure (pure_ranges; range_pop int32 [long_parts] on the device) where
    long_parts allows, else range k = target * long_parts // n (range_pop None)."""
    c = rt.canon
    if c is None:
        raise RuntimeError("the adjoint needs the canonical synapse table (Network.keep_canonical = True)")
    dev = rt.device
    pf = torch.as_tensor(np.asarray(rt.pop_first, dtype=np.int64), device=dev)
    tgt = c.tgt.to(dev).to(torch.int64)
    src = torch.repeat_interleave(torch.arange(rt.n, device=dev), c.out_degree().to(dev))
    pt = torch.searchsorted(pf, tgt, right=True) - 1
    ps = torch.searchsorted(pf, src, right=True) - 1
    x = tgt | ((c.d.to(dev).to(torch.int64) - 1) << 17) | (pt << 24)
    scale = alpha[pt, ps]
    if bool((scale == 0).any()):
        raise ValueError("a weight scale of a connected population pair is 0: its base weights are unknown")
    base = c.w.to(dev) / scale
    links = torch.stack([x.to(torch.int32), base.contiguous().view(torch.int32)], dim=1).contiguous()
    is_long = c.d.to(dev) >= long_delay
    pure = pure_ranges(rt.pop_first, long_parts) if (pure and long_parts > 1) else None
    out = []
    for keep, parts in ((~is_long, 1), (is_long, long_parts)):
        if parts > 1 and pure is not None:
            rf = torch.as_tensor(pure[0], device=dev)
            rk = torch.searchsorted(rf, tgt[keep], right=True) - 1
        else:
            rk = (tgt[keep] * parts) // rt.n
        key = src[keep] * parts + rk
        cnt = torch.bincount(key, minlength=rt.n * parts)
        ptr = torch.zeros(rt.n * parts + 1, dtype=torch.int64, device=dev)
        ptr[1:] = cnt.cumsum(0)
        if bool(keep.any()):
            sel = torch.empty_like(links[keep])
            sel[_stable_order_by_key(key, rt.n * parts)] = links[keep]  # canonical order within each run
        else:
            sel = torch.zeros(1, 2, dtype=torch.int32, device=dev)  # never empty
        out += [ptr, sel.contiguous()]
    out.append(None if pure is None else torch.as_tensor(pure[1], device=dev))
    return tuple(out)


class RateLossGrad:
    def __init__(self, runtimes: list, alpha, checkpoint: int = 500, kappa: float = 25.0,
                 cutoff: float = float("inf"), gamma: float = 1.0, graph_window: int = 100, tuning=None,
                 long_delay: int = 4, long_parts: int = 24, pure_ranges: bool = False, overlap: bool = True):
        if not runtimes or len(runtimes) > 16:
            raise ValueError("RateLossGrad: 1 to 16 networks")
        for rt in runtimes:
            if rt.model != MODEL_LIF_EXP:
                raise ValueError("RateLossGrad: exponential-synapse LIF networks only")
            if rt.edges:
                raise ValueError("RateLossGrad: static edges into the current only (no edge kinds or ports)")
            if len(rt.pops) > 8:
                raise ValueError("RateLossGrad: at most 8 populations")
            if not rt.net.record:
                raise ValueError("RateLossGrad: the networks must record their spikes (Network.record)")
            if rt.n > (1 << 17):
                raise ValueError("RateLossGrad: at most 2^17 neurons a network")
            if any(p.stride > 1 for p in rt.params):
                raise ValueError("RateLossGrad: every population must update every step (no timestep of its own)")
        if len({rt.step for rt in runtimes}) != 1:
            raise ValueError("RateLossGrad: the networks are at different steps")
        if checkpoint < 1:
            raise ValueError("checkpoint must be positive")
        if not 1 <= long_delay <= 8:
            raise ValueError("long_delay: 1 to 8 steps")
        if not 1 <= long_parts <= 64:
            raise ValueError("long_parts: 1 to 64 target ranges")
        self.runtimes = list(runtimes)
        self.dev = runtimes[0].device
        P = len(runtimes[0].pops)
        if any(len(rt.pops) != P for rt in runtimes):
            raise ValueError("RateLossGrad: every network needs the same populations")
        self.P = P
        self.alpha = torch.as_tensor(np.asarray(alpha, dtype=np.float32), device=self.dev).reshape(P, P).contiguous()
        # the batched step's launch mapping (forward and recompute): by default each XCD runs a contiguous
        # eighth of the groups, so that an XCD's groups share their offset rows and link segments in its L2
        # (4 PD14s: forward 219-222 -> 210 ms against block b running group b)
        self.tuning = tuning if tuning is not None else Tuning(order=0, xcd=8, stage=0, run=0)
        self.checkpoint, self.window = int(checkpoint), int(graph_window)
        self.sur = Surrogate(float(kappa), float(cutoff), float(gamma), 0.0)
        self.long_delay = int(long_delay)
        self.long_parts = int(long_parts)
        self.overlap = bool(overlap)
        self.tables = [out_links(rt, self.alpha, self.long_delay, self.long_parts, pure_ranges) for rt in runtimes]
        self.pop_first = [torch.as_tensor(np.asarray(rt.pop_first, dtype=np.int32), device=self.dev) for rt in runtimes]
        self.ring_slots = max(rt.d_slots for rt in runtimes) + 1
        if any(int(rt.canon.d.max()) >= self.ring_slots for rt in runtimes if rt.canon.d.numel()):
            raise ValueError("RateLossGrad: a delay reaches the ring length")
        self.last = {}

    def _counts(self, rt, s0: int, T: int) -> np.ndarray:
        """Spikes per population over steps [s0, s0 + T)."""
        gof, lof = rt.layout.group_of.to(self.dev), rt.layout.local_of.to(self.dev)
        per_n = torch.zeros(rt.n, dtype=torch.int64, device=self.dev)
        for a in range(s0, s0 + T, 500):
            b = min(s0 + T, a + 500)
            per_n += ((rt.raster_bits(a, b)[:, gof] >> lof) & 1).sum(dim=0)
        pf = np.asarray(rt.pop_first, dtype=np.int64)
        cs = torch.cat([torch.zeros(1, dtype=torch.int64, device=self.dev), per_n.cumsum(0)]).cpu().numpy()
        return (cs[pf[1:]] - cs[pf[:-1]]).astype(np.float64)

    def __call__(self, n_steps: int, targets):
        rts, K, T = self.runtimes, self.checkpoint, int(n_steps)
        if T < 1:
            raise ValueError("n_steps must be positive")
        targets = np.asarray(targets, dtype=np.float64).reshape(len(rts), self.P)
        lib = _bind(_lib.load())
        s0 = rts[0].step
        rec_from = [rt.record_from for rt in rts]
        # ---- forward, checkpointed ----
        t_fw = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_fw[0].record()
        # the compensation state (what fp32 rounding dropped from V and I), tracked through the forward and
        # checkpointed with the state, so that the recompute traces compensated distances to the threshold
        lo = [torch.zeros(2, rt.n, device=self.dev) for rt in rts]
        for rt, l in zip(rts, lo):
            _lib.check(lib.n2h_set_vtrace(rt._plan, None, 0, 0, l[0].data_ptr(), l[1].data_ptr()), "n2h_set_vtrace")
        fr = FusedRound(rts, self.tuning, reserve_steps=T + 1)
        # the traced distances to the threshold and the spike flags of a segment, two sets that alternate
        # between segments in the backward
        L0 = min(K, T)
        bufs = [[{"vtr": torch.empty(L0, rt.n, device=self.dev),
                  "spk": torch.empty(L0, rt.n, dtype=torch.uint8, device=self.dev)} for rt in rts] for _ in range(2)]
        n_seg = (T + K - 1) // K
        ckpts = []
        for k, k0 in enumerate(range(0, T, K)):
            if k < n_seg - 1:
                ckpts.append([(rt.snapshot(), l.clone()) for rt, l in zip(rts, lo)])
            else:
                # the last segment is traced as it runs: the backward reverses it first, and need not recompute
                # it from its checkpoint (the trace does not change the step's results)
                ckpts.append(None)
                for rt, b, l in zip(rts, bufs[k % 2], lo):
                    _lib.check(lib.n2h_set_vtrace(rt._plan, b["vtr"].data_ptr(), s0 + k0, T - k0, l[0].data_ptr(),
                                                  l[1].data_ptr()), "n2h_set_vtrace")
                fr.sync()
            fr.launch(min(K, T - k0), self.window)
        fr.close()
        for rt in rts:
            _lib.check(lib.n2h_set_vtrace(rt._plan, None, 0, 0, None, None), "n2h_set_vtrace")
        t_fw[1].record()
        final = [rt.snapshot() for rt in rts]
        dt_s = rts[0].dt * 1e-3
        loss, c_pops = 0.0, []
        for i, rt in enumerate(rts):
            sizes = np.diff(np.asarray(rt.pop_first, dtype=np.int64)).astype(np.float64)
            rates = self._counts(rt, s0, T) / (sizes * T * dt_s)
            loss += float(((rates - targets[i]) ** 2).sum())
            c_pops.append(2.0 * (rates - targets[i]) / (sizes * T * dt_s))
        # ---- backward ----
        t_bw = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_bw[0].record()
        R, W = self.ring_slots, self.long_delay
        st = []
        for i, rt in enumerate(rts):
            coef = np.zeros((self.P, 8), dtype=np.float32)
            for p, prm in enumerate(rt.params):
                coef[p] = [prm.decay, prm.p_i, prm.p_vi, prm.v_reset, prm.v_th, prm.v_th - prm.v_reset, c_pops[i][p], 0.0]
            st.append({
                "coef": torch.as_tensor(coef, device=self.dev),
                "gx": torch.zeros(rt.n, R + W - 1, device=self.dev), "gv": torch.zeros(rt.n, device=self.dev),
                "lb": torch.zeros(W, rt.n, 8, device=self.dev), "gr": torch.zeros(W, rt.n, device=self.dev),
                "gi": torch.zeros(rt.n, device=self.dev),
                "gsum": torch.zeros(rt.n, self.P, 2, device=self.dev),  # compensated: hi, lo
            })
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        nets = (AdjNet * len(rts))()
        for i, (rt, sd, (sptr, sl, lptr, ll, rpop)) in enumerate(zip(rts, st, self.tables)):
            nets[i] = AdjNet(rt.n, self.P, R, W, self.long_parts, 0, self.pop_first[i].data_ptr(), sptr.data_ptr(), sl.data_ptr(),
                             lptr.data_ptr(), ll.data_ptr(), self.alpha.data_ptr(), sd["coef"].data_ptr(), 0, 0,
                             sd["gx"].data_ptr(), sd["gr"].data_ptr(), 0, sd["gv"].data_ptr(), sd["gi"].data_ptr(),
                             sd["gsum"].data_ptr(), rpop.data_ptr() if rpop is not None else 0)
        sur = ctypes.byref(self.sur)
        sizes = (ctypes.c_int32 * len(rts))(*[rt.n for rt in rts])
        words = lib.n2h_adj_work_words(sizes, len(rts), W, self.long_parts)
        if words < 0:
            raise RuntimeError("n2h_adj_work_words: bad network sizes")
        work = torch.zeros(words, dtype=torch.int32, device=self.dev)
        wp, wn = work.data_ptr(), work.numel()
        marks = []  # (phase, event) in stream order: the backward's GPU time by phase

        def mark(phase):
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            marks.append((phase, e))

        mark("setup")
        # The recompute of segment k - 1 (its checkpoint restored, its steps traced, its spikes extracted) runs
        # on a stream of its own while segment k's reverse steps run on this one: the two touch disjoint
        # buffers (the runtimes' state against the adjoint's), and two trace/spike buffers alternate between
        # segments.  One batch recomputes every segment: n2h_batch_sync brings its step and its copy of the
        # plans' trace bindings in line after each restore, on the recompute's stream.
        rev_stream = torch.cuda.current_stream(self.dev)
        fr_rec = FusedRound(rts, self.tuning)
        rec_stream = torch.cuda.Stream(device=self.dev) if self.overlap else rev_stream
        rec_stream.wait_stream(rev_stream)
        ev_rec, ev_rev, t_rec, t_rev = [None] * n_seg, [None] * n_seg, [], []

        def timed_event(stream):
            e = torch.cuda.Event(enable_timing=True)
            e.record(stream)
            return e

        def spike_flags(k):
            start, L = s0 + k * K, min(K, T - k * K)
            for rt, b in zip(rts, bufs[k % 2]):
                gof, lof = rt.layout.group_of.to(self.dev), rt.layout.local_of.to(self.dev)
                b["spk"][:L] = ((rt.raster_bits(start, start + L)[:, gof] >> lof) & 1).to(torch.uint8)

        def recompute(k):
            start, L = s0 + k * K, min(K, T - k * K)
            with torch.cuda.stream(rec_stream):
                e0 = timed_event(rec_stream)
                for rt, (ck, lk), b, l in zip(rts, ckpts[k], bufs[k % 2], lo):
                    rt.restore(ck)
                    l.copy_(lk)
                    _lib.check(lib.n2h_set_vtrace(rt._plan, b["vtr"].data_ptr(), start, L, l[0].data_ptr(),
                                                  l[1].data_ptr()), "n2h_set_vtrace")
                fr_rec.sync()
                fr_rec.launch(L, 0)
                for rt in rts:
                    _lib.check(lib.n2h_set_vtrace(rt._plan, None, 0, 0, None, None), "n2h_set_vtrace")
                spike_flags(k)
                ev_rec[k] = timed_event(rec_stream)
            t_rec.append((e0, ev_rec[k]))

        def reverse(k):
            start, L = s0 + k * K, min(K, T - k * K)
            rev_stream.wait_event(ev_rec[k])
            e0 = timed_event(rev_stream)
            # windows of W steps from the segment's end: [t0, t1)
            # row addresses by arithmetic on the base pointers (a tensor view per row costs microseconds of host
            # time a step, as much as the step's kernels): vtrace fp32 [K][n], spikes u8 [K][n], long_bins [W][n][8]
            base = [(b["vtr"].data_ptr(), b["spk"].data_ptr(), sd["lb"].data_ptr(), rt.n)
                    for b, sd, rt in zip(bufs[k % 2], st, rts)]
            for t1 in range(L, 0, -W):
                t0 = max(0, t1 - W)
                for i, (vb, sb, lb, n) in enumerate(base):
                    nets[i].vtrace, nets[i].spikes, nets[i].long_bins = vb + 4 * t0 * n, sb + t0 * n, lb
                _check(lib.n2h_adj_window(nets, len(rts), start + t0, t1 - t0, sur, wp, wn, stream), "n2h_adj_window")
                for t in reversed(range(t0, t1)):
                    for i, (vb, sb, lb, n) in enumerate(base):
                        nets[i].vtrace, nets[i].spikes, nets[i].long_bins = vb + 4 * t * n, sb + t * n, lb + 32 * (t - t0) * n
                    _check(lib.n2h_adj_step(nets, len(rts), start + t, t - t0, sur, wp, wn, stream), "n2h_adj_step")
            ev_rev[k] = timed_event(rev_stream)
            t_rev.append((e0, ev_rev[k]))

        # the last segment was traced by the forward: only its spike flags are extracted before its reverse,
        # and before the recompute stream's first restore replaces the bit history they are read from
        e0 = timed_event(rev_stream)
        spike_flags(n_seg - 1)
        ev_rec[n_seg - 1] = timed_event(rev_stream)
        t_rec.append((e0, ev_rec[n_seg - 1]))
        rec_stream.wait_event(ev_rec[n_seg - 1])
        for k in reversed(range(n_seg)):
            if k >= 1:
                if k + 1 < n_seg:
                    rec_stream.wait_event(ev_rev[k + 1])  # segment k + 1's buffers become segment k - 1's
                recompute(k - 1)
            reverse(k)
        rev_stream.wait_stream(rec_stream)
        t_bw[1].record()
        fr_rec.close()
        for rt, sn, rf in zip(rts, final, rec_from):
            rt.restore(sn)
            rt.record_from = rf  # the history still holds every step since (the final snapshot's bits)
        # dL/dalpha[a][b] = sum over networks and source neurons j of population b of g_sum[j][a] (hi + lo), in
        # fp32 tree sums on the device
        g = torch.zeros(self.P, self.P, device=self.dev)
        for rt, sd in zip(rts, st):
            pf = np.asarray(rt.pop_first, dtype=np.int64)
            for b in range(self.P):
                blk = sd["gsum"][pf[b]:pf[b + 1]]
                g[:, b] += blk[:, :, 0].sum(dim=0) + blk[:, :, 1].sum(dim=0)
        torch.cuda.synchronize(self.dev)
        phases = {"setup": t_bw[0].elapsed_time(marks[0][1]),
                  "recompute": sum(a.elapsed_time(b) for a, b in t_rec),   # on its own stream when overlapped
                  "reverse": sum(a.elapsed_time(b) for a, b in t_rev),
                  "first_segment": t_rec[0][0].elapsed_time(t_rec[0][1])}  # its spike flags: nothing overlaps them
        self.last = {"forward_ms": t_fw[0].elapsed_time(t_fw[1]), "backward_ms": t_bw[0].elapsed_time(t_bw[1]),
                     "backward_phases_ms": phases, "c_pops": c_pops}
        return loss, g.cpu().numpy().astype(np.float64)
