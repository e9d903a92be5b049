"""Event-driven, exact-time networks: Brunel's model A as the reference runs it, built into one n2h_ev plan
of libnaer2_hip.so and advanced window by window on the GPU.

Reference: test/networks/Brunel/neuron.py:6-45 and poisson.py:7-27 (the node types), naer2/population.py
(a population whose incoming connections each have one delay gets queues, SPIKES.md), naer2/network.py:80-152
(Network.step: every population runs every D, the sources' derived timestep), naer2/history.py:119-124
(a run's spikes as (e, node)).

Every connection has one delay value, a whole multiple (its lag) of the smallest one, the window W.  The
spikes of window k are the inputs of window k + lag through each connection, so a window is two kernel
launches (deliver, sort) with no host involvement; windows replay in captured hipGraphs.  There is no CPU fallback: without the HIP library or a GPU this raises.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import EV_NEURON, EV_POISSON, EV_TRAIN, GROUP
from .topology import canonical, group_layout, segments

EVENT_FAMILIES = {"event_lif": EV_NEURON, "event_poisson": EV_POISSON, "event_train": EV_TRAIN}


def is_event_net(net) -> bool:
    fams = {p.family for p in net.populations}
    ev = fams & set(EVENT_FAMILIES)
    if ev and fams - set(EVENT_FAMILIES):
        raise NotImplementedError(f"event-driven and time-driven node types in one network: {sorted(fams)}")
    return bool(ev)


def _uniform(t: torch.Tensor, what: str) -> float:
    if t.numel() == 0:
        return float("nan")
    v = float(t.reshape(-1)[0])
    if not bool((t == v).all()):
        raise NotImplementedError(f"{what}: every link of a connection needs the same delay here")
    return v


def region_cap(pop, window: float) -> int:
    """Spikes one group of 64 may emit in one window.  Neurons: 64 (W // t_ref + 1) -- a neuron fires at most
    once per refractory period -- unless the population sets history_cap (spikes per neuron and window).
    Poisson: the group's mean plus 10 standard deviations (the chance of more is below 1e-20 per window)."""
    override = pop.__dict__.get("history_cap")
    if override is not None:
        return max(1, min(int(override) * GROUP, 1 << 20))
    inst = pop.instance
    fam = pop.family
    if fam == "event_lif":
        per = int(window // inst.t_ref) + 1 if inst.t_ref > 0 else 16
        return min(GROUP * per, 1 << 20)
    if fam == "event_poisson":
        mean = GROUP * inst.rate * window
        return int(min(max(GROUP, math.ceil((mean + 10.0 * math.sqrt(mean) + 32) / GROUP) * GROUP), 1 << 20))
    # spike trains: the most any group lists in one window
    W = np.float32(window)
    most = 1
    times = inst.times
    for g0 in range(0, len(times), GROUP):
        ts = np.concatenate([np.asarray(t, np.float32) for t in times[g0:g0 + GROUP]] or [np.zeros(0, np.float32)])
        if ts.size:
            k = np.floor(ts.astype(np.float64) / np.float64(W)).astype(np.int64)
            most = max(most, int(np.bincount(k - k.min()).max()) + GROUP)  # (window edges rounded: one spare)
    return min(most, 1 << 20)


@dataclass
class EventPlan:
    """The host-side description of an event-driven network (no GPU needed): populations in creation order,
    the window, each connection's delay and lag, the input streams and the region capacities."""
    pops: list
    pop_first: np.ndarray
    window: float
    conn_delay: list
    conn_lag: list
    streams: list          # per stream: ((source population, delay, lag), ...) in connection order
    stream_of: dict        # neuron population -> its stream (-1: no inputs)
    caps: list             # region capacity per population

    @property
    def n(self) -> int:
        return int(self.pop_first[-1])


def describe(net) -> EventPlan:
    """Validate an event-driven network and plan its windows and streams (raises ValueError /
    NotImplementedError for what the event path does not run)."""
    pops = sorted(net.populations, key=lambda p: p._order)
    fams = {p.family for p in pops}
    if fams - set(EVENT_FAMILIES):
        raise NotImplementedError(f"event-driven and time-driven node types in one network: {sorted(fams)}")
    if len(pops) > _lib.MAX_POPS:
        raise ValueError(f"at most {_lib.MAX_POPS} populations")
    clocked = [p for p in pops if getattr(p, "timestep", None) is not None]
    if clocked:
        raise NotImplementedError(f"{clocked[0]!r}: event-driven populations run at their inputs' exact times, "
                                  f"not on a timestep of their own")
    sizes = [p.n for p in pops]
    if any(sz >= 1 << 24 for sz in sizes):
        raise ValueError("at most 2^24 - 1 neurons per population")
    pop_first = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    index = {p: i for i, p in enumerate(pops)}
    conns = list(net.connections)
    pairs, raw = set(), []
    for c in conns:
        if c._weight is None or c._delay is None:
            raise ValueError(f"{c!r}: weight and delay must be set")
        if c.target.family != "event_lif":
            raise ValueError(f"{c!r}: only event_lif populations receive inputs")
        if not c.plain:
            raise NotImplementedError(f"{c!r}: event-driven networks take static edges only (no edge kinds or ports)")
        key = (index[c.source], index[c.target])
        if key in pairs:
            raise NotImplementedError("at most one connection per (source, target) population pair")
        pairs.add(key)
        raw.append(float(np.float32(_uniform(c.delay, repr(c)))))
    # the window: the smallest delay; every connection's delay a whole number (its lag) of windows
    delays = [d for d in raw if not math.isnan(d)]
    W = float(min(delays)) if delays else float(np.float32(getattr(net, "window", 1.0)))
    conn_delay, conn_lag = [], []
    for c, d in zip(conns, raw):
        if math.isnan(d):
            d = W
        lag = int(round(d / W))
        if lag < 1 or lag > _lib.EV_MAX_LAG or abs(lag * W - d) > 1e-5 * d:
            raise NotImplementedError(f"{c!r}: delay {d} ms is not a whole number (<= {_lib.EV_MAX_LAG}) of "
                                      f"windows of {W} ms (the smallest delay)")
        conn_delay.append(d)
        conn_lag.append(lag)
    # input streams: one per distinct list of (source population, delay, lag), in connection order
    streams: list = []
    stream_of = {}
    for p in pops:
        if p.family != "event_lif":
            continue
        srcs_p = tuple((index[c.source], conn_delay[ci], conn_lag[ci]) for ci, c in enumerate(conns) if c.target is p)
        if not srcs_p:
            stream_of[p] = -1
            continue
        if len(srcs_p) > _lib.EV_MAX_SOURCES:
            raise NotImplementedError(f"at most {_lib.EV_MAX_SOURCES} source populations per target")
        if srcs_p not in streams:
            streams.append(srcs_p)
        stream_of[p] = streams.index(srcs_p)
    for p in pops:
        if p.family == "event_train" and len(p.instance.times) != p.n:
            raise ValueError(f"{p!r}: {len(p.instance.times)} spike trains for {p.n} sources")
    return EventPlan(pops, pop_first, W, conn_delay, conn_lag, streams, stream_of, [region_cap(p, W) for p in pops])


class EventRuntime:
    """The built event-driven network: device tensors, the C plan, captured windows, the drained record."""

    def __init__(self, net):
        self.net = net
        self.device = torch.device(net.device)
        if self.device.type != "cuda":
            raise RuntimeError("naer2_amd runs on the GPU: no CUDA/HIP device is available")
        dev = self.device
        plan = describe(net)
        pops = self.pops = plan.pops
        self.pop_first = plan.pop_first
        self.n = n = plan.n
        W = self.window = plan.window
        self.conn_delay, self.conn_lag = plan.conn_delay, plan.conn_lag
        sizes = [p.n for p in pops]
        first_of = {p: int(f) for p, f in zip(pops, self.pop_first[:-1])}
        conns = list(net.connections)
        srcs = [c.src.to(dev, torch.int64) + first_of[c.source] for c in conns]
        tgts = [c.tgt.to(dev, torch.int64) + first_of[c.target] for c in conns]
        ws = [c.weight.to(dev) for c in conns]
        cat = (lambda xs, t: torch.cat(xs) if xs else torch.zeros(0, dtype=t, device=dev))
        e_all = cat(srcs, torch.int64)
        canon = canonical(n, e_all, cat(tgts, torch.int64), cat(ws, torch.float32),
                          torch.ones(e_all.numel(), dtype=torch.int32, device=dev))
        del srcs, tgts, ws, e_all
        if not net.keep_description:
            for c in conns:
                c.release()
        self.layout = group_layout(sizes, dev)
        self.n_groups = self.layout.n_groups
        seg = segments(canon, self.layout)
        self.seg_off, self.links = seg.seg_off, seg.links
        self.n_links = canon.n_links
        self.out_degree = canon.out_degree()
        self.canon = canon if net.keep_canonical else None
        del seg
        self.streams = plan.streams
        stream_of = plan.stream_of
        self.caps = plan.caps
        # state
        self.v = torch.zeros(n, dtype=torch.float32, device=dev)
        self.last = torch.zeros(n, dtype=torch.float32, device=dev)
        self.tre = torch.zeros(n, dtype=torch.float32, device=dev)
        self.nxt = torch.full((n,), -1.0, dtype=torch.float32, device=dev)
        for p, lo in zip(pops, self.pop_first[:-1]):
            sl = slice(int(lo), int(lo) + p.n)
            for name, t in (("V", self.v), ("last_time", self.last), ("t_ref_end", self.tre),
                            ("next_spike", self.nxt)):
                if name in p._init:
                    t[sl] = p._init[name].to(dev, torch.float32)
                elif hasattr(p.instance, name) and name in getattr(p.instance, "state", ()):
                    t[sl] = float(getattr(p.instance, name))
        # spike trains
        self.train_ptr = self.train_t = None
        if any(p.family == "event_train" for p in pops):
            counts = np.zeros(n, np.int64)
            chunks = []
            for p, lo in zip(pops, self.pop_first[:-1]):
                if p.family != "event_train":
                    continue
                if len(p.instance.times) != p.n:
                    raise ValueError(f"{p!r}: {len(p.instance.times)} spike trains for {p.n} sources")
                for i, ts in enumerate(p.instance.times):
                    counts[int(lo) + i] = len(ts)
            ptr = np.zeros(n + 1, np.int64)
            ptr[1:] = np.cumsum(counts)
            flat = np.zeros(int(ptr[-1]), np.float32)
            for p, lo in zip(pops, self.pop_first[:-1]):
                if p.family == "event_train":
                    for i, ts in enumerate(p.instance.times):
                        flat[ptr[int(lo) + i]:ptr[int(lo) + i + 1]] = np.asarray(ts, np.float32)
            del chunks
            self.train_ptr = torch.from_numpy(ptr).to(dev)
            self.train_t = torch.from_numpy(flat).to(dev)
        self._stream_of = stream_of
        self._log = None
        self._graphs: dict = {}
        self._plan = None
        self._create_plan()
        # deliver-kernel mapping: runs of 16 consecutive groups per XCD (8 XCDs on MI355X): neighbouring groups
        # share their segments' cache lines in one L2 (measured 2-4 % faster than one block per group in order)
        self.set_xcd(8)
        self.record_capacity = 0
        self._rec = None
        self._rec_nodes: list = []
        self._rec_times: list = []
        self._rec_windows: list = []
        self.window_index = 0
        self.record_from = 0
        for p, lo in zip(pops, self.pop_first[:-1]):
            object.__setattr__(p, "_net_slice", (self, int(lo)))

    # -- the C plan
    def _create_plan(self) -> None:
        lib = _lib.load()
        P = len(self.pops)
        pops = (_lib.EvPop * P)()
        for i, p in enumerate(self.pops):
            e = pops[i]
            e.kind = EVENT_FAMILIES[p.family]
            e.region_cap = self.caps[i]
            e.stream = -1
            if p.family == "event_lif":
                inst = p.instance
                e.tau_inv = float(np.float32(1.0 / inst.tau))
                e.theta, e.v_reset, e.t_ref = inst.theta, inst.V_r, inst.t_ref
                e.stream = self._stream_of[p]
            elif p.family == "event_poisson":
                e.rate_inv = float(np.float32(1.0 / p.instance.rate)) if p.instance.rate > 0 else float("inf")
        streams = (_lib.EvStream * max(len(self.streams), 1))()
        for i, srcs in enumerate(self.streams):
            streams[i].n_sources = len(srcs)
            for j, (q, d, lag) in enumerate(srcs):
                streams[i].source_pop[j] = q
                streams[i].delay[j] = d
                streams[i].lag[j] = lag
        first = (ctypes.c_int32 * (P + 1))(*[int(x) for x in self.pop_first])
        d = _lib.EvDesc(n_neurons=self.n, n_pops=P, n_groups=self.n_groups, n_streams=len(self.streams),
                        pop_first=first, pops=pops, streams=streams, window=self.window,
                        instance=int(self.net.instance) & 0xFFFFFFFF, seed=int(self.net.seed) & (2 ** 64 - 1),
                        group_first=self.layout.group_first.data_ptr(), group_info=self.layout.group_info.data_ptr(),
                        seg_off=self.seg_off.data_ptr(), links=self.links.data_ptr(), n_links=self.n_links,
                        train_ptr=self.train_ptr.data_ptr() if self.train_ptr is not None else None,
                        train_t=self.train_t.data_ptr() if self.train_t is not None else None)
        plan = ctypes.c_void_p()
        _lib.ev_check(lib.n2h_ev_create(ctypes.byref(d), ctypes.byref(plan)), "n2h_ev_create")
        self._plan = plan
        _lib.ev_check(lib.n2h_ev_bind_state(plan, self.v.data_ptr(), self.last.data_ptr(), self.tre.data_ptr(),
                                            self.nxt.data_ptr()), "n2h_ev_bind_state")

    def __del__(self):
        try:
            lib = _lib.load()
            for g in self._graphs.values():
                lib.n2h_ev_graph_destroy(g)
            if self._plan is not None:
                lib.n2h_ev_destroy(self._plan)
        except Exception:
            pass

    @property
    def spikes_bound(self) -> int:
        """Spikes one window can emit at most (every group's region full)."""
        return int(sum(c * g for c, g in zip(self.caps, self._pop_groups())))

    def _pop_groups(self) -> list:
        return [(p.n + GROUP - 1) // GROUP for p in self.pops]

    def _bind_record(self, windows: int) -> None:
        """Record buffers for `windows` windows at the worst case (every region full)."""
        lib = _lib.load()
        cap = max(self.spikes_bound * windows, 1)
        dev = self.device
        if self._rec is None or self._rec[0].numel() < cap or self._rec[3].numel() < windows:
            node = torch.zeros(cap, dtype=torch.int32, device=dev)
            time = torch.zeros(cap, dtype=torch.float32, device=dev)
            count = torch.zeros(1, dtype=torch.int32, device=dev)
            wcount = torch.zeros(windows, dtype=torch.int32, device=dev)
            self._rec = (node, time, count, wcount)
            _lib.ev_check(lib.n2h_ev_bind_record(self._plan, node.data_ptr(), time.data_ptr(), count.data_ptr(), cap,
                                                 wcount.data_ptr(), windows), "n2h_ev_bind_record")
            self._drop_graphs()

    def _unbind_record(self) -> None:
        if self._rec is not None:
            _lib.ev_check(_lib.load().n2h_ev_bind_record(self._plan, None, None, None, 0, None, 1),
                          "n2h_ev_bind_record")
            self._rec = None
            self._drop_graphs()

    def _drop_graphs(self) -> None:
        lib = _lib.load()
        for g in self._graphs.values():
            lib.n2h_ev_graph_destroy(g)
        self._graphs = {}

    def _drain(self, k0: int, windows: int) -> None:
        node, time, count, wcount = self._rec
        torch.cuda.synchronize(self.device)
        c = int(count.item())
        slots = wcount.numel()
        self._rec_nodes.append(node[:c].cpu().numpy().copy())
        self._rec_times.append(time[:c].cpu().numpy().copy())
        wc = wcount.cpu().numpy()
        self._rec_windows.extend(int(wc[(k0 + i) % slots]) for i in range(windows))
        count.zero_()

    def _launch(self, n: int, graph_window: int, stream) -> None:
        lib = _lib.load()
        done = 0
        if graph_window and n >= graph_window:
            g = self._graphs.get(graph_window)
            if g is None:
                gp = ctypes.c_void_p()
                _lib.ev_check(lib.n2h_ev_graph_create(self._plan, graph_window, stream, ctypes.byref(gp)),
                              "n2h_ev_graph_create")
                g = self._graphs[graph_window] = gp
            while n - done >= graph_window:
                _lib.ev_check(lib.n2h_ev_graph_launch(g, stream), "n2h_ev_graph_launch")
                done += graph_window
        if n > done:
            _lib.ev_check(lib.n2h_ev_run(self._plan, n - done, stream), "n2h_ev_run")

    def run(self, n_windows: int, graph_window: int = 0) -> None:
        """Advance by n_windows windows.  With recording on, the spikes are copied to the host every chunk
        of windows (a synchronization each)."""
        if n_windows < 0:
            raise ValueError("n_windows must be >= 0")
        if n_windows == 0:
            return
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        if not self.net.record:
            self._unbind_record()
            self._launch(n_windows, graph_window, stream)
            self.window_index += n_windows
            self.record_from = self.window_index
            return
        gw = graph_window if 0 < graph_window <= 200 else 0
        chunk = min(n_windows, (200 // gw) * gw if gw else 200)
        self._bind_record(chunk)
        done = 0
        while done < n_windows:
            w = min(chunk, n_windows - done)
            self._launch(w, graph_window if w >= graph_window else 0, stream)
            self._drain(self.window_index, w)
            self.window_index += w
            done += w

    @property
    def max_lag(self) -> int:
        return max(self.conn_lag) if self.conn_lag else 1

    def advance(self, t_ms: float) -> None:
        """Advance by t_ms, a whole number of windows."""
        k = t_ms / self.window
        if abs(k - round(k)) > 1e-6 * max(1.0, abs(k)):
            raise ValueError(f"{t_ms} ms is not a whole number of windows of {self.window} ms")
        self.run(int(round(k)), self.net.graph_window)

    def log_deliveries(self, capacity: int | None) -> None:
        """Log every delivered input of the windows that follow (the reference's delivery log); None: off.
        deliveries() reads it."""
        lib = _lib.load()
        self._drop_graphs()
        if capacity is None:
            _lib.ev_check(lib.n2h_ev_bind_log(self._plan, None, None, None, None, 0), "n2h_ev_bind_log")
            self._log = None
            return
        P = len(self.pops)
        index = {p: i for i, p in enumerate(self.pops)}
        conn = np.full(P * P, -1, np.int32)
        for ci, c in enumerate(self.net.connections):
            conn[index[c.target] * P + index[c.source]] = ci
        rec = torch.zeros(max(int(capacity), 1), 4, dtype=torch.int32, device=self.device)
        arr = torch.zeros(max(int(capacity), 1), dtype=torch.float32, device=self.device)
        cnt = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._log = (rec, arr, cnt, conn)
        _lib.ev_check(lib.n2h_ev_bind_log(self._plan, conn.ctypes.data_as(ctypes.c_void_p), rec.data_ptr(),
                                          arr.data_ptr(), cnt.data_ptr(), int(capacity)), "n2h_ev_bind_log")

    def deliveries(self) -> list[tuple[int, float, int, int, int]]:
        """The delivery log since log_deliveries(): (window, arrival, connection, source, target), source and
        target as indices in their populations, per target in delivery order (naer2/runtime.py:747-752)."""
        if self._log is None:
            raise RuntimeError("no delivery log bound (log_deliveries)")
        rec, arr, cnt, _ = self._log
        torch.cuda.synchronize(self.device)
        n = int(cnt.item())
        if n > rec.shape[0]:
            raise RuntimeError(f"the delivery log overflowed ({n} inputs, capacity {rec.shape[0]})")
        r = rec[:n].cpu().tolist()
        a = arr[:n].cpu().tolist()
        return [(k, ai, c, s, t) for (k, c, s, t), ai in zip(r, a)]

    def set_xcd(self, k: int) -> None:
        """Deliver-kernel mapping: 0 one block per group in order; k > 1 runs of 16 consecutive groups per
        XCD of k (results do not change)."""
        t = _lib.Tuning(order=0, xcd=int(k))
        _lib.ev_check(_lib.load().n2h_ev_set_tuning(self._plan, ctypes.byref(t)), "n2h_ev_set_tuning")
        self.xcd = int(k)
        self._drop_graphs()

    def counters(self) -> tuple[int, int]:
        """(inputs delivered, spikes emitted) since the build (synchronizes)."""
        c = (ctypes.c_int64 * 2)()
        _lib.ev_check(_lib.load().n2h_ev_counters(self._plan, c, 2), "n2h_ev_counters")
        return int(c[0]), int(c[1])

    def run_timed(self, n_windows: int) -> tuple[float, float]:
        """Run n_windows eager windows with HIP events around each kernel: (deliver ms, sort ms) summed.
        Recording is switched off for them."""
        self._unbind_record()
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        d, s = ctypes.c_double(), ctypes.c_double()
        _lib.ev_check(_lib.load().n2h_ev_run_timed(self._plan, int(n_windows), stream, ctypes.byref(d),
                                                   ctypes.byref(s)), "n2h_ev_run_timed")
        self.window_index += n_windows
        self.record_from = self.window_index
        return float(d.value), float(s.value)

    TRACE_PHASES = ("entries", "staging", "inbox places", "inbox fill", "walk")
    SORT_EVENTS = ("start", "group prefix", "record", "first stream sorted", "streams written", "first stream keyed",
                   "first stream binned", "end")

    def trace(self, n_windows: int = 1) -> dict:
        """Run n_windows eager windows with the in-kernel phase trace on; for the last one: per phase of the
        neuron groups the mean and max over groups of its summed duration (us), the rounds, the groups'
        start spread and their longest run, and the sort kernel's phase durations (us)."""
        lib = _lib.load()
        buf = torch.zeros(self.n_groups + 1, 8, dtype=torch.int64, device=self.device)
        self._unbind_record()
        self._drop_graphs()
        _lib.ev_check(lib.n2h_ev_set_trace(self._plan, buf.data_ptr()), "n2h_ev_set_trace")
        try:
            self.run(n_windows, 0)
            torch.cuda.synchronize(self.device)
        finally:
            _lib.ev_check(lib.n2h_ev_set_trace(self._plan, None), "n2h_ev_set_trace")
        t = buf.cpu().numpy().astype(np.int64)
        rows = t[:-1]
        used = rows[:, 7] > 0
        out = {}
        for j, name in enumerate(self.TRACE_PHASES):
            d = rows[used, 1 + j] / 100.0
            out[name] = (float(d.mean()), float(d.max())) if d.size else (0.0, 0.0)
        if used.any():
            out["rounds"] = (float(rows[used, 6].mean()), int(rows[used, 6].max()))
            out["group starts spread"] = float((rows[used, 0].max() - rows[used, 0].min()) / 100.0)
            out["group run"] = (float(((rows[used, 7] - rows[used, 0]) / 100.0).mean()),
                                float(((rows[used, 7] - rows[used, 0]) / 100.0).max()))
        srow = t[-1]
        out["sort"] = {f"{self.SORT_EVENTS[j - 1]} -> {self.SORT_EVENTS[j]}": float((srow[j] - srow[j - 1]) / 100.0)
                       for j in (1, 2, 3, 4)}
        out["sort"]["streams written -> end"] = float((srow[7] - srow[4]) / 100.0)
        if srow[5] and srow[6]:  # (the first stream's keys, then its binning, inside record -> sorted)
            out["sort"]["record -> first stream keyed"] = float((srow[5] - srow[2]) / 100.0)
            out["sort"]["keyed -> binned"] = float((srow[6] - srow[5]) / 100.0)
            out["sort"]["binned -> first stream sorted"] = float((srow[3] - srow[6]) / 100.0)
        if used.any():
            out["deliver end -> sort start"] = float((srow[0] - rows[used, 7].max()) / 100.0)
        return out

    def check(self) -> None:
        """Raise if a kernel flagged lost spikes or links (the reference's check(), runtime.py:724-735)."""
        flags = (ctypes.c_int32 * _lib.EV_FLAGS.__len__())()
        _lib.ev_check(_lib.load().n2h_ev_status(self._plan, flags, len(_lib.EV_FLAGS)), "n2h_ev_status")
        bad = [name for name, f in zip(_lib.EV_FLAGS, flags) if f]
        if bad:
            raise RuntimeError(f"event-driven run lost data: {', '.join(bad)} "
                               f"(region: raise Population.history_cap; segment: at most 2048 links from one "
                               f"source into one group of 64)")

    # -- state
    def _tensor(self, name: str) -> torch.Tensor:
        return {"V": self.v, "last_time": self.last, "t_ref_end": self.tre, "next_spike": self.nxt}[name]

    def _read_state(self, pop, name):
        lo = dict(zip(self.pops, self.pop_first[:-1]))[pop]
        return self._tensor(name)[int(lo):int(lo) + pop.n]

    def _write_state(self, pop, name, value):
        self._read_state(pop, name).copy_(value.to(self.device))

    # -- spikes
    def spikes(self, pop=None) -> tuple[np.ndarray, np.ndarray]:
        """(times in ms, neuron ids) of every recorded spike, window by window, each window in (population,
        time, neuron) order; ids local to pop when given."""
        nodes = np.concatenate(self._rec_nodes) if self._rec_nodes else np.zeros(0, np.int32)
        times = np.concatenate(self._rec_times) if self._rec_times else np.zeros(0, np.float32)
        if pop is not None:
            first = int(dict(zip(self.pops, self.pop_first[:-1]))[pop])
            m = (nodes >= first) & (nodes < first + pop.n)
            return times[m], nodes[m] - first
        return times, nodes

    def window_counts(self) -> np.ndarray:
        return np.asarray(self._rec_windows, np.int64)

    def rates(self, t_skip_ms: float = 0.0) -> dict:
        times, nodes = self.spikes()
        t_end = self.window_index * self.window
        dur = (t_end - max(t_skip_ms, self.record_from * self.window)) / 1000.0
        out = {}
        keep = times >= t_skip_ms
        for p, lo in zip(self.pops, self.pop_first[:-1]):
            m = keep & (nodes >= lo) & (nodes < lo + p.n)
            out[p] = float(m.sum()) / p.n / dur if dur > 0 else float("nan")
        return out
