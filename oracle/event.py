"""TEST INFRASTRUCTURE ONLY: the event-driven, exact-time oracle (oracle/event_oracle.c through ctypes), its
fp32 math restated in NumPy, and an independent global event-queue simulation that pins it.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the product
(naer2_amd) never does.

Reference behaviour restated (MatteoSchmider/naer2): test/networks/Brunel/neuron.py:6-45 (the neuron),
test/networks/Brunel/poisson.py:7-27 (the source), naer2/transpiler/templates/fused.py:1300-1370 and
1421-1526 (merge of the connections' queues by arrival, ties to the lower connection; per target the inputs
at one time received one by one, then send once).  The contract is spelled out in event_oracle.h.
"""
from __future__ import annotations

import ctypes
import heapq
from dataclasses import dataclass, field

import numpy as np

from . import oracle as _o

NEURON, POISSON, TRAIN = 0, 1, 2
F32 = np.float32


class EvPop(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("tau_inv", ctypes.c_float),
                ("theta", ctypes.c_float), ("v_reset", ctypes.c_float), ("t_ref", ctypes.c_float),
                ("rate_inv", ctypes.c_float), ("reserved_f", ctypes.c_float)]


_ready = False


def _lib():
    global _ready
    lib = _o.load()
    if not _ready:
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        lib.or_ev_exp.argtypes = [ctypes.c_float]
        lib.or_ev_exp.restype = ctypes.c_float
        lib.or_ev_log.argtypes = [ctypes.c_float]
        lib.or_ev_log.restype = ctypes.c_float
        lib.or_ev_interval.argtypes = [ctypes.c_uint32, i64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_float]
        lib.or_ev_interval.restype = ctypes.c_float
        lib.or_ev_run.argtypes = [i32, i32, P, P, i32, P, P, P, P, ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32,
                                  P, P, P, P, P, P, P, P, P, P, P, P, P, i64, i64, i32, P, P, i64, P, P, P, P, P, i64,
                                  P]
        lib.or_ev_run.restype = ctypes.c_int
        _ready = True
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class EvNet:
    """A network description in the oracle's terms (host numpy arrays)."""
    pop_first: np.ndarray           # int32 [P + 1]
    pops: list                      # EvPop per population
    conn_src: np.ndarray            # int32 [C]: source population of each connection (creation order)
    conn_tgt: np.ndarray            # int32 [C]
    window: float                   # the smallest delay, ms (every delay a whole multiple of it)
    src_ptr: np.ndarray             # int64 [n + 1]: canonical CSR by source
    tgt: np.ndarray                 # int32 [E]
    w: np.ndarray                   # float32 [E]
    seed: int = 0
    instance: int = 0
    train_ptr: np.ndarray | None = None  # int64 [n + 1]
    train_t: np.ndarray | None = None    # float32
    conn_delay: np.ndarray | None = None  # float32 [C]: each connection's delay (None: all the window)

    @property
    def n(self) -> int:
        return int(self.pop_first[-1])

    def delays(self) -> tuple[np.ndarray, np.ndarray]:
        """(delay fp32, lag = delay / window) per connection."""
        d = (np.full(len(self.conn_src), self.window, np.float32) if self.conn_delay is None
             else np.ascontiguousarray(self.conn_delay, np.float32))
        lag = np.rint(d.astype(np.float64) / np.float64(np.float32(self.window))).astype(np.int32)
        return d, lag


@dataclass
class EvState:
    v: np.ndarray
    last: np.ndarray
    tre: np.ndarray
    nxt: np.ndarray
    prev_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    prev_e: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    window: int = 0
    prev_win: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    @staticmethod
    def initial(net: EvNet, v0) -> "EvState":
        n = net.n
        return EvState(np.ascontiguousarray(v0, dtype=np.float32).copy(), np.zeros(n, np.float32),
                       np.zeros(n, np.float32), np.full(n, -1.0, np.float32))

    def copy(self) -> "EvState":
        return EvState(self.v.copy(), self.last.copy(), self.tre.copy(), self.nxt.copy(), self.prev_node.copy(),
                       self.prev_e.copy(), self.window, self.prev_win.copy())


def run(net: EvNet, st: EvState, n_windows: int, record: bool = True, log: list | None = None,
        log_cap: int = 0):
    """Advance st by n_windows windows (in place).  Returns (nodes, times, per-window counts, events).  log: a
    list that receives the delivery log (window, arrival, connection, source, target), per target in order."""
    lib = _lib()
    n = net.n
    pops = (EvPop * len(net.pops))(*net.pops)
    pf = np.ascontiguousarray(net.pop_first, np.int32)
    cs = np.ascontiguousarray(net.conn_src, np.int32)
    ct = np.ascontiguousarray(net.conn_tgt, np.int32)
    sp = np.ascontiguousarray(net.src_ptr, np.int64)
    tg = np.ascontiguousarray(net.tgt, np.int32)
    w = np.ascontiguousarray(net.w, np.float32)
    tp = None if net.train_ptr is None else np.ascontiguousarray(net.train_ptr, np.int64)
    tt = None if net.train_t is None else np.ascontiguousarray(net.train_t, np.float32)
    dly, lag = net.delays()
    M = int(lag.max()) if lag.size else 1
    cap = max(n * 8, 1024) + (0 if tt is None else int(tt.size))
    pcap = cap * M
    prev_node = np.zeros(pcap, np.int32)
    prev_e = np.zeros(pcap, np.float32)
    prev_win = np.zeros(pcap, np.int64)
    prev_node[:st.prev_node.size] = st.prev_node
    prev_e[:st.prev_e.size] = st.prev_e
    if st.prev_win.size == st.prev_node.size:
        prev_win[:st.prev_win.size] = st.prev_win
    else:  # (the spikes of the window before: a state made by hand)
        prev_win[:st.prev_node.size] = st.window - 1
    prev_n = ctypes.c_int64(st.prev_node.size)
    rec_cap = cap * max(n_windows, 1) if record else 0
    rec_node = np.zeros(rec_cap, np.int32) if record else None
    rec_e = np.zeros(rec_cap, np.float32) if record else None
    wcount = np.zeros(max(n_windows, 1), np.int32)
    rec_n = ctypes.c_int64(0)
    events = ctypes.c_int64(0)
    log_rec = np.zeros((max(log_cap, 1), 4), np.int32) if log is not None else None
    log_a = np.zeros(max(log_cap, 1), np.float32) if log is not None else None
    log_n = ctypes.c_int64(0)
    rc = lib.or_ev_run(n, len(net.pops), _p(pf), pops, len(cs), _p(cs), _p(ct), _p(dly), _p(lag),
                       ctypes.c_float(net.window), ctypes.c_uint64(net.seed), ctypes.c_uint32(net.instance), _p(sp),
                       _p(tg), _p(w), _p(tp), _p(tt), _p(st.v), _p(st.last), _p(st.tre), _p(st.nxt),
                       ctypes.byref(prev_n), _p(prev_node), _p(prev_e), _p(prev_win), pcap, st.window, n_windows,
                       _p(rec_node), _p(rec_e), rec_cap, ctypes.byref(rec_n),
                       _p(wcount), ctypes.byref(events), _p(log_rec), _p(log_a), int(log_cap), ctypes.byref(log_n))
    if rc != 0:
        raise RuntimeError(f"or_ev_run failed ({rc})")
    if log is not None:
        if log_n.value > log_cap:
            raise RuntimeError(f"the delivery log overflowed ({log_n.value} > {log_cap})")
        for (k, c, s_, t), a in zip(log_rec[:log_n.value].tolist(), log_a[:log_n.value].tolist()):
            log.append((k, a, c, s_, t))
    st.prev_node = prev_node[:prev_n.value].copy()
    st.prev_e = prev_e[:prev_n.value].copy()
    st.prev_win = prev_win[:prev_n.value].copy()
    st.window += n_windows
    k = rec_n.value
    nodes = rec_node[:k].copy() if record else None
    times = rec_e[:k].copy() if record else None
    return nodes, times, wcount[:n_windows].copy(), int(events.value)


# ---------------------------------------------------------------- the contract's fp32 math in NumPy

def ev_exp(x) -> np.float32:
    """The same operations as or_ev_exp, in float32 scalars."""
    x = F32(x)
    if x < F32(-87.0):
        return F32(0.0)
    n = np.rint(F32(x * F32(1.44269502)))
    r = F32(x - F32(n * F32(0.693145751953125)))
    r = F32(r - F32(n * F32(1.42860677e-06)))
    r2 = F32(r * r)
    r4 = F32(r2 * r2)
    q01 = F32(F32(1.0) + F32(F32(1.0) * r))
    q23 = F32(F32(0.5) + F32(F32(1.66666672e-01) * r))
    q45 = F32(F32(4.16666679e-02) + F32(F32(8.33333377e-03) * r))
    q67 = F32(F32(1.38888892e-03) + F32(F32(1.98412701e-04) * r))
    q03 = F32(q01 + F32(r2 * q23))
    q47 = F32(q45 + F32(r2 * q67))
    return F32(np.ldexp(F32(q03 + F32(r4 * q47)), int(n)))


def ev_log(u) -> np.float32:
    """The same operations as or_ev_log, in float32 scalars."""
    b = int(np.array(u, np.float32).view(np.uint32))
    e = ((b >> 23) & 0xFF) - 127
    m = np.array((b & 0x7FFFFF) | 0x3F800000, np.uint32).view(np.float32)[()]
    if m > F32(1.41421354):
        m = F32(m * F32(0.5))
        e += 1
    t = F32(m - F32(1.0))
    p = F32(0.0674023256)
    for c in (-0.117272295, 0.119239263, -0.123992212, 0.142126068, -0.166680425, 0.200025767, -0.250000238,
              0.333333015, -0.5, 1.0):
        p = F32(F32(p * t) + F32(c))
    fe = F32(e)
    r = F32(t * p)
    r = F32(r + F32(fe * F32(1.42860677e-06)))
    return F32(r + F32(fe * F32(0.693145751953125)))


def interval(node: int, window: int, m: int, seed: int, instance: int, rate_inv: float) -> np.float32:
    return F32(_lib().or_ev_interval(node, window, m, seed, instance, rate_inv))


# ---------------------------------------------------------------- an independent formulation: one global queue

def heap_run(net: EvNet, v0, n_windows: int, log: list | None = None):
    """The same network as ONE global event queue instead of windows: every input is a heap entry keyed
    (arrival, connection, source, link); a neuron's pending threshold check runs as soon as the
    queue's next arrival is later than its last input (before that arrival is taken), so its spike's inputs
    enter the queue in time.  Sources are generated window by window as in the contract; spikes emitted in
    window j are delivered only when j + 1 < n_windows (what the windowed run covers).  Pure Python: tiny
    networks only.  Returns (nodes, times) sorted by (population, time, node) within each window, and
    per-window counts; log (a list) receives every delivered input as (window, arrival, connection, source,
    target) in global delivery order."""
    n = net.n
    pf = np.asarray(net.pop_first, np.int64)
    pop_of = np.searchsorted(pf, np.arange(n), side="right") - 1
    conn_of = {(int(s), int(t)): c for c, (s, t) in enumerate(zip(net.conn_src, net.conn_tgt))}
    W = F32(net.window)
    dly, lag = net.delays()
    v = np.asarray(v0, np.float32).copy()
    last = np.zeros(n, np.float32)
    tre = np.zeros(n, np.float32)
    spikes: list[list] = [[] for _ in range(n_windows)]
    heap: list = []

    def emit(node: int, e, win: int) -> None:
        spikes[win].append((int(pop_of[node]), float(e), node))
        for link in range(int(net.src_ptr[node]), int(net.src_ptr[node + 1])):
            t = int(net.tgt[link])
            c = conn_of[(int(pop_of[node]), int(pop_of[t]))]
            if win + int(lag[c]) >= n_windows:
                continue
            a = F32(F32(e) + F32(dly[c]))
            heapq.heappush(heap, (float(a), c, node, link, t, win + int(lag[c])))

    for k in range(n_windows):
        t_start = F32(np.float64(k) * np.float64(W))
        t_end = F32(np.float64(k + 1) * np.float64(W))
        for p, P in enumerate(net.pops):
            lo, hi = int(pf[p]), int(pf[p + 1])
            if P.kind == TRAIN:
                for i in range(lo, hi):
                    for j in range(int(net.train_ptr[i]), int(net.train_ptr[i + 1])):
                        t = F32(net.train_t[j])
                        if t_start <= t < t_end:
                            emit(i, t, k)
    for p, P in enumerate(net.pops):
        if P.kind != POISSON:
            continue
        for i in range(int(pf[p]), int(pf[p + 1])):
            nxt = F32(-1.0)
            for k in range(n_windows):
                t_end = F32(np.float64(k + 1) * np.float64(W))
                m = 0
                if nxt < 0:
                    nxt = interval(i, k, m, net.seed, net.instance, P.rate_inv)
                    m += 1
                while nxt < t_end:
                    emit(i, nxt, k)
                    nxt = F32(nxt + interval(i, k, m, net.seed, net.instance, P.rate_inv))
                    m += 1

    pending: dict[int, tuple] = {}  # neuron -> (its last input's arrival, that input's window)

    def flush() -> None:
        for i, (t, win) in sorted(pending.items()):
            P = net.pops[int(pop_of[i])]
            if v[i] >= F32(P.theta):
                v[i] = F32(P.v_reset)
                tre[i] = F32(t + F32(P.t_ref))
                last[i] = tre[i]
                emit(i, t, win)
        pending.clear()

    while heap:
        a = F32(heap[0][0])
        if pending and any(a > t for t, _ in pending.values()):
            # every pending check is at the current time (all earlier ones were flushed): run them first
            flush()
            continue
        a_, c, node, link, t, win = heapq.heappop(heap)
        P = net.pops[int(pop_of[t])]
        if log is not None:
            log.append((win, float(a), c, node - int(pf[pop_of[node]]), t - int(pf[pop_of[t]])))
        if a >= tre[t]:
            v[t] = F32(F32(v[t] * ev_exp(F32(F32(last[t] - a) * F32(P.tau_inv)))) + F32(net.w[link]))
            last[t] = a
        pending[t] = (a, win)
    flush()
    nodes, times, counts = [], [], []
    for lst in spikes:
        lst.sort()
        counts.append(len(lst))
        nodes += [s[2] for s in lst]
        times += [s[1] for s in lst]
    return np.asarray(nodes, np.int32), np.asarray(times, np.float32), np.asarray(counts, np.int32), v


# ---------------------------------------------------------------- builders for tests

def pops_from(kinds_params) -> list:
    """[(kind, dict)] -> EvPop list.  Neuron: tau, theta, v_reset, t_ref; Poisson: rate (per ms)."""
    out = []
    for kind, prm in kinds_params:
        p = EvPop()
        p.kind = kind
        if kind == NEURON:
            p.tau_inv = float(F32(1.0 / prm["tau"]))
            p.theta, p.v_reset, p.t_ref = prm["theta"], prm["v_reset"], prm["t_ref"]
        elif kind == POISSON:
            p.rate_inv = float(F32(1.0 / prm["rate"])) if prm["rate"] > 0 else float("inf")
        out.append(p)
    return out
