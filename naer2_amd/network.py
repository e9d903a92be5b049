"""Network: the populations and connections of one model, built into one plan of libnaer2_hip.so and
stepped on the GPU.

Reference: naer2/network.py:17-152 (Network, step), naer2/runtime.py:213-526 (PopulationRuntime: state,
ring, history, launch), naer2/graph.py:35-136 (captured windows).  The reference schedules every
population separately (one kernel and one stream per population and round, or one fused kernel per round,
naer2/fusion.py).  Every network here is time-driven on one clock dt (`resolution`), so one step of the
whole network is one fused kernel launch over all target groups of all populations; windows of steps are
captured into hipGraphs and replayed.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import GROUP, MODEL_LIF_DELTA, MODEL_LIF_EXP
from .topology import Canonical, GroupLayout, canonical, delay_steps, group_layout, rows, segments

_FAMILY_MODEL = {"lif_delta": MODEL_LIF_DELTA, "lif_exp": MODEL_LIF_EXP}
_EPS = 1e-9


def default_device() -> torch.device:
    return torch.device("cuda" if torch.cuda.is_available() else "cpu")


class Network:
    def __init__(self, population=None):
        self.populations: list = [population] if population is not None else []
        self.connections: list = []
        self.resolution: float = 0.1      # dt, ms
        self.seed: int = 0                # connectivity draws and the Philox key
        self.instance: int = 0            # Philox counter word: distinct streams for batched copies
        self.device: torch.device = default_device()
        self.graph_window: int = 100      # steps per captured hipGraph in run() (0: no graphs)
        self.record: bool = True          # keep every step's spike bitmap (the raster)
        self.keep_canonical: bool = True  # keep the canonical synapse table on the device after the build
        # row intake (target-major pull over every in-link, for steps after many spikes): "auto" builds its
        # table where the spike bitmap fits in LDS (<= 1535 groups); dense_min_spikes None: n_neurons // 20
        self.row_intake = "auto"
        self.row_format = "auto"          # "auto" (slot16, else compact where it fits), "wide", "compact", "slot16"
        self.dense_min_spikes: int | None = None
        # keep the connections' per-link tensors after the build (False: release them -- the runtime holds
        # its own tables; connection weights and delays can then no longer be read or changed)
        self.keep_description: bool = True
        self.t: float = 0.0
        self._generator = None
        self._fuse_rounds = True
        self._rt: Runtime | None = None

    # -- composition (naer2/network.py:60-66)
    def merge(self, other: "Network") -> "Network":
        if other is self:
            return self
        for pop in other.populations:
            object.__setattr__(pop, "network", self)
            if pop not in self.populations:
                self.populations.append(pop)
        self.connections += [c for c in other.connections if c not in self.connections]
        self._invalidate()
        return self

    __add__ = __iadd__ = merge

    @property
    def generator(self) -> torch.Generator:
        if self._generator is None:
            self._generator = torch.Generator(device=self.device).manual_seed(int(self.seed))
        return self._generator

    @property
    def fuse_rounds(self) -> bool:
        """Kept from naer2/network.py:42-51: here every step is already one launch for all populations."""
        return self._fuse_rounds

    @fuse_rounds.setter
    def fuse_rounds(self, on: bool) -> None:
        self._fuse_rounds = bool(on)

    def _invalidate(self) -> None:
        if self._rt is not None:
            raise RuntimeError("the network is built: change populations and connections before its first step")

    # -- build and run
    @property
    def runtime(self) -> "Runtime":
        if self._rt is None:
            from .event import EventRuntime, is_event_net
            self._rt = EventRuntime(self) if is_event_net(self) else Runtime(self)
        return self._rt

    @property
    def event_driven(self) -> bool:
        """True for networks of the event-driven, exact-time node types (naer2_amd/event.py)."""
        from .event import is_event_net
        return is_event_net(self)

    def build(self) -> "Runtime":
        return self.runtime

    @property
    def step_index(self) -> int:
        return self._rt.step if self._rt is not None else 0

    def step(self, t_step: float) -> None:
        """Advance the network by t_step ms: a whole number of steps of `resolution`, or for an event-driven
        network a whole number of windows (its common delay)."""
        if self.event_driven:
            rt = self.runtime
            rt.advance(t_step)
            self.t = rt.window_index * rt.window
            return
        n = t_step / self.resolution
        if abs(n - round(n)) > 1e-6 * max(1.0, abs(n)):
            raise ValueError(f"t_step {t_step} is not a whole number of steps of {self.resolution} ms")
        self.run(int(round(n)))

    def run(self, n_steps: int, graph_window: int | None = None) -> None:
        """n_steps steps of `resolution` (event-driven networks: windows)."""
        rt = self.runtime
        rt.run(n_steps, self.graph_window if graph_window is None else graph_window)
        self.t = rt.window_index * rt.window if self.event_driven else rt.step * self.resolution

    def plan(self, max_launches: int = 1000, until: float | None = None) -> list[list[tuple]]:
        """A dry run of the next launches (naer2/network.py:154-204): each launch the list of
        (population, t_start, t_end, gated) whose window it integrates, at most max_launches of them or up to
        network time `until`; nothing runs and nothing changes.

        Time-driven: one launch per step of `resolution`, every population's intake in every launch; a
        population with a timestep of its own (k steps) integrates its window [t_end - k dt, t_end) in the
        launch of the window's last step, at the steps s with (s + 1) % k == 0.  Event-driven: one launch per
        window of the smallest delay; receiving populations are gated (a group with no input arriving in the
        window does no work), sources emit every window."""
        pops = sorted(self.populations, key=lambda p: p._order)
        if self.event_driven:
            from .event import describe
            W = describe(self).window
            k0 = self._rt.window_index if self._rt is not None else 0
            out = []
            while len(out) < max_launches:
                t0, t1 = k0 * W, (k0 + 1) * W
                if until is not None and t1 > until + 1e-9 * max(1.0, abs(until)):
                    break
                out.append([(p, t0, t1, p.family == "event_lif") for p in pops])
                k0 += 1
            return out
        dt = float(self.resolution)
        strides = [pop_stride(p, dt) for p in pops]
        s = self.step_index
        out = []
        while len(out) < max_launches:
            if until is not None and (s + 1) * dt > until + 1e-9 * max(1.0, abs(until)):
                break
            out.append([(p, (s + 1 - k) * dt, (s + 1) * dt, False) for p, k in zip(pops, strides)
                        if (s + 1) % k == 0])
            s += 1
        return out

    def run_planned(self, max_launches: int = 1000, until: float | None = None) -> int:
        """The planned launches (plan()) run eagerly, one kernel sequence each; returns the launches made."""
        n = len(self.plan(max_launches, until))
        if n:
            self.run(n, graph_window=0)
        return n

    def run_graph(self, max_launches: int = 1000, until: float | None = None) -> int:
        """As run_planned, replayed from captured hipGraph windows (naer2/network.py:262-321): one capture
        per window length, every later window a replay."""
        n = len(self.plan(max_launches, until))
        if n:
            self.run(n)
        return n

    def spikes(self, pop=None) -> tuple[np.ndarray, np.ndarray]:
        """(steps, neuron ids) of every recorded spike, ascending by step then id; neuron ids are the
        population's own (local) indices when pop is given, the network's otherwise.  Event-driven networks:
        (spike times in ms, neuron ids), window by window."""
        return self.runtime.spikes(pop)

    def rates(self, t_skip_ms: float = 0.0) -> dict:
        """Mean rate in Hz of each population over the recorded steps after t_skip_ms."""
        rt = self.runtime
        if self.event_driven:
            return rt.rates(t_skip_ms)
        steps, ids = rt.spikes(None)
        first = int(round(t_skip_ms / self.resolution))
        dur = (rt.step - first) * self.resolution / 1000.0
        out = {}
        keep = steps >= first
        ids = ids[keep]
        for pop, lo in zip(rt.pops, rt.pop_first[:-1]):
            n = ((ids >= lo) & (ids < lo + pop.n)).sum()
            out[pop] = float(n) / pop.n / dur if dur > 0 else float("nan")
        return out


@dataclass
class Assembled:
    """A network's host-side description for one device: populations in creation order, their
    parameters, the canonical synapse table and the initial state."""
    pops: list
    family: str
    model: int
    dt: float
    pop_first: np.ndarray
    params: list
    canon: Canonical
    d_slots: int
    v: torch.Tensor
    isyn: torch.Tensor | None
    ref: torch.Tensor
    edges: bool = False  # some connection has a non-static edge kind or the potential port

    @property
    def n(self) -> int:
        return int(self.pop_first[-1])


# the edge bits of a link's packed word (n2h_bind_edges): kind in bits 8-9, port in bit 10
EDGE_KIND_BITS = {"static": 0, "gain": 1 << 8, "plastic": 2 << 8}
PORT_POTENTIAL_BIT = 1 << 10
# the rows of the per-link payload (`canon.extra`, int32 [4, n_links])
PAYLOAD_CONN, PAYLOAD_FLAGS, PAYLOAD_ATTR, PAYLOAD_INPUT = range(4)


def assemble(net: Network, device=None, with_connections: bool = False) -> Assembled:
    """Everything the plan needs, on `device` (default: the network's), without touching the GPU.
    With with_connections, or when some connection has a non-static edge kind or the potential port, the
    canonical table carries an int32 [4, n_links] payload (`canon.extra`, one row per field, each row indexed
    as a 1-D tensor: torch's row indexing of a [n_links, 4] table gives wrong rows past about 2 GiB on this
    ROCm build): per link its connection index
    (creation order), its edge bits (EDGE_KIND_BITS | PORT_POTENTIAL_BIT), the bits of its second attribute
    (gain or factor; 0.0 for static edges) and its place in the connections' links concatenated in creation
    order."""
    pops = sorted(net.populations, key=lambda p: p._order)
    fams = {p.family for p in pops}
    if None in fams:
        bad = [type(p.instance).__name__ for p in pops if p.family is None]
        raise NotImplementedError(f"no hand-written kernel for node types {bad}: the reference's hook "
                                  f"transpiler is not part of this backend (nodes.py lists the families)")
    if len(fams) != 1:
        raise NotImplementedError(f"one model family per network (got {sorted(fams)})")
    family = fams.pop()
    model = _FAMILY_MODEL[family]
    if len(pops) > _lib.MAX_POPS:
        raise ValueError(f"at most {_lib.MAX_POPS} populations")
    dt = float(net.resolution)
    dev = torch.device(device if device is not None else net.device)
    sizes = [p.n for p in pops]
    pop_first = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(pop_first[-1])
    first_of = {p: int(f) for p, f in zip(pops, pop_first[:-1])}
    strides = [pop_stride(p, dt) for p in pops]
    params = []
    for p, k in zip(pops, strides):
        par = p.instance.params(k * dt)  # propagators, Poisson mean and refractory updates for its interval
        par.stride = k
        params.append(par)
    edges = any(not c.plain for c in net.connections)
    if edges and model != MODEL_LIF_EXP:
        raise NotImplementedError("edge kinds and ports need the exponential-synapse model (lif_exp)")
    payload = with_connections or edges
    srcs, tgts, ws, ds, cs = [], [], [], [], []
    base = 0
    for ci, c in enumerate(net.connections):
        if c._weight is None or c._delay is None:
            raise ValueError(f"{c!r}: weight and delay must be set")
        srcs.append(c.src.to(dev, torch.int64) + first_of[c.source])
        tgts.append(c.tgt.to(dev, torch.int64) + first_of[c.target])
        ws.append(c.weight.to(dev))
        ds.append(delay_steps(c.delay.to(dev), dt))
        if payload:
            m = srcs[-1].numel()
            col = torch.zeros(4, m, dtype=torch.int32, device=dev)
            col[PAYLOAD_CONN] = ci
            col[PAYLOAD_FLAGS] = EDGE_KIND_BITS[c.kind] | (PORT_POTENTIAL_BIT if c.port == "potential" else 0)
            if c.second_attr is not None:
                col[PAYLOAD_ATTR] = c.second_attr.to(dev, torch.float32).view(torch.int32)
            col[PAYLOAD_INPUT] = torch.arange(base, base + m, dtype=torch.int32, device=dev)
            base += m
            cs.append(col)
    cat = (lambda xs, dt_: torch.cat(xs) if xs else torch.zeros(0, dtype=dt_, device=dev))
    canon = canonical(n, cat(srcs, torch.int64), cat(tgts, torch.int64), cat(ws, torch.float32),
                      cat(ds, torch.int32), (torch.cat(cs, dim=1) if cs else torch.zeros(4, 0, dtype=torch.int32, device=dev))
                      if payload else None)
    del srcs, tgts, ws, ds, cs
    v = torch.empty(n, dtype=torch.float32, device=dev)
    isyn = torch.zeros(n, dtype=torch.float32, device=dev) if model == MODEL_LIF_EXP else None
    ref = torch.zeros(n, dtype=torch.int32, device=dev)
    vname = "V" if model == MODEL_LIF_DELTA else "potential"
    for p, lo in zip(pops, pop_first[:-1]):
        sl = slice(int(lo), int(lo) + p.n)
        v[sl] = p._init.get(vname, torch.tensor(float(getattr(p.instance, vname)))).to(dev, torch.float32)
        ref[sl] = p._init.get("ref", torch.tensor(0)).to(dev, torch.int32)
        if isyn is not None:
            isyn[sl] = p._init.get("current", torch.tensor(float(p.instance.current))).to(dev, torch.float32)
    # a window's slots stay unwritten by later arrivals until the window's update reads them
    d_slots = canon.max_delay + max(strides) - 1
    if d_slots > _lib.MAX_SLOTS:
        raise ValueError(f"largest delay {canon.max_delay} steps + largest stride {max(strides)} - 1 exceeds "
                         f"{_lib.MAX_SLOTS} ring slots")
    return Assembled(pops, family, model, dt, pop_first, params, canon, d_slots, v, isyn, ref, edges)


def pop_stride(pop, dt: float) -> int:
    """Steps of `dt` per update of `pop`: 1 without a timestep of its own, else timestep / dt, which must be
    a whole number (naer2/population.py:191-204: a clocked population runs windows of its timestep)."""
    ts = getattr(pop, "timestep", None)
    if ts is None:
        return 1
    k = float(ts) / dt
    if not math.isfinite(k) or k < 1 or abs(k - round(k)) > 1e-6 * k:
        raise ValueError(f"{pop!r}: timestep {ts} is not a whole multiple (>= 1) of the resolution {dt}")
    k = int(round(k))
    if k > _lib.MAX_STRIDE:
        raise ValueError(f"{pop!r}: timestep {ts} is more than {_lib.MAX_STRIDE} steps of {dt}")
    return k


class Runtime:
    """The built network: device tensors, the C plan, captured graphs."""

    def __init__(self, net: Network):
        self.net = net
        self.device = torch.device(net.device)
        if self.device.type != "cuda":
            raise RuntimeError("naer2_amd runs on the GPU: no CUDA/HIP device is available")
        a = assemble(net, self.device)
        if not net.keep_description:
            for c in net.connections:
                c.release()
        self.pops, self.family, self.model, self.dt = a.pops, a.family, a.model, a.dt
        self.pop_first, self.params, self.canon, self.d_slots = a.pop_first, a.params, a.canon, a.d_slots
        self.n = a.n
        self.v, self.isyn, self.ref = a.v, a.isyn, a.ref
        self.layout: GroupLayout = group_layout([p.n for p in self.pops], self.device)
        self.n_groups = self.layout.n_groups
        seg = segments(self.canon, self.layout)
        self.seg_off, self.links = seg.seg_off, seg.links
        # edge kinds and ports (n2h_bind_edges): the edge bits ORed into each link's packed word, the second
        # attribute beside the table, and the potential port's own ring
        self.edges = a.edges
        self.edge_attr = self.ring_v = self.link_input = None
        self.v_slots = 0
        if self.edges:
            if net.row_intake is True:
                raise ValueError("edge kinds and ports step without the row intake")
            x = seg.extra
            flags = x[PAYLOAD_FLAGS]
            self.links[:, 0] |= flags
            kinds = flags & (3 << 8)
            if bool((kinds != 0).any()):  # gain or plastic links read their second attribute
                self.edge_attr = x[PAYLOAD_ATTR].contiguous().view(torch.float32)
            self.link_input = x[PAYLOAD_INPUT].contiguous()
            self._plastic = bool((kinds == EDGE_KIND_BITS["plastic"]).any())
            # the potential port: its own ring, and LDS cells up to its links' largest delay
            pot = (flags & PORT_POTENTIAL_BIT) != 0
            self.v_slots = int(((self.links[:, 0] >> 16) * pot).max()) if bool(pot.any()) else 0
            if self.v_slots:
                self.ring_v = torch.zeros(self.d_slots, self.n, dtype=torch.float32, device=self.device)
            del x, flags, kinds, pot
        self.conn_links = [c.shape[0] for c in net.connections]
        self.rows = None
        if net.row_intake and self.n_groups <= _lib.DENSE_GROUPS and not self.edges:
            self.rows = rows(self.canon, self.layout, fmt=net.row_format)
        elif net.row_intake is True:
            raise ValueError(f"row intake needs at most {_lib.DENSE_GROUPS} groups ({self.n_groups} here)")
        # default: the row intake from 5 % of the sources on.  It reads every in-link (4-8 bytes, ~0.1 VALU
        # instruction per in-link and wave) where the segment intake reads only the delivered links (~3 VALU
        # instructions each): cheaper in instructions above ~3 % activity, in HBM bytes only far above; 5 %
        # measured best over 256 batched Brunel networks (2 %: 15 % slower, 10 %: 12 % slower)
        self.dense_min = int(net.dense_min_spikes) if net.dense_min_spikes is not None else max(65, self.n // 20)
        self.n_links = self.canon.n_links
        self.out_degree = self.canon.out_degree()
        # every link's delay in steps when all are equal (one partial sum per target in the step), else 0
        d = self.canon.d
        self.uniform_delay = int(d[0]) if d.numel() and bool((d == d[0]).all()) and not self.edges else 0
        if not net.keep_canonical:
            self.canon = None
        del seg
        self.ring = torch.zeros(self.d_slots, self.n, dtype=torch.float32, device=self.device)
        self.bit_slots = 2
        self.bits = torch.zeros(self.bit_slots, self.n_groups, dtype=torch.int64, device=self.device)
        # the spike lists of the last three steps: count, then source slots (64 * group + lane)
        self.spike_list = torch.zeros(4, 1 + GROUP * self.n_groups, dtype=torch.int32, device=self.device)
        self.record_from = 0
        self.step = 0
        self._batch_capacity = None
        self._graphs: dict = {}
        self._plan = None
        self._create_plan()
        for p, lo in zip(self.pops, self.pop_first[:-1]):
            object.__setattr__(p, "_net_slice", (self, int(lo)))

    # -- the C plan
    def _create_plan(self) -> None:
        lib = _lib.load()
        pops_arr = (_lib.PopParams * len(self.params))(*self.params)
        first_arr = (ctypes.c_int32 * (len(self.pops) + 1))(*[int(x) for x in self.pop_first])
        desc = _lib.ModelDesc(model=self.model, n_neurons=self.n, n_pops=len(self.pops), d_slots=self.d_slots,
                              pop_first=first_arr, pops=pops_arr, seed=int(self.net.seed) & (2**64 - 1),
                              instance=int(self.net.instance) & 0xFFFFFFFF)
        tab = _lib.Tables(n_groups=self.n_groups, uniform_delay=self.uniform_delay,
                          group_first=self.layout.group_first.data_ptr(),
                          group_info=self.layout.group_info.data_ptr(), seg_off=self.seg_off.data_ptr(),
                          links=self.links.data_ptr(), n_links=self.n_links)
        plan = ctypes.c_void_p()
        _lib.check(lib.n2h_plan_create(ctypes.byref(desc), ctypes.byref(tab), ctypes.byref(plan)), "n2h_plan_create")
        self._plan = plan
        _lib.check(lib.n2h_bind_state(plan, self.v.data_ptr(), self.isyn.data_ptr() if self.isyn is not None else None,
                                      self.ref.data_ptr()), "n2h_bind_state")
        self._bind_rows()
        self._bind_ring()
        if self.edges:
            ptr = (lambda t: t.data_ptr() if t is not None else None)
            _lib.check(lib.n2h_bind_edges(plan, ptr(self.edge_attr), ptr(self.ring_v), self.v_slots),
                       "n2h_bind_edges")

    def _bind_rows(self) -> None:
        lib = _lib.load()
        r = self.rows
        if r is None:
            _lib.check(lib.n2h_bind_rows(self._plan, None, None, None, 0, None, 0), "n2h_bind_rows")
        else:
            wtab = r.wtab.data_ptr() if r.wtab is not None else None
            _lib.check(lib.n2h_bind_rows(self._plan, r.rows.data_ptr(), r.row_base.data_ptr(), r.row_depth.data_ptr(),
                                         int(r.fmt), wtab, int(self.dense_min)),
                       "n2h_bind_rows")

    def set_dense_min(self, spikes: int) -> None:
        """Steps that follow at least `spikes` spikes use the row intake (results do not change)."""
        if self.rows is None:
            raise RuntimeError("no row-intake table was built for this network")
        self.dense_min = int(spikes)
        self._bind_rows()
        lib = _lib.load()
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def _bind_ring(self) -> None:
        lib = _lib.load()
        _lib.check(lib.n2h_bind_ring(self._plan, self.ring.data_ptr(), self.bits.data_ptr(), self.bit_slots,
                                     self.spike_list.data_ptr()), "n2h_bind_ring")
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            lib = _lib.load()
            for g in self._graphs.values():
                lib.n2h_graph_destroy(g)
            if self._plan is not None:
                lib.n2h_plan_destroy(self._plan)
        except Exception:
            pass

    def _ensure_slots(self, n_more: int) -> None:
        """With recording on, keep every step's bitmap since record_from (and the step before it, which
        the first recorded step's delivery reads): grow the history to hold the steps to come."""
        if not self.net.record:
            return
        need = self.step + n_more - self.record_from + 1
        if need <= self.bit_slots:
            return
        new_slots = max(need, 2 * self.bit_slots)
        new = torch.zeros(new_slots, self.n_groups, dtype=torch.int64, device=self.device)
        keep = torch.arange(max(self.record_from - 1, 0), self.step, device=self.device)
        if keep.numel():
            new[keep % new_slots] = self.bits[keep % self.bit_slots]
        self.bits, self.bit_slots = new, new_slots
        self._bind_ring()

    @staticmethod
    def schedule(step: int, n_steps: int, graph_window: int) -> list[tuple[str, int, int]]:
        """How run(n_steps, graph_window) from `step` launches: a list of ("eager", steps, phase) and
        ("graph", steps, phase) entries, phase = the step phase (mod 4) the entry starts at.  A graph replays
        only at the phase it was captured at.  graph_window == n_steps: one graph of exactly n_steps (any
        length).  Otherwise windows are a multiple of 4 steps: eager steps up to phase 0, whole windows, an
        eager tail."""
        if n_steps <= 0:
            return []
        if not graph_window:
            return [("eager", n_steps, step % 4)]
        if graph_window == n_steps:
            return [("graph", n_steps, step % 4)]
        if graph_window % 4:
            raise ValueError("graph windows must be a multiple of 4 steps (the spike lists' phase) unless the "
                             "window is the whole run")
        out, done = [], 0
        if step % 4 and n_steps >= graph_window + 4:
            lead = 4 - step % 4  # eager steps up to the captured phase
            out.append(("eager", lead, step % 4))
            done = lead
        while n_steps - done >= graph_window:
            out.append(("graph", graph_window, (step + done) % 4))
            done += graph_window
        if n_steps > done:
            out.append(("eager", n_steps - done, (step + done) % 4))
        return out

    def _graph(self, n: int, phase: int, at: int, sv) -> ctypes.c_void_p:
        """The graph of n steps replayed at `phase`; captured now (the plan's step mirror is at step `at`)
        when missing."""
        key = (n, phase)
        g = self._graphs.get(key)
        if g is None:
            lib = _lib.load()
            if at % 4 != phase:
                raise RuntimeError(f"capture at phase {phase} from step {at}")
            gp = ctypes.c_void_p()
            _lib.check(lib.n2h_graph_create(self._plan, n, sv, ctypes.byref(gp)), "n2h_graph_create")
            g = self._graphs[key] = gp
        return g

    def prepare(self, n_steps: int, graph_window: int = 0) -> int:
        """Capture (without running) every graph that run(n_steps, graph_window) will replay from the current
        step, so that a timed run captures nothing; returns the graphs captured.  Graphs are captured at the
        current step's phase: each is keyed by (length, phase), and the plan's step mirror is moved to the
        entry's phase for its capture and back."""
        plan = self.schedule(self.step, n_steps, graph_window)
        self._ensure_slots(n_steps)  # a history grown later would rebind the ring and drop the graphs
        lib = _lib.load()
        sv = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        made = 0
        for kind, k, phase in plan:
            if kind != "graph" or (k, phase) in self._graphs:
                continue
            # the capture records kernels whose arguments follow the host mirror's phase only
            at = self.step
            shift = (phase - at % 4) % 4
            if shift:
                _lib.check(lib.n2h_set_step(self._plan, at + shift, sv), "n2h_set_step")
            try:
                gp = ctypes.c_void_p()
                _lib.check(lib.n2h_graph_create(self._plan, k, sv, ctypes.byref(gp)), "n2h_graph_create")
                self._graphs[(k, phase)] = gp
                made += 1
            finally:
                if shift:
                    _lib.check(lib.n2h_set_step(self._plan, at, sv), "n2h_set_step")
        return made

    def run(self, n_steps: int, graph_window: int = 0) -> None:
        if n_steps < 0:
            raise ValueError("n_steps must be >= 0")
        if n_steps == 0:
            return
        plan = self.schedule(self.step, n_steps, graph_window)
        self._ensure_slots(n_steps)
        lib = _lib.load()
        sv = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        at = self.step
        for kind, k, phase in plan:
            if kind == "eager":
                _lib.check(lib.n2h_step(self._plan, k, sv), "n2h_step")
            else:
                _lib.check(lib.n2h_graph_launch(self._graph(k, phase, at, sv), sv), "n2h_graph_launch")
            at += k
        self.step += n_steps
        if not self.net.record:
            self.record_from = max(self.step - 1, 0)

    def force_spikes(self, step: int, spiking: np.ndarray) -> None:
        """Replace the spikes of `step` (bool [n]) before step + 1 runs -- the host-commanded firing of
        the reference's test fixtures (test/runtimes.py:24-44): bitmap word and spike list both."""
        words = torch.from_numpy(raster_to_bits(np.asarray(spiking, bool)[None], self.layout)[0]).to(self.device)
        self.bits[step % self.bit_slots] = words
        ids = np.nonzero(np.asarray(spiking, bool))[0]
        slots = (self.layout.group_of.cpu().numpy()[ids] * GROUP + self.layout.local_of.cpu().numpy()[ids])
        row = torch.zeros(1 + GROUP * self.n_groups, dtype=torch.int32)
        row[0] = len(slots)
        row[1:1 + len(slots)] = torch.from_numpy(slots[::-1].astype(np.int32).copy())  # any order
        self.spike_list[step % 4] = row.to(self.device)

    # -- readback formats: the delivery log and status rows (naer2/runtime.py:747-752, fused.py:865-883)
    def set_delivery_log(self, capacity: int | None) -> None:
        """Record every link the following steps deliver (n2h_bind_log), up to `capacity` records; None:
        off.  Needs the connections' link tables (Network.keep_description).  Results do not change."""
        lib = _lib.load()
        if capacity is None:
            _lib.check(lib.n2h_bind_log(self._plan, None, None, None, 0), "n2h_bind_log")
            self._log = None
        else:
            if not self.net.keep_description:
                raise RuntimeError("the delivery log needs the connections' links (Network.keep_description)")
            if len(self.net.connections) > 32767:
                raise ValueError("at most 32767 connections carry their index in the delivery log")
            a = assemble(self.net, self.device, with_connections=True)
            conn = segments(a.canon, self.layout).extra[PAYLOAD_CONN].to(torch.int16).contiguous()
            del a
            rec = torch.zeros(int(capacity), 5, dtype=torch.int32, device=self.device)
            count = torch.zeros(1, dtype=torch.int32, device=self.device)
            self._log = (conn, rec, count)
            _lib.check(lib.n2h_bind_log(self._plan, conn.data_ptr(), rec.data_ptr() if capacity else count.data_ptr(),
                                        count.data_ptr(), int(capacity)), "n2h_bind_log")
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def deliveries(self) -> list[tuple[int, int, int, int, int]]:
        """The delivery log in log order: (arrival step, emission step, connection, source, target), network
        neuron ids; raises if records were lost to the capacity."""
        conn, rec, count = self._log
        torch.cuda.synchronize(self.device)
        n = int(count.item())
        if n > rec.shape[0]:
            raise RuntimeError(f"the delivery log overflowed: {n} records, capacity {rec.shape[0]}")
        return [tuple(r) for r in rec[:n].cpu().tolist()]

    def delivery_count(self) -> int:
        torch.cuda.synchronize(self.device)
        return int(self._log[2].item())

    def set_status_rows(self, on: bool = True) -> None:
        """Have every step write its per-group status rows (n2h_set_status)."""
        lib = _lib.load()
        self._status = torch.full((self.n_groups, 4), -1, dtype=torch.int32, device=self.device) if on else None
        _lib.check(lib.n2h_set_status(self._plan, self._status.data_ptr() if on else None), "n2h_set_status")
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def set_launch_mapping(self, xcd: int, run: int = 0) -> None:
        """Launch mapping of this runtime's own step kernels (n2h_set_tuning): xcd = 0 runs group b in block b;
        k > 1 gives the blocks the hardware deals to each of k XCDs consecutive groups: one chunk each (run 0,
        the default with k = 8) or runs of `run` groups dealt to the XCDs in turn.  Results do not change;
        captured graphs are dropped so that the next run captures with the new mapping."""
        lib = _lib.load()
        t = _lib.Tuning(order=0, xcd=int(xcd), stage=0, run=int(run))
        _lib.check(lib.n2h_set_tuning(self._plan, ctypes.byref(t)), "n2h_set_tuning")
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def status_rows(self) -> np.ndarray:
        """int32 [n_groups, 4]: per target group, the last step it ran, the spikes it emitted then, its intake
        path (0 none, 1 sorted list, 2 bitmap scan, 3 row intake) and the spiking sources it took in."""
        torch.cuda.synchronize(self.device)
        return self._status.cpu().numpy()

    TRACE_EVENTS = ("entry", "list read", "sources ranked", "segments", "links staged", "walked", "updated",
                    "exit")

    def trace(self, n_steps: int = 1) -> dict:
        """Run n_steps eager steps with in-kernel phase stamps on; return, for the last of them, the mean and
        max over groups of each phase's duration in microseconds (stamps: 100 MHz wall clock) and the
        spread of group start and end times."""
        buf = torch.zeros(self.n_groups, 8, dtype=torch.int64, device=self.device)
        lib = _lib.load()
        _lib.check(lib.n2h_set_trace(self._plan, buf.data_ptr()), "n2h_set_trace")
        try:
            self.run(n_steps, graph_window=0)
            torch.cuda.synchronize(self.device)
        finally:
            _lib.check(lib.n2h_set_trace(self._plan, None), "n2h_set_trace")
        t = buf.cpu().numpy().astype(np.int64)
        out = {}
        for j in range(1, 8):
            d = (t[:, j] - t[:, j - 1]) / 100.0
            ok = (t[:, j] > 0) & (t[:, j - 1] > 0) & (d >= 0)
            if ok.any():
                out[f"{self.TRACE_EVENTS[j - 1]} -> {self.TRACE_EVENTS[j]}"] = (float(d[ok].mean()), float(d[ok].max()))
        out["group starts spread"] = float((t[:, 0].max() - t[:, 0].min()) / 100.0)
        out["entry -> exit"] = (float(((t[:, 7] - t[:, 0]) / 100.0).mean()), float(((t[:, 7] - t[:, 0]) / 100.0).max()))
        out["first entry -> last exit"] = float((t[:, 7].max() - t[:, 0].min()) / 100.0)
        return out

    def trace_potentials(self, buffer: torch.Tensor | None, step0: int = 0,
                         compensation: torch.Tensor | None = None) -> None:
        """Write each neuron's distance to the threshold as the threshold sees it (V before reset minus V_th;
        +inf on refractory steps) for steps step0 .. step0 + len(buffer) - 1 into buffer [steps, n] (float32, on
        the device); None: off.  compensation (exponential-synapse networks): a float32 [2, n] device tensor
        holding what fp32 rounding dropped from V and I (zeros: none yet), updated every step while set; the
        trace is then the compensated distance.  Results are unchanged; captured windows are dropped so that
        the next run sees the setting."""
        lib = _lib.load()
        lo = (None, None)
        if compensation is not None:
            if (compensation.dtype != torch.float32 or tuple(compensation.shape) != (2, self.n)
                    or not compensation.is_contiguous()):
                raise ValueError("compensation: contiguous float32 [2, n_neurons]")
            lo = (compensation[0].data_ptr(), compensation[1].data_ptr())
        if buffer is None:
            _lib.check(lib.n2h_set_vtrace(self._plan, None, 0, 0, *lo), "n2h_set_vtrace")
        else:
            if buffer.dtype != torch.float32 or buffer.dim() != 2 or buffer.shape[1] != self.n or not buffer.is_contiguous():
                raise ValueError("buffer: contiguous float32 [steps, n_neurons]")
            _lib.check(lib.n2h_set_vtrace(self._plan, buffer.data_ptr(), int(step0), int(buffer.shape[0]), *lo),
                       "n2h_set_vtrace")
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}

    def snapshot(self) -> tuple:
        """A device copy of the dynamic state (step, V, I, refractory counters, ring, spike bits and lists;
        with edges, the potential port's ring and, with plastic edges, the link weights)."""
        edge = None
        if self.edges:
            edge = (None if self.ring_v is None else self.ring_v.clone(),
                    self.links[:, 1].clone() if self._plastic else None)
        return (self.step, self.v.clone(), None if self.isyn is None else self.isyn.clone(), self.ref.clone(),
                self.ring.clone(), self.bits.clone(), self.spike_list.clone(), edge)

    def restore(self, snap: tuple) -> None:
        """Back to a snapshot() of this runtime; recording restarts at its step.  A bit history grown since
        is replaced by the snapshot's (not while a FusedRound holds this runtime: it keeps the addresses)."""
        step, v, isyn, ref, ring, bits, lst, edge = snap
        if edge is not None:
            if edge[0] is not None:
                self.ring_v.copy_(edge[0])
            if edge[1] is not None:
                self.links[:, 1] = edge[1]
        self.v.copy_(v)
        if isyn is not None:
            self.isyn.copy_(isyn)
        self.ref.copy_(ref)
        self.ring.copy_(ring)
        if bits.shape == self.bits.shape:
            self.bits.copy_(bits)
        else:
            if getattr(self, "_batch_capacity", None) is not None:
                raise RuntimeError("restore: the bit history moved under a FusedRound")
            self.bits, self.bit_slots = bits.clone(), int(bits.shape[0])
            self._bind_ring()
        self.spike_list.copy_(lst)
        self.set_step(step)

    def set_step(self, step: int) -> None:
        """Move the step counter (the Philox counter and ring slot follow the absolute step)."""
        lib = _lib.load()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(lib.n2h_set_step(self._plan, int(step), ctypes.c_void_p(stream)), "n2h_set_step")
        self.step = int(step)
        self.record_from = int(step)

    def link_weights(self, conn=None) -> torch.Tensor:
        """The current link weights (plastic edges update them): of the connections' links concatenated in
        creation order, or of `conn` alone, each in its own link order (float32, on the device)."""
        w_seg = self.links[:, 1].contiguous().view(torch.float32)
        if self.link_input is None:
            raise RuntimeError("link weights in connection order need an edge plan (the weights of static "
                               "links do not change: Connection.weight holds them)")
        out = torch.empty_like(w_seg)
        out[self.link_input.to(torch.int64)] = w_seg
        if conn is None:
            return out
        ci = self.net.connections.index(conn)
        a = sum(self.conn_links[:ci])
        return out[a:a + self.conn_links[ci]]

    # -- state access for populations
    def _tensor(self, name: str):
        if name in ("V", "potential"):
            return self.v
        if name == "current":
            if self.isyn is None:
                raise AttributeError("current: no synaptic current in the delta-synapse model")
            return self.isyn
        if name == "ref":
            return self.ref
        raise AttributeError(name)

    def _read_state(self, pop, name):
        lo = dict(zip(self.pops, self.pop_first[:-1]))[pop]
        return self._tensor(name)[int(lo):int(lo) + pop.n]

    def _write_state(self, pop, name, value):
        self._read_state(pop, name).copy_(value.to(self.device))

    # -- raster
    def raster_bits(self, step_lo: int | None = None, step_hi: int | None = None) -> torch.Tensor:
        """int64 [steps, n_groups] spike words of the recorded steps [step_lo, step_hi)."""
        lo = self.record_from if step_lo is None else step_lo
        hi = self.step if step_hi is None else step_hi
        if lo < self.record_from or hi > self.step or (hi - lo) > self.bit_slots:
            raise ValueError(f"steps [{lo}, {hi}) are not recorded")
        idx = torch.arange(lo, hi, device=self.device) % self.bit_slots
        return self.bits[idx]

    def raster(self, step_lo: int | None = None, step_hi: int | None = None) -> np.ndarray:
        """bool [steps, n] spikes per step and neuron."""
        words = self.raster_bits(step_lo, step_hi).cpu().numpy()
        return bits_to_raster(words, self.layout, self.n)

    def spikes(self, pop=None):
        lo = self.record_from
        r = self.raster()
        steps, ids = np.nonzero(r)
        steps = steps + lo
        if pop is not None:
            first = int(dict(zip(self.pops, self.pop_first[:-1]))[pop])
            m = (ids >= first) & (ids < first + pop.n)
            return steps[m], ids[m] - first
        return steps, ids


def bits_to_raster(words: np.ndarray, layout: GroupLayout, n: int) -> np.ndarray:
    """bool [steps, n] from spike words [steps, n_groups] (bit b of group g = neuron group_first[g] + b)."""
    words = np.ascontiguousarray(words).astype(np.uint64)
    bits = np.unpackbits(words.view(np.uint8).reshape(words.shape[0], -1), axis=1, bitorder="little")
    bits = bits.reshape(words.shape[0], -1, GROUP)
    gf = layout.group_first.cpu().numpy().astype(np.int64)
    cnt = (layout.group_info.cpu().numpy() & 0xFFFF).astype(np.int64)
    out = np.zeros((words.shape[0], n), dtype=bool)
    for g in range(len(gf)):
        out[:, gf[g]:gf[g] + cnt[g]] = bits[:, g, :cnt[g]]
    return out


def raster_to_bits(raster: np.ndarray, layout: GroupLayout) -> np.ndarray:
    """Spike words [steps, n_groups] from a bool raster [steps, n]."""
    raster = np.asarray(raster, dtype=bool)
    gf = layout.group_first.cpu().numpy().astype(np.int64)
    cnt = (layout.group_info.cpu().numpy() & 0xFFFF).astype(np.int64)
    out = np.zeros((raster.shape[0], len(gf)), dtype=np.uint64)
    weights = (np.uint64(1) << np.arange(GROUP, dtype=np.uint64))
    for g in range(len(gf)):
        blk = raster[:, gf[g]:gf[g] + cnt[g]].astype(np.uint64)
        out[:, g] = (blk * weights[:cnt[g]]).sum(axis=1, dtype=np.uint64)
    return out.view(np.int64)
