"""naer2's network-description API, kept as the surface of this backend.

    E = LIFDelta(...) * 2000                      # node instance * size -> Population
    conn = E >> FixedInDegree(200) >> I          # source >> pattern >> target -> Connection
    conn.weight = 0.2                             # per-link tensor or one number
    conn.delay = 1.5                              # ms; or: E >> pattern >> I | 1.5
    net = E.network                               # every connected population's Network

Reference: naer2/decorators/node.py:10-73 (node, `instance * n`, `>>`, `@`), naer2/decorators/pattern.py
(pattern), naer2/decorators/edge.py (edge), naer2/_partial.py (edge @ pattern), naer2/connection.py:43-309
(ConnectionBuilder, Connection, `| delay`), naer2/population.py:22-143 (Population), naer2/types/region.py
and points.py (Region, Points).

What differs: the reference transpiles a node's update/send/receive hooks into a Gluon kernel
(naer2/transpiler); this backend runs hand-written gfx950 kernels, one per model family, so a node type
names its family (`family = "lif_delta"` or `"lif_exp"`, see nodes.py).  A node class without one can still
be declared and wired, and Network.build raises NotImplementedError for it.
"""
from __future__ import annotations

import itertools

import torch

_NODE_TYPES: dict[type, str | None] = {}
_EDGE_TYPES: dict[type, str] = {}
# edge kinds the step kernel runs (n2h_bind_edges) and the name of each kind's second per-link attribute
EDGE_KINDS = {"static": None, "gain": "gain", "plastic": "factor"}
# receive ports of the exponential-synapse model: the synaptic current or the membrane potential
PORTS = ("current", "potential")
_PATTERN_TYPES: set = set()
_ORDER = itertools.count()


# ---------------------------------------------------------------- regions (naer2/types)

class Region(tuple):
    """Index bounds of a population: ((lo, hi),) -- one dimension here."""

    def __new__(cls, shape):
        if isinstance(shape, Region):
            return shape
        if isinstance(shape, int):
            if shape < 0:
                raise ValueError("a population size must be >= 0")
            return super().__new__(cls, ((0, shape),))
        return super().__new__(cls, tuple((int(lo), int(hi)) for lo, hi in shape))

    def __len__(self) -> int:  # the number of neurons
        lo, hi = tuple.__getitem__(self, 0)
        return hi - lo


class Points:
    """A pattern's result for one source: target indices, shape [k, 1] (naer2/types/points.py)."""

    def __init__(self, points: torch.Tensor):
        self.points = points.reshape(-1, 1).to(torch.int64)

    def __len__(self) -> int:
        return int(self.points.shape[0])


# ---------------------------------------------------------------- decorators

def node(cls=None, *, family: str | None = None):
    """Register a node class (naer2/decorators/node.py).  `family` (or a class attribute of that name)
    names the hand-written kernel that runs it."""

    def wrap(c):
        if not isinstance(c, type):
            raise ValueError(f"node decorator can only be applied to types (got {type(c)})")
        _NODE_TYPES[c] = family if family is not None else getattr(c, "family", None)
        c.__mul__ = lambda self, amount: Population(self, amount)
        c.__rshift__ = lambda self, right: ConnectionBuilder(Population(self, 1)) >> right
        return c

    return wrap(cls) if cls is not None else wrap


def edge(cls=None, *, kind: str | None = None, **_):
    """Register an edge (synapse) class (naer2/decorators/edge.py).  `kind` (or a class attribute of that
    name; default "static") names the edge forward the step kernel runs for its links, as `family` names a
    node's kernel: "static" delivers the weight; "gain" delivers weight * gain; "plastic" first adds factor
    to the stored weight (the update persists), then delivers the new weight -- the reference's
    plastic_synapse forward (test/edges/plastic_synapse.py:10-11).  weight and the kind's attribute
    (gain, factor) become per-link tensors; backward hooks and hooks reading source or target state (the
    reference's open decisions D-9 to D-11, TODO.md:196-204) are not run."""

    def wrap(c):
        k = kind if kind is not None else getattr(c, "kind", "static")
        if k not in EDGE_KINDS:
            raise ValueError(f"edge kind {k!r}: one of {sorted(EDGE_KINDS)}")
        _EDGE_TYPES[c] = k
        c.__matmul__ = lambda self, other: _PartialConnection(self) @ other
        return c

    return wrap(cls) if cls is not None else wrap


def edge_kind(instance) -> str:
    """The edge kind of an edge instance ("static" for none)."""
    return "static" if instance is None else _EDGE_TYPES.get(type(instance), "static")


def pattern(cb):
    """Register a connection pattern (naer2/decorators/pattern.py).  A pattern is either the reference's
    per-source callable `(position, source_region, target_region) -> Points`, or offers the vectorized
    `links(n_src, n_tgt, generator) -> (src, tgt)` (int64, source-major order) used where present."""
    if not callable(cb):
        raise ValueError(f"pattern decorator can only be applied to callables (got {type(cb).__name__})")
    _PATTERN_TYPES.add(cb)
    return cb


def family_of(instance) -> str | None:
    if type(instance) not in _NODE_TYPES:
        raise ValueError(f"{type(instance).__name__} is not a registered node type (use @node)")
    return _NODE_TYPES[type(instance)]


# ---------------------------------------------------------------- populations

class Population:
    """`size` neurons of one node type (naer2/population.py:22-143).  Its dynamic state lives in the
    network's device tensors once built; before that, values set here are kept as the initial state."""

    _STATE = {"V", "potential", "current", "ref", "last_time", "t_ref_end", "next_spike"}

    def __init__(self, instance, shape):
        family_of(instance)
        object.__setattr__(self, "instance", instance)
        object.__setattr__(self, "shape", Region(shape))
        object.__setattr__(self, "incoming", [])
        object.__setattr__(self, "_init", {})
        object.__setattr__(self, "_order", next(_ORDER))
        object.__setattr__(self, "_net_slice", None)  # (network, first neuron) once built
        # its own clock (naer2/population.py:191-204): None updates every step of the network's resolution;
        # a whole multiple of it updates once per that many steps
        object.__setattr__(self, "timestep", None)
        from .network import Network
        object.__setattr__(self, "network", Network(self))

    @property
    def n(self) -> int:
        return len(self.shape)

    @property
    def family(self) -> str | None:
        return family_of(self.instance)

    def __len__(self) -> int:
        return self.n

    def __setattr__(self, name, value):
        if name in self._STATE:
            t = torch.as_tensor(value)
            if t.dim() == 0:
                t = t.expand(self.n)
            if t.shape != (self.n,):
                raise ValueError(f"{name} must have shape ({self.n},)")
            built = self._net_slice
            if built is not None:
                built[0]._write_state(self, name, t)
            else:
                self._init[name] = t.clone()
            return
        if name == "timestep" and self._net_slice is not None and value != self.timestep:
            raise RuntimeError("the network is built: set timestep before its first step")
        object.__setattr__(self, name, value)

    def __getattr__(self, name):
        if name in Population._STATE:
            built = self.__dict__.get("_net_slice")
            if built is not None:
                return built[0]._read_state(self, name)
            init = self.__dict__["_init"]
            if name in init:
                return init[name]
        raise AttributeError(name)

    def __rshift__(self, right):
        return ConnectionBuilder(self) >> right

    def __repr__(self) -> str:
        return f"Population({type(self.instance).__name__} * {self.n})"

    __hash__ = object.__hash__


# ---------------------------------------------------------------- connections

class _PartialConnection:
    """`edge @ "port" @ pattern` in any order, before the source is known (naer2/_partial.py:1-44): a string
    names the target port, anything else but a mapping is the pattern."""

    def __init__(self, edge_instance, pattern_=None):
        self.edge = edge_instance
        self.pattern = pattern_
        self.port = None

    def __matmul__(self, other):
        if isinstance(other, str):
            if self.port is not None:
                raise ValueError(f"target port already set to {self.port!r}")
            self.port = other
            return self
        if isinstance(other, dict):
            raise ValueError("attribute mappings rename hook-facing names: this backend runs no hooks")
        if self.pattern is not None:
            raise ValueError(f"pattern already set to {self.pattern!r}")
        self.pattern = other
        return self


def _links_of(pattern_, n_src: int, n_tgt: int, generator) -> tuple[torch.Tensor, torch.Tensor]:
    """(src, tgt) int64 in source-major order."""
    vec = getattr(pattern_, "links", None)
    if vec is not None:
        src, tgt = vec(n_src, n_tgt, generator)
        return src.to(torch.int64), tgt.to(torch.int64)
    srcs, tgts = [], []
    src_region, tgt_region = Region(n_src), Region(n_tgt)
    for s in range(n_src):
        out = pattern_((s,), src_region, tgt_region)
        pts = out.points if isinstance(out, Points) else torch.as_tensor(out).reshape(-1, 1)
        k = int(pts.shape[0])
        srcs.append(torch.full((k,), s, dtype=torch.int64))
        tgts.append(pts.reshape(-1).to(torch.int64))
    if not srcs:
        return torch.zeros(0, dtype=torch.int64), torch.zeros(0, dtype=torch.int64)
    return torch.cat(srcs), torch.cat(tgts)


class ConnectionBuilder:
    """`source >> [edge @] pattern >> target` (naer2/connection.py:43-218)."""

    def __init__(self, source: Population):
        self._source = source
        self._pattern = None
        self._edge = None
        self._port = None

    def __rshift__(self, to):
        if isinstance(to, _PartialConnection):
            self._edge, self._pattern, self._port = to.edge, to.pattern, to.port
            return self
        if isinstance(to, Population):
            return self._target(to)
        if type(to) in _NODE_TYPES:
            raise ValueError("the target must be a Population (node * size)")
        if type(to) in _EDGE_TYPES:
            self._edge = to
            return self
        if to in _PATTERN_TYPES or type(to) in _PATTERN_TYPES or callable(to):
            self._pattern = to
            return self
        raise ValueError(f"Cannot connect to {to!r}")

    def _target(self, target: Population) -> "Connection":
        if self._pattern is None:
            raise ValueError("a connection needs a pattern (source >> pattern >> target)")
        net = self._source.network
        net.merge(target.network)
        src, tgt = _links_of(self._pattern, self._source.n, target.n, net.generator)
        if tgt.numel() and (int(tgt.min()) < 0 or int(tgt.max()) >= target.n):
            raise ValueError("the pattern produced a target outside the target population")
        if src.numel() > 1 and not bool((src[1:] >= src[:-1]).all()):
            order = torch.argsort(src, stable=True)  # the connection's link order: source-major, stable
            src, tgt = src[order], tgt[order]
        conn = Connection(self._source, target, src, tgt, self._edge, self._port)
        net.connections.append(conn)
        target.incoming.append(conn)
        return conn


class Connection:
    """A finished connection: per-link source/target (local indices, source-major), weight and delay
    (naer2/connection.py:221-309).  `conn.weight = x`, `conn.delay = x` (ms), `conn | delay`; gain and
    plastic edges add `conn.gain` / `conn.factor`; `conn.port` is the target port ("current" or
    "potential"; None: "current")."""

    def __init__(self, source: Population, target: Population, src: torch.Tensor, tgt: torch.Tensor,
                 edge_instance=None, port: str | None = None):
        self.source, self.target = source, target
        # local indices as int32 (populations hold fewer than 2^31 neurons): 8 bytes per link less than int64
        self._src, self._tgt = src.to(torch.int32), tgt.to(torch.int32)
        self._n_links = int(tgt.numel())
        self._released = False
        self.edge = edge_instance
        self._kind = edge_kind(edge_instance)
        self._port = None
        self.port = port
        self._weight = None
        self._delay = None
        self._attr2 = None
        if edge_instance is not None:
            names = ("weight", "delay") + ((EDGE_KINDS[self.kind],) if EDGE_KINDS[self.kind] else ())
            for name in names:
                if hasattr(edge_instance, name):
                    setattr(self, name, getattr(edge_instance, name))

    @property
    def shape(self) -> tuple[int]:
        return (self._n_links,)

    def _check_held(self) -> None:
        if self._released:
            raise RuntimeError("the connection's links were released after the build "
                               "(Network.keep_description = False)")

    @property
    def src(self) -> torch.Tensor:
        """int32 [n_links] local source index of each link."""
        self._check_held()
        return self._src

    @property
    def tgt(self) -> torch.Tensor:
        """int32 [n_links] local target index of each link."""
        self._check_held()
        return self._tgt

    def release(self) -> None:
        """Drop the per-link tensors (the built runtime keeps its own tables)."""
        self._src = self._tgt = self._weight = self._delay = self._attr2 = None
        self._released = True

    def _per_link(self, value, name: str) -> torch.Tensor:
        """fp32 [n_links] on the device of the links (a number is broadcast)."""
        t = torch.as_tensor(value, dtype=torch.float32)
        if t.dim() == 0:
            return torch.full(self.shape, float(t), dtype=torch.float32, device=self.tgt.device)
        if tuple(t.shape) != self.shape:
            raise ValueError(f"{name} tensor must be shape {self.shape}")
        return t.detach().to(self.tgt.device, torch.float32)

    @property
    def weight(self) -> torch.Tensor:
        self._check_held()
        if self._weight is None:
            raise ValueError("weight not set")
        return self._weight

    @weight.setter
    def weight(self, value) -> None:
        self._weight = self._per_link(value, "weight")
        self.source.network._invalidate()

    @property
    def delay(self) -> torch.Tensor:
        self._check_held()
        if self._delay is None:
            raise ValueError("delay not set")
        return self._delay

    @delay.setter
    def delay(self, value) -> None:
        d = self._per_link(value, "delay")
        if d.numel() and float(d.min()) <= 0:
            raise ValueError("delays must be positive")
        self._delay = d
        self.source.network._invalidate()

    def __or__(self, value) -> "Connection":
        self.delay = value
        return self

    @property
    def kind(self) -> str:
        """The edge kind the step runs for this connection's links (EDGE_KINDS); set from the edge instance,
        or directly (the kind's second attribute then has to be set again)."""
        return self._kind

    @kind.setter
    def kind(self, value: str) -> None:
        if value not in EDGE_KINDS:
            raise ValueError(f"edge kind {value!r}: one of {sorted(EDGE_KINDS)}")
        if value != self._kind:
            self._kind = value
            self._attr2 = None
            self.source.network._invalidate()

    @property
    def port(self) -> str | None:
        return self._port

    @port.setter
    def port(self, value: str | None) -> None:
        if value is not None and value not in PORTS:
            raise ValueError(f"port {value!r}: this backend's targets have the ports {PORTS}")
        self._port = value
        self.source.network._invalidate()

    def _second(self, name: str) -> None:
        if EDGE_KINDS[self.kind] != name:
            raise AttributeError(f"{name}: only {[k for k, a in EDGE_KINDS.items() if a == name][0]} edges have it "
                                 f"(this connection's edge is {self.kind})")

    @property
    def gain(self) -> torch.Tensor:
        """Per-link gain of a gain edge (delivers weight * gain)."""
        self._second("gain")
        self._check_held()
        if self._attr2 is None:
            raise ValueError("gain not set")
        return self._attr2

    @gain.setter
    def gain(self, value) -> None:
        self._second("gain")
        self._attr2 = self._per_link(value, "gain")
        self.source.network._invalidate()

    @property
    def factor(self) -> torch.Tensor:
        """Per-link increment of a plastic edge (weight += factor at every delivery, before it delivers)."""
        self._second("factor")
        self._check_held()
        if self._attr2 is None:
            raise ValueError("factor not set")
        return self._attr2

    @factor.setter
    def factor(self, value) -> None:
        self._second("factor")
        self._attr2 = self._per_link(value, "factor")
        self.source.network._invalidate()

    @property
    def second_attr(self) -> torch.Tensor | None:
        """The edge kind's second per-link attribute (gain or factor), or None for static edges."""
        return None if EDGE_KINDS[self.kind] is None else getattr(self, EDGE_KINDS[self.kind])

    @property
    def plain(self) -> bool:
        """A static edge into the current port: what the plain step delivers."""
        return self.kind == "static" and self._port in (None, "current")

    def __repr__(self) -> str:
        return f"Connection({self.source!r} -> {self.target!r}, {self.shape[0]} links)"
