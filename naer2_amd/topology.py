"""Synapse tables: the canonical CSR by source, the target-group layout, and the group segments the step
kernel reads.

Reference: naer2/topology.py:98-183 (Topology, canonical CSR by source) and :291-557
(SimulationTopology: links sorted by (source, target group, delay bucket) into segments with a directory
per source).  Here delays are whole steps fixed on the host (d_steps = max(1, round(d / dt))) instead of
bucket means, a target group is one wavefront of 64 neurons, and the per-source directory is a dense
(source, group) offset table -- 8 bytes per (source, group) pair, which HBM holds easily, and two adjacent
loads per spiking source instead of the reference's three dependent ones (fused.py:252-267).

Canonical link order: links sorted stably by source; for one source, the order in which its connections
were made and, inside one connection, the order its pattern produced them.  The summation order of the
step (ascending canonical link index per target) is defined on this order.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from ._lib import GROUP, MAX_SLOTS


@dataclass
class Canonical:
    """Canonical CSR by source: the links of source s are [src_ptr[s], src_ptr[s+1])."""
    n: int
    src_ptr: torch.Tensor  # int64 [n + 1]
    tgt: torch.Tensor      # int32 [E]
    w: torch.Tensor        # float32 [E]
    d: torch.Tensor        # int32 [E], delay in steps, >= 1
    extra: torch.Tensor | None = None  # an optional per-link payload in the same order: [E] or [k, E] (k rows)

    @property
    def n_links(self) -> int:
        return int(self.tgt.numel())

    @property
    def max_delay(self) -> int:
        return int(self.d.max()) if self.n_links else 1

    def out_degree(self) -> torch.Tensor:
        return self.src_ptr[1:] - self.src_ptr[:-1]

    def to(self, device) -> "Canonical":
        return Canonical(self.n, self.src_ptr.to(device), self.tgt.to(device), self.w.to(device),
                         self.d.to(device), None if self.extra is None else self.extra.to(device))


def delay_steps(delay_ms: torch.Tensor, dt: float) -> torch.Tensor:
    """int32 delay steps max(1, round(d / dt)), rounded half to even in float64."""
    steps = torch.round(delay_ms.to(torch.float64) / dt).clamp_(min=1)
    if steps.numel() and float(steps.max()) > MAX_SLOTS:
        raise ValueError(f"a delay of {float(steps.max())} steps exceeds the ring's {MAX_SLOTS} slots")
    return steps.to(torch.int32)


SORT_CHUNK = 1 << 30  # links sorted at once (torch's sort takes at most 2^31 - 1 elements)


def _stable_order_by_key(key: torch.Tensor, n_keys: int, chunk: int | None = None) -> torch.Tensor:
    """The positions a stable sort by key (int64 in [0, n_keys)) gives each element: chunk by chunk, a stable
    sort within the chunk, then each key's run placed after the same key's elements of earlier chunks."""
    chunk = SORT_CHUNK if chunk is None else chunk
    E = key.numel()
    dev = key.device
    if E <= chunk:
        order = torch.argsort(key, stable=True)
        pos = torch.empty_like(order)
        pos[order] = torch.arange(E, device=dev)
        return pos
    counts = torch.zeros(n_keys, dtype=torch.int64, device=dev)
    for a in range(0, E, chunk):
        counts += torch.bincount(key[a:a + chunk], minlength=n_keys)
    base = torch.cumsum(counts, 0) - counts  # the first position of each key
    pos = torch.empty(E, dtype=torch.int64, device=dev)
    for a in range(0, E, chunk):
        k = key[a:a + chunk]
        order = torch.argsort(k, stable=True)
        ks = k[order]
        cnt = torch.bincount(ks, minlength=n_keys)
        first = torch.cumsum(cnt, 0) - cnt
        pos[a + order] = base[ks] + (torch.arange(ks.numel(), device=dev) - first[ks])
        base += cnt
        del order, ks, cnt, first
    return pos


def canonical(n: int, src: torch.Tensor, tgt: torch.Tensor, w: torch.Tensor, d: torch.Tensor,
              extra: torch.Tensor | None = None) -> Canonical:
    """The canonical CSR of links given in any source order (a stable sort by source keeps each source's
    links in their given order); extra, a per-link payload ([E], or [k, E]: k rows of E), follows the links."""
    src = src.to(torch.int64)
    if src.numel() and (int(src.min()) < 0 or int(src.max()) >= n):
        raise ValueError("link source out of range")
    if tgt.numel() and (int(tgt.min()) < 0 or int(tgt.max()) >= n):
        raise ValueError("link target out of range")
    if d.numel() and int(d.min()) < 1:
        raise ValueError("link delays must be at least one step")
    counts = torch.bincount(src, minlength=n)
    src_ptr = torch.zeros(n + 1, dtype=torch.int64, device=src.device)
    src_ptr[1:] = torch.cumsum(counts, 0)
    pos = _stable_order_by_key(src, n)
    out_t = torch.empty(tgt.numel(), dtype=torch.int32, device=src.device)
    out_w = torch.empty(tgt.numel(), dtype=torch.float32, device=src.device)
    out_d = torch.empty(tgt.numel(), dtype=torch.int32, device=src.device)
    out_t[pos] = tgt.to(torch.int32)
    out_w[pos] = w.to(torch.float32)
    out_d[pos] = d.to(torch.int32)
    out_x = None
    if extra is not None:
        out_x = torch.empty_like(extra)
        for row_out, row in _payload_rows(out_x, extra):
            row_out[pos] = row
    return Canonical(n, src_ptr, out_t, out_w, out_d, out_x)


@dataclass
class GroupLayout:
    """Target groups of at most 64 neurons of one population each."""
    n: int
    group_first: torch.Tensor  # int32 [G]
    group_info: torch.Tensor   # int32 [G]: count | pop << 16
    group_of: torch.Tensor     # int64 [n]
    local_of: torch.Tensor     # int64 [n]

    @property
    def n_groups(self) -> int:
        return int(self.group_first.numel())


def group_layout(pop_sizes: list[int], device="cpu") -> GroupLayout:
    firsts, infos = [], []
    base = 0
    for p, size in enumerate(pop_sizes):
        for g0 in range(0, size, GROUP):
            firsts.append(base + g0)
            infos.append(min(GROUP, size - g0) | (p << 16))
        base += size
    group_first = torch.tensor(firsts, dtype=torch.int32)
    group_info = torch.tensor(infos, dtype=torch.int32)
    counts = (group_info & 0xFFFF).to(torch.int64)
    group_of = torch.repeat_interleave(torch.arange(len(firsts), dtype=torch.int64), counts)
    local_of = torch.arange(base, dtype=torch.int64) - group_first.to(torch.int64)[group_of]
    return GroupLayout(base, group_first.to(device), group_info.to(device), group_of.to(device),
                       local_of.to(device))


@dataclass
class Segments:
    seg_off: torch.Tensor  # int64 [64 * G * G + 1]
    links: torch.Tensor    # int32 [E, 2]: (delay << 16 | local target, weight bits)
    extra: torch.Tensor | None = None  # the canonical table's per-link payload ([E] or [k, E]), in segment order


def slot_of(layout: GroupLayout) -> torch.Tensor:
    """int64 [n]: a neuron's source slot, 64 * (its group) + (its index in the group) -- the id the step's
    spike lists and bitmaps carry."""
    return layout.group_of * GROUP + layout.local_of


def _payload_rows(out: torch.Tensor, src: torch.Tensor):
    """(out row, source row) pairs of a per-link payload: itself when 1-D, else its k rows.  Each row is
    permuted as a 1-D tensor: torch's row gather and scatter on a contiguous [E, 4] int32 table return wrong
    rows once the table passes about 2 GiB on this ROCm build, where 1-D indexing is exact."""
    return ((out, src),) if src.dim() == 1 else zip(out, src)


def segments(canon: Canonical, layout: GroupLayout) -> Segments:
    """Group segments: the links of source slot p into group g at [seg_off[p*G + g], seg_off[p*G + g + 1]),
    in canonical order (a stable partition of each source's links by target group); padding slots of
    partly filled groups have no links."""
    dev = canon.tgt.device
    n, G = canon.n, layout.n_groups
    group_of = layout.group_of.to(dev)
    local_of = layout.local_of.to(dev)
    slots = GROUP * G
    src = torch.repeat_interleave(slot_of(layout).to(dev), canon.out_degree())
    tg = canon.tgt.to(torch.int64)
    key = src * G + group_of[tg]
    del src
    counts = torch.zeros(slots * G, dtype=torch.int64, device=dev)
    for a in range(0, key.numel(), SORT_CHUNK):
        counts += torch.bincount(key[a:a + SORT_CHUNK], minlength=slots * G)
    seg_off = torch.zeros(slots * G + 1, dtype=torch.int64, device=dev)
    seg_off[1:] = torch.cumsum(counts, 0)
    del counts
    # (links are grouped by source already: a stable sort by (source slot, group) chunk by chunk of sources)
    links = torch.empty(key.numel(), 2, dtype=torch.int32, device=dev)
    extra = torch.empty_like(canon.extra) if canon.extra is not None else None
    E = key.numel()
    sp = canon.src_ptr
    a = 0
    while a < E:
        # a chunk of whole sources, at most SORT_CHUNK links (one source's links always fit)
        b = min(E, a + SORT_CHUNK)
        if b < E:
            b = int(sp[int(torch.searchsorted(sp, torch.tensor(b, device=dev), right=True)) - 1])
            if b <= a:
                raise ValueError(f"a source with more than {SORT_CHUNK} links")
        k = key[a:b]
        order = torch.argsort(k, stable=True)
        t = tg[a:b][order]
        x = (canon.d[a:b].to(torch.int64)[order] << 16) | local_of[t]
        links[a:b, 0] = x.to(torch.int32)
        links[a:b, 1] = canon.w[a:b][order].contiguous().view(torch.int32)
        if extra is not None:
            for row_out, row in _payload_rows(extra, canon.extra):
                row_out[a:b] = row[a:b][order]
        del k, order, t, x
        a = b
    del key
    return Segments(seg_off, links, extra)


ROWS_WIDE, ROWS_COMPACT, ROWS_SLOT16 = 0, 1, 2
COMPACT_WEIGHTS = 256      # distinct weights a compact row table can index
COMPACT_SLOTS = 1 << 17    # source slots a compact entry can name


@dataclass
class Rows:
    """Row-intake tables: each target's in-links in contract order (ascending canonical link index).

    Each group is one block of 16-byte cells [depth][64]: cell q of lane l is at cell index
    row_base[g] + 64 q + l and holds the lane's entries per*q .. per*q + per - 1, so that every lane loads
    its row 16 bytes at a time and a wavefront's load is 1 KB of contiguous memory.  Two entry formats:
      wide    (per = 2): two int32 words (source slot << 7 | (delay - 1), weight bits);
      compact (per = 4): one int32 word  source slot << 15 | (delay - 1) << 8 | weight index, the weight
                         being wtab[index] -- for networks with at most 256 distinct weights and
                         2^17 source slots (Brunel-type networks: half the bytes).
      slot16  (per = 8): the 16-bit source slot alone (entry 2k in the low half of word k), for networks
                         with one delay for every link whose weights are set by the source's side of one
                         slot split (two source populations: Brunel-type networks) and 64 G <= 65535; wtab
                         holds {w_lo, w_hi, split slot bits}.  Half the compact entry: the row intake of a
                         bursting network streams half the bytes.
    Entries past a target's in-links name the sentinel source slot 64 G, which never spikes (the kernel
    keeps a zero bitmap word for it), so the walk needs no bounds test.
    """
    rows: torch.Tensor       # int32 [cells * 4]
    row_base: torch.Tensor   # int64 [G]: first cell of each group
    row_len: torch.Tensor    # int32 [64 G]: in-links per target slot
    row_depth: torch.Tensor  # int32 [G]: cells per lane
    fmt: int                 # ROWS_WIDE, ROWS_COMPACT or ROWS_SLOT16
    wtab: torch.Tensor | None  # float32 [COMPACT_WEIGHTS] (compact)
    uniform_delay: int       # the delay in steps of every link when all are equal, else 0


def _as_u32_bits(v: torch.Tensor) -> torch.Tensor:
    """int64 values in [0, 2^32) as the int32 tensor with the same 32 bits."""
    return torch.where(v >= 2 ** 31, v - 2 ** 32, v).to(torch.int32)


def _slot_split(canon: Canonical, layout: GroupLayout):
    """(split slot, w_lo bits, w_hi bits) when every link's weight is set by its source slot's side of one
    split (w_lo below, w_hi from it on) and every link has one delay, else None."""
    E = canon.n_links
    if E == 0 or not bool((canon.d == canon.d[0]).all()):
        return None
    slot = torch.repeat_interleave(slot_of(layout).to(canon.tgt.device), canon.out_degree())
    wbits = canon.w.contiguous().view(torch.int32)
    uniq = torch.unique(wbits)
    if uniq.numel() > 2:
        return None
    first = int(torch.argmin(slot))
    w_lo = int(wbits[first])
    other = wbits != w_lo
    if not bool(other.any()):
        return 0, w_lo, w_lo
    split = int(slot[other].min())
    if int(slot[~other].max()) >= split:
        return None
    return split, w_lo, int(uniq[uniq != w_lo][0])


def rows(canon: Canonical, layout: GroupLayout, fmt: str = "auto") -> Rows:
    """fmt: "auto" (slot16 where it applies, else compact where it fits), "wide", "compact" or "slot16"."""
    dev = canon.tgt.device
    n, G = canon.n, layout.n_groups
    if fmt in ("auto", "slot16"):
        sp = _slot_split(canon, layout) if GROUP * G <= 0xffff else None
        if sp is not None:
            return _rows_slot16(canon, layout, *sp)
        if fmt == "slot16":
            raise ValueError("slot16 rows need one delay for every link, weights set by one source-slot split "
                             "and at most 65,535 source slots")
    wbits = canon.w.contiguous().view(torch.int32)
    uniq = None
    if fmt in ("auto", "compact") and GROUP * G <= COMPACT_SLOTS:
        uniq = torch.unique(wbits)
        if uniq.numel() > COMPACT_WEIGHTS:
            uniq = None
    if fmt == "compact" and uniq is None:
        raise ValueError("compact rows need at most 256 distinct weights and 2^17 source slots")
    compact = uniq is not None and fmt != "wide"
    per = 4 if compact else 2
    slot = slot_of(layout).to(dev)
    group_of = layout.group_of.to(dev)
    local_of = layout.local_of.to(dev)
    tg = canon.tgt.to(torch.int64)
    E = int(tg.numel())
    deg = torch.bincount(tg, minlength=n)
    row_len = torch.zeros(GROUP * G, dtype=torch.int64, device=dev)
    row_len[slot] = deg
    depth = (row_len.view(G, GROUP).max(dim=1).values + per - 1) // per
    row_base = torch.zeros(G, dtype=torch.int64, device=dev)
    if G > 1:
        row_base[1:] = torch.cumsum(depth * GROUP, 0)[:-1]
    cells = int((depth * GROUP).sum())
    order = torch.argsort(tg, stable=True)  # by target, canonical order within each target
    t_sorted = tg[order]
    row_start = torch.cumsum(deg, 0) - deg
    rank = torch.arange(E, device=dev) - row_start[t_sorted]
    cell = row_base[group_of[t_sorted]] + (rank // per) * GROUP + local_of[t_sorted]
    word = cell * 4 + (rank % per) * (1 if compact else 2)
    del rank, row_start, cell, t_sorted
    src = torch.repeat_interleave(slot, canon.out_degree())[order]
    dm1 = canon.d.to(torch.int64)[order] - 1
    sentinel = GROUP * G
    out = torch.zeros(cells * 4, dtype=torch.int32, device=dev)
    fill = (sentinel << 15) if compact else (sentinel << 7)
    fill = fill - 2 ** 32 if fill >= 2 ** 31 else fill  # the same 32 bits as int32
    if compact:
        out[:] = fill
    else:
        out.view(-1, 2)[:, 0] = fill
    wtab = None
    if compact:
        widx = torch.searchsorted(uniq, wbits[order].contiguous())
        out[word] = _as_u32_bits((src << 15) | (dm1 << 8) | widx)
        wtab = torch.zeros(COMPACT_WEIGHTS, dtype=torch.int32, device=dev)
        wtab[:uniq.numel()] = uniq
        wtab = wtab.view(torch.float32)
    else:
        out[word] = _as_u32_bits((src << 7) | dm1)
        out[word + 1] = wbits[order]
    del src, dm1, order, word
    uniform = 0
    if E and bool((canon.d == canon.d[0]).all()):
        uniform = int(canon.d[0])
    return Rows(out, row_base, row_len.to(torch.int32), depth.to(torch.int32),
                ROWS_COMPACT if compact else ROWS_WIDE, wtab, uniform)


def _rows_slot16(canon: Canonical, layout: GroupLayout, split: int, w_lo: int, w_hi: int) -> Rows:
    """The slot16 row table (see Rows): eight 16-bit source slots a cell, in contract order per target."""
    dev = canon.tgt.device
    n, G = canon.n, layout.n_groups
    per = 8
    slot = slot_of(layout).to(dev)
    group_of = layout.group_of.to(dev)
    local_of = layout.local_of.to(dev)
    tg = canon.tgt.to(torch.int64)
    E = int(tg.numel())
    deg = torch.bincount(tg, minlength=n)
    row_len = torch.zeros(GROUP * G, dtype=torch.int64, device=dev)
    row_len[slot] = deg
    depth = (row_len.view(G, GROUP).max(dim=1).values + per - 1) // per
    row_base = torch.zeros(G, dtype=torch.int64, device=dev)
    if G > 1:
        row_base[1:] = torch.cumsum(depth * GROUP, 0)[:-1]
    cells = int((depth * GROUP).sum())
    order = torch.argsort(tg, stable=True)  # by target, canonical order within each target
    t_sorted = tg[order]
    row_start = torch.cumsum(deg, 0) - deg
    rank = torch.arange(E, device=dev) - row_start[t_sorted]
    cell = row_base[group_of[t_sorted]] + (rank // per) * GROUP + local_of[t_sorted]
    half = cell * per + rank % per  # the entry's 16-bit position: entry 2k is the low half of word k
    del rank, row_start, cell, t_sorted
    src = torch.repeat_interleave(slot, canon.out_degree())[order]
    sentinel = GROUP * G
    out16 = torch.full((cells * per,), sentinel - 65536 if sentinel >= 32768 else sentinel, dtype=torch.int16,
                       device=dev)
    out16[half] = torch.where(src >= 32768, src - 65536, src).to(torch.int16)
    del src, order, half
    wtab = torch.tensor([w_lo, w_hi, split], dtype=torch.int32, device=dev).view(torch.float32)
    return Rows(out16.view(torch.int32), row_base, row_len.to(torch.int32), depth.to(torch.int32), ROWS_SLOT16,
                wtab, int(canon.d[0]))
