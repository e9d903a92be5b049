"""Host-side logic on the CPU: the description API, connectivity rules, delay rounding, the group layout,
the group segments the kernel reads, raster <-> spike-word conversion, and the C-ABI's exports and
argument checks (no compute call: no GPU here)."""
from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import naer2_amd as n2
from naer2_amd import _lib
from naer2_amd.network import assemble, bits_to_raster, raster_to_bits
from naer2_amd.networks import PD14, brunel, pd14
from naer2_amd.nodes import FixedInDegree, FixedTotalNumber
from naer2_amd.topology import (ROWS_COMPACT, ROWS_SLOT16, ROWS_WIDE, canonical, delay_steps, group_layout, rows,
                                segments, slot_of)

ROOT = Path(__file__).resolve().parent.parent


def test_header_declares_exactly_the_exports():
    text = (ROOT / "include" / "naer2_hip.h").read_text()
    declared = set(re.findall(r"^(?:int|int64_t)\s+(n2h_\w+)\(", text, flags=re.M))
    assert declared == set(_lib.EXPORTS)
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.n2h_version() == 1


def test_plan_create_rejects_bad_arguments():
    lib = _lib.load()
    plan = ctypes.c_void_p()
    first = (ctypes.c_int32 * 2)(0, 10)
    pops = (_lib.PopParams * 1)()
    tab = _lib.Tables(n_groups=1, group_first=1, group_info=1, seg_off=1, links=1, n_links=1)
    cases = [
        (dict(model=3, n_neurons=10, n_pops=1, d_slots=1), "unknown model"),
        (dict(model=1, n_neurons=0, n_pops=1, d_slots=1), "n_neurons"),
        (dict(model=1, n_neurons=10, n_pops=0, d_slots=1), "n_pops"),
        (dict(model=1, n_neurons=10, n_pops=1, d_slots=129), "d_slots"),
    ]
    for kw, msg in cases:
        desc = _lib.ModelDesc(pop_first=first, pops=pops, **kw)
        rc = lib.n2h_plan_create(ctypes.byref(desc), ctypes.byref(tab), ctypes.byref(plan))
        assert rc == -1 and msg in _lib.last_error()
    bad = (ctypes.c_int32 * 2)(0, 9)
    desc = _lib.ModelDesc(model=1, n_neurons=10, n_pops=1, d_slots=1, pop_first=bad, pops=pops)
    assert lib.n2h_plan_create(ctypes.byref(desc), ctypes.byref(tab), ctypes.byref(plan)) == -1
    assert "pop_first" in _lib.last_error()
    pops[0].poisson_thr[3] = 5  # not non-decreasing after zeros? make it decreasing
    pops[0].poisson_thr[4] = 1
    desc = _lib.ModelDesc(model=1, n_neurons=10, n_pops=1, d_slots=1, pop_first=first, pops=pops)
    assert lib.n2h_plan_create(ctypes.byref(desc), ctypes.byref(tab), ctypes.byref(plan)) == -1
    assert "poisson_thr" in _lib.last_error()
    assert lib.n2h_step(None, 1, None) == -1
    assert lib.n2h_bind_ring(None, None, None, 2, None) == -1
    t = _lib.Tuning(order=0, xcd=8, stage=0, run=0)
    assert lib.n2h_set_tuning(None, ctypes.byref(t)) == -1 and "null" in _lib.last_error()


def test_fixed_in_degree_gives_every_target_exactly_k_inputs():
    net, pops = brunel(N_E=400, N_I=100, C_E=40, C_I=10, device="cpu")
    E, I = pops["E"], pops["I"]
    for c in net.connections:
        k = 40 if c.source is E else 10
        counts = torch.bincount(c.tgt, minlength=c.target.n)
        assert bool((counts == k).all())
        assert bool((c.src[1:] >= c.src[:-1]).all())  # source-major link order


def test_fixed_total_number_out_degree_is_the_reference_rule():
    p = 0.1009
    nx, ny = 2068, 2068
    q = math.log(1 - p) / math.log(1 - 1 / (nx * ny))
    assert FixedTotalNumber(p).per_source(nx, ny) == round(q / nx)
    total = 0
    for si, s in enumerate(PD14["size"]):
        for ti, t in enumerate(PD14["size"]):
            total += s * FixedTotalNumber(PD14["probabilities"][si][ti]).per_source(s, t)
    assert total == 298822986  # ~298.8 M synapses at full scale


def test_per_source_pattern_callable_is_supported():
    @n2.pattern
    def ring_next(position, source_region, target_region=None):
        n = len(target_region)
        return n2.Points(torch.tensor([[(position[0] + 1) % n]]))

    A = n2.LIFDelta() * 5
    conn = A >> ring_next >> A
    assert conn.tgt.tolist() == [1, 2, 3, 4, 0]
    conn.weight, conn.delay = 1.0, 0.1
    assert (A >> FixedInDegree(2) >> A | 0.5).delay.shape == (10,)


def test_delay_steps_round_and_clip():
    d = delay_steps(torch.tensor([0.1, 0.14, 0.15, 0.25, 1.5, 0.01], dtype=torch.float32), 0.1)
    assert d.tolist() == [1, 1, 2, 2, 15, 1]
    with pytest.raises(ValueError):
        delay_steps(torch.tensor([13.0]), 0.1)


def test_group_layout_never_straddles_populations():
    lay = group_layout([130, 64, 1])
    assert lay.group_first.tolist() == [0, 64, 128, 130, 194]
    assert (lay.group_info & 0xFFFF).tolist() == [64, 64, 2, 64, 1]
    assert (lay.group_info >> 16).tolist() == [0, 0, 0, 1, 2]
    assert lay.group_of[129].item() == 2 and lay.local_of[129].item() == 1
    assert slot_of(lay)[129].item() == 2 * 64 + 1


def test_segments_hold_each_sources_links_per_group_in_canonical_order():
    rng = np.random.default_rng(0)
    n = 300
    src = torch.from_numpy(rng.integers(0, n, 5000))
    tgt = torch.from_numpy(rng.integers(0, n, 5000))
    w = torch.from_numpy(rng.standard_normal(5000).astype(np.float32))
    d = torch.from_numpy(rng.integers(1, 20, 5000).astype(np.int32))
    can = canonical(n, src, tgt, w, d)
    lay = group_layout([200, 100])
    seg = segments(can, lay)
    G = lay.n_groups
    links = seg.links.numpy()
    gf = lay.group_first.numpy()
    sp = can.src_ptr.numpy()
    slot = (lay.group_of * 64 + lay.local_of).numpy()
    assert seg.seg_off.numel() == 64 * G * G + 1
    for s in range(n):
        row = range(sp[s], sp[s + 1])
        for g in range(G):
            p = slot[s]
            lo, hi = int(seg.seg_off[p * G + g]), int(seg.seg_off[p * G + g + 1])
            expect = [l for l in row if lay.group_of[can.tgt[l]].item() == g]
            got = links[lo:hi]
            assert len(got) == len(expect)
            for (x, y), l in zip(got, expect):
                assert gf[g] + (x & 0xFFFF) == can.tgt[l].item()
                assert (x >> 16) == can.d[l].item()
                assert np.int32(y).view(np.float32) == can.w[l].item()


def test_raster_bits_round_trip():
    lay = group_layout([70, 3, 64])
    r = np.random.default_rng(1).random((9, 137)) < 0.3
    assert np.array_equal(bits_to_raster(raster_to_bits(r, lay), lay, 137), r)


def test_assemble_pd14_small():
    net, pops = pd14(scale=0.02, device="cpu")
    a = assemble(net)
    assert a.n == sum(p.n for p in pops.values())
    assert a.isyn is not None and a.family == "lif_exp"
    assert 1 <= int(a.canon.d.min()) and a.d_slots <= 128
    # inhibitory weights negative, excitatory non-negative
    e23 = pops["E_23"]
    c = [c for c in net.connections if c.source is pops["I_23"]][0]
    assert float(c.weight.max()) <= 0.0
    c = [c for c in net.connections if c.source is e23][0]
    assert float(c.weight.min()) >= 0.0


def test_unknown_node_family_is_rejected_at_build():
    @n2.node
    class Custom:
        def __init__(self):
            self.x = 0.0

    A = Custom() * 3
    with pytest.raises(NotImplementedError):
        assemble(A.network)


@pytest.mark.parametrize("fmt", ["wide", "compact"])
def test_rows_hold_each_targets_in_links_in_canonical_order(fmt):
    rng = np.random.default_rng(3)
    n = 150
    src = torch.from_numpy(rng.integers(0, n, 4000))
    tgt = torch.from_numpy(rng.integers(0, n, 4000))
    if fmt == "wide":
        w = torch.from_numpy(rng.standard_normal(4000).astype(np.float32))
    else:  # a few distinct weights, as in Brunel-type networks (and -0.0 apart from +0.0)
        w = torch.from_numpy(rng.choice(np.array([0.2, -1.0, 0.0, -0.0, 3.5e-3], np.float32), 4000))
    d = torch.from_numpy(rng.integers(1, 129, 4000).astype(np.int32))
    can = canonical(n, src, tgt, w, d)
    lay = group_layout([100, 50])
    r = rows(can, lay, fmt="auto")
    assert r.fmt == (ROWS_WIDE if fmt == "wide" else ROWS_COMPACT)
    per = 2 if fmt == "wide" else 4
    slot = (lay.group_of * 64 + lay.local_of).numpy()
    src_of = np.repeat(np.arange(n), np.diff(can.src_ptr.numpy()))
    tg = can.tgt.numpy()
    words = r.rows.numpy().view(np.uint32)
    sentinel = 64 * lay.n_groups
    assert r.uniform_delay == 0
    for g in range(lay.n_groups):  # padding entries name the never-spiking sentinel slot
        for lane in range(64):
            ln = int(r.row_len[g * 64 + lane])
            for k in range(ln, int(r.row_depth[g]) * per):
                cell = int(r.row_base[g]) + 64 * (k // per) + lane
                x = words[cell * 4 + (k % per) * (2 if fmt == "wide" else 1)]
                assert (x >> (7 if fmt == "wide" else 15)) == sentinel
    for i in range(n):
        g, lane = int(lay.group_of[i]), int(lay.local_of[i])
        expect = np.nonzero(tg == i)[0]  # ascending canonical link index
        assert int(r.row_len[g * 64 + lane]) == len(expect)
        assert int(r.row_depth[g]) * per >= len(expect)
        for k, l in enumerate(expect):
            cell = int(r.row_base[g]) + 64 * (k // per) + lane
            if fmt == "wide":
                x, y = words[cell * 4 + 2 * (k % per)], words[cell * 4 + 2 * (k % per) + 1]
                assert (x >> 7) == slot[src_of[l]] and (x & 127) + 1 == can.d[l].item()
                assert np.uint32(y).view(np.float32).view(np.uint32) == can.w[l:l + 1].numpy().view(np.uint32)[0]
            else:
                x = words[cell * 4 + (k % per)]
                assert (x >> 15) == slot[src_of[l]] and ((x >> 8) & 127) + 1 == can.d[l].item()
                wt = r.wtab.numpy().view(np.uint32)[x & 255]
                assert wt == can.w[l:l + 1].numpy().view(np.uint32)[0]  # the same bits, -0.0 included


def test_released_connections_refuse_link_access():
    A = n2.LIFDelta() * 10
    B = n2.LIFDelta() * 10
    conn = A >> FixedInDegree(3) >> B
    conn.weight = 0.5
    conn.delay = 1.0
    assert conn.shape == (30,) and conn.src.dtype == torch.int32
    conn.release()
    assert conn.shape == (30,)
    for name in ("src", "tgt", "weight", "delay"):
        with pytest.raises(RuntimeError):
            getattr(conn, name)
    with pytest.raises(RuntimeError):
        conn.weight = 1.0


def test_rows_flag_a_uniform_delay():
    rng = np.random.default_rng(4)
    n = 90
    src = torch.from_numpy(rng.integers(0, n, 900))
    tgt = torch.from_numpy(rng.integers(0, n, 900))
    w = torch.full((900,), 0.25)
    d = torch.full((900,), 15, dtype=torch.int32)
    can = canonical(n, src, tgt, w, d)
    r = rows(can, group_layout([n]), fmt="compact")
    assert r.uniform_delay == 15 and r.fmt == ROWS_COMPACT
    r = rows(can, group_layout([n]))  # one weight, one delay: the 16-bit format, split 0
    assert r.uniform_delay == 15 and r.fmt == ROWS_SLOT16
    assert r.wtab.numpy().view(np.uint32).tolist() == [np.float32(0.25).view(np.uint32)] * 2 + [0]


def _brunel_like(seed, n_e=150, n_i=40, m=6000):
    """Links from E (weight 0.2) and I (weight -1.0) sources, one delay of 15 steps."""
    rng = np.random.default_rng(seed)
    n = n_e + n_i
    src = rng.integers(0, n, m)
    tgt = rng.integers(0, n, m)
    w = np.where(src < n_e, np.float32(0.2), np.float32(-1.0)).astype(np.float32)
    return n, torch.from_numpy(src), torch.from_numpy(tgt), torch.from_numpy(w), torch.full((m,), 15, dtype=torch.int32)


def test_slot16_rows_hold_each_targets_source_slots_in_canonical_order():
    n, src, tgt, w, d = _brunel_like(5)
    can = canonical(n, src, tgt, w, d)
    lay = group_layout([150, 40])
    r = rows(can, lay)
    assert r.fmt == ROWS_SLOT16 and r.uniform_delay == 15
    split = 64 * 3  # the I population's first group
    wt = r.wtab.numpy().view(np.uint32)
    assert wt.tolist() == [np.float32(0.2).view(np.uint32), np.float32(-1.0).view(np.uint32), split]
    slot = (lay.group_of * 64 + lay.local_of).numpy()
    src_of = np.repeat(np.arange(n), np.diff(can.src_ptr.numpy()))
    half = r.rows.numpy().view(np.uint16)  # little-endian: entry 2k is the low half of word k
    sentinel = 64 * lay.n_groups
    for i in range(n):
        g, lane = int(lay.group_of[i]), int(lay.local_of[i])
        expect = np.nonzero(can.tgt.numpy() == i)[0]
        assert int(r.row_len[g * 64 + lane]) == len(expect)
        got = [int(half[(int(r.row_base[g]) + 64 * (k // 8) + lane) * 8 + k % 8]) for k in range(int(r.row_depth[g]) * 8)]
        assert got[:len(expect)] == [int(slot[src_of[l]]) for l in expect]
        assert all(x == sentinel for x in got[len(expect):])
        # the weight the walk takes for each entry is the link's
        for x, l in zip(got, expect):
            assert (wt[1] if x >= wt[2] else wt[0]) == can.w[l:l + 1].numpy().view(np.uint32)[0]


def test_slot16_rows_only_where_they_apply():
    n, src, tgt, w, d = _brunel_like(6)
    lay = group_layout([150, 40])
    bad_w = w.clone()
    bad_w[0] = 0.3  # a third weight
    mixed = w.clone()
    mixed[src == 0] = -1.0  # an E source with the I weight: no single split
    bad_d = d.clone()
    bad_d[0] = 3  # a second delay
    for ww, dd in ((bad_w, d), (mixed, d), (w, bad_d)):
        can = canonical(n, src, tgt, ww, dd)
        assert rows(can, lay).fmt == ROWS_COMPACT
        with pytest.raises(ValueError, match="slot16"):
            rows(can, lay, fmt="slot16")
    # the sentinel slot must fit 16 bits: 1,024 groups are too many
    big = group_layout([64 * 1024])
    m = 2000
    can = canonical(64 * 1024, torch.arange(m), torch.arange(m), torch.full((m,), 0.5), torch.full((m,), 2, dtype=torch.int32))
    assert rows(can, big).fmt == ROWS_COMPACT


def test_tables_built_in_chunks_equal_whole(monkeypatch):
    """Past 2^30 links the canonical and segment tables are sorted chunk by chunk (torch's sort takes at most
    2^31 - 1 elements): forced on a small network, the tables are identical."""
    from naer2_amd import topology as T

    g = torch.Generator().manual_seed(0)
    n, E = 500, 20000
    src = torch.randint(0, n, (E,), generator=g)
    tgt = torch.randint(0, n, (E,), generator=g)
    w = torch.randn(E, generator=g)
    d = torch.randint(1, 5, (E,), generator=g).to(torch.int32)
    lay = T.group_layout([300, 200])
    c1 = T.canonical(n, src, tgt, w, d)
    s1 = T.segments(c1, lay)
    monkeypatch.setattr(T, "SORT_CHUNK", 997)
    c2 = T.canonical(n, src, tgt, w, d)
    s2 = T.segments(c2, lay)
    for a, b in ((c1.src_ptr, c2.src_ptr), (c1.tgt, c2.tgt), (c1.w, c2.w), (c1.d, c2.d),
                 (s1.seg_off, s2.seg_off), (s1.links, s2.links)):
        assert torch.equal(a, b)


def test_rlif_layer_splits_the_decays_exactly_enough():
    """The BPTT layer's decays go to the kernels as float + float remainder: hi + lo equals the double-
    precision decay to ~2^-48, hi alone only to ~2^-24."""
    import math

    from naer2_amd.rlif import RLIFConfig, layer
    cfg = RLIFConfig()
    lay = layer(128, 1024, 100, cfg)
    for hi, lo, exact in ((lay.alpha, lay.alpha_lo, math.exp(-1 / 5)), (lay.beta, lay.beta_lo, math.exp(-1 / 20))):
        assert abs(hi - exact) > 1e-10  # the float alone is off in its last bits
        assert abs((hi + lo) - exact) < 1e-14 * exact
    assert (lay.batch, lay.n_hidden, lay.n_steps, lay.checkpoint) == (128, 1024, 100, 25)
