"""Host side of the rate-loss adjoint (naer2_amd/adjoint.py, n2h_adj_*) on the CPU: the out-link tables the
windowed reverse step walks, decoded link by link against the canonical CSR they come from, and the C-ABI's
work sizing and argument errors (calls that fail validation before any GPU work)."""
from __future__ import annotations

import ctypes
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from naer2_amd import _lib
from naer2_amd.adjoint import AdjNet, Surrogate, _bind, out_links, pure_ranges
from naer2_amd.topology import canonical

N2H_EINVAL = -1  # include/naer2_hip.h: a bad argument


def _fake_runtime(seed=4, n=70, e=900, d_max=12):
    """A canonical CSR of e links given in a shuffled source order (so canonical order within a source is
    not target order), 3 populations, delays 1..d_max."""
    g = torch.Generator().manual_seed(seed)
    src = torch.randint(0, n, (e,), generator=g)
    tgt = torch.randint(0, n, (e,), generator=g)
    w = torch.randn(e, generator=g) * 10.0
    d = torch.randint(1, d_max + 1, (e,), generator=g)
    pop_first = [0, 25, 50, n]
    return SimpleNamespace(canon=canonical(n, src, tgt, w, d), device="cpu", n=n, pop_first=pop_first)


def _decode(links):
    x = links[:, 0].numpy().astype(np.int64) & 0xFFFFFFFF
    return x & 0x1FFFF, ((x >> 17) & 127) + 1, x >> 24, links[:, 1].contiguous().view(torch.float32).numpy()


@pytest.mark.parametrize("long_delay,parts,want_pure", [(1, 1, False), (4, 3, False), (4, 8, False), (4, 8, True),
                                                       (13, 2, False), (2, 16, True), (4, 12, True)])
def test_out_link_tables_split_every_link_once_in_canonical_order(long_delay, parts, want_pure):
    rt = _fake_runtime()
    c = rt.canon
    alpha = torch.tensor(np.random.default_rng(1).uniform(0.5, 2.0, (3, 3)), dtype=torch.float32)
    sptr, sl, lptr, ll, rpop = out_links(rt, alpha, long_delay, parts, pure=want_pure)
    n = rt.n
    pure = pure_ranges(rt.pop_first, parts) if (want_pure and parts > 1) else None
    # population-pure ranges where asked for and possible (a multiple of 8, at least one per population)
    assert (rpop is None) == (pure is None) and (pure is not None) == (want_pure and parts in (8, 16))
    if pure is not None:
        assert np.array_equal(rpop.numpy(), pure[1])
        range_of = lambda t: np.searchsorted(pure[0], t, side="right") - 1
    else:
        range_of = lambda t: (t * parts) // n
    assert sptr.shape == (n + 1,) and lptr.shape == (n * parts + 1,)
    pop = np.repeat(np.arange(3), np.diff(rt.pop_first))
    ptr, tgt, dl, w = c.src_ptr.numpy(), c.tgt.numpy(), c.d.numpy(), c.w.numpy()
    seen = np.zeros(c.n_links, dtype=int)
    tables = ((sptr.numpy(), sl, 1, lambda dd: dd < long_delay), (lptr.numpy(), ll, parts, lambda dd: dd >= long_delay))
    for tptr, links, P, keep in tables:
        if tptr[-1] == 0:
            continue  # an empty table (one padding row)
        t_tgt, t_d, t_pop, t_w = _decode(links)
        for j in range(n):
            canon_idx = np.arange(ptr[j], ptr[j + 1])
            mine = canon_idx[keep(dl[canon_idx])]
            got = []
            for k in range(P):
                run = range(tptr[j * P + k], tptr[j * P + k + 1])
                # the run: source j's links into target range k, in canonical order
                want = mine[(range_of(tgt[mine]) if P > 1 else 0 * tgt[mine]) == k]
                assert len(run) == len(want)
                for e, ci in zip(run, want):
                    assert t_tgt[e] == tgt[ci] and t_d[e] == dl[ci] and t_pop[e] == pop[tgt[ci]]
                    base = np.float32(w[ci]) / np.float32(alpha[pop[tgt[ci]], pop[j]].item())
                    assert t_w[e] == base  # the base weight, bit for bit
                    seen[ci] += 1
                got += list(want)
            assert got == sorted(got) or P > 1  # one run: canonical order throughout
    assert (seen == 1).all()  # every link in exactly one table, once


def test_work_words_follow_the_list_layout():
    lib = _bind(_lib.load())
    sizes = [77169, 77169, 77169, 77169]
    arr = (ctypes.c_int32 * 4)(*sizes)
    blocks = sum((s + 255) // 256 for s in sizes)
    # a sub-list takes the appends of two wavefronts (one parity) of the mark workgroups dealt to its XCD, 64
    # each: with 4 networks, each network's workgroups over 2 XCDs
    sub_cap = 64 * 2 * ((blocks // 4 + 1) // 2)
    lists = {4: 9 * 16 * 32 + 5 * 16 * sub_cap, 8: 9 * 16 * 32 + 9 * 16 * sub_cap, 3: 9 * 16 * 32 + 4 * 16 * sub_cap}
    one = (ctypes.c_int32 * 1)(212)  # one workgroup: its two wavefronts of a parity fill one sub-list
    assert lib.n2h_adj_work_words(one, 1, 4, 1) == 9 * 16 * 32 + 5 * 16 * 128
    # uneven networks: 391 workgroups over XCDs 0-3 (98 each at most), 4 over XCDs 4-7
    uneven = (ctypes.c_int32 * 2)(100000, 1000)
    assert lib.n2h_adj_work_words(uneven, 2, 4, 1) == 9 * 16 * 32 + 5 * 16 * 128 * 98
    # more networks than XCDs: network q on XCD q mod 8 (XCDs 0 and 1 take two networks of 2 workgroups)
    ten = (ctypes.c_int32 * 10)(*[300] * 10)
    assert lib.n2h_adj_work_words(ten, 10, 4, 1) == 9 * 16 * 32 + 5 * 16 * 128 * 4
    assert lib.n2h_adj_work_words(arr, 4, 4, 1) == lists[4]  # one launch a range: the lists only
    assert lib.n2h_adj_work_words(arr, 4, 8, 5) == lists[8]
    n = sum(sizes)
    assert lib.n2h_adj_work_words(arr, 4, 4, 16) == lists[4] + n * 16 * 8 * 4  # + the range sums
    assert lib.n2h_adj_work_words(arr, 4, 3, 8) == lists[3] + n * 8 * 8 * 4   # 3 steps padded to 4
    assert lib.n2h_adj_work_words(arr, 0, 4, 8) == -1
    assert lib.n2h_adj_work_words(arr, 4, 9, 8) == -1
    assert lib.n2h_adj_work_words(arr, 4, 4, 65) == -1
    bad = (ctypes.c_int32 * 1)(0)
    assert lib.n2h_adj_work_words(bad, 1, 4, 8) == -1


def test_adjoint_calls_refuse_bad_arguments_before_any_launch():
    lib = _bind(_lib.load())
    sur = Surrogate(25.0, 0.2, 0.01, 0.0)
    nets = (AdjNet * 1)()
    nets[0] = AdjNet(100, 8, 40, 4, 8, 0)  # every pointer null
    buf = ctypes.create_string_buffer(512)
    assert lib.n2h_adj_window(nets, 1, 0, 4, ctypes.byref(sur), None, 0, None) == N2H_EINVAL
    lib.n2h_adj_last_error(buf, len(buf))
    assert b"null pointer" in buf.value
    nets[0].long_delay = 9
    assert lib.n2h_adj_step(nets, 1, 0, 0, ctypes.byref(sur), None, 0, None) == N2H_EINVAL
    lib.n2h_adj_last_error(buf, len(buf))
    assert b"long_delay" in buf.value
    nets[0].long_delay, nets[0].long_parts = 4, 65
    assert lib.n2h_adj_step(nets, 1, 0, 0, ctypes.byref(sur), None, 0, None) == N2H_EINVAL
    lib.n2h_adj_last_error(buf, len(buf))
    assert b"long_parts" in buf.value
    bad_sur = Surrogate(0.0, 0.2, 0.01, 0.0)
    assert lib.n2h_adj_window(nets, 1, 0, 4, ctypes.byref(bad_sur), None, 0, None) == N2H_EINVAL
    assert lib.n2h_adj_window(nets, 0, 0, 4, ctypes.byref(sur), None, 0, None) == N2H_EINVAL


@pytest.mark.parametrize("sizes,parts", [([20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948], 8),
                                         ([20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948], 24),
                                         ([10, 0, 20], 8), ([1, 1, 1, 1, 1, 1, 1, 1], 16), ([5, 7], 40)])
def test_pure_ranges_cover_every_population_in_order(sizes, parts):
    pf = np.concatenate([[0], np.cumsum(sizes)])
    first, pop = pure_ranges(pf, parts)
    assert len(pop) == parts and first[0] == 0 and first[-1] == pf[-1]
    assert (np.diff(first) >= 0).all()  # ranges of populations smaller than their share may be empty
    assert (np.diff(pop) >= 0).all()  # ranges in target order
    for k in range(parts):  # every range inside one population
        assert pf[pop[k]] <= first[k] and first[k + 1] <= pf[pop[k] + 1]
    assert set(pop.tolist()) == {p for p, s in enumerate(sizes) if s > 0}
    assert pure_ranges(pf, 12) is None  # not a multiple of 8
    assert pure_ranges(pf, 0) is None or len(sizes) == 0
