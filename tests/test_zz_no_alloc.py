"""n2h_rlif_backward_noalloc (the BPTT backward with the library's static scratch instead of a work area)
against n2h_rlif_backward (a caller's work area), through the C-ABI: identical input adjoints, bit for bit,
on two calls in a row, the second smaller; and sizes beyond the scratch refused."""
from __future__ import annotations

import ctypes
import math
import sys

import pytest
import torch

from naer2_amd import rlif as rl
from naer2_amd.rlif import RLIFConfig, poisson_input

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bind():
    lib = rl._bind(rl._lib.load())
    vp = ctypes.c_void_p
    lib.n2h_rlif_backward_noalloc.argtypes = [ctypes.POINTER(rl.RLIFLayer), vp, vp, vp, vp, vp, vp, vp]
    lib.n2h_rlif_backward_noalloc.restype = ctypes.c_int32
    return lib


def _inputs(B, C, H, T, K, seed):
    """Forward through the C-ABI: (layer, u, w_rec_t, spikes, checkpoints, gso)."""
    lib = _bind()
    g = torch.Generator(device="cuda").manual_seed(seed)
    cfg = RLIFConfig(checkpoint=K)
    x = poisson_input(B, C, T, 1.0, g)
    w_in = torch.randn(H, C, generator=g, device="cuda") / math.sqrt(C)
    w_rec = torch.randn(H, H, generator=g, device="cuda") / math.sqrt(H)
    u = (x.reshape(T * B, C) @ w_in.t()).reshape(T, B, H).contiguous()
    w_rec_t = w_rec.t().contiguous()
    spikes = torch.empty(T, B, H, device="cuda")
    ckpt = torch.empty((T + K - 1) // K, 4, B, H, device="cuda")
    state = torch.empty(4, B, H, device="cuda")
    lay = rl.layer(B, H, T, cfg)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rl._check(lib.n2h_rlif_forward(ctypes.byref(lay), rl._ptr(u), rl._ptr(w_rec_t), rl._ptr(state), rl._ptr(spikes),
                                   rl._ptr(ckpt), stream), "n2h_rlif_forward")
    gso = torch.randn(T, B, H, generator=g, device="cuda") * 0.1
    return lay, u, w_rec_t, spikes, ckpt, gso


def _backward(lay, u, w_rec_t, spikes, ckpt, gso, noalloc):
    lib = _bind()
    T, B, H = gso.shape
    g_i = torch.full((T, B, H), float("nan"), device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = lib.n2h_rlif_backward_noalloc(ctypes.byref(lay), rl._ptr(u), rl._ptr(w_rec_t), rl._ptr(spikes),
                                           rl._ptr(ckpt), rl._ptr(gso), rl._ptr(g_i), stream)
        rl._check(rc, "n2h_rlif_backward_noalloc")
    else:
        work = torch.full((int(lib.n2h_rlif_work_floats(ctypes.byref(lay))),), float("nan"), device="cuda")
        rl._check(lib.n2h_rlif_backward(ctypes.byref(lay), rl._ptr(u), rl._ptr(w_rec_t), rl._ptr(spikes),
                                        rl._ptr(ckpt), rl._ptr(gso), rl._ptr(g_i), rl._ptr(work), stream),
                  "n2h_rlif_backward")
    torch.cuda.synchronize()
    return g_i


def test_noalloc_backward_equals_the_work_area_backward():
    for B, C, H, T, K, seed in ((128, 700, 1024, 100, 25, 11), (24, 40, 160, 37, 9, 12)):
        _say(f"case: rlif backward noalloc vs work area B={B} H={H} T={T} K={K}")
        args = _inputs(B, C, H, T, K, seed)
        a = _backward(*args, noalloc=False)
        b = _backward(*args, noalloc=True)
        assert torch.isfinite(a).all() and float(a.abs().max()) > 0
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_noalloc_backward_refuses_work_beyond_its_scratch():
    _say("case: rlif backward noalloc refuses an oversized layer")
    lib = _bind()
    lay, u, w_rec_t, spikes, ckpt, gso = _inputs(8, 16, 64, 10, 5, 13)
    big = rl.layer(128, 1024, 100, RLIFConfig(checkpoint=100))  # (5 + 100) * 128 * 1024 floats > 2^22
    assert lib.n2h_rlif_work_floats(ctypes.byref(big)) > (1 << 22)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.n2h_rlif_backward_noalloc(ctypes.byref(big), rl._ptr(u), rl._ptr(w_rec_t), rl._ptr(spikes),
                                       rl._ptr(ckpt), rl._ptr(gso), rl._ptr(gso), stream)
    assert rc == -1
    buf = ctypes.create_string_buffer(512)
    lib.n2h_rlif_last_error(buf, len(buf))
    assert b"scratch" in buf.value
