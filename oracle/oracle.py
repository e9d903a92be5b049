"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (oracle/_build/liboracle.so) and the small
NumPy brute-force restatements that pin it.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the product
(naer2_amd) never does.  Every function cites the reference lines it restates.
"""
from __future__ import annotations

import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
LIF_DELTA, LIF_EXP = 1, 2
POISSON_K = 32

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def use_native() -> Path:
    """Build the oracle for this host (-O3 -march=native, always rebuilt: a build made on another machine
    may use instructions this one lacks) and load that build instead of the portable one.  For the CPU
    baseline's timing only; it must be called before the first load()."""
    global LIB
    native = HERE / "_build" / "liboracle_native.so"
    if LIB == native:
        return LIB
    if _lib is not None:
        raise RuntimeError("use_native() after the portable oracle was loaded")
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "native"], check=True)
    LIB = native
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        lib.or_philox4x32_10.argtypes = [P, P, P]
        lib.or_philox4x32_10.restype = None
        lib.or_poisson_count.argtypes = [ctypes.c_uint32, P]
        lib.or_poisson_count.restype = ctypes.c_int
        lib.or_run.argtypes = [ctypes.c_int, ctypes.c_int32, ctypes.c_int32, P, P, ctypes.c_int32, ctypes.c_uint64,
                               ctypes.c_uint32, P, P, P, P, P, P, P, P, P, ctypes.c_int64, ctypes.c_int32, P,
                               ctypes.c_int32, P]
        lib.or_run.restype = ctypes.c_int
        lib.or_run_edges.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, ctypes.c_int32, ctypes.c_uint64,
                                     ctypes.c_uint32, P, P, P, P, P, P, P, P, P, P, P, P, P, ctypes.c_int64,
                                     ctypes.c_int32, P, P]
        lib.or_run_edges.restype = ctypes.c_int
        lib.or_mt_create.argtypes = [ctypes.c_int32, P, P, P, P, ctypes.c_int32]
        lib.or_mt_create.restype = P
        lib.or_mt_destroy.argtypes = [P]
        lib.or_mt_destroy.restype = None
        lib.or_mt_run.argtypes = [P, ctypes.c_int, ctypes.c_int32, P, P, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint32,
                                  P, P, P, P, P, ctypes.c_int64, ctypes.c_int32, P, P]
        lib.or_mt_run.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of one block (Salmon et al., SC'11)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().or_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def poisson_count(u: int, thr: np.ndarray) -> int:
    t = np.ascontiguousarray(thr, dtype=np.uint32)
    return int(load().or_poisson_count(ctypes.c_uint32(u), _ptr(t)))


class State:
    """The oracle's copy of a network's state (numpy, host)."""

    def __init__(self, n: int, d_slots: int, v, isyn=None, ref=None, ring=None, prev_spk=None):
        self.v = np.ascontiguousarray(v, dtype=np.float32).copy()
        self.isyn = None if isyn is None else np.ascontiguousarray(isyn, dtype=np.float32).copy()
        self.ref = np.zeros(n, np.int32) if ref is None else np.ascontiguousarray(ref, dtype=np.int32).copy()
        self.ring = (np.zeros((d_slots, n), np.float32) if ring is None
                     else np.ascontiguousarray(ring, dtype=np.float32).copy())
        self.prev_spk = (np.zeros(n, np.uint8) if prev_spk is None
                         else np.ascontiguousarray(prev_spk, dtype=np.uint8).copy())


def run(model: int, pop_first, pops_bytes: bytes, d_slots: int, seed: int, instance: int,
        src_ptr, tgt, w, delay, state: State, step0: int, n_steps: int, record: bool = True,
        accum_mode: int = 0):
    """Runs the oracle in place on `state`; returns (raster bool [n_steps, n] or None, events)."""
    lib = load()
    pf = np.ascontiguousarray(pop_first, dtype=np.int32)
    n = int(pf[-1])
    pops = ctypes.create_string_buffer(pops_bytes, len(pops_bytes))
    sp = np.ascontiguousarray(src_ptr, dtype=np.int64)
    tg = np.ascontiguousarray(tgt, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.float32)
    dd = np.ascontiguousarray(delay, dtype=np.int32)
    raster = np.zeros((n_steps, n), dtype=np.uint8) if record else None
    ev = ctypes.c_int64(0)
    rc = lib.or_run(model, n, len(pf) - 1, _ptr(pf), ctypes.cast(pops, ctypes.c_void_p), d_slots, seed, instance,
                    _ptr(sp), _ptr(tg), _ptr(ww), _ptr(dd), _ptr(state.v), _ptr(state.isyn), _ptr(state.ref),
                    _ptr(state.ring), _ptr(state.prev_spk), step0, n_steps, _ptr(raster), accum_mode,
                    ctypes.byref(ev))
    if rc != 0:
        raise RuntimeError(f"or_run failed ({rc})")
    return (raster.astype(bool) if record else None), int(ev.value)


def run_edges(pop_first, pops_bytes: bytes, d_slots: int, seed: int, instance: int, src_ptr, tgt, w, delay,
              kind, port, attr2, state: State, ring_v, step0: int, n_steps: int, record: bool = True):
    """The exp-LIF step with edge kinds (0 static, 1 gain w * attr2, 2 plastic w += attr2 before delivery) and
    ports (0 current, 1 potential through ring_v [d_slots, n]), in place on `state`, `w` (plastic links; must be
    a contiguous float32 array) and `ring_v`; returns (raster or None, events)."""
    lib = load()
    pf = np.ascontiguousarray(pop_first, dtype=np.int32)
    n = int(pf[-1])
    pops = ctypes.create_string_buffer(pops_bytes, len(pops_bytes))
    sp = np.ascontiguousarray(src_ptr, dtype=np.int64)
    tg = np.ascontiguousarray(tgt, dtype=np.int32)
    dd = np.ascontiguousarray(delay, dtype=np.int32)
    kk = np.ascontiguousarray(kind, dtype=np.uint8)
    pp = np.ascontiguousarray(port, dtype=np.uint8)
    aa = np.ascontiguousarray(attr2, dtype=np.float32)
    for a in (w, ring_v):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise ValueError("w and ring_v are updated in place: contiguous float32 arrays")
    raster = np.zeros((n_steps, n), dtype=np.uint8) if record else None
    ev = ctypes.c_int64(0)
    rc = lib.or_run_edges(n, len(pf) - 1, _ptr(pf), ctypes.cast(pops, ctypes.c_void_p), d_slots, seed, instance,
                          _ptr(sp), _ptr(tg), _ptr(w), _ptr(dd), _ptr(kk), _ptr(pp), _ptr(aa), _ptr(ring_v),
                          _ptr(state.v), _ptr(state.isyn), _ptr(state.ref), _ptr(state.ring), _ptr(state.prev_spk),
                          step0, n_steps, _ptr(raster), ctypes.byref(ev))
    if rc != 0:
        raise RuntimeError(f"or_run_edges failed ({rc})")
    return (raster.astype(bool) if record else None), int(ev.value)


class Parallel:
    """The oracle's step on `threads` host threads (OpenMP, contiguous target ranges): the same results as
    run(), for timing the CPU baseline at every core.  Setup (splitting the links) is done once here."""

    def __init__(self, n: int, src_ptr, tgt, w, delay, threads: int):
        lib = load()
        self._keep = [np.ascontiguousarray(src_ptr, dtype=np.int64), np.ascontiguousarray(tgt, dtype=np.int32),
                      np.ascontiguousarray(w, dtype=np.float32), np.ascontiguousarray(delay, dtype=np.int32)]
        self.n, self.threads = int(n), int(threads)
        self._h = lib.or_mt_create(self.n, *[_ptr(a) for a in self._keep], self.threads)
        if not self._h:
            raise RuntimeError("or_mt_create failed")
        self._keep = None  # the handle holds its own split copies

    def run(self, model: int, pop_first, pops_bytes: bytes, d_slots: int, seed: int, instance: int, state: State,
            step0: int, n_steps: int, record: bool = True):
        lib = load()
        pf = np.ascontiguousarray(pop_first, dtype=np.int32)
        pops = ctypes.create_string_buffer(pops_bytes, len(pops_bytes))
        raster = np.zeros((n_steps, self.n), dtype=np.uint8) if record else None
        ev = ctypes.c_int64(0)
        rc = lib.or_mt_run(self._h, model, len(pf) - 1, _ptr(pf), ctypes.cast(pops, ctypes.c_void_p), d_slots, seed,
                           instance, _ptr(state.v), _ptr(state.isyn), _ptr(state.ref), _ptr(state.ring),
                           _ptr(state.prev_spk), step0, n_steps, _ptr(raster), ctypes.byref(ev))
        if rc != 0:
            raise RuntimeError(f"or_mt_run failed ({rc})")
        return (raster.astype(bool) if record else None), int(ev.value)

    def close(self):
        if self._h:
            load().or_mt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- NumPy brute force (tiny sizes)

def philox_py(ctr, key) -> list[int]:
    """Philox4x32-10 in pure Python integers (an independent restatement of the same published method)."""
    c = [int(x) & 0xFFFFFFFF for x in ctr]
    k = [int(x) & 0xFFFFFFFF for x in key]
    for _ in range(10):
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
        k = [(k[0] + 0x9E3779B9) & 0xFFFFFFFF, (k[1] + 0xBB67AE85) & 0xFFFFFFFF]
    return c


def ring_inputs_dense(n: int, src_ptr, tgt, w, delay, raster: np.ndarray, prev_spk0=None) -> np.ndarray:
    """float64 [n_steps, n]: the recurrent input that arrives at each step, summed in float64 from a dense
    per-delay weight matrix and the raster (test/test_population_fused_gpu.py:36-44, _expected: every
    link delivered once, at its source's spike step + its delay).  Spikes of step -1 come from prev_spk0."""
    n_steps = raster.shape[0]
    sp = np.asarray(src_ptr)
    src = np.repeat(np.arange(n), np.diff(sp))
    out = np.zeros((n_steps, n), dtype=np.float64)
    full = np.zeros((n_steps + 1, n), dtype=bool)  # row 0 = step -1
    if prev_spk0 is not None:
        full[0] = np.asarray(prev_spk0, dtype=bool)
    full[1:] = raster
    for d in np.unique(delay):
        m = np.asarray(delay) == d
        W = np.zeros((n, n), dtype=np.float64)  # [target, source]
        np.add.at(W, (np.asarray(tgt)[m], src[m]), np.asarray(w, dtype=np.float64)[m])
        # input at step n from spikes at n - d (row n - d + 1 of full)
        for s_row in range(0, n_steps + 1):
            arr = s_row - 1 + d
            if 0 <= arr < n_steps and full[s_row].any():
                out[arr] += W @ full[s_row].astype(np.float64)
    return out


def exact_lif_exp(v0: float, i0: float, t: float, t_mem: float, t_syn: float, r_mem: float, v_rest: float):
    """Closed-form solution (float64) of tau_m dV/dt = -(V - V_rest) + R I, dI/dt = -I/tau_s: the exact
    propagator of Rotter & Diesmann (1999) the reference integrates (test/networks/PD14/neuron.py:37-50)."""
    pv = math.exp(-t / t_mem)
    pi = math.exp(-t / t_syn)
    return v_rest + pv * (v0 - v_rest) + r_mem * t_syn / (t_mem - t_syn) * (pv - pi) * i0, pi * i0
