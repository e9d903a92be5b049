"""ctypes binding of libnaer2_hip.so (the C-ABI declared in include/naer2_hip.h).

The library is built in-tree (naer2_amd/libnaer2_hip.so, see naer2_amd/build.py).  There is no CPU
fallback: loading fails loudly when the library is missing, and every call that returns a negative status
raises RuntimeError with the library's own error text.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnaer2_hip.so"

N2H_OK = 0
MODEL_LIF_DELTA = 1
MODEL_LIF_EXP = 2
GROUP = 64
POISSON_K = 32
MAX_POPS = 64
MAX_SLOTS = 128
MAX_STRIDE = 64
DENSE_GROUPS = 1535  # row intake: spike bitmaps of at most this many groups fit in LDS

# every symbol include/naer2_hip.h declares
EXPORTS = (
    "n2h_version", "n2h_last_error", "n2h_plan_create", "n2h_plan_destroy", "n2h_bind_state",
    "n2h_bind_ring", "n2h_set_step", "n2h_get_step", "n2h_step", "n2h_graph_create", "n2h_graph_launch",
    "n2h_graph_destroy", "n2h_status", "n2h_batch_create", "n2h_batch_destroy", "n2h_batch_step", "n2h_batch_sync",
    "n2h_batch_graph_create", "n2h_bind_rows", "n2h_set_vtrace", "n2h_set_trace", "n2h_adj_last_error", "n2h_adj_step", "n2h_adj_window", "n2h_adj_work_words", "n2h_rlif_last_error", "n2h_rlif_forward", "n2h_rlif_readout",
    "n2h_rlif_readout_adjoint", "n2h_rlif_backward", "n2h_rlif_delivery_adjoint", "n2h_rlif_work_floats",
    "n2h_rlif_backward_noalloc", "n2h_bind_log", "n2h_set_status", "n2h_bind_edges", "n2h_set_tuning",
    "n2h_ev_last_error", "n2h_ev_create", "n2h_ev_destroy", "n2h_ev_bind_state", "n2h_ev_bind_record",
    "n2h_ev_set_window", "n2h_ev_get_window", "n2h_ev_run", "n2h_ev_graph_create", "n2h_ev_graph_launch",
    "n2h_ev_graph_destroy", "n2h_ev_status", "n2h_ev_bind_log", "n2h_ev_set_tuning", "n2h_ev_set_trace", "n2h_ev_counters", "n2h_ev_run_timed", "n2h_ev_math",
)

EV_NEURON, EV_POISSON, EV_TRAIN = 0, 1, 2
EV_MAX_SOURCES = 8
EV_MAX_LAG = 64
EV_FLAGS = ("region", "segment", "record", "sort")


class PopParams(ctypes.Structure):
    _fields_ = [
        ("ref_steps", ctypes.c_int32), ("stride", ctypes.c_int32),
        ("decay", ctypes.c_float), ("p_i", ctypes.c_float), ("p_vi", ctypes.c_float),
        ("v_rest", ctypes.c_float), ("v_th", ctypes.c_float), ("v_reset", ctypes.c_float),
        ("j_ext", ctypes.c_float), ("reserved_f", ctypes.c_float),
        ("poisson_thr", ctypes.c_uint32 * POISSON_K),
    ]


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("n_neurons", ctypes.c_int32), ("n_pops", ctypes.c_int32),
        ("d_slots", ctypes.c_int32), ("pop_first", ctypes.POINTER(ctypes.c_int32)),
        ("pops", ctypes.POINTER(PopParams)), ("seed", ctypes.c_uint64), ("instance", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class Tables(ctypes.Structure):
    _fields_ = [
        ("n_groups", ctypes.c_int32), ("uniform_delay", ctypes.c_int32),
        ("group_first", ctypes.c_void_p), ("group_info", ctypes.c_void_p), ("seg_off", ctypes.c_void_p),
        ("links", ctypes.c_void_p), ("n_links", ctypes.c_int64),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("xcd", ctypes.c_int32), ("stage", ctypes.c_int32),
                ("run", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {"order": int(self.order), "xcd": int(self.xcd), "stage": int(self.stage), "run": int(self.run)}


class EvPop(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("region_cap", ctypes.c_int32), ("tau_inv", ctypes.c_float),
                ("theta", ctypes.c_float), ("v_reset", ctypes.c_float), ("t_ref", ctypes.c_float),
                ("rate_inv", ctypes.c_float), ("stream", ctypes.c_int32)]


class EvStream(ctypes.Structure):
    _fields_ = [("n_sources", ctypes.c_int32), ("source_pop", ctypes.c_int32 * EV_MAX_SOURCES),
                ("delay", ctypes.c_float * EV_MAX_SOURCES), ("lag", ctypes.c_int32 * EV_MAX_SOURCES)]


class EvDesc(ctypes.Structure):
    _fields_ = [
        ("n_neurons", ctypes.c_int32), ("n_pops", ctypes.c_int32), ("n_groups", ctypes.c_int32),
        ("n_streams", ctypes.c_int32), ("pop_first", ctypes.POINTER(ctypes.c_int32)),
        ("pops", ctypes.POINTER(EvPop)), ("streams", ctypes.POINTER(EvStream)), ("window", ctypes.c_float),
        ("instance", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("group_first", ctypes.c_void_p),
        ("group_info", ctypes.c_void_p), ("seg_off", ctypes.c_void_p), ("links", ctypes.c_void_p),
        ("n_links", ctypes.c_int64), ("train_ptr", ctypes.c_void_p), ("train_t", ctypes.c_void_p),
    ]


_lib = None


def load(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """The loaded library (once per process).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"naer2_amd: {p} is not built (run __graft_entry__.build() or naer2_amd.build)")
    lib = ctypes.CDLL(str(p))
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sig = {
        "n2h_version": ([], i32),
        "n2h_last_error": ([ctypes.c_char_p, ctypes.c_size_t], i32),
        "n2h_plan_create": ([ctypes.POINTER(ModelDesc), ctypes.POINTER(Tables), ctypes.POINTER(vp)], i32),
        "n2h_plan_destroy": ([vp], i32),
        "n2h_bind_state": ([vp, vp, vp, vp], i32),
        "n2h_bind_ring": ([vp, vp, vp, i64, vp], i32),
        "n2h_set_step": ([vp, i64, vp], i32),
        "n2h_get_step": ([vp, ctypes.POINTER(i64)], i32),
        "n2h_step": ([vp, i32, vp], i32),
        "n2h_graph_create": ([vp, i32, vp, ctypes.POINTER(vp)], i32),
        "n2h_graph_launch": ([vp, vp], i32),
        "n2h_graph_destroy": ([vp], i32),
        "n2h_status": ([vp, ctypes.POINTER(i32), i32], i32),
        "n2h_batch_create": ([ctypes.POINTER(vp), i32, ctypes.POINTER(Tuning), ctypes.POINTER(vp)], i32),
        "n2h_batch_destroy": ([vp], i32),
        "n2h_set_trace": ([vp, vp], i32),
        "n2h_bind_log": ([vp, vp, vp, vp, i64], i32),
        "n2h_set_status": ([vp, vp], i32),
        "n2h_set_tuning": ([vp, ctypes.POINTER(Tuning)], i32),
        "n2h_bind_edges": ([vp, vp, vp, i32], i32),
        "n2h_bind_rows": ([vp, vp, vp, vp, i32, vp, i32], i32),
        "n2h_set_vtrace": ([vp, vp, i64, i32, vp, vp], i32),
        "n2h_batch_step": ([vp, i32, vp], i32),
    "n2h_batch_sync": ([vp, vp], i32),
        "n2h_batch_graph_create": ([vp, i32, vp, ctypes.POINTER(vp)], i32),
        "n2h_ev_last_error": ([ctypes.c_char_p, ctypes.c_size_t], i32),
        "n2h_ev_create": ([ctypes.POINTER(EvDesc), ctypes.POINTER(vp)], i32),
        "n2h_ev_destroy": ([vp], i32),
        "n2h_ev_bind_state": ([vp, vp, vp, vp, vp], i32),
        "n2h_ev_bind_record": ([vp, vp, vp, vp, i64, vp, i32], i32),
        "n2h_ev_set_window": ([vp, i64, i32, vp], i32),
        "n2h_ev_get_window": ([vp, ctypes.POINTER(i64)], i32),
        "n2h_ev_run": ([vp, i32, vp], i32),
        "n2h_ev_graph_create": ([vp, i32, vp, ctypes.POINTER(vp)], i32),
        "n2h_ev_graph_launch": ([vp, vp], i32),
        "n2h_ev_graph_destroy": ([vp], i32),
        "n2h_ev_status": ([vp, ctypes.POINTER(i32), i32], i32),
        "n2h_ev_bind_log": ([vp, vp, vp, vp, vp, i64], i32),
        "n2h_ev_set_tuning": ([vp, ctypes.POINTER(Tuning)], i32),
        "n2h_ev_set_trace": ([vp, vp], i32),
        "n2h_ev_counters": ([vp, ctypes.POINTER(i64), i32], i32),
        "n2h_ev_run_timed": ([vp, i32, vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], i32),
        "n2h_ev_math": ([vp, vp, vp, vp, i64, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load().n2h_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != N2H_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def ev_check(rc: int, what: str) -> None:
    """check() for the n2h_ev_* entry points (their own error text)."""
    if rc != N2H_OK:
        buf = ctypes.create_string_buffer(512)
        load().n2h_ev_last_error(buf, len(buf))
        raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")
