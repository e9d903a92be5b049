"""The fusion plan: one launch for the runs of several networks (naer2/fusion.py:31-141, FusedRound).

The populations of one network already share one launch per step (their target groups are programs of one
grid).  FusedRound puts the target groups of up to 256 independent built networks -- a batched sweep --
into one launch per step (n2h_batch_*), with the launch mapping (plan order, XCD-aware block -> group
mapping) chosen by naer2_amd/autotune.py.  Windows of steps are captured into hipGraphs and replayed.

    FusedRound(runtimes, tuning, weights=None)   the batch over these runtimes (in this order unless tuning
                                    reorders: order 1 most links first, order 2 heaviest weight first)
    .covers(runtimes)               whether a round's runs are exactly these runtimes'
    .launch(n_steps, graph_window)  n_steps steps of every runtime (torch's current stream)

Results do not depend on the batch or its mapping: every runtime computes exactly what it computes alone.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import Tuning


class FusedRound:
    def __init__(self, runtimes: list, tuning: Tuning | dict | None = None, reserve_steps: int = 0,
                 weights: list | None = None):
        if not runtimes:
            raise ValueError("FusedRound: no runtimes")
        if len(runtimes) > 256:
            raise ValueError("FusedRound: at most 256 runtimes in one launch")
        if len({id(r) for r in runtimes}) != len(runtimes):
            raise ValueError("FusedRound: a runtime is listed twice")
        steps = {r.step for r in runtimes}
        if len(steps) != 1:
            raise ValueError(f"FusedRound: the runtimes are at different steps ({sorted(steps)})")
        if isinstance(tuning, dict):
            tuning = Tuning(order=int(tuning.get("order", 0)), xcd=int(tuning.get("xcd", 0)),
                            stage=int(tuning.get("stage", 0)), run=int(tuning.get("run", 0)))
        self.tuning = tuning if tuning is not None else Tuning(order=0, xcd=0, stage=0)
        self.runtimes = list(runtimes)
        launch_tuning = self.tuning
        if self.tuning.order == 2:  # heaviest groups first, by a measured weight per runtime
            if weights is None or len(weights) != len(self.runtimes):
                raise ValueError("FusedRound: order 2 needs one weight per runtime (autotune.measure_work)")
            perm = sorted(range(len(self.runtimes)), key=lambda i: -float(weights[i]))
            self.runtimes = [self.runtimes[i] for i in perm]
            launch_tuning = Tuning(order=0, xcd=self.tuning.xcd, stage=self.tuning.stage, run=self.tuning.run)
        for rt in self.runtimes:  # the bit histories must not move while the batch holds their addresses
            rt._ensure_slots(reserve_steps)
            rt._batch_capacity = rt.step + reserve_steps if rt.net.record else None
        self.n_programs = sum(rt.n_groups for rt in self.runtimes)
        self.launches = 0
        self.device = self.runtimes[0].device
        lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.runtimes))(*[rt._plan.value for rt in self.runtimes])
        out = ctypes.c_void_p()
        _lib.check(lib.n2h_batch_create(arr, len(self.runtimes), ctypes.byref(launch_tuning), ctypes.byref(out)),
                   "n2h_batch_create")
        self._batch = out
        self._graphs: dict = {}

    def covers(self, runtimes) -> bool:
        return len(runtimes) == len(self.runtimes) and {id(r) for r in runtimes} == {id(r) for r in self.runtimes}

    @property
    def step(self) -> int:
        return self.runtimes[0].step

    def prepare(self, n_steps: int, graph_window: int) -> int:
        """Capture (without running) the graphs launch(n_steps, graph_window) will replay from the current
        step, so that a timed launch captures nothing; returns the graphs captured."""
        from .network import Runtime

        lib = _lib.load()
        sv = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        made = 0
        for kind, k, phase in Runtime.schedule(self.step, n_steps, graph_window):
            if kind != "graph" or (k, phase) in self._graphs:
                continue
            if k % 4 or phase != self.step % 4:
                continue  # launch() replays whole windows from phase 0 only; captured when reached
            gp = ctypes.c_void_p()
            _lib.check(lib.n2h_batch_graph_create(self._batch, k, sv, ctypes.byref(gp)), "n2h_batch_graph_create")
            self._graphs[(k, phase)] = gp
            made += 1
        return made

    def launch(self, n_steps: int = 1, graph_window: int = 0) -> None:
        """n_steps steps of every covered runtime in batched launches (graph windows of graph_window steps,
        a multiple of 4, where they fit; eager launches for the rest)."""
        if n_steps < 0:
            raise ValueError("n_steps must be >= 0")
        for rt in self.runtimes:
            cap = getattr(rt, "_batch_capacity", None)
            if cap is not None and rt.step + n_steps > cap:
                raise RuntimeError("FusedRound: more steps than reserve_steps recorded; rebuild with a larger "
                                   "reserve (the bit histories cannot move under a batch)")
        if graph_window and graph_window % 4:
            raise ValueError("graph windows must be a multiple of 4 steps")
        lib = _lib.load()
        sv = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        done = 0
        if graph_window and self.step % 4 and n_steps >= graph_window + 4:
            lead = 4 - self.step % 4
            _lib.check(lib.n2h_batch_step(self._batch, lead, sv), "n2h_batch_step")
            self._advance(lead)
            done = lead
        if graph_window and n_steps - done >= graph_window:
            key = (graph_window, self.step % 4)
            g = self._graphs.get(key)
            if g is None:
                gp = ctypes.c_void_p()
                _lib.check(lib.n2h_batch_graph_create(self._batch, graph_window, sv, ctypes.byref(gp)),
                           "n2h_batch_graph_create")
                g = self._graphs[key] = gp
            while n_steps - done >= graph_window:
                _lib.check(lib.n2h_graph_launch(g, sv), "n2h_graph_launch")
                self._advance(graph_window)
                done += graph_window
        if n_steps > done:
            _lib.check(lib.n2h_batch_step(self._batch, n_steps - done, sv), "n2h_batch_step")
            self._advance(n_steps - done)
        self.launches += 1

    def sync(self) -> None:
        """Follow the runtimes again after they were moved (restore, set_step) or rebound (trace_potentials):
        the batch's step and every plan's launch arguments, refreshed on the current stream (n2h_batch_sync)."""
        steps = {rt.step for rt in self.runtimes}
        if len(steps) != 1:
            raise ValueError(f"FusedRound: the runtimes are at different steps ({sorted(steps)})")
        lib = _lib.load()
        sv = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(lib.n2h_batch_sync(self._batch, sv), "n2h_batch_sync")

    def _advance(self, n: int) -> None:
        for rt in self.runtimes:
            rt.step += n
            if not rt.net.record:
                rt.record_from = max(rt.step - 1, 0)

    def close(self) -> None:
        lib = _lib.load()
        for g in self._graphs.values():
            lib.n2h_graph_destroy(g)
        self._graphs = {}
        if self._batch is not None:
            lib.n2h_batch_destroy(self._batch)
            self._batch = None
        for rt in self.runtimes:
            rt._batch_capacity = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
