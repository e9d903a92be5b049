"""The fusion plan: one launch for the runs of several populations (naer2/fusion.py:31-141, FusedRound).

In this backend the populations of one network already share one launch per step (their target groups
are programs of one grid).  FusedRound keeps the reference's interface -- FusedRound(runtimes),
.covers(runtimes), .launch(...) -- over built network runtimes.
"""
from __future__ import annotations


class FusedRound:
    """One launch per step over the given runtimes' target groups, in this order."""

    def __init__(self, runtimes: list):
        if not runtimes:
            raise ValueError("FusedRound: no runtimes")
        if len(runtimes) != 1:
            raise NotImplementedError("FusedRound over several networks needs the batched step (not built yet)")
        self.runtimes = list(runtimes)
        self.n_programs = sum(rt.n_groups for rt in self.runtimes)
        self.launches = 0

    def covers(self, runtimes) -> bool:
        return len(runtimes) == len(self.runtimes) and {id(r) for r in runtimes} == {id(r) for r in self.runtimes}

    def launch(self, n_steps: int = 1, graph_window: int = 0) -> None:
        """n_steps steps of every covered runtime (stream: torch's current stream)."""
        self.launches += 1
        for rt in self.runtimes:
            rt.run(n_steps, graph_window)
