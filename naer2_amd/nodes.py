"""The node types and connection patterns this backend runs, with the reference models' names and
parameters.

Node families (one hand-written kernel each, naer2_amd/csrc/naer2_hip.hip):
  LIFDelta  "lif_delta"  current-based delta-synapse LIF, the time-driven Brunel model A of
                         tools/brunel_reference.py:14-48 (Brunel_neuron's parameters, test/networks/Brunel/
                         neuron.py:23-30), with a Poisson drive of rate_ext events/ms of weight J_ext.
  LIFExp    "lif_exp"    LIF with exponential current synapses and the exact propagator
                         (test/networks/PD14/neuron.py:6-60), with a Poisson background of K_ext inputs at
                         nu_ext of weight J_ext instead of the reference's DC shift of V_rest (neuron.py:28-29).
  LIFDeltaExact  "event_lif"      Brunel_neuron (test/networks/Brunel/neuron.py:6-45): the event-driven delta LIF
                                  with exact input and spike times (naer2_amd/event.py, csrc/event.hip)
  PoissonExact   "event_poisson"  Brunel_poisson (test/networks/Brunel/poisson.py:7-27): exponential intervals,
                                  exact spike times
  SpikeTrains    "event_train"    listed spike times per source (the commanded sources of the reference's
                                  tests, test/nodes/commanded.py)
Edges (one edge forward each, run by the step kernel's edge variant; exponential-synapse networks):
  StaticSynapse(weight)           "static"   delivers weight (test/edges/static_synapse.py:5-21)
  GainSynapse(weight, gain)       "gain"     delivers weight * gain: two per-link attributes (TODO.md:189-192)
  PlasticSynapse(weight, factor)  "plastic"  weight += factor, stored, then delivers the new weight
                                             (test/edges/plastic_synapse.py:5-14)
Patterns:
  FixedInDegree(K)          every target draws K sources uniformly with replacement
                            (test/networks/Brunel/connection.py:7-33, tools/brunel_reference.py:20-21)
  FixedTotalNumber(p)       every source draws round(Q / N_x) targets uniformly with replacement,
                            Q = ln(1 - p) / ln(1 - 1/(N_x N_y)) (test/networks/PD14/connection.py:9-31)
"""
from __future__ import annotations

import math

import torch

from .dsl import edge, node, pattern
from .params import lif_delta_params, lif_exp_params


@node
class LIFDelta:
    family = "lif_delta"
    state = ("V", "ref")

    def __init__(self, tau: float = 20.0, theta: float = 20.0, V_r: float = 10.0, t_ref: float = 2.0,
                 J_ext: float = 0.0, rate_ext: float = 0.0):
        self.tau, self.theta, self.V_r, self.t_ref = tau, theta, V_r, t_ref
        self.J_ext = J_ext          # mV per background event
        self.rate_ext = rate_ext    # background events per ms (all external inputs together)
        self.V = 0.0
        self.ref = 0

    def params(self, dt: float):
        return lif_delta_params(dt, self.tau, self.theta, self.V_r, self.t_ref, self.J_ext, self.rate_ext * dt)


@node
class LIFExp:
    family = "lif_exp"
    state = ("potential", "current", "ref")

    def __init__(self, K_ext: int = 0, t_mem: float = 10.0, t_syn: float = 0.5, t_ref: float = 2.0,
                 R_mem: float = 40.0, E_leak: float = -65.0, V_thresh: float = -50.0, V_reset: float = -65.0,
                 nu_ext: float = 8e-3, J_ext: float = 0.08781):
        self.K_ext = K_ext
        self.t_mem, self.t_syn, self.t_ref, self.R_mem = t_mem, t_syn, t_ref, R_mem
        self.E_leak, self.V_thresh, self.V_reset = E_leak, V_thresh, V_reset
        self.nu_ext = nu_ext        # spikes/ms per background input
        self.J_ext = J_ext          # nA per background spike
        self.potential = V_reset
        self.current = 0.0
        self.ref = 0

    def params(self, dt: float):
        return lif_exp_params(dt, self.t_mem, self.t_syn, self.R_mem, self.t_ref, self.E_leak, self.V_thresh,
                              self.V_reset, self.J_ext, self.K_ext * self.nu_ext * dt)


@node
class LIFDeltaExact:
    """Event-driven delta-synapse LIF with exact times: receive decays V to the input's arrival time and adds
    the weight unless refractory; after the inputs at one time, V >= theta fires at that time."""
    family = "event_lif"
    state = ("V", "last_time", "t_ref_end")

    def __init__(self, tau: float = 20.0, theta: float = 20.0, V_r: float = 10.0, t_ref: float = 2.0):
        self.tau, self.theta, self.V_r, self.t_ref = tau, theta, V_r, t_ref
        self.V = 0.0
        self.last_time = 0.0
        self.t_ref_end = 0.0


@node
class PoissonExact:
    """A Poisson source with exact spike times: rate in spikes per ms."""
    family = "event_poisson"
    state = ("next_spike",)

    def __init__(self, rate: float):
        if rate < 0:
            raise ValueError("rate must be >= 0")
        self.rate = rate
        self.next_spike = -1.0


@node
class SpikeTrains:
    """Sources firing at listed times: times[i] the (ascending, >= 0) spike times of source i, ms."""
    family = "event_train"
    state = ()

    def __init__(self, times):
        self.times = [sorted(float(t) for t in ts) for ts in times]
        if any(t < 0 for ts in self.times for t in ts):
            raise ValueError("spike times must be >= 0")


@pattern
class FixedInDegree:
    def __init__(self, K: int):
        if K < 0:
            raise ValueError("K must be >= 0")
        self.K = int(K)

    def links(self, n_src: int, n_tgt: int, generator: torch.Generator):
        dev = generator.device
        if n_src == 0 or self.K == 0:
            z = torch.zeros(0, dtype=torch.int64, device=dev)
            return z, z
        src = torch.randint(0, n_src, (n_tgt * self.K,), generator=generator, device=dev)
        tgt = torch.arange(n_tgt, device=dev).repeat_interleave(self.K)
        return src, tgt

    def __call__(self, position, source_region, target_region=None):
        raise NotImplementedError("FixedInDegree draws per target: use links()")


@pattern
class FixedTotalNumber:
    def __init__(self, probability: float):
        if not 0.0 <= probability < 1.0:
            raise ValueError("probability must be in [0, 1)")
        self.probability = probability

    def per_source(self, n_src: int, n_tgt: int) -> int:
        if self.probability == 0.0 or n_src == 0 or n_tgt == 0:
            return 0
        q = math.log(1 - self.probability) / math.log(1 - 1 / (n_src * n_tgt))
        return round(q / n_src)

    def links(self, n_src: int, n_tgt: int, generator: torch.Generator):
        dev = generator.device
        k = self.per_source(n_src, n_tgt)
        src = torch.arange(n_src, device=dev).repeat_interleave(k)
        tgt = torch.randint(0, max(n_tgt, 1), (n_src * k,), generator=generator, device=dev)
        return src, tgt

    def __call__(self, position, source_region, target_region=None):
        from .dsl import Points
        n_src = len(source_region)
        n_tgt = len(target_region if target_region is not None else source_region)
        k = self.per_source(n_src, n_tgt)
        return Points(torch.randint(0, max(n_tgt, 1), (k, 1)))


@edge(kind="static")
class StaticSynapse:
    """A fixed weight (test/edges/static_synapse.py:5-21)."""

    def __init__(self, weight: float):
        self.weight = weight


@edge(kind="gain")
class GainSynapse:
    """Delivers weight * gain, both per-link attributes (the reference's two-attribute receive,
    test/test_fused_generated_gpu.py:373-381)."""

    def __init__(self, weight: float, gain: float = 1.0):
        self.weight = weight
        self.gain = gain


@edge(kind="plastic")
class PlasticSynapse:
    """weight += factor at every delivery, before the weight is delivered; the update persists
    (test/edges/plastic_synapse.py:5-14)."""

    def __init__(self, weight: float, factor: float):
        self.weight = weight
        self.factor = factor
