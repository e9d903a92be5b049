"""naer2_amd: an MI355X-native execution backend for naer2's spiking networks.

naer2's description API (node, pattern, edge, Population, `>>`, `| delay`, Network) on top of a C-ABI
(include/naer2_hip.h) that launches hand-written gfx950 HIP kernels: one fused kernel per network step
covering pull intake, delivery into the delayed-input ring, and the LIF update; and, for event-driven
networks with exact spike times (Brunel's model A as the reference runs it), two kernels per window of the
common delay (naer2_amd/event.py).
"""
from .dsl import Connection, ConnectionBuilder, Points, Population, Region, edge, node, pattern
from .fusion import FusedRound
from .network import Network, Runtime
from .event import EventRuntime
from .nodes import (FixedInDegree, FixedTotalNumber, GainSynapse, LIFDelta, LIFDeltaExact, LIFExp, PlasticSynapse,
                    PoissonExact, SpikeTrains, StaticSynapse)

__all__ = [
    "Connection", "ConnectionBuilder", "Points", "Population", "Region", "edge", "node", "pattern",
    "FusedRound", "Network", "Runtime", "EventRuntime", "FixedInDegree", "FixedTotalNumber", "LIFDelta", "LIFExp",
    "LIFDeltaExact", "PoissonExact", "SpikeTrains", "StaticSynapse", "GainSynapse", "PlasticSynapse",
]
