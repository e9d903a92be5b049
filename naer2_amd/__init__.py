"""naer2_amd: an MI355X-native execution backend for naer2's spiking networks.

naer2's description API (node, pattern, edge, Population, `>>`, `| delay`, Network) on top of a C-ABI
(include/naer2_hip.h) that launches hand-written gfx950 HIP kernels: one fused kernel per network step
covering pull intake, delivery into the delayed-input ring, and the LIF update.
"""
from .dsl import Connection, ConnectionBuilder, Points, Population, Region, edge, node, pattern
from .fusion import FusedRound
from .network import Network, Runtime
from .nodes import FixedInDegree, FixedTotalNumber, LIFDelta, LIFExp

__all__ = [
    "Connection", "ConnectionBuilder", "Points", "Population", "Region", "edge", "node", "pattern",
    "FusedRound", "Network", "Runtime", "FixedInDegree", "FixedTotalNumber", "LIFDelta", "LIFExp",
]
