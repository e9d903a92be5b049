"""Shared test helpers: tiny seeded topologies and single-population parameter blocks."""
from __future__ import annotations

import numpy as np

from naer2_amd.params import lif_delta_params, lif_exp_params, params_bytes


def tiny_topology(n: int, n_links: int, max_delay: int, seed: int, wscale: float = 1.0):
    """A random canonical CSR by source (numpy): src_ptr int64 [n+1], tgt int32, w float32, d int32.
    Weights of varied magnitude and sign, so a changed summation order changes fp32 results."""
    rng = np.random.default_rng(seed)
    src = np.sort(rng.integers(0, n, n_links))
    tgt = rng.integers(0, n, n_links).astype(np.int32)
    w = (rng.standard_normal(n_links) * wscale * np.exp(rng.uniform(-3, 3, n_links))).astype(np.float32)
    d = rng.integers(1, max_delay + 1, n_links).astype(np.int32)
    src_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(src_ptr, src + 1, 1)
    src_ptr = np.cumsum(src_ptr)
    return src_ptr, tgt, w, d


def probe_params(dt: float = 0.1):
    """A delta-LIF block whose potential after a step equals that step's ring input: decay 0, no
    threshold crossing, no background (V = V*0 + x)."""
    p = lif_delta_params(dt, tau_m=20.0, theta=1e30, v_reset=0.0, t_ref=2.0, j_ext=0.0, lam=0.0)
    p.decay = 0.0
    return p


def brunel_params(dt=0.1, lam=1.0, j=0.2):
    return lif_delta_params(dt, 20.0, 20.0, 10.0, 2.0, j, lam)


def pd14_params(dt=0.1, lam=1.28, v_thresh=-50.0):
    return lif_exp_params(dt, 10.0, 0.5, 40.0, 2.0, -65.0, v_thresh, -65.0, 0.08781, lam)


__all__ = ["tiny_topology", "probe_params", "brunel_params", "pd14_params", "params_bytes"]
