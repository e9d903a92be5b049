"""Committed oracle fixtures (tests/fixtures/*.npz + *.json, written by tests/fixtures/make_fixtures.py):
small seeded networks stored explicitly -- populations, the canonical synapse table, the initial state --
with the oracle's raster and final state after a fixed number of steps.  A change to the oracle, to the host
assembly or to the kernels that keeps every mathematical pin but changes results shows up against them."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import torch

from naer2_amd import LIFDelta, LIFExp
from tests.gpu_util import Given

FIXTURES = Path(__file__).resolve().parent / "fixtures"
NODES = {"LIFDelta": LIFDelta, "LIFExp": LIFExp}


def names() -> list[str]:
    return sorted(p.stem for p in FIXTURES.glob("*.json"))


def load(name: str) -> tuple[dict, dict]:
    meta = json.loads((FIXTURES / f"{name}.json").read_text())
    with np.load(FIXTURES / f"{name}.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    return meta, arrays


def raster(meta: dict, arrays: dict) -> np.ndarray:
    out = np.zeros((meta["steps"], int(arrays["pop_first"][-1])), dtype=bool)
    out[arrays["spike_step"].astype(np.int64), arrays["spike_id"].astype(np.int64)] = True
    return out


def build_network(meta: dict, arrays: dict, device):
    """The fixture's network through the description API: its populations in order, one connection per
    (source population, target population) pair carrying that pair's links in canonical order (connections
    created source-population-major, so the canonical table comes out identical), its initial state."""
    pf = arrays["pop_first"].astype(np.int64)
    pops = []
    net = None
    for spec in meta["pops"]:
        pop = NODES[spec["node"]](**spec["kwargs"]) * int(spec["n"])
        if net is None:
            net = pop.network
        else:
            net.merge(pop.network)
        pops.append(pop)
    net.resolution, net.seed, net.instance = meta["dt"], meta["seed"], meta["instance"]
    net.device = torch.device(device)
    net._generator = None
    src_ptr = arrays["src_ptr"].astype(np.int64)
    n = int(pf[-1])
    src = np.repeat(np.arange(n), np.diff(src_ptr))
    tgt, w, d = arrays["tgt"].astype(np.int64), arrays["w"], arrays["d"]
    pop_of = np.repeat(np.arange(len(pops)), np.diff(pf))
    for si, sp in enumerate(pops):
        for ti, tp in enumerate(pops):
            m = (pop_of[src] == si) & (pop_of[tgt] == ti)
            if not m.any():
                continue
            conn = sp >> Given(src[m] - pf[si], tgt[m] - pf[ti]) >> tp
            conn.weight = torch.as_tensor(w[m])
            conn.delay = torch.as_tensor(d[m].astype(np.float64) * meta["dt"], dtype=torch.float32)
    for p, pop, lo, hi in zip(meta["pops"], pops, pf[:-1], pf[1:]):
        vname = "V" if p["node"] == "LIFDelta" else "potential"
        setattr(pop, vname, torch.as_tensor(arrays["v0"][lo:hi]))
        if p["node"] == "LIFExp":
            pop.current = torch.as_tensor(arrays["i0"][lo:hi])
    return net
