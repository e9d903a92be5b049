"""Writes the oracle fixtures of tests/fixture_util.py: small seeded networks built on the host (torch's CPU
generator), stored explicitly, and the CPU oracle's results after `steps` steps from their initial state.

    python tests/fixtures/make_fixtures.py

Each fixture is <name>.json (populations, seed, sizes, totals) and <name>.npz (the canonical synapse table,
the population parameter blocks, the initial state, the raster as (step, neuron) pairs, the final potentials,
currents, refractory counters and ring), under 1 MB together."""
from __future__ import annotations

import inspect
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from naer2_amd.network import assemble  # noqa: E402
from naer2_amd.networks import brunel, pd14  # noqa: E402
from naer2_amd.params import params_bytes  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = {
    # config 0's model at a fifth of its size: delta-synapse LIF, one delay (15 steps)
    "brunel_500": (lambda: brunel(N_E=400, N_I=100, C_E=40, C_I=10, seed=11, device="cpu"), 1000),
    # PD14's model: exponential synapses, 8 populations, per-link delays of 1-~40 steps
    "pd14_502": (lambda: pd14(scale=0.0065, seed=12, device="cpu"), 1000),
}


def node_spec(pop) -> dict:
    inst = pop.instance
    keys = [k for k in inspect.signature(type(inst).__init__).parameters if k != "self"]
    return {"node": type(inst).__name__, "n": int(pop.n), "kwargs": {k: getattr(inst, k) for k in keys}}


def write(name: str, make, steps: int) -> dict:
    net, _ = make()
    a = assemble(net, "cpu")
    c = a.canon
    v0 = a.v.numpy().copy()
    i0 = a.isyn.numpy().copy() if a.isyn is not None else np.zeros(0, np.float32)
    st = oracle.State(a.n, a.d_slots, v0, a.isyn.numpy() if a.isyn is not None else None, a.ref.numpy())
    model = oracle.LIF_DELTA if a.family == "lif_delta" else oracle.LIF_EXP
    pb = params_bytes(a.params)
    r, events = oracle.run(model, a.pop_first, pb, a.d_slots, net.seed, net.instance, c.src_ptr.numpy(),
                           c.tgt.numpy(), c.w.numpy(), c.d.numpy(), st, 0, steps)
    s_step, s_id = np.nonzero(r)
    assert a.n < 32768 and steps < 32768
    np.savez_compressed(
        HERE / f"{name}.npz", pop_first=a.pop_first.astype(np.int32), params=np.frombuffer(pb, dtype=np.uint8),
        src_ptr=c.src_ptr.numpy(), tgt=c.tgt.numpy(), w=c.w.numpy(), d=c.d.numpy(), v0=v0, i0=i0,
        spike_step=s_step.astype(np.int16), spike_id=s_id.astype(np.int16), v=st.v,
        isyn=st.isyn if st.isyn is not None else np.zeros(0, np.float32), ref=st.ref, ring=st.ring)
    meta = {"name": name, "model": a.family, "seed": int(net.seed), "instance": int(net.instance),
            "dt": float(a.dt), "steps": steps, "d_slots": int(a.d_slots), "n_neurons": int(a.n),
            "n_links": int(c.n_links), "spikes": int(len(s_id)), "events": int(events),
            "pops": [node_spec(p) for p in a.pops]}
    (HERE / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
    return meta


if __name__ == "__main__":
    for name, (make, steps) in CASES.items():
        m = write(name, make, steps)
        print(f"{name}: {m['n_neurons']} neurons, {m['n_links']} links, {m['spikes']} spikes in {steps} steps")
