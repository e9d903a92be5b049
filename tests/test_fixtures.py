"""The committed oracle fixtures (tests/fixture_util.py) on the CPU: the oracle reproduces every one bit for
bit from its stored network and initial state, and the description API rebuilds its network exactly (the
canonical synapse table, the population parameter blocks, the initial state)."""
from __future__ import annotations

import numpy as np
import pytest

from naer2_amd.network import assemble
from naer2_amd.params import params_bytes
from oracle import oracle
from tests import fixture_util as fx

NAMES = fx.names()


def test_fixtures_are_present_and_small():
    assert set(NAMES) >= {"brunel_500", "pd14_502"}
    total = sum(p.stat().st_size for p in fx.FIXTURES.glob("*.npz"))
    assert total < (1 << 20)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_the_fixture(name):
    meta, a = fx.load(name)
    n = int(a["pop_first"][-1])
    isyn = a["i0"] if a["i0"].size else None
    st = oracle.State(n, meta["d_slots"], a["v0"], isyn, np.zeros(n, np.int32))
    model = oracle.LIF_DELTA if meta["model"] == "lif_delta" else oracle.LIF_EXP
    r, events = oracle.run(model, a["pop_first"], a["params"].tobytes(), meta["d_slots"], meta["seed"],
                           meta["instance"], a["src_ptr"], a["tgt"], a["w"], a["d"], st, 0, meta["steps"])
    assert events == meta["events"] and int(r.sum()) == meta["spikes"] > 100
    assert np.array_equal(r, fx.raster(meta, a))
    assert np.array_equal(st.v.view(np.uint32), a["v"].view(np.uint32))
    assert np.array_equal(st.ref, a["ref"])
    assert np.array_equal(st.ring.view(np.uint32), a["ring"].view(np.uint32))
    if isyn is not None:
        assert np.array_equal(st.isyn.view(np.uint32), a["isyn"].view(np.uint32))


@pytest.mark.parametrize("name", NAMES)
def test_description_api_rebuilds_the_fixture_network(name):
    meta, a = fx.load(name)
    net = fx.build_network(meta, a, "cpu")
    asm = assemble(net, "cpu")
    c = asm.canon
    assert np.array_equal(asm.pop_first, a["pop_first"])
    assert params_bytes(asm.params) == a["params"].tobytes()
    assert asm.d_slots == meta["d_slots"]
    assert np.array_equal(c.src_ptr.numpy(), a["src_ptr"]) and np.array_equal(c.tgt.numpy(), a["tgt"])
    assert np.array_equal(c.w.numpy().view(np.uint32), a["w"].view(np.uint32))
    assert np.array_equal(c.d.numpy(), a["d"])
    assert np.array_equal(asm.v.numpy().view(np.uint32), a["v0"].view(np.uint32))
    if a["i0"].size:
        assert np.array_equal(asm.isyn.numpy().view(np.uint32), a["i0"].view(np.uint32))
