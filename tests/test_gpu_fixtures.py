"""The committed oracle fixtures (tests/fixture_util.py) on the GPU, through the C-ABI: each stored network
rebuilt through the description API runs its stored number of steps and gives the stored raster, potentials,
currents, refractory counters and ring bit for bit."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from naer2_amd.params import params_bytes
from tests import fixture_util as fx
from tests.gpu_util import bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", fx.names())
@pytest.mark.parametrize("graph_window", [0, 100])
def test_gpu_reproduces_the_fixture(name, graph_window):
    print(f"case: fixture {name}, graph window {graph_window}", flush=True)
    sys.stdout.flush()
    meta, a = fx.load(name)
    net = fx.build_network(meta, a, "cuda")
    rt = net.build()
    assert params_bytes(rt.params) == a["params"].tobytes()
    net.run(meta["steps"], graph_window=graph_window)
    torch.cuda.synchronize()
    assert np.array_equal(rt.raster(), fx.raster(meta, a))
    assert bits_equal(rt.v.cpu().numpy(), a["v"])
    assert np.array_equal(rt.ref.cpu().numpy(), a["ref"])
    assert bits_equal(rt.ring.cpu().numpy(), a["ring"])
    if a["isyn"].size:
        assert bits_equal(rt.isyn.cpu().numpy(), a["isyn"])
