"""Golden fixtures (tests/golden/tiny_couette.npz, written by tests/golden/make_golden.py
from the oracle): the oracle reproduces them bit for bit; the HIP library matches them to
the parity bars (N to 1e-12, the state to 1e-12 after one step and 1e-10 after ten)."""
import os
import sys

import numpy as np
import pytest

import rpcf
from helpers import rel

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
from make_golden import IC, PARAMS  # noqa: E402


def run(lib):
    g = np.load(os.path.join(HERE, "tiny_couette.npz"))
    prm = rpcf.make_params(**PARAMS)
    with rpcf.Simulation(lib, prm) as sim:
        v0, e0 = rpcf.seed_state(sim.dll, prm, *IC)
        sim.set_state(v0, e0)
        sim.bootstrap()
        s1 = sim.get_state()
        h = sim.get_nonlinear(prev=True)
        sim.step(9)
        s10 = sim.get_state()
    return g, (v0, e0), s1, h, s10


def test_oracle_reproduces_golden():
    g, s0, s1, h, s10 = run(rpcf.ORACLE_LIB)
    for name, a in zip(["v0", "e0", "v1", "e1", "nv_hist", "ne_hist", "v10", "e10"], [*s0, *s1, *h, *s10]):
        assert np.array_equal(a, g[name]), name


@pytest.mark.gpu
def test_gpu_matches_golden():
    g, s0, s1, h, s10 = run(rpcf.HIP_LIB)
    assert np.array_equal(s0[0], g["v0"]) and np.array_equal(s0[1], g["e0"])
    assert rel(h[0], g["nv_hist"]) < 1e-12 and rel(h[1], g["ne_hist"]) < 1e-12
    assert rel(s1[0], g["v1"]) < 1e-12 and rel(s1[1], g["e1"]) < 1e-12
    assert rel(s10[0], g["v10"]) < 1e-10 and rel(s10[1], g["e10"]) < 1e-10
