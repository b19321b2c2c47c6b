"""The tile autotuner (rpcf_hip.hip, autotune_tiles): the tile of a transform kernel only
regroups independent pencils, rows or columns, so the rule-based tiles (RPCF_AUTOTUNE=0),
the halved candidates (RPCF_AUTOTUNE=half) and the timed choice give bit-identical states,
in FP64 and in the mixed form."""
import os

import numpy as np
import pytest

import rpcf
from helpers import IC, params

pytestmark = pytest.mark.gpu


def run(prm, mode, nsteps=3):
    old = os.environ.get("RPCF_AUTOTUNE")
    os.environ["RPCF_AUTOTUNE"] = mode
    try:
        sim = rpcf.Simulation(rpcf.HIP_LIB, prm)
    finally:
        if old is None:
            del os.environ["RPCF_AUTOTUNE"]
        else:
            os.environ["RPCF_AUTOTUNE"] = old
    with sim:
        v0, e0 = rpcf.seed_state(sim.dll, prm, 0, *IC["C0"])
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(nsteps)
        return sim.get_state()


@pytest.mark.parametrize("precision", [0, 1])
def test_tiles_do_not_change_results(precision):
    prm = params("C0")
    prm.precision = precision
    ref = run(prm, "0")
    for mode in ("half", "1"):
        got = run(prm, mode)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), mode
