"""The mixed-precision form (precision = 1): the transform intermediates and the x / span
FFTs and the Lamb vector in FP32, the Chebyshev transforms, the nonlinear combination,
the operators and the solves in FP64.  Bar: after 100 steps (200 for the rotating Couette
case) every field stays within 1e-6 relative (max |a-b| / max |b|) of the FP64 run from
the same initial field, and differs from it (the FP32 path ran)."""
import ctypes

import numpy as np
import pytest

import rpcf
from helpers import IC, params, rel

pytestmark = pytest.mark.gpu

CASES = {
    # name: (params, initial field, steps)
    "C0": (params("C0"), IC["C0"], 100),
    "C1": (params("C1"), IC["C1"], 100),
    "C2": (rpcf.make_params(mx=384, my=384, mz=384, n_pi_x=4.0, n_pi_y=2.0, Re=1300.0, Ro=0.1, dt=0.002,
                            u_profile=(0.0, 1.0, 0.0)), (0.05, 16, 24, 5.0 / 6.0), 200),
    # the longest compiled Chebyshev length (nz = 193) with an even profile: 97 parity pairs
    "DEEP": (rpcf.make_params(mx=24, my=24, mz=768, n_pi_x=2.0, n_pi_y=1.0, Re=2000.0, Ro=0.0, dt=0.002,
                              u_profile=(1.0, 0.0, -1.0), flow_control=1), (0.05, 8, 8, 0.0), 50),
}


def run(prm, ic, steps):
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        sim.set_state(*rpcf.seed_state(sim.dll, prm, 0, *ic))
        sim.bootstrap()
        sim.step(steps - 1)
        state = sim.get_state()
        fields = [sim.get_field(f) for f in range(6)]
    return state, fields


def fresh(prm):
    """a copy of a parameter set (the cases are shared between tests)"""
    q = type(prm)()
    ctypes.pointer(q)[0] = prm
    return q


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_within_1e6_of_fp64(name):
    prm, ic, steps = CASES[name]
    prm = fresh(prm)
    (v64, e64), f64 = run(prm, ic, steps)
    prm.precision = 1
    (v32, e32), f32 = run(prm, ic, steps)
    errs = [rel(v32, v64), rel(e32, e64)] + [rel(a, b) for a, b in zip(f32, f64) if np.abs(b).max() > 0]
    print(name, "max relative error per field (v, eta, u, v, w, ox, oy, oz):", ["%.2e" % x for x in errs])
    assert max(errs) <= 1e-6
    assert max(errs) > 1e-13


def test_group_precision_must_match():
    prm = params("TINY")
    with rpcf.Group(rpcf.HIP_LIB, prm, 1) as g:
        q = params("TINY")
        q.precision = 1
        with pytest.raises(rpcf.RpcfError):
            g.init(0, q)


def test_fp64_parity_solves():
    """precision 2: FP64 with the reflection-parity blocks of the inverses for an even base
    profile -- within 1e-10 of the full-inverse FP64 state after one step and 1e-8 after 20
    (the blocks drop the inverse's own rounding asymmetry); a Couette (odd) profile runs the
    full inverses, bit for bit as precision 0"""
    prm, ic, _ = CASES["C1"]
    prm = fresh(prm)
    a, _ = run(prm, ic, 1)
    a20, _ = run(prm, ic, 20)
    prm.precision = 2
    b, _ = run(prm, ic, 1)
    b20, _ = run(prm, ic, 20)
    e1, e20 = max(rel(b[0], a[0]), rel(b[1], a[1])), max(rel(b20[0], a20[0]), rel(b20[1], a20[1]))
    print("precision 2 vs 0: 1 step %.2e, 20 steps %.2e" % (e1, e20))
    assert 0 < e1 < 1e-10 and e20 < 1e-8
    prm0, ic0, _ = CASES["C0"]  # plane Couette
    prm0 = fresh(prm0)
    c, _ = run(prm0, ic0, 3)
    prm0.precision = 2
    d, _ = run(prm0, ic0, 3)
    assert np.array_equal(c[0], d[0]) and np.array_equal(c[1], d[1])
