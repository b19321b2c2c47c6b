"""The mixed-precision form (precision = 1): the transform intermediates, the x / span
FFTs and the Lamb vector in FP32; the wall-normal Chebyshev transforms as split-FP16 products
on the matrix cores with FP32 accumulation (the padded s2p of the six fields; the padded p2s,
derivative, nonlinear combination and unpadded s2p folded into three products); velocity
recovery, the operators, the solves and the mean flow in FP64.  Bars: after 100 steps (200
for the rotating Couette case) every field stays within 1e-6 relative (max |a-b| / max |b|)
of the FP64 run from the same initial field, and differs from it (the FP32 path ran); the
nonlinear term of one evaluation within 2e-6 (N_eta) and 2e-4 (N_v) of the CPU oracle: FP32
round-off of the Lamb products, which the wall-normal derivative in N_v amplifies up to ~N^2
for the high Chebyshev modes (the FFT-based FP32 form measures 6.5e-5 there at
128x129x128; those modes are the ones the implicit fourth-order solve damps most)."""
import ctypes
import os

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


NL_CASES = {
    # 2N = 64 (N = 32), Ro != 0, Couette
    "TINY": (params("TINY"), IC["TINY"]),
    # 2N = 256 (N = 128), channel, flow control
    "C1": (params("C1"), IC["C1"]),
    # 2N = 96 (N = 48): a length whose halves are not powers of two
    "N48": (rpcf.make_params(mx=48, my=48, mz=192, n_pi_x=2.0, n_pi_y=1.0, Re=1500.0, Ro=0.3, dt=0.002,
                             u_profile=(1.0, 0.0, -1.0), flow_control=1), (0.05, 6, 10, 0.0, True)),
    # 2N = 24 (N = 12): the shortest compiled length, tiles of 7 rows
    "N12": (rpcf.make_params(mx=24, my=24, mz=48, n_pi_x=2.0, n_pi_y=1.0, Re=800.0, Ro=0.1, dt=0.002,
                             u_profile=(0.0, 1.0, 0.0)), (0.05, 4, 6, 0.0, True)),
    # 2N = 384 (N = 192): the longest compiled Chebyshev length (seven row tiles per block)
    "DEEP": (CASES["DEEP"][0], (0.05, 8, 8, 0.0, True)),
}


def nonlinear_of(lib, prm, ic, mfma=True):
    old = os.environ.get("RPCF_MIXED_MFMA")
    os.environ["RPCF_MIXED_MFMA"] = "1" if mfma else "0"
    try:
        sim = rpcf.Simulation(lib, prm)
    finally:
        if old is None:
            del os.environ["RPCF_MIXED_MFMA"]
        else:
            os.environ["RPCF_MIXED_MFMA"] = old
    with sim:
        sim.set_state(*rpcf.seed_state(sim.dll, prm, 3, *ic))
        sim.nonlinear()
        return sim.get_nonlinear()


@pytest.mark.parametrize("name", list(NL_CASES))
def test_mixed_nonlinear_term_against_oracle(name):
    """one nonlinear evaluation (six backward transforms, Lamb + Coriolis, three forward
    transforms, rhsNonlinear, unpadded s2p) of the mixed form on the matrix cores against the
    FP64 oracle, and against the mixed form with the FFT-based Chebyshev kernels"""
    prm, ic = NL_CASES[name]
    prm = fresh(prm)
    ov, oe = nonlinear_of(rpcf.ORACLE_LIB, prm, ic)
    prm.precision = 1
    mv, me = nonlinear_of(rpcf.HIP_LIB, prm, ic, mfma=True)
    fv, fe = nonlinear_of(rpcf.HIP_LIB, prm, ic, mfma=False)
    errs = {"N_v": rel(mv, ov), "N_eta": rel(me, oe), "N_v (FFT form)": rel(fv, ov), "N_eta (FFT form)": rel(fe, oe),
            "N_v MFMA vs FFT form": rel(mv, fv), "N_eta MFMA vs FFT form": rel(me, fe)}
    print(name, {k: "%.2e" % v for k, v in errs.items()})
    assert errs["N_v"] < 2e-4 and errs["N_eta"] < 2e-6
    assert errs["N_v MFMA vs FFT form"] < 2e-4 and errs["N_eta MFMA vs FFT form"] < 2e-6
    assert not np.array_equal(mv, fv)  # the two Chebyshev paths both ran


def test_group_precision_must_match():
    prm = params("TINY")
    with rpcf.Group(rpcf.HIP_LIB, prm, 1) as g:
        q = params("TINY")
        q.precision = 1
        with pytest.raises(rpcf.RpcfError):
            g.init(0, q)


@pytest.mark.parametrize("name", ["C1", "DEEP"])
def test_fp64_solve_remainders_restore_full_inverse(name, monkeypatch):
    """precision 0 with an even base profile solves with the parity blocks and their BF16
    remainders (the v blocks only; 9/16 of the full inverses' bytes in all), which restore
    the full inverses' results:
    within 2e-13 of the full-inverse state (RPCF_FULL_INVERSE=1) after one step and 1e-11
    after 20, and at least 50 times closer to it than the blocks alone (precision 2) after
    one step: the BF16 rounding (2^-9 relative) of the remainders leaves ~2^-9 of the gap
    the blocks alone open (1.4e-11 at 128x129x128), 7e-14 measured.  (The numpy
    reconstruction of one solve on the oracle's 128x129x128 inverses measured 3e-16
    relative with exact remainders, <= 2.5e-15 with BF16 ones and up to 7.6e-13 without
    them.)"""
    prm, ic, _ = CASES[name]
    a, _ = run(fresh(prm), ic, 1)
    a20, _ = run(fresh(prm), ic, 20)
    monkeypatch.setenv("RPCF_FULL_INVERSE", "1")
    f, _ = run(fresh(prm), ic, 1)
    f20, _ = run(fresh(prm), ic, 20)
    monkeypatch.delenv("RPCF_FULL_INVERSE")
    q = fresh(prm)
    q.precision = 2
    b, _ = run(q, ic, 1)
    e1 = max(rel(a[i], f[i]) for i in (0, 1))
    e20 = max(rel(a20[i], f20[i]) for i in (0, 1))
    eb = max(rel(b[i], f[i]) for i in (0, 1))
    print("%s: remainders vs full inverse %.2e (1 step) %.2e (20 steps); blocks alone %.2e" % (name, e1, e20, eb))
    assert e1 < 2e-13 and e20 < 1e-11
    assert eb > 50 * e1


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
