"""Parity of the HIP library (librpcf_hip.so, MI355X) with the CPU oracle on the same
seeded inputs, through the C ABI.  Metric: per field max|a-b|/max|b| over all pencils,
the carried Nyquist pencils included.  Bars: 1e-12 after one step, 1e-9 after 100
steps (FP64); matrices and velocity recovery bit for bit."""
import os

import numpy as np
import pytest

import rpcf
from helpers import IC, grid, params, pencil_index, pencils_to_phys, rel

pytestmark = pytest.mark.gpu


def pair(prm):
    return rpcf.Simulation(rpcf.HIP_LIB, prm), rpcf.Simulation(rpcf.ORACLE_LIB, prm)


@pytest.mark.parametrize("name", ["C0", "TINY"])
def test_inverse_matrices_bit_identical(name):
    g, o = pair(params(name))
    s = g.s
    for eq in (0, 1):
        for p in range(s["P"]):
            assert np.array_equal(g.get_inverse(eq, p), o.get_inverse(eq, p)), (eq, p)
    g.close(), o.close()


def test_inverse_matrices_c1_sample():
    g, o = pair(params("C1"))
    s = g.s
    for eq in (0, 1):
        for kx, ky in [(0, 0), (1, 0), (0, 1), (64, 64), (17, 100), (33, 5)]:
            p = pencil_index(s, kx, ky)
            assert np.array_equal(g.get_inverse(eq, p), o.get_inverse(eq, p)), (eq, kx, ky)
    g.close(), o.close()


ODD = dict(mx=60, my=30, mz=80, n_pi_x=3.0, n_pi_y=1.5, Re=400.0, Ro=0.1, dt=0.01, u_profile=(0.0, 1.0, 0.0))
# compiled Chebyshev length (2N = 64) with ny = 24 span modes: the Chebyshev kernels' tiles
# (32 pencils for s2p, 16 for velocity and p2s) end in a partial tile
RAGGED = dict(mx=48, my=36, mz=128, n_pi_x=2.0, n_pi_y=1.0, Re=800.0, Ro=0.2, dt=0.005,
              u_profile=(1.0, 0.0, -1.0), flow_control=1)


@pytest.mark.parametrize("name", ["C0", "REF128", "TINY", "ODD", "RAGGED"])
def test_transforms_match_oracle(name):
    prm = rpcf.make_params(**{"ODD": ODD, "RAGGED": RAGGED}[name]) if name in ("ODD", "RAGGED") else params(name)
    g, o = pair(prm)
    s = g.s
    rng = np.random.default_rng(1)
    for padded in (True, False):
        planes = s["pz"] if padded else s["nz"]
        spec = (rng.standard_normal((s["P"], s["nz"])) + 1j * rng.standard_normal((s["P"], s["nz"])))
        assert rel(g.spec_to_phys(spec, padded), o.spec_to_phys(spec, padded)) < 1e-13
        phys = rng.standard_normal((planes, prm.my, prm.mx))
        assert rel(g.phys_to_spec(phys, padded), o.phys_to_spec(phys, padded)) < 1e-13
    g.close(), o.close()


def test_closed_forms_on_gpu():
    """the reference's GPU closed-form tests (test_transform.cpp, test_nonlinear.cpp,
    test_getUVW.cpp) run against the HIP library, tolerance 1e-8"""
    prm = params("REF128")
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        s = sim.s
        Z, Y, X = grid(prm, s["pz"])
        Zn, Yn, Xn = grid(prm, s["nz"])
        f = (1 - Z * Z) * np.sin(Y) * np.cos(X)
        spec = sim.phys_to_spec(f, True)
        assert np.abs(sim.spec_to_phys(spec, True) - f).max() < 1e-8
        assert np.abs(sim.spec_to_phys(spec, False) - (1 - Zn * Zn) * np.sin(Yn) * np.cos(Xn)).max() < 1e-8
        zero = np.zeros_like(Z)
        fields = [f, zero, zero, zero, -2 * Z * np.sin(Y) * np.cos(X), -(1 - Z * Z) * np.cos(Y) * np.cos(X)]
        for i, fl in enumerate(fields):
            sim.set_field(i, sim.phys_to_spec(fl, True))
        sim.nonlinear()
        nv, ne = sim.get_nonlinear()
        z, y, x = Zn, Yn, Xn
        ex_v = -(1 - z * z) * (2 * z) * (np.cos(2 * x) - np.cos(2 * y) - 2 * np.cos(2 * y) * np.cos(2 * x)) + 2 * (
            1 - z * z) * (-2 * z) * np.cos(2 * y) * np.cos(x) ** 2
        ex_e = -(1 - z * z) ** 2 * np.cos(y) * np.sin(y) * 2 * np.cos(x) * np.sin(x)
        assert np.abs(pencils_to_phys(prm, s, nv) - ex_v).max() < 1e-8
        assert np.abs(pencils_to_phys(prm, s, ne) - ex_e).max() < 1e-8
        flds = [np.sin(X), np.cos(Y), -Z * np.cos(X) + Z * np.sin(Y), -Z * np.cos(Y), Z * np.sin(X), zero]
        sp = [sim.phys_to_spec(fl, True) for fl in flds]
        sim.set_state(sp[2], sp[5])
        for i, fl in enumerate(flds):
            assert np.abs(sim.spec_to_phys(sim.get_field(i), True) - fl).max() < 1e-8, i


@pytest.mark.parametrize("name", ["C0", "REFCOUETTE", "TINY"])
def test_velocity_and_nonlinear_match(name):
    prm = params(name)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 0, *IC[name])
    g.set_state(v0, e0), o.set_state(v0, e0)
    # the kernels take the Chebyshev derivatives by a parity-class scan, the reference by the
    # ascending sum of ddz_sm: the fields agree to round-off
    for f in range(6):
        assert rel(g.get_field(f), o.get_field(f)) < 1e-13, f
    g.nonlinear(), o.nonlinear()
    a, b = g.get_nonlinear(), o.get_nonlinear()
    # the kernels sum the FFTs and the Chebyshev derivative in another order than the
    # oracle (mixed-radix Stockham stages; a parity-class scan instead of the ascending
    # sum of ddz_sm), so N agrees to round-off, not bit for bit
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.close(), o.close()


@pytest.mark.parametrize("name", ["REFCOUETTE", "TINY"])  # C0: tests/test_configs.py
def test_one_and_hundred_steps(name):
    prm = params(name)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 0, *IC[name])
    for sim in (g, o):
        sim.set_state(v0, e0)
        sim.bootstrap()
    a, b = g.get_state(), o.get_state()
    print(name, "1 step", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.step(99), o.step(99)
    a, b = g.get_state(), o.get_state()
    print(name, "100 steps", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-9 and rel(a[1], b[1]) < 1e-9
    g.close(), o.close()


def test_orr_sommerfeld_growth_rate_gpu():
    from os_mode import growth_rate_run

    sigma = growth_rate_run(rpcf.HIP_LIB)
    print("TS growth rate", sigma)
    assert abs(sigma / 0.00373967 - 1) < 1e-4


def test_gpu_errors_are_reported():
    with rpcf.Simulation(rpcf.HIP_LIB, params("TINY")) as sim:
        with pytest.raises(rpcf.RpcfError, match="no state"):
            sim.bootstrap()
        with pytest.raises(rpcf.RpcfError, match="bad field"):
            sim.set_field(9, np.zeros((sim.s["P"], sim.s["nz"]), np.complex128))


@pytest.mark.gpu
def test_channel_workload_stays_bounded():
    """the 128x129x128 channel workload (constant flow rate, seeded perturbation with peak
    velocity bound 0.1) runs its 500 steps without blowing up: the state stays finite and
    the perturbation stays of the order of its initial size"""
    prm = params("C1")
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        v0, e0 = rpcf.seed_state(sim.dll, prm, 0, *IC["C1"])
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(499)
        v, e = sim.get_state()
    assert np.isfinite(v).all() and np.isfinite(e).all()
    assert np.abs(v[1:]).max() < 10 * np.abs(v0[1:]).max()


def test_ragged_tiles_step_parity():
    """partial Chebyshev tiles through the whole step: one step to 1e-12, ten to 1e-10"""
    prm = rpcf.make_params(**RAGGED)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 5, 0.05, 4, 8, 0.0)
    for sim in (g, o):
        sim.set_state(v0, e0)
        sim.bootstrap()
    a, b = g.get_state(), o.get_state()
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.step(9), o.step(9)
    a, b = g.get_state(), o.get_state()
    assert rel(a[0], b[0]) < 1e-10 and rel(a[1], b[1]) < 1e-10
    g.close(), o.close()


def test_fallback_lengths_step_parity():
    """lengths without compiled plans (60x30x80: 2N = 40, mx = 60, my = 30) take the runtime
    FFT plans, the generic velocity, Lamb and p2s kernels and the operator products off the
    matrix cores: one step to 1e-12, ten to 1e-10"""
    prm = rpcf.make_params(**ODD)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 5, 0.05, 4, 8, 0.0)
    for sim in (g, o):
        sim.set_state(v0, e0)
        sim.bootstrap()
    a, b = g.get_state(), o.get_state()
    print("ODD 1 step", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.step(9), o.step(9)
    a, b = g.get_state(), o.get_state()
    print("ODD 10 steps", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-10 and rel(a[1], b[1]) < 1e-10
    g.close(), o.close()


# the longest compiled Chebyshev length, 2N = 384 (nz = 193, pz = 385: three-stage FFT plans,
# the largest register tiles of the Chebyshev kernels and 25 k-steps of the operator products)
DEEP = dict(mx=24, my=24, mz=768, n_pi_x=2.0, n_pi_y=1.0, Re=2000.0, Ro=0.0, dt=0.002,
            u_profile=(1.0, 0.0, -1.0), flow_control=1)


def test_longest_compiled_chebyshev_step_parity():
    """one step to 1e-12 and five to 1e-11 at nz = 193"""
    prm = rpcf.make_params(**DEEP)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 7, 0.05, 8, 8, 0.0)
    for sim in (g, o):
        sim.set_state(v0, e0)
        sim.bootstrap()
    a, b = g.get_state(), o.get_state()
    print("DEEP 1 step", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.step(4), o.step(4)
    a, b = g.get_state(), o.get_state()
    print("DEEP 5 steps", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-11 and rel(a[1], b[1]) < 1e-11
    g.close(), o.close()


# an even base profile at odd N = 45 (nz = 46): the parity pairs of the solves have no centre
# row, and 2N = 90, 4N = 180 run the runtime FFT plans and the generic transform kernels
ODDN = dict(mx=24, my=24, mz=180, n_pi_x=2.0, n_pi_y=1.0, Re=1500.0, Ro=0.1, dt=0.002,
            u_profile=(1.0, 0.0, -1.0), flow_control=1)


def test_channel_odd_n_step_parity(monkeypatch):
    """the parity-block solves with the v remainders at odd N: one step to 1e-12 and ten to
    1e-10 against the oracle, ten within 1e-12 of the full-inverse solve (the mixed form runs
    compiled Chebyshev lengths only, all of even N)"""
    prm = rpcf.make_params(**ODDN)
    g, o = pair(prm)
    v0, e0 = rpcf.seed_state(o.dll, prm, 5, 0.05, 8, 8, 0.0)
    for sim in (g, o):
        sim.set_state(v0, e0)
        sim.bootstrap()
    a, b = g.get_state(), o.get_state()
    print("odd N 1 step", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-12 and rel(a[1], b[1]) < 1e-12
    g.step(9), o.step(9)
    a, b = g.get_state(), o.get_state()
    print("odd N 10 steps", rel(a[0], b[0]), rel(a[1], b[1]))
    assert rel(a[0], b[0]) < 1e-10 and rel(a[1], b[1]) < 1e-10
    g.close(), o.close()
    monkeypatch.setenv("RPCF_FULL_INVERSE", "1")
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as f:
        f.set_state(v0, e0)
        f.bootstrap()
        f.step(9)
        c = f.get_state()
    monkeypatch.delenv("RPCF_FULL_INVERSE")
    print("odd N remainders vs full inverse after 10 steps", rel(a[0], c[0]), rel(a[1], c[1]))
    assert rel(a[0], c[0]) < 1e-12 and rel(a[1], c[1]) < 1e-12


# one grid per compiled Chebyshev length of the FP64 matrix-core p2s (2N = 24 .. 384)
P2S_GRIDS = {
    "2N24": (dict(mx=24, my=24, mz=48, n_pi_x=2.0, n_pi_y=1.0, Re=800.0, Ro=0.1, dt=0.002,
                  u_profile=(0.0, 1.0, 0.0)), (0.05, 4, 6, 0.0, True)),
    "2N64": (dict(mx=48, my=36, mz=128, n_pi_x=2.0, n_pi_y=1.0, Re=800.0, Ro=0.2, dt=0.005,
                  u_profile=(1.0, 0.0, -1.0), flow_control=1), (0.05, 4, 8, 0.0, True)),
    "2N96": (dict(mx=48, my=48, mz=192, n_pi_x=2.0, n_pi_y=1.0, Re=1500.0, Ro=0.3, dt=0.002,
                  u_profile=(1.0, 0.0, -1.0), flow_control=1), (0.05, 6, 10, 0.0, True)),
    "2N256": (None, None),  # the 128x129x128 channel
    "2N384": (DEEP, (0.05, 8, 8, 0.0, True)),
}


def nonlinear_with(lib, prm, ic, p2s_mfma):
    old = os.environ.get("RPCF_P2S_MFMA")
    os.environ["RPCF_P2S_MFMA"] = "1" if p2s_mfma else "0"
    try:
        sim = rpcf.Simulation(lib, prm)
    finally:
        if old is None:
            del os.environ["RPCF_P2S_MFMA"]
        else:
            os.environ["RPCF_P2S_MFMA"] = old
    with sim:
        sim.set_state(*rpcf.seed_state(sim.dll, prm, 3, *ic))
        sim.nonlinear()
        return sim.get_nonlinear()


@pytest.mark.parametrize("name", list(P2S_GRIDS))
def test_fp64_p2s_products_against_oracle(name):
    """the nonlinear term with the padded p2s, derivative, rhsNonlinear and unpadded s2p as
    FP64 matrix-core products (k_p2s_dmfma) against the FFT-based kernel on the same Lamb
    components (N_eta within 1e-13) and against the oracle: N_eta within 1e-12, N_v no farther
    from the oracle than the FFT-based kernel (plus 1e-13), which is within 5e-11.  The N_v bar is
    wider than N_eta's because N_v = k^2 L_z + i(alpha dL_x + beta dL_y) multiplies the
    round-off of the x/span FFTs (whose stage order differs from the oracle's) by k^2 and by
    the derivative: its largest difference sits on the span Nyquist pencil (kx = 0,
    ky = ny/2), 4.7e-13, 6.6e-12 and 1.5e-11 on the 2N = 24, 64, 96 grids, the same in both
    kernels.  At 2N = 384 the derivative's own amplification (~N^2) shows: 3.6e-13 from the
    oracle for the products, 5.6e-13 for the FFT-based kernel, 4.0e-13 between them"""
    kw, ic = P2S_GRIDS[name]
    prm = params("C1") if kw is None else rpcf.make_params(**kw)
    ic = IC["C1"] if ic is None else ic
    ov, oe = nonlinear_with(rpcf.ORACLE_LIB, prm, ic, False)
    mv, me = nonlinear_with(rpcf.HIP_LIB, prm, ic, True)
    fv, fe = nonlinear_with(rpcf.HIP_LIB, prm, ic, False)
    errs = {"N_v": rel(mv, ov), "N_eta": rel(me, oe), "N_v (FFT)": rel(fv, ov), "N_eta (FFT)": rel(fe, oe),
            "N_v MFMA vs FFT": rel(mv, fv), "N_eta MFMA vs FFT": rel(me, fe)}
    at = np.unravel_index(np.argmax(np.abs(mv - ov)), mv.shape)
    print(name, {k: "%.2e" % v for k, v in errs.items()}, "largest N_v difference at pencil, row", at)
    assert errs["N_eta MFMA vs FFT"] < 1e-13
    assert errs["N_eta"] < 1e-12 and errs["N_eta (FFT)"] < 1e-12
    assert errs["N_v"] <= errs["N_v (FFT)"] + 1e-13 and errs["N_v (FFT)"] < 5e-11
    assert not np.array_equal(mv, fv)  # both kernels ran
