"""Pins of the oracle's rotating-frame terms, independent of the reference's code:

  * the Coriolis terms of the Lamb vector (source/nonlinear.cu:421-423) against the closed form
    of the forcing f = -2 Omega x u on analytic fields (the reference's own nonlinear test field
    of tests/test_nonlinear.cpp:527-603 plus a wall-normal velocity);
  * the linear step at beta != 0 and Ro != 0 -- Squire advection, lift-up coupling
    (source/linear.cpp:137-153), Coriolis coupling, Crank-Nicolson / Adams-Bashforth assembly --
    against the independent numpy model of tests/rotation_model.py, step by step;
  * the model and the oracle against the literature: rotating plane Couette flow at Ro = 0.5
    is neutrally stable at Re = 20.66 (Lezius & Johnston 1976, J. Fluid Mech. 77, 153; the
    streamwise-invariant problem is Rayleigh-Benard convection with Ra = 16 Ro (1 - Ro) Re^2,
    Ra_c = 1707.76 at gap wavenumber 3.1163).

Each pin is shown to fail on a planted mistake (rpcf_oracle_plant: the sign of either Coriolis
term, of the lift-up coupling, of the Squire advection).  CPU only.
"""
import contextlib
import ctypes

import numpy as np
import pytest

import rpcf
from helpers import grid, params, pencil_index, pencils_to_phys, rel
from rotation_model import Mode


@contextlib.contextmanager
def planted(lib, which):
    flag = ctypes.c_int.in_dll(lib, "rpcf_oracle_plant")
    flag.value = which
    try:
        yield
    finally:
        flag.value = 0


def coriolis_response(prm_fn, Ro):
    """N_v, N_eta at the collocation points with Ro minus those with Ro = 0, for the same
    six fields set directly (so the Lamb products cancel)"""
    out = []
    for r in (Ro, 0.0):
        prm = prm_fn(r)
        with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
            s = sim.s
            Z, Y, X = grid(prm, s["pz"])
            zero = np.zeros_like(Z)
            u = (1 - Z * Z) * np.sin(Y) * np.cos(X)
            wn = (1 - Z * Z) ** 2 * np.sin(X) * np.cos(Y)  # wall-normal velocity
            fields = [u, zero, wn, zero, -2 * Z * np.sin(Y) * np.cos(X), -(1 - Z * Z) * np.cos(Y) * np.cos(X)]
            for i, fl in enumerate(fields):
                sim.set_field(i, sim.phys_to_spec(fl, True))
            sim.nonlinear()
            nv, ne = sim.get_nonlinear()
            out.append((pencils_to_phys(prm, s, nv), pencils_to_phys(prm, s, ne)))
    return (out[0][0] - out[1][0], out[0][1] - out[1][1]), s


def test_coriolis_terms_closed_form(oracle_lib):
    """With u = (1-y^2) sin z cos x and wall-normal v = (1-y^2)^2 sin x cos z (unit wavenumbers
    in the 4 pi x 2 pi box), f = -2 Omega x u = Ro (v, -u, 0) gives
        h_v   = -lap_h f_y - d/dy d/dx f_x = 2 Ro u + 4 Ro y (1-y^2) cos x cos z,
        h_eta = d/dz f_x                   = -Ro (1-y^2)^2 sin x sin z
    at every collocation point; the reference's tolerance 1e-8.  Each planted sign error of a
    Coriolis term moves the response by O(Ro)."""
    Ro = 0.3
    prm_fn = lambda r: params("REF128", Ro=r)
    (dv, de), s = coriolis_response(prm_fn, Ro)
    prm = prm_fn(Ro)
    z, y, x = grid(prm, s["nz"])
    u = (1 - z * z) * np.sin(y) * np.cos(x)
    ex_v = 2 * Ro * u + 4 * Ro * z * (1 - z * z) * np.cos(x) * np.cos(y)
    ex_e = -Ro * (1 - z * z) ** 2 * np.sin(x) * np.sin(y)
    assert np.abs(dv - ex_v).max() < 1e-8
    assert np.abs(de - ex_e).max() < 1e-8
    for which in (1, 2):
        with planted(oracle_lib, which):
            (pv, pe), _ = coriolis_response(prm_fn, Ro)
        assert max(np.abs(pv - ex_v).max(), np.abs(pe - ex_e).max()) > 0.1, which


# pencils (kx, signed ky) of the linearised check: a streamwise-invariant mode with its
# Hermitian partner, oblique modes with beta > 0 and beta < 0, and a two-dimensional one
MODES = [(0, 1), (0, -1), (1, 2), (3, -2), (2, 0)]


def _linear_case():
    # quadratic base profile: U, U' and U'' all nonzero
    prm = params("TINY", Ro=0.3, Re=400.0, dt=0.01, u_profile=(0.2, 0.7, -0.3))
    s = rpcf.sizes(rpcf.load(rpcf.ORACLE_LIB), prm)
    nz, ny = s["nz"], s["ny"]
    rng = np.random.default_rng(3)
    zc = np.polynomial.chebyshev
    one_m_y2 = np.array([0.5, 0.0, -0.5])  # 1 - y^2
    init = {}
    for kx, kys in MODES:
        if (kx, kys) == (0, -1):
            continue
        gv = rng.standard_normal(7) + 1j * rng.standard_normal(7)
        ge = rng.standard_normal(7) + 1j * rng.standard_normal(7)
        v = zc.chebmul(zc.chebmul(gv, one_m_y2), one_m_y2)
        e = zc.chebmul(ge, one_m_y2)
        vv, ee = np.zeros(nz, complex), np.zeros(nz, complex)
        vv[:len(v)], ee[:len(e)] = v, e
        init[(kx, kys)] = (1e-7 * vv, 1e-7 * ee)
    init[(0, -1)] = tuple(np.conj(a) for a in init[(0, 1)])
    return prm, s, init


def _model_run(prm, s, init, nsteps):
    c = prm.u_profile
    out = {}
    for (kx, kys), (v, e) in init.items():
        a = kx / (prm.n_pi_x / 2)
        b = kys / (prm.n_pi_y / 2)
        m = Mode(s["nz"], a, b, prm.Re, prm.Ro, lambda z: c[0] + c[1] * z + c[2] * z * z,
                 lambda z: c[1] + 2 * c[2] * z, lambda z: 2 * c[2] + 0 * z)
        out[(kx, kys)] = m.run(v, e, nsteps, prm.dt)
    return out


def _oracle_run(prm, s, init, checkpoints):
    hnx, ny = s["hnx"], s["ny"]
    v0 = np.zeros((s["P"], s["nz"]), complex)
    e0 = np.zeros_like(v0)
    idx = {}
    for (kx, kys), (v, e) in init.items():
        p = pencil_index(s, kx, kys % ny)
        idx[(kx, kys)] = p
        v0[p], e0[p] = v, e
    res = {}
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        sim.set_state(v0, e0)
        sim.bootstrap()
        done = 1
        for n in checkpoints:
            if n > done:
                sim.step(n - done)
                done = n
            res[n] = sim.get_state()
    return res, idx


def _worst(model, oracle, idx, checkpoints):
    worst = 0.0
    for n in checkpoints:
        v, e = oracle[n]
        for key, p in idx.items():
            mv, me = model[key][n - 1]
            worst = max(worst, rel(v[p], mv), rel(e[p], me))
    return worst


def test_linearised_steps_match_independent_model(oracle_lib):
    """A small perturbation (1e-7) on pencils with beta != 0 and Ro != 0 under a base profile
    with U, U', U'' != 0: after 1 and 40 steps every seeded pencil of the oracle agrees with
    the numpy model (Coriolis coupling, Squire advection, lift-up, CN/AB2 assembly) to 1e-10
    relative (about 6e-12 measured).  Two effects the model leaves out bound the amplitude
    from both sides: the products of the perturbation with itself reach the seeded pencils
    only through the mean flow, at O(amp^2) relative; and the reference's base-flow Coriolis
    term Ro (u + U) (source/nonlinear.cu:400-403) rounds u against the O(1) profile, noise of
    O(eps |U| / amp) relative (6e-10 after 40 steps at amp = 1e-9).  Every planted sign error
    breaks the agreement."""
    prm, s, init = _linear_case()
    cps = [1, 40]
    model = _model_run(prm, s, init, max(cps))
    oracle, idx = _oracle_run(prm, s, init, cps)
    worst = _worst(model, oracle, idx, cps)
    print("oracle vs linear model, worst relative difference %.2e" % worst)
    assert worst < 1e-10
    for which in (1, 2, 3, 4):
        with planted(oracle_lib, which):
            bad, _ = _oracle_run(prm, s, init, cps)
        assert _worst(model, bad, idx, cps) > 1e-4, which


# rotating plane Couette, U = y, Ro = 0.5: the least stable streamwise-invariant mode
LJ_BETA = 3.1163 / 2.0  # the Rayleigh-Benard critical wavenumber on the half gap


def _lj_mode(Re, nz=33):
    return Mode(nz, 0.0, LJ_BETA, Re, 0.5, lambda z: z, lambda z: 1.0 + 0 * z, lambda z: 0 * z)


def test_model_neutral_at_lezius_johnston_critical_reynolds():
    """the model's streamwise-invariant rolls at Ro = 0.5 are neutral at Re = 20.66 and the
    growth rate changes sign there (1707.76 = 16 Ro (1 - Ro) Re_c^2)"""
    lo, hi = 19.0, 23.0
    assert _lj_mode(lo).growth_rate()[0].real < 0 < _lj_mode(hi).growth_rate()[0].real
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if _lj_mode(mid).growth_rate()[0].real > 0:
            hi = mid
        else:
            lo = mid
    assert abs(0.5 * (lo + hi) - 20.66) < 0.005
    assert abs(0.5 * (lo + hi) - np.sqrt(1707.762 / 16 / 0.25)) < 0.005


def _oracle_growth(Re, nsteps=2000, dt=0.005):
    """growth rate of |v| of the seeded least-stable mode in the oracle (kx = 0, ky = +-1)"""
    prm = rpcf.make_params(6, 6, 128, n_pi_x=2.0, n_pi_y=2.0 / LJ_BETA, Re=Re, Ro=0.5, dt=dt,
                           u_profile=(0.0, 1.0, 0.0))
    m = _lj_mode(Re)
    sigma, v, e = m.growth_rate()
    scale = 1e-8 / np.abs(v).max()
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        v0 = np.zeros((s["P"], s["nz"]), complex)
        e0 = np.zeros_like(v0)
        p, q = pencil_index(s, 0, 1), pencil_index(s, 0, s["ny"] - 1)
        v0[p], e0[p] = v * scale, e * scale
        v0[q], e0[q] = np.conj(v0[p]), np.conj(e0[p])
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(nsteps // 2 - 1)
        n0 = np.linalg.norm(sim.get_state()[0][p])
        sim.step(nsteps // 2)
        n1 = np.linalg.norm(sim.get_state()[0][p])
    return np.log(n1 / n0) / (nsteps // 2 * dt), sigma.real


def test_oracle_rotating_couette_growth_rate(oracle_lib):
    """the oracle's growth rate of the Lezius-Johnston rolls at Re = 26 (unstable) and 17
    (stable) matches the model's eigenvalue within 2e-3 relative (time-step error at
    dt = 0.005); with either Coriolis sign or the lift-up sign planted, the growth rate at
    Re = 26 is off by more than 10 % (the coupling product beta^2 Ro (Ro - U') of the v and eta
    equations becomes -0.75 beta^2 or +0.25 / +0.75 beta^2 instead of -0.25 beta^2)"""
    for Re in (26.0, 17.0):
        got, want = _oracle_growth(Re)
        print("Re %g: oracle %.6f, model %.6f" % (Re, got, want))
        assert abs(got / want - 1) < 2e-3
    for which in (1, 2, 3):
        with planted(oracle_lib, which):
            got, want = _oracle_growth(26.0, nsteps=400)
        assert abs(got / want - 1) > 0.1, which
