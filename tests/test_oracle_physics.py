"""Pins of the oracle's linear step, mean-flow equations, bootstrap and extensions by
properties the mathematics fixes -- the reference's own tests pin none of these
(tests/test_coef.cpp's fixture files are absent; tests/test_rhs.cpp asserts nothing).
CPU only."""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import rpcf
from helpers import IC, params, pencil_index, rel


def cheb_op(nz):
    """Chebyshev collocation operators from numpy's Chebyshev algebra: value, 2nd and 4th
    derivative of T_j at z_i = cos(pi i/N)."""
    N = nz - 1
    z = np.cos(np.pi * np.arange(nz) / N)
    E = np.eye(nz)
    T0 = np.stack([C.chebval(z, E[j]) for j in range(nz)], 1)
    T2 = np.stack([C.chebval(z, C.chebder(E[j], 2)) for j in range(nz)], 1)
    T4 = np.stack([C.chebval(z, C.chebder(E[j], 4)) if j >= 4 else np.zeros(nz) for j in range(nz)], 1)
    return z, T0, T2, T4


def test_inverse_times_crank_nicolson_operator_is_identity():
    """The v and eta matrices are the Crank-Nicolson Orr-Sommerfeld / Squire operators
    (I - dt/2 L) collocated at interior points with clamped / Dirichlet rows; built here
    from numpy's Chebyshev derivatives, A @ inverse == I."""
    prm = params("TINY", Re=700.0, dt=0.02, u_profile=(0.3, 0.8, -0.4))
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        nz = s["nz"]
        z, T0, T2, T4 = cheb_op(nz)
        c = prm.u_profile
        U, dU, ddU = c[0] + c[1] * z + c[2] * z * z, c[1] + 2 * c[2] * z, 2 * c[2] + 0 * z
        Re, dt = prm.Re, prm.dt
        j = np.arange(nz)
        for kx, ky in [(1, 0), (3, 2), (0, 5), (2, s["ny"] - 3)]:
            a = kx / (prm.n_pi_x / 2)
            b = (ky if ky <= s["ny"] // 2 else ky - s["ny"]) / (prm.n_pi_y / 2)
            k2 = a * a + b * b
            p = pencil_index(s, kx, ky)
            # v: (1 - dt/2 L_OS) (D2 - k2) v with L_OS = -i a U (D2-k2) + i a U'' + (D2-k2)^2/Re
            Av = np.zeros((nz, nz), complex)
            for r in range(4, nz):
                q = r - 2
                lap = T2[q] - k2 * T0[q]
                bil = T4[q] - 2 * k2 * T2[q] + k2 * k2 * T0[q]
                Av[r] = lap - dt / 2 * (-1j * a * U[q] * lap + 1j * a * ddU[q] * T0[q] + bil / Re)
            Av[0], Av[1], Av[2], Av[3] = 1.0, (-1.0) ** j, j * j, -((-1.0) ** j) * j * j
            Ao = np.zeros((nz, nz), complex)
            for r in range(2, nz):
                q = r - 1
                Ao[r] = T0[q] - dt / 2 * (-1j * a * U[q] * T0[q] + (T2[q] - k2 * T0[q]) / Re)
            Ao[0], Ao[1] = 1.0, (-1.0) ** j
            for A, eq in ((Av, 0), (Ao, 1)):
                X = sim.get_inverse(eq, p)
                assert np.abs(A @ X - np.eye(nz)).max() < 1e-9, (kx, ky, eq)


def test_matrices_shared_by_plus_minus_ky_and_free_of_Ro():
    """Operators depend on kx and k^2 only (source/coefficient.cu:7-64): pencils ky and
    ny-ky carry bit-identical inverses, and Ro never enters them."""
    a = rpcf.Simulation(rpcf.ORACLE_LIB, params("TINY", Ro=0.0))
    b = rpcf.Simulation(rpcf.ORACLE_LIB, params("TINY", Ro=0.37))
    s = a.s
    for kx in (0, 1, 4):
        for ky in (1, 3, 7):
            p, m = pencil_index(s, kx, ky), pencil_index(s, kx, s["ny"] - ky)
            for eq in (0, 1):
                assert np.array_equal(a.get_inverse(eq, p), a.get_inverse(eq, m))
                assert np.array_equal(a.get_inverse(eq, p), b.get_inverse(eq, p))
    a.close()
    b.close()


@pytest.mark.parametrize("profile,fc", [((0.0, 1.0, 0.0), 0), ((0.5, 0.5, 0.0), 0), ((1.0, 0.0, -1.0), 1)])
def test_laminar_base_flow_is_a_fixed_point(profile, fc):
    """Zero perturbation of laminar Couette / Poiseuille stays zero to machine precision,
    with rotation (the Coriolis term of the base flow only reaches the unused L_z(0,0))."""
    prm = params("TINY", u_profile=profile, flow_control=fc, Ro=0.3)
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        z = np.zeros((sim.s["P"], sim.s["nz"]), np.complex128)
        sim.set_state(z, z)
        sim.bootstrap()
        sim.step(20)
        v, e = sim.get_state()
        assert np.abs(v).max() < 1e-14 and np.abs(e).max() < 1e-14


def cheb_coeffs(f, nz):
    z = np.cos(np.pi * np.arange(nz) / (nz - 1))
    return C.chebfit(z, f(z), nz - 1)


@pytest.mark.parametrize("eq", [0, 1])
def test_mean_mode_diffuses_with_crank_nicolson_factor(eq):
    """u0 (eq 0) or w0 (eq 1) = sin(n pi (y+1)/2) alone: the (0,0) equation is
    du/dt = u''/Re with Dirichlet walls, so after k steps the mode is scaled by
    g^k, g = (1 - lam dt/2)/(1 + lam dt/2), lam = (n pi/2)^2/Re."""
    prm = params("TINY", Ro=0.0, Re=50.0, dt=0.01)
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        nz = s["nz"]
        n = 2
        a = cheb_coeffs(lambda y: np.sin(n * np.pi * (y + 1) / 2), nz)
        v = np.zeros((s["P"], nz), np.complex128)
        e = np.zeros_like(v)
        (e if eq else v)[0] = a
        sim.set_state(v, e)
        sim.bootstrap()
        sim.step(49)
        lam = (n * np.pi / 2) ** 2 / prm.Re
        g = (1 - lam * prm.dt / 2) / (1 + lam * prm.dt / 2)
        out = sim.get_state()[eq][0]
        assert rel(out, a * g ** 50) < 1e-9


def test_constant_flow_rate_holds():
    """flow_control = 1: int u0 dy keeps its initial value while the nonlinear mean
    stress drives u0."""
    prm = params("TINY", u_profile=(1.0, 0.0, -1.0), flow_control=1, Ro=0.0)
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        v0, e0 = rpcf.seed_state(sim.dll, prm, 11, 0.5, 4, 8, 0.0)
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(30)
        v, _ = sim.get_state()
        k = np.arange(s["nz"])
        w = np.array([2.0 / (1.0 - kk * kk) if kk % 2 == 0 else 0.0 for kk in k])
        assert np.abs(v[0]).max() > 1e-6  # the mean flow did change
        assert abs(np.dot(w, v[0])) < 1e-15
    prm0 = params("TINY", u_profile=(1.0, 0.0, -1.0), flow_control=0, Ro=0.0)
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm0) as sim:
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(30)
        v, _ = sim.get_state()
        assert abs(np.dot(w, v[0])) > 1e-9  # without control the bulk drifts


def test_bootstrap_uses_euler_first_step():
    """main_loop.cpp:39-52: after the extra step the history holds N(state0); the real
    first step then sees N^{n-1} = N^n, and both histories equal N(state0)."""
    prm = params("TINY")
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        v0, e0 = rpcf.seed_state(sim.dll, prm, 2, *IC["TINY"])
        sim.set_state(v0, e0)
        sim.nonlinear()
        n0 = sim.get_nonlinear()
        sim.set_state(v0, e0)
        sim.bootstrap()
        cur, prev = sim.get_nonlinear(), sim.get_nonlinear(prev=True)
        for a, b, c in zip(n0, cur, prev):
            assert np.array_equal(a, b) and np.array_equal(a, c)


def test_seeded_state_properties(oracle_lib):
    """v = dv/dy = 0 and eta = 0 at both walls (the BC rows of the collocation system),
    Hermitian symmetry on kx = 0, Nyquist and (0,0) pencils zero, degree <= D+4,
    deterministic in the seed."""
    prm = params("C0")
    v, e = rpcf.seed_state(oracle_lib, prm, 0, *IC["C0"])
    s = rpcf.sizes(oracle_lib, prm)
    nz, j = s["nz"], np.arange(s["nz"])
    deg = IC["C0"][2]
    assert np.abs(v @ np.ones(nz)).max() < 1e-16 and np.abs(v @ (-1.0) ** j).max() < 1e-16
    assert np.abs(v @ (j * j)).max() < 1e-14 and np.abs(v @ ((-1.0) ** j * j * j)).max() < 1e-14
    assert np.abs(e @ np.ones(nz)).max() < 1e-16 and np.abs(e @ (-1.0) ** j).max() < 1e-16
    assert np.all(v[:, deg + 5:] == 0) and np.all(e[:, deg + 3:] == 0)
    assert np.all(v[0] == 0) and np.all(e[0] == 0)
    hnx, ny = s["hnx"], s["ny"]
    for ky in range(1, ny // 2):
        assert np.array_equal(v[ky * hnx], np.conj(v[(ny - ky) * hnx]))
    assert np.all(v[(ny // 2) * hnx:(ny // 2 + 1) * hnx] == 0)
    assert np.all(v[:, :][hnx - 1::hnx] == 0)
    v2, e2 = rpcf.seed_state(oracle_lib, prm, 0, *IC["C0"])
    assert np.array_equal(v, v2) and np.array_equal(e, e2)
    v3, _ = rpcf.seed_state(oracle_lib, prm, 1, *IC["C0"])
    assert not np.array_equal(v, v3)
    nonzero = np.abs(v).max(axis=1) > 0
    assert nonzero.sum() == 4 * 9 + 8  # kx 1..4 x ky -4..4, plus kx = 0 with ky = +-1..4


@pytest.mark.slow
def test_orr_sommerfeld_growth_rate():
    """Plane Poiseuille U = 1 - y^2, Re = 10000, alpha = 1, 2D mode: the least stable
    Orr-Sommerfeld eigenvalue c = 0.23752649 + 0.00373967i (Orszag 1971) gives the growth
    rate alpha c_i = 0.00373967; the DNS at tiny amplitude must reproduce it within 1e-4
    relative."""
    from os_mode import growth_rate_run

    sigma = growth_rate_run(rpcf.ORACLE_LIB)
    assert abs(sigma / 0.00373967 - 1) < 1e-4, sigma


def test_inverse_parity_structure():
    """the structure the mixed form's parity-block solves rely on (k_parity_blocks): with an
    even base profile the inverse of either collocation system satisfies
    M[j][i'] = (-1)^j s M[j][i] for the row pairs (i, i', s) of the wall reflection -- v: rows
    0/1 (+), 2/3 (slopes, -), i/N+4-i; eta: rows 0/1, i/N+2-i -- to round-off; an odd profile
    (Couette) breaks it"""
    def pairs(eq, N):
        if eq == 0:
            pr = [(0, 1, 1), (2, 3, -1)] + [(i, N + 4 - i, 1) for i in range(4, N + 5 - 4) if i <= N + 4 - i]
        else:
            pr = [(0, 1, 1)] + [(i, N + 2 - i, 1) for i in range(2, N + 1) if i <= N + 2 - i]
        assert len(pr) == N // 2 + 1
        return pr

    def asym(prm, eq):
        with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
            s = sim.s
            N = s["nz"] - 1
            worst = 0.0
            sign = (-1.0) ** np.arange(N + 1)
            for p in (1, s["hnx"] + 3, 5 * s["hnx"] + 2):
                M = sim.get_inverse(eq, p)
                dev = max(np.abs(M[:, i2] - sign * sg * M[:, i]).max() for i, i2, sg in pairs(eq, N))
                worst = max(worst, dev / np.abs(M).max())
            return worst

    channel = params("C0", u_profile=(1.0, 0.0, -1.0), flow_control=1)
    couette = params("C0", u_profile=(0.0, 1.0, 0.0))
    for eq in (0, 1):
        assert asym(channel, eq) < 1e-12
    assert asym(couette, 0) > 1e-6
