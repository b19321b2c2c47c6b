"""Pins of the CPU oracle (oracle/rpcf_oracle.c) against the reference's own known
answers and against what the mathematics fixes.  CPU only.

Reference tests restated here (their assertions, not their code):
  tests/test_coef.cpp:8-32        T4 known answers at N = 100
  tests/test_inverse.cpp:6-48     10*I inverse
  tests/test_transform.cpp:13-151, 218-255   padded round trip / padded->unpadded
  tests/test_nonlinear.cpp:103-143, 527-603  N_v, N_eta closed forms (Ro = 0)
  tests/test_getUVW.cpp:61-205               velocity/vorticity recovery
The reference's tolerance is isEqual(a, b, 1e-8) (source/operation.cpp:98-111).
"""
import ctypes

import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import rpcf
from helpers import IC, grid, params, pencil_index, pencils_to_phys, rel

_dp = ctypes.POINTER(ctypes.c_double)


def t_matrices(lib, N):
    fn = lib.rpcf_oracle_T_matrix
    fn.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    fn.restype = None
    T0, T2, T4 = (np.zeros((N + 1, N + 1)) for _ in range(3))
    fn(N, *(t.ctypes.data_as(_dp) for t in (T0, T2, T4)))
    return T0, T2, T4


def test_T4_reference_known_answers(oracle_lib):
    """test_coef.cpp:21-24.  The first two values carry the FP64 recurrence's rounding:
    the exact derivatives are 1.846564249246214e9 and 8.842927486045031e7 (6e-12 and
    1.7e-10 away), so agreeing with the reference's printed digits to ~1e-13 pins the
    recurrence order, not just the mathematics.  The remaining 7e-14 on the second value
    is the host libm's cos (the reference's printed value came from another C runtime)."""
    T0, T2, T4 = t_matrices(oracle_lib, 100)
    assert abs(T4[16, 100] / 1e9 - 1.846564249235145) < 2e-15
    assert abs(T4[42, 95] / 1e7 - 8.842927487588026) < 1e-12
    assert abs(T4[42, 95] / 1e7 - 8.842927486045031) > 1e-9  # the recurrence's rounding is present
    assert T4[0, 4] == 192.0
    assert T4[100, 5] == -1920.0
    assert np.all(T4[:, :4] == 0.0)


@pytest.mark.parametrize("N", [8, 32, 64])
def test_T_matrices_are_chebyshev_derivatives(oracle_lib, N):
    """T0 = T_j(z_i), T2/T4 = d2/d4 of T_j at z_i = cos(pi i/N) (numpy's Chebyshev
    algebra, independent of the recurrence)."""
    T0, T2, T4 = t_matrices(oracle_lib, N)
    z = np.cos(np.pi * np.arange(N + 1) / N)
    for j in range(N + 1):
        c = np.zeros(N + 1)
        c[j] = 1
        assert np.allclose(T0[:, j], C.chebval(z, c), atol=1e-13)
        d2 = C.chebval(z, C.chebder(c, 2))
        d4 = C.chebval(z, C.chebder(c, 4))
        assert np.abs(T2[:, j] - d2).max() <= 1e-11 * max(1.0, np.abs(d2).max())
        assert np.abs(T4[:, j] - d4).max() <= 1e-10 * max(1.0, np.abs(d4).max())


def oracle_inverse(lib, a):
    fn = lib.rpcf_oracle_inverse
    fn.argtypes = [_dp, ctypes.c_int]
    fn.restype = ctypes.c_int
    a = np.ascontiguousarray(a, np.complex128).copy()
    rc = fn(a.view(np.float64).ctypes.data_as(_dp), a.shape[0])
    return rc, a


def test_inverse_of_10I(oracle_lib):
    """test_inverse.cpp:27-48"""
    rc, inv = oracle_inverse(oracle_lib, 10 * np.eye(100))
    assert rc == 0
    assert np.array_equal(inv, 0.1 * np.eye(100))


def test_inverse_vs_lapack(oracle_lib):
    rng = np.random.default_rng(3)
    for n in (1, 5, 33):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        rc, inv = oracle_inverse(oracle_lib, a)
        assert rc == 0
        assert rel(inv, np.linalg.inv(a)) < 1e-12
    rc, _ = oracle_inverse(oracle_lib, np.zeros((4, 4)))
    assert rc == -1


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 12, 48, 96, 128, 192, 256, 384, 512, 7, 60])
def test_fft_vs_numpy(oracle_lib, n):
    fn = oracle_lib.rpcf_oracle_fft
    fn.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inv in (0, 1):
        out = np.zeros(n, np.complex128)
        assert fn(n, x.view(np.float64).ctypes.data_as(_dp), out.view(np.float64).ctypes.data_as(_dp), inv) == 0
        ref = np.fft.ifft(x) * n if inv else np.fft.fft(x)
        assert rel(out, ref) < 1e-14


def test_chebyshev_transforms_brute_force(oracle_lib):
    """s2p/p2s on the (0,0) pencil reduce to sum_k a_k cos(pi k q/(P-1)) at every (x,y)."""
    prm = params("TINY")
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        rng = np.random.default_rng(5)
        a = rng.standard_normal(s["nz"])
        spec = np.zeros((s["P"], s["nz"]), np.complex128)
        spec[0] = a
        for padded, P in ((True, s["pz"]), (False, s["nz"])):
            q = np.arange(P)
            ref = np.cos(np.pi * np.outer(q, np.arange(s["nz"])) / (P - 1)) @ a
            phys = sim.spec_to_phys(spec, padded)
            assert np.abs(phys - ref[:, None, None]).max() < 1e-13 * np.abs(a).sum()
            back = sim.phys_to_spec(phys, padded)
            assert rel(back, spec) < 1e-13


@pytest.fixture(scope="module")
def ref128():
    prm = params("REF128")
    sim = rpcf.Simulation(rpcf.ORACLE_LIB, prm)
    yield prm, sim
    sim.close()


def test_transform_closed_forms(ref128):
    """test_transform.cpp:13-151 (padded forward+backward round trip) and 218-255
    (padded forward, unpadded backward on the nz collocation planes); tolerance 1e-8."""
    prm, sim = ref128
    s = sim.s
    Z, Y, X = grid(prm, s["pz"])
    Zn, Yn, Xn = grid(prm, s["nz"])
    cases = [
        ((1 - Z * Z) * np.sin(Y) * np.cos(X), (1 - Zn * Zn) * np.sin(Yn) * np.cos(Xn)),
        (-2 * Z * np.sin(Y) * np.cos(X), -2 * Zn * np.sin(Yn) * np.cos(Xn)),
        (-(1 - Z * Z) * np.cos(Y) * np.cos(X), -(1 - Zn * Zn) * np.cos(Yn) * np.cos(Xn)),
    ]
    for f, fn in cases:
        spec = sim.phys_to_spec(f, True)
        assert np.abs(sim.spec_to_phys(spec, True) - f).max() < 1e-8
        assert np.abs(sim.spec_to_phys(spec, False) - fn).max() < 1e-8


def test_nonlinear_closed_form(ref128):
    """test_nonlinear.cpp:75-86 and 527-603: u=(1-z^2) sin y cos x, omega_y = -2z sin y
    cos x, omega_z = -(1-z^2) cos y cos x, v = w = omega_x = 0, Ro = 0."""
    prm, sim = ref128
    s = sim.s
    Z, Y, X = grid(prm, s["pz"])
    zero = np.zeros_like(Z)
    fields = [(1 - Z * Z) * np.sin(Y) * np.cos(X), zero, zero, zero, -2 * Z * np.sin(Y) * np.cos(X),
              -(1 - Z * Z) * np.cos(Y) * np.cos(X)]
    for i, f in enumerate(fields):
        sim.set_field(i, sim.phys_to_spec(f, True))
    sim.nonlinear()
    nv, ne = sim.get_nonlinear()
    z, y, x = grid(prm, s["nz"])
    ex_v = -(1 - z * z) * (2 * z) * (np.cos(2 * x) - np.cos(2 * y) - 2 * np.cos(2 * y) * np.cos(2 * x)) + 2 * (
        1 - z * z) * (-2 * z) * np.cos(2 * y) * np.cos(x) ** 2
    ex_e = -(1 - z * z) ** 2 * np.cos(y) * np.sin(y) * 2 * np.cos(x) * np.sin(x)
    assert np.abs(pencils_to_phys(prm, s, nv) - ex_v).max() < 1e-8
    assert np.abs(pencils_to_phys(prm, s, ne) - ex_e).max() < 1e-8


def test_getuvw_closed_form(ref128):
    """test_getUVW.cpp:15-59, 148-205: u = sin x, v = cos y, w = -z cos x + z sin y,
    omega_x = -z cos y, omega_y = z sin x, omega_z = 0 rebuilt from (w, omega_z)."""
    prm, sim = ref128
    s = sim.s
    Z, Y, X = grid(prm, s["pz"])
    flds = [np.sin(X), np.cos(Y), -Z * np.cos(X) + Z * np.sin(Y), -Z * np.cos(Y), Z * np.sin(X), np.zeros_like(Z)]
    spec = [sim.phys_to_spec(f, True) for f in flds]
    sim.set_state(spec[2], spec[5])
    for i, f in enumerate(flds):
        assert np.abs(sim.spec_to_phys(sim.get_field(i), True) - f).max() < 1e-8, i


def test_velocity_invariants(oracle_lib):
    """continuity i alpha u + D w + i beta v = 0 and omega_z = i beta u - i alpha v per
    pencil (code names: u streamwise, v spanwise, w wall-normal)."""
    prm = params("C0")
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        s = sim.s
        v0, e0 = rpcf.seed_state(oracle_lib, prm, 7, 1.0, 6, 10, 0.0)
        sim.set_state(v0, e0)
        u, vs, w, ox, oy, oz = (sim.get_field(i) for i in range(6))
        nz = s["nz"]
        D = np.zeros((nz, nz))
        for j in range(nz):
            c = np.zeros(nz)
            c[j] = 1
            d = C.chebder(c)
            D[:len(d), j] = d
        aphi, beta = prm.n_pi_x / 2, prm.n_pi_y / 2
        for ky in range(s["ny"]):
            for kx in range(s["hnx"]):
                if kx == 0 and ky == 0:
                    continue
                p = pencil_index(s, kx, ky)
                a = kx / aphi
                b = (ky if ky <= s["ny"] // 2 else ky - s["ny"]) / beta
                div = 1j * a * u[p] + D @ w[p] + 1j * b * vs[p]
                assert np.abs(div).max() < 1e-12 * max(1.0, np.abs(w[p]).max() * nz * nz)
                assert np.abs(oz[p] - (1j * b * u[p] - 1j * a * vs[p])).max() < 1e-12 * max(1.0, np.abs(oz[p]).max())
                # vorticity definitions in code convention omega_code = -curl(u):
                # ox = D v - i beta w, oy = -D u + i alpha w
                assert np.abs(ox[p] - (D @ vs[p] - 1j * b * w[p])).max() < 1e-11 * max(1.0, np.abs(ox[p]).max())
                assert np.abs(oy[p] - (-(D @ u[p]) + 1j * a * w[p])).max() < 1e-11 * max(1.0, np.abs(oy[p]).max())
