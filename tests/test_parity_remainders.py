"""The solve of an even base profile with the reflection-parity blocks of its inverses
(csrc/k_linear.hpp, k_parity_blocks / gemv_sel), restated in numpy on the oracle's inverses
(CPU only): with H = (M[j][i] + s_j M[j][i'])/2 and E = (M[j][i] - s_j M[j][i'])/2,
s_j = (-1)^j s, the sum H (b_i + s_j b_i') + E (b_i - s_j b_i') is M b term for term.
Pinned here: that identity, that BF16 remainders keep the solve at the FP64 rounding level,
and the reason only the v blocks carry remainders -- dropping the eta remainders moves an
eta solve and its first two y-derivatives by round-off only, while dropping the v remainders
moves the v solve's y-derivatives (which velocity recovery takes) far more than their BF16
rounding does."""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import rpcf
from helpers import pencil_index


def pairs(eq, N):
    """row pairs (i, i', s) of the v (eq 0) and eta (eq 1) systems: the points z_p, z_{N-p},
    the wall rows of each condition (the slope rows of v with s = -1)"""
    if eq == 0:
        pr = [(0, 1, 1), (2, 3, -1)] + [(i, N + 4 - i, 1) for i in range(4, N + 4) if i <= N + 4 - i]
    else:
        pr = [(0, 1, 1)] + [(i, N + 2 - i, 1) for i in range(2, N + 1) if i <= N + 2 - i]
    return pr


def blocks(M, pr):
    nz = M.shape[0]
    H = np.zeros((nz, len(pr)), complex)
    E = np.zeros_like(H)
    for c, (i, i2, s) in enumerate(pr):
        for j in range(nz):
            sj = s * (-1) ** j
            if i == i2:  # centre row: even j in H, odd j in E
                (H if j % 2 == 0 else E)[j, c] = M[j, i]
            else:
                H[j, c] = 0.5 * (M[j, i] + sj * M[j, i2])
                E[j, c] = 0.5 * (M[j, i] - sj * M[j, i2])
    return H, E


def solve(H, E, pr, b):
    cp = np.array([b[i] + s * b[i2] if i != i2 else b[i] for (i, i2, s) in pr])
    cm = np.array([b[i] - s * b[i2] if i != i2 else 0 for (i, i2, s) in pr])
    return np.array([H[j] @ (cp if j % 2 == 0 else cm) + E[j] @ (cm if j % 2 == 0 else cp)
                     for j in range(H.shape[0])])


def bf16(x):
    """round to BF16 (nearest even) through FP32, as k_parity_blocks stores the remainders"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def channel():
    prm = rpcf.make_params(mx=48, my=48, mz=512, n_pi_x=4.0, n_pi_y=4.0 / 3.0, Re=3300.0, Ro=0.0, dt=0.002,
                           u_profile=(1.0, 0.0, -1.0), flow_control=1)
    sim = rpcf.Simulation(rpcf.ORACLE_LIB, prm)
    yield sim
    sim.close()


def test_remainders_restore_the_inverse_and_only_v_needs_them(channel):
    s = channel.s
    nz = s["nz"]
    N = nz - 1
    z = np.cos(np.pi * np.arange(nz) / N)
    eye = np.eye(nz)
    D = [np.stack([C.chebval(z, C.chebder(eye[j], m) if m else eye[j]) for j in range(nz)], 1) for m in range(3)]
    rng = np.random.default_rng(3)
    worst = {}
    for kx, ky in ((1, 0), (1, 1), (3, 5), (7, 2)):
        p = pencil_index(s, kx, ky)
        for eq in (0, 1):
            M = channel.get_inverse(eq, p)
            pr = pairs(eq, N)
            H, E = blocks(M, pr)
            b = np.zeros(nz, complex)
            lo = 4 if eq == 0 else 2  # the wall rows of the right-hand sides are zero
            for _ in range(3):
                b[lo:] += (rng.standard_normal() + 1j * rng.standard_normal()) * np.cos(
                    rng.integers(1, 20) * np.arccos(z[:nz - lo]))
            x = M @ b

            def err(Ek, m):
                d = D[m] @ (solve(H, Ek, pr, b) - x)
                return np.abs(d).max() / np.abs(D[m] @ x).max()

            Eb = bf16(E.real) + 1j * bf16(E.imag)
            for key, Ek in (("exact", E), ("bf16", Eb), ("none", 0 * E)):
                for m in range(3):
                    worst[(eq, key, m)] = max(worst.get((eq, key, m), 0.0), err(Ek, m))
    print({k: "%.1e" % v for k, v in sorted(worst.items())})
    for eq in (0, 1):
        for m in range(3):
            assert worst[(eq, "exact", m)] < 1e-13   # the identity: FP64 rounding only
            assert worst[(eq, "bf16", m)] < 2e-13    # BF16 remainders: the same level
            assert worst[(1, "none", m)] < 1e-13     # eta: no remainders needed
    # v: without its remainders the solve moves far more than with BF16 ones (measured at
    # nz = 129: 1.5e-12 against 5.5e-15 for x, 3.4e-13 against 4.7e-14 for its second
    # y-derivative, where the FP64 rounding of the blocks already stands at 5e-14)
    assert worst[(0, "none", 0)] > 20 * worst[(0, "bf16", 0)]
    assert worst[(0, "none", 2)] > 5 * worst[(0, "bf16", 2)]
