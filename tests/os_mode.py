"""Orr-Sommerfeld check of the time stepper: plane Poiseuille flow U = 1 - y^2 at
Re = 10000, streamwise wavenumber alpha = 1 (Lx = 2 pi), spanwise-uniform mode.

The least stable eigenvalue c = 0.23752649 + 0.00373967 i (Orszag 1971, J. Fluid
Mech. 50, 689) gives the temporal growth rate alpha c_i = 0.00373967.  The eigenvector
is computed here with numpy/scipy on the same Chebyshev collocation (interior points
cos(pi p/N), p = 2..N-2, clamped walls), seeded into pencil (kx=1, ky=0) at tiny
amplitude, and the simulation's growth of |v| over T time units is measured.
"""
import numpy as np
import numpy.polynomial.chebyshev as C
import scipy.linalg

import rpcf

RE = 10000.0
ALPHA = 1.0
MZ = 320          # nz = 81 Chebyshev coefficients
DT = 0.01
NSTEPS = 2000
AMP = 1e-8


def os_eigenmode(nz, re=RE, alpha=ALPHA):
    N = nz - 1
    z = np.cos(np.pi * np.arange(nz) / N)
    E = np.eye(nz)
    T0 = np.stack([C.chebval(z, E[j]) for j in range(nz)], 1)
    T2 = np.stack([C.chebval(z, C.chebder(E[j], 2)) for j in range(nz)], 1)
    T4 = np.stack([C.chebval(z, C.chebder(E[j], 4)) if j >= 4 else np.zeros(nz) for j in range(nz)], 1)
    U, ddU = 1 - z * z, -2.0 + 0 * z
    L = T2 - alpha ** 2 * T0
    L2 = T4 - 2 * alpha ** 2 * T2 + alpha ** 4 * T0
    A = U[:, None] * L - ddU[:, None] * T0 - L2 / (1j * alpha * re)
    B = L.astype(complex)
    j = np.arange(nz)
    rows = list(range(2, N - 1))  # interior points p = 2..N-2
    A = np.vstack([A[rows], np.ones(nz), (-1.0) ** j, j * j, -((-1.0) ** j) * j * j])
    B = np.vstack([B[rows], np.zeros((4, nz))])
    w, V = scipy.linalg.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 10)
    i = np.argmax(np.where(ok, w.imag, -np.inf))
    return w[i], V[:, i]


def growth_rate_run(lib_path, nsteps=NSTEPS, dt=DT):
    prm = rpcf.make_params(6, 6, MZ, n_pi_x=2.0, n_pi_y=2.0, Re=RE, Ro=0.0, dt=dt, u_profile=(1.0, 0.0, -1.0),
                           flow_control=0)
    with rpcf.Simulation(lib_path, prm) as sim:
        s = sim.s
        c, a = os_eigenmode(s["nz"])
        a = a / np.abs(a).max() * AMP
        v = np.zeros((s["P"], s["nz"]), np.complex128)
        e = np.zeros_like(v)
        v[1] = a  # pencil kx = 1, ky = 0
        sim.set_state(v, e)
        sim.bootstrap()
        n0 = np.linalg.norm(sim.get_state()[0][1])
        sim.step(nsteps)
        n1 = np.linalg.norm(sim.get_state()[0][1])
    return np.log(n1 / n0) / (nsteps * dt)
