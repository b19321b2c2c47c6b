"""An independent model of the linearised rotating-frame velocity-vorticity equations, used to
pin the oracle's Coriolis terms (source/nonlinear.cu:421-423) and the eta right-hand side at
beta != 0 (Squire advection and lift-up coupling, source/linear.cpp:137-153).

Nothing here is read from the oracle or the reference's code.  The equations are derived from
the Navier-Stokes equations in a frame rotating about the spanwise axis, in the right-handed
frame (x streamwise, y wall-normal, z spanwise) of the Kim-Moin-Moser formulation:

    du/dt = H + f - grad p + lap u / Re,    H = u x omega,    f = -2 Omega x u,
    Omega = (Ro / 2) e_z   (rotation about the span; Ro > 0 is anti-cyclonic for U' > 0).

Perturbations (u, v, w) of a base flow U(y) e_x, Fourier modes exp(i(alpha x + beta z)),
k2 = alpha^2 + beta^2, D = d/dy.  Taking curl curl and curl of the momentum equation:

    d/dt lap v = -i alpha U lap v + i alpha U'' v + lap^2 v / Re + h_v
    d/dt eta   = -i alpha U eta - i beta U' v + lap eta / Re + h_eta        (eta = i beta u - i alpha w)
    h_v   = -k2 f_y - D(i alpha f_x + i beta f_z)        (the wall-normal part of lap f - grad div f)
    h_eta = i beta f_x - i alpha f_z                     (the wall-normal part of curl f)

u and w follow from continuity i alpha u + D v + i beta w = 0 and the definition of eta.  The
spectral variable names of the library map onto these as: library u <- u, library w (wall-normal)
<- v, library v (spanwise) <- w, library omega_z <- eta.

Discretisation, as the reference states its scheme: Chebyshev coefficients in y, collocation at
z_p = cos(pi p/N); the v equation at the interior points p = 2..N-2 with v = Dv = 0 at both
walls, the eta equation at p = 1..N-1 with eta = 0 at the walls; Crank-Nicolson for every
linear term (the lift-up coupling with v^n and v^{n+1}), second-order Adams-Bashforth for the
forcing h evaluated at the collocation points, and a first step with h^{n-1} = h^n.  The linear
operators come from numpy's Chebyshev algebra, the solves from numpy.linalg.
"""
import numpy as np
import numpy.polynomial.chebyshev as C


def tables(nz):
    """z_p and the value / 1st / 2nd / 4th derivative of T_j at z_p (row p, column j)"""
    N = nz - 1
    z = np.cos(np.pi * np.arange(nz) / N)
    E = np.eye(nz)
    T = [np.stack([C.chebval(z, C.chebder(E[j], m)) if j >= m else np.zeros(nz) for j in range(nz)], 1)
         for m in (0, 1, 2, 4)]
    return z, T[0], T[1], T[2], T[3]


def deriv_matrix(nz):
    """coefficients of D a from the coefficients a (numpy chebder, padded to nz)"""
    Dm = np.zeros((nz, nz))
    for j in range(nz):
        e = np.zeros(nz)
        e[j] = 1
        d = C.chebder(e)
        Dm[:len(d), j] = d
    return Dm


class Mode:
    """One Fourier mode (alpha, beta) of the linear problem on nz Chebyshev coefficients."""

    def __init__(self, nz, alpha, beta, Re, Ro, U, dU, ddU):
        self.nz, self.N = nz, nz - 1
        self.a, self.b, self.Re, self.Ro = alpha, beta, Re, Ro
        self.k2 = alpha * alpha + beta * beta
        self.z, self.T0, self.T1, self.T2, self.T4 = tables(nz)
        self.Dm = deriv_matrix(nz)
        self.U, self.dU, self.ddU = U(self.z), dU(self.z), ddU(self.z)

    def forcing(self, v, eta):
        """Coriolis forcing h_v, h_eta (Chebyshev coefficients) of the state (v, eta)"""
        a, b = self.a, self.b
        Dv = self.Dm @ v
        # continuity and the definition of eta, solved for (u, w) coefficient by coefficient
        M = np.array([[1j * a, 1j * b], [1j * b, -1j * a]])
        u, w = np.linalg.solve(M, np.stack([-Dv, eta]))
        vel = np.stack([u, v, w])                       # (x, y, z) components
        omega = np.array([0.0, 0.0, self.Ro / 2.0])     # rotation vector
        f = -2.0 * np.cross(omega[:, None], vel, axis=0)
        h_v = -self.k2 * f[1] - self.Dm @ (1j * a * f[0] + 1j * b * f[2])
        h_e = 1j * b * f[0] - 1j * a * f[2]
        return h_v, h_e

    def operators(self):
        """lap, the Orr-Sommerfeld operator L_OS and the Squire operator L_SQ at the points"""
        a, k2, Re = self.a, self.k2, self.Re
        T0, T2, T4 = self.T0, self.T2, self.T4
        lap = T2 - k2 * T0
        bil = T4 - 2 * k2 * T2 + k2 * k2 * T0
        los = -1j * a * self.U[:, None] * lap + 1j * a * self.ddU[:, None] * T0 + bil / Re
        lsq = -1j * a * self.U[:, None] * T0 + (T2 - k2 * T0) / Re
        return lap, los, lsq

    def wall_rows(self, order):
        j = np.arange(self.nz)
        rows = [np.ones(self.nz), (-1.0) ** j]
        if order == 4:
            rows += [j * j, -((-1.0) ** j) * j * j]
        return np.array(rows, complex)

    def stepper(self, dt):
        """the matrices of one Crank-Nicolson / Adams-Bashforth step"""
        lap, los, lsq = self.operators()
        N = self.N
        pv, pe = np.arange(2, N - 1), np.arange(1, N)
        Av = np.vstack([self.wall_rows(4), (lap - dt / 2 * los)[pv]])
        Bv = np.vstack([np.zeros((4, self.nz)), (lap + dt / 2 * los)[pv]])
        Ae = np.vstack([self.wall_rows(2), (self.T0 - dt / 2 * lsq)[pe]])
        Be = np.vstack([np.zeros((2, self.nz)), (self.T0 + dt / 2 * lsq)[pe]])
        return Av, Bv, Ae, Be, pv, pe

    def run(self, v, eta, nsteps, dt):
        """nsteps steps from (v, eta); returns the state after each step"""
        Av, Bv, Ae, Be, pv, pe = self.stepper(dt)
        lift = -1j * self.b * self.dU[:, None] * self.T0  # -i beta U' v at the points
        hv0, he0 = self.forcing(v, eta)
        hv_prev, he_prev = self.T0 @ hv0, self.T0 @ he0
        out = []
        for _ in range(nsteps):
            hv, he = self.forcing(v, eta)
            hv, he = self.T0 @ hv, self.T0 @ he
            rv = Bv @ v
            rv[4:] += dt * (1.5 * hv[pv] - 0.5 * hv_prev[pv])
            v1 = np.linalg.solve(Av, rv)
            re = Be @ eta
            re[2:] += (dt / 2 * lift @ (v + v1))[pe] + dt * (1.5 * he[pe] - 0.5 * he_prev[pe])
            eta = np.linalg.solve(Ae, re)
            v = v1
            hv_prev, he_prev = hv, he
            out.append((v.copy(), eta.copy()))
        return out

    def growth_rate(self):
        """the eigenvalue of largest real part of the continuous-time problem"""
        lap, los, lsq = self.operators()
        nz, N = self.nz, self.N
        pv, pe = np.arange(2, N - 1), np.arange(1, N)
        # the forcing as a matrix on the state (v, eta)
        Hv = np.zeros((nz, 2 * nz), complex)
        He = np.zeros((nz, 2 * nz), complex)
        for c in range(2 * nz):
            x = np.zeros(2 * nz, complex)
            x[c] = 1.0
            hv, he = self.forcing(x[:nz], x[nz:])
            Hv[:, c], He[:, c] = self.T0 @ hv, self.T0 @ he
        A = np.zeros((2 * nz, 2 * nz), complex)
        B = np.zeros((2 * nz, 2 * nz), complex)
        A[:4, :nz] = self.wall_rows(4)
        A[4:nz, :nz] = los[pv]
        A[4:nz] += Hv[pv]
        B[4:nz, :nz] = lap[pv]
        A[nz:nz + 2, nz:] = self.wall_rows(2)
        A[nz + 2:, nz:] = lsq[pe]
        A[nz + 2:, :nz] = (-1j * self.b * self.dU[:, None] * self.T0)[pe]
        A[nz + 2:] += He[pe]
        B[nz + 2:, nz:] = self.T0[pe]
        import scipy.linalg

        w, V = scipy.linalg.eig(A, B)
        ok = np.isfinite(w) & (np.abs(w) < 1e3)
        i = np.argmax(np.where(ok, w.real, -np.inf))
        return w[i], V[:nz, i], V[nz:, i]
