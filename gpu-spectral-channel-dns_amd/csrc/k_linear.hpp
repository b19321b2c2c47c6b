// k_linear.hpp -- HIP kernels of the per-mode linear algebra: collocation matrix build
// + inversion, the Crank-Nicolson/AB2 right-hand sides and solves, velocity recovery.
//
// Inverse matrices are stored column-major, one per slot (kx, kyh) with
// kyh = 0..ny/2: pencils ky and ny-ky have bit-identical operators (they depend on
// kx and k^2 only, source/coefficient.cu:7-64), so one matrix read serves both.
// Element (r, j) of slot s of equation e sits at minv[(e*nslot + s)*nz*nz + j*nz + r],
// so thread r of a solve reads consecutive addresses for every j.
//
// Arithmetic follows the reference expressions in the same order (the oracle restates
// them identically), so the matrices and the linear update are reproducible bit for bit
// from identical inputs when compiled without FP contraction.
#pragma once
#include "rpcf_common.hpp"

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmulr(double2 a, double r) { return make_double2(a.x * r, a.y * r); }

struct LinParams {
    double aphi, beta, Re, dt;
};

// get_ialpha_ibeta: source/RPCFKernels.cu:522-532
__device__ __forceinline__ void ab_of(const Dims& d, const LinParams& lp, int kx, int ky, double* ia, double* ib) {
    *ia = (double)kx / lp.aphi;
    *ib = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / lp.beta : (double)ky / lp.beta;
}

// -------------------------------------------------------------------------------------
// Build the collocation matrix of one slot and invert it (initAuxMatrix,
// source/solver.cpp:260-320): _get_coefficient_v / _omega / _get_coef_u0
// (source/coefficient.cu:7-86), then LU with partial pivoting (largest |a|^2, first on
// ties) and column-by-column solves of A X = I -- the oracle's restatement of Eigen's
// inverse (source/matrix_op.cpp:22-27).  One block per (slot, equation); the LU works on
// a row-major copy in `work` (nz*nz per block), the inverse lands column-major in minv.
// -------------------------------------------------------------------------------------
__global__ void k_build_inverse(double2* __restrict__ minv, double2* __restrict__ work, int slot0, Dims d,
                                LinParams lp, const double* __restrict__ T0, const double* __restrict__ T2,
                                const double* __restrict__ T4, const double* __restrict__ U,
                                const double* __restrict__ ddU, int* __restrict__ status) {
    const int slot = slot0 + blockIdx.x, eq = blockIdx.y;
    if (slot >= d.nslot) return;
    const int n = d.nz, N = d.N, tid = threadIdx.x, nthr = blockDim.x;
    const long nn = (long)n * n;
    double2* A = work + ((long)blockIdx.y * gridDim.x + blockIdx.x) * nn;
    double2* X = minv + ((long)eq * d.nslot + slot) * nn;
    const int kyh = slot / d.hnx, kx = slot - kyh * d.hnx;
    const double Re = lp.Re, dt = lp.dt;
    __shared__ int piv[512];
    __shared__ int s_pr;
    __shared__ double s_best;

    if (kx == 0 && kyh == 0) {
        for (long e = tid; e < nn; e += nthr) {
            const int i = (int)(e / n), j = (int)(e - (long)i * n);
            if (i >= 2) A[e] = make_double2(-T2[(long)n * (i - 1) + j] * dt * 0.5 / Re + T0[(long)n * (i - 1) + j], 0.0);
            else if (i == 0) A[e] = make_double2(1.0, 0.0);
            else A[e] = make_double2((j % 2 == 0) ? 1.0 : -1.0, 0.0);
        }
    } else {
        double alpha, beta;
        ab_of(d, lp, kx, kyh, &alpha, &beta);
        const double kmn = alpha * alpha + beta * beta;
        for (long e = tid; e < nn; e += nthr) {
            const int i = (int)(e / n), j = (int)(e - (long)i * n);
            double2 a;
            if (eq == 0) {
                if (i >= 4) {
                    const long i2 = (long)n * (i - 2) + j;
                    a.x = -T4[i2] * dt * 0.5 / Re + (1 + kmn * dt / Re) * T2[i2] - kmn * (1 + kmn * dt * 0.5 / Re) * T0[i2];
                    a.y = alpha * dt * U[i - 2] * 0.5 * T2[i2] -
                          (kmn * alpha * dt * 0.5 * U[i - 2] + alpha * dt * 0.5 * ddU[i - 2]) * T0[i2];
                } else if (i == 0) a = make_double2(1.0, 0.0);
                else if (i == 1) a = make_double2((j % 2 == 0) ? 1.0 : -1.0, 0.0);
                else if (i == 2) a = make_double2((double)(j * j), 0.0);
                else a = make_double2(((j % 2 == 0) ? -1.0 : 1.0) * j * j, 0.0);
            } else {
                if (i >= 2) {
                    const long i1 = (long)n * (i - 1) + j;
                    a.x = -T2[i1] * dt * 0.5 / Re + (1 + kmn * dt * 0.5 / Re) * T0[i1];
                    a.y = alpha * dt * 0.5 * U[i - 1] * T0[i1];
                } else if (i == 0) a = make_double2(1.0, 0.0);
                else a = make_double2((j % 2 == 0) ? 1.0 : -1.0, 0.0);
            }
            A[e] = a;
        }
    }
    (void)N;
    __syncthreads();
    int failed = 0;
    for (int k = 0; k < n; k++) {
        if (tid == 0) {
            int pr = k;
            double2 a = A[(long)n * k + k];
            double best = a.x * a.x + a.y * a.y;
            for (int r = k + 1; r < n; r++) {
                double2 b = A[(long)n * r + k];
                double v = b.x * b.x + b.y * b.y;
                if (v > best) { best = v; pr = r; }
            }
            s_pr = pr;
            s_best = best;
            piv[k] = pr;
        }
        __syncthreads();
        const int pr = s_pr;
        if (s_best == 0.0) { failed = 1; break; }
        if (pr != k)
            for (int j = tid; j < n; j += nthr) {
                double2 t = A[(long)n * k + j];
                A[(long)n * k + j] = A[(long)n * pr + j];
                A[(long)n * pr + j] = t;
            }
        __syncthreads();
        const double2 dk = A[(long)n * k + k];
        const double dd = dk.x * dk.x + dk.y * dk.y;
        const double2 dinv = make_double2(dk.x / dd, -dk.y / dd);
        for (int r = k + 1 + tid; r < n; r += nthr) A[(long)n * r + k] = zmul(A[(long)n * r + k], dinv);
        __syncthreads();
        const int m = n - k - 1;
        for (long e = tid; e < (long)m * m; e += nthr) {
            const int r = k + 1 + (int)(e / m), j = k + 1 + (int)(e - (e / m) * m);
            A[(long)n * r + j] = zsub(A[(long)n * r + j], zmul(A[(long)n * r + k], A[(long)n * k + j]));
        }
        __syncthreads();
    }
    if (failed) {
        if (tid == 0) atomicAdd(status, 1);
        return;
    }
    // X = A^{-1}: thread c solves column c; column c of X is contiguous (column-major).
    for (int c = tid; c < n; c += nthr) {
        double2* col = X + (long)c * n;
        for (int i = 0; i < n; i++) col[i] = make_double2(i == c ? 1.0 : 0.0, 0.0);
        for (int k = 0; k < n; k++)
            if (piv[k] != k) {
                double2 t = col[k];
                col[k] = col[piv[k]];
                col[piv[k]] = t;
            }
        for (int i = 0; i < n; i++) {
            double2 s = col[i];
            for (int j = 0; j < i; j++) s = zsub(s, zmul(A[(long)n * i + j], col[j]));
            col[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double2 s = col[i];
            for (int j = i + 1; j < n; j++) s = zsub(s, zmul(A[(long)n * i + j], col[j]));
            const double2 dk = A[(long)n * i + i];
            const double dd = dk.x * dk.x + dk.y * dk.y;
            col[i] = make_double2((s.x * dk.x + s.y * dk.y) / dd, (s.y * dk.x - s.x * dk.y) / dd);
        }
    }
}

// Chebyshev derivative of one complex coefficient: ddz (source/RPCFKernels.cu:476-520)
__device__ __forceinline__ double2 zddz(const double2* a, int k, int n) {
    double re = 0.0, im = 0.0;
    for (int j = k + 1; j < n; j += 2) {
        const double dm = 2 * (double)j;
        re = re + a[j].x * dm;
        im = im + a[j].y * dm;
    }
    return k == 0 ? make_double2(re * 0.5, im * 0.5) : make_double2(re, im);
}

// -------------------------------------------------------------------------------------
// Velocity recovery: getVelocityKernel_sm (source/velocity.cu:90-170) after the state
// copy of source/solver.cpp:95-96.  One block per pencil, thread k = coefficient k.
// F[6][P][nz] <- (u, v, w, ox, oy, oz) from (v-hat, eta-hat).
// -------------------------------------------------------------------------------------
__global__ void k_velocity(const double2* __restrict__ sv, const double2* __restrict__ se, double2* __restrict__ F,
                           long fstride, Dims d, LinParams lp) {
    __shared__ double2 sw[256], soz[256], su[256], sv2[256];
    const int p = blockIdx.x, kx = p % d.hnx, ky = p / d.hnx, k = threadIdx.x, nz = d.nz;
    const long o = (long)p * nz + k;
    if (k < nz) {
        sw[k] = sv[o];
        soz[k] = se[o];
    }
    __syncthreads();
    if (kx == 0 && ky == 0) {
        if (k < nz) {
            su[k] = sw[k];
            sv2[k] = soz[k];
        }
        __syncthreads();
        if (k < nz) {
            const double2 du = zddz(su, k, nz), dv = zddz(sv2, k, nz);
            F[0 * fstride + o] = su[k];
            F[1 * fstride + o] = sv2[k];
            F[2 * fstride + o] = make_double2(0.0, 0.0);
            F[3 * fstride + o] = dv;
            F[4 * fstride + o] = zmulr(du, -1.0);
            F[5 * fstride + o] = make_double2(0.0, 0.0);
        }
        return;
    }
    double ia, ib;
    ab_of(d, lp, kx, ky, &ia, &ib);
    const double kmn = ia * ia + ib * ib;
    const double kmn1 = 1.0 / kmn;
    if (k < nz) {
        const double2 t = zddz(sw, k, nz);
        su[k] = zsub(zmul(make_double2(0.0, ia * kmn1), t), zmul(make_double2(0.0, ib * kmn1), soz[k]));
        sv2[k] = zadd(zmul(make_double2(0.0, ib * kmn1), t), zmul(make_double2(0.0, ia * kmn1), soz[k]));
    }
    __syncthreads();
    if (k < nz) {
        const double2 dv = zddz(sv2, k, nz);
        const double2 ox = zadd(dv, zmul(make_double2(0.0, -1.0 * ib), sw[k]));
        const double2 du = zddz(su, k, nz);
        const double2 oy = zadd(zmulr(du, -1.0), zmul(make_double2(0.0, ia), sw[k]));
        F[0 * fstride + o] = su[k];
        F[1 * fstride + o] = sv2[k];
        F[2 * fstride + o] = sw[k];
        F[3 * fstride + o] = ox;
        F[4 * fstride + o] = oy;
        F[5 * fstride + o] = soz[k];
    }
}

// r-th entry of (minv of slot) * x, j ascending (m_multi_v, source/matrix_op.cpp:68-78)
__device__ __forceinline__ double2 gemv_row(const double2* __restrict__ M, const double2* x, int r, int n) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < n; j++) acc = zadd(acc, zmul(M[(long)j * n + r], x[j]));
    return acc;
}

struct LinState {
    const double2* nlv;   // N_v at collocation points, current / previous step
    const double2* nlvp;
    const double2* nle;   // N_eta
    const double2* nlep;
    double2* sv;          // v-hat (u0 in pencil 0), updated in place
    double2* se;          // eta-hat (w0 in pencil 0)
    const double2* minv;  // [2][nslot][nz*nz] column-major inverses
    double qw_target_re, qw_target_im;  // flow-rate target (flow_control)
    int flow_control;
};

// -------------------------------------------------------------------------------------
// One Crank-Nicolson/AB2 update of both prognostic equations for the one or two pencils
// sharing a matrix slot: _get_linear_v + AB2 (source/linear.cpp:86-125,
// source/rhs.cpp:205-215), v solve (source/solver.cpp:107-123), _get_linear_omega_y
// (source/linear.cpp:127-171), eta solve; the slot (0,0) runs the mean-flow equations
// _get_linear_u0/_w0 (source/linear.cpp:173-205) and the constant-flow-rate correction.
// One block per slot, thread r = row r.
// -------------------------------------------------------------------------------------
__global__ void k_linear(LinState st, Dims d, LinParams lp, const double* __restrict__ T0,
                         const double* __restrict__ T2, const double* __restrict__ T4, const double* __restrict__ U,
                         const double* __restrict__ dU, const double* __restrict__ ddU,
                         const double* __restrict__ qw) {
    __shared__ double2 vn[2][256], vnew[2][256], rhs[2][256], en[2][256];
    const int slot = blockIdx.x;
    const int kyh = slot / d.hnx, kx = slot - kyh * d.hnx;
    const int n = d.nz, N = d.N, r = threadIdx.x;
    const double Re = lp.Re, dt = lp.dt;
    const long nn = (long)n * n;
    const double2* Mv = st.minv + (long)slot * nn;
    const double2* Mo = st.minv + ((long)d.nslot + slot) * nn;

    if (kx == 0 && kyh == 0) {
        // mean flow: u0 with L_x(0,0), w0 with L_y(0,0) -- the (0,0) pencils of N_v, N_eta
        for (int e = 0; e < 2; e++) {
            double2* s = e ? st.se : st.sv;
            const double2* l0 = e ? st.nle : st.nlv;
            const double2* l0p = e ? st.nlep : st.nlvp;
            const double2* M = e ? Mo : Mv;
            if (r < n) vn[0][r] = s[r];
            __syncthreads();
            if (r < n) {
                double2 v = make_double2(0.0, 0.0);
                if (r >= 2) {
                    double2 acc = make_double2(0.0, 0.0);
                    for (int j = 0; j <= N; j++) {
                        const long i1 = (long)n * (r - 1) + j;
                        acc = zadd(acc, zmul(vn[0][j], make_double2(T0[i1] + dt * 0.5 / Re * T2[i1], 0.0)));
                    }
                    v = zsub(acc, zmulr(zsub(zmulr(l0[r - 1], 1.5), zmulr(l0p[r - 1], 0.5)), dt));
                }
                rhs[0][r] = v;
                // unit-forcing response for the flow-rate control: (0, 0, dt, ..., dt)
                rhs[1][r] = make_double2(r >= 2 ? dt : 0.0, 0.0);
            }
            __syncthreads();
            double2 x = make_double2(0.0, 0.0), uh = make_double2(0.0, 0.0);
            if (r < n) {
                x = gemv_row(M, rhs[0], r, n);
                uh = gemv_row(Mv, rhs[1], r, n);
                vnew[0][r] = x;
                vnew[1][r] = uh;
            }
            __syncthreads();
            if (e == 0 && st.flow_control == 1) {
                __shared__ double2 s_f;
                if (r == 0) {
                    double2 qp = make_double2(0.0, 0.0);
                    double qh = 0.0;
                    for (int k = 0; k < n; k++) {
                        qp = zadd(qp, zmulr(vnew[0][k], qw[k]));
                        qh = qh + vnew[1][k].x * qw[k];
                    }
                    s_f = make_double2((st.qw_target_re - qp.x) / qh, (st.qw_target_im - qp.y) / qh);
                }
                __syncthreads();
                if (r < n) x = zadd(x, zmul(s_f, uh));
            }
            if (r < n) s[r] = x;
            __syncthreads();
        }
        return;
    }

    // two pencils per slot: ky = kyh and ky = ny - kyh (same matrices); one if kyh is 0 or ny/2
    const int np = (kyh == 0 || kyh == d.ny / 2) ? 1 : 2;
    long po[2];
    po[0] = ((long)kyh * d.hnx + kx) * n;
    po[1] = ((long)(d.ny - kyh) * d.hnx + kx) * n;
    double alpha, beta0;
    ab_of(d, lp, kx, kyh, &alpha, &beta0);
    const double kmn = alpha * alpha + beta0 * beta0;
    double betas[2];
    {
        double a2, b2;
        ab_of(d, lp, kx, d.ny - kyh, &a2, &b2);
        betas[0] = beta0;
        betas[1] = b2;
    }
    if (r < n)
        for (int q = 0; q < np; q++) {
            vn[q][r] = st.sv[po[q] + r];
            en[q][r] = st.se[po[q] + r];
        }
    __syncthreads();
    // v right-hand side, rows 4..N at collocation point p = r-2; rows 0..3 (BCs) zero
    if (r < n) {
        double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        if (r >= 4) {
            const int pr = r - 2;
            for (int j = 0; j <= N; j++) {
                const long i2 = (long)n * pr + j;
                const double2 b = make_double2(
                    T4[i2] * dt * 0.5 / Re + (1 - kmn * dt / Re) * T2[i2] - kmn * (1 - kmn * dt * 0.5 / Re) * T0[i2],
                    -alpha * dt * 0.5 * U[pr] * T2[i2] +
                        (kmn * alpha * dt * 0.5 * U[pr] + alpha * dt * 0.5 * ddU[pr]) * T0[i2]);
                for (int q = 0; q < np; q++) acc[q] = zadd(acc[q], zmul(vn[q][j], b));
            }
            for (int q = 0; q < np; q++)
                acc[q] = zadd(acc[q], zmulr(zsub(zmulr(st.nlv[po[q] + pr], 1.5), zmulr(st.nlvp[po[q] + pr], 0.5)), dt));
        }
        for (int q = 0; q < np; q++) rhs[q][r] = acc[q];
    }
    __syncthreads();
    if (r < n)
        for (int q = 0; q < np; q++) vnew[q][r] = gemv_row(Mv, rhs[q], r, n);
    __syncthreads();
    // eta right-hand side, rows 2..N at p = r-1; rows 0,1 zero
    if (r < n) {
        double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        if (r >= 2) {
            const int pr = r - 1;
            for (int j = 0; j <= N; j++) {
                const long i1 = (long)n * pr + j;
                const double2 c = make_double2(T2[i1] * 0.5 / Re * dt + (1 - kmn * dt * 0.5 / Re) * T0[i1],
                                               -alpha * dt * 0.5 * U[pr] * T0[i1]);
                for (int q = 0; q < np; q++) {
                    acc[q] = zadd(acc[q], zmul(en[q][j], c));
                    acc[q] = zadd(acc[q], zmul(zadd(vn[q][j], vnew[q][j]),
                                               make_double2(0.0, -0.5 * betas[q] * dU[pr] * dt * T0[i1])));
                }
            }
            for (int q = 0; q < np; q++)
                acc[q] = zadd(acc[q], zmulr(zsub(zmulr(st.nle[po[q] + pr], 1.5), zmulr(st.nlep[po[q] + pr], 0.5)), dt));
        }
        for (int q = 0; q < np; q++) rhs[q][r] = acc[q];
    }
    __syncthreads();
    if (r < n)
        for (int q = 0; q < np; q++) {
            st.se[po[q] + r] = gemv_row(Mo, rhs[q], r, n);
            st.sv[po[q] + r] = vnew[q][r];
        }
}
