// k_linear.hpp -- HIP kernels of the per-mode linear algebra: collocation matrix build
// + inversion, the Crank-Nicolson/AB2 right-hand sides and solves, velocity recovery.
//
// Inverse matrices are stored column-major, one per slot (kx, kyh) with
// kyh = 0..ny/2: pencils ky and ny-ky have bit-identical operators (they depend on
// kx and k^2 only, source/coefficient.cu:7-64), so one matrix read serves both.
// Element (r, j) of slot s of equation e sits at minv[(e*nslot + s)*nz*nz + j*nz + r],
// so thread r of a solve reads consecutive addresses for every j.
//
// The matrix build, the inverse, the mean-flow update and the velocity recovery follow
// the reference expressions in the same order as the oracle (no FP contraction), so they
// reproduce it bit for bit from identical inputs.  The per-mode update regroups the
// right-hand-side sums (three dot products per row, fused multiply-adds) and splits the
// solves over four waves; it matches the oracle to rounding (about 1e-16 relative).
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
#pragma unroll 16
    for (int j = 0; j < n; j++) acc = zadd(acc, zmul(M[(long)j * n + r], x[j]));
    return acc;
}

// both pencils of a slot against one matrix: the matrix is read once
__device__ __forceinline__ void gemv_row2(const double2* __restrict__ M, const double2* x0, const double2* x1, int r,
                                          int n, double2* y0, double2* y1) {
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll 16
    for (int j = 0; j < n; j++) {
        const double2 m = M[(long)j * n + r];
        a0 = zadd(a0, zmul(m, x0[j]));
        a1 = zadd(a1, zmul(m, x1[j]));
    }
    *y0 = a0;
    *y1 = a1;
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

// Batched GEMV of the two pencils of a slot: y_q = M x_q, M column-major n x n.
// Wave w of the block sums the columns j in its quarter of the range for the RPL rows
// lane, lane+64, ... (coalesced 1 KB loads, RPL*J/4 of them independent per lane); the
// four partial sums are added through LDS (part[4][2][nz]) in wave order.
template <int RPL>
__device__ __forceinline__ void gemv_slot(const double2* __restrict__ M, const double2* x0, const double2* x1, int n,
                                          double2* part, double2* y0, double2* y1) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int jb = (n * w) / 4, je = (n * (w + 1)) / 4;
    double2 a0[RPL], a1[RPL];
#pragma unroll
    for (int i = 0; i < RPL; i++) a0[i] = a1[i] = make_double2(0.0, 0.0);
#pragma unroll 2
    for (int j = jb; j < je; j++) {
        const double2 u0 = x0[j], u1 = x1[j];
        const double2* col = M + (long)j * n;
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            const int r = lane + 64 * i;
            if (r < n) {
                const double2 m = col[r];
                a0[i].x = fma(m.x, u0.x, fma(-m.y, u0.y, a0[i].x));
                a0[i].y = fma(m.x, u0.y, fma(m.y, u0.x, a0[i].y));
                a1[i].x = fma(m.x, u1.x, fma(-m.y, u1.y, a1[i].x));
                a1[i].y = fma(m.x, u1.y, fma(m.y, u1.x, a1[i].y));
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RPL; i++) {
        const int r = lane + 64 * i;
        if (r < n) {
            part[(w * 2 + 0) * n + r] = a0[i];
            part[(w * 2 + 1) * n + r] = a1[i];
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double2 s0 = part[r], s1 = part[n + r];
        for (int q = 1; q < 4; q++) {
            s0 = zadd(s0, part[(q * 2) * n + r]);
            s1 = zadd(s1, part[(q * 2 + 1) * n + r]);
        }
        y0[r] = s0;
        y1[r] = s1;
    }
    __syncthreads();
}

// Mean-flow equations of the (0,0) slot: _get_linear_u0/_w0 (source/linear.cpp:173-205),
// their solves, and the constant-flow-rate correction of u0 (u0 += f u_h with u_h the
// response to a unit uniform forcing and f restoring the initial bulk).  Runs as block 0
// of k_linear so that it overlaps the per-mode blocks.
template <int RPL>
__device__ void linear_mean_block(const LinState& st, const Dims& d, const LinParams& lp, const double* __restrict__ T0t,
                                  const double* __restrict__ T2t, const double* __restrict__ qw, double2* lds) {
    const int n = d.nz, N = d.N, r = threadIdx.x;
    const double Re = lp.Re, dt = lp.dt;
    const long nn = (long)n * n;
    double2* vn = lds;
    double2* rhs0 = lds + n;
    double2* rhs1 = lds + 2 * n;
    double2* x0 = lds + 3 * n;
    double2* x1 = lds + 4 * n;
    double2* sf = lds + 5 * n;
    double2* part = lds + 8 * n;
    for (int e = 0; e < 2; e++) {
        double2* s = e ? st.se : st.sv;
        const double2* l0 = e ? st.nle : st.nlv;
        const double2* l0p = e ? st.nlep : st.nlvp;
        if (r < n) vn[r] = s[r];
        __syncthreads();
        if (r < n) {
            double2 v = make_double2(0.0, 0.0);
            if (r >= 2) {
                double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
                for (int j = 0; j <= N; j++) {
                    const long i1 = (long)n * j + (r - 1);
                    acc = zadd(acc, zmul(vn[j], make_double2(T0t[i1] + dt * 0.5 / Re * T2t[i1], 0.0)));
                }
                v = zsub(acc, zmulr(zsub(zmulr(l0[r - 1], 1.5), zmulr(l0p[r - 1], 0.5)), dt));
            }
            rhs0[r] = v;
            rhs1[r] = make_double2(r >= 2 ? dt : 0.0, 0.0);  // unit uniform forcing
        }
        __syncthreads();
        gemv_slot<RPL>(st.minv + (e ? (long)d.nslot * nn : 0), rhs0, rhs1, n, part, x0, x1);
        if (e == 0 && st.flow_control == 1) {
            if (r == 0) {
                double2 qp = make_double2(0.0, 0.0);
                double qh = 0.0;
                for (int k = 0; k < n; k++) {
                    qp = zadd(qp, zmulr(x0[k], qw[k]));
                    qh = qh + x1[k].x * qw[k];
                }
                sf[0] = make_double2((st.qw_target_re - qp.x) / qh, (st.qw_target_im - qp.y) / qh);
            }
            __syncthreads();
            if (r < n) x0[r] = zadd(x0[r], zmul(sf[0], x1[r]));
        }
        if (r < n) s[r] = x0[r];
        __syncthreads();
    }
}

// -------------------------------------------------------------------------------------
// One Crank-Nicolson/AB2 update of both prognostic equations for the one or two pencils
// sharing a matrix slot (every slot but (0,0)): _get_linear_v + AB2
// (source/linear.cpp:86-125, source/rhs.cpp:205-215), v solve (source/solver.cpp:107-123),
// _get_linear_omega_y (source/linear.cpp:127-171), eta solve.
// The right-hand sides use the linearity of the reference's per-element operator,
//   sum_j v_j B_rj = (T4q v)_p + c2 (T2 v)_p - c0 (T0 v)_p + i [a2_p (T2 v)_p + a0_p (T0 v)_p],
// so each row needs three real dot products per pencil (fused multiply-adds) instead of
// rebuilding B_rj; T0, T2 and the quotients T4q = T4*dt*0.5/Re, T2q = T2*0.5/Re*dt arrive
// transposed (element (p, j) at j*nz + p: consecutive rows, consecutive addresses).
// Block of 256 threads; rows r < nz map to threads.
// -------------------------------------------------------------------------------------
// LDS (dynamic, 16*nz*16 bytes): vn, en, rhs, vnew (two pencils each) and part[4][2][nz].
// Block 0 runs the mean-flow equations of slot (0,0); block b > 0 runs slot b.
template <int RPL>
__global__ __launch_bounds__(256, 4) void k_linear(LinState st, Dims d, LinParams lp, const double* __restrict__ T0,
                                                   const double* __restrict__ T2, const double* __restrict__ T4q,
                                                   const double* __restrict__ T2q, const double* __restrict__ U,
                                                   const double* __restrict__ dU, const double* __restrict__ ddU,
                                                   const double* __restrict__ qw) {
    extern __shared__ __align__(16) unsigned char lin_lds_raw[];
    const int n = d.nz, N = d.N, r = threadIdx.x;
    double2* lds = (double2*)lin_lds_raw;
    double2* vn[2] = {lds, lds + n};
    double2* en[2] = {lds + 2 * n, lds + 3 * n};
    double2* rhs[2] = {lds + 4 * n, lds + 5 * n};
    double2* vnew[2] = {lds + 6 * n, lds + 7 * n};
    double2* part = lds + 8 * n;
    if (blockIdx.x == 0) {
        linear_mean_block<RPL>(st, d, lp, T0, T2, qw, lds);
        return;
    }
    const int slot = blockIdx.x;
    const int kyh = slot / d.hnx, kx = slot - kyh * d.hnx;
    const double Re = lp.Re, dt = lp.dt;
    const long nn = (long)n * n;
    const double2* Mv = st.minv + (long)slot * nn;
    const double2* Mo = st.minv + ((long)d.nslot + slot) * nn;
    // pencils ky = kyh and ky = ny - kyh share the slot; with kyh = 0 or ny/2 the slot has
    // one pencil and the second lane repeats it without being written back
    const bool two = !(kyh == 0 || kyh == d.ny / 2);
    const long po0 = ((long)kyh * d.hnx + kx) * n;
    const long po1 = two ? ((long)(d.ny - kyh) * d.hnx + kx) * n : po0;
    double alpha, beta0, beta1, a_unused;
    ab_of(d, lp, kx, kyh, &alpha, &beta0);
    ab_of(d, lp, kx, two ? d.ny - kyh : kyh, &a_unused, &beta1);
    const double kmn = alpha * alpha + beta0 * beta0;
    if (r < n) {
        vn[0][r] = st.sv[po0 + r];
        en[0][r] = st.se[po0 + r];
        vn[1][r] = st.sv[po1 + r];
        en[1][r] = st.se[po1 + r];
    }
    __syncthreads();
    // v right-hand side, rows 4..N at collocation point p = r-2; rows 0..3 (BCs) zero
    if (r < n) {
        double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
        if (r >= 4) {
            const int pr = r - 2;
            double s0r = 0, s0i = 0, s2r = 0, s2i = 0, s4r = 0, s4i = 0;  // pencil 0: T0 v, T2 v, T4q v
            double t0r = 0, t0i = 0, t2r = 0, t2i = 0, t4r = 0, t4i = 0;  // pencil 1
#pragma unroll 2
            for (int j = 0; j <= N; j++) {
                const long ij = (long)n * j + pr;
                const double c0 = T0[ij], c2 = T2[ij], c4 = T4q[ij];
                const double2 x = vn[0][j], y = vn[1][j];
                s0r = fma(c0, x.x, s0r); s0i = fma(c0, x.y, s0i);
                s2r = fma(c2, x.x, s2r); s2i = fma(c2, x.y, s2i);
                s4r = fma(c4, x.x, s4r); s4i = fma(c4, x.y, s4i);
                t0r = fma(c0, y.x, t0r); t0i = fma(c0, y.y, t0i);
                t2r = fma(c2, y.x, t2r); t2i = fma(c2, y.y, t2i);
                t4r = fma(c4, y.x, t4r); t4i = fma(c4, y.y, t4i);
            }
            const double k2 = (1 - kmn * dt / Re), k0 = kmn * (1 - kmn * dt * 0.5 / Re);
            const double i2 = -alpha * dt * 0.5 * U[pr];
            const double i0 = (kmn * alpha * dt * 0.5 * U[pr] + alpha * dt * 0.5 * ddU[pr]);
            // (B v) = (T4q v + k2 T2 v - k0 T0 v) + i (i2 T2 v + i0 T0 v), with complex v
            const double br0 = s4r + k2 * s2r - k0 * s0r, bi0 = s4i + k2 * s2i - k0 * s0i;
            const double br1 = t4r + k2 * t2r - k0 * t0r, bi1 = t4i + k2 * t2i - k0 * t0i;
            a0 = make_double2(br0 - (i2 * s2i + i0 * s0i), bi0 + (i2 * s2r + i0 * s0r));
            a1 = make_double2(br1 - (i2 * t2i + i0 * t0i), bi1 + (i2 * t2r + i0 * t0r));
            a0 = zadd(a0, zmulr(zsub(zmulr(st.nlv[po0 + pr], 1.5), zmulr(st.nlvp[po0 + pr], 0.5)), dt));
            a1 = zadd(a1, zmulr(zsub(zmulr(st.nlv[po1 + pr], 1.5), zmulr(st.nlvp[po1 + pr], 0.5)), dt));
        }
        rhs[0][r] = a0;
        rhs[1][r] = a1;
    }
    __syncthreads();
    gemv_slot<RPL>(Mv, rhs[0], rhs[1], n, part, vnew[0], vnew[1]);
    // eta right-hand side, rows 2..N at p = r-1; rows 0,1 zero:
    //   sum_j eta_j C_rj + (v_j + vnew_j) E_rj with C = T2q + c0 T0 + i ci T0, E = i e T0
    if (r < n) {
        double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
        if (r >= 2) {
            const int pr = r - 1;
            double s0r = 0, s0i = 0, s2r = 0, s2i = 0, w0r = 0, w0i = 0;  // pencil 0: T0 eta, T2q eta, T0 (v+vnew)
            double t0r = 0, t0i = 0, t2r = 0, t2i = 0, z0r = 0, z0i = 0;  // pencil 1
#pragma unroll 2
            for (int j = 0; j <= N; j++) {
                const long ij = (long)n * j + pr;
                const double c0 = T0[ij], c2 = T2q[ij];
                const double2 x = en[0][j], y = en[1][j];
                const double2 vx = zadd(vn[0][j], vnew[0][j]), vy = zadd(vn[1][j], vnew[1][j]);
                s0r = fma(c0, x.x, s0r); s0i = fma(c0, x.y, s0i);
                s2r = fma(c2, x.x, s2r); s2i = fma(c2, x.y, s2i);
                w0r = fma(c0, vx.x, w0r); w0i = fma(c0, vx.y, w0i);
                t0r = fma(c0, y.x, t0r); t0i = fma(c0, y.y, t0i);
                t2r = fma(c2, y.x, t2r); t2i = fma(c2, y.y, t2i);
                z0r = fma(c0, vy.x, z0r); z0i = fma(c0, vy.y, z0i);
            }
            const double c0k = (1 - kmn * dt * 0.5 / Re), ci = -alpha * dt * 0.5 * U[pr];
            const double e0 = -0.5 * beta0 * dU[pr] * dt, e1 = -0.5 * beta1 * dU[pr] * dt;
            // C eta = (T2q eta + c0k T0 eta) + i ci T0 eta ; E w = i e T0 w
            a0 = make_double2(s2r + c0k * s0r - ci * s0i - e0 * w0i, s2i + c0k * s0i + ci * s0r + e0 * w0r);
            a1 = make_double2(t2r + c0k * t0r - ci * t0i - e1 * z0i, t2i + c0k * t0i + ci * t0r + e1 * z0r);
            a0 = zadd(a0, zmulr(zsub(zmulr(st.nle[po0 + pr], 1.5), zmulr(st.nlep[po0 + pr], 0.5)), dt));
            a1 = zadd(a1, zmulr(zsub(zmulr(st.nle[po1 + pr], 1.5), zmulr(st.nlep[po1 + pr], 0.5)), dt));
        }
        rhs[0][r] = a0;
        rhs[1][r] = a1;
    }
    __syncthreads();
    gemv_slot<RPL>(Mo, rhs[0], rhs[1], n, part, en[0], en[1]);
    if (r < n) {
        st.se[po0 + r] = en[0][r];
        st.sv[po0 + r] = vnew[0][r];
        if (two) {
            st.se[po1 + r] = en[1][r];
            st.sv[po1 + r] = vnew[1][r];
        }
    }
}
