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
// copy of source/solver.cpp:95-96.  One block per pencil (x) and member (y), thread k =
// coefficient k.  F[6][P][nz] <- (u, v, w, ox, oy, oz) from (v-hat, eta-hat).
// -------------------------------------------------------------------------------------
__global__ void k_velocity(const double2* __restrict__ sv, const double2* __restrict__ se, long sS,
                           double2* __restrict__ F, long fstride, long sF, Dims d, Members mb) {
    __shared__ double2 sw[256], soz[256], su[256], sv2[256];
    const int slot = mb.ids[blockIdx.y];
    const MemberParams& mp = mb.mp[slot];
    LinParams lp;
    lp.aphi = mp.aphi;
    lp.beta = mp.beta;
    lp.Re = mp.Re;
    lp.dt = mp.dt;
    sv += slot * sS;
    se += slot * sS;
    F += slot * sF;
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
    const double2* minvh; // [2][nslot][np][nz] their parity blocks (k_parity_blocks; mixed form: v only)
    const unsigned* minve; // [nslot][np][nz] FP64 form: the v blocks' remainders; mixed form: the
                           // low parts of the eta blocks (BF16 (re, im) pairs)
    const float2* minvf;   // [nslot][np][nz] mixed form: the high parts of the eta blocks (FP32)
    double qw_target_re, qw_target_im;  // flow-rate target (flow_control)
    int flow_control;
    int parity;           // 1: solve with the parity blocks, 2: with the blocks and remainders
};

// -------------------------------------------------------------------------------------
// Operator products of the Crank-Nicolson right-hand sides for every pencil:
//   W[0] = T0 v, W[1] = T2 v, W[2] = T4q v, W[3] = T0 eta, W[4] = T2q eta   (W[m][P][nz])
// with T4q = T4*dt*0.5/Re and T2q = T2*0.5/Re*dt (the quotients of source/linear.cpp:100,
// 142).  The reference rebuilds B_rj = f(T0, T2, T4) per element (source/linear.cpp:95-107,
// 137-153); by linearity sum_j v_j B_rj is a combination of these row products, so one
// pass over the operator tables serves the NPB pencils of a block (real-matrix times
// complex-vector products, fused multiply-adds).  Register tile: a thread owns 4 rows x
// 8 pencils of one product m, so each vector element read from LDS feeds 8 FMAs
// (tables transposed: Tt[j*nz + p], vectors in LDS as [j][pencil]).
// -------------------------------------------------------------------------------------
constexpr int kTvRows = 4, kTvCols = 8;

template <int NPB>
__global__ __launch_bounds__(384) void k_lin_tv(const double2* __restrict__ sv, const double2* __restrict__ se, long sS,
                                                double2* __restrict__ W, long sWl, Dims d, const double* __restrict__ T0t,
                                                const double* __restrict__ T2t, const double* __restrict__ T4q,
                                                const double* __restrict__ T2q, long sQ, const int* __restrict__ ids) {
    {
        const long slot = ids[blockIdx.y];
        sv += slot * sS;
        se += slot * sS;
        W += slot * sWl;
        T4q += slot * sQ;
        T2q += slot * sQ;
    }
    static_assert(NPB % kTvCols == 0, "pencil tile");
    extern __shared__ __align__(16) unsigned char tv_lds_raw[];
    constexpr int LDX = NPB + 1;  // padded row: the transposing stores hit different banks
    double2* xv = (double2*)tv_lds_raw;  // [nz][LDX]
    double2* xe = xv + (size_t)d.nz * LDX;
    const int n = d.nz, p0 = blockIdx.x * NPB;
    const int np = min(NPB, d.P - p0);
    for (int i = threadIdx.x; i < NPB * n; i += blockDim.x) {
        const int q = i / n, j = i - q * n;
        const bool ok = q < np;
        xv[j * LDX + q] = ok ? sv[(long)(p0 + q) * n + j] : make_double2(0.0, 0.0);
        xe[j * LDX + q] = ok ? se[(long)(p0 + q) * n + j] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    constexpr int NH = NPB / kTvCols;
    const int ngr = (n + kTvRows - 1) / kTvRows;
    for (int u = threadIdx.x; u < 5 * ngr * NH; u += blockDim.x) {
        const int m = u / (ngr * NH), rem = u - m * ngr * NH;
        const int g = rem / NH, h = rem - g * NH;
        const int pb = g * kTvRows, qb = h * kTvCols;
        const double* T = m == 0 || m == 3 ? T0t : (m == 1 ? T2t : (m == 2 ? T4q : T2q));
        const double2* X = (m < 3 ? xv : xe) + qb;
        double ar[kTvRows][kTvCols], ai[kTvRows][kTvCols];
#pragma unroll
        for (int i = 0; i < kTvRows; i++)
#pragma unroll
            for (int q = 0; q < kTvCols; q++) ar[i][q] = ai[i][q] = 0.0;
        // the table column of j+1 is loaded while column j is used
        double cn[kTvRows];
#pragma unroll
        for (int i = 0; i < kTvRows; i++) cn[i] = pb + i < n ? T[pb + i] : 0.0;
        for (int j = 0; j < n; j++) {
            double c[kTvRows];
#pragma unroll
            for (int i = 0; i < kTvRows; i++) {
                c[i] = cn[i];
                cn[i] = (j + 1 < n && pb + i < n) ? T[(long)(j + 1) * n + pb + i] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < kTvCols; q++) {
                const double2 x = X[j * LDX + q];
#pragma unroll
                for (int i = 0; i < kTvRows; i++) {
                    ar[i][q] = fma(c[i], x.x, ar[i][q]);
                    ai[i][q] = fma(c[i], x.y, ai[i][q]);
                }
            }
        }
        double2* out = W + (long)m * d.P * n;
#pragma unroll
        for (int q = 0; q < kTvCols; q++)
            if (qb + q < np)
#pragma unroll
                for (int i = 0; i < kTvRows; i++)
                    if (pb + i < n) out[(long)(p0 + qb + q) * n + pb + i] = make_double2(ar[i][q], ai[i][q]);
    }
}

// -------------------------------------------------------------------------------------
// The same operator products on the matrix cores: W[m] = T_m X with X the nz x 2P real
// matrix of (Re, Im) columns, by v_mfma_f64_16x16x4_f64 (A: lane l holds A[l%16][l/16],
// B: lane l holds B[l/16][l%16], D: item r of lane l is D[l/16 + 4r][l%16]).
// The Chebyshev tables have the reflection parity T_m[N-i][j] = (-1)^j T_m[i][j] (m = 0, 2,
// 4 at the symmetric points), so with E_i = sum_{j even} T_m[i][j] x_j and
// O_i = sum_{j odd} T_m[i][j] x_j for the rows i <= N/2,
//   W[i] = E_i + O_i,   W[N-i] = E_i - O_i
// -- half the matrix-core work of the full product.  Tf[m][par][rt][s][64]: the 16x4 tile
// (rows 16rt.., columns 4s.. of the even (par 0: j = 2k) or odd (par 1: j = 2k+1) columns)
// of T_m in A-operand lane order, zero outside, so a chunk is one contiguous copy.
// Block: operator m = blockIdx.y % 5 of member blockIdx.y / 5, 32 pencils, 4 waves of 8
// pencils (16 columns); a wave keeps its B fragments (even and odd coefficients, S k-steps
// each) in registers; the A fragments of kTvChunk row tiles of both parities are staged in
// LDS once per block and chunk.
// -------------------------------------------------------------------------------------
typedef double rpcf_v4d __attribute__((ext_vector_type(4)));
constexpr int kTvChunk = 3;
constexpr int kTvBlockPencils = 32;

// three waves per SIMD (168 registers, no accumulation registers) up to S = 17: the
// B-fragment loads of more waves in flight; two waves (136 + 48) took 76 us against 68 at
// 128x129x128.  Longer Chebyshev lengths (S = 25, 33) hold 4 S B-fragment registers and
// would spill at three waves, so they run two.
constexpr int tv_waves(int s) { return s <= 17 ? 3 : 2; }
template <int S>  // k-steps per parity: S >= ceil((N/2 + 1)/4)
__global__ __launch_bounds__(256, tv_waves(S)) void k_lin_tv_mfma(const double2* __restrict__ sv, const double2* __restrict__ se,
                                                     long sS, double2* __restrict__ W, long sWl, Dims d,
                                                     const double* __restrict__ Tf, long sTf, int RT,
                                                     const int* __restrict__ ids) {
    extern __shared__ __align__(16) double tvA[];  // [2][kTvChunk][S][64]
    const int m = blockIdx.y % 5, tid = threadIdx.x, lane = tid & 63;
    {
        const long slot = ids[blockIdx.y / 5];
        sv += slot * sS;
        se += slot * sS;
        W += slot * sWl;
        Tf += slot * sTf;
    }
    const int n = d.nz, N = d.N, col = lane & 15, kq = lane >> 4;
    const int p = blockIdx.x * kTvBlockPencils + (tid >> 6) * 8 + (col >> 1);
    const bool okp = p < d.P;
    const double* xp = (const double*)((m < 3 ? sv : se) + (long)(okp ? p : 0) * n) + (col & 1);
    double be[S], bo[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        // unconditional loads from clamped addresses, masked by a multiply (a select lets
        // the compiler sink the loads into a branch that waits on them)
        const int je = 2 * (4 * s + kq), jo = je + 1;
        be[s] = xp[2 * min(je, n - 1)] * ((okp && je < n) ? 1.0 : 0.0);
        bo[s] = xp[2 * min(jo, n - 1)] * ((okp && jo < n) ? 1.0 : 0.0);
    }
    const double2* A = (const double2*)(Tf + (long)m * 2 * RT * S * 64);
    double* Wm = (double*)(W + (long)m * d.P * n) + (long)p * 2 * n + (col & 1);
    const int rows = N / 2 + 1;
    for (int rc = 0; rc < RT; rc += kTvChunk) {
        const int ct = min(kTvChunk, RT - rc);
        const int cnt = ct * S * 32;  // double2 elements of one parity's chunk
        __syncthreads();
        constexpr int KC = (kTvChunk * S * 32 + 255) / 256;
#pragma unroll
        for (int par = 0; par < 2; par++) {
            const double2* src = A + ((long)par * RT + rc) * S * 32;
            double vx[KC], vy[KC];
#pragma unroll
            for (int u = 0; u < KC; u++) {
                const int i = min(tid + u * 256, cnt - 1);
                const double2 t = src[i];
                vx[u] = t.x;
                vy[u] = t.y;
            }
#pragma unroll
            for (int u = 0; u < KC; u++)
                if (tid + u * 256 < cnt)
                    ((double2*)tvA)[par * kTvChunk * S * 32 + tid + u * 256] = make_double2(vx[u], vy[u]);
        }
        __syncthreads();
        // tiles beyond RT hold stale values; their rows are > N/2 and are not stored
        rpcf_v4d ae[kTvChunk], ao[kTvChunk];
#pragma unroll
        for (int t = 0; t < kTvChunk; t++) ae[t] = ao[t] = rpcf_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < S; s++)
#pragma unroll
            for (int t = 0; t < kTvChunk; t++) {
                ae[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tvA[(t * S + s) * 64 + lane], be[s], ae[t], 0, 0, 0);
                ao[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tvA[((kTvChunk + t) * S + s) * 64 + lane], bo[s], ao[t],
                                                             0, 0, 0);
            }
        if (okp) {
#pragma unroll
            for (int t = 0; t < kTvChunk; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = 16 * (rc + t) + kq + 4 * r;
                    if (row < rows) {
                        Wm[2 * row] = ae[t][r] + ao[t][r];
                        if (N - row != row) Wm[2 * (N - row)] = ae[t][r] - ao[t][r];
                    }
                }
        }
    }
}

// The same products from persistent blocks: block (i, 5 j + m) copies the table of operator
// m of member j (both parities, all RT row tiles: 2 RT S 64 doubles) into LDS once, then its
// four waves take the groups of 8 pencils g = 4 i + w, 4 i + w + 4 G, .. (G blocks per
// operator), the next group's even and odd coefficients loaded while the current group's
// RT x 2 chains of S products run (one block, one wave per SIMD per compute unit).  The table is read from L2 once per block instead of once
// per 32 pencils.  Same tiles and summation order as k_lin_tv_mfma: bit-identical results.
template <int S, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_lin_tv_stream(const double2* __restrict__ sv, const double2* __restrict__ se,
                                                       long sS, double2* __restrict__ W, long sWl, Dims d,
                                                       const double* __restrict__ Tf, long sTf, int RT,
                                                       const int* __restrict__ ids) {
    extern __shared__ __align__(16) double tvT[];  // [2][RT][S][64]
    const int m = blockIdx.y % 5, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const long slot = ids[blockIdx.y / 5];
        sv += slot * sS;
        se += slot * sS;
        W += slot * sWl;
        Tf += slot * sTf;
    }
    const int cnt = RT * S * 64;  // double2 elements of the operator's table (both parities)
    {
        const double2* src = (const double2*)(Tf + (long)m * 2 * RT * S * 64);
        for (int i = tid; i < cnt; i += blockDim.x) ((double2*)tvT)[i] = src[i];
    }
    __syncthreads();
    const int n = d.nz, N = d.N, col = lane & 15, kq = lane >> 4, rows = N / 2 + 1;
    const int ngroups = (d.P + 7) / 8, stride = gridDim.x * NW;
    const double2* x = m < 3 ? sv : se;
    double* Wm0 = (double*)(W + (long)m * d.P * n) + (col & 1);
    double be[S], bo[S];
    auto load = [&](int g) {
        const int p = g * 8 + (col >> 1);
        const bool okp = g < ngroups && p < d.P;
        const double* xp = (const double*)(x + (long)(okp ? p : 0) * n) + (col & 1);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int je = 2 * (4 * s + kq), jo = je + 1;
            be[s] = xp[2 * min(je, n - 1)] * ((okp && je < n) ? 1.0 : 0.0);
            bo[s] = xp[2 * min(jo, n - 1)] * ((okp && jo < n) ? 1.0 : 0.0);
        }
    };
    int g = blockIdx.x * NW + wave;
    if constexpr (NW == 4) load(g);
    for (; g < ngroups; g += stride) {
        double ce[S], co[S];
        if constexpr (NW == 4) {
#pragma unroll
            for (int s = 0; s < S; s++) ce[s] = be[s], co[s] = bo[s];
            load(g + stride);  // the next group's coefficients, in flight during this group's products
        } else {
            // two waves per SIMD: the other wave's products hide this group's loads
            load(g);
#pragma unroll
            for (int s = 0; s < S; s++) ce[s] = be[s], co[s] = bo[s];
        }
        const int p = g * 8 + (col >> 1);
        double* Wm = Wm0 + (long)p * 2 * n;
        // a last row tile that holds only the centre row p = N/2 (N/2 a multiple of 16) is a
        // dot product on the vector units: lane (column, kq) sums its k-steps' row-0 entries
        // times its coefficients, then the four kq lanes of the column add up
        const int RTm = 16 * (RT - 1) + 1 == rows ? RT - 1 : RT;
        if (RTm < RT) {
            double e = 0.0, o = 0.0;
#pragma unroll
            for (int s = 0; s < S; s++) {
                e = fma(tvT[(RTm * S + s) * 64 + 16 * kq], ce[s], e);
                o = fma(tvT[((RT + RTm) * S + s) * 64 + 16 * kq], co[s], o);
            }
            e += __shfl_xor(e, 16);
            o += __shfl_xor(o, 16);
            e += __shfl_xor(e, 32);
            o += __shfl_xor(o, 32);
            if (kq == 0 && p < d.P) Wm[2 * (N / 2)] = e + o;
        }
        for (int t = 0; t < RTm; t++) {
            // the row tile's table fragments first, so that no product waits on an LDS read
            double fe[S], fo[S];
#pragma unroll
            for (int s = 0; s < S; s++) {
                fe[s] = tvT[(t * S + s) * 64 + lane];
                fo[s] = tvT[((RT + t) * S + s) * 64 + lane];
            }
            rpcf_v4d ae = {0.0, 0.0, 0.0, 0.0}, ao = ae;
#pragma unroll
            for (int s = 0; s < S; s++) {
                ae = __builtin_amdgcn_mfma_f64_16x16x4f64(fe[s], ce[s], ae, 0, 0, 0);
                ao = __builtin_amdgcn_mfma_f64_16x16x4f64(fo[s], co[s], ao, 0, 0, 0);
            }
            if (p < d.P) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = 16 * t + kq + 4 * r;
                    if (row < rows) {
                        Wm[2 * row] = ae[r] + ao[r];
                        if (N - row != row) Wm[2 * (N - row)] = ae[r] - ao[r];
                    }
                }
            }
        }
    }
}

// y_q = M x_q for the two vectors of a slot with every thread of the block: G = nthr / n
// groups of n threads (thread t = g*n + r owns row r), group g takes the columns
// j = g, g + G, g + 2G, .. (fused multiply-adds, j ascending), so that the block reads the
// column-major matrix (M[j*n + r] = M[r][j]) as one flat, coalesced stream; the G partial
// sums meet in part[2][kSolveThreads] in group order.  M complex (double2) or real
// (double).  The rows land in threads tid < n; every thread of the block calls it (it
// synchronises).
constexpr int kSolveThreads = 576;
// matrix loads per thread in flight in the solve's GEMVs: with 4 the solve needs 80 registers
// and no spills, six waves per SIMD; at 128x129x128 the parity-block solve took 282 us
// against 299 with 8 (its fixed phases weigh more), the full-inverse solve the same 474 us.
// The summation order is the same for any batch.
constexpr int kGemvBatch = 4;
__device__ __forceinline__ void gemv_acc(double2 m, double2 u0, double2 u1, double* a) {
    a[0] = fma(m.x, u0.x, fma(-m.y, u0.y, a[0]));
    a[1] = fma(m.x, u0.y, fma(m.y, u0.x, a[1]));
    a[2] = fma(m.x, u1.x, fma(-m.y, u1.y, a[2]));
    a[3] = fma(m.x, u1.y, fma(m.y, u1.x, a[3]));
}
__device__ __forceinline__ void gemv_acc(double c, double2 u0, double2 u1, double* a) {
    a[0] = fma(c, u0.x, a[0]);
    a[1] = fma(c, u0.y, a[1]);
    a[2] = fma(c, u1.x, a[2]);
    a[3] = fma(c, u1.y, a[3]);
}
// Columns j < j0 are skipped (the caller's vectors vanish there: the boundary rows of the
// right-hand sides).  The last, partial batch of B columns loads clamped addresses and
// masks the vector by a multiply, so that its loads stay unconditional and in flight together.
template <typename MT>
__device__ __forceinline__ void gemv_grouped(const MT* __restrict__ M, const double2* x0, const double2* x1, int n,
                                             double2* part, int tid, int nthr, double2* y0, double2* y1, int j0 = 0) {
    constexpr int B = kGemvBatch;
    const int G = nthr / n;
    const int g = tid / n, r = tid - g * n;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < G) {
        int j = j0 + g;
        for (; j + G * (B - 1) < n; j += G * B) {
            MT m[B];
#pragma unroll
            for (int u = 0; u < B; u++) m[u] = M[(long)(j + G * u) * n + r];
#pragma unroll
            for (int u = 0; u < B; u++) gemv_acc(m[u], x0[j + G * u], x1[j + G * u], a);
        }
        if (j < n) {
            MT m[B];
#pragma unroll
            for (int u = 0; u < B; u++) m[u] = M[(long)min(j + G * u, n - 1) * n + r];
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int jj = min(j + G * u, n - 1);
                const double k = j + G * u < n ? 1.0 : 0.0;
                const double2 u0 = x0[jj], u1 = x1[jj];
                gemv_acc(m[u], make_double2(u0.x * k, u0.y * k), make_double2(u1.x * k, u1.y * k), a);
            }
        }
        part[tid] = make_double2(a[0], a[1]);
        part[kSolveThreads + tid] = make_double2(a[2], a[3]);
    }
    __syncthreads();
    if (tid < n) {
        double2 s0 = part[tid], s1 = part[kSolveThreads + tid];
        for (int q = 1; q < G; q++) {
            s0 = zadd(s0, part[q * n + tid]);
            s1 = zadd(s1, part[kSolveThreads + q * n + tid]);
        }
        *y0 = s0;
        *y1 = s1;
    }
}

// A complex BF16 pair (re in the low half) as a double2, exactly.
__device__ __forceinline__ double2 bf16x2_to_double2(unsigned w) {
    return make_double2((double)__uint_as_float(w << 16), (double)__uint_as_float(w & 0xffff0000u));
}

// The same with a rows x cols matrix (column-major, M[c*rows + r]) whose rows r < split
// take the vectors xa and the others xb; columns from j0.  With CORR the remainders E (same
// layout, BF16 pairs) multiply the other vectors (rows r < split xb, the others xa), and
// each column adds M x + E x' in that order.
template <bool CORR>
__device__ __forceinline__ void gemv_sel(const double2* __restrict__ M, const unsigned* __restrict__ E,
                                         const double2* xa0, const double2* xa1, const double2* xb0,
                                         const double2* xb1, int rows, int cols, int split, double2* part, int tid,
                                         int nthr, double2* y0, double2* y1, int j0) {
    constexpr int B = kGemvBatch;
    const int G = nthr / rows;
    const int g = tid / rows, r = tid - g * rows;
    const double2* x0 = r < split ? xa0 : xb0;
    const double2* x1 = r < split ? xa1 : xb1;
    const double2* o0 = r < split ? xb0 : xa0;
    const double2* o1 = r < split ? xb1 : xa1;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < G) {
        int j = j0 + g;
        for (; j + G * (B - 1) < cols; j += G * B) {
            double2 m[B];
            unsigned e[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                m[u] = M[(long)(j + G * u) * rows + r];
                if (CORR) e[u] = E[(long)(j + G * u) * rows + r];
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                gemv_acc(m[u], x0[j + G * u], x1[j + G * u], a);
                if (CORR) gemv_acc(bf16x2_to_double2(e[u]), o0[j + G * u], o1[j + G * u], a);
            }
        }
        if (j < cols) {
            double2 m[B];
            unsigned e[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                m[u] = M[(long)min(j + G * u, cols - 1) * rows + r];
                if (CORR) e[u] = E[(long)min(j + G * u, cols - 1) * rows + r];
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int jj = min(j + G * u, cols - 1);
                const double k = j + G * u < cols ? 1.0 : 0.0;
                const double2 u0 = x0[jj], u1 = x1[jj];
                gemv_acc(m[u], make_double2(u0.x * k, u0.y * k), make_double2(u1.x * k, u1.y * k), a);
                if (CORR) {
                    const double2 w0 = o0[jj], w1 = o1[jj];
                    gemv_acc(bf16x2_to_double2(e[u]), make_double2(w0.x * k, w0.y * k),
                             make_double2(w1.x * k, w1.y * k), a);
                }
            }
        }
        part[tid] = make_double2(a[0], a[1]);
        part[kSolveThreads + tid] = make_double2(a[2], a[3]);
    }
    __syncthreads();
    if (tid < rows) {
        double2 s0 = part[tid], s1 = part[kSolveThreads + tid];
        for (int q = 1; q < G; q++) {
            s0 = zadd(s0, part[q * rows + tid]);
            s1 = zadd(s1, part[kSolveThreads + q * rows + tid]);
        }
        *y0 = s0;
        *y1 = s1;
    }
}

// gemv_sel (no remainders) with the matrix as FP32 high parts Mf and BF16 low parts Ml
// (same layout); each entry enters as the exact double hi + lo.
__device__ __forceinline__ void gemv_sel_hilo(const float2* __restrict__ Mf, const unsigned* __restrict__ Ml,
                                              const double2* xa0, const double2* xa1, const double2* xb0,
                                              const double2* xb1, int rows, int cols, int split, double2* part,
                                              int tid, int nthr, double2* y0, double2* y1, int j0) {
    constexpr int B = kGemvBatch;
    const int G = nthr / rows;
    const int g = tid / rows, r = tid - g * rows;
    const double2* x0 = r < split ? xa0 : xb0;
    const double2* x1 = r < split ? xa1 : xb1;
    auto entry = [](float2 h, unsigned l) {
        const double2 lo = bf16x2_to_double2(l);
        return make_double2((double)h.x + lo.x, (double)h.y + lo.y);
    };
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < G) {
        int j = j0 + g;
        for (; j + G * (B - 1) < cols; j += G * B) {
            float2 h[B];
            unsigned l[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                h[u] = Mf[(long)(j + G * u) * rows + r];
                l[u] = Ml[(long)(j + G * u) * rows + r];
            }
#pragma unroll
            for (int u = 0; u < B; u++) gemv_acc(entry(h[u], l[u]), x0[j + G * u], x1[j + G * u], a);
        }
        if (j < cols) {
            float2 h[B];
            unsigned l[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                h[u] = Mf[(long)min(j + G * u, cols - 1) * rows + r];
                l[u] = Ml[(long)min(j + G * u, cols - 1) * rows + r];
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int jj = min(j + G * u, cols - 1);
                const double k = j + G * u < cols ? 1.0 : 0.0;
                const double2 u0 = x0[jj], u1 = x1[jj];
                gemv_acc(entry(h[u], l[u]), make_double2(u0.x * k, u0.y * k), make_double2(u1.x * k, u1.y * k), a);
            }
        }
        part[tid] = make_double2(a[0], a[1]);
        part[kSolveThreads + tid] = make_double2(a[2], a[3]);
    }
    __syncthreads();
    if (tid < rows) {
        double2 s0 = part[tid], s1 = part[kSolveThreads + tid];
        for (int q = 1; q < G; q++) {
            s0 = zadd(s0, part[q * rows + tid]);
            s1 = zadd(s1, part[kSolveThreads + q * rows + tid]);
        }
        *y0 = s0;
        *y1 = s1;
    }
}

// -------------------------------------------------------------------------------------
// Parity blocks of the inverses of one slot (block (matrix slot, equation)).  With an even
// base profile the collocation system A of either equation commutes with the wall reflection:
// A R = Q A, (R c)_j = (-1)^j c_j on the Chebyshev coefficients and Q the signed pairing of
// the rows (pairs[p] = (i, i', s): the rows of z_p and z_{N-p}, the two wall rows of each
// condition, s = -1 for the slope rows).  Hence M = A^-1 has M[j][i'] = (-1)^j s M[j][i]:
// the even coefficients of x = M b depend on b_i + s b_i' only and the odd ones on
// b_i - s b_i', and
//   x_j = sum_p H[j][p] (b_i + (-1)^j s b_i')   (a centre row i = i' enters once, and only
//                                                 for even j)
// with H[j][p] = (M[j][i] + (-1)^j s M[j][i'])/2 -- half the bytes of M.  Computed from the
// full inverse, so the blocks agree with it to its own (round-off) asymmetry.  Layout of a
// block: np columns (pairs) of nz rows, the even j = 0, 2, .. first, then the odd j.
//
// The remainders E[j][p] = (M[j][i] - (-1)^j s M[j][i'])/2 (for the centre row: M[j][i] at
// odd j, 0 at even j) restore the full inverse: H + E = M[j][i], H - E = (-1)^j s M[j][i'],
// so  x_j = sum_p H[j][p] (b_i + (-1)^j s b_i') + E[j][p] (b_i - (-1)^j s b_i')  is M b
// exactly.  E is the inverse's round-off asymmetry (~1e-14 of its largest entry at
// 128x129x128, the difference of two nearly equal doubles, so computed exactly) and is
// stored as BF16: its 2^-9 relative rounding is below the FP64 rounding of M itself.
// Dropping E moves a v solve by up to ~1e-12 (relative; its y-derivatives, which velocity
// recovery takes, more), but an eta solve and its first two y-derivatives by <= 4e-15 (the
// eta system is well conditioned), so only the v blocks keep remainders (same layout as H):
// the solves read (2 x 16 + 4)/64 of the full inverses' bytes and keep their results.
//
// The mixed form (bar 1e-6) solves with the blocks alone, the eta blocks split into an FP32
// high part and a BF16 low part (hi + lo = H to ~2^-33 relative, 12 of 16 bytes): the eta
// system is well conditioned, so its solve stays within ~1e-10 of the FP64 blocks'; the v
// system (condition ~5e7 at 128x129x128) keeps FP64 blocks.
// minve: the v remainders (FP64 form, minvf null) or the eta low parts (mixed form, with
// minvf the eta high parts); both may be null.
// -------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned bf16_bits(double x) {
    unsigned u = __float_as_uint((float)x);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return u >> 16;
}
__global__ void k_parity_blocks(const double2* __restrict__ minv, double2* __restrict__ minvh,
                                unsigned* __restrict__ minve, float2* __restrict__ minvf,
                                const int4* __restrict__ pairs, int np, Dims d) {
    const int slot = blockIdx.x, eq = blockIdx.y, n = d.nz, ne = d.N / 2 + 1;
    const double2* M = minv + ((long)eq * d.nslot + slot) * n * n;  // M[j][i] at i*n + j
    // the mixed form keeps no FP64 eta blocks (its solves read the split parts)
    double2* Hm = minvf && eq == 1 ? nullptr : minvh + ((long)eq * d.nslot + slot) * (long)n * np;
    unsigned* Em = minve && !minvf && eq == 0 ? minve + (long)slot * n * np : nullptr;  // v remainders
    unsigned* Lm = minve && minvf && eq == 1 ? minve + (long)slot * n * np : nullptr;   // eta low parts
    float2* Fm = minvf && eq == 1 ? minvf + (long)slot * n * np : nullptr;               // eta high parts
    const int4* pr = pairs + eq * np;
    for (long e = threadIdx.x; e < (long)n * np; e += blockDim.x) {
        const int p = (int)(e / n), r = (int)(e - (long)p * n);
        const int4 q = pr[p];
        const bool even = r < ne;
        const int j = even ? 2 * r : 2 * (r - ne) + 1;
        double2 v, w;
        if (q.x == q.y) {
            const double2 a = M[(long)q.x * n + j];
            v = even ? a : make_double2(0.0, 0.0);
            w = even ? make_double2(0.0, 0.0) : a;
        } else {
            const double2 a = M[(long)q.x * n + j], b = M[(long)q.y * n + j];
            const double sg = even ? (double)q.z : -(double)q.z;
            v = make_double2(0.5 * (a.x + sg * b.x), 0.5 * (a.y + sg * b.y));
            w = make_double2(0.5 * (a.x - sg * b.x), 0.5 * (a.y - sg * b.y));
        }
        if (Hm) Hm[e] = v;
        if (Em) Em[e] = bf16_bits(w.x) | (bf16_bits(w.y) << 16);
        if (Fm) {
            const float2 hi = make_float2((float)v.x, (float)v.y);
            Fm[e] = hi;
            Lm[e] = bf16_bits(v.x - (double)hi.x) | (bf16_bits(v.y - (double)hi.y) << 16);
        }
    }
}

// y_q = T0 x_q for the two vectors of a slot from the rows p <= N/2 of T0 (column-major,
// T0t[j*nz + p] = T_j(z_p)) and the reflection parity T_j(z_{N-p}) = (-1)^j T_j(z_p):
// E_p, O_p = the sums over the even and odd columns, y[p] = E_p + O_p, y[N-p] = E_p - O_p.
// G (even) groups of N/2+1 threads; group g takes the columns j = g, g + G, .., all of the
// parity of g.  The results land in y0[], y1[] (LDS, nz entries each).  Every thread of the
// block calls it (it synchronises).
__device__ __forceinline__ void gemv_t0_reflect(const double* __restrict__ T0t, const double2* x0, const double2* x1,
                                                int n, int N, double2* part, int tid, int nthr, double2* y0,
                                                double2* y1) {
    constexpr int B = kGemvBatch;
    const int h = N / 2 + 1, G = (nthr / h) & ~1;
    const int g = tid / h, r = tid - g * h;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < G) {
        int j = g;
        for (; j + G * (B - 1) < n; j += G * B) {
            double m[B];
#pragma unroll
            for (int u = 0; u < B; u++) m[u] = T0t[(long)(j + G * u) * n + r];
#pragma unroll
            for (int u = 0; u < B; u++) gemv_acc(m[u], x0[j + G * u], x1[j + G * u], a);
        }
        for (; j < n; j += G) gemv_acc(T0t[(long)j * n + r], x0[j], x1[j], a);
        part[tid] = make_double2(a[0], a[1]);
        part[kSolveThreads + tid] = make_double2(a[2], a[3]);
    }
    __syncthreads();
    if (tid < h) {
        double2 e0 = part[tid], e1 = part[kSolveThreads + tid];
        double2 o0 = part[h + tid], o1 = part[kSolveThreads + h + tid];
        for (int q = 2; q < G; q += 2) {
            e0 = zadd(e0, part[q * h + tid]);
            e1 = zadd(e1, part[kSolveThreads + q * h + tid]);
            o0 = zadd(o0, part[(q + 1) * h + tid]);
            o1 = zadd(o1, part[kSolveThreads + (q + 1) * h + tid]);
        }
        y0[tid] = zadd(e0, o0);
        y1[tid] = zadd(e1, o1);
        if (N - tid != tid) {
            y0[N - tid] = zsub(e0, o0);
            y1[N - tid] = zsub(e1, o1);
        }
    }
}

// x = M b for the two right-hand sides rhs[0..1][nz] of a slot: with the full inverse M, or,
// for a member with an even base profile, with its parity blocks H (k_parity_blocks): the
// combined sides b_i + s b_i' and b_i - s b_i' per row pair, then one nz x np GEMV whose even
// rows take the first and odd rows the second (parity 2 and Em given: and the remainders E
// the other; Hf given: the blocks as FP32 high parts Hf and BF16 low parts Em).  b vanishes
// on the rows below j0 (an even count: whole row pairs), whose columns are skipped.  Thread
// r < nz returns x at coefficient *jrow.  Every thread of the block calls it (it synchronises).
__device__ __forceinline__ void solve_slot(const double2* M, const double2* Hm, const unsigned* Em, const float2* Hf,
                                           const int4* pairs, int np, int parity, double2 (*rhs)[256],
                                           double2 (*bc)[2][128], double2* part, int n, int N, int tid, int nthr,
                                           double2* y0, double2* y1, int* jrow, int j0) {
    if (!parity) {
        gemv_grouped(M, rhs[0], rhs[1], n, part, tid, nthr, y0, y1, j0);
        *jrow = tid;
        return;
    }
    if (tid < np) {
        const int4 q = pairs[tid];
        for (int v = 0; v < 2; v++) {
            const double2 a = rhs[v][q.x], b = rhs[v][q.y];
            if (q.x == q.y) {
                bc[v][0][tid] = a;
                bc[v][1][tid] = make_double2(0.0, 0.0);
            } else {
                const double s = (double)q.z;
                bc[v][0][tid] = make_double2(a.x + s * b.x, a.y + s * b.y);
                bc[v][1][tid] = make_double2(a.x - s * b.x, a.y - s * b.y);
            }
        }
    }
    __syncthreads();
    const int ne = N / 2 + 1;
    if (Hf)
        gemv_sel_hilo(Hf, Em, bc[0][0], bc[1][0], bc[0][1], bc[1][1], n, np, ne, part, tid, nthr, y0, y1, j0 / 2);
    else if (parity == 2 && Em)
        gemv_sel<true>(Hm, Em, bc[0][0], bc[1][0], bc[0][1], bc[1][1], n, np, ne, part, tid, nthr, y0, y1, j0 / 2);
    else
        gemv_sel<false>(Hm, Em, bc[0][0], bc[1][0], bc[0][1], bc[1][1], n, np, ne, part, tid, nthr, y0, y1, j0 / 2);
    *jrow = tid < ne ? 2 * tid : 2 * (tid - ne) + 1;
}

// -------------------------------------------------------------------------------------
// Crank-Nicolson/AB2 update of the one or two pencils of a matrix slot from the operator
// products W (k_lin_tv): v right-hand side (_get_linear_v + AB2, source/linear.cpp:86-125,
// source/rhs.cpp:205-215), v solve (solveEq, source/solver.cpp:107-123), T0 v^{n+1} (one
// pass over T0), eta right-hand side (_get_linear_omega_y, source/linear.cpp:127-171), eta
// solve.  Block 0 runs the mean flow of slot (0,0) (_get_linear_u0/_w0,
// source/linear.cpp:173-205) and the constant-flow-rate correction.  kSolveThreads
// threads: thread r < nz assembles row r, the products (v solve, T0 v^{n+1}, eta solve) use every thread
// (gemv_grouped); matrix slot blockIdx.x, member blockIdx.y.
//   v rows r = 4..N, p = r-2:  (T4q v + k2 T2 v - k0 T0 v) + i (a2 T2 v + a0 T0 v)
//   eta rows r = 2..N, p = r-1: (T2q eta + c0 T0 eta) + i ci T0 eta + i e T0 (v^n + v^{n+1})
// -------------------------------------------------------------------------------------
__global__ __launch_bounds__(kSolveThreads, 6) void k_lin_solve(double2* __restrict__ SV, double2* __restrict__ SE, long sS,
                                                   const double2* __restrict__ NL, long sNL,
                                                   const double2* __restrict__ MINV, long sM,
                                                   const double2* __restrict__ MINVH, long sMH,
                                                   const unsigned* __restrict__ MINVE, long sME,
                                                   const float2* __restrict__ MINVF,
                                                   const int4* __restrict__ pairs, int np,
                                                   const double2* __restrict__ W, long sWl, Dims d, Members mb,
                                                   const double* __restrict__ T0t, const double* __restrict__ U,
                                                   const double* __restrict__ dU, const double* __restrict__ ddU, long sU,
                                                   const double* __restrict__ qw) {
    __shared__ double2 rhs[2][256], vnew[2][256], part[2][kSolveThreads];
    __shared__ double2 bc[2][2][128];  // parity form: [vector][b_i + s b_i', b_i - s b_i'][pair]; T0 v^{n+1}
    __shared__ double2 s_f;
    LinState st;
    LinParams lp;
    {
        const int ms = mb.ids[blockIdx.y];
        const MemberParams& mp = mb.mp[ms];
        const int ph = mb.phase[ms];
        const long PS = (long)d.P * d.nz;
        const double2* h = NL + ms * sNL;
        st.nlv = h + ph * PS;
        st.nlvp = h + (1 - ph) * PS;
        st.nle = h + (2 + ph) * PS;
        st.nlep = h + (3 - ph) * PS;
        st.sv = SV + ms * sS;
        st.se = SE + ms * sS;
        st.minv = MINV + ms * sM;
        st.minvh = MINVH + ms * sMH;
        st.minve = MINVE ? MINVE + ms * sME : nullptr;
        st.minvf = MINVF ? MINVF + ms * sME : nullptr;
        st.parity = mp.parity;
        st.qw_target_re = mp.q_re;
        st.qw_target_im = mp.q_im;
        st.flow_control = mp.flow_control;
        lp.aphi = mp.aphi;
        lp.beta = mp.beta;
        lp.Re = mp.Re;
        lp.dt = mp.dt;
        W += ms * sWl;
        U += ms * sU;
        dU += ms * sU;
        ddU += ms * sU;
    }
    const int slot = blockIdx.x;
    const int kyh = slot / d.hnx, kx = slot - kyh * d.hnx;
    const int n = d.nz, r = threadIdx.x;
    const double Re = lp.Re, dt = lp.dt;
    const long nn = (long)n * n, PS = (long)d.P * n;
    const double2* Mv = st.minv + (long)slot * nn;
    const double2* Mo = st.minv + ((long)d.nslot + slot) * nn;
    const double2 *W0 = W, *W1 = W + PS, *W2 = W + 2 * PS, *W3 = W + 3 * PS, *W4 = W + 4 * PS;

    if (slot == 0) {
        // u0: rhs_r = (T0 u0 + dt/2/Re T2 u0)_{r-1} - dt (1.5 Lx0 - 0.5 Lx0_prev)_{r-1}, rows 0,1 = 0;
        // w0 likewise with eta's products and Ly0.  Second vector: unit uniform forcing.
        const double h = dt * 0.5 / Re;
        for (int e = 0; e < 2; e++) {
            double2* s = e ? st.se : st.sv;
            const double2* l0 = e ? st.nle : st.nlv;
            const double2* l0p = e ? st.nlep : st.nlvp;
            if (r < n) {
                double2 v = make_double2(0.0, 0.0);
                if (r >= 2) {
                    const double2 acc = e ? zadd(W3[r - 1], W4[r - 1]) : zadd(W0[r - 1], zmulr(W1[r - 1], h));
                    v = zsub(acc, zmulr(zsub(zmulr(l0[r - 1], 1.5), zmulr(l0p[r - 1], 0.5)), dt));
                }
                rhs[0][r] = v;
                rhs[1][r] = make_double2(r >= 2 ? dt : 0.0, 0.0);
            }
            __syncthreads();
            double2 x = make_double2(0.0, 0.0), uh = make_double2(0.0, 0.0);
            gemv_grouped(e ? Mo : Mv, rhs[0], rhs[1], n, &part[0][0], r, blockDim.x, &x, &uh, 2);
            if (r < n) {
                vnew[0][r] = x;
                vnew[1][r] = uh;
            }
            __syncthreads();
            if (e == 0 && st.flow_control == 1) {
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

    const bool two = !(kyh == 0 || kyh == d.ny / 2);
    const long c0 = (long)kyh * d.hnx + kx;
    const long c1 = two ? (long)(d.ny - kyh) * d.hnx + kx : c0;
    const long po0 = c0 * n, po1 = c1 * n;
    double alpha, beta0, beta1, a_unused;
    ab_of(d, lp, kx, kyh, &alpha, &beta0);
    ab_of(d, lp, kx, two ? d.ny - kyh : kyh, &a_unused, &beta1);
    const double kmn = alpha * alpha + beta0 * beta0;
    if (r < n) {
        double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
        if (r >= 4) {
            const int p = r - 2;
            const double k2 = (1 - kmn * dt / Re), k0 = kmn * (1 - kmn * dt * 0.5 / Re);
            const double i2 = -alpha * dt * 0.5 * U[p];
            const double i0 = (kmn * alpha * dt * 0.5 * U[p] + alpha * dt * 0.5 * ddU[p]);
            const double2 s0 = W0[po0 + p], s2 = W1[po0 + p], s4 = W2[po0 + p];
            const double2 t0 = W0[po1 + p], t2 = W1[po1 + p], t4 = W2[po1 + p];
            a0 = make_double2(s4.x + k2 * s2.x - k0 * s0.x - (i2 * s2.y + i0 * s0.y),
                              s4.y + k2 * s2.y - k0 * s0.y + (i2 * s2.x + i0 * s0.x));
            a1 = make_double2(t4.x + k2 * t2.x - k0 * t0.x - (i2 * t2.y + i0 * t0.y),
                              t4.y + k2 * t2.y - k0 * t0.y + (i2 * t2.x + i0 * t0.x));
            a0 = zadd(a0, zmulr(zsub(zmulr(st.nlv[po0 + p], 1.5), zmulr(st.nlvp[po0 + p], 0.5)), dt));
            a1 = zadd(a1, zmulr(zsub(zmulr(st.nlv[po1 + p], 1.5), zmulr(st.nlvp[po1 + p], 0.5)), dt));
        }
        rhs[0][r] = a0;
        rhs[1][r] = a1;
    }
    __syncthreads();
    const long nh = (long)n * np;
    {
        double2 y0, y1;
        int j;
        const unsigned* Ev = st.minve && !st.minvf ? st.minve + (long)slot * nh : nullptr;
        solve_slot(Mv, st.minvh + (long)slot * nh, Ev, nullptr, pairs, np, st.parity, rhs, bc, &part[0][0], n, d.N,
                   r, blockDim.x, &y0, &y1, &j, 4);
        if (r < n) vnew[0][j] = y0, vnew[1][j] = y1;
    }
    __syncthreads();
    // T0 v^{n+1} at the points p -> tv0[p], tv1[p] (in bc, free until the eta solve)
    double2 *tv0 = &bc[0][0][0], *tv1 = &bc[1][0][0];
    gemv_t0_reflect(T0t, vnew[0], vnew[1], n, d.N, &part[0][0], r, blockDim.x, tv0, tv1);
    __syncthreads();
    if (r < n) {
        double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
        if (r >= 2) {
            const int p = r - 1;
            const double w0r = tv0[p].x, w0i = tv0[p].y, w1r = tv1[p].x, w1i = tv1[p].y;
            const double c0k = (1 - kmn * dt * 0.5 / Re), ci = -alpha * dt * 0.5 * U[p];
            const double e0 = -0.5 * beta0 * dU[p] * dt, e1 = -0.5 * beta1 * dU[p] * dt;
            const double2 s0 = W3[po0 + p], s2 = W4[po0 + p], sv0 = W0[po0 + p];
            const double2 t0 = W3[po1 + p], t2 = W4[po1 + p], sv1 = W0[po1 + p];
            const double2 u0 = make_double2(sv0.x + w0r, sv0.y + w0i), u1 = make_double2(sv1.x + w1r, sv1.y + w1i);
            a0 = make_double2(s2.x + c0k * s0.x - ci * s0.y - e0 * u0.y, s2.y + c0k * s0.y + ci * s0.x + e0 * u0.x);
            a1 = make_double2(t2.x + c0k * t0.x - ci * t0.y - e1 * u1.y, t2.y + c0k * t0.y + ci * t0.x + e1 * u1.x);
            a0 = zadd(a0, zmulr(zsub(zmulr(st.nle[po0 + p], 1.5), zmulr(st.nlep[po0 + p], 0.5)), dt));
            a1 = zadd(a1, zmulr(zsub(zmulr(st.nle[po1 + p], 1.5), zmulr(st.nlep[po1 + p], 0.5)), dt));
        }
        rhs[0][r] = a0;
        rhs[1][r] = a1;
    }
    __syncthreads();
    double2 e0, e1;
    int je;
    // eta: the FP64 blocks without remainders, or (mixed form) their FP32 + BF16 split
    const bool split = st.minvf && st.parity == 1;
    solve_slot(Mo, split ? nullptr : st.minvh + ((long)d.nslot + slot) * nh, split ? st.minve + (long)slot * nh : nullptr,
               split ? st.minvf + (long)slot * nh : nullptr, pairs + np, np, st.parity, rhs, bc, &part[0][0], n, d.N,
               r, blockDim.x, &e0, &e1, &je, 2);
    if (r < n) {
        st.se[po0 + je] = e0;
        st.sv[po0 + r] = vnew[0][r];
        if (two) {
            st.se[po1 + je] = e1;
            st.sv[po1 + r] = vnew[1][r];
        }
    }
}
