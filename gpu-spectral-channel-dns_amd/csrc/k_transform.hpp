// k_transform.hpp -- HIP kernels of the spectral transforms and the nonlinear term.
//
// Pipeline of one nonlinear evaluation (the reference's getNonlinear,
// source/nonlinear.cu:25-108), FP64, all intermediates in HBM:
//
//  F[6][P][nz]       spectral u,v,w,ox,oy,oz pencils (Chebyshev index contiguous)
//   k_cheb_s2p    -> Y[6][hnx][pz][ny]     wall-normal values, ky contiguous
//   k_span_inv    -> W[6][pz][my][nxh]     span-inverse FFT, kx contiguous
//   k_lamb_phys   -> Z2[3][pz][my][hnx]    x C2R, Lamb + Coriolis, x R2C, 1/(mx my)
//   k_span_fwd    -> Z1[3][hnx][pz][ny]    span-forward FFT + truncation to ny modes
//   k_p2s_nl      -> NLv, NLe [P][nz]      Chebyshev p2s, rhsNonlinear, unpadded s2p
//
// Every kernel reads one layout and writes the next with the contiguous index of the
// write side coalesced through an LDS tile; physical fields never reach HBM because the
// x transforms, the Lamb vector and the Coriolis term are fused in k_lamb_phys.
#pragma once
#include "fft_lds.hpp"
#include "rpcf_common.hpp"

#define RPCF_LDS(T) extern __shared__ __align__(16) unsigned char rpcf_lds_raw[]; T* lds = (T*)rpcf_lds_raw

// -------------------------------------------------------------------------------------
// Chebyshev spectral -> physical along the wall-normal direction, f(z_q) = sum_k a_k T_k(z_q).
// PAD=0: the nz points cos(pi p/N): cheby_pre_s2p_noPad + length-2N FFT of the even
//        extension (source/transform.cu:455-491, 716-733).
// PAD=1: the pz = 2N+1 points cos(pi q/(2N)), which the reference takes from a length-4N
//        FFT of the zero-padded even extension (source/transform.cu:416-453, 679-698).  Here
//        the even points q = 2m are the PAD=0 transform of the same coefficients, and the odd
//        points q = 2m+1 are g_m = sum_k a_k cos(pi k (2m+1)/(2N)) = FFT_2N(d)_m with
//        d_0 = a_0, d_k = a_k/2 e^{-i pi k/(2N)}, d_{2N-k} = a_k/2 e^{+i pi k/(2N)} (0<k<N),
//        d_N = 0 -- two length-2N FFTs in a 2N-point LDS tile instead of one of length 4N.
// tw4: exp(-2 pi i k/(4N)), k < 4N.  Tile: TB pencils (kx fixed); output Y[f][kx][q][ky].
// -------------------------------------------------------------------------------------
__device__ __forceinline__ void s2p_even_ext(double2* s, double2 a, int k, int N) {
    const int L = 2 * N;
    if (k < N) {
        const double2 t = cscale(a, 0.5);
        if (k == 0) s[0] = cscale(t, 2.0);
        else { s[k] = t; s[L - k] = t; }
    } else {
        s[N] = a;
    }
}

__device__ __forceinline__ void s2p_odd_ext(double2* s, double2 a, int k, int N, const double2* tw4) {
    const int L = 2 * N;
    if (k == 0) s[0] = a;
    else if (k < N) {
        const double2 t = cscale(a, 0.5), w = tw4[k];
        s[k] = cmul(t, w);
        s[L - k] = cmul(t, make_double2(w.x, -w.y));
    } else {
        s[N] = make_double2(0.0, 0.0);
    }
}

template <int PAD, int FN = 0>  // FN: compile-time 2N or 0
__global__ __launch_bounds__(256) void k_cheb_s2p(const double2* __restrict__ F, long fstride, double2* __restrict__ Y,
                                                  long ystride, Dims d, FftPlan pl, const double2* __restrict__ tw4,
                                                  int TB, int ld) {
    RPCF_LDS(double2);
    const int f = blockIdx.z, kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0);
    const int N = d.N, nz = d.nz, nout = PAD ? d.pz : d.nz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const double2* in = F + f * fstride + ((long)ky0 * d.hnx + kx) * nz;
    double2* out = Y + f * ystride + (long)kx * nout * d.ny + ky0;
    for (int half = 0; half < (PAD ? 2 : 1); half++) {
        for (int i = tid; i < nb * nz; i += nthr) {
            const int b = i / nz, k = i - b * nz;
            const double2 a = in[(long)b * d.hnx * nz + k];
            if (half == 0) s2p_even_ext(lds + b * ld, a, k, N);
            else s2p_odd_ext(lds + b * ld, a, k, N, tw4);
        }
        __syncthreads();
        fft_any<FN, -1>(lds, ld, nb, pl, tid, nthr);
        // even half: points q = m (PAD=0) or q = 2m (PAD=1), m = 0..N; odd half: q = 2m+1, m < N
        const int cnt = half ? N : N + 1;
        for (int i = tid; i < cnt * nb; i += nthr) {
            const int m = i / nb, b = i - m * nb;
            const int q = PAD ? 2 * m + half : m;
            out[(long)q * d.ny + b] = lds[b * ld + m];
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------------------------
// Span inverse FFT with the 3/2 zero padding: the span C2C half of the reference's 2D
// Z2D (source/transform.cu:176) after transpose_backward + setZeros
// (source/transpose.cu:358-381, source/transform.cu:355-386).  Tile: TK kx columns of one
// plane.  Y[f][kx][q][ky] -> W[f][q][s][kx] for kx < nxh (the x-Nyquist column is zero).
// -------------------------------------------------------------------------------------
template <int FN = 0>
__global__ __launch_bounds__(256) void k_span_inv(const double2* __restrict__ Y, long ystride, double2* __restrict__ W, long wstride, Dims d,
                           int planes, FftPlan pl, int TK, int ld) {
    RPCF_LDS(double2);
    const int f = blockIdx.z, q = blockIdx.y, kx0 = blockIdx.x * TK;
    const int nk = min(TK, d.nxh - kx0);
    const int my = d.my, tid = threadIdx.x, nthr = blockDim.x;
    const double2* in = Y + f * ystride;
    for (int i = tid; i < nk * my; i += nthr) {
        const int kxi = i / my, s = i - kxi * my;
        const int ky = ky_of_span(d, s);
        lds[kxi * ld + s] = ky >= 0 ? in[((long)(kx0 + kxi) * planes + q) * d.ny + ky] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_any<FN, +1>(lds, ld, nk, pl, tid, nthr);
    double2* out = W + f * wstride;
    for (int i = tid; i < my * nk; i += nthr) {
        const int s = i / nk, kxi = i - s * nk;
        out[((long)q * my + s) * d.nxh + kx0 + kxi] = lds[kxi * ld + s];
    }
}

// Load two half-spectrum rows (kx < nxh) as one packed length-mx spectrum
// Z = X_A + i X_B, so that one inverse complex FFT returns row A in .x and row B in .y.
// The kx = 0 entries contribute their real parts, as a C2R transform does.
__device__ __forceinline__ double2 c2r_pack(const double2* rA, const double2* rB, int k, int mx, int nxh) {
    if (k < nxh) {
        double2 a = rA[k], b = rB[k];
        if (k == 0) { a.y = 0.0; b.y = 0.0; }
        return make_double2(a.x - b.y, a.y + b.x);
    }
    if (k > mx - nxh) {
        const int kk = mx - k;
        double2 a = rA[kk], b = rB[kk];
        return make_double2(a.x + b.y, -a.y + b.x);
    }
    return make_double2(0.0, 0.0);
}

// Split the FFT Z of a packed real pair z = a + i b into the half spectra of a and b.
__device__ __forceinline__ void r2c_unpack(const double2* z, int k, int mx, double fac, double2* xa, double2* xb) {
    const double2 zk = z[k], zm = z[k == 0 ? 0 : mx - k];
    *xa = make_double2((zk.x + zm.x) * 0.5 * fac, (zk.y - zm.y) * 0.5 * fac);
    *xb = make_double2((zk.y + zm.y) * 0.5 * fac, -(zk.x - zm.x) * 0.5 * fac);
}

// -------------------------------------------------------------------------------------
// Fused physical-space stage: x C2R of the six fields, Lamb vector + Coriolis term
// (computeLambDevice, source/nonlinear.cu:409-428, base flow at padded point q as in
// source/nonlinear.cu:400-403), x R2C of the three Lamb components and the
// 1/(mx my) normalisation of source/transform.cu:267.  One block: one plane q, RP
// pairs of span rows.  W[6][pz][my][nxh] -> Z2[3][pz][my][hnx].
// -------------------------------------------------------------------------------------
template <int FN = 0>
__global__ __launch_bounds__(256) void k_lamb_phys(const double2* __restrict__ W, long wstride, double2* __restrict__ Z2, long zstride,
                            const double* __restrict__ meanU, double Ro, double fac, Dims d, FftPlan pl, int RP,
                            int ld) {
    RPCF_LDS(double2);
    const int q = blockIdx.y, rp0 = blockIdx.x * RP;
    const int nrp = min(RP, d.my / 2 - rp0);
    const int mx = d.mx, my = d.my, nxh = d.nxh, tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < 6 * nrp * mx; i += nthr) {
        const int seq = i / mx, k = i - seq * mx;
        const int f = seq / nrp, ip = seq - f * nrp;
        const double2* rA = W + f * wstride + ((long)q * my + 2 * (rp0 + ip)) * nxh;
        lds[seq * ld + k] = c2r_pack(rA, rA + nxh, k, mx, nxh);
    }
    __syncthreads();
    fft_any<FN, +1>(lds, ld, 6 * nrp, pl, tid, nthr);
    const double mU = meanU[q];
    for (int i = tid; i < nrp * mx; i += nthr) {
        const int ip = i / mx, x = i - ip * mx;
        const double2 u = lds[(0 * nrp + ip) * ld + x], v = lds[(1 * nrp + ip) * ld + x];
        const double2 w = lds[(2 * nrp + ip) * ld + x], ox = lds[(3 * nrp + ip) * ld + x];
        const double2 oy = lds[(4 * nrp + ip) * ld + x], oz = lds[(5 * nrp + ip) * ld + x];
        const double lxa = oz.x * v.x - oy.x * w.x - Ro * w.x, lxb = oz.y * v.y - oy.y * w.y - Ro * w.y;
        const double lya = ox.x * w.x - oz.x * u.x, lyb = ox.y * w.y - oz.y * u.y;
        const double lza = oy.x * u.x - ox.x * v.x + Ro * (u.x + mU), lzb = oy.y * u.y - ox.y * v.y + Ro * (u.y + mU);
        lds[(0 * nrp + ip) * ld + x] = make_double2(lxa, lxb);
        lds[(1 * nrp + ip) * ld + x] = make_double2(lya, lyb);
        lds[(2 * nrp + ip) * ld + x] = make_double2(lza, lzb);
    }
    __syncthreads();
    fft_any<FN, -1>(lds, ld, 3 * nrp, pl, tid, nthr);
    const int hnx = d.hnx;
    for (int i = tid; i < 3 * nrp * hnx; i += nthr) {
        const int seq = i / hnx, k = i - seq * hnx;
        const int g = seq / nrp, ip = seq - g * nrp;
        double2 xa, xb;
        r2c_unpack(lds + seq * ld, k, mx, fac, &xa, &xb);
        double2* o = Z2 + g * zstride + ((long)q * my + 2 * (rp0 + ip)) * hnx;
        o[k] = xa;
        o[hnx + k] = xb;
    }
}

// x C2R of one field (no Lamb): W[q][s][nxh] -> phys[q][s][mx] (rpcf_spec_to_phys).
__global__ __launch_bounds__(256) void k_x_c2r(const double2* __restrict__ W, double* __restrict__ phys, Dims d, FftPlan pl, int RP, int ld) {
    RPCF_LDS(double2);
    const int q = blockIdx.y, rp0 = blockIdx.x * RP;
    const int nrp = min(RP, d.my / 2 - rp0);
    const int mx = d.mx, my = d.my, nxh = d.nxh, tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < nrp * mx; i += nthr) {
        const int ip = i / mx, k = i - ip * mx;
        const double2* rA = W + ((long)q * my + 2 * (rp0 + ip)) * nxh;
        lds[ip * ld + k] = c2r_pack(rA, rA + nxh, k, mx, nxh);
    }
    __syncthreads();
    fft_run<+1>(lds, ld, nrp, pl, tid, nthr);
    for (int i = tid; i < nrp * mx; i += nthr) {
        const int ip = i / mx, x = i - ip * mx;
        double* rowA = phys + ((long)q * my + 2 * (rp0 + ip)) * mx;
        const double2 z = lds[ip * ld + x];
        rowA[x] = z.x;
        rowA[mx + x] = z.y;
    }
}

// x R2C of one field: phys[q][s][mx] -> Z2[q][s][hnx] with 1/(mx my) (rpcf_phys_to_spec).
__global__ __launch_bounds__(256) void k_x_r2c(const double* __restrict__ phys, double2* __restrict__ Z2, double fac, Dims d, FftPlan pl,
                        int RP, int ld) {
    RPCF_LDS(double2);
    const int q = blockIdx.y, rp0 = blockIdx.x * RP;
    const int nrp = min(RP, d.my / 2 - rp0);
    const int mx = d.mx, my = d.my, tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < nrp * mx; i += nthr) {
        const int ip = i / mx, x = i - ip * mx;
        const double* rowA = phys + ((long)q * my + 2 * (rp0 + ip)) * mx;
        lds[ip * ld + x] = make_double2(rowA[x], rowA[mx + x]);
    }
    __syncthreads();
    fft_run<-1>(lds, ld, nrp, pl, tid, nthr);
    const int hnx = d.hnx;
    for (int i = tid; i < nrp * hnx; i += nthr) {
        const int ip = i / hnx, k = i - ip * hnx;
        double2 xa, xb;
        r2c_unpack(lds + ip * ld, k, mx, fac, &xa, &xb);
        double2* o = Z2 + ((long)q * my + 2 * (rp0 + ip)) * hnx;
        o[k] = xa;
        o[hnx + k] = xb;
    }
}

// -------------------------------------------------------------------------------------
// Span forward FFT + truncation to the ny kept modes: the span C2C half of the
// reference's D2Z (source/transform.cu:240) and transpose_forward
// (source/transpose.cu:223-257).  Z2[g][q][s][kx] -> Z1[g][kx][q][ky].
// -------------------------------------------------------------------------------------
template <int FN = 0>
__global__ __launch_bounds__(256) void k_span_fwd(const double2* __restrict__ Z2, long zstride, double2* __restrict__ Z1, long ostride, Dims d,
                           int planes, FftPlan pl, int TK, int ld) {
    RPCF_LDS(double2);
    const int g = blockIdx.z, q = blockIdx.y, kx0 = blockIdx.x * TK;
    const int nk = min(TK, d.hnx - kx0);
    const int my = d.my, hnx = d.hnx, tid = threadIdx.x, nthr = blockDim.x;
    const double2* in = Z2 + g * zstride;
    for (int i = tid; i < my * nk; i += nthr) {
        const int s = i / nk, kxi = i - s * nk;
        lds[kxi * ld + s] = in[((long)q * my + s) * hnx + kx0 + kxi];
    }
    __syncthreads();
    fft_any<FN, -1>(lds, ld, nk, pl, tid, nthr);
    double2* out = Z1 + g * ostride;
    for (int i = tid; i < nk * d.ny; i += nthr) {
        const int kxi = i / d.ny, ky = i - kxi * d.ny;
        out[((long)(kx0 + kxi) * planes + q) * d.ny + ky] = lds[kxi * ld + span_of(d, ky)];
    }
}

// Load one field's wall-normal values (P points, P = pz or nz) of TB pencils as the
// even extension of length 2(P-1): cheby_pre_p2s (source/transform.cu:494-510).
__device__ __forceinline__ void load_p2s(double2* lds, int ld, const double2* Z1, int kx, int ky0, int nb, int P,
                                         int ny, int tid, int nthr) {
    const int L = 2 * (P - 1);
    for (int i = tid; i < P * nb; i += nthr) {
        const int q = i / nb, b = i - q * nb;
        const double2 v = Z1[((long)kx * P + q) * ny + ky0 + b];
        double2* s = lds + b * ld;
        s[q] = v;
        if (q >= 1 && q <= P - 2) s[L - q] = v;
    }
}

// Chebyshev physical -> spectral of one field (rpcf_phys_to_spec): cheby_p2s
// (source/transform.cu:538-658, post-processing 513-536).  Z1[kx][q][ky] -> F[p][k<nz].
__global__ __launch_bounds__(256) void k_cheb_p2s(const double2* __restrict__ Z1, double2* __restrict__ F, Dims d, int planes, FftPlan pl,
                           int TB, int ld) {
    RPCF_LDS(double2);
    const int kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0), nz = d.nz, tid = threadIdx.x, nthr = blockDim.x;
    load_p2s(lds, ld, Z1, kx, ky0, nb, planes, d.ny, tid, nthr);
    __syncthreads();
    fft_run<-1>(lds, ld, nb, pl, tid, nthr);
    const double factor = (1.0 / (planes - 1));
    for (int i = tid; i < nb * nz; i += nthr) {
        const int b = i / nz, k = i - b * nz;
        double2 t = cscale(lds[b * ld + k], factor);
        if (k == 0 || k == planes - 1) t = cscale(t, 0.5);
        F[((long)(ky0 + b) * d.hnx + kx) * nz + k] = t;
    }
}

// Chebyshev derivative of one coefficient: ddz_sm (source/RPCFKernels.cu:476-520),
// b_k = sum_{j=k+1,k+3,..<n} 2j a_j, b_0 halved.
__device__ __forceinline__ double2 ddz_at(const double2* a, int k, int n) {
    double re = 0.0, im = 0.0;
    for (int j = k + 1; j < n; j += 2) {
        const double dm = 2 * (double)j;
        re = re + a[j].x * dm;
        im = im + a[j].y * dm;
    }
    return k == 0 ? make_double2(re * 0.5, im * 0.5) : make_double2(re, im);
}

// -------------------------------------------------------------------------------------
// Forward Chebyshev transform of the three Lamb components, rhsNonlinearKernel
// (source/nonlinear.cu:146-242) and the unpadded s2p of N_v and N_eta
// (source/rhs.cpp:96-97).  Tile: TB pencils (kx fixed).  Z1[3][hnx][pz][ny] ->
// NLv, NLe [P][nz] (values at the nz collocation points; pencil (0,0) keeps L_x, L_y).
//
// The reference's p2s is a length-4N FFT of the even extension u of the pz values
// (cheby_pre_p2s/cheby_post_p2s, source/transform.cu:494-536).  Split by the parity of
// the sample index (one radix-2 decimation-in-time step):
//   F_k = E_k + w4^k H_k,  E = FFT_2N(even extension of f_0, f_2, .., f_2N),
//                          H = FFT_2N(f_1, f_3, .., f_{2N-1}, f_{2N-1}, .., f_1),
// w4 = exp(-2 pi i/(4N)); a_k = F_k/(2N), a_0 halved; only k < nz is kept.
// LDS: FFT tile TB*ld (ld >= 2N+1), then coefficient arrays Lx, Ly, Lz [TB][nz].
// -------------------------------------------------------------------------------------
template <int FN = 0>  // FN: compile-time 2N or 0
__global__ __launch_bounds__(256) void k_p2s_nl(const double2* __restrict__ Z1, long zstride, double2* __restrict__ NLv,
                                                double2* __restrict__ NLe, Dims d, FftPlan plN,
                                                const double2* __restrict__ tw4, double aphi, double beta, int TB,
                                                int ld) {
    RPCF_LDS(double2);
    const int kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0), nz = d.nz, N = d.N, ny = d.ny;
    const int L = 2 * N;
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2* fb = lds;
    double2* coef = lds + TB * ld;  // [3][TB][nz]
    const double factor = (1.0 / (d.pz - 1));
    for (int g = 0; g < 3; g++) {
        const double2* in = Z1 + g * zstride + (long)kx * d.pz * ny + ky0;
        for (int half = 0; half < 2; half++) {
            const int cnt = half ? N : N + 1;  // samples f_{2m+half}
            for (int i = tid; i < cnt * nb; i += nthr) {
                const int m = i / nb, b = i - m * nb;
                const double2 v = in[(long)(2 * m + half) * ny + b];
                double2* sq = fb + b * ld;
                if (half == 0) {
                    sq[m] = v;
                    if (m >= 1 && m <= N - 1) sq[L - m] = v;
                } else {
                    sq[m] = v;
                    sq[L - 1 - m] = v;
                }
            }
            __syncthreads();
            fft_any<FN, -1>(fb, ld, nb, plN, tid, nthr);
            for (int i = tid; i < nb * nz; i += nthr) {
                const int b = i / nz, k = i - b * nz;
                double2* c = coef + (g * TB + b) * nz + k;
                if (half == 0) {
                    *c = fb[b * ld + k];
                } else {
                    double2 t = cscale(cadd(*c, cmul(tw4[k], fb[b * ld + k])), factor);
                    if (k == 0) t = cscale(t, 0.5);
                    *c = t;
                }
            }
            __syncthreads();
        }
    }
    // rhsNonlinear: N_v -> the Lz slot (read only by its own thread), N_eta -> the FFT tile
    for (int i = tid; i < nb * nz; i += nthr) {
        const int b = i / nz, k = i - b * nz, ky = ky0 + b;
        const double2* Lx = coef + (0 * TB + b) * nz;
        const double2* Ly = coef + (1 * TB + b) * nz;
        double2* Lz = coef + (2 * TB + b) * nz;
        double2 nv, ne;
        if (kx == 0 && ky == 0) {
            nv = Lx[k];
            ne = Ly[k];
        } else {
            const double ia = (double)kx / aphi;
            const double ib = ky >= ny / 2 + 1 ? (double)(ky - ny) / beta : (double)ky / beta;
            const double kmn = ia * ia + ib * ib;
            const double2 dx = ddz_at(Lx, k, nz), dy = ddz_at(Ly, k, nz);
            double ure = kmn * Lz[k].x - ia * dx.y;
            double uim = kmn * Lz[k].y + ia * dx.x;
            ure = ure - ib * dy.y;
            uim = uim + ib * dy.x;
            const double wre = -1.0 * ia * Ly[k].y + ib * Lx[k].y;
            const double wim = ia * Ly[k].x - ib * Lx[k].x;
            nv = make_double2(ure, uim);
            ne = make_double2(wre, wim);
        }
        Lz[k] = nv;
        fb[b * ld + k] = ne;
    }
    __syncthreads();
    for (int i = tid; i < nb * nz; i += nthr) {
        const int b = i / nz, k = i - b * nz;
        coef[(0 * TB + b) * nz + k] = fb[b * ld + k];  // N_eta -> the Lx slot
    }
    __syncthreads();
    for (int h = 0; h < 2; h++) {
        const double2* src = coef + (h ? 0 : 2) * TB * nz;  // N_v in the Lz slot, N_eta in the Lx slot
        for (int i = tid; i < nb * nz; i += nthr) {
            const int b = i / nz, k = i - b * nz;
            s2p_even_ext(fb + b * ld, src[b * nz + k], k, N);
        }
        __syncthreads();
        fft_any<FN, -1>(fb, ld, nb, plN, tid, nthr);
        double2* out = h ? NLe : NLv;
        for (int i = tid; i < nb * nz; i += nthr) {
            const int b = i / nz, p = i - b * nz;
            out[((long)(ky0 + b) * d.hnx + kx) * nz + p] = fb[b * ld + p];
        }
        __syncthreads();
    }
}
