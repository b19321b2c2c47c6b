// k_cheb_mfma.hpp -- the wall-normal Chebyshev transforms of the mixed-precision form as
// dense matrix products on the matrix cores, with FP32 accumulation of split FP16 products.
//
// Reference counterparts: the padded s2p of the six velocity/vorticity fields
// (cheby_s2p, source/transform.cu:679-698, 716-733) and the padded p2s of the three Lamb
// components (cheby_p2s, source/transform.cu:538-658) followed by rhsNonlinear
// (source/nonlinear.cu:146-242) and the unpadded s2p of N_v, N_eta (source/rhs.cpp:96-97).
//
// Split products.  A value x (FP32 or FP64) scaled by a power of two 2^e into [2^14, 2^15)
// is held as x_hi = fp16(x 2^e) and x_lo = fp16(x 2^e - x_hi): 22 significant bits.  A
// table entry t is held the same way with one power-of-two scale per table.  The product
// sum_k t_k x_k is taken as three FP16 products on v_mfma_f32_16x16x32_f16,
// t_hi x_hi + t_hi x_lo + t_lo x_hi, accumulated in FP32 (the dropped t_lo x_lo is 2^-22 of
// t x); the result is scaled back by 2^-(e + table scale), exactly.  The scale e is chosen
// per pencil (per component), so pencils of small amplitude keep their relative accuracy.
//
// Reflection parity.  At the Gauss-Lobatto points z_{M-p} = -z_p, so T_k(z_{M-p}) =
// (-1)^k T_k(z_p): every product is split into the half of the rows (p <= M/2) against the
// even and against the odd coefficients (s2p), or against the sums and differences of the
// mirrored samples (p2s), and the mirrored row is their difference.
//
// MFMA operand layouts of v_mfma_f32_16x16x32_f16 (a 16 x 32 by 32 x 16 product into a
// 16 x 16 FP32 tile; checked on the GPU): lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16] for i < 8, and D[4 (l / 16) + r][l % 16] for r < 4.  Tables
// (the A operands) are stored in this lane order, fragment (row tile, k-step) as
// [hi, lo][64 lanes][8 halves]; the data (the B operands) go through LDS as
// [column][k] halves, 16 pencils' real parts in columns 0..15 and imaginary parts in
// 16..31, so that the lane holding a real part also holds the imaginary part.
#pragma once
#include "k_transform.hpp"

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

namespace cheb_mfma {

constexpr int kPencils = 16;  // pencils per block: one 16-column tile of real parts, one of imaginary parts
constexpr int kCols = 2 * kPencils;
constexpr int kTableScale = 14;  // s2p tables: |entry| <= 1 scaled by 2^14

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
constexpr int round_up(int a, int b) { return ceil_div(a, b) * b; }

// Sizes of the products of one compiled length 2N.
template <int FN>
struct Shape {
    static constexpr int N = FN / 2, nz = N + 1, pz = FN + 1;
    // s2p: rows q = 0..N against the even coefficients (N/2 + 1), rows q < N against the odd (N/2)
    static constexpr int RTE = ceil_div(N + 1, 16), KSE = ceil_div(N / 2 + 1, 32);
    static constexpr int RTO = ceil_div(N, 16), KSO = ceil_div(N / 2, 32);
    static constexpr int LDE = KSE * 32 + 8, LDO = KSO * 32 + 8;  // LDS row strides (halves)
    // p2s: rows p = 0..N/2 ("sym", mirrored rows equal) and p < N/2 ("anti", mirrored rows
    // negated), against G+ (N + 1 entries: sums of mirrored samples and the centre) or G- (N)
    static constexpr int RTS = ceil_div(N / 2 + 1, 16), RTA = ceil_div(N / 2, 16);
    static constexpr int KSP = ceil_div(N + 1, 32), KSM = ceil_div(N, 32);
    static constexpr int LDP = KSP * 32 + 8, LDM = KSM * 32 + 8;
    // fragments of each table, in the order they are stored
    static constexpr int FE = RTE * KSE, FO = RTO * KSO;
    static constexpr int F0S = RTS * KSP, F0A = RTA * KSM, F1S = RTS * KSM, F1A = RTA * KSP;
    static constexpr int frags_s2p = FE + FO;
    static constexpr int frags_p2s = F0S + F0A + F1S + F1A;
};

__device__ __forceinline__ f4_t mma3(const h8_t& ah, const h8_t& al, const h8_t& bh, const h8_t& bl, f4_t acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
    return acc;
}

// power-of-two exponent e with m 2^e in [2^14, 2^15) (0 for m = 0)
__device__ __forceinline__ int scale_exp(float m) {
    if (!(m > 0.f)) return 0;
    int ex;
    (void)frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
    return 15 - ex;
}

__device__ __forceinline__ void split_store(_Float16* hi, _Float16* lo, double x) {
    const _Float16 h = (_Float16)(float)x;
    *hi = h;
    *lo = (_Float16)(float)(x - (double)(float)h);
}

// LDS max of non-negative floats through their bit patterns (ordered like the values)
__device__ __forceinline__ void lds_max(unsigned* slot, float v) { atomicMax(slot, __float_as_uint(v)); }

}  // namespace cheb_mfma

// -------------------------------------------------------------------------------------
// Padded s2p of the six spectral fields (mixed form): F[f][P][nz] (FP64) ->
// Y[f][kx][q][ky] (FP32), q = 0..2N.  f_q = sum_k a_k cos(pi k q / (2N)); with
// E_q = sum_{k even} a_k cos(pi k q/(2N)) (q <= N) and O_q the odd sum (q < N):
// f_q = E_q + O_q, f_{2N-q} = E_q - O_q, f_N = E_N.  Block: 16 pencils (ky0.., kx fixed) of
// one field (blockIdx.z = member * 6 + field), 256 threads; wave w takes the row tiles
// w, w + 4, .. of E and the same row tiles of O.  tab: the E fragments, then the O
// fragments.
// -------------------------------------------------------------------------------------
template <int FN>
__global__ __launch_bounds__(256) void k_s2p_mfma(const double2* __restrict__ F, long fstride, long sF,
                                                  float2* __restrict__ Y, long ystride, long sY, Dims d,
                                                  const h8_t* __restrict__ tab, const int* __restrict__ ids) {
    using namespace cheb_mfma;
    using Sh = Shape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, pz = Sh::pz;
    constexpr int KC = ceil_div(kPencils * nz, 256);
    __shared__ __align__(16) _Float16 bE[2][kCols * Sh::LDE];  // [hi, lo][column][even k / 2]
    __shared__ __align__(16) _Float16 bO[2][kCols * Sh::LDO];  // [hi, lo][column][(odd k - 1) / 2]
    __shared__ unsigned smax[kPencils];
    __shared__ int sexp[kPencils];
    const int slot = ids[blockIdx.z / 6], field = blockIdx.z % 6;
    F += slot * sF + field * fstride;
    Y += slot * sY + field * ystride;
    const int kx = blockIdx.y, ky0 = blockIdx.x * kPencils;
    const int nb = min(kPencils, d.ny - ky0);
    const int tid = threadIdx.x, total = nb * nz;
    const long pstride = (long)d.hnx * nz, pin = ((long)ky0 * d.hnx + kx) * nz;
    double2 a[KC];
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * 256;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            a[u] = F[pin + b * pstride + k];
        }
    }
    // zero the operand tiles (padding k and ragged columns stay zero)
    {
        uint4* z = (uint4*)&bE[0][0];
        constexpr int nE = 2 * kCols * Sh::LDE * 2 / 16, nO = 2 * kCols * Sh::LDO * 2 / 16;
        for (int i = tid; i < nE; i += 256) z[i] = make_uint4(0, 0, 0, 0);
        z = (uint4*)&bO[0][0];
        for (int i = tid; i < nO; i += 256) z[i] = make_uint4(0, 0, 0, 0);
        if (tid < kPencils) smax[tid] = 0u;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * 256;
        if (i < total) lds_max(&smax[i / nz], (float)fmax(fabs(a[u].x), fabs(a[u].y)));
    }
    __syncthreads();
    if (tid < kPencils) sexp[tid] = scale_exp(__uint_as_float(smax[tid]));
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * 256;
        if (i < total) {
            const int b = i / nz, k = i - b * nz, e = sexp[b], j = k >> 1;
            const double xr = ldexp(a[u].x, e), xi = ldexp(a[u].y, e);
            if (k & 1) {
                split_store(&bO[0][b * Sh::LDO + j], &bO[1][b * Sh::LDO + j], xr);
                split_store(&bO[0][(kPencils + b) * Sh::LDO + j], &bO[1][(kPencils + b) * Sh::LDO + j], xi);
            } else {
                split_store(&bE[0][b * Sh::LDE + j], &bE[1][b * Sh::LDE + j], xr);
                split_store(&bE[0][(kPencils + b) * Sh::LDE + j], &bE[1][(kPencils + b) * Sh::LDE + j], xi);
            }
        }
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, lc = lane & 15, lk = 8 * (lane >> 4);
    const h8_t* tabE = tab;
    const h8_t* tabO = tab + 2 * 64 * Sh::FE;
    const int b = lc;
    const float sc = b < nb ? ldexpf(1.f, -sexp[b] - kTableScale) : 0.f;
    for (int rt = wave; rt < Sh::RTE; rt += 4) {
        f4_t eR = {0, 0, 0, 0}, eI = eR, oR = eR, oI = eR;
#pragma unroll
        for (int s = 0; s < Sh::KSE; s++) {
            const h8_t ah = tabE[((rt * Sh::KSE + s) * 2 + 0) * 64 + lane];
            const h8_t al = tabE[((rt * Sh::KSE + s) * 2 + 1) * 64 + lane];
            const int o = s * 32 + lk;
            const h8_t brh = *(const h8_t*)&bE[0][lc * Sh::LDE + o], brl = *(const h8_t*)&bE[1][lc * Sh::LDE + o];
            const h8_t bih = *(const h8_t*)&bE[0][(16 + lc) * Sh::LDE + o];
            const h8_t bil = *(const h8_t*)&bE[1][(16 + lc) * Sh::LDE + o];
            eR = mma3(ah, al, brh, brl, eR);
            eI = mma3(ah, al, bih, bil, eI);
        }
        if (rt < Sh::RTO) {
#pragma unroll
            for (int s = 0; s < Sh::KSO; s++) {
                const h8_t ah = tabO[((rt * Sh::KSO + s) * 2 + 0) * 64 + lane];
                const h8_t al = tabO[((rt * Sh::KSO + s) * 2 + 1) * 64 + lane];
                const int o = s * 32 + lk;
                const h8_t brh = *(const h8_t*)&bO[0][lc * Sh::LDO + o], brl = *(const h8_t*)&bO[1][lc * Sh::LDO + o];
                const h8_t bih = *(const h8_t*)&bO[0][(16 + lc) * Sh::LDO + o];
                const h8_t bil = *(const h8_t*)&bO[1][(16 + lc) * Sh::LDO + o];
                oR = mma3(ah, al, brh, brl, oR);
                oI = mma3(ah, al, bih, bil, oI);
            }
        }
        if (b < nb) {
            float2* out = Y + (long)kx * pz * d.ny + ky0 + b;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int q = rt * 16 + (lane >> 4) * 4 + r;
                if (q < N) {
                    out[(long)q * d.ny] = make_float2((eR[r] + oR[r]) * sc, (eI[r] + oI[r]) * sc);
                    out[(long)(FN - q) * d.ny] = make_float2((eR[r] - oR[r]) * sc, (eI[r] - oI[r]) * sc);
                } else if (q == N) {
                    out[(long)q * d.ny] = make_float2(eR[r] * sc, eI[r] * sc);
                }
            }
        }
    }
}

// -------------------------------------------------------------------------------------
// Padded p2s of the three Lamb components, rhsNonlinear and the unpadded s2p of N_v, N_eta
// (mixed form) as three products of combined, folded samples.  With the samples f_q
// (q = 0..2N) of a component, its coefficients a = P f (the reference's p2s:
// a_k = (1/2N)(f_0 + (-1)^k f_2N + 2 sum_{0<q<2N} f_q cos(pi k q/(2N))), a_0 halved, k <= N),
// their derivative D a (ddz: D_k = c_k sum_{j>k, j-k odd} 2 j a_j, c_0 = 1/2) and values at
// the nz collocation points S a (S_pk = cos(pi k p/N)), the values of a component are
// M0 f (M0 = S P) and those of its derivative M1 f (M1 = S D P).  rhsNonlinear
// (source/nonlinear.cu:146-242) is linear per pencil, so it commutes with these products:
//   N_v   = M0 h1 + M1 h2,  h1 = k^2 f_z,  h2 = i alpha f_x + i beta f_y,
//   N_eta = M0 h3,          h3 = i alpha f_y - i beta f_x,
// and for pencil (0,0) N_v = M0 f_x, N_eta = M0 f_y (h1 = f_x, h2 = 0, h3 = f_y).
// Folding with G+_q = h_q + h_{2N-q}, G-_q = h_q - h_{2N-q} (q < N), G+_N = h_N:
//   M0 keeps the reflection parity: (M0 h)_p = E_p + A_p, (M0 h)_{N-p} = E_p - A_p,
//     E = Q0s G+ (rows p <= N/2), A = Q0a G- (rows p < N/2);
//   M1 flips it: (M1 h)_p = E_p + A_p, (M1 h)_{N-p} = E_p - A_p,
//     E = Q1s G- (rows p <= N/2), A = Q1a G+ (rows p < N/2);
// Q0s[p][q] = (M0[p][q] + M0[p][2N-q])/2, Q0s[p][N] = M0[p][N], Q0a the half difference;
// Q1a the half sums of M1 (with M1[p][N]), Q1s its half differences.
// Block: 16 pencils (kx fixed) of one member (blockIdx.z), one wave per row tile of the
// p <= N/2 rows (64 * RTS threads); wave w also holds row tile w of the p < N/2 rows.  The
// inputs h3, h2, h1 are taken one at a time: fold and combine in FP32, one scale per pencil,
// split into LDS, products.  Output: N_v, N_eta (FP64) of the slot's current history buffers.
// tab: fragments of Q0s, Q0a, Q1s, Q1a (in this order); s0, s1: scales of the M0 and M1 tables.
// -------------------------------------------------------------------------------------
template <int FN>
__global__ __launch_bounds__(64 * cheb_mfma::Shape<FN>::RTS) void k_p2s_mfma(
    const float2* __restrict__ Z1, long zstride, long sY, double2* __restrict__ NL, long sNL, Dims d,
    const h8_t* __restrict__ tab, int s0, int s1, Members mb) {
    using namespace cheb_mfma;
    using Sh = Shape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, pz = Sh::pz;
    constexpr int NT = 64 * Sh::RTS;               // threads
    constexpr int KQ = ceil_div(kPencils * (N + 1), NT);
    __shared__ __align__(16) _Float16 bP[2][kCols * Sh::LDP];  // G+ [hi, lo][column][q]
    __shared__ __align__(16) _Float16 bM[2][kCols * Sh::LDM];  // G-
    __shared__ unsigned smax[3][kPencils];
    __shared__ int sexp[3][kPencils];
    const int mslot = mb.ids[blockIdx.z];
    const int ph = mb.phase[mslot];
    const double aphi = mb.mp[mslot].aphi, beta = mb.mp[mslot].beta;
    const long PS = (long)d.P * nz;
    Z1 += mslot * sY;
    double2* NLv = NL + mslot * sNL + ph * PS;        // N_v, current
    double2* NLe = NL + mslot * sNL + (2 + ph) * PS;  // N_eta, current
    const int kx = blockIdx.y, ky0 = blockIdx.x * kPencils;
    const int nb = min(kPencils, d.ny - ky0);
    const int tid = threadIdx.x, total = nb * (N + 1);
    const int lane = tid & 63, rt = tid >> 6, lc = lane & 15, lk = 8 * (lane >> 4);
    const h8_t* t0s = tab;
    const h8_t* t0a = t0s + 2 * 64 * Sh::F0S;
    const h8_t* t1s = t0a + 2 * 64 * Sh::F0A;
    const h8_t* t1a = t1s + 2 * 64 * Sh::F1S;
    const bool anti = rt < Sh::RTA;
    if (tid < 3 * kPencils) (&smax[0][0])[tid] = 0u;
    // pencil scalars of an item's pencil b: i alpha, i beta, k^2; pencil (0,0) flagged
    auto scal = [&](int b, float* ia, float* ib, float* k2) {
        const int ky = ky0 + b;
        const double a_ = (double)kx / aphi;
        const double b_ = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / beta : (double)ky / beta;
        *ia = (float)a_;
        *ib = (float)b_;
        *k2 = (float)(a_ * a_ + b_ * b_);
        return kx == 0 && ky == 0;
    };
    // samples q and 2N - q of one component for every item (q = i / nb, b = i % nb)
    auto load = [&](int g, float2* fa, float2* fb) {
        const float2* in = Z1 + g * zstride + (long)kx * pz * d.ny + ky0;
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb;
                fa[u] = in[(long)q * d.ny + b];
                fb[u] = in[(long)(FN - q) * d.ny + b];
            }
        }
    };
    // accumulators (re, im tiles): N_eta sym/anti via M0, N_v sym/anti via M0 and via M1
    f4_t es[2], ea[2], v0s[2], v0a[2], v1s[2], v1a[2];
#pragma unroll
    for (int c = 0; c < 2; c++) es[c] = ea[c] = v0s[c] = v0a[c] = v1s[c] = v1a[c] = f4_t{0, 0, 0, 0};
    // one input h: hp, hm its folded G+, G- per item; products sym (table ts, K = ks blocks of
    // the G+ tile if sym_plus) and anti into the given accumulators
    auto stage = [&](int hidx, const float2* hp, const float2* hm, const h8_t* ts, const h8_t* ta, bool sym_plus,
                     f4_t* accs, f4_t* acca) {
        {
            uint4* z = (uint4*)&bP[0][0];
            constexpr int nP = 2 * kCols * Sh::LDP * 2 / 16, nM = 2 * kCols * Sh::LDM * 2 / 16;
            for (int i = tid; i < nP; i += NT) z[i] = make_uint4(0, 0, 0, 0);
            z = (uint4*)&bM[0][0];
            for (int i = tid; i < nM; i += NT) z[i] = make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int b = i % nb;
                lds_max(&smax[hidx][b], fmaxf(fmaxf(fabsf(hp[u].x), fabsf(hp[u].y)), fmaxf(fabsf(hm[u].x), fabsf(hm[u].y))));
            }
        }
        __syncthreads();
        if (tid < kPencils) sexp[hidx][tid] = scale_exp(__uint_as_float(smax[hidx][tid]));
        __syncthreads();
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb, e = sexp[hidx][b];
                split_store(&bP[0][b * Sh::LDP + q], &bP[1][b * Sh::LDP + q], ldexpf(hp[u].x, e));
                split_store(&bP[0][(kPencils + b) * Sh::LDP + q], &bP[1][(kPencils + b) * Sh::LDP + q], ldexpf(hp[u].y, e));
                if (q < N) {
                    split_store(&bM[0][b * Sh::LDM + q], &bM[1][b * Sh::LDM + q], ldexpf(hm[u].x, e));
                    split_store(&bM[0][(kPencils + b) * Sh::LDM + q], &bM[1][(kPencils + b) * Sh::LDM + q],
                                ldexpf(hm[u].y, e));
                }
            }
        }
        __syncthreads();
        auto prod = [&](const h8_t* t, int ks, const _Float16* bh, const _Float16* bl, int ld, f4_t* acc) {
            for (int s = 0; s < ks; s++) {
                const h8_t ah = t[((rt * ks + s) * 2 + 0) * 64 + lane];
                const h8_t al = t[((rt * ks + s) * 2 + 1) * 64 + lane];
                const int o = s * 32 + lk;
                acc[0] = mma3(ah, al, *(const h8_t*)&bh[lc * ld + o], *(const h8_t*)&bl[lc * ld + o], acc[0]);
                acc[1] = mma3(ah, al, *(const h8_t*)&bh[(16 + lc) * ld + o], *(const h8_t*)&bl[(16 + lc) * ld + o], acc[1]);
            }
        };
        // sym rows from G+ (M0) or G- (M1); anti rows from the other one
        if (sym_plus) prod(ts, Sh::KSP, bP[0], bP[1], Sh::LDP, accs);
        else prod(ts, Sh::KSM, bM[0], bM[1], Sh::LDM, accs);
        if (anti) {
            if (sym_plus) prod(ta, Sh::KSM, bM[0], bM[1], Sh::LDM, acca);
            else prod(ta, Sh::KSP, bP[0], bP[1], Sh::LDP, acca);
        }
        __syncthreads();
    };
    float2 xa[KQ], xb[KQ], ya[KQ], yb[KQ];
    __syncthreads();  // smax cleared
    load(0, xa, xb);
    load(1, ya, yb);
    {
        // h3 = i alpha f_y - i beta f_x (pencil (0,0): f_y);  h2 = i alpha f_x + i beta f_y (0)
        float2 hp[KQ], hm[KQ];
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb;
                float ia, ib, k2;
                const bool mean = scal(b, &ia, &ib, &k2);
                const float2 xp = q < N ? make_float2(xa[u].x + xb[u].x, xa[u].y + xb[u].y) : xa[u];
                const float2 xm = q < N ? make_float2(xa[u].x - xb[u].x, xa[u].y - xb[u].y) : make_float2(0.f, 0.f);
                const float2 yp = q < N ? make_float2(ya[u].x + yb[u].x, ya[u].y + yb[u].y) : ya[u];
                const float2 ym = q < N ? make_float2(ya[u].x - yb[u].x, ya[u].y - yb[u].y) : make_float2(0.f, 0.f);
                // i a y - i b x = (-a y.im + b x.im, a y.re - b x.re)
                hp[u] = mean ? yp : make_float2(-ia * yp.y + ib * xp.y, ia * yp.x - ib * xp.x);
                hm[u] = mean ? ym : make_float2(-ia * ym.y + ib * xm.y, ia * ym.x - ib * xm.x);
                // keep h2 in the x registers: i a x + i b y = (-a x.im - b y.im, a x.re + b y.re)
                xa[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xp.y - ib * yp.y, ia * xp.x + ib * yp.x);
                xb[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xm.y - ib * ym.y, ia * xm.x + ib * ym.x);
            }
        }
        load(2, ya, yb);  // f_z for h1
        stage(0, hp, hm, t0s, t0a, true, es, ea);
    }
    stage(1, xa, xb, t1s, t1a, false, v1s, v1a);
    {
        // h1 = k^2 f_z (pencil (0,0): f_x, reloaded)
        const bool has_mean = kx == 0 && ky0 == 0;
        if (has_mean) load(0, xa, xb);
        float2 hp[KQ], hm[KQ];
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb;
                float ia, ib, k2;
                const bool mean = scal(b, &ia, &ib, &k2);
                const float2 fa = mean ? xa[u] : ya[u], fb = mean ? xb[u] : yb[u];
                const float2 zp = q < N ? make_float2(fa.x + fb.x, fa.y + fb.y) : fa;
                const float2 zm = q < N ? make_float2(fa.x - fb.x, fa.y - fb.y) : make_float2(0.f, 0.f);
                hp[u] = mean ? zp : make_float2(k2 * zp.x, k2 * zp.y);
                hm[u] = mean ? zm : make_float2(k2 * zm.x, k2 * zm.y);
            }
        }
        stage(2, hp, hm, t0s, t0a, true, v0s, v0a);
    }
    // combine per pencil: lane holds pencil b = lc, rows 16 rt + 4 (lane / 16) + r
    const int b = lc;
    if (b >= nb) return;
    const double se = ldexp(1.0, -sexp[0][b] - s0), sv1 = ldexp(1.0, -sexp[1][b] - s1),
                 sv0 = ldexp(1.0, -sexp[2][b] - s0);
    const long po = ((long)(ky0 + b) * d.hnx + kx) * nz;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int p = rt * 16 + (lane >> 4) * 4 + r;
        if (2 * p > N) continue;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            if (side == 1 && 2 * p == N) continue;
            const double sg = side ? -1.0 : 1.0;
            auto val = [&](const f4_t* s_, const f4_t* a_, int c) {
                return (double)s_[c][r] + (anti ? sg * (double)a_[c][r] : 0.0);
            };
            const double2 ne = make_double2(val(es, ea, 0) * se, val(es, ea, 1) * se);
            const double2 nv = make_double2(val(v0s, v0a, 0) * sv0 + val(v1s, v1a, 0) * sv1,
                                            val(v0s, v0a, 1) * sv0 + val(v1s, v1a, 1) * sv1);
            const int row = side ? N - p : p;
            NLv[po + row] = nv;
            NLe[po + row] = ne;
        }
    }
}
