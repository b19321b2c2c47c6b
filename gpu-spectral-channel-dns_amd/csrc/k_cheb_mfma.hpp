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
    // LDS row strides of the p2s operand tiles: padded by 8 halves against bank conflicts,
    // except at 2N = 384, where three operand sets would not fit in 160 KB with the padding
    static constexpr int LDP = KSP * 32 + (FN >= 384 ? 0 : 8), LDM = KSM * 32 + (FN >= 384 ? 0 : 8);
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
// Padded s2p of the six spectral fields (mixed form): F[f][P][nz] (FP32, as the velocity
// recovery of this form stores them) -> Y[f][kx][q][ky] (FP32), q = 0..2N.
// f_q = sum_k a_k cos(pi k q / (2N)); with E_q = sum_{k even} a_k cos(pi k q/(2N)) (q <= N)
// and O_q the odd sum (q < N): f_q = E_q + O_q, f_{2N-q} = E_q - O_q, f_N = E_N.
// Persistent blocks of max(8, RTE) waves: wave w keeps the table fragments of row tile w
// (E, and O when w < RTO) in registers for the whole launch, so the tables are read once
// per wave.  Items (member, field, kx, 16 consecutive ky) go round-robin over the grid in
// pairs: thread t < 512 loads for item 2 p + t / 256 the run of coefficients
// k = 16 m .. 16 m + 15 of pencil b ((b, m) = ((t % 256) / 16, t % 16), contiguous), takes
// the pencil's largest magnitude from its 16 lanes (shuffles) and writes the run, split
// and scaled, as 16-byte rows of the operand tiles (zeros past nz and for missing pencils).
// Operand tiles and scales are double buffered: the next pair's runs are in flight while
// the current pair's products and stores run; one barrier per pair.
// LDS (dynamic): [2 buffers][2 items][E hi | E lo | O hi | O lo][column][k / 2] halves.
// -------------------------------------------------------------------------------------
template <int FN>
struct S2pMfma {
    using Sh = cheb_mfma::Shape<FN>;
    static constexpr int waves = Sh::RTE > 8 ? Sh::RTE : 8;
    static constexpr int threads = 64 * waves;
    static constexpr int tile = 2 * cheb_mfma::kCols * (Sh::LDE + Sh::LDO);  // halves of one item's operands
    static constexpr size_t lds = sizeof(_Float16) * 4 * tile;
};

template <int FN>
__global__ __launch_bounds__(S2pMfma<FN>::threads) void k_s2p_mfma(const float2* __restrict__ F, long fstride,
                                                                   long sF, float2* __restrict__ Y, long ystride,
                                                                   long sY, Dims d, const h8_t* __restrict__ tab,
                                                                   const int* __restrict__ ids, int nmem) {
    using namespace cheb_mfma;
    using Sh = Shape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, pz = Sh::pz;
    constexpr int TILE = S2pMfma<FN>::tile;
    static_assert(nz <= 16 * 16 && 8 * 16 >= Sh::KSE * 32, "one 16-coefficient run per lane");
    RPCF_LDS(_Float16);  // [2][2][TILE]
    __shared__ int sexp[2][2][kPencils];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lc = lane & 15, lk = 8 * (lane >> 4);
    const int tiles_y = (d.ny + kPencils - 1) / kPencils;
    const int items = nmem * 6 * d.hnx * tiles_y, pairs = (items + 1) / 2;
    // this wave's table fragments, for the whole launch
    h8_t aeh[Sh::KSE], ael[Sh::KSE], aoh[Sh::KSO], aol[Sh::KSO];
    const int rt = wave;
    const bool has_e = rt < Sh::RTE, has_o = rt < Sh::RTO;
#pragma unroll
    for (int s = 0; s < Sh::KSE; s++) {
        aeh[s] = has_e ? tab[((rt * Sh::KSE + s) * 2 + 0) * 64 + lane] : h8_t{};
        ael[s] = has_e ? tab[((rt * Sh::KSE + s) * 2 + 1) * 64 + lane] : h8_t{};
    }
    const h8_t* tabO = tab + 2 * 64 * Sh::FE;
#pragma unroll
    for (int s = 0; s < Sh::KSO; s++) {
        aoh[s] = has_o ? tabO[((rt * Sh::KSO + s) * 2 + 0) * 64 + lane] : h8_t{};
        aol[s] = has_o ? tabO[((rt * Sh::KSO + s) * 2 + 1) * 64 + lane] : h8_t{};
    }
    // item -> (member slot, field, kx, ky0)
    auto decode = [&](int it, int* slot, int* field, int* kx, int* ky0) {
        const int t = it % tiles_y, r = it / tiles_y;
        *kx = r % d.hnx;
        const int mf = r / d.hnx;
        *field = mf % 6;
        *slot = ids[mf / 6];
        *ky0 = t * kPencils;
    };
    const int lu = tid >> 8, lb = (tid >> 4) & 15, lm = tid & 15, k0 = 16 * lm;  // loader: item, pencil, run
    float2 a[16];
    auto load = [&](int pr) {
#pragma unroll
        for (int i = 0; i < 16; i++) a[i] = make_float2(0.f, 0.f);
        const int it = 2 * pr + lu;
        if (tid >= 512 || pr >= pairs || it >= items) return;
        int slot, field, kx, ky0;
        decode(it, &slot, &field, &kx, &ky0);
        if (ky0 + lb >= d.ny) return;
        const float2* src = F + slot * sF + field * fstride + ((long)(ky0 + lb) * d.hnx + kx) * nz + k0;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (k0 + i < nz) a[i] = src[i];
    };
    auto split = [&](int pb) {
        if (tid >= 512) return;
        float mx = 0.f;
#pragma unroll
        for (int i = 0; i < 16; i++) mx = fmaxf(mx, fmaxf(fabsf(a[i].x), fabsf(a[i].y)));
        mx = fmaxf(mx, __shfl_xor(mx, 1));
        mx = fmaxf(mx, __shfl_xor(mx, 2));
        mx = fmaxf(mx, __shfl_xor(mx, 4));
        mx = fmaxf(mx, __shfl_xor(mx, 8));
        const int e = scale_exp(mx);
        if (lm == 0) sexp[pb][lu][lb] = e;
        h8_t erh, erl, eih, eil, orh, orl, oih, oil;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float xr = ldexpf(a[2 * i].x, e), xi = ldexpf(a[2 * i].y, e);
            const float yr = ldexpf(a[2 * i + 1].x, e), yi = ldexpf(a[2 * i + 1].y, e);
            erh[i] = (_Float16)xr;
            erl[i] = (_Float16)(xr - (float)erh[i]);
            eih[i] = (_Float16)xi;
            eil[i] = (_Float16)(xi - (float)eih[i]);
            orh[i] = (_Float16)yr;
            orl[i] = (_Float16)(yr - (float)orh[i]);
            oih[i] = (_Float16)yi;
            oil[i] = (_Float16)(yi - (float)oih[i]);
        }
        _Float16* eh = lds + (pb * 2 + lu) * TILE;
        _Float16* el = eh + kCols * Sh::LDE;
        _Float16* oh = el + kCols * Sh::LDE;
        _Float16* ol = oh + kCols * Sh::LDO;
        const int j0 = 8 * lm;
        if (j0 < Sh::KSE * 32) {
            *(h8_t*)&eh[lb * Sh::LDE + j0] = erh;
            *(h8_t*)&el[lb * Sh::LDE + j0] = erl;
            *(h8_t*)&eh[(kPencils + lb) * Sh::LDE + j0] = eih;
            *(h8_t*)&el[(kPencils + lb) * Sh::LDE + j0] = eil;
        }
        if (j0 < Sh::KSO * 32) {
            *(h8_t*)&oh[lb * Sh::LDO + j0] = orh;
            *(h8_t*)&ol[lb * Sh::LDO + j0] = orl;
            *(h8_t*)&oh[(kPencils + lb) * Sh::LDO + j0] = oih;
            *(h8_t*)&ol[(kPencils + lb) * Sh::LDO + j0] = oil;
        }
    };
    int pr = blockIdx.x, pb = 0;
    load(pr);
    split(pb);
    __syncthreads();
    for (; pr < pairs; pr += gridDim.x, pb ^= 1) {
        load(pr + gridDim.x);  // the next pair's runs, in flight during this pair's products
        if (has_e) {
#pragma unroll 1
            for (int u = 0; u < 2; u++) {
                const int it = 2 * pr + u;
                if (it >= items) break;
                int slot, field, kx, ky0;
                decode(it, &slot, &field, &kx, &ky0);
                const int nb = min(kPencils, d.ny - ky0);
                const _Float16* eh = lds + (pb * 2 + u) * TILE;
                const _Float16* el = eh + kCols * Sh::LDE;
                const _Float16* oh = el + kCols * Sh::LDE;
                const _Float16* ol = oh + kCols * Sh::LDO;
                f4_t eR = {0, 0, 0, 0}, eI = eR, oR = eR, oI = eR;
#pragma unroll
                for (int s = 0; s < Sh::KSE; s++) {
                    const int o = s * 32 + lk;
                    eR = mma3(aeh[s], ael[s], *(const h8_t*)&eh[lc * Sh::LDE + o], *(const h8_t*)&el[lc * Sh::LDE + o],
                              eR);
                    eI = mma3(aeh[s], ael[s], *(const h8_t*)&eh[(16 + lc) * Sh::LDE + o],
                              *(const h8_t*)&el[(16 + lc) * Sh::LDE + o], eI);
                }
                if (has_o) {
#pragma unroll
                    for (int s = 0; s < Sh::KSO; s++) {
                        const int o = s * 32 + lk;
                        oR = mma3(aoh[s], aol[s], *(const h8_t*)&oh[lc * Sh::LDO + o],
                                  *(const h8_t*)&ol[lc * Sh::LDO + o], oR);
                        oI = mma3(aoh[s], aol[s], *(const h8_t*)&oh[(16 + lc) * Sh::LDO + o],
                                  *(const h8_t*)&ol[(16 + lc) * Sh::LDO + o], oI);
                    }
                }
                const int b = lc;
                if (b < nb) {
                    const float sc = ldexpf(1.f, -sexp[pb][u][b] - kTableScale);
                    float2* out = Y + slot * sY + field * ystride + (long)kx * pz * d.ny + ky0 + b;
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
        split(pb ^ 1);  // the other buffers: their last readers finished before the previous barrier
        __syncthreads();
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
// Block: 16 pencils (kx fixed) of one member (blockIdx.z); two waves per row tile of the
// p <= N/2 rows (128 * RTS threads), wave 2 rt + c holding the real (c = 0) or imaginary
// (c = 1) parts of the row tile rt of every product: after the per-pencil combinations
// the products act on real and imaginary parts alike.  One stage: the three components
// are loaded together, folded and combined in FP32 into h3, h2, h1, each with one
// power-of-two scale per pencil (largest magnitude by shuffles and one LDS atomic per wave
// and pencil), split into three operand sets in LDS; every Q0 fragment read from the table
// then serves both h3 (N_eta) and h1 (N_v), the Q1 fragments h2.  The results go through
// LDS to contiguous rows of N_v, N_eta (FP64) in the slot's current history buffers.
// tab: fragments of Q0s, Q0a, Q1s, Q1a (in this order); s0, s1: scales of the M0 and M1 tables.
// LDS (dynamic): three operand sets [set][hi, lo][32 columns][LDP (G+) | LDM (G-)] halves.
// -------------------------------------------------------------------------------------
template <int FN>
constexpr size_t p2s_mfma_lds() {
    using Sh = cheb_mfma::Shape<FN>;
    return sizeof(_Float16) * 3 * 2 * cheb_mfma::kCols * (Sh::LDP + Sh::LDM);
}

template <int FN>
__global__ __launch_bounds__(128 * cheb_mfma::Shape<FN>::RTS) void k_p2s_mfma(
    const float2* __restrict__ Z1, long zstride, long sY, double2* __restrict__ NL, long sNL, Dims d,
    const h8_t* __restrict__ tab, int s0, int s1, Members mb) {
    using namespace cheb_mfma;
    using Sh = Shape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, pz = Sh::pz;
    constexpr int NT = 128 * Sh::RTS;  // threads
    constexpr int KQ = ceil_div(kPencils * (N + 1), NT);
    constexpr int SET = 2 * kCols * (Sh::LDP + Sh::LDM);  // halves per operand set
    static_assert(2 * kPencils * nz * sizeof(double2) <= p2s_mfma_lds<FN>(), "output staging fits");
    RPCF_LDS(_Float16);
    __shared__ unsigned smax[3][kPencils];
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
    const int lane = tid & 63, wave = tid >> 6, rt = wave >> 1, ct = wave & 1, lc = lane & 15, lk = 8 * (lane >> 4);
    const bool anti = rt < Sh::RTA;
    const bool has_mean = kx == 0 && ky0 == 0;
    // operand tiles of set s: G+ hi, G+ lo, G- hi, G- lo
    auto gp_hi = [&](int st) { return lds + st * SET; };
    auto gp_lo = [&](int st) { return lds + st * SET + kCols * Sh::LDP; };
    auto gm_hi = [&](int st) { return lds + st * SET + 2 * kCols * Sh::LDP; };
    auto gm_lo = [&](int st) { return lds + st * SET + 2 * kCols * Sh::LDP + kCols * Sh::LDM; };
    // samples q and 2N - q of component g for every item (q = i / nb, b = i % nb)
    float2 xa[KQ], xb[KQ], ya[KQ], yb[KQ], za[KQ], zb[KQ];
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
    load(0, xa, xb);
    load(1, ya, yb);
    load(2, za, zb);
    {
        // zero the three sets: padding k and ragged columns stay zero, the rest is overwritten
        uint4* z = (uint4*)lds;
        for (int i = tid; i < 3 * SET * 2 / 16; i += NT) z[i] = make_uint4(0, 0, 0, 0);
        if (tid < 3 * kPencils) (&smax[0][0])[tid] = 0u;
    }
    auto fold = [&](float2 a, float2 b, int q, float2* gp, float2* gm) {
        *gp = q < N ? make_float2(a.x + b.x, a.y + b.y) : a;
        *gm = q < N ? make_float2(a.x - b.x, a.y - b.y) : make_float2(0.f, 0.f);
    };
    auto mag = [](float2 a, float2 b) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(b.x), fabsf(b.y))); };
    // largest magnitude per pencil of input h over the block: with nb == 16 every item of a
    // thread is pencil tid % 16, so a thread's maximum, then the lanes of its pencil in the
    // wave (xor 16, 32), then one atomic per wave and pencil
    auto pencil_max = [&](int h, const float* m) {
        if (nb == kPencils) {
            float v = 0.f;
#pragma unroll
            for (int u = 0; u < KQ; u++) v = fmaxf(v, m[u]);
            v = fmaxf(v, __shfl_xor(v, 16));
            v = fmaxf(v, __shfl_xor(v, 32));
            if (lane < 16) lds_max(&smax[h][lane], v);
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                const int i = tid + u * NT;
                if (i < total) lds_max(&smax[h][i % nb], m[u]);
            }
        }
    };
    auto split_set = [&](int st, const float2* hp, const float2* hm) {
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb, e = scale_exp(__uint_as_float(smax[st][b]));
                split_store(&gp_hi(st)[b * Sh::LDP + q], &gp_lo(st)[b * Sh::LDP + q], ldexpf(hp[u].x, e));
                split_store(&gp_hi(st)[(kPencils + b) * Sh::LDP + q], &gp_lo(st)[(kPencils + b) * Sh::LDP + q],
                            ldexpf(hp[u].y, e));
                if (q < N) {
                    split_store(&gm_hi(st)[b * Sh::LDM + q], &gm_lo(st)[b * Sh::LDM + q], ldexpf(hm[u].x, e));
                    split_store(&gm_hi(st)[(kPencils + b) * Sh::LDM + q], &gm_lo(st)[(kPencils + b) * Sh::LDM + q],
                                ldexpf(hm[u].y, e));
                }
            }
        }
    };
    // h3 = i alpha f_y - i beta f_x -> set 0; h2 = i alpha f_x + i beta f_y -> set 1;
    // h1 = k^2 f_z -> set 2.  Pencil (0,0): h3 = f_y, h2 = 0, h1 = f_x.
    float2 h3p[KQ], h3m[KQ], h2p[KQ], h2m[KQ], h1p[KQ], h1m[KQ];
    {
        float m3[KQ], m2[KQ], m1[KQ];
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            m3[u] = m2[u] = m1[u] = 0.f;
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb, ky = ky0 + b;
                const double a_ = (double)kx / aphi;
                const double b_ = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / beta : (double)ky / beta;
                const float ia = (float)a_, ib = (float)b_, k2 = (float)(a_ * a_ + b_ * b_);
                const bool mean = has_mean && b == 0;
                float2 xp, xm, yp, ym, zp, zm;
                fold(xa[u], xb[u], q, &xp, &xm);
                fold(ya[u], yb[u], q, &yp, &ym);
                fold(za[u], zb[u], q, &zp, &zm);
                // i a y - i b x = (-a y.im + b x.im, a y.re - b x.re)
                h3p[u] = mean ? yp : make_float2(-ia * yp.y + ib * xp.y, ia * yp.x - ib * xp.x);
                h3m[u] = mean ? ym : make_float2(-ia * ym.y + ib * xm.y, ia * ym.x - ib * xm.x);
                // i a x + i b y = (-a x.im - b y.im, a x.re + b y.re)
                h2p[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xp.y - ib * yp.y, ia * xp.x + ib * yp.x);
                h2m[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xm.y - ib * ym.y, ia * xm.x + ib * ym.x);
                h1p[u] = mean ? xp : make_float2(k2 * zp.x, k2 * zp.y);
                h1m[u] = mean ? xm : make_float2(k2 * zm.x, k2 * zm.y);
                m3[u] = mag(h3p[u], h3m[u]);
                m2[u] = mag(h2p[u], h2m[u]);
                m1[u] = mag(h1p[u], h1m[u]);
            }
        }
        __syncthreads();  // sets zeroed, maxima cleared
        pencil_max(0, m3);
        pencil_max(1, m2);
        pencil_max(2, m1);
    }
    __syncthreads();
    split_set(0, h3p, h3m);
    split_set(1, h2p, h2m);
    split_set(2, h1p, h1m);
    __syncthreads();
    // N_eta: sym rows Q0s G+, anti rows Q0a G- of h3 (set 0); N_v: the same tables on h1 (set 2)
    // and sym rows Q1s G-, anti rows Q1a G+ of h2 (set 1)
    const h8_t* t0s = tab;
    const h8_t* t0a = t0s + 2 * 64 * Sh::F0S;
    const h8_t* t1s = t0a + 2 * 64 * Sh::F0A;
    const h8_t* t1a = t1s + 2 * 64 * Sh::F1S;
    const int col = ct * kPencils + lc;
    const f4_t zero4 = {0, 0, 0, 0};
    f4_t es = zero4, ea = zero4, v1s = zero4, v1a = zero4, v0s = zero4, v0a = zero4;
#pragma unroll
    for (int s = 0; s < Sh::KSP; s++) {
        const h8_t ah = t0s[((rt * Sh::KSP + s) * 2 + 0) * 64 + lane];
        const h8_t al = t0s[((rt * Sh::KSP + s) * 2 + 1) * 64 + lane];
        const int o = col * Sh::LDP + s * 32 + lk;
        es = mma3(ah, al, *(const h8_t*)&gp_hi(0)[o], *(const h8_t*)&gp_lo(0)[o], es);
        v0s = mma3(ah, al, *(const h8_t*)&gp_hi(2)[o], *(const h8_t*)&gp_lo(2)[o], v0s);
    }
#pragma unroll
    for (int s = 0; s < Sh::KSM; s++) {
        const h8_t ah = t1s[((rt * Sh::KSM + s) * 2 + 0) * 64 + lane];
        const h8_t al = t1s[((rt * Sh::KSM + s) * 2 + 1) * 64 + lane];
        const int o = col * Sh::LDM + s * 32 + lk;
        v1s = mma3(ah, al, *(const h8_t*)&gm_hi(1)[o], *(const h8_t*)&gm_lo(1)[o], v1s);
    }
    if (anti) {
#pragma unroll
        for (int s = 0; s < Sh::KSM; s++) {
            const h8_t ah = t0a[((rt * Sh::KSM + s) * 2 + 0) * 64 + lane];
            const h8_t al = t0a[((rt * Sh::KSM + s) * 2 + 1) * 64 + lane];
            const int o = col * Sh::LDM + s * 32 + lk;
            ea = mma3(ah, al, *(const h8_t*)&gm_hi(0)[o], *(const h8_t*)&gm_lo(0)[o], ea);
            v0a = mma3(ah, al, *(const h8_t*)&gm_hi(2)[o], *(const h8_t*)&gm_lo(2)[o], v0a);
        }
#pragma unroll
        for (int s = 0; s < Sh::KSP; s++) {
            const h8_t ah = t1a[((rt * Sh::KSP + s) * 2 + 0) * 64 + lane];
            const h8_t al = t1a[((rt * Sh::KSP + s) * 2 + 1) * 64 + lane];
            const int o = col * Sh::LDP + s * 32 + lk;
            v1a = mma3(ah, al, *(const h8_t*)&gp_hi(1)[o], *(const h8_t*)&gp_lo(1)[o], v1a);
        }
    }
    // combine: lane holds part ct of pencil b = lc, rows 16 rt + 4 (lane / 16) + r; results go
    // to LDS as [N_v, N_eta][pencil][row] complex doubles, then out in rows of nz
    const int b = lc;
    const double se = ldexp(1.0, -scale_exp(__uint_as_float(smax[0][b])) - s0);
    const double sv1 = ldexp(1.0, -scale_exp(__uint_as_float(smax[1][b])) - s1);
    const double sv0 = ldexp(1.0, -scale_exp(__uint_as_float(smax[2][b])) - s0);
    __syncthreads();  // operand sets no longer read: the output staging reuses them
    double* outv = (double*)lds;                        // [pencil][nz][2]
    double* oute = outv + 2 * kPencils * nz;
    if (b < nb) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int p = rt * 16 + (lane >> 4) * 4 + r;
            if (2 * p > N) continue;
#pragma unroll
            for (int side = 0; side < 2; side++) {
                if (side == 1 && 2 * p == N) continue;
                const double sg = side ? -1.0 : 1.0;
                const double ev = (double)es[r] + (anti ? sg * (double)ea[r] : 0.0);
                const double vv0 = (double)v0s[r] + (anti ? sg * (double)v0a[r] : 0.0);
                const double vv1 = (double)v1s[r] + (anti ? sg * (double)v1a[r] : 0.0);
                const int row = side ? N - p : p;
                oute[(b * nz + row) * 2 + ct] = ev * se;
                outv[(b * nz + row) * 2 + ct] = vv0 * sv0 + vv1 * sv1;
            }
        }
    }
    __syncthreads();
    const double2* ov = (const double2*)outv;
    const double2* oe = (const double2*)oute;
    const long po = ((long)ky0 * d.hnx + kx) * nz, pstride = (long)d.hnx * nz;
    for (int i = tid; i < nb * nz; i += NT) {
        const int bb = i / nz, row = i - bb * nz;
        NLv[po + bb * pstride + row] = ov[i];
        NLe[po + bb * pstride + row] = oe[i];
    }
}
