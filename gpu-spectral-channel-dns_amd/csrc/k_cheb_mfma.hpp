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
// Padded s2p of the six spectral fields (mixed form): F[f][P][nz] (FP64, or FP32 when the
// velocity recovery stores it so) -> Y[f][kx][q][ky] (FP32), q = 0..2N.
// f_q = sum_k a_k cos(pi k q / (2N)); with E_q = sum_{k even} a_k cos(pi k q/(2N)) (q <= N)
// and O_q the odd sum (q < N): f_q = E_q + O_q, f_{2N-q} = E_q - O_q, f_N = E_N.
// Block: 16 pencils (ky0.., kx fixed) of one field (blockIdx.z = member * 6 + field), 256
// threads.  Thread (b, m) = (tid / 16, tid % 16) loads the coefficients k = 16 m .. 16 m + 15
// of pencil b (a contiguous 256- or 128-byte run), takes the pencil's largest magnitude
// over its 16 lanes (shuffles, no LDS atomics) and writes the even and the odd eight of its
// run, split and scaled, as 16-byte rows of the operand tiles (zeros past nz and for
// missing pencils, so every tile entry a product reads is written).  Wave w then takes the
// row tiles w, w + 4, .. of E and the same row tiles of O.  tab: the E fragments, then the
// O fragments.
// -------------------------------------------------------------------------------------
template <int FN, typename TI = double2>
__global__ __launch_bounds__(256) void k_s2p_mfma(const TI* __restrict__ F, long fstride, long sF,
                                                  float2* __restrict__ Y, long ystride, long sY, Dims d,
                                                  const h8_t* __restrict__ tab, const int* __restrict__ ids) {
    using namespace cheb_mfma;
    using Sh = Shape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, pz = Sh::pz;
    static_assert(nz <= 16 * 16 && 8 * 16 >= Sh::KSE * 32, "one 16-coefficient run per lane");
    __shared__ __align__(16) _Float16 bE[2][kCols * Sh::LDE];  // [hi, lo][column][even k / 2]
    __shared__ __align__(16) _Float16 bO[2][kCols * Sh::LDO];  // [hi, lo][column][(odd k - 1) / 2]
    __shared__ int sexp[kPencils];
    const int slot = ids[blockIdx.z / 6], field = blockIdx.z % 6;
    F += slot * sF + field * fstride;
    Y += slot * sY + field * ystride;
    const int kx = blockIdx.y, ky0 = blockIdx.x * kPencils;
    const int nb = min(kPencils, d.ny - ky0);
    const int tid = threadIdx.x;
    {
        const int b = tid >> 4, m = tid & 15, k0 = 16 * m;
        const bool have = b < nb;
        TI a[16];
        const TI* src = F + ((long)(ky0 + b) * d.hnx + kx) * nz + k0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            a[i].x = 0;
            a[i].y = 0;
            if (have && k0 + i < nz) a[i] = src[i];
        }
        float mx = 0.f;
#pragma unroll
        for (int i = 0; i < 16; i++) mx = fmaxf(mx, (float)fmax(fabs((double)a[i].x), fabs((double)a[i].y)));
        mx = fmaxf(mx, __shfl_xor(mx, 1));
        mx = fmaxf(mx, __shfl_xor(mx, 2));
        mx = fmaxf(mx, __shfl_xor(mx, 4));
        mx = fmaxf(mx, __shfl_xor(mx, 8));
        const int e = scale_exp(mx);
        if (m == 0) sexp[b] = e;
        // runs j = 8 m .. 8 m + 7 of the even (k = 2 j) and odd (k = 2 j + 1) coefficients
        h8_t erh, erl, eih, eil, orh, orl, oih, oil;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const double xr = ldexp((double)a[2 * i].x, e), xi = ldexp((double)a[2 * i].y, e);
            const double yr = ldexp((double)a[2 * i + 1].x, e), yi = ldexp((double)a[2 * i + 1].y, e);
            erh[i] = (_Float16)(float)xr;
            erl[i] = (_Float16)(float)(xr - (double)(float)erh[i]);
            eih[i] = (_Float16)(float)xi;
            eil[i] = (_Float16)(float)(xi - (double)(float)eih[i]);
            orh[i] = (_Float16)(float)yr;
            orl[i] = (_Float16)(float)(yr - (double)(float)orh[i]);
            oih[i] = (_Float16)(float)yi;
            oil[i] = (_Float16)(float)(yi - (double)(float)oih[i]);
        }
        const int j0 = 8 * m;
        if (j0 < Sh::KSE * 32) {
            *(h8_t*)&bE[0][b * Sh::LDE + j0] = erh;
            *(h8_t*)&bE[1][b * Sh::LDE + j0] = erl;
            *(h8_t*)&bE[0][(kPencils + b) * Sh::LDE + j0] = eih;
            *(h8_t*)&bE[1][(kPencils + b) * Sh::LDE + j0] = eil;
        }
        if (j0 < Sh::KSO * 32) {
            *(h8_t*)&bO[0][b * Sh::LDO + j0] = orh;
            *(h8_t*)&bO[1][b * Sh::LDO + j0] = orl;
            *(h8_t*)&bO[0][(kPencils + b) * Sh::LDO + j0] = oih;
            *(h8_t*)&bO[1][(kPencils + b) * Sh::LDO + j0] = oil;
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
// Block: 16 pencils (kx fixed) of one member (blockIdx.z); two waves per row tile of the
// p <= N/2 rows (128 * RTS threads), wave 2 rt + c holding the real (c = 0) or imaginary
// (c = 1) parts of the row tile rt of every product: after the per-pencil combinations
// the products act on real and imaginary parts alike.  Two stages: h3 and h2 (from f_x,
// f_y) into two operand sets, then h1 (from f_z, loaded during the first products).  Each
// stage: fold and combine in FP32, one power-of-two scale per pencil and input (largest
// magnitude through LDS), split into LDS, products.  The results go through LDS to
// coalesced rows of N_v, N_eta (FP64) in the slot's current history buffers.
// tab: fragments of Q0s, Q0a, Q1s, Q1a (in this order); s0, s1: scales of the M0 and M1 tables.
// LDS (dynamic): two operand sets [set][hi, lo][32 columns][LDP (G+) | LDM (G-)] halves.
// -------------------------------------------------------------------------------------
template <int FN>
constexpr size_t p2s_mfma_lds() {
    using Sh = cheb_mfma::Shape<FN>;
    return sizeof(_Float16) * 2 * 2 * cheb_mfma::kCols * (Sh::LDP + Sh::LDM);
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
    {
        // zero both sets once: padding k and ragged columns stay zero, the rest is overwritten
        uint4* z = (uint4*)lds;
        for (int i = tid; i < 2 * SET * 2 / 16; i += NT) z[i] = make_uint4(0, 0, 0, 0);
        if (tid < 3 * kPencils) (&smax[0][0])[tid] = 0u;
    }
    // pencil scalars of pencil b: alpha, beta, k^2; true for pencil (0,0)
    auto scal = [&](int b, float* ia, float* ib, float* k2) {
        const int ky = ky0 + b;
        const double a_ = (double)kx / aphi;
        const double b_ = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / beta : (double)ky / beta;
        *ia = (float)a_;
        *ib = (float)b_;
        *k2 = (float)(a_ * a_ + b_ * b_);
        return has_mean && b == 0;
    };
    // samples q and 2N - q of component g for every item (q = i / nb, b = i % nb)
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
    auto split_set = [&](int st, int h, const float2* hp, const float2* hm) {
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb, e = scale_exp(__uint_as_float(smax[h][b]));
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
    // acc += table (row tile rt, ks k-steps) x operand columns of this wave's part ct
    auto prod = [&](const h8_t* t, int ks, const _Float16* bh, const _Float16* bl, int ld, f4_t acc) {
        const int col = ct * kPencils + lc;
#pragma unroll
        for (int s = 0; s < ks; s++) {
            const h8_t ah = t[((rt * ks + s) * 2 + 0) * 64 + lane];
            const h8_t al = t[((rt * ks + s) * 2 + 1) * 64 + lane];
            const int o = col * ld + s * 32 + lk;
            acc = mma3(ah, al, *(const h8_t*)&bh[o], *(const h8_t*)&bl[o], acc);
        }
        return acc;
    };
    const h8_t* t0s = tab;
    const h8_t* t0a = t0s + 2 * 64 * Sh::F0S;
    const h8_t* t1s = t0a + 2 * 64 * Sh::F0A;
    const h8_t* t1a = t1s + 2 * 64 * Sh::F1S;
    const f4_t zero4 = {0, 0, 0, 0};
    f4_t es = zero4, ea = zero4, v1s = zero4, v1a = zero4, v0s = zero4, v0a = zero4;
    float2 za[KQ], zb[KQ];
    __syncthreads();  // tiles zeroed, maxima cleared
    {
        // stage 1: h3 = i alpha f_y - i beta f_x (pencil (0,0): f_y) -> set 0;
        //          h2 = i alpha f_x + i beta f_y (pencil (0,0): 0)   -> set 1
        float2 h3p[KQ], h3m[KQ], h2p[KQ], h2m[KQ];
        {
            float2 xa[KQ], xb[KQ], ya[KQ], yb[KQ];
            load(0, xa, xb);
            load(1, ya, yb);
            float m3[KQ], m2[KQ];
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                m3[u] = m2[u] = 0.f;
                const int i = tid + u * NT;
                if (i < total) {
                    const int q = i / nb, b = i - q * nb;
                    float ia, ib, k2;
                    const bool mean = scal(b, &ia, &ib, &k2);
                    float2 xp, xm, yp, ym;
                    fold(xa[u], xb[u], q, &xp, &xm);
                    fold(ya[u], yb[u], q, &yp, &ym);
                    // i a y - i b x = (-a y.im + b x.im, a y.re - b x.re)
                    h3p[u] = mean ? yp : make_float2(-ia * yp.y + ib * xp.y, ia * yp.x - ib * xp.x);
                    h3m[u] = mean ? ym : make_float2(-ia * ym.y + ib * xm.y, ia * ym.x - ib * xm.x);
                    // i a x + i b y = (-a x.im - b y.im, a x.re + b y.re)
                    h2p[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xp.y - ib * yp.y, ia * xp.x + ib * yp.x);
                    h2m[u] = mean ? make_float2(0.f, 0.f) : make_float2(-ia * xm.y - ib * ym.y, ia * xm.x + ib * ym.x);
                    m3[u] = mag(h3p[u], h3m[u]);
                    m2[u] = mag(h2p[u], h2m[u]);
                }
            }
            pencil_max(0, m3);
            pencil_max(1, m2);
        }
        __syncthreads();
        split_set(0, 0, h3p, h3m);
        split_set(1, 1, h2p, h2m);
    }
    load(2, za, zb);  // f_z for h1, in flight during the first products
    __syncthreads();
    // N_eta: sym rows Q0s G+, anti rows Q0a G- of h3 (set 0); N_v (M1 part): sym rows Q1s G-,
    // anti rows Q1a G+ of h2 (set 1)
    es = prod(t0s, Sh::KSP, gp_hi(0), gp_lo(0), Sh::LDP, es);
    v1s = prod(t1s, Sh::KSM, gm_hi(1), gm_lo(1), Sh::LDM, v1s);
    if (anti) {
        ea = prod(t0a, Sh::KSM, gm_hi(0), gm_lo(0), Sh::LDM, ea);
        v1a = prod(t1a, Sh::KSP, gp_hi(1), gp_lo(1), Sh::LDP, v1a);
    }
    {
        // stage 2: h1 = k^2 f_z (pencil (0,0): f_x) -> set 0
        float2 h1p[KQ], h1m[KQ];
        float m1[KQ];
        const float2* x0 = Z1 + (long)kx * pz * d.ny + ky0;  // f_x of pencil (0,0)
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            m1[u] = 0.f;
            const int i = tid + u * NT;
            if (i < total) {
                const int q = i / nb, b = i - q * nb;
                float ia, ib, k2;
                const bool mean = scal(b, &ia, &ib, &k2);
                float2 zp, zm;
                if (mean) fold(x0[(long)q * d.ny], x0[(long)(FN - q) * d.ny], q, &zp, &zm);
                else fold(za[u], zb[u], q, &zp, &zm);
                h1p[u] = mean ? zp : make_float2(k2 * zp.x, k2 * zp.y);
                h1m[u] = mean ? zm : make_float2(k2 * zm.x, k2 * zm.y);
                m1[u] = mag(h1p[u], h1m[u]);
            }
        }
        pencil_max(2, m1);
        __syncthreads();  // also: set 0 is no longer read by the stage-1 products
        split_set(0, 2, h1p, h1m);
    }
    __syncthreads();
    v0s = prod(t0s, Sh::KSP, gp_hi(0), gp_lo(0), Sh::LDP, v0s);
    if (anti) v0a = prod(t0a, Sh::KSM, gm_hi(0), gm_lo(0), Sh::LDM, v0a);
    // combine: lane holds part ct of pencil b = lc, rows 16 rt + 4 (lane / 16) + r; results go
    // to LDS as [N_v, N_eta][pencil][row] complex doubles, then out in rows of nz
    const int b = lc;
    const float e3 = (float)__uint_as_float(smax[0][b]), e2 = __uint_as_float(smax[1][b]), e1 = __uint_as_float(smax[2][b]);
    const double se = ldexp(1.0, -scale_exp(e3) - s0), sv1 = ldexp(1.0, -scale_exp(e2) - s1),
                 sv0 = ldexp(1.0, -scale_exp(e1) - s0);
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
