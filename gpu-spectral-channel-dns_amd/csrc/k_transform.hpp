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

// Pencils per block of the Chebyshev kernels: 8 x 16 B = one 128-byte line per point row
// of the [kx][q][ky] layouts.  Columns per block of the span kernels.
constexpr int kChebTile = 8;
// Pencils per block of k_cheb_s2p for a compiled 2N: shorter transforms take wider tiles,
// so that a block keeps about the same work (and LDS) as at 2N = 256.
// Pencils per block of the kernels with the derivative scan (k_velocity_tile, k_p2s_nl2):
// 16 (scan with 4 chunks per class) for 2N <= 128, else 8.
constexpr int cheb_nl_tile_max(int fn) { return fn > 0 && fn <= 128 ? 2 * kChebTile : kChebTile; }
constexpr int s2p_tile_max(int fn) { return fn >= 192 || fn == 0 ? kChebTile : fn >= 96 ? 2 * kChebTile : 4 * kChebTile; }
constexpr int kSpanTile = 16;
constexpr int kLambPairs = 2;  // pairs of span rows per block of k_lamb_phys
// ... and their bounds for a compiled transform length: shorter rows and columns take wider
// tiles, so that a block keeps about the work and LDS it has at length 192.
constexpr int span_tile_max(int fn) { return fn >= 192 || fn == 0 ? kSpanTile : 2 * kSpanTile; }
constexpr int lamb_pairs_max(int fn) { return fn >= 192 || fn == 0 ? kLambPairs : fn >= 96 ? 2 * kLambPairs : 4 * kLambPairs; }

template <int V> struct IntC { static constexpr int value = V; };

// complex of the storage precision of the transform intermediates (FP64, or FP32 in the
// mixed-precision form), and conversions at the FP64 boundaries
template <typename R> using cplx = typename cplx_t<R>::type;
template <typename C> __device__ __forceinline__ C to_c(double2 a) { return cmake<C>(a.x, a.y); }
template <typename C> __device__ __forceinline__ double2 to_d2(C a) { return make_double2(a.x, a.y); }

// Move `total` elements through registers with K loads in flight per thread before any is
// consumed: a loop that loads, waits and stores one element at a time pays one memory
// latency per element.
template <int K, typename Load, typename Use>
__device__ __forceinline__ void staged(int total, int tid, int nthr, Load load, Use use) {
    using V = decltype(load(0));
    for (int base = tid; base < total; base += K * nthr) {
        V v[K];
#pragma unroll
        for (int u = 0; u < K; u++) {
            const int i = base + u * nthr;
            if (i < total) v[u] = load(i);
        }
#pragma unroll
        for (int u = 0; u < K; u++) {
            const int i = base + u * nthr;
            if (i < total) use(i, v[u]);
        }
    }
}

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
// blockIdx.z = member * 6 + field; per-slot arrays at F + slot*sF, Y + slot*sY.
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

template <int PAD, int FN = 0, typename RO = double>  // FN: compile-time 2N or 0; RO: precision of Y
__global__ __launch_bounds__(256) void k_cheb_s2p(const double2* __restrict__ F, long fstride, long sF,
                                                  cplx<RO>* __restrict__ Y, long ystride, long sY, Dims d, FftPlan pl,
                                                  const double2* __restrict__ tw4, int TB, int ld,
                                                  const int* __restrict__ ids) {
    RPCF_LDS(double2);
    const int f = blockIdx.z % 6, kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const long slot = ids[blockIdx.z / 6];
    F += slot * sF;
    Y += slot * sY;
    const int nb_rt = min(TB, d.ny - ky0);
    // A full tile of a compiled length fixes the tile width, nz and the LDS stride (2N + 2),
    // so the index arithmetic reduces to constant multiplies and shifts; offsets inside a
    // slot's field fit 32 bits.
    auto body = [&](auto nbc) {
        constexpr int NB = decltype(nbc)::value;
        const int nb = NB ? NB : nb_rt;
        const int N = FN ? FN / 2 : d.N, nz = N + 1, nout = PAD ? 2 * N + 1 : nz;
        const int ldl = FN ? FN + 2 : ld;
        const int tid = threadIdx.x, nthr = blockDim.x, ny = d.ny;
        const int pstride = d.hnx * nz;
        const double2* in = F + (f * fstride + (ky0 * d.hnx + kx) * nz);
        cplx<RO>* out = Y + (f * ystride + kx * nout * ny + ky0);
        const int total = nb * nz;
        auto load = [&](int i) {
            const unsigned b = (unsigned)i / nz, k = i - b * nz;
            return in[b * pstride + k];
        };
        auto ext = [&](int half, int i, double2 a) {
            const unsigned b = (unsigned)i / nz, k = i - b * nz;
            if (half == 0) s2p_even_ext(lds + b * ldl, a, k, N);
            else s2p_odd_ext(lds + b * ldl, a, k, N, tw4);
        };
        // with a compiled length and TB <= s2p_tile_max(FN) the tile's coefficients stay in
        // registers for both halves
        constexpr int TBX = s2p_tile_max(FN);
        constexpr int KC = FN ? (TBX * (FN / 2 + 1) + 255) / 256 : 1;
        const bool inreg = FN > 0 && total <= KC * nthr;
        double2 a[KC];
        if (inreg) {
#pragma unroll
            for (int u = 0; u < KC; u++) {
                const int i = tid + u * nthr;
                if (i < total) a[u] = load(i);
            }
        }
#pragma unroll
        for (int half = 0; half < (PAD ? 2 : 1); half++) {
            if (inreg) {
#pragma unroll
                for (int u = 0; u < KC; u++) {
                    const int i = tid + u * nthr;
                    if (i < total) ext(half, i, a[u]);
                }
            } else {
                staged<8>(total, tid, nthr, load, [&](int i, double2 v) { ext(half, i, v); });
            }
            __syncthreads();
            fft_any<FN, -1, TBX, 0>(lds, ldl, nb, pl, tid, nthr);
            // even half: points q = m (PAD=0) or q = 2m (PAD=1), m = 0..N; odd half: q = 2m+1
            const int cnt = half ? N : N + 1;
            staged<8>(
                cnt * nb, tid, nthr,
                [&](int i) {
                    const unsigned m = (unsigned)i / nb, b = i - m * nb;
                    return lds[b * ldl + m];
                },
                [&](int i, double2 v) {
                    const unsigned m = (unsigned)i / nb, b = i - m * nb;
                    out[(PAD ? 2 * m + half : m) * ny + b] = to_c<cplx<RO>>(v);
                });
            __syncthreads();
        }
    };
    if (FN > 0 && nb_rt == s2p_tile_max(FN)) body(IntC<s2p_tile_max(FN)>{});
    else body(IntC<0>{});
}

// LDS elements (double2) of k_cheb_s2p_h for a tile of tb pencils
__host__ __device__ constexpr size_t s2p_h_lds(int fn, int tb) {
    return 2 * (size_t)tb * (fn / 2 + 2) + (size_t)tb * (fn / 4 + 2) + ct_tw_lds_len(fn / 2) + ct_tw_lds_len(fn / 4) +
           fn / 2;
}

// -------------------------------------------------------------------------------------
// k_cheb_s2p (PAD = 1) for compiled 2N >= 64 with half-length transforms.  With a_k, k <= N,
// the padded values f_q = sum_k a_k cos(pi k q/(2N)), q = 0..2N, split as
//   q odd, q = 2m+1:  O_m = DCT-III_N(a_0..a_{N-1})_m                 (a_N drops out)
//   q = 4p:           E_2p = DCT-I_{N/2+1}(b)_p,  b_0 = a_0 + a_N, b_k = a_k + a_{N-k}, b_{N/2} = a_{N/2}
//   q = 4p+2:         E_2p+1 = DCT-III_{N/2}(c)_p, c_k = a_k - a_{N-k}
// DCT-I_{M+1}(b) is the length-2M FFT of the even extension (s_0 = b_0, s_M = b_M,
// s_k = s_{2M-k} = b_k/2).  DCT-III_M(X)_n = sum_{k<M} X_k cos(pi k (2n+1)/(2M)) for complex X
// is one length-M inverse FFT (Makhoul's reordering): Z_0 = 2 X_0,
// Z_k = e^{i pi k/(2M)} (X_k - i X_{M-k}); z = sum_k Z_k e^{2 pi i k n/M}; then
// x_2n = z_n / 2, x_2n+1 = z_{M-1-n} / 2.  FFT work per pencil: N + N + N/2 points in place of
// 2N + 2N.  The twiddle tables of lengths N and N/2 are taken from tw4 (exp(-2 pi i k/(4N)))
// into LDS.  LDS: A [2 TB][N+2] (rows [0, nb): Z of the odd points; rows [nb, 2nb): the even
// extension of b; rows [TB, TB+nb) stage the coefficients first), B [TB][N/2+2], tables.
// -------------------------------------------------------------------------------------
template <int FN, typename RO = double>
__global__ __launch_bounds__(256) void k_cheb_s2p_h(const double2* __restrict__ F, long fstride, long sF,
                                                    cplx<RO>* __restrict__ Y, long ystride, long sY, Dims d,
                                                    const double2* __restrict__ tw4, int TB,
                                                    const int* __restrict__ ids) {
    static_assert(FN >= 64 && FN % 8 == 0, "2N >= 64, N divisible by 4");
    RPCF_LDS(double2);
    constexpr int N = FN / 2, H = N / 2, nz = N + 1, pz = FN + 1, LA = N + 2, LB = H + 2;
    constexpr int TBX = s2p_tile_max(FN);
    const int f = blockIdx.z % 6, kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const long slot = ids[blockIdx.z / 6];
    F += slot * sF;
    Y += slot * sY;
    const int nb_rt = min(TB, d.ny - ky0), ny = d.ny, tid = threadIdx.x, nthr = blockDim.x;
    // a full tile of width s2p_tile_max(FN) makes the tile width a constant
    auto body = [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    const int nb = NB ? NB : nb_rt;
    double2* A = lds;
    double2* B = A + 2 * TB * LA;
    double2* twN = B + TB * LB;  // the LDS twiddle copies of the length N and N/2 plans
    double2* twH = twN + ct_tw_lds_len(N);
    double2* pre = twH + ct_tw_lds_len(H);  // tw4[k], k < N (conjugated where used)
    for (int i = tid; i < N; i += nthr) pre[i] = tw4[i];
    for (int i = tid; i < ct_tw_lds_len(N); i += nthr) twN[i] = tw4[4 * ct_tw_lds_index(N, i)];
    for (int i = tid; i < ct_tw_lds_len(H); i += nthr) twH[i] = tw4[8 * ct_tw_lds_index(H, i)];
    const int pstride = d.hnx * nz;
    const double2* in = F + (f * fstride + (ky0 * d.hnx + kx) * nz);
    double2* stg = A + TB * LA;
    staged<8>(
        nb * nz, tid, nthr,
        [&](int i) {
            const unsigned b = (unsigned)i / nz, k = i - b * nz;
            return in[b * pstride + k];
        },
        [&](int i, double2 v) {
            const unsigned b = (unsigned)i / nz, k = i - b * nz;
            stg[b * LA + k] = v;
        });
    __syncthreads();
    // Z of the odd points (M = N) and Z' of the points 4p+2 (M = N/2); the even extension
    // of b into registers (its rows overlap the staging area)
    for (int i = tid; i < nb * N; i += nthr) {
        const unsigned b = (unsigned)i / N, k = i - b * N;
        const double2* a = stg + b * LA;
        double2 z;
        if (k == 0) {
            z = make_double2(2.0 * a[0].x, 2.0 * a[0].y);
        } else {
            const double2 p = a[k], m = a[N - k], w = pre[k];
            const double2 x = make_double2(p.x + m.y, p.y - m.x);  // X_k - i X_{N-k}
            z = make_double2(w.x * x.x + w.y * x.y, w.x * x.y - w.y * x.x);  // conj(w) x
        }
        A[b * LA + k] = z;
    }
    for (int i = tid; i < nb * H; i += nthr) {
        const unsigned b = (unsigned)i / H, k = i - b * H;
        const double2* a = stg + b * LA;
        double2 z;
        if (k == 0) {
            z = make_double2(2.0 * (a[0].x - a[N].x), 2.0 * (a[0].y - a[N].y));
        } else {
            const double2 ck = make_double2(a[k].x - a[N - k].x, a[k].y - a[N - k].y);
            const double2 cm = make_double2(a[H - k].x - a[H + k].x, a[H - k].y - a[H + k].y);
            const double2 x = make_double2(ck.x + cm.y, ck.y - cm.x);
            const double2 w = pre[2 * k];
            z = make_double2(w.x * x.x + w.y * x.y, w.x * x.y - w.y * x.x);
        }
        B[b * LB + k] = z;
    }
    constexpr int KS = (TBX * N + 255) / 256;  // even-extension elements per thread
    double2 sv[KS];
#pragma unroll
    for (int u = 0; u < KS; u++) {
        const int i = tid + u * nthr;
        if (i < nb * N) {
            const unsigned b = (unsigned)i / N, k = i - b * N;
            const double2* a = stg + b * LA;
            if (k == 0) sv[u] = make_double2(a[0].x + a[N].x, a[0].y + a[N].y);
            else if (k == H) sv[u] = a[H];
            else sv[u] = make_double2(0.5 * (a[k].x + a[N - k].x), 0.5 * (a[k].y + a[N - k].y));
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KS; u++) {
        const int i = tid + u * nthr;
        if (i < nb * N) {
            const unsigned b = (unsigned)i / N, k = i - b * N;
            A[(nb + b) * LA + k] = sv[u];
        }
    }
    __syncthreads();
    FftPlan pn{}, ph{};
    pn.tw = twN;
    ph.tw = twH;
    fft_any<N, +1, 2 * TBX, 1, true>(A, LA, 2 * nb, pn, tid, nthr);
    fft_any<H, +1, TBX, 1, true>(B, LB, nb, ph, tid, nthr);
    cplx<RO>* out = Y + (f * ystride + kx * pz * ny + ky0);
    staged<8>(
        pz * nb, tid, nthr,
        [&](int i) {
            const unsigned q = (unsigned)i / nb, b = i - q * nb;
            if (q & 1) {
                const unsigned m = q >> 1;
                const double2 z = A[b * LA + ((m & 1) ? N - 1 - (m >> 1) : (m >> 1))];
                return make_double2(0.5 * z.x, 0.5 * z.y);
            }
            const unsigned p = q >> 2;
            if ((q & 3) == 0) return A[(nb + b) * LA + p];
            const double2 z = B[b * LB + ((p & 1) ? H - 1 - (p >> 1) : (p >> 1))];
            return make_double2(0.5 * z.x, 0.5 * z.y);
        },
        [&](int i, double2 v) {
            const unsigned q = (unsigned)i / nb, b = i - q * nb;
            out[q * ny + b] = to_c<cplx<RO>>(v);
        });
    };
    if (nb_rt == TBX) body(IntC<TBX>{});
    else body(IntC<0>{});
}

// -------------------------------------------------------------------------------------
// Span inverse FFT with the 3/2 zero padding: the span C2C half of the reference's 2D
// Z2D (source/transform.cu:176) after transpose_backward + setZeros
// (source/transpose.cu:358-381, source/transform.cu:355-386).  blockIdx.z = member * 6 +
// field.  Tile: TK kx columns of one
// plane.  Y[f][kx][q][ky] -> W[f][q][s][kx] for kx < nxh (the x-Nyquist column is zero).
// -------------------------------------------------------------------------------------
template <int FN = 0, typename R = double>  // R: precision of Y, W and the FFT
__global__ __launch_bounds__(256) void k_span_inv(const cplx<R>* __restrict__ Y, long ystride, long sY, cplx<R>* __restrict__ W,
                                                  long wstride, long sW, Dims d, int planes, FftPlan pl, int TK, int ld,
                                                  const int* __restrict__ ids) {
    using C2 = cplx<R>;
    RPCF_LDS(C2);
    const int f = blockIdx.z % 6, q = blockIdx.y, kx0 = blockIdx.x * TK;
    const long slot = ids[blockIdx.z / 6];
    Y += slot * sY;
    W += slot * sW;
    const int nk_rt = min(TK, d.nxh - kx0);
    // a full tile of a compiled length: constant tile width, length and LDS stride (my + 1);
    // offsets inside a slot's field fit 32 bits
    auto body = [&](auto nkc) {
        constexpr int NK = decltype(nkc)::value;
        const int nk = NK ? NK : nk_rt;
        const int my = FN ? FN : d.my, ldl = FN ? FN + 1 : ld, tid = threadIdx.x, nthr = blockDim.x;
        const int ny = d.ny, nxh = d.nxh;
        const C2* in = Y + f * ystride;
        staged<12>(
            nk * my, tid, nthr,
            [&](int i) {
                const unsigned kxi = (unsigned)i / my, s = i - kxi * my;
                const int ky = ky_of_span(d, s);
                return ky >= 0 ? in[((kx0 + kxi) * planes + q) * ny + ky] : cmake<C2>(0, 0);
            },
            [&](int i, C2 v) {
                const unsigned kxi = (unsigned)i / my, s = i - kxi * my;
                lds[kxi * ldl + s] = v;
            });
        FftPlan pll = pl;  // compiled lengths: twiddles from an LDS copy behind the tile
        if constexpr (FN > 0) {
            C2* twl = lds + TK * ldl;
            for (int i = tid; i < ct_tw_lds_len(FN); i += nthr) twl[i] = ((const C2*)pl.tw)[ct_tw_lds_src(FN, i)];
            pll.tw = twl;
        }
        __syncthreads();
        fft_any<FN, +1, span_tile_max(FN), 1, true>(lds, ldl, nk, pll, tid, nthr);
        C2* out = W + f * wstride;
        staged<12>(
            my * nk, tid, nthr,
            [&](int i) {
                const unsigned s = (unsigned)i / nk, kxi = i - s * nk;
                return lds[kxi * ldl + s];
            },
            [&](int i, C2 v) {
                const unsigned s = (unsigned)i / nk, kxi = i - s * nk;
                out[(q * my + s) * nxh + kx0 + kxi] = v;
            });
    };
    if (FN > 0 && nk_rt == span_tile_max(FN)) body(IntC<span_tile_max(FN)>{});
    else body(IntC<0>{});
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
template <typename C>
__device__ __forceinline__ void r2c_unpack(const C* z, int k, int mx, decltype(C().x) fac, C* xa, C* xb) {
    using T = decltype(C().x);
    const C zk = z[k], zm = z[k == 0 ? 0 : mx - k];
    const T h = T(0.5);
    *xa = cmake<C>((zk.x + zm.x) * h * fac, (zk.y - zm.y) * h * fac);
    *xb = cmake<C>((zk.y + zm.y) * h * fac, -(zk.x - zm.x) * h * fac);
}

// -------------------------------------------------------------------------------------
// Fused physical-space stage: x C2R of the six fields, Lamb vector + Coriolis term
// (computeLambDevice, source/nonlinear.cu:409-428, base flow at padded point q as in
// source/nonlinear.cu:400-403), x R2C of the three Lamb components and the
// 1/(mx my) normalisation of source/transform.cu:267.  Member blockIdx.z (its Ro and
// padded base profile meanU).  One block: one plane q, RP
// pairs of span rows.  W[6][pz][my][nxh] -> Z2[3][pz][my][hnx].  LDS: [6 RP][ld] sequences,
// then (compiled lengths) the FN twiddles.
// -------------------------------------------------------------------------------------
template <int FN = 0, typename R = double>  // R: precision of W, Z2, the FFTs and the products
__global__ __launch_bounds__(256) void k_lamb_phys(const cplx<R>* __restrict__ W, long wstride, long sW,
                                                   cplx<R>* __restrict__ Z2, long zstride, long sZ, int zld,
                                                   const double* __restrict__ meanU, long sPz, double fac, Dims d,
                                                   FftPlan pl, int RP, int ld, Members mb) {
    using C2 = cplx<R>;
    RPCF_LDS(C2);
    const int q = blockIdx.y, rp0 = blockIdx.x * RP;
    const int slot = mb.ids[blockIdx.z];
    const R Ro = (R)mb.mp[slot].Ro;
    W += slot * sW;
    Z2 += slot * sZ;
    meanU += slot * sPz;
    const int nrp_rt = min(RP, d.my / 2 - rp0);
    // The compiled lengths fix mx, nxh = (mx/3*2)/2, hnx = nxh + 1 and the LDS stride
    // mx + 1; a full tile fixes nrp = lamb_pairs_max(mx) -- the index arithmetic then reduces to
    // constant multiplies and shifts.  Offsets inside a slot fit 32 bits.
    auto body = [&](auto nc) {
        constexpr int NC = decltype(nc)::value;
        const int nrp = NC ? NC : nrp_rt;
        const int mx = FN ? FN : d.mx, nxh = FN ? (FN / 3 * 2) / 2 : d.nxh, hnx = nxh + 1;
        const int lds_ld = FN ? FN + 1 : ld;
        const int my = d.my, tid = threadIdx.x, nthr = blockDim.x;
        const int ws = (int)wstride, zs = (int)zstride;
        // packed spectra Z = X_A + i X_B of the row pairs (c2r_pack, element by element):
        // each kept coefficient pair (A_k, B_k), k < nxh, is loaded once and gives Z_k and
        // Z_{mx-k}; Z_k = 0 for nxh <= k <= mx - nxh
        constexpr int KL = 3;  // loads per thread in flight (6 * 2 * 64 / 256 at mx = 192)
        for (int base = tid; base < 6 * nrp * nxh; base += KL * nthr) {
            C2 va[KL], vb[KL];
#pragma unroll
            for (int u = 0; u < KL; u++) {
                const int i = base + u * nthr;
                if (i < 6 * nrp * nxh) {
                    const int seq = i / nxh, k = i - seq * nxh;
                    const int f = seq / nrp, ip = seq - f * nrp;
                    const C2* rA = W + (f * ws + (q * my + 2 * (rp0 + ip)) * nxh);
                    va[u] = rA[k];
                    vb[u] = rA[nxh + k];
                }
            }
#pragma unroll
            for (int u = 0; u < KL; u++) {
                const int i = base + u * nthr;
                if (i < 6 * nrp * nxh) {
                    const int seq = i / nxh, k = i - seq * nxh;
                    C2 a = va[u], b = vb[u];
                    C2* z = lds + seq * lds_ld;
                    if (k == 0) {
                        a.y = 0;
                        b.y = 0;
                    } else {
                        z[mx - k] = cmake<C2>(a.x + b.y, -a.y + b.x);
                    }
                    z[k] = cmake<C2>(a.x - b.y, a.y + b.x);
                }
            }
        }
        const int nzero = mx - 2 * nxh + 1;
        for (int i = tid; i < 6 * nrp * nzero; i += nthr) {
            const int seq = i / nzero, k = i - seq * nzero;
            lds[seq * lds_ld + nxh + k] = cmake<C2>(0, 0);
        }
        // compiled lengths: the FFT twiddles from an LDS copy behind the sequences, so that
        // no butterfly waits on a global load
        FftPlan pll = pl;
        if constexpr (FN > 0) {
            C2* twl = lds + 6 * RP * lds_ld;
            const C2* twg = (const C2*)pl.tw;
            for (int i = tid; i < ct_tw_lds_len(FN); i += nthr) twl[i] = twg[ct_tw_lds_src(FN, i)];
            pll.tw = twl;
        }
        __syncthreads();
        fft_any<FN, +1, 6 * lamb_pairs_max(FN), 1, true>(lds, lds_ld, 6 * nrp, pll, tid, nthr);
        // Ro U(y) is uniform in x and span, so it reaches only the (0,0) pencil of L_z, which
        // rhsNonlinear does not use: the FP32 form leaves it out (in FP32 its round-off would
        // swamp the perturbation products); the FP64 form keeps the reference's expression.
        const R mU = sizeof(R) == 8 ? (R)meanU[q] : R(0);
        for (int i = tid; i < nrp * mx; i += nthr) {
            const int ip = i / mx, x = i - ip * mx;
            const C2 u = lds[(0 * nrp + ip) * lds_ld + x], v = lds[(1 * nrp + ip) * lds_ld + x];
            const C2 w = lds[(2 * nrp + ip) * lds_ld + x], ox = lds[(3 * nrp + ip) * lds_ld + x];
            const C2 oy = lds[(4 * nrp + ip) * lds_ld + x], oz = lds[(5 * nrp + ip) * lds_ld + x];
            const R lxa = oz.x * v.x - oy.x * w.x - Ro * w.x, lxb = oz.y * v.y - oy.y * w.y - Ro * w.y;
            const R lya = ox.x * w.x - oz.x * u.x, lyb = ox.y * w.y - oz.y * u.y;
            const R lza = oy.x * u.x - ox.x * v.x + Ro * (u.x + mU), lzb = oy.y * u.y - ox.y * v.y + Ro * (u.y + mU);
            lds[(0 * nrp + ip) * lds_ld + x] = cmake<C2>(lxa, lxb);
            lds[(1 * nrp + ip) * lds_ld + x] = cmake<C2>(lya, lyb);
            lds[(2 * nrp + ip) * lds_ld + x] = cmake<C2>(lza, lzb);
        }
        __syncthreads();
        fft_any<FN, -1, 3 * lamb_pairs_max(FN), 1, true>(lds, lds_ld, 3 * nrp, pll, tid, nthr);
        for (int i = tid; i < 3 * nrp * hnx; i += nthr) {
            const int seq = i / hnx, k = i - seq * hnx;
            const int g = seq / nrp, ip = seq - g * nrp;
            C2 xa, xb;
            r2c_unpack(lds + seq * lds_ld, k, mx, (R)fac, &xa, &xb);
            C2* o = Z2 + (g * zs + (q * my + 2 * (rp0 + ip)) * zld);
            o[k] = xa;
            o[zld + k] = xb;
        }
    };
    if (FN > 0 && nrp_rt == lamb_pairs_max(FN)) body(IntC<lamb_pairs_max(FN)>{});
    else body(IntC<0>{});
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
__global__ __launch_bounds__(256) void k_x_r2c(const double* __restrict__ phys, double2* __restrict__ Z2, int zld, double fac, Dims d,
                        FftPlan pl, int RP, int ld) {
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
        double2* o = Z2 + ((long)q * my + 2 * (rp0 + ip)) * zld;
        o[k] = xa;
        o[zld + k] = xb;
    }
}

// -------------------------------------------------------------------------------------
// Span forward FFT + truncation to the ny kept modes: the span C2C half of the
// reference's D2Z (source/transform.cu:240) and transpose_forward
// (source/transpose.cu:223-257).  Z2[g][q][s][kx] -> Z1[g][kx][q][ky]; blockIdx.z =
// member * 3 + component.
// -------------------------------------------------------------------------------------
template <int FN = 0, typename R = double>  // R: precision of Z2, Z1 and the FFT
__global__ __launch_bounds__(256) void k_span_fwd(const cplx<R>* __restrict__ Z2, long zstride, long sZ, int zld,
                                                  cplx<R>* __restrict__ Z1, long ostride, long sY, Dims d, int planes,
                                                  FftPlan pl, int TK, int ld, const int* __restrict__ ids,
                                                  int* __restrict__ phase = nullptr) {
    using C2 = cplx<R>;
    RPCF_LDS(C2);
    const int g = blockIdx.z % 3, q = blockIdx.y, kx0 = blockIdx.x * TK;
    const long slot = ids[blockIdx.z / 3];
    // the AB2 history swap of the step (source/rhs.cpp:107-108), by one thread per member:
    // nothing in this kernel reads the phase, and the next kernel (the p2s) writes the new N
    // into the buffers the flipped phase names current
    if (phase && blockIdx.x == 0 && blockIdx.y == 0 && g == 0 && threadIdx.x == 0) phase[slot] ^= 1;
    Z2 += slot * sZ;
    Z1 += slot * sY;
    const int nk_rt = min(TK, d.hnx - kx0);
    // a full tile of a compiled length: constant tile width, length and LDS stride (my + 1);
    // offsets inside a slot's field fit 32 bits
    auto body = [&](auto nkc) {
        constexpr int NK = decltype(nkc)::value;
        const int nk = NK ? NK : nk_rt;
        const int my = FN ? FN : d.my, ldl = FN ? FN + 1 : ld, ny = d.ny, tid = threadIdx.x, nthr = blockDim.x;
        const C2* in = Z2 + g * zstride;
        staged<12>(
            my * nk, tid, nthr,
            [&](int i) {
                const unsigned s = (unsigned)i / nk, kxi = i - s * nk;
                return in[(q * my + s) * zld + kx0 + kxi];
            },
            [&](int i, C2 v) {
                const unsigned s = (unsigned)i / nk, kxi = i - s * nk;
                lds[kxi * ldl + s] = v;
            });
        FftPlan pll = pl;  // compiled lengths: twiddles from an LDS copy behind the tile
        if constexpr (FN > 0) {
            C2* twl = lds + TK * ldl;
            for (int i = tid; i < ct_tw_lds_len(FN); i += nthr) twl[i] = ((const C2*)pl.tw)[ct_tw_lds_src(FN, i)];
            pll.tw = twl;
        }
        __syncthreads();
        fft_any<FN, -1, span_tile_max(FN), 1, true>(lds, ldl, nk, pll, tid, nthr);
        C2* out = Z1 + g * ostride;
        staged<8>(
            nk * ny, tid, nthr,
            [&](int i) {
                const unsigned kxi = (unsigned)i / ny, ky = i - kxi * ny;
                return lds[kxi * ldl + span_of(d, ky)];
            },
            [&](int i, C2 v) {
                const unsigned kxi = (unsigned)i / ny, ky = i - kxi * ny;
                out[((kx0 + kxi) * planes + q) * ny + ky] = v;
            });
    };
    if (FN > 0 && nk_rt == span_tile_max(FN)) body(IntC<span_tile_max(FN)>{});
    else body(IntC<0>{});
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

// Chebyshev derivative of ddz_sm (source/RPCFKernels.cu:476-520), D_k = sum_{j = k+1,
// k+3, .. <= N} 2j a_j with D_0 halved, of the two coefficient arrays A_c = fb + b*ld + c*nz
// (c = 0, 1) of each of the nb <= 1024 / (4 NCH) pencils of a tile, in place.  D is a suffix
// sum over each parity class of j: two-level scan with NCH chunks per (pencil, array,
// class) on 256 threads (NCH = 8: 8 pencils, NCH = 4: 16 pencils), written over the
// coefficients of the other class (read before the barrier).  Callers synchronise before
// (arrays written) and after (D read).  tot: 256 LDS entries.
template <int CHM, int NCH = 8>  // CHM: chunk length bound, >= ceil((N/2 + 1) / NCH)
__device__ __forceinline__ void cheb_deriv_scan(double2* fb, int ld, int nz, int N, int nb, double2* tot, int tid) {
    static_assert(NCH == 4 || NCH == 8, "4 or 8 chunks per parity class");
    constexpr int SH = NCH == 8 ? 3 : 2;
    const int b = tid >> (SH + 2), c2 = (tid >> (SH + 1)) & 1, par = (tid >> SH) & 1, c = tid & (NCH - 1);
    const int Lc = (N - par) / 2 + 1;  // class size: j = par, par+2, .. <= N
    const int CH = (Lc + NCH - 1) / NCH;
    const int t0 = c * CH, t1 = min(t0 + CH, Lc);
    double2* A = fb + b * ld + c2 * nz;
    double2 cj[CHM];
    double sre = 0.0, sim = 0.0;
    if (b < nb) {
#pragma unroll
        for (int e = 0; e < CHM; e++) {
            const int t = t0 + e;
            if (t < t1) {
                const int j = 2 * t + par;
                const double dm = 2 * (double)j;
                cj[e] = make_double2(A[j].x * dm, A[j].y * dm);
                sre += cj[e].x;
                sim += cj[e].y;
            }
        }
    }
    tot[tid] = make_double2(sre, sim);
    __syncthreads();
    if (b < nb) {
        double rre = 0.0, rim = 0.0;
        for (int cc = NCH - 1; cc > c; cc--) {
            const double2 o = tot[(tid & ~(NCH - 1)) | cc];
            rre += o.x;
            rim += o.y;
        }
#pragma unroll
        for (int e = CHM - 1; e >= 0; e--) {
            const int t = t0 + e;
            if (t < t1) {
                const int j = 2 * t + par;
                rre += cj[e].x;
                rim += cj[e].y;
                if (j >= 1) A[j - 1] = j == 1 ? make_double2(rre * 0.5, rim * 0.5) : make_double2(rre, rim);
            }
        }
        if (c == 0 && par == (N & 1)) A[N] = make_double2(0.0, 0.0);  // no j > N
    }
}

// -------------------------------------------------------------------------------------
// Velocity recovery, getVelocityKernel_sm (source/velocity.cu:90-170), on tiles of
// TB <= kChebTile pencils (kx fixed) with the Chebyshev derivatives from cheb_deriv_scan
// (the reference sums each D_k directly, O(nz) per coefficient).  256 threads, member
// blockIdx.z; thread-owned elements (pencil b = i / nz, k = i % nz) as in k_p2s_nl.
// Pencil != (0,0):  u = (i a/k2) Dv - (i b/k2) eta,  w_span = (i b/k2) Dv + (i a/k2) eta,
//                   ox = D w_span - i b v,  oy = -D u + i a v,  oz = eta;
// pencil (0,0):     u = u0, w_span = w0, v = 0, ox = D w0, oy = -D u0, oz = 0.
// -------------------------------------------------------------------------------------
template <int FN, typename TO = double2>  // TO: storage of the six fields (FP32 for the mixed form's MFMA path)
__global__ __launch_bounds__(256) void k_velocity_tile(const double2* __restrict__ sv, const double2* __restrict__ se,
                                                       long sS, TO* __restrict__ F, long fstride, long sF, Dims d,
                                                       Members mb, int TB, int ld) {
    static_assert(FN > 0, "compiled lengths only");
    RPCF_LDS(double2);
    constexpr int N = FN / 2, nz = N + 1;
    constexpr int TBC = cheb_nl_tile_max(FN), NCH = TBC > kChebTile ? 4 : 8;
    constexpr int KC = (TBC * nz + 255) / 256;
    constexpr int CHM = (FN / 4 + 1 + NCH - 1) / NCH;
    const int slot = mb.ids[blockIdx.z];
    const MemberParams& mp = mb.mp[slot];
    sv += slot * sS;
    se += slot * sS;
    F += slot * sF;
    const int kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0);
    const int tid = threadIdx.x, nthr = blockDim.x, total = nb * nz;
    double2* fb = lds;
    double2* tot = lds + TB * ld;
    const long pstride = (long)d.hnx * nz;
    const long pin = ((long)ky0 * d.hnx + kx) * nz;
    double2 rv[KC], re[KC];
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            rv[u] = sv[pin + b * pstride + k];
            re[u] = se[pin + b * pstride + k];
        }
    }
    auto mean = [&](int i) { return kx == 0 && ky0 + i / nz == 0; };
    auto alpha_beta = [&](int b, double* ia, double* ib) {
        const int ky = ky0 + b;
        *ia = (double)kx / mp.aphi;
        *ib = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / mp.beta : (double)ky / mp.beta;
    };
    // D v
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            fb[b * ld + k] = rv[u];
        }
    }
    __syncthreads();
    cheb_deriv_scan<CHM, NCH>(fb, ld, nz, N, nb, tot, tid);
    __syncthreads();
    double2 su[KC], sw[KC];
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            if (mean(i)) {
                su[u] = rv[u];
                sw[u] = re[u];
            } else {
                double ia, ib;
                alpha_beta(b, &ia, &ib);
                const double kmn1 = 1.0 / (ia * ia + ib * ib);
                const double2 t = fb[b * ld + k];
                const double ca = ia * kmn1, cb = ib * kmn1;
                // (i ca) t - (i cb) eta and (i cb) t + (i ca) eta
                su[u] = make_double2((0.0 * t.x - ca * t.y) - (0.0 * re[u].x - cb * re[u].y),
                                     (0.0 * t.y + ca * t.x) - (0.0 * re[u].y + cb * re[u].x));
                sw[u] = make_double2((0.0 * t.x - cb * t.y) + (0.0 * re[u].x - ca * re[u].y),
                                     (0.0 * t.y + cb * t.x) + (0.0 * re[u].y + ca * re[u].x));
            }
        }
    }
    __syncthreads();
    // D u, D w_span
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            fb[b * ld + k] = su[u];
            fb[b * ld + nz + k] = sw[u];
        }
    }
    __syncthreads();
    cheb_deriv_scan<CHM, NCH>(fb, ld, nz, N, nb, tot, tid);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            const double2 du = fb[b * ld + k], dw = fb[b * ld + nz + k];
            double2 ox, oy, vv = rv[u], ee = re[u];
            if (mean(i)) {
                ox = dw;                               // D w0
                oy = make_double2(-du.x, -du.y);       // -D u0
                vv = ee = make_double2(0.0, 0.0);
            } else {
                double ia, ib;
                alpha_beta(b, &ia, &ib);
                const double nb_ = -1.0 * ib;
                ox = make_double2(dw.x + (0.0 * vv.x - nb_ * vv.y), dw.y + (0.0 * vv.y + nb_ * vv.x));
                oy = make_double2(du.x * -1.0 + (0.0 * vv.x - ia * vv.y), du.y * -1.0 + (0.0 * vv.y + ia * vv.x));
            }
            const long o = pin + b * pstride + k;
            F[0 * fstride + o] = to_c<TO>(su[u]);
            F[1 * fstride + o] = to_c<TO>(sw[u]);
            F[2 * fstride + o] = to_c<TO>(vv);
            F[3 * fstride + o] = to_c<TO>(ox);
            F[4 * fstride + o] = to_c<TO>(oy);
            F[5 * fstride + o] = to_c<TO>(ee);
        }
    }
}

// -------------------------------------------------------------------------------------
// Forward Chebyshev transform of the three Lamb components, rhsNonlinearKernel
// (source/nonlinear.cu:146-242) and the unpadded s2p of N_v and N_eta
// (source/rhs.cpp:96-97).  Tile: TB <= kChebTile pencils (kx fixed), 256 threads; member
// blockIdx.z.  Z1[3][hnx][pz][ny] -> the current N_v, N_eta [P][nz] of the slot's history
// NL[4][P][nz] (values at the nz collocation points; pencil (0,0) keeps L_x, L_y).
//
// The reference's p2s is a length-4N FFT of the even extension u of the pz values
// (cheby_pre_p2s/cheby_post_p2s, source/transform.cu:494-536).  Split by the parity of
// the sample index (one radix-2 decimation-in-time step):
//   F_k = E_k + w4^k H_k,  E = FFT_2N(even extension of f_0, f_2, .., f_2N),
//                          H = FFT_2N(f_1, f_3, .., f_{2N-1}, f_{2N-1}, .., f_1),
// w4 = exp(-2 pi i/(4N)); a_k = F_k/(2N), a_0 halved; only k < nz is kept.
//
// Thread-owned elements: i = tid + u*256 <-> (pencil b = i / nz, coefficient k = i % nz);
// the coefficients of the three components stay in registers; the derivatives of L_x and
// L_y come from cheb_deriv_scan.  LDS: FFT tile [TB][ld] (ld >= 2N+2, also holding L_x, L_y
// as [TB][2][nz]) and 256 chunk totals.
// -------------------------------------------------------------------------------------
template <int FN = 0, typename RI = double>  // FN: compile-time 2N or 0; RI: precision of Z1
__global__ __launch_bounds__(256) void k_p2s_nl(const cplx<RI>* __restrict__ Z1, long zstride, long sY,
                                                double2* __restrict__ NL, long sNL, Dims d, FftPlan plN,
                                                const double2* __restrict__ tw4, int TB, int ld, Members mb) {
    RPCF_LDS(double2);
    const int mslot = mb.ids[blockIdx.z];
    const int ph = mb.phase[mslot];
    const double aphi = mb.mp[mslot].aphi, beta = mb.mp[mslot].beta;
    const long PS = (long)d.P * d.nz;
    Z1 += mslot * sY;
    double2* NLv = NL + mslot * sNL + ph * PS;        // N_v, current
    double2* NLe = NL + mslot * sNL + (2 + ph) * PS;  // N_eta, current
    constexpr int KC = FN ? (kChebTile * (FN / 2 + 1) + 255) / 256 : (kChebTile * 257 + 255) / 256;
    constexpr int CHM = FN ? (FN / 4 + 1 + 7) / 8 : (128 + 1 + 7) / 8;  // chunk length bound
    const int kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0);
    const int N = FN ? FN / 2 : d.N, nz = N + 1, ny = d.ny, L = 2 * N;
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2* fb = lds;
    double2* tot = lds + TB * ld;
    const double factor = (1.0 / (d.pz - 1));
    const int total = nb * nz;
    double2 ck[3][KC];
    // one Lamb component g (compile-time, so that ck[g][.] stays in registers)
    auto component = [&](auto gc) {
        constexpr int g = decltype(gc)::value;
        const cplx<RI>* in = Z1 + g * zstride + (long)kx * d.pz * ny + ky0;
        for (int half = 0; half < 2; half++) {
            // samples f_{2m+half}: row 2m+half of the plane stack, nb consecutive ky
            staged<KC>(
                (half ? N : N + 1) * nb, tid, nthr,
                [&](int i) {
                    const int m = i / nb, b = i - m * nb;
                    return to_d2(in[(long)(2 * m + half) * ny + b]);
                },
                [&](int i, double2 v) {
                    const int m = i / nb, b = i - m * nb;
                    double2* sq = fb + b * ld;
                    sq[m] = v;
                    if (half == 0) {
                        if (m >= 1 && m <= N - 1) sq[L - m] = v;
                    } else {
                        sq[L - 1 - m] = v;
                    }
                });
            __syncthreads();
            fft_any<FN, -1, kChebTile>(fb, ld, nb, plN, tid, nthr);
#pragma unroll
            for (int u = 0; u < KC; u++) {
                const int i = tid + u * nthr;
                if (i < total) {
                    const int b = i / nz, k = i - b * nz;
                    const double2 x = fb[b * ld + k];
                    if (half == 0) {
                        ck[g][u] = x;
                    } else {
                        double2 t = cscale(cadd(ck[g][u], cmul(tw4[k], x)), factor);
                        if (k == 0) t = cscale(t, 0.5);
                        ck[g][u] = t;
                    }
                }
            }
            __syncthreads();
        }
    };
    component(IntC<0>{});
    component(IntC<1>{});
    component(IntC<2>{});
    // L_x, L_y -> LDS [b][0..1][nz]
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            fb[b * ld + k] = ck[0][u];
            fb[b * ld + nz + k] = ck[1][u];
        }
    }
    __syncthreads();
    cheb_deriv_scan<CHM>(fb, ld, nz, N, nb, tot, tid);
    __syncthreads();
    // rhsNonlinear: N_v, N_eta per (pencil, k) into the registers of L_x, L_y
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz, ky = ky0 + b;
            const double2 lx = ck[0][u], ly = ck[1][u], lz = ck[2][u];
            double2 nv, ne;
            if (kx == 0 && ky == 0) {
                nv = lx;
                ne = ly;
            } else {
                const double ia = (double)kx / aphi;
                const double ib = ky >= ny / 2 + 1 ? (double)(ky - ny) / beta : (double)ky / beta;
                const double kmn = ia * ia + ib * ib;
                const double2 dx = fb[b * ld + k], dy = fb[b * ld + nz + k];
                double ure = kmn * lz.x - ia * dx.y;
                double uim = kmn * lz.y + ia * dx.x;
                ure = ure - ib * dy.y;
                uim = uim + ib * dy.x;
                const double wre = -1.0 * ia * ly.y + ib * lx.y;
                const double wim = ia * ly.x - ib * lx.x;
                nv = make_double2(ure, uim);
                ne = make_double2(wre, wim);
            }
            ck[0][u] = nv;
            ck[1][u] = ne;
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int u = 0; u < KC; u++) {
            const int i = tid + u * nthr;
            if (i < total) {
                const int b = i / nz, k = i - b * nz;
                s2p_even_ext(fb + b * ld, ck[h][u], k, N);
            }
        }
        __syncthreads();
        fft_any<FN, -1, kChebTile>(fb, ld, nb, plN, tid, nthr);
        double2* out = h ? NLe : NLv;
        staged<8>(
            total, tid, nthr, [&](int i) {
                const int b = i / nz, p = i - b * nz;
                return fb[b * ld + p];
            },
            [&](int i, double2 v) {
                const int b = i / nz, p = i - b * nz;
                out[((long)(ky0 + b) * d.hnx + kx) * nz + p] = v;
            });
        __syncthreads();
    }
}

// -------------------------------------------------------------------------------------
// k_p2s_nl for the compiled lengths, with the work of a tile batched: the E and H sequences
// of a Lamb component (both parity halves of its pz samples) go through one FFT call over
// 2*TB sequences, as do N_v and N_eta at the end -- four FFT calls per tile instead of
// eight, each with a full block of butterflies.  The samples of component g+1 are loaded
// into registers while component g is transformed; the FFT and p2s twiddles are read from
// LDS copies, so that no wait for a global load of theirs waits for those samples too.
// Results are bit-identical to k_p2s_nl (same butterflies, same operation order).
// LDS: [2*TB][ld] sequences, 256 scan totals, FN FFT twiddles, nz p2s twiddles.
// -------------------------------------------------------------------------------------
// three waves per SIMD (168 registers instead of 172 at two waves) up to 2N = 256; the
// longer lengths keep more samples and coefficients per thread and run two waves, 2N = 512
// one (no spills)
constexpr int p2s_waves(int fn) { return fn <= 256 ? 3 : fn <= 384 ? 2 : 1; }
template <int FN, typename RI = double>
__global__ __launch_bounds__(256, p2s_waves(FN)) void k_p2s_nl2(const cplx<RI>* __restrict__ Z1, long zstride, long sY,
                                                 double2* __restrict__ NL, long sNL, Dims d, FftPlan plN,
                                                 const double2* __restrict__ tw4, int TB, int ld, Members mb) {
    static_assert(FN > 0, "compiled lengths only");
    RPCF_LDS(double2);
    constexpr int N = FN / 2, nz = N + 1, L = FN, pz = FN + 1;
    constexpr int TBC = cheb_nl_tile_max(FN), NCH = TBC > kChebTile ? 4 : 8;
    constexpr int KC = (TBC * nz + 255) / 256;
    constexpr int KL = (TBC * pz + 255) / 256;
    constexpr int CHM = (FN / 4 + 1 + NCH - 1) / NCH;
    const int mslot = mb.ids[blockIdx.z];
    const int ph = mb.phase[mslot];
    const double aphi = mb.mp[mslot].aphi, beta = mb.mp[mslot].beta;
    const long PS = (long)d.P * nz;
    Z1 += mslot * sY;
    double2* NLv = NL + mslot * sNL + ph * PS;        // N_v, current
    double2* NLe = NL + mslot * sNL + (2 + ph) * PS;  // N_eta, current
    const int kx = blockIdx.y, ky0 = blockIdx.x * TB;
    const int nb = min(TB, d.ny - ky0), ny = d.ny;
    const int tid = threadIdx.x, nthr = blockDim.x;
    double2* fb = lds;
    double2* tot = lds + 2 * TB * ld;
    double2* twl = tot + 256;
    double2* tw4l = twl + ct_tw_lds_len(FN);
    {
        const double2* twg = (const double2*)plN.tw;
        for (int i = tid; i < ct_tw_lds_len(FN); i += nthr) twl[i] = twg[ct_tw_lds_src(FN, i)];
        for (int i = tid; i < nz; i += nthr) tw4l[i] = tw4[i];
    }
    FftPlan pll = plN;
    pll.tw = twl;
    const double factor = (1.0 / (d.pz - 1));
    const int total = nb * nz, nsamp = nb * pz;
    const cplx<RI>* in0 = Z1 + (long)kx * d.pz * ny + ky0;
    // samples are held as loaded (FP32 in the mixed form) and widened where they are used, so
    // that the prefetch of the next component does not wait for its loads
    cplx<RI> smp[KL];
    auto load = [&](int g) {
        const cplx<RI>* in = in0 + g * zstride;
#pragma unroll
        for (int u = 0; u < KL; u++) {
            const int i = tid + u * nthr;
            if (i < nsamp) {
                const int row = i / nb, b = i - row * nb;
                smp[u] = in[(long)row * ny + b];
            }
        }
    };
    load(0);
    double2 ck[3][KC];
    auto component = [&](auto gc) {
        constexpr int g = decltype(gc)::value;
        // even rows f_2m -> E of pencil b (sequence b), odd rows f_2m+1 -> H (sequence nb+b)
#pragma unroll
        for (int u = 0; u < KL; u++) {
            const int i = tid + u * nthr;
            if (i < nsamp) {
                const int row = i / nb, b = i - row * nb, m = row >> 1;
                const double2 v = to_d2(smp[u]);
                if ((row & 1) == 0) {
                    double2* sq = fb + b * ld;
                    sq[m] = v;
                    if (m >= 1 && m <= N - 1) sq[L - m] = v;
                } else {
                    double2* sq = fb + (nb + b) * ld;
                    sq[m] = v;
                    sq[L - 1 - m] = v;
                }
            }
        }
        if constexpr (g < 2) load(g + 1);
        __syncthreads();
        fft_any<FN, -1, 2 * TBC, 1, true>(fb, ld, 2 * nb, pll, tid, nthr);
#pragma unroll
        for (int u = 0; u < KC; u++) {
            const int i = tid + u * nthr;
            if (i < total) {
                const int b = i / nz, k = i - b * nz;
                const double2 e = fb[b * ld + k], h = fb[(nb + b) * ld + k];
                double2 t = cscale(cadd(e, cmul(tw4l[k], h)), factor);
                if (k == 0) t = cscale(t, 0.5);
                ck[g][u] = t;
            }
        }
        __syncthreads();
    };
    component(IntC<0>{});
    component(IntC<1>{});
    component(IntC<2>{});
    // L_x, L_y -> LDS [b][0..1][nz] for the derivative scan
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            fb[b * ld + k] = ck[0][u];
            fb[b * ld + nz + k] = ck[1][u];
        }
    }
    __syncthreads();
    cheb_deriv_scan<CHM, NCH>(fb, ld, nz, N, nb, tot, tid);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz, ky = ky0 + b;
            const double2 lx = ck[0][u], ly = ck[1][u], lz = ck[2][u];
            double2 nv, ne;
            if (kx == 0 && ky == 0) {
                nv = lx;
                ne = ly;
            } else {
                const double ia = (double)kx / aphi;
                const double ib = ky >= ny / 2 + 1 ? (double)(ky - ny) / beta : (double)ky / beta;
                const double kmn = ia * ia + ib * ib;
                const double2 dx = fb[b * ld + k], dy = fb[b * ld + nz + k];
                double ure = kmn * lz.x - ia * dx.y;
                double uim = kmn * lz.y + ia * dx.x;
                ure = ure - ib * dy.y;
                uim = uim + ib * dy.x;
                const double wre = -1.0 * ia * ly.y + ib * lx.y;
                const double wim = ia * ly.x - ib * lx.x;
                nv = make_double2(ure, uim);
                ne = make_double2(wre, wim);
            }
            ck[0][u] = nv;
            ck[1][u] = ne;
        }
    }
    __syncthreads();
    // unpadded s2p of N_v (sequences b) and N_eta (sequences nb+b) in one FFT call
#pragma unroll
    for (int u = 0; u < KC; u++) {
        const int i = tid + u * nthr;
        if (i < total) {
            const int b = i / nz, k = i - b * nz;
            s2p_even_ext(fb + b * ld, ck[0][u], k, N);
            s2p_even_ext(fb + (nb + b) * ld, ck[1][u], k, N);
        }
    }
    __syncthreads();
    fft_any<FN, -1, 2 * TBC, 1, true>(fb, ld, 2 * nb, pll, tid, nthr);
    staged<8>(
        2 * total, tid, nthr,
        [&](int i) {
            const int h = i >= total, j = i - h * total, b = j / nz, p = j - b * nz;
            return fb[(h * nb + b) * ld + p];
        },
        [&](int i, double2 v) {
            const int h = i >= total, j = i - h * total, b = j / nz, p = j - b * nz;
            (h ? NLe : NLv)[((long)(ky0 + b) * d.hnx + kx) * nz + p] = v;
        });
}
