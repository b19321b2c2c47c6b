// fft_lds.hpp -- batched mixed-radix (8/4/2/3/5) Stockham FFT on sequences held in
// LDS, FP64 (double2) and FP32 (float2).  One workgroup transforms B sequences of
// length n stored at buf + b*ld; every stage reads R inputs per butterfly into
// registers, synchronises, and writes the R outputs back in place (self-sorting
// Stockham order, Govindaraju et al., SC'08).  Unnormalised: DIR = -1 computes
// sum_j x_j e^{-2 pi i jk/n}, DIR = +1 the inverse sign -- the DFTs the reference
// takes from cuFFT (source/transform.cu:49-71).
#pragma once
#include <hip/hip_runtime.h>

// A thread holds at most RPCF_REGC complex values of one stage in registers: a radix-R
// stage lets each thread process RPCF_REGC/R butterflies, so one block transforms at
// most RPCF_REGC*nthr complex elements (B*n <= fft_capacity(plan, nthr), checked on the
// host before every launch).
#ifndef RPCF_REGC
#define RPCF_REGC 16
#endif

__host__ __device__ inline int fft_capacity(const int* rad, int nst, int nthr) {
    int cap = 1 << 30;
    for (int s = 0; s < nst; s++) {
        const int R = rad[s];
        const int c = (RPCF_REGC / R) * R * nthr;
        cap = c < cap ? c : cap;
    }
    return cap;
}

struct FftPlan {
    int n;
    int nst;
    int rad[16];
    const void* tw;  // n twiddles exp(-2 pi i k/n), double2 or float2 (then, for the
                     // two-pass lengths, the stage-1 table: ct_tw1_count below)
};

template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };

// The complex helpers of the transforms allow multiply-add contraction (the library is
// otherwise built with -ffp-contract=off): a product feeding a sum becomes one fused
// multiply-add, which the FP64 butterflies are bound by.  The transforms then differ from
// an uncontracted evaluation by round-off only.
template <typename C> __device__ __forceinline__ C cmake(decltype(C().x) a, decltype(C().x) b) { C c; c.x = a; c.y = b; return c; }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
#pragma clang fp contract(fast)
    return cmake<C>(a.x + b.x, a.y + b.y);
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
#pragma clang fp contract(fast)
    return cmake<C>(a.x - b.x, a.y - b.y);
}
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
#pragma clang fp contract(fast)
    return cmake<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C().x) s) {
#pragma clang fp contract(fast)
    return cmake<C>(a.x * s, a.y * s);
}
// multiply by i (DIR=+1) or -i (DIR=-1)
template <int DIR, typename C> __device__ __forceinline__ C cjrot(C a) {
    return DIR > 0 ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<2, DIR, C> {
    static __device__ __forceinline__ void run(C* v) {
        C a = v[0];
        v[0] = cadd(a, v[1]);
        v[1] = csub(a, v[1]);
    }
};

template <int DIR, typename C> struct Dft<4, DIR, C> {
    static __device__ __forceinline__ void run(C* v) {
        C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        C jd = cjrot<DIR>(d13);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, jd);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, jd);
    }
};

template <int DIR, typename C> struct Dft<3, DIR, C> {
    static __device__ __forceinline__ void run(C* v) {
        using R = decltype(C().x);
        const R h = (R)0.86602540378443864676;  // sqrt(3)/2
        C t1 = cadd(v[1], v[2]);
        C t2 = csub(v[0], cscale(t1, (R)0.5));
        C d = cscale(csub(v[1], v[2]), h);
        C jd = cjrot<DIR>(d);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, jd);
        v[2] = csub(t2, jd);
    }
};

template <int DIR, typename C> struct Dft<5, DIR, C> {
    static __device__ __forceinline__ void run(C* v) {
        using R = decltype(C().x);
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        C a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
        C a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
        C b1 = cadd(cscale(t3, s1), cscale(t4, s2));
        C b2 = csub(cscale(t3, s2), cscale(t4, s1));
        C jb1 = cjrot<DIR>(b1), jb2 = cjrot<DIR>(b2);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, jb1);
        v[4] = csub(a1, jb1);
        v[2] = cadd(a2, jb2);
        v[3] = csub(a2, jb2);
    }
};

template <int DIR, typename C> struct Dft<8, DIR, C> {
    static __device__ __forceinline__ void run(C* v) {
        using R = decltype(C().x);
        const R h = (R)0.70710678118654752440;  // sqrt(2)/2
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, DIR, C>::run(e);
        Dft<4, DIR, C>::run(o);
        // o_k *= exp(DIR * i pi k / 4)
        C o1 = cmul(o[1], cmake<C>(h, (R)DIR * h));
        C o2 = cjrot<DIR>(o[2]);
        C o3 = cmul(o[3], cmake<C>(-h, (R)DIR * h));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// cos(2 pi m / N), sin(2 pi m / N) for the composite radices (N = 12, 16)
template <int N> struct TwTab;
template <> struct TwTab<16> {
    static constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                     0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                     -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                                     0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
    static constexpr double s[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                                     1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                     0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                     -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};
};
template <> struct TwTab<12> {
    static constexpr double c[12] = {1.0, 0.86602540378443864676, 0.5, 0.0, -0.5, -0.86602540378443864676,
                                     -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5, 0.86602540378443864676};
    static constexpr double s[12] = {0.0, 0.5, 0.86602540378443864676, 1.0, 0.86602540378443864676, 0.5,
                                     0.0, -0.5, -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5};
};

// Radix R1*R2 butterfly in registers (Cooley-Tukey, n = R2 n1 + n2, k = k1 + R1 k2):
// R2 DFTs of length R1, the twiddles exp(DIR 2 pi i n2 k1 / (R1 R2)), R1 DFTs of length R2.
template <int R1, int R2, int DIR, typename C> struct DftComposite {
    static __device__ __forceinline__ void run(C* v) {
        using T = decltype(C().x);
        constexpr int N = R1 * R2;
        C y[R2][R1];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) {
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) y[n2][n1] = v[R2 * n1 + n2];
            Dft<R1, DIR, C>::run(y[n2]);
#pragma unroll
            for (int k1 = 1; k1 < R1; k1++) {
                const int m = (n2 * k1) % N;
                if (m != 0) {
                    const C w = cmake<C>((T)TwTab<N>::c[m], (T)(DIR * TwTab<N>::s[m]));
                    y[n2][k1] = cmul(y[n2][k1], w);
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            C t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) t[n2] = y[n2][k1];
            Dft<R2, DIR, C>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = t[k2];
        }
    }
};
template <int DIR, typename C> struct Dft<16, DIR, C> : DftComposite<4, 4, DIR, C> {};
template <int DIR, typename C> struct Dft<12, DIR, C> : DftComposite<4, 3, DIR, C> {};

template <int R, int DIR, typename C>
__device__ __forceinline__ void fft_stage(C* buf, int ld, int B, int n, int Ns, const C* __restrict__ tw, int tid,
                                          int nthr) {
    const int nR = n / R;
    const int tot = B * nR;
    const int tstep = n / (Ns * R);
    constexpr int MAXPT = RPCF_REGC / R;
    C v[MAXPT][R];
#pragma unroll
    for (int it = 0; it < MAXPT; it++) {
        int t = tid + it * nthr;
        if (t < tot) {
            int b = t / nR, j = t - b * nR, k = j % Ns;
            const C* base = buf + b * ld;
#pragma unroll
            for (int r = 0; r < R; r++) {
                C x = base[j + r * nR];
                if (r > 0) {
                    C w = tw[k * r * tstep];
                    if (DIR > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[it][r] = x;
            }
            Dft<R, DIR, C>::run(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXPT; it++) {
        int t = tid + it * nthr;
        if (t < tot) {
            int b = t / nR, j = t - b * nR, k = j % Ns;
            C* base = buf + b * ld;
            int d = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) base[d + r * Ns] = v[it][r];
        }
    }
    __syncthreads();
}

// Transform B sequences buf[b*ld .. b*ld+n) in place.  Callers synchronise before
// (data written) and may read the result right after (the last stage ends with a barrier).
template <int DIR, typename C>
__device__ void fft_run(C* buf, int ld, int B, const FftPlan& pl, int tid, int nthr) {
    const C* tw = (const C*)pl.tw;
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int R = pl.rad[s];
        switch (R) {
            case 8: fft_stage<8, DIR>(buf, ld, B, pl.n, Ns, tw, tid, nthr); break;
            case 4: fft_stage<4, DIR>(buf, ld, B, pl.n, Ns, tw, tid, nthr); break;
            case 2: fft_stage<2, DIR>(buf, ld, B, pl.n, Ns, tw, tid, nthr); break;
            case 3: fft_stage<3, DIR>(buf, ld, B, pl.n, Ns, tw, tid, nthr); break;
            default: fft_stage<5, DIR>(buf, ld, B, pl.n, Ns, tw, tid, nthr); break;
        }
        Ns *= R;
    }
}

// ---- compile-time plans ---------------------------------------------------------------
// The same factorisation as the host (rpcf_host::factorize): as many 8s as possible, a
// 4 (or 4,4 in place of an 8 and a 2), then 3s and 5s.  With the length, the radices and
// every stage stride known at compile time, the index arithmetic of a stage reduces to
// constant multiplies and shifts.
constexpr int ct_pow2(int n) { int a = 0; while (n % 2 == 0) { n /= 2; a++; } return a; }
constexpr int ct_odd(int n) { while (n % 2 == 0) n /= 2; return n; }
constexpr int ct_nst(int n) {
    int a = ct_pow2(n), q = a / 3, r = a % 3, k = 0;
    if (r == 1 && q >= 1) k = (q - 1) + 2;
    else k = q + (r != 0 ? 1 : 0);
    int o = ct_odd(n);
    while (o % 3 == 0) { o /= 3; k++; }
    while (o % 5 == 0) { o /= 5; k++; }
    return k;
}
constexpr int ct_rad(int n, int s) {
    int a = ct_pow2(n), q = a / 3, r = a % 3;
    int rad[16] = {0}, k = 0;
    if (r == 1 && q >= 1) { for (int i = 0; i < q - 1; i++) rad[k++] = 8; rad[k++] = 4; rad[k++] = 4; }
    else { for (int i = 0; i < q; i++) rad[k++] = 8; if (r == 1) rad[k++] = 2; if (r == 2) rad[k++] = 4; }
    int o = ct_odd(n);
    while (o % 3 == 0) { o /= 3; rad[k++] = 3; }
    while (o % 5 == 0) { o /= 5; rad[k++] = 5; }
    return rad[s];
}

// BT > 0: at most BT sequences and kFftBlock threads, so a thread owns at most
// ceil(BT*N/R / kFftBlock) butterflies of a radix-R stage; BT = 0: RPCF_REGC/R.
constexpr int kFftBlock = 256;

// Two-pass plans for the compiled lengths (radix 16 and 12 butterflies held in registers,
// one LDS exchange per pass); other lengths use ct_rad's radix-8 plan.
constexpr int ct2_nst(int n) {
    return (n == 48 || n == 64 || n == 96 || n == 128 || n == 192 || n == 256) ? 2 : (n == 384 ? 3 : 0);
}
constexpr int ct2_rad(int n, int s) {
    return n == 48 ? (s == 0 ? 16 : 3)
         : n == 64 ? (s == 0 ? 16 : 4)
         : n == 96 ? (s == 0 ? 8 : 12)
         : n == 128 ? (s == 0 ? 16 : 8)
         : n == 192 ? (s == 0 ? 16 : 12)
         : n == 256 ? 16
         : n == 384 ? (s == 0 ? 16 : (s == 1 ? 12 : 2))
         : 0;
}
// Stage 1 of a two-pass plan multiplies input r of butterfly k (k < NS, NS the first radix)
// by w^(k r tstep), tstep = n/(NS R).  Read as tw[k r tstep], the lanes k = 0, 1, .. of a
// wave hit LDS at stride r: 2-, 4- or 8-way bank conflicts for even r.  The plan's twiddle
// table therefore carries, behind its n entries exp(-2 pi i m/n), the same values laid out
// [r - 1][k] (entry i: m = ct_tw1_index(n, i)), read at consecutive addresses.
constexpr int ct_tw1_count(int n) { return ct2_nst(n) >= 2 ? (ct2_rad(n, 1) - 1) * ct2_rad(n, 0) : 0; }
constexpr int ct_tw_len(int n) { return n + ct_tw1_count(n); }
constexpr int ct_tw1_index(int n, int i) {
    return (i % ct2_rad(n, 0)) * (i / ct2_rad(n, 0) + 1) * (n / (ct2_rad(n, 0) * ct2_rad(n, 1)));
}
// The twiddles a compiled plan reads, as kernels copy them into LDS: the plain entries
// m < ct_tw_plain(n) its later stages read (none for a two-stage plan, whose stage 0 has no
// twiddles; m < n/2 for the last, radix-2 stage of the three-stage one; all n for the
// radix-8 plans), then the stage-1 table -- fewer entries than n for the two-pass plans.
constexpr int ct_tw_plain(int n) { return ct2_nst(n) == 2 ? 0 : ct2_nst(n) == 3 ? n / ct2_rad(n, 2) : n; }
constexpr int ct_tw_lds_len(int n) { return ct_tw_plain(n) + ct_tw1_count(n); }
// the plan-table entry that entry i of the LDS copy holds, and its twiddle index m
constexpr int ct_tw_lds_src(int n, int i) { return i < ct_tw_plain(n) ? i : n + (i - ct_tw_plain(n)); }
constexpr int ct_tw_lds_index(int n, int i) {
    return i < ct_tw_plain(n) ? i : ct_tw1_index(n, i - ct_tw_plain(n));
}
constexpr int ct_plan_nst(int n) { return ct2_nst(n) ? ct2_nst(n) : ct_nst(n); }
constexpr int ct_plan_rad(int n, int s) { return ct2_nst(n) ? ct2_rad(n, s) : ct_rad(n, s); }

// LDS address swizzle of a sequence element: a ^ ((a / Q) % Q) for Q = 8 or 16 permutes
// the elements inside each aligned group of Q (a bijection on [0, n) when Q divides n).  The
// first stage of a plan writes its butterfly outputs at stride R (j*R + r): unswizzled, the
// lanes of a wave hit the same LDS banks; swizzled by Q = R they spread over all banks.  The
// second stage reads through the same swizzle.
template <int Q> __device__ __forceinline__ unsigned swz(unsigned a) {
    if constexpr (Q == 0) return a;
    else return a ^ ((a / Q) % Q);
}

template <int R, int DIR, int N, int NS, int BT, typename C, int QW = 0, int QR = 0, bool TW1 = false, int T1OFF = N>
__device__ __forceinline__ void fft_stage_ct(C* buf, int ld, int B, const C* __restrict__ tw, int tid, int nthr) {
    // unsigned index arithmetic: divisions by the compile-time strides become shifts and
    // multiplies without sign fix-ups
    constexpr unsigned nR = N / R;
    constexpr unsigned tstep = N / (NS * R);
    constexpr int MAXPT = BT ? (BT * nR + kFftBlock - 1) / kFftBlock : RPCF_REGC / R;
    const unsigned tot = B * nR;
    C v[MAXPT][R];
#pragma unroll
    for (int it = 0; it < MAXPT; it++) {
        const unsigned t = tid + it * nthr;
        if (t < tot) {
            const unsigned b = t / nR, j = t - b * nR, k = j % NS;
            const C* base = buf + b * ld;
#pragma unroll
            for (int r = 0; r < R; r++) {
                C x = base[swz<QR>(j + r * nR)];
                if (r > 0 && NS > 1) {
                    C w = TW1 ? tw[T1OFF + (r - 1) * NS + k] : tw[k * r * tstep];
                    if (DIR > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[it][r] = x;
            }
            Dft<R, DIR, C>::run(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXPT; it++) {
        const unsigned t = tid + it * nthr;
        if (t < tot) {
            const unsigned b = t / nR, j = t - b * nR, k = j % NS;
            C* base = buf + b * ld;
            const unsigned dst = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) base[swz<QW>(dst + r * NS)] = v[it][r];
        }
    }
    __syncthreads();
}

// TWO = 1: the two-pass plan where one exists; TWO = 0: the radix-8 plan (fewer registers).
// T1OFF: the first stage-1 twiddle in tw (N in a plan table, ct_tw_plain(N) in an LDS copy).
template <int N, int DIR, int BT, int TWO, typename C, int T1OFF = N, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_ct(C* buf, int ld, int B, const C* tw, int tid, int nthr) {
    constexpr int nst = TWO ? ct_plan_nst(N) : ct_nst(N);
    if constexpr (S < nst) {
        constexpr int R = TWO ? ct_plan_rad(N, S) : ct_rad(N, S);
        constexpr int R0 = TWO ? ct_plan_rad(N, 0) : ct_rad(N, 0);
        constexpr int Q = (nst >= 2 && (R0 == 8 || R0 == 16)) ? R0 : 0;  // swizzle between stages 0 and 1
        constexpr bool TW1 = TWO && S == 1 && ct2_nst(N) >= 2;  // stage 1 from the [r-1][k] table
        fft_stage_ct<R, DIR, N, NS, BT, C, S == 0 ? Q : 0, S == 1 ? Q : 0, TW1, T1OFF>(buf, ld, B, tw, tid, nthr);
        fft_ct<N, DIR, BT, TWO, C, T1OFF, S + 1, NS * R>(buf, ld, B, tw, tid, nthr);
    }
}

// FN > 0: the compile-time plan of length FN (must equal pl.n); FN = 0: the runtime plan.
// BT > 0 (compiled plans only): B <= BT sequences and nthr == kFftBlock.  LDSTW: pl.tw is
// the LDS copy of the plan's twiddles (ct_tw_lds_len entries, ct_tw_lds_src).
template <int FN, int DIR, int BT = 0, int TWO = 1, bool LDSTW = false, typename C>
__device__ __forceinline__ void fft_any(C* buf, int ld, int B, const FftPlan& pl, int tid, int nthr) {
    static_assert(FN == 0 || BT > 0, "compiled plans need the tile bound (their radix-12/16 stages size registers by it)");
    if constexpr (FN > 0)
        fft_ct<FN, DIR, BT, TWO, C, LDSTW ? ct_tw_plain(FN) : FN>(buf, ld, B, (const C*)pl.tw, tid, nthr);
    else fft_run<DIR>(buf, ld, B, pl, tid, nthr);
}

// lengths with compiled plans (every grid of the benchmark, ensemble and serving sets)
#define RPCF_FFT_LENGTHS(X) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
