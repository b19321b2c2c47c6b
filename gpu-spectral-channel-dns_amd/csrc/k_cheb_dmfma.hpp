// k_cheb_dmfma.hpp -- the padded p2s of the FP64 form as dense FP64 products on the matrix
// cores (v_mfma_f64_16x16x4_f64), with the Chebyshev derivative, the nonlinear combination
// and the unpadded s2p folded into the tables.
//
// Reference counterparts: the padded p2s of the three Lamb components (cheby_p2s,
// source/transform.cu:538-658), rhsNonlinear (source/nonlinear.cu:146-242) and the unpadded
// s2p of N_v, N_eta (source/rhs.cpp:96-97); the (0,0) pencil keeps L_x, L_y at the nz points
// (source/rhs.cpp:229-236).
//
// With P the padded p2s (pz = 2N + 1 samples -> nz coefficients), D the Chebyshev
// derivative and S the values at the nz collocation points, M0 = S P and M1 = S D P map the
// samples of a component straight to the nz points.  rhsNonlinear is linear in the three
// components, so per pencil (alpha, beta, k^2 = alpha^2 + beta^2)
//   N_v   = M0 h1 + M1 h2,   h1 = k^2 f_z,   h2 = i alpha f_x + i beta f_y
//   N_eta = M0 h3,           h3 = i alpha f_y - i beta f_x
// (pencil (0,0): h1 = f_x, h2 = 0, h3 = f_y).  Reflection parity folds every product: with
// G+_q = f_q + f_{2N-q} (q < N; G+_N = f_N) and G-_q = f_q - f_{2N-q}, row p <= N/2 of a
// product is S_p + A_p and row N - p is S_p - A_p, where S = Q0s G+ (or Q1s G- for M1) over
// the "sym" rows p <= N/2 and A = Q0a G- (Q1a G+) over the "anti" rows p < N/2.  The tables
// are computed once per length on the host in long double and rounded to FP64.
//
// MFMA operand layouts of v_mfma_f64_16x16x4_f64 (checked on the GPU): lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16], and D[l / 16 + 4 r][l % 16] for r < 4.  Tables are
// stored as fragments (row tile, k-step) of 64 doubles in this lane order.
#pragma once
#include "k_transform.hpp"

namespace cheb_dmfma {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kPencils = 16;  // pencils per item: columns 0..15 real parts, 16..31 imaginary parts
constexpr int kCols = 2 * kPencils;
constexpr int kChunk = 32;    // samples q per chunk of the product's inner dimension, at most

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }

template <int FN>
struct DShape {
    static constexpr int N = FN / 2, nz = N + 1, pz = FN + 1;
    static constexpr int RS = N / 2 + 1, RA = N / 2;  // sym rows p <= N/2, anti rows p < N/2
    // a centre row N/2 that starts a row tile of its own is a dot product on the vector units
    static constexpr bool CENTRE = (N / 2) % 16 == 0;
    static constexpr int RTS = ceil_div(RS, 16), RTA = ceil_div(RA, 16);
    static constexpr int RTM = CENTRE ? RTS - 1 : RTS;  // sym row tiles on the matrix cores
    static constexpr int NPAIR = RTM > RTA ? RTM : RTA;  // row tiles of each kind
    static_assert(RTM == RTA, "the sym and anti row tiles pair up");
    // a block takes TPB sym and TPB anti row tiles of an item (H blocks per item): one wave
    // per tile, each wave over both column tiles (real and imaginary parts)
    static constexpr int TPB = NPAIR > 2 && NPAIR % 2 == 0 ? NPAIR / 2 : NPAIR;
    static constexpr int H = NPAIR / TPB;
    static constexpr int NT = 128 * TPB;
    static constexpr int KP = ceil_div(N + 1, 4), KM = ceil_div(N, 4);  // k-steps of the tables over G+, G-
    // The chunks cover the samples q < N (N/4 k-steps of both G+ and G-); the centre sample
    // q = N (G+ only) is a rank-one update on the vector units.  Samples per chunk: the largest
    // multiple of 4 dividing N, at most 32.
    static constexpr int chunk_of(int n) {
        for (int c = kChunk; c >= 4; c -= 4)
            if (n % c == 0) return c;
        return 0;
    }
    static constexpr int CHK = chunk_of(N);
    static_assert(CHK > 0, "N divisible by 4");
    static constexpr int KPC = CHK / 4;  // k-steps per chunk
    static constexpr int NCH = N / CHK;
    static constexpr int UPT = ceil_div(CHK * kPencils, NT);  // (q, pencil) units per thread and chunk
    // fragment offsets (in fragments of 64 doubles): Q0s [RTS][KP], Q0a [RTA][KM], Q1s [RTS][KM],
    // Q1a [RTA][KP]; then the centre rows Q0s[N/2][0..N] and Q1s[N/2][0..N-1], and the
    // columns of the centre sample Q0s[0..N/2][N] and Q1a[0..N/2-1][N], as plain doubles
    static constexpr long F0S = 0, F0A = F0S + (long)RTS * KP, F1S = F0A + (long)RTA * KM,
                          F1A = F1S + (long)RTS * KM, FEND = F1A + (long)RTA * KP;
    static constexpr long CENTRE_OFF = FEND * 64;
    static constexpr long COLN_OFF = CENTRE_OFF + (N + 1) + N;
    static constexpr long table_doubles = COLN_OFF + RS + RA;
    // LDS: the operand buffer [set h3, h2, h1][G+, G-][CHK][kCols] doubles, reused after the
    // last chunk to stage the block's output rows [pencil][SR][re, im] of N_v, then of N_eta:
    // rows p of its tiles, their mirrors N - p and the centre row
    static constexpr int BUF = 3 * 2 * CHK * kCols;
    static constexpr int SR = 32 * TPB + 1;
    static constexpr size_t lds_bytes() {
        const size_t ops = (size_t)BUF * sizeof(double);
        const size_t out = (size_t)kPencils * SR * 2 * sizeof(double);
        return ops > out ? ops : out;
    }
    // blocks for nitems items: groups of 8 items, the H blocks of an item 8 ids apart (the
    // same XCD, dispatched together, so the item's samples are read from HBM once)
    static int blocks(int nitems) { return (nitems + 7) / 8 * 8 * H; }
};

}  // namespace cheb_dmfma

// -------------------------------------------------------------------------------------
// Padded p2s + derivative + rhsNonlinear + unpadded s2p of the FP64 form: Z1[3][kx][q][ky]
// (the three Lamb components after the span transform, q = 0..2N) -> N_v, N_eta at the nz
// points, FP64, in the slot's current history buffers.  An item is (member, kx, 16
// consecutive ky); block (item, h) computes the rows of sym and anti row tiles
// h TPB .. h TPB + TPB - 1 (and their mirrors N - p), in NCH chunks of CHK samples q < N
// (with their mirrors 2N - q).  Per chunk: thread units (q, pencil) fold the six samples of
// x, y, z (loaded one chunk ahead) into G+ and G-, combine them into h3, h2, h1 and write them
// as B operands to LDS; the waves load their table fragments (from L2); after a barrier each
// wave runs the chunk's k-steps for its tile over both column tiles -- Q0 fragments serve h3
// and h1, Q1 fragments h2 -- while the next chunk's samples are in flight.  The centre
// sample q = N enters afterwards as a rank-one update of the accumulators on the vector
// units.  Three blocks share a compute unit, so one block's products overlap another's
// loads.  At the end the
// sym waves write S to rows p and N - p of an LDS stage, the anti waves add A to row p and
// subtract it from row N - p, and the rows leave as contiguous runs of the pencils.  A
// centre row N/2 of its own runs on the vector units from the same LDS operands.
// -------------------------------------------------------------------------------------
template <int FN>
__global__ __launch_bounds__(cheb_dmfma::DShape<FN>::NT, 3) void k_p2s_dmfma(
    const double2* __restrict__ Z1, long zstride, long sY, double2* __restrict__ NL, long sNL, Dims d,
    const double* __restrict__ tab, Members mb, int nitems, int nyb) {
    using namespace cheb_dmfma;
    using Sh = DShape<FN>;
    constexpr int N = Sh::N, nz = Sh::nz, NT = Sh::NT, UPT = Sh::UPT, CHK = Sh::CHK, TPB = Sh::TPB, SR = Sh::SR;
    RPCF_LDS(double);
    __shared__ double cpart[2 * TPB][2][32];  // centre row: per wave, [N_eta, N_v][column]
    // centre row weights Q0s[N/2][0..N], Q1s[N/2][0..N-1]
    __shared__ double ctr[Sh::CENTRE ? 2 * N + 1 : 1];
    // block -> (item, h): groups of 8 items, the H blocks of an item 8 ids apart
    const int grp = blockIdx.x / (8 * Sh::H), rem8 = blockIdx.x % (8 * Sh::H);
    const int item = grp * 8 + rem8 % 8, h = rem8 / 8;
    if (item >= nitems) return;
    const bool centre = Sh::CENTRE && h == Sh::H - 1;  // this block also computes row N/2
    // waves 0 .. TPB-1: sym row tiles h TPB + wave; waves TPB .. 2 TPB-1: the anti row tiles
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: branches on it are scalar
    const bool anti = wave >= TPB;
    const int rt = h * TPB + (anti ? wave - TPB : wave);
    const int kq = lane >> 4, lc = lane & 15;
    const v4d zero4 = {0.0, 0.0, 0.0, 0.0};

    int slot, kx, ky0, nb;
    {
        const int per = d.hnx * nyb;
        const int m = item / per, r = item - m * per;
        slot = mb.ids[m];
        kx = r / nyb;
        ky0 = (r - kx * nyb) * kPencils;
        nb = min(kPencils, d.ny - ky0);
    }
    const double2* in = Z1 + slot * sY + (long)kx * Sh::pz * d.ny + ky0;
    // registers: the samples q and 2N - q of x, y, z for this thread's units of one chunk
    double2 sa[UPT][3], sb[UPT][3];
    auto load = [&](int c) {
#pragma unroll
        for (int u = 0; u < UPT; u++) {
            const int i = tid + u * NT, q = c * CHK + i / kPencils, b = i % kPencils;
            const bool ok = i < CHK * kPencils && b < nb;
            const long oa = ok ? (long)q * d.ny + b : 0, ob = ok ? (long)(FN - q) * d.ny + b : 0;
#pragma unroll
            for (int g = 0; g < 3; g++) {
                sa[u][g] = in[g * zstride + oa];
                sb[u][g] = in[g * zstride + ob];
            }
        }
    };
    const double aphi = mb.mp[slot].aphi, beta = mb.mp[slot].beta;
    // h3 = i alpha y - i beta x, h2 = i alpha x + i beta y, h1 = k^2 z of pencil ky0 + b
    // (pencil (0,0): y, 0, x)
    auto combine = [&](double2 x, double2 y, double2 z, int b, double2* h3, double2* h2, double2* h1) {
        const int ky = ky0 + b;
        const double ia = (double)kx / aphi;
        const double ib = ky >= d.ny / 2 + 1 ? (double)(ky - d.ny) / beta : (double)ky / beta;
        const double k2 = ia * ia + ib * ib;
        const bool mean = kx == 0 && ky == 0;
        // i a y - i b x = (-a y.im + b x.im, a y.re - b x.re); i a x + i b y likewise
        *h3 = mean ? y : make_double2(-ia * y.y + ib * x.y, ia * y.x - ib * x.x);
        *h2 = mean ? make_double2(0.0, 0.0) : make_double2(-ia * x.y - ib * y.y, ia * x.x + ib * y.x);
        *h1 = mean ? x : make_double2(k2 * z.x, k2 * z.y);
    };
    // the centre sample q = N of every pencil, combined: hN[set h3, h2, h1][column]
    __shared__ double hN[3][kCols];
    if (tid < kPencils) {
        double2 h3 = make_double2(0.0, 0.0), h2 = h3, h1 = h3;
        if (tid < nb) {
            const long o = (long)N * d.ny + tid;
            combine(in[o], in[zstride + o], in[2 * zstride + o], tid, &h3, &h2, &h1);
        }
        hN[0][tid] = h3.x, hN[0][kPencils + tid] = h3.y;
        hN[1][tid] = h2.x, hN[1][kPencils + tid] = h2.y;
        hN[2][tid] = h1.x, hN[2][kPencils + tid] = h1.y;
    }
    // fold, combine and write the B operands of chunk c
    auto stage = [&](int c) {
#pragma unroll
        for (int u = 0; u < UPT; u++) {
            const int i = tid + u * NT;
            if (i >= CHK * kPencils) continue;
            const int qq = i / kPencils, b = i % kPencils;
            const bool ok = b < nb;
            double2 gp[3], gm[3];
#pragma unroll
            for (int g = 0; g < 3; g++) {
                const double2 a = sa[u][g], m = sb[u][g];
                gp[g] = ok ? make_double2(a.x + m.x, a.y + m.y) : make_double2(0.0, 0.0);
                gm[g] = ok ? make_double2(a.x - m.x, a.y - m.y) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int pm = 0; pm < 2; pm++) {
                double2 h3, h2, h1;
                combine(pm ? gm[0] : gp[0], pm ? gm[1] : gp[1], pm ? gm[2] : gp[2], b, &h3, &h2, &h1);
                double* o = lds + (pm * CHK + qq) * kCols + b;
                o[0] = h3.x;
                o[kPencils] = h3.y;
                o[2 * CHK * kCols] = h2.x;
                o[2 * CHK * kCols + kPencils] = h2.y;
                o[4 * CHK * kCols] = h1.x;
                o[4 * CHK * kCols + kPencils] = h1.y;
            }
        }
    };
    // this wave's tables: sym: Q0 = Q0s over G+, Q1 = Q1s over G-; anti: Q0 = Q0a over G-,
    // Q1 = Q1a over G+
    const double* tq0 = tab + (anti ? Sh::F0A + (long)rt * Sh::KM : Sh::F0S + (long)rt * Sh::KP) * 64 + lane;
    const double* tq1 = tab + (anti ? Sh::F1A + (long)rt * Sh::KP : Sh::F1S + (long)rt * Sh::KM) * 64 + lane;
    const int pm0 = anti ? 1 : 0;  // side of h3, h1 (Q0) for this wave; h2 (Q1) takes the other
    // B operand of lane (kq, lc) at local k-step sl of set s, side pm, column tile ct
    auto bop = [&](int s, int pm, int sl, int ct) {
        return lds[((s * 2 + pm) * CHK + 4 * sl + kq) * kCols + ct * kPencils + lc];
    };
    if (centre) {
        for (int i = tid; i < 2 * N + 1; i += NT) ctr[i] = tab[Sh::CENTRE_OFF + i];
    }
    // accumulators [column tile] of N_eta and N_v over this wave's rows (S or A)
    v4d accE[2] = {zero4, zero4}, accV[2] = {zero4, zero4};
    double cE = 0.0, cV = 0.0;  // centre row partial sums
    load(0);
    for (int c = 0; c < Sh::NCH; c++) {
        stage(c);
        // the chunk's fragments, issued before the next chunk's samples
        double f0[Sh::KPC], f1[Sh::KPC];
#pragma unroll
        for (int sl = 0; sl < Sh::KPC; sl++) {
            f0[sl] = tq0[(long)(c * Sh::KPC + sl) * 64];
            f1[sl] = tq1[(long)(c * Sh::KPC + sl) * 64];
        }
        __syncthreads();
        if (c + 1 < Sh::NCH) load(c + 1);  // in flight during this chunk's products
        // every chunk is KPC k-steps of both tables: no branch, so that the operand reads
        // are issued ahead of the products
#pragma unroll
        for (int sl = 0; sl < Sh::KPC; sl++) {
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                accE[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(f0[sl], bop(0, pm0, sl, ct), accE[ct], 0, 0, 0);
                accV[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(f0[sl], bop(2, pm0, sl, ct), accV[ct], 0, 0, 0);
            }
#pragma unroll
            for (int ct = 0; ct < 2; ct++)
                accV[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(f1[sl], bop(1, 1 - pm0, sl, ct), accV[ct], 0, 0, 0);
        }
        if (centre) {
            // column tid % 32, samples qq = tid / 32 + (NT / 32) j of the chunk
            const int col = tid & 31;
            for (int qq = tid >> 5; qq < CHK; qq += NT / 32) {
                const int q = c * CHK + qq;
                const double w0 = ctr[q];
                cE = fma(w0, lds[(0 * CHK + qq) * kCols + col], cE);
                cV = fma(w0, lds[(4 * CHK + qq) * kCols + col], cV);
                cV = fma(ctr[N + 1 + q], lds[(3 * CHK + qq) * kCols + col], cV);
            }
        }
        __syncthreads();  // the operands are read before the next chunk (or the stage) overwrites them
    }
    // output stage [pencil][SR][re, im], one output at a time (N_v, then N_eta): local row
    // j < 16 TPB is p = p0 + j, row 16 TPB + j its mirror N - p, row 32 TPB the centre
    const int p0 = 16 * h * TPB;
    double* out = lds;
    double cen[2] = {0.0, 0.0};  // centre row sums of N_v, N_eta (threads tid < kCols)
    if (centre) {
        cE += __shfl_xor(cE, 32);
        cV += __shfl_xor(cV, 32);
        if (lane < 32) cpart[wave][0][lane] = cE, cpart[wave][1][lane] = cV;
    }
    __syncthreads();
    if (centre && tid < kCols) {
        for (int w = 0; w < 2 * TPB; w++) cen[1] += cpart[w][0][tid], cen[0] += cpart[w][1][tid];
        cen[1] = fma(ctr[N], hN[0][tid], cen[1]);
        cen[0] = fma(ctr[N], hN[2][tid], cen[0]);
    }
    // the centre sample q = N (G+ only): sym rows Q0s[p][N] h3_N, h1_N; anti rows Q1a[p][N] h2_N
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int p = 16 * rt + kq + 4 * r;
        const double w = anti ? (p < Sh::RA ? tab[Sh::COLN_OFF + Sh::RS + p] : 0.0)
                              : (p < Sh::RS ? tab[Sh::COLN_OFF + p] : 0.0);
#pragma unroll
        for (int ct = 0; ct < 2; ct++) {
            if (anti) {
                accV[ct][r] = fma(w, hN[1][ct * kPencils + lc], accV[ct][r]);
            } else {
                accE[ct][r] = fma(w, hN[0][ct * kPencils + lc], accE[ct][r]);
                accV[ct][r] = fma(w, hN[2][ct * kPencils + lc], accV[ct][r]);
            }
        }
    }
    const int ph = mb.phase[slot];
    const long PS = (long)d.P * nz;
    const long po = ((long)ky0 * d.hnx + kx) * nz, pstride = (long)d.hnx * nz;
#pragma unroll
    for (int o = 0; o < 2; o++) {
        const v4d* acc = o ? accE : accV;
        // sym waves write S to rows p and N - p, then anti waves add A to row p and
        // subtract it from row N - p
        if (!anti) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = 16 * wave + kq + 4 * r, p = p0 + j;
                if (p >= Sh::RS || (Sh::CENTRE && p == N / 2)) continue;
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    out[(lc * SR + j) * 2 + ct] = acc[ct][r];
                    out[(lc * SR + 16 * TPB + j) * 2 + ct] = acc[ct][r];
                }
            }
        }
        if (centre && tid < kCols) out[((tid % kPencils) * SR + 32 * TPB) * 2 + tid / kPencils] = cen[o];
        __syncthreads();
        if (anti) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = 16 * (wave - TPB) + kq + 4 * r, p = p0 + j;
                if (p >= Sh::RA) continue;
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    out[(lc * SR + j) * 2 + ct] += acc[ct][r];
                    out[(lc * SR + 16 * TPB + j) * 2 + ct] -= acc[ct][r];
                }
            }
        }
        __syncthreads();
        double2* dst = NL + slot * sNL + (o ? 2 + ph : ph) * PS;
        const double2* ov = (const double2*)out;
        for (int i = tid; i < nb * SR; i += NT) {
            const int bb = i / SR, j = i - bb * SR;
            int row;
            if (j < 16 * TPB) {
                const int p = p0 + j;
                row = p < Sh::RS && !(Sh::CENTRE && p == N / 2) ? p : -1;
            } else if (j < 32 * TPB) {
                const int p = p0 + j - 16 * TPB;
                row = p < Sh::RA ? N - p : -1;
            } else {
                row = centre ? N / 2 : -1;
            }
            if (row >= 0) dst[po + bb * pstride + row] = ov[i];
        }
        if (o == 0) __syncthreads();  // the stage is read before N_eta overwrites it
    }
}
