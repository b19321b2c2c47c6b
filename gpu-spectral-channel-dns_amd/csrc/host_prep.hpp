// host_prep.hpp -- host-side set-up computed once per context: Chebyshev operator
// matrices, base profile, FFT factorisations and twiddles, the seeded initial field.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace rpcf_host {

// get_T_matrix: source/coefficient.cu:88-147.  T0[i][j] = T_j(z_i) at z_i = cos(pi i/N),
// T2/T4 the 2nd/4th derivatives from the same FP64 recurrence, in the same order.
inline void t_matrices(int N, double* T0, double* T2, double* T4) {
    const double PI = 4.0 * atan(1.0);
    const size_t n1 = (size_t)(N + 1);
    std::vector<double> T1(n1 * n1), T3(n1 * n1);
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) T0[n1 * i + j] = cos(j * PI * i / N);
    for (int j = 0; j <= 3; j++)
        for (int i = 0; i <= N; i++) {
            size_t inc = n1 * i + j;
            T1[inc] = 0.0; T2[inc] = 0.0; T3[inc] = 0.0; T4[inc] = 0.0;
        }
    for (int i = 0; i <= N; i++) T1[n1 * i + 1] = T0[n1 * i + 0];
    for (int i = 0; i <= N; i++) {
        T2[n1 * i + 2] = 4 * T0[n1 * i + 0];
        T1[n1 * i + 2] = 4 * T0[n1 * i + 1];
    }
    for (int j = 3; j <= N; j++)
        for (int i = 0; i <= N; i++) {
            size_t J = n1 * i + j, J1 = J - 1, J2 = J - 2;
            T1[J] = 2 * j * T0[J1] + j * T1[J2] / (j - 2);
            T2[J] = 2 * j * T1[J1] + j * T2[J2] / (j - 2);
            T3[J] = 2 * j * T2[J1] + j * T3[J2] / (j - 2);
            T4[J] = 2 * j * T3[J1] + j * T4[J2] / (j - 2);
        }
}

// get_U (source/coefficient.cu:149-159) generalised to U = c0 + c1 y + c2 y^2 at the
// nz collocation points.
inline void base_profile(int n, const double* c, double* U, double* dU, double* ddU) {
    const double PI = 4.0 * atan(1.0);
    for (int i = 0; i < n; i++) {
        double z = cos((double)i / (n - 1) * PI);
        U[i] = c[0] + c[1] * z + c[2] * (z * z);
        dU[i] = c[1] + 2.0 * c[2] * z;
        ddU[i] = 2.0 * c[2];
    }
}

// base flow at the pz padded points for the Coriolis term (source/nonlinear.cu:400-403)
inline void base_profile_padded(int pz, const double* c, double* mU) {
    const double PI = 4.0 * atan(1.0);
    for (int q = 0; q < pz; q++) {
        double y = cos((double)q / ((double)(pz - 1)) * PI);
        mU[q] = c[0] + c[1] * y + c[2] * (y * y);
    }
}

// radices of a Stockham plan: as many 8s as possible, then 4 (or 4,4 for a leftover
// 2 when an 8 can be split), then 3s and 5s.  Returns false if n has another factor.
inline bool factorize(int n, int* rad, int* nst) {
    int a = 0, k = 0;
    while (n % 2 == 0) { n /= 2; a++; }
    int q = a / 3, r = a % 3;
    if (r == 1 && q >= 1) { q -= 1; for (int i = 0; i < q; i++) rad[k++] = 8; rad[k++] = 4; rad[k++] = 4; }
    else { for (int i = 0; i < q; i++) rad[k++] = 8; if (r == 1) rad[k++] = 2; if (r == 2) rad[k++] = 4; }
    while (n % 3 == 0) { n /= 3; rad[k++] = 3; }
    while (n % 5 == 0) { n /= 5; rad[k++] = 5; }
    *nst = k;
    return n == 1 && k <= 16;
}

inline int min_radix(const int* rad, int nst) {
    int m = 64;
    for (int i = 0; i < nst; i++) m = rad[i] < m ? rad[i] : m;
    return m;
}

// exp(-2 pi i k / n), the same expression as the oracle's twiddles
inline void twiddles(int n, double* tw) {
    const double PI = 4.0 * atan(1.0);
    for (int j = 0; j < n; j++) {
        double ang = -2.0 * PI * (double)j / (double)n;
        tw[2 * j] = cos(ang);
        tw[2 * j + 1] = sin(ang);
    }
}

// ---- seeded initial field (see rpcf_seed_state in include/rpcf.h) -------------------
inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline double gauss(uint64_t* s) {
    const double PI = 4.0 * atan(1.0);
    double u1 = 1.0 - (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    double u2 = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
}

// b = (1 - y^2) a on Chebyshev coefficients (complex as re,im pairs)
inline void mul_one_minus_y2(const double* a, double* b, int n) {
    std::vector<double> ya(2 * n, 0.0), yya(2 * n, 0.0);
    for (int pass = 0; pass < 2; pass++) {
        const double* src = pass ? ya.data() : a;
        double* dst = pass ? yya.data() : ya.data();
        for (int k = 0; k < 2 * n; k++) dst[k] = 0.0;
        for (int k = 0; k < n; k++) {
            if (k == 0) {
                if (n > 1) { dst[2] = dst[2] + src[0]; dst[3] = dst[3] + src[1]; }
                continue;
            }
            double hr = src[2 * k] * 0.5, hi = src[2 * k + 1] * 0.5;
            if (k + 1 < n) { dst[2 * (k + 1)] = dst[2 * (k + 1)] + hr; dst[2 * (k + 1) + 1] = dst[2 * (k + 1) + 1] + hi; }
            dst[2 * (k - 1)] = dst[2 * (k - 1)] + hr;
            dst[2 * (k - 1) + 1] = dst[2 * (k - 1) + 1] + hi;
        }
    }
    for (int k = 0; k < 2 * n; k++) b[k] = a[k] - yya[k];
}

// ---- tables of the split-FP16 Chebyshev products (mixed form, k_cheb_mfma.hpp) ---------
// cos(pi n / m) with n reduced modulo 2m first
inline long double cospi_ratio(long n, long m) {
    const long double PI = 3.141592653589793238462643383279502884L;
    n %= 2 * m;
    return cosl(PI * (long double)n / (long double)m);
}

// power-of-two exponent e with max|T| 2^e in [2^14, 2^15)
inline int table_scale(const std::vector<long double>& T) {
    long double m = 0;
    for (long double v : T) m = fabsl(v) > m ? fabsl(v) : m;
    if (!(m > 0)) return 0;
    int ex;
    (void)frexpl(m, &ex);
    return 15 - ex;
}

// Append the fragments of a rows x cols table (row-major) scaled by 2^scale, in the A-operand
// lane order of v_mfma_f32_16x16x32_f16: fragment (row tile rt, k-step s) is
// [hi, lo][lane l][i] = T[16 rt + l % 16][32 s + 8 (l / 16) + i] (zero outside the table), with
// hi = fp16(x), lo = fp16(x - hi).
inline void mfma_fragments(const std::vector<long double>& T, int rows, int cols, int scale,
                           std::vector<_Float16>& out) {
    const int RT = (rows + 15) / 16, KS = (cols + 31) / 32;
    for (int rt = 0; rt < RT; rt++)
        for (int s = 0; s < KS; s++)
            for (int part = 0; part < 2; part++)
                for (int l = 0; l < 64; l++)
                    for (int i = 0; i < 8; i++) {
                        const int r = 16 * rt + l % 16, k = 32 * s + 8 * (l / 16) + i;
                        const double x = r < rows && k < cols ? (double)ldexpl(T[(size_t)r * cols + k], scale) : 0.0;
                        const _Float16 hi = (_Float16)(float)x;
                        out.push_back(part == 0 ? hi : (_Float16)(float)(x - (double)(float)hi));
                    }
}

// s2p of length 2N: E[q][j] = cos(pi 2j q/(2N)) (q <= N, j <= N/2) and
// O[q][j] = cos(pi (2j+1) q/(2N)) (q < N, j < N/2), scaled by 2^14.
inline void s2p_tables(int N, std::vector<_Float16>& out) {
    const int ne = N / 2 + 1, no = N / 2;
    std::vector<long double> E((size_t)(N + 1) * ne), O((size_t)N * no);
    for (int q = 0; q <= N; q++)
        for (int j = 0; j < ne; j++) E[(size_t)q * ne + j] = cospi_ratio((long)2 * j * q, 2L * N);
    for (int q = 0; q < N; q++)
        for (int j = 0; j < no; j++) O[(size_t)q * no + j] = cospi_ratio((long)(2 * j + 1) * q, 2L * N);
    mfma_fragments(E, N + 1, ne, 14, out);
    mfma_fragments(O, N, no, 14, out);
}

// p2s + derivative + unpadded s2p of length 2N (k_p2s_mfma, k_p2s_dmfma): M0 = S P,
// M1 = S D P with P the padded p2s (source/transform.cu:513-536, 538-658), D the Chebyshev
// derivative (ddz_sm, source/RPCFKernels.cu:476-520), S the values at the nz collocation
// points; folded by reflection parity into Q0s, Q0a, Q1s, Q1a (rows p <= N/2, p < N/2).
struct P2sFolded {
    std::vector<long double> q0s, q0a, q1s, q1a;  // [N/2+1][N+1], [N/2][N], [N/2+1][N], [N/2][N+1]
};
inline P2sFolded p2s_folded(int N) {
    const int nz = N + 1, pz = 2 * N + 1;
    std::vector<long double> P((size_t)nz * pz), DP((size_t)nz * pz, 0.0L), M0((size_t)nz * pz), M1((size_t)nz * pz);
    for (int k = 0; k < nz; k++)
        for (int q = 0; q < pz; q++) {
            const long double w = (q == 0 || q == 2 * N) ? 1.0L : 2.0L;
            P[(size_t)k * pz + q] = w * cospi_ratio((long)k * q, 2L * N) / (2.0L * N) * (k == 0 ? 0.5L : 1.0L);
        }
    for (int k = 0; k < nz; k++)
        for (int j = k + 1; j < nz; j += 2)
            for (int q = 0; q < pz; q++) DP[(size_t)k * pz + q] += 2.0L * j * P[(size_t)j * pz + q] * (k == 0 ? 0.5L : 1.0L);
    for (int p = 0; p < nz; p++)
        for (int q = 0; q < pz; q++) {
            long double a = 0, b = 0;
            for (int k = 0; k < nz; k++) {
                const long double c = cospi_ratio((long)k * p, (long)N);
                a += c * P[(size_t)k * pz + q];
                b += c * DP[(size_t)k * pz + q];
            }
            M0[(size_t)p * pz + q] = a;
            M1[(size_t)p * pz + q] = b;
        }
    const int rs = N / 2 + 1, ra = N / 2;
    P2sFolded f;
    f.q0s.resize((size_t)rs * (N + 1));
    f.q0a.resize((size_t)ra * N);
    f.q1s.resize((size_t)rs * N);
    f.q1a.resize((size_t)ra * (N + 1));
    for (int p = 0; p < rs; p++)
        for (int q = 0; q <= N; q++) {
            const long double m0 = M0[(size_t)p * pz + q], m0r = M0[(size_t)p * pz + 2 * N - q];
            const long double m1 = M1[(size_t)p * pz + q], m1r = M1[(size_t)p * pz + 2 * N - q];
            f.q0s[(size_t)p * (N + 1) + q] = q < N ? (m0 + m0r) / 2 : m0;
            if (q < N) f.q1s[(size_t)p * N + q] = (m1 - m1r) / 2;
            if (p < ra) {
                if (q < N) f.q0a[(size_t)p * N + q] = (m0 - m0r) / 2;
                f.q1a[(size_t)p * (N + 1) + q] = q < N ? (m1 + m1r) / 2 : m1;
            }
        }
    return f;
}

// split-FP16 fragments of the folded tables (k_p2s_mfma); s0, s1: scales of the M0 and M1 tables
inline void p2s_tables(int N, std::vector<_Float16>& out, int* s0, int* s1) {
    const P2sFolded f = p2s_folded(N);
    const int rs = N / 2 + 1, ra = N / 2;
    std::vector<long double> all0(f.q0s), all1(f.q1s);
    all0.insert(all0.end(), f.q0a.begin(), f.q0a.end());
    all1.insert(all1.end(), f.q1a.begin(), f.q1a.end());
    *s0 = table_scale(all0);
    *s1 = table_scale(all1);
    mfma_fragments(f.q0s, rs, N + 1, *s0, out);
    mfma_fragments(f.q0a, ra, N, *s0, out);
    mfma_fragments(f.q1s, rs, N, *s1, out);
    mfma_fragments(f.q1a, ra, N + 1, *s1, out);
}

// Append the fragments of a rows x cols table (row-major) in the A-operand lane order of
// v_mfma_f64_16x16x4_f64: fragment (row tile rt, k-step s) is [lane l] = T[16 rt + l % 16][4 s + l / 16]
// (zero outside the table).
inline void dmfma_fragments(const std::vector<long double>& T, int rows, int cols, std::vector<double>& out) {
    const int RT = (rows + 15) / 16, KS = (cols + 3) / 4;
    for (int rt = 0; rt < RT; rt++)
        for (int s = 0; s < KS; s++)
            for (int l = 0; l < 64; l++) {
                const int r = 16 * rt + l % 16, k = 4 * s + l / 16;
                out.push_back(r < rows && k < cols ? (double)T[(size_t)r * cols + k] : 0.0);
            }
}

// FP64 fragments of the folded tables (k_p2s_dmfma): Q0s, Q0a, Q1s, Q1a, then as plain doubles
// the centre rows Q0s[N/2][0..N] and Q1s[N/2][0..N-1] and the columns of the centre sample
// Q0s[0..N/2][N] and Q1a[0..N/2-1][N]
inline void p2s_tables_f64(int N, std::vector<double>& out) {
    const P2sFolded f = p2s_folded(N);
    const int rs = N / 2 + 1, ra = N / 2;
    dmfma_fragments(f.q0s, rs, N + 1, out);
    dmfma_fragments(f.q0a, ra, N, out);
    dmfma_fragments(f.q1s, rs, N, out);
    dmfma_fragments(f.q1a, ra, N + 1, out);
    for (int q = 0; q <= N; q++) out.push_back((double)f.q0s[(size_t)(N / 2) * (N + 1) + q]);
    for (int q = 0; q < N; q++) out.push_back((double)f.q1s[(size_t)(N / 2) * N + q]);
    for (int p = 0; p < rs; p++) out.push_back((double)f.q0s[(size_t)p * (N + 1) + N]);
    for (int p = 0; p < ra; p++) out.push_back((double)f.q1a[(size_t)p * (N + 1) + N]);
}

}  // namespace rpcf_host
