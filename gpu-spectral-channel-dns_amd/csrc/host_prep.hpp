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

}  // namespace rpcf_host
