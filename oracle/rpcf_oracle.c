/*
 * rpcf_oracle.c -- CPU FP64 restatement of one time step of the reference DNS
 * (Riveryh/gpu-spectral-channel-DNS).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline:
 * it is built into oracle/_build/librpcf_oracle.so and may be loaded only by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg -- always as
 * the checker, never as the code path being measured or shipped.
 *
 * It exports the C ABI of include/rpcf.h.  Every function cites the reference
 * file:line it restates.  Arithmetic is written in the reference's expression order
 * and compiled with -ffp-contract=off so that no fused multiply-add changes the
 * rounding.  Third-party arithmetic the reference calls and that is not available
 * here is restated from its published definition:
 *   - cuFFT (CUDA >= 9, no version pinned; call sites source/transform.cu:49-71,
 *     176, 240, 581, 628, 696, 731): unnormalised DFTs e^{-2 pi i jk/n} (forward)
 *     and e^{+...} (inverse), computed here by a mixed-radix decimation-in-time FFT.
 *   - Eigen 3.2+ Matrix::inverse() (PartialPivLU; source/matrix_op.cpp:19-27) and
 *     the matrix-vector product (source/matrix_op.cpp:68-78): restated as LU with
 *     partial pivoting + forward/back substitution, and a row-by-row dot product with
 *     j ascending.  Eigen's own rounding order is NOT reproduced: the inverse and GEMV
 *     are "parity unpinned" against the reference binary; they are pinned by A*A^-1=I,
 *     by a dense solve on small systems and by the physics checks in tests/.
 *
 * Parity status of each function is listed in the table at the top of DESIGN.md.
 *
 * rpcf_oracle_plant (exported, 0 by default) plants a known mistake for the tests that
 * show a pin fails on it: 1 flips the sign of the Coriolis term of L_x, 2 that of L_z,
 * 3 the sign of the lift-up coupling of the eta right-hand side, 4 the sign of the
 * Squire advection term -i alpha U eta.  With 0 every factor is exactly 1.0, so the
 * arithmetic is bit for bit the unplanted one.
 */
#include "../include/rpcf.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------ */
/* complex type and the operators of cuRPCF::complex (include/data.h:37-79)        */
/* ------------------------------------------------------------------------------ */
typedef struct { double re, im; } cx;

static inline cx cmk(double r, double i) { cx c; c.re = r; c.im = i; return c; }
/* operator*(REAL) include/data.h:53-57 */
static inline cx cmulr(cx a, double r) { return cmk(a.re * r, a.im * r); }
/* operator*(complex) include/data.h:58-62 */
static inline cx cmul(cx a, cx c) {
    double tre = a.re * c.re - a.im * c.im;
    double tim = a.re * c.im + a.im * c.re;
    return cmk(tre, tim);
}
/* operator+ / operator- include/data.h:63-72 */
static inline cx cadd(cx a, cx c) { return cmk(a.re + c.re, a.im + c.im); }
static inline cx csub(cx a, cx c) { return cmk(a.re - c.re, a.im - c.im); }
static inline cx cconj(cx a) { return cmk(a.re, -a.im); }

/* ------------------------------------------------------------------------------ */
/* FFT: mixed-radix recursive decimation in time, unnormalised                      */
/* Restates the DFT that cuFFT computes (plans at source/transform.cu:49-71).       */
/* ------------------------------------------------------------------------------ */
typedef struct {
    int n;
    int nfac;
    int fac[64];   /* pairs (radix p, remaining length m) */
    cx* twf;       /* exp(-2 pi i k / n) */
    cx* twb;       /* exp(+2 pi i k / n) */
} fftplan;

static int fft_plan_init(fftplan* pl, int n) {
    memset(pl, 0, sizeof(*pl));
    pl->n = n;
    int rem = n, p = 4, k = 0;
    while (rem > 1) {
        while (rem % p) {
            if (p == 4) p = 2;
            else if (p == 2) p = 3;
            else p += 2;
            if (p > 64) return -1;
        }
        rem /= p;
        pl->fac[2 * k] = p;
        pl->fac[2 * k + 1] = rem;
        k++;
    }
    pl->nfac = k;
    pl->twf = (cx*)malloc(sizeof(cx) * (size_t)n);
    pl->twb = (cx*)malloc(sizeof(cx) * (size_t)n);
    if (!pl->twf || !pl->twb) return -1;
    const double PI = 4.0 * atan(1.0);
    for (int j = 0; j < n; j++) {
        double ang = -2.0 * PI * (double)j / (double)n;
        pl->twf[j] = cmk(cos(ang), sin(ang));
        pl->twb[j] = cmk(cos(ang), -sin(ang));
    }
    return 0;
}

static void fft_plan_free(fftplan* pl) {
    free(pl->twf);
    free(pl->twb);
    pl->twf = pl->twb = NULL;
}

static void fft_work(cx* out, const cx* in, size_t fstride, const int* fac, const cx* tw, int n, int inv) {
    const int p = fac[0], m = fac[1];
    if (m == 1) {
        for (int q = 0; q < p; q++) out[q] = in[(size_t)q * fstride];
    } else {
        for (int q = 0; q < p; q++) fft_work(out + (size_t)q * m, in + (size_t)q * fstride, fstride * (size_t)p, fac + 2, tw, n, inv);
    }
    /* combine: X[u + q1 m] = sum_q W_{pm}^{q (u + q1 m)} Y_q[u]; W_{pm} = tw[. * fstride] */
    if (p == 2) {
        for (int u = 0; u < m; u++) {
            cx t = cmul(out[u + m], tw[(size_t)u * fstride]);
            cx a = out[u];
            out[u] = cadd(a, t);
            out[u + m] = csub(a, t);
        }
    } else if (p == 4) {
        for (int u = 0; u < m; u++) {
            cx t0 = out[u];
            cx t1 = cmul(out[u + m], tw[(size_t)u * fstride]);
            cx t2 = cmul(out[u + 2 * m], tw[(size_t)2 * u * fstride]);
            cx t3 = cmul(out[u + 3 * m], tw[(size_t)3 * u * fstride]);
            cx s02 = cadd(t0, t2), d02 = csub(t0, t2);
            cx s13 = cadd(t1, t3), d13 = csub(t1, t3);
            /* -i*d13 for the forward transform, +i*d13 for the inverse */
            cx jd = inv ? cmk(-d13.im, d13.re) : cmk(d13.im, -d13.re);
            out[u] = cadd(s02, s13);
            out[u + m] = cadd(d02, jd);
            out[u + 2 * m] = csub(s02, s13);
            out[u + 3 * m] = csub(d02, jd);
        }
    } else {
        cx scratch[64];
        for (int u = 0; u < m; u++) {
            for (int q = 0; q < p; q++) scratch[q] = out[u + q * m];
            for (int q1 = 0; q1 < p; q1++) {
                size_t k = (size_t)u + (size_t)q1 * m;
                cx acc = scratch[0];
                size_t tidx = 0;
                for (int q = 1; q < p; q++) {
                    tidx += fstride * k;
                    tidx %= (size_t)n;
                    acc = cadd(acc, cmul(scratch[q], tw[tidx]));
                }
                out[k] = acc;
            }
        }
    }
}

/* out = DFT(in) with sign - (inv=0) or + (inv=1); in != out */
static void fft_exec(const fftplan* pl, const cx* in, cx* out, int inv) {
    if (pl->n == 1) { out[0] = in[0]; return; }
    fft_work(out, in, 1, pl->fac, inv ? pl->twb : pl->twf, pl->n, inv);
}

/* ------------------------------------------------------------------------------ */
/* context                                                                          */
/* ------------------------------------------------------------------------------ */
int rpcf_oracle_plant = 0;
static double plant_sign(int which) { return rpcf_oracle_plant == which ? -1.0 : 1.0; }

struct rpcf_ctx {
    rpcf_params prm;
    int mx, my, mz, nx, ny, nz, pz, hnx, N, P;
    double aphi, beta, lx, ly, Re, Ro, dt;
    double *T0, *T2, *T4;        /* [nz][nz], row = collocation point, col = degree */
    double *U, *dU, *ddU;        /* base profile at the nz collocation points */
    double* meanU_pad;           /* base profile at the pz padded points (Coriolis) */
    cx *mat_v, *mat_o;           /* [P][nz*nz] inverse collocation matrices */
    cx *rhs_v, *rhs_v_p, *rhs_o; /* [P][nz] prognostic state / rhs (host rhs_v ...) */
    cx *nl_v, *nl_vp, *nl_o, *nl_op; /* [P][nz] N_v, N_eta at collocation points */
    cx *lambx0, *lambz0, *lambx0_p, *lambz0_p, *tv0, *tomega_y_0; /* [nz] */
    cx* fld[6];                  /* [P][nz] spectral u,v,w,ox,oy,oz (device t-fields) */
    /* flow-rate control extension */
    cx* u_h;                     /* response of the (0,0) u0 solve to unit forcing */
    double* qw;                  /* integral weights of T_k over [-1,1] */
    cx q_target;
    fftplan plan_x, plan_y, plan_zpad, plan_znop;
    int have_state;
    char err[256];
};

static int set_err(rpcf_ctx* c, int code, const char* msg) {
    if (c) snprintf(c->err, sizeof(c->err), "%s", msg);
    return code;
}

/* problem(RPCF_Paras) + problem::initVars: include/data.h:312-328, source/data.cpp:3-14 */
static int check_sizes(const rpcf_params* prm, int* nx, int* ny, int* nz, int* pz, int* hnx) {
    if (!prm) return RPCF_EINVAL;
    int mx = prm->mx, my = prm->my, mz = prm->mz;
    if (mx < 6 || my < 6 || mz < 16) return RPCF_EINVAL;
    if (mx % 2 || my % 2 || mz % 4) return RPCF_EINVAL;
    if (!(prm->Re > 0) || !(prm->dt > 0) || !(prm->n_pi_x > 0) || !(prm->n_pi_y > 0)) return RPCF_EINVAL;
    if (nx) *nx = mx / 3 * 2;
    if (ny) *ny = my / 3 * 2;
    if (nz) *nz = mz / 4 + 1;
    if (pz) *pz = mz / 2 + 1;
    if (hnx) *hnx = (mx / 3 * 2) / 2 + 1;
    return RPCF_OK;
}

int rpcf_sizes(const rpcf_params* prm, int* nx, int* ny, int* nz, int* pz, int* hnx) {
    return check_sizes(prm, nx, ny, nz, pz, hnx);
}

/* ------------------------------------------------------------------------------ */
/* operators                                                                        */
/* ------------------------------------------------------------------------------ */

/* get_T_matrix: source/coefficient.cu:88-147.  T0[i][j] = T_j(z_i); T2, T4 are the
 * 2nd and 4th derivative values, built by the same FP64 recurrence in the same order. */
void rpcf_oracle_T_matrix(int N, double* T0, double* T2, double* T4) {
    const double PI = 4.0 * atan(1.0);
    const size_t n1 = (size_t)(N + 1);
    double* T1 = (double*)malloc(sizeof(double) * n1 * n1);
    double* T3 = (double*)malloc(sizeof(double) * n1 * n1);
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
    free(T1);
    free(T3);
}

/* get_U: source/coefficient.cu:149-159, with U(y) = c0 + c1 y + c2 y^2 (reference
 * c = {0.5, 0.5, 0} gives 0.5*(1+y) bit for bit because scaling by 0.5 is exact). */
static void get_U(int n, const double* c, double* U, double* dU, double* ddU) {
    const double PI = 4.0 * atan(1.0);
    for (int i = 0; i < n; i++) {
        double z = cos((double)i / (n - 1) * PI);
        U[i] = c[0] + c[1] * z + c[2] * (z * z);
        dU[i] = c[1] + 2.0 * c[2] * z;
        ddU[i] = 2.0 * c[2];
    }
}

/* _get_coefficient_v: source/coefficient.cu:7-37 */
static void coef_v(cx* m, int N, const double* U, const double* ddU, const double* T0, const double* T2,
                   const double* T4, double Re, double dt, double kmn, double alpha) {
    const size_t n1 = (size_t)(N + 1);
    for (int i = 4; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            size_t inc = n1 * i + j, inc2 = n1 * (i - 2) + j;
            m[inc].re = -T4[inc2] * dt * 0.5 / Re + (1 + kmn * dt / Re) * T2[inc2]
                        - kmn * (1 + kmn * dt * 0.5 / Re) * T0[inc2];
            m[inc].im = alpha * dt * U[i - 2] * 0.5 * T2[inc2]
                        - (kmn * alpha * dt * 0.5 * U[i - 2] + alpha * dt * 0.5 * ddU[i - 2]) * T0[inc2];
        }
    for (int j = 0; j <= N; j++) {
        m[n1 * 0 + j] = cmk(1.0, 0.0);
        m[n1 * 1 + j] = cmk((j % 2 == 0) ? 1.0 : -1.0, 0.0);
        m[n1 * 2 + j] = cmk((double)(j * j), 0.0);
        m[n1 * 3 + j] = cmk(((j % 2 == 0) ? -1.0 : 1.0) * j * j, 0.0);
    }
}

/* _get_coefficient_omega: source/coefficient.cu:39-64 */
static void coef_o(cx* m, int N, const double* U, const double* T0, const double* T2, double Re, double dt,
                   double kmn, double alpha) {
    const size_t n1 = (size_t)(N + 1);
    for (int i = 2; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            size_t inc = n1 * i + j, inc1 = n1 * (i - 1) + j;
            m[inc].re = -T2[inc1] * dt * 0.5 / Re + (1 + kmn * dt * 0.5 / Re) * T0[inc1];
            m[inc].im = alpha * dt * 0.5 * U[i - 1] * T0[inc1];
        }
    for (int j = 0; j <= N; j++) {
        m[n1 * 0 + j] = cmk(1.0, 0.0);
        m[n1 * 1 + j] = cmk((j % 2 == 0) ? 1.0 : -1.0, 0.0);
    }
}

/* _get_coef_u0 / _get_coef_w0: source/coefficient.cu:66-86 */
static void coef_u0(cx* m, int N, const double* T0, const double* T2, double Re, double dt) {
    const size_t n1 = (size_t)(N + 1);
    for (int i = 2; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            size_t inc = n1 * i + j, inc1 = n1 * (i - 1) + j;
            m[inc] = cmk(-T2[inc1] * dt * 0.5 / Re + T0[inc1], 0.0);
        }
    for (int j = 0; j <= N; j++) {
        m[n1 * 0 + j] = cmk(1.0, 0.0);
        m[n1 * 1 + j] = cmk((j % 2 == 0) ? 1.0 : -1.0, 0.0);
    }
}

/* inverse: source/matrix_op.cpp:22-27 (Eigen PartialPivLU inverse), restated as LU
 * with partial pivoting (pivot = largest |re|^2+|im|^2, first on ties) followed by
 * solving A X = I column by column.  Returns -1 on an exactly singular pivot. */
int rpcf_oracle_inverse(cx* a, int n) {
    const size_t N = (size_t)n;
    cx* lu = (cx*)malloc(sizeof(cx) * N * N);
    int* piv = (int*)malloc(sizeof(int) * N);
    cx* col = (cx*)malloc(sizeof(cx) * N);
    if (!lu || !piv || !col) { free(lu); free(piv); free(col); return -1; }
    memcpy(lu, a, sizeof(cx) * N * N);
    int rc = 0;
    for (int k = 0; k < n; k++) {
        int pr = k;
        double best = lu[N * k + k].re * lu[N * k + k].re + lu[N * k + k].im * lu[N * k + k].im;
        for (int r = k + 1; r < n; r++) {
            double v = lu[N * r + k].re * lu[N * r + k].re + lu[N * r + k].im * lu[N * r + k].im;
            if (v > best) { best = v; pr = r; }
        }
        piv[k] = pr;
        if (best == 0.0) { rc = -1; break; }
        if (pr != k)
            for (int j = 0; j < n; j++) { cx t = lu[N * k + j]; lu[N * k + j] = lu[N * pr + j]; lu[N * pr + j] = t; }
        cx d = lu[N * k + k];
        double dd = d.re * d.re + d.im * d.im;
        cx dinv = cmk(d.re / dd, -d.im / dd);
        for (int r = k + 1; r < n; r++) {
            cx l = cmul(lu[N * r + k], dinv);
            lu[N * r + k] = l;
            for (int j = k + 1; j < n; j++) lu[N * r + j] = csub(lu[N * r + j], cmul(l, lu[N * k + j]));
        }
    }
    if (rc == 0) {
        for (int c = 0; c < n; c++) {
            for (int i = 0; i < n; i++) col[i] = cmk(i == c ? 1.0 : 0.0, 0.0);
            for (int k = 0; k < n; k++)
                if (piv[k] != k) { cx t = col[k]; col[k] = col[piv[k]]; col[piv[k]] = t; }
            for (int i = 0; i < n; i++) {
                cx s = col[i];
                for (int j = 0; j < i; j++) s = csub(s, cmul(lu[N * i + j], col[j]));
                col[i] = s;
            }
            for (int i = n - 1; i >= 0; i--) {
                cx s = col[i];
                for (int j = i + 1; j < n; j++) s = csub(s, cmul(lu[N * i + j], col[j]));
                cx d = lu[N * i + i];
                double dd = d.re * d.re + d.im * d.im;
                col[i] = cmk((s.re * d.re + s.im * d.im) / dd, (s.im * d.re - s.re * d.im) / dd);
            }
            for (int i = 0; i < n; i++) a[N * i + c] = col[i];
        }
    }
    free(lu); free(piv); free(col);
    return rc;
}

/* m_multi_v: source/matrix_op.cpp:68-78 -- v <- M v, row dot products, j ascending */
static void m_multi_v(const cx* m, cx* v, int n, cx* tmp) {
    for (int i = 0; i < n; i++) {
        cx acc = cmk(0.0, 0.0);
        const cx* row = m + (size_t)n * i;
        for (int j = 0; j < n; j++) acc = cadd(acc, cmul(row[j], v[j]));
        tmp[i] = acc;
    }
    for (int i = 0; i < n; i++) v[i] = tmp[i];
}

/* get_ialpha_ibeta: source/RPCFKernels.cu:522-532 */
static void ialpha_ibeta(const rpcf_ctx* c, int kx, int ky, double* ia, double* ib) {
    *ia = (double)kx / c->aphi;
    *ib = (double)ky / c->beta;
    if (ky >= c->ny / 2 + 1) *ib = (double)(ky - c->ny) / c->beta;
}

/* ------------------------------------------------------------------------------ */
/* create / destroy                                                                 */
/* ------------------------------------------------------------------------------ */
static void* zalloc(size_t bytes) { return calloc(1, bytes ? bytes : 1); }

void rpcf_destroy(rpcf_ctx* c) {
    if (!c) return;
    free(c->T0); free(c->T2); free(c->T4);
    free(c->U); free(c->dU); free(c->ddU); free(c->meanU_pad);
    free(c->mat_v); free(c->mat_o);
    free(c->rhs_v); free(c->rhs_v_p); free(c->rhs_o);
    free(c->nl_v); free(c->nl_vp); free(c->nl_o); free(c->nl_op);
    free(c->lambx0); free(c->lambz0); free(c->lambx0_p); free(c->lambz0_p);
    free(c->tv0); free(c->tomega_y_0);
    for (int f = 0; f < 6; f++) free(c->fld[f]);
    free(c->u_h); free(c->qw);
    fft_plan_free(&c->plan_x); fft_plan_free(&c->plan_y);
    fft_plan_free(&c->plan_zpad); fft_plan_free(&c->plan_znop);
    free(c);
}

const char* rpcf_last_error(const rpcf_ctx* c) { return c ? c->err : "null context"; }

/* initAuxMatrix: source/solver.cpp:260-320 (T, U, per-pencil matrices, inverses) */
int rpcf_create(const rpcf_params* prm, rpcf_ctx** out) {
    if (!out) return RPCF_EINVAL;
    *out = NULL;
    int nx, ny, nz, pz, hnx;
    if (check_sizes(prm, &nx, &ny, &nz, &pz, &hnx) != RPCF_OK) return RPCF_EINVAL;
    rpcf_ctx* c = (rpcf_ctx*)zalloc(sizeof(rpcf_ctx));
    if (!c) return RPCF_ENOMEM;
    c->prm = *prm;
    c->mx = prm->mx; c->my = prm->my; c->mz = prm->mz;
    c->nx = nx; c->ny = ny; c->nz = nz; c->pz = pz; c->hnx = hnx; c->N = nz - 1;
    c->P = hnx * ny;
    const double PI = 4.0 * atan(1.0);
    /* include/data.h:321-322, source/data.cpp:5-6 */
    c->aphi = prm->n_pi_x / 2;
    c->beta = prm->n_pi_y / 2;
    c->lx = 2 * PI * c->aphi;
    c->ly = 2 * PI * c->beta;
    c->Re = prm->Re; c->Ro = prm->Ro; c->dt = prm->dt;

    const size_t P = (size_t)c->P, NZ = (size_t)nz;
    int ok = 1;
#define ALLOC(ptr, type, n) do { (ptr) = (type*)zalloc(sizeof(type) * (n)); if (!(ptr)) ok = 0; } while (0)
    ALLOC(c->T0, double, NZ * NZ); ALLOC(c->T2, double, NZ * NZ); ALLOC(c->T4, double, NZ * NZ);
    ALLOC(c->U, double, NZ); ALLOC(c->dU, double, NZ); ALLOC(c->ddU, double, NZ);
    ALLOC(c->meanU_pad, double, (size_t)pz);
    ALLOC(c->mat_v, cx, P * NZ * NZ); ALLOC(c->mat_o, cx, P * NZ * NZ);
    ALLOC(c->rhs_v, cx, P * NZ); ALLOC(c->rhs_v_p, cx, P * NZ); ALLOC(c->rhs_o, cx, P * NZ);
    ALLOC(c->nl_v, cx, P * NZ); ALLOC(c->nl_vp, cx, P * NZ); ALLOC(c->nl_o, cx, P * NZ); ALLOC(c->nl_op, cx, P * NZ);
    ALLOC(c->lambx0, cx, NZ); ALLOC(c->lambz0, cx, NZ); ALLOC(c->lambx0_p, cx, NZ); ALLOC(c->lambz0_p, cx, NZ);
    ALLOC(c->tv0, cx, NZ); ALLOC(c->tomega_y_0, cx, NZ);
    for (int f = 0; f < 6; f++) ALLOC(c->fld[f], cx, P * NZ);
    ALLOC(c->u_h, cx, NZ); ALLOC(c->qw, double, NZ);
#undef ALLOC
    if (!ok) { rpcf_destroy(c); return RPCF_ENOMEM; }
    if (fft_plan_init(&c->plan_x, c->mx) || fft_plan_init(&c->plan_y, c->my) ||
        fft_plan_init(&c->plan_zpad, 4 * c->N) || fft_plan_init(&c->plan_znop, 2 * c->N)) {
        rpcf_destroy(c);
        return RPCF_EINVAL;
    }

    rpcf_oracle_T_matrix(c->N, c->T0, c->T2, c->T4);
    get_U(nz, prm->u_profile, c->U, c->dU, c->ddU);
    /* Coriolis base flow at the padded points: source/nonlinear.cu:400-403 */
    for (int q = 0; q < pz; q++) {
        double y = cos((double)q / ((double)(pz - 1)) * PI);
        c->meanU_pad[q] = prm->u_profile[0] + prm->u_profile[1] * y + prm->u_profile[2] * (y * y);
    }

    int fail = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : fail)
    for (int kx = 0; kx < hnx; kx++) {
        for (int ky = 0; ky < ny; ky++) {
            size_t inc = NZ * NZ * ((size_t)hnx * ky + kx);
            cx* mv = c->mat_v + inc;
            cx* mo = c->mat_o + inc;
            if (kx == 0 && ky == 0) {
                coef_u0(mv, c->N, c->T0, c->T2, c->Re, c->dt);
                coef_u0(mo, c->N, c->T0, c->T2, c->Re, c->dt);
            } else {
                double ia, ib;
                ialpha_ibeta(c, kx, ky, &ia, &ib);
                double kmn = ia * ia + ib * ib;
                coef_v(mv, c->N, c->U, c->ddU, c->T0, c->T2, c->T4, c->Re, c->dt, kmn, ia);
                coef_o(mo, c->N, c->U, c->T0, c->T2, c->Re, c->dt, kmn, ia);
            }
            if (rpcf_oracle_inverse(mo, nz)) fail = 1;
            if (rpcf_oracle_inverse(mv, nz)) fail = 1;
        }
    }
    if (fail) { rpcf_destroy(c); return RPCF_EINVAL; }

    /* flow-rate extension: u_h = A_u0^{-1} (0, 0, dt, ..., dt), weights of int T_k */
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
        for (int r = 0; r < nz; r++) c->u_h[r] = cmk(r >= 2 ? c->dt : 0.0, 0.0);
        m_multi_v(c->mat_v, c->u_h, nz, tmp);
        free(tmp);
        for (int k = 0; k < nz; k++) c->qw[k] = (k % 2 == 0) ? 2.0 / (1.0 - (double)k * (double)k) : 0.0;
    }
    *out = c;
    return RPCF_OK;
}

/* ------------------------------------------------------------------------------ */
/* transforms: transform_3d_one (source/transform.cu:95-319)                         */
/* ------------------------------------------------------------------------------ */

/* cheby_pre_s2p_pad + Z2Z planZ_pad (source/transform.cu:416-453, 679-698):
 * pencil of nz coefficients -> values at the pz padded points. */
static void s2p_pad(const rpcf_ctx* c, const cx* a, cx* vals, cx* buf, cx* fo) {
    const int N = c->N, L = 4 * N;
    for (int i = 0; i < L; i++) buf[i] = cmk(0.0, 0.0);
    for (int iz = 0; iz <= N; iz++) {
        cx t = cmulr(a[iz], 0.5);
        if (iz == 0) buf[0] = cmulr(t, 2.0);
        else { buf[iz] = t; buf[L - iz] = t; }
    }
    fft_exec(&c->plan_zpad, buf, fo, 0);
    for (int q = 0; q < c->pz; q++) vals[q] = fo[q];
}

/* cheby_pre_s2p_noPad + Z2Z planZ_no_pad (source/transform.cu:455-491, 716-733):
 * pencil of nz coefficients -> values at the nz collocation points. */
static void s2p_nopad(const rpcf_ctx* c, const cx* a, cx* vals, cx* buf, cx* fo) {
    const int N = c->N, L = 2 * N;
    for (int i = 0; i < N; i++) buf[i] = cmulr(a[i], 0.5);
    buf[N] = a[N];
    for (int i = 0; i < N; i++) {
        if (i >= 1) buf[L - i] = buf[i];
        else buf[0] = cmulr(buf[0], 2.0);
    }
    fft_exec(&c->plan_znop, buf, fo, 0);
    for (int q = 0; q < c->nz; q++) vals[q] = fo[q];
}

/* cheby_pre_p2s + Z2Z + cheby_post_p2s (source/transform.cu:494-536, 538-658):
 * values at P = pz (padded) or nz points -> Chebyshev coefficients; keeps k < nz. */
static void p2s(const rpcf_ctx* c, const cx* vals, cx* a, int padded, cx* buf, cx* fo) {
    const int P = padded ? c->pz : c->nz;
    const int L = 2 * (P - 1);
    for (int q = 0; q < P; q++) buf[q] = vals[q];
    for (int q = 1; q < P - 1; q++) buf[L - q] = buf[q];
    fft_exec(padded ? &c->plan_zpad : &c->plan_znop, buf, fo, 0);
    double factor = (1.0 / (P - 1));
    for (int k = 0; k < c->nz; k++) {
        cx t = cmulr(fo[k], factor);
        if (k == 0 || k == P - 1) t = cmulr(t, 0.5);
        a[k] = t;
    }
}

/* physical span index of pencil ky (source/transpose.cu:238-240, 373-375) */
static inline int span_of(const rpcf_ctx* c, int ky) { return ky > c->ny / 2 ? ky + (c->my - c->ny) : ky; }

/* BACKWARD: spectral pencils -> physical planes [planes][my][mx]
 * cheby_s2p, transpose_backward, setZeros, Z2D (source/transform.cu:124-210, 355-386;
 * source/transpose.cu:358-381).  The C2R uses the real parts of the kx = 0 and
 * kx = mx/2 half-spectrum entries. */
static void backward_field(const rpcf_ctx* c, const cx* spec, size_t sstride, double* phys, int padded) {
    const int planes = padded ? c->pz : c->nz;
    const int mx = c->mx, my = c->my, hmx = mx / 2 + 1, hnx = c->hnx, ny = c->ny;
    const size_t P = (size_t)c->P;
    cx* vals = (cx*)malloc(sizeof(cx) * P * planes);
    cx* H = (cx*)malloc(sizeof(cx) * (size_t)planes * my * hmx);
#pragma omp parallel
    {
        cx* buf = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
        cx* fo = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
#pragma omp for schedule(static)
        for (size_t p = 0; p < P; p++) {
            if (padded) s2p_pad(c, spec + p * sstride, vals + p * planes, buf, fo);
            else s2p_nopad(c, spec + p * sstride, vals + p * planes, buf, fo);
        }
        free(buf); free(fo);
    }
    /* transpose_backward + setZeros */
#pragma omp parallel for schedule(static)
    for (int q = 0; q < planes; q++) {
        cx* Hq = H + (size_t)q * my * hmx;
        for (size_t i = 0; i < (size_t)my * hmx; i++) Hq[i] = cmk(0.0, 0.0);
        for (int ky = 0; ky < ny; ky++)
            for (int kx = 0; kx < hnx; kx++)
                Hq[(size_t)span_of(c, ky) * hmx + kx] = vals[((size_t)ky * hnx + kx) * planes + q];
        for (int s = 0; s < my; s++)
            for (int kx = 0; kx < hmx; kx++) {
                int zero = (s >= ny / 2 && s <= my - ny / 2) || kx >= c->nx / 2;
                if (zero) Hq[(size_t)s * hmx + kx] = cmk(0.0, 0.0);
            }
    }
    /* Z2D: inverse C2C along the span for every kx, then C2R along x */
#pragma omp parallel
    {
        int nmax = mx > my ? mx : my;
        cx* bi = (cx*)malloc(sizeof(cx) * nmax);
        cx* bo = (cx*)malloc(sizeof(cx) * nmax);
#pragma omp for schedule(static)
        for (int q = 0; q < planes; q++) {
            cx* Hq = H + (size_t)q * my * hmx;
            for (int kx = 0; kx < hmx; kx++) {
                for (int s = 0; s < my; s++) bi[s] = Hq[(size_t)s * hmx + kx];
                fft_exec(&c->plan_y, bi, bo, 1);
                for (int s = 0; s < my; s++) Hq[(size_t)s * hmx + kx] = bo[s];
            }
            for (int s = 0; s < my; s++) {
                const cx* X = Hq + (size_t)s * hmx;
                bi[0] = cmk(X[0].re, 0.0);
                for (int k = 1; k < mx / 2; k++) { bi[k] = X[k]; bi[mx - k] = cconj(X[k]); }
                bi[mx / 2] = cmk(X[mx / 2].re, 0.0);
                fft_exec(&c->plan_x, bi, bo, 1);
                double* row = phys + ((size_t)q * my + s) * mx;
                for (int x = 0; x < mx; x++) row[x] = bo[x].re;
            }
        }
        free(bi); free(bo);
    }
    free(vals);
    free(H);
}

/* FORWARD: physical planes -> spectral pencils (k < nz kept)
 * D2Z, normalize, transpose_forward, cheby_p2s (source/transform.cu:211-315, 388-412;
 * source/transpose.cu:223-257). */
static void forward_field(const rpcf_ctx* c, const double* phys, cx* spec, size_t sstride, int padded) {
    const int planes = padded ? c->pz : c->nz;
    const int mx = c->mx, my = c->my, hmx = mx / 2 + 1, hnx = c->hnx;
    const size_t P = (size_t)c->P;
    cx* vals = (cx*)malloc(sizeof(cx) * P * planes);
    cx* H = (cx*)malloc(sizeof(cx) * (size_t)planes * my * hmx);
    const double factor = 1.0 / mx / my;
#pragma omp parallel
    {
        int nmax = mx > my ? mx : my;
        cx* bi = (cx*)malloc(sizeof(cx) * nmax);
        cx* bo = (cx*)malloc(sizeof(cx) * nmax);
#pragma omp for schedule(static)
        for (int q = 0; q < planes; q++) {
            cx* Hq = H + (size_t)q * my * hmx;
            for (int s = 0; s < my; s++) {
                const double* row = phys + ((size_t)q * my + s) * mx;
                for (int x = 0; x < mx; x++) bi[x] = cmk(row[x], 0.0);
                fft_exec(&c->plan_x, bi, bo, 0);
                for (int k = 0; k < hmx; k++) Hq[(size_t)s * hmx + k] = bo[k];
            }
            for (int kx = 0; kx < hmx; kx++) {
                for (int s = 0; s < my; s++) bi[s] = Hq[(size_t)s * hmx + kx];
                fft_exec(&c->plan_y, bi, bo, 0);
                for (int s = 0; s < my; s++) Hq[(size_t)s * hmx + kx] = cmulr(bo[s], factor);
            }
        }
        free(bi); free(bo);
    }
#pragma omp parallel
    {
        cx* buf = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
        cx* fo = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
#pragma omp for schedule(static)
        for (size_t p = 0; p < P; p++) {
            int ky = (int)(p / hnx), kx = (int)(p % hnx);
            cx* v = vals + p * planes;
            for (int q = 0; q < planes; q++) v[q] = H[((size_t)q * my + span_of(c, ky)) * hmx + kx];
            p2s(c, v, spec + p * sstride, padded, buf, fo);
        }
        free(buf); free(fo);
    }
    free(vals);
    free(H);
}

/* ddz / ddz_sm on one complex pencil: source/RPCFKernels.cu:476-520
 * b_k = sum_{j=k+1,k+3,..<n} 2j a_j ; b_0 halved. */
static void ddz(const cx* u, cx* out, int n) {
    for (int k = 0; k < n; k++) {
        double bre = 0.0, bim = 0.0;
        for (int j = k + 1; j < n; j += 2) {
            double dmat = 2 * (double)j;
            bre = bre + u[j].re * dmat;
            bim = bim + u[j].im * dmat;
        }
        if (k == 0) out[0] = cmk(bre * 0.5, bim * 0.5);
        else out[k] = cmk(bre, bim);
    }
}

/* ------------------------------------------------------------------------------ */
/* nonlinear term                                                                   */
/* ------------------------------------------------------------------------------ */

/* computeLambDevice: source/nonlinear.cu:409-428, base flow at padded plane q */
static void lamb(const rpcf_ctx* c, double** ph) {
    const int pz = c->pz, my = c->my, mx = c->mx;
    const double Ro = c->Ro, Rx = plant_sign(1) * Ro, Rz = plant_sign(2) * Ro;
#pragma omp parallel for schedule(static)
    for (int q = 0; q < pz; q++) {
        double meanU = c->meanU_pad[q];
        for (size_t i = (size_t)q * my * mx; i < (size_t)(q + 1) * my * mx; i++) {
            double u = ph[0][i], v = ph[1][i], w = ph[2][i];
            double ox = ph[3][i], oy = ph[4][i], oz = ph[5][i];
            double lx = oz * v - oy * w - Rx * w;
            double ly = ox * w - oz * u;
            double lz = oy * u - ox * v + Rz * (u + meanU);
            ph[0][i] = lx; ph[1][i] = ly; ph[2][i] = lz;
        }
    }
}

/* rhsNonlinearKernel: source/nonlinear.cu:146-242 on one pencil (k < nz). */
static void rhs_nonlinear_pencil(const rpcf_ctx* c, int kx, int ky, cx* Lx, cx* Ly, const cx* Lz, cx* dbuf) {
    if (kx == 0 && ky == 0) return;
    const int nz = c->nz;
    double ia, ib;
    ialpha_ibeta(c, kx, ky, &ia, &ib);
    double kmn = ia * ia + ib * ib;
    cx* dx = dbuf;
    cx* dy = dbuf + nz;
    ddz(Lx, dx, nz);
    ddz(Ly, dy, nz);
    for (int i = 0; i < nz; i++) {
        double ure = kmn * Lz[i].re - ia * dx[i].im;
        double uim = kmn * Lz[i].im + ia * dx[i].re;
        ure = ure - ib * dy[i].im;
        uim = uim + ib * dy[i].re;
        double wre = -1.0 * ia * Ly[i].im + ib * Lx[i].im;
        double wim = ia * Ly[i].re - ib * Lx[i].re;
        dbuf[2 * nz + i] = cmk(ure, uim);
        dbuf[3 * nz + i] = cmk(wre, wim);
    }
    for (int i = 0; i < nz; i++) { Lx[i] = dbuf[2 * nz + i]; Ly[i] = dbuf[3 * nz + i]; }
}

static void swap_ptr(cx** a, cx** b) { cx* t = *a; *a = *b; *b = t; }

/* getNonlinear (source/nonlinear.cu:25-108) followed by the helper-thread tail
 * (source/rhs.cpp:96-127): unpadded s2p of N_v and N_eta, history swap, D2H and
 * save_zero_wave_number_lamb (source/rhs.cpp:229-236). */
int rpcf_nonlinear(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return set_err(c, RPCF_ESTATE, "rpcf_nonlinear: no state set");
    const size_t P = (size_t)c->P, NZ = (size_t)c->nz;
    const size_t plane = (size_t)c->my * c->mx;
    double* ph[6];
    for (int f = 0; f < 6; f++) {
        ph[f] = (double*)malloc(sizeof(double) * plane * c->pz);
        if (!ph[f]) { for (int g = 0; g < f; g++) free(ph[g]); return set_err(c, RPCF_ENOMEM, "oom"); }
    }
    /* transform(BACKWARD): source/transform.cu:333-342 */
    for (int f = 0; f < 6; f++) backward_field(c, c->fld[f], NZ, ph[f], 1);
    /* computeLambVector: source/nonlinear.cu:307-364 (L overwrites u, v, w) */
    lamb(c, ph);
    /* transform(FORWARD) of lamb_z, lamb_y, lamb_x: source/transform.cu:343-349 */
    cx* tL[3];
    for (int g = 0; g < 3; g++) tL[g] = (cx*)malloc(sizeof(cx) * P * NZ);
    forward_field(c, ph[2], tL[2], NZ, 1);
    forward_field(c, ph[1], tL[1], NZ, 1);
    forward_field(c, ph[0], tL[0], NZ, 1);
    for (int f = 0; f < 6; f++) free(ph[f]);
    /* rhsNonlinear: source/nonlinear.cu:130-145 */
#pragma omp parallel
    {
        cx* dbuf = (cx*)malloc(sizeof(cx) * 4 * NZ);
#pragma omp for schedule(static)
        for (size_t p = 0; p < P; p++) {
            int ky = (int)(p / c->hnx), kx = (int)(p % c->hnx);
            rhs_nonlinear_pencil(c, kx, ky, tL[0] + p * NZ, tL[1] + p * NZ, tL[2] + p * NZ, dbuf);
        }
        free(dbuf);
    }
    /* history swap (source/rhs.cpp:107-108), then unpadded s2p straight into the new
     * current buffers (source/rhs.cpp:96-97, 116-117) */
    swap_ptr(&c->nl_o, &c->nl_op);
    swap_ptr(&c->nl_v, &c->nl_vp);
#pragma omp parallel
    {
        cx* buf = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
        cx* fo = (cx*)malloc(sizeof(cx) * 4 * (size_t)c->N);
#pragma omp for schedule(static)
        for (size_t p = 0; p < P; p++) {
            s2p_nopad(c, tL[0] + p * NZ, c->nl_v + p * NZ, buf, fo);
            s2p_nopad(c, tL[1] + p * NZ, c->nl_o + p * NZ, buf, fo);
        }
        free(buf); free(fo);
    }
    for (int g = 0; g < 3; g++) free(tL[g]);
    /* save_zero_wave_number_lamb */
    swap_ptr(&c->lambx0, &c->lambx0_p);
    swap_ptr(&c->lambz0, &c->lambz0_p);
    for (int i = 0; i < c->nz; i++) { c->lambx0[i] = c->nl_v[i]; c->lambz0[i] = c->nl_o[i]; }
    return RPCF_OK;
}

/* ------------------------------------------------------------------------------ */
/* linear part                                                                      */
/* ------------------------------------------------------------------------------ */

/* _get_linear_v: source/linear.cpp:86-125 */
static void linear_v_pencil(cx* rhs_v, cx* rhs_v_p, int N, const double* U, const double* ddU, const double* T0,
                            const double* T2, const double* T4, double Re, double dt, double kmn, double alpha,
                            cx* tmp) {
    const size_t n1 = (size_t)(N + 1);
    for (int i = 4; i <= N; i++) {
        tmp[i] = cmk(0.0, 0.0);
        for (int j = 0; j <= N; j++) {
            size_t inc2 = n1 * (i - 2) + j;
            cx b = cmk(T4[inc2] * dt * 0.5 / Re + (1 - kmn * dt / Re) * T2[inc2]
                           - kmn * (1 - kmn * dt * 0.5 / Re) * T0[inc2],
                       -alpha * dt * 0.5 * U[i - 2] * T2[inc2]
                           + (kmn * alpha * dt * 0.5 * U[i - 2] + alpha * dt * 0.5 * ddU[i - 2]) * T0[inc2]);
            tmp[i] = cadd(tmp[i], cmul(rhs_v[j], b));
        }
    }
    for (int i = 0; i <= N; i++) {
        rhs_v_p[i] = rhs_v[i];
        rhs_v[i] = tmp[i];
    }
    rhs_v[0] = rhs_v[1] = rhs_v[2] = rhs_v[3] = cmk(0.0, 0.0);
}

/* _get_linear_omega_y: source/linear.cpp:127-171 */
static void linear_o_pencil(cx* rhs_o, const cx* nl_o, const cx* nl_op, const cx* rhs_v, const cx* rhs_v_p, int N,
                            const double* U, const double* dU, const double* T0, const double* T2, double Re,
                            double dt, double kmn, double alpha, double beta, cx* tmp) {
    const size_t n1 = (size_t)(N + 1);
    const double sq = plant_sign(4), lu = plant_sign(3);
    for (int i = 2; i <= N; i++) {
        tmp[i] = cmk(0.0, 0.0);
        for (int j = 0; j <= N; j++) {
            size_t inc1 = n1 * (i - 1) + j;
            tmp[i] = cadd(tmp[i], cmul(rhs_o[j], cmk(T2[inc1] * 0.5 / Re * dt + (1 - kmn * dt * 0.5 / Re) * T0[inc1],
                                                     -(sq * alpha) * dt * 0.5 * U[i - 1] * T0[inc1])));
            tmp[i] = cadd(tmp[i], cmul(cadd(rhs_v_p[j], rhs_v[j]),
                                       cmk(0.0, -0.5 * (lu * beta) * dU[i - 1] * dt * T0[inc1])));
        }
    }
    for (int i = 2; i <= N; i++)
        rhs_o[i] = cadd(tmp[i], cmulr(csub(cmulr(nl_o[i - 1], 1.5), cmulr(nl_op[i - 1], 0.5)), dt));
    rhs_o[0] = rhs_o[1] = cmk(0.0, 0.0);
}

/* _get_linear_u0 / _get_linear_w0: source/linear.cpp:173-205 */
static void linear_u0(cx* rhs_u0, const cx* lamb0, const cx* lamb0_p, int N, const double* T0, const double* T2,
                      double Re, double dt, cx* tmp) {
    const size_t n1 = (size_t)(N + 1);
    for (int i = 2; i <= N; i++) {
        tmp[i] = cmk(0.0, 0.0);
        for (int j = 0; j <= N; j++) {
            size_t inc1 = n1 * (i - 1) + j;
            tmp[i] = cadd(tmp[i], cmul(rhs_u0[j], cmk(T0[inc1] + dt * 0.5 / Re * T2[inc1], 0.0)));
        }
    }
    for (int i = 2; i <= N; i++)
        rhs_u0[i] = csub(tmp[i], cmulr(csub(cmulr(lamb0[i - 1], 1.5), cmulr(lamb0_p[i - 1], 0.5)), dt));
    rhs_u0[0] = rhs_u0[1] = cmk(0.0, 0.0);
}

/* constant-flow-rate extension (not in the reference): after the (0,0) solve,
 * u0 <- u0 + f u_h with f chosen so that int u0 dy keeps its value at set_state. */
static void flow_control(rpcf_ctx* c) {
    if (c->prm.flow_control != 1) return;
    cx qp = cmk(0.0, 0.0);
    double qh = 0.0;
    for (int k = 0; k < c->nz; k++) {
        qp = cadd(qp, cmulr(c->rhs_v[k], c->qw[k]));
        qh = qh + c->u_h[k].re * c->qw[k];
    }
    cx f = cmk((c->q_target.re - qp.re) / qh, (c->q_target.im - qp.im) / qh);
    for (int k = 0; k < c->nz; k++) c->rhs_v[k] = cadd(c->rhs_v[k], cmul(f, c->u_h[k]));
}

/* nextStep minus the nonlinear part: source/solver.cpp:43-89 with get_rhs_v
 * (source/rhs.cpp:185-219), get_linear_v/omega_y (source/linear.cpp:7-84),
 * solveEq (source/solver.cpp:107-123), save_0_v_omega_y (source/solver.cpp:332-338). */
int rpcf_linear(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return set_err(c, RPCF_ESTATE, "rpcf_linear: no state set");
    const int hnx = c->hnx, ny = c->ny, N = c->N, nz = c->nz;
    const size_t NZ = (size_t)nz;
    /* get_linear_v */
#pragma omp parallel
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
#pragma omp for schedule(static)
        for (int i = 0; i < hnx; i++)
            for (int j = 0; j < ny; j++) {
                if (i == 0 && j == 0) continue;
                size_t inc = NZ * ((size_t)j * hnx + i);
                double ia, ib;
                ialpha_ibeta(c, i, j, &ia, &ib);
                double kmn = ia * ia + ib * ib;
                linear_v_pencil(c->rhs_v + inc, c->rhs_v_p + inc, N, c->U, c->ddU, c->T0, c->T2, c->T4, c->Re,
                                c->dt, kmn, ia, tmp);
            }
        free(tmp);
    }
    /* AB2 term: source/rhs.cpp:205-215 */
    for (int i = 0; i < hnx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 4; k < nz; k++) {
                if (i == 0 && j == 0) continue;
                size_t inc = NZ * ((size_t)hnx * j + i) + k;
                c->rhs_v[inc] = cadd(c->rhs_v[inc],
                                     cmulr(csub(cmulr(c->nl_v[inc - 2], 1.5), cmulr(c->nl_vp[inc - 2], 0.5)), c->dt));
            }
    /* get_linear_zero_wave_u_w: source/linear.cpp:43-50 */
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
        linear_u0(c->tv0, c->lambx0, c->lambx0_p, N, c->T0, c->T2, c->Re, c->dt, tmp);
        linear_u0(c->tomega_y_0, c->lambz0, c->lambz0_p, N, c->T0, c->T2, c->Re, c->dt, tmp);
        for (int k = 0; k < nz; k++) { c->rhs_v[k] = c->tv0[k]; c->rhs_o[k] = c->tomega_y_0[k]; }
        free(tmp);
    }
    /* solveEq(matrix_coeff_v, rhs_v) */
#pragma omp parallel
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
#pragma omp for schedule(static)
        for (int p = 0; p < c->P; p++) m_multi_v(c->mat_v + NZ * NZ * p, c->rhs_v + NZ * p, nz, tmp);
        free(tmp);
    }
    flow_control(c);
    /* get_rhs_omega -> get_linear_omega_y */
#pragma omp parallel
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
#pragma omp for schedule(static)
        for (int i = 0; i < hnx; i++)
            for (int j = 0; j < ny; j++) {
                if (i == 0 && j == 0) continue;
                size_t inc = NZ * ((size_t)j * hnx + i);
                double ia, ib;
                ialpha_ibeta(c, i, j, &ia, &ib);
                double kmn = ia * ia + ib * ib;
                linear_o_pencil(c->rhs_o + inc, c->nl_o + inc, c->nl_op + inc, c->rhs_v + inc, c->rhs_v_p + inc, N,
                                c->U, c->dU, c->T0, c->T2, c->Re, c->dt, kmn, ia, ib, tmp);
            }
        free(tmp);
    }
    /* solveEq(matrix_coeff_omega, rhs_omega_y) */
#pragma omp parallel
    {
        cx* tmp = (cx*)malloc(sizeof(cx) * NZ);
#pragma omp for schedule(static)
        for (int p = 0; p < c->P; p++) m_multi_v(c->mat_o + NZ * NZ * p, c->rhs_o + NZ * p, nz, tmp);
        free(tmp);
    }
    /* save_0_v_omega_y */
    for (int k = 0; k < nz; k++) { c->tv0[k] = c->rhs_v[k]; c->tomega_y_0[k] = c->rhs_o[k]; }
    return RPCF_OK;
}

/* ------------------------------------------------------------------------------ */
/* velocity recovery                                                                */
/* ------------------------------------------------------------------------------ */

/* getVelocityKernel_sm: source/velocity.cu:90-170 on one pencil */
static void uvw_pencil(const rpcf_ctx* c, int kx, int ky, cx* u, cx* v, cx* w, cx* ox, cx* oy, cx* oz, cx* t) {
    const int nz = c->nz;
    cx* d = t + nz;
    if (kx == 0 && ky == 0) {
        for (int k = 0; k < nz; k++) { u[k] = w[k]; v[k] = oz[k]; w[k] = cmk(0.0, 0.0); }
        ddz(u, t, nz);
        ddz(v, d, nz);
        for (int k = 0; k < nz; k++) { ox[k] = d[k]; oy[k] = cmulr(t[k], -1.0); oz[k] = cmk(0.0, 0.0); }
        return;
    }
    double ia, ib;
    ialpha_ibeta(c, kx, ky, &ia, &ib);
    double kmn = ia * ia + ib * ib;
    double kmn1 = 1.0 / kmn;
    ddz(w, t, nz);
    for (int i = 0; i < nz; i++) {
        u[i] = csub(cmul(cmk(0.0, ia * kmn1), t[i]), cmul(cmk(0.0, ib * kmn1), oz[i]));
        v[i] = cadd(cmul(cmk(0.0, ib * kmn1), t[i]), cmul(cmk(0.0, ia * kmn1), oz[i]));
    }
    ddz(v, t, nz);
    for (int i = 0; i < nz; i++) ox[i] = cadd(t[i], cmul(cmk(0.0, -1.0 * ib), w[i]));
    ddz(u, t, nz);
    for (int i = 0; i < nz; i++) oy[i] = cadd(cmulr(t[i], -1.0), cmul(cmk(0.0, ia), w[i]));
}

static void get_uvw(rpcf_ctx* c) {
    const size_t NZ = (size_t)c->nz;
#pragma omp parallel
    {
        cx* t = (cx*)malloc(sizeof(cx) * 2 * NZ);
#pragma omp for schedule(static)
        for (int p = 0; p < c->P; p++) {
            size_t o = NZ * p;
            uvw_pencil(c, p % c->hnx, p / c->hnx, c->fld[0] + o, c->fld[1] + o, c->fld[2] + o, c->fld[3] + o,
                       c->fld[4] + o, c->fld[5] + o, t);
        }
        free(t);
    }
}

/* H2D copy rhs_v -> tw, rhs_omega_y -> tomega_z, getUVW: source/solver.cpp:95-98 */
int rpcf_velocity(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return set_err(c, RPCF_ESTATE, "rpcf_velocity: no state set");
    memcpy(c->fld[2], c->rhs_v, sizeof(cx) * c->P * c->nz);
    memcpy(c->fld[5], c->rhs_o, sizeof(cx) * c->P * c->nz);
    get_uvw(c);
    return RPCF_OK;
}

/* nextStep: source/solver.cpp:35-101 */
int rpcf_step(rpcf_ctx* c, int nsteps) {
    if (!c) return RPCF_EINVAL;
    for (int s = 0; s < nsteps; s++) {
        int rc = rpcf_nonlinear(c);
        if (!rc) rc = rpcf_linear(c);
        if (!rc) rc = rpcf_velocity(c);
        if (rc) return rc;
    }
    return RPCF_OK;
}

/* source/main_loop.cpp:39-52: the throw-away step leaves N(state0) in the history;
 * re-initialisation restores the state; the real first step follows. */
int rpcf_bootstrap(rpcf_ctx* c) {
    int rc = rpcf_nonlinear(c);
    if (rc) return rc;
    return rpcf_step(c, 1);
}

/* ------------------------------------------------------------------------------ */
/* state exchange                                                                   */
/* ------------------------------------------------------------------------------ */
static int copy_in(rpcf_ctx* c, cx* dst, const double* src, long stride) {
    if (!src || stride < c->nz) return RPCF_EINVAL;
    for (size_t p = 0; p < (size_t)c->P; p++)
        for (int k = 0; k < c->nz; k++) {
            const double* s = src + 2 * ((size_t)stride * p + k);
            dst[(size_t)c->nz * p + k] = cmk(s[0], s[1]);
        }
    return RPCF_OK;
}

static int copy_out(const rpcf_ctx* c, double* dst, const cx* src, long stride) {
    if (!dst || stride < c->nz) return RPCF_EINVAL;
    for (size_t p = 0; p < (size_t)c->P; p++)
        for (int k = 0; k < c->nz; k++) {
            double* d = dst + 2 * ((size_t)stride * p + k);
            cx v = src[(size_t)c->nz * p + k];
            d[0] = v.re; d[1] = v.im;
        }
    return RPCF_OK;
}

/* initFlow tail: source/RPCFKernels.cu:409-427 */
int rpcf_set_state(rpcf_ctx* c, const double* v_hat, const double* eta_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (copy_in(c, c->rhs_v, v_hat, stride) || copy_in(c, c->rhs_o, eta_hat, stride))
        return set_err(c, RPCF_EINVAL, "rpcf_set_state: bad array or stride");
    c->have_state = 1;
    rpcf_velocity(c);
    for (int k = 0; k < c->nz; k++) { c->tv0[k] = c->rhs_v[k]; c->tomega_y_0[k] = c->rhs_o[k]; }
    memcpy(c->rhs_v_p, c->rhs_v, sizeof(cx) * c->P * c->nz);
    c->q_target = cmk(0.0, 0.0);
    for (int k = 0; k < c->nz; k++) c->q_target = cadd(c->q_target, cmulr(c->rhs_v[k], c->qw[k]));
    return RPCF_OK;
}

int rpcf_get_state(rpcf_ctx* c, double* v_hat, double* eta_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (copy_out(c, v_hat, c->rhs_v, stride) || copy_out(c, eta_hat, c->rhs_o, stride))
        return set_err(c, RPCF_EINVAL, "rpcf_get_state: bad array or stride");
    return RPCF_OK;
}

int rpcf_set_field(rpcf_ctx* c, int field, const double* f_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (field < 0 || field > 5) return set_err(c, RPCF_EINVAL, "rpcf_set_field: bad field id");
    if (copy_in(c, c->fld[field], f_hat, stride)) return set_err(c, RPCF_EINVAL, "rpcf_set_field: bad array");
    c->have_state = 1;
    return RPCF_OK;
}

int rpcf_get_field(rpcf_ctx* c, int field, double* f_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (field < 0 || field > 5) return set_err(c, RPCF_EINVAL, "rpcf_get_field: bad field id");
    if (copy_out(c, f_hat, c->fld[field], stride)) return set_err(c, RPCF_EINVAL, "rpcf_get_field: bad array");
    return RPCF_OK;
}

int rpcf_get_nonlinear(rpcf_ctx* c, int prev, double* n_v, double* n_eta, long stride) {
    if (!c) return RPCF_EINVAL;
    if (copy_out(c, n_v, prev ? c->nl_vp : c->nl_v, stride) || copy_out(c, n_eta, prev ? c->nl_op : c->nl_o, stride))
        return set_err(c, RPCF_EINVAL, "rpcf_get_nonlinear: bad array");
    return RPCF_OK;
}

/* read_recover_data (source/output.cpp:141-178): nonlinear_* and the (0,0) Lamb copies */
int rpcf_set_nonlinear(rpcf_ctx* c, int prev, const double* n_v, const double* n_eta, long stride) {
    if (!c) return RPCF_EINVAL;
    if (copy_in(c, prev ? c->nl_vp : c->nl_v, n_v, stride) || copy_in(c, prev ? c->nl_op : c->nl_o, n_eta, stride))
        return set_err(c, RPCF_EINVAL, "rpcf_set_nonlinear: bad array");
    cx* lx = prev ? c->lambx0_p : c->lambx0;
    cx* lz = prev ? c->lambz0_p : c->lambz0;
    for (int i = 0; i < c->nz; i++) { lx[i] = (prev ? c->nl_vp : c->nl_v)[i]; lz[i] = (prev ? c->nl_op : c->nl_o)[i]; }
    return RPCF_OK;
}

int rpcf_phys_to_spec(rpcf_ctx* c, const double* phys, double* spec, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!phys || !spec || stride < c->nz) return set_err(c, RPCF_EINVAL, "rpcf_phys_to_spec: bad array");
    cx* tmp = (cx*)malloc(sizeof(cx) * c->P * c->nz);
    forward_field(c, phys, tmp, (size_t)c->nz, padded ? 1 : 0);
    copy_out(c, spec, tmp, stride);
    free(tmp);
    return RPCF_OK;
}

int rpcf_spec_to_phys(rpcf_ctx* c, const double* spec, double* phys, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!phys || !spec || stride < c->nz) return set_err(c, RPCF_EINVAL, "rpcf_spec_to_phys: bad array");
    cx* tmp = (cx*)malloc(sizeof(cx) * c->P * c->nz);
    copy_in(c, tmp, spec, stride);
    backward_field(c, tmp, (size_t)c->nz, phys, padded ? 1 : 0);
    free(tmp);
    return RPCF_OK;
}

/* inverse matrix of one pencil (eq 0: v, 1: eta), row-major nz*nz complex */
int rpcf_get_inverse(rpcf_ctx* c, int eq, int pencil, double* out) {
    if (!c || !out || pencil < 0 || pencil >= c->P || eq < 0 || eq > 1) return RPCF_EINVAL;
    const cx* m = (eq ? c->mat_o : c->mat_v) + (size_t)c->nz * c->nz * pencil;
    memcpy(out, m, sizeof(cx) * c->nz * c->nz);
    return RPCF_OK;
}

/* exposed for tests: unnormalised DFT of length n */
int rpcf_oracle_fft(int n, const double* in, double* out, int inverse) {
    fftplan pl;
    if (fft_plan_init(&pl, n)) { fft_plan_free(&pl); return RPCF_EINVAL; }
    fft_exec(&pl, (const cx*)in, (cx*)out, inverse);
    fft_plan_free(&pl);
    return RPCF_OK;
}

/* ------------------------------------------------------------------------------ */
/* seeded initial field (extension; see include/rpcf.h)                             */
/* ------------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double gauss(uint64_t* s) {
    const double PI = 4.0 * atan(1.0);
    double u1 = 1.0 - (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    double u2 = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
}

/* b = (1 - y^2) a on Chebyshev coefficients, y T_0 = T_1, y T_k = (T_{k+1} + T_{k-1})/2 */
static void mul_one_minus_y2(const cx* a, cx* b, int n) {
    cx* ya = (cx*)calloc((size_t)n, sizeof(cx));
    cx* yya = (cx*)calloc((size_t)n, sizeof(cx));
    for (int pass = 0; pass < 2; pass++) {
        const cx* src = pass ? ya : a;
        cx* dst = pass ? yya : ya;
        for (int k = 0; k < n; k++) dst[k] = cmk(0.0, 0.0);
        for (int k = 0; k < n; k++) {
            if (k == 0) { if (n > 1) dst[1] = cadd(dst[1], src[0]); continue; }
            cx h = cmulr(src[k], 0.5);
            if (k + 1 < n) dst[k + 1] = cadd(dst[k + 1], h);
            dst[k - 1] = cadd(dst[k - 1], h);
        }
    }
    for (int k = 0; k < n; k++) b[k] = csub(a[k], yya[k]);
    free(ya);
    free(yya);
}

/* Bound of the peak perturbation velocity of a spectral state: per velocity component
 * (u, v, w of getVelocityKernel_sm, source/velocity.cu:90-170), the sum of the magnitudes of
 * its Chebyshev-Fourier coefficients over all pencils, the conjugate pencil of kx > 0
 * counted (|T_k| <= 1, so this bounds the maximum over the domain); the largest of the
 * three.  D v by the descending recurrence b_{k} = b_{k+2} + 2(k+1) a_{k+1}, b_0 halved. */
static double seed_velocity_bound(const rpcf_params* prm, const double* v, const double* e, long stride, int nz,
                                  int ny, int hnx) {
    const double aphi = prm->n_pi_x / 2, beta = prm->n_pi_y / 2;
    double su = 0.0, sv = 0.0, sw = 0.0;
    double* b = (double*)calloc(2 * (size_t)nz + 4, sizeof(double));
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < hnx; kx++) {
            if (kx == 0 && ky == 0) continue;
            const size_t p = (size_t)ky * hnx + kx;
            const double* a = v + 2 * (size_t)stride * p;
            const double* h = e + 2 * (size_t)stride * p;
            const double mult = kx == 0 ? 1.0 : 2.0;
            const double al = (double)kx / aphi;
            const double be = (ky >= ny / 2 + 1 ? (double)(ky - ny) : (double)ky) / beta;
            const double k2 = al * al + be * be;
            for (int k = 0; k < 2 * nz + 4; k++) b[k] = 0.0;
            for (int k = nz - 2; k >= 0; k--) {
                b[2 * k] = b[2 * (k + 2)] + 2.0 * (k + 1) * a[2 * (k + 1)];
                b[2 * k + 1] = b[2 * (k + 2) + 1] + 2.0 * (k + 1) * a[2 * (k + 1) + 1];
            }
            b[0] = b[0] * 0.5;
            b[1] = b[1] * 0.5;
            for (int k = 0; k < nz; k++) {
                const double ur = -(al * b[2 * k + 1] - be * h[2 * k + 1]) / k2;
                const double ui = (al * b[2 * k] - be * h[2 * k]) / k2;
                const double wr = -(be * b[2 * k + 1] + al * h[2 * k + 1]) / k2;
                const double wi = (be * b[2 * k] + al * h[2 * k]) / k2;
                su = su + mult * sqrt(ur * ur + ui * ui);
                sv = sv + mult * sqrt(a[2 * k] * a[2 * k] + a[2 * k + 1] * a[2 * k + 1]);
                sw = sw + mult * sqrt(wr * wr + wi * wi);
            }
        }
    free(b);
    double m = su > sv ? su : sv;
    return m > sw ? m : sw;
}

int rpcf_seed_state(const rpcf_params* prm, uint64_t seed, double amplitude, int kmax, int degmax, double kexp,
                    int peak_bound, double* v_hat, double* eta_hat, long stride) {
    int nx, ny, nz, pz, hnx;
    if (check_sizes(prm, &nx, &ny, &nz, &pz, &hnx) != RPCF_OK) return RPCF_EINVAL;
    if (!v_hat || !eta_hat || stride < nz || degmax < 0 || degmax + 4 >= nz || kmax < 0) return RPCF_EINVAL;
    const size_t P = (size_t)hnx * ny;
    memset(v_hat, 0, sizeof(double) * 2 * (size_t)stride * P);
    memset(eta_hat, 0, sizeof(double) * 2 * (size_t)stride * P);
    uint64_t s = seed;
    cx* g = (cx*)calloc((size_t)nz, sizeof(cx));
    cx* t = (cx*)calloc((size_t)nz, sizeof(cx));
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < hnx; kx++) {
            int kys = ky <= ny / 2 ? ky : ky - ny;
            if (kx == 0 && ky == 0) continue;
            if (kx == nx / 2 || ky == ny / 2) continue;
            if (kx > kmax || abs(kys) > kmax) continue;
            if (kx == 0 && kys < 0) continue;
            double scale = pow(1.0 + sqrt((double)kx * kx + (double)kys * kys), -kexp);
            size_t p = (size_t)ky * hnx + kx;
            for (int field = 0; field < 2; field++) {
                for (int k = 0; k < nz; k++) g[k] = cmk(0.0, 0.0);
                for (int k = 0; k <= degmax; k++) {
                    double re = gauss(&s);
                    double im = gauss(&s);
                    g[k] = cmk(re * scale, im * scale);
                }
                mul_one_minus_y2(g, t, nz);
                if (field == 0) mul_one_minus_y2(t, g, nz);
                const cx* res = field == 0 ? g : t;
                double* dst = (field == 0 ? v_hat : eta_hat) + 2 * (size_t)stride * p;
                for (int k = 0; k < nz; k++) { dst[2 * k] = res[k].re; dst[2 * k + 1] = res[k].im; }
            }
        }
    /* Hermitian symmetry on kx = 0 */
    for (int ky = 1; ky < ny / 2; ky++) {
        size_t pp = (size_t)ky * hnx, pm = (size_t)(ny - ky) * hnx;
        for (int f = 0; f < 2; f++) {
            double* a = f ? eta_hat : v_hat;
            for (int k = 0; k < nz; k++) {
                a[2 * ((size_t)stride * pm + k)] = a[2 * ((size_t)stride * pp + k)];
                a[2 * ((size_t)stride * pm + k) + 1] = -a[2 * ((size_t)stride * pp + k) + 1];
            }
        }
    }
    /* amplitude: the factor of the Gaussian coefficients, or (peak_bound) the bound of the
     * peak perturbation velocity */
    const double bound = peak_bound ? seed_velocity_bound(prm, v_hat, eta_hat, stride, nz, ny, hnx) : 1.0;
    if (bound > 0.0) {
        const double f = peak_bound ? amplitude / bound : amplitude;
        for (size_t p = 0; p < P; p++)
            for (int k = 0; k < 2 * nz; k++) {
                v_hat[2 * (size_t)stride * p + k] *= f;
                eta_hat[2 * (size_t)stride * p + k] *= f;
            }
    }
    free(g);
    free(t);
    return RPCF_OK;
}

/* profiling symbols of include/rpcf.h: the oracle has no device kernels */
int rpcf_timing(rpcf_ctx* c, int enable) { (void)enable; return c ? RPCF_OK : RPCF_EINVAL; }
int rpcf_timing_kernels(rpcf_ctx* c, unsigned mask) { (void)mask; return c ? RPCF_OK : RPCF_EINVAL; }
int rpcf_timing_get(rpcf_ctx* c, int kid, double* total_ms, long* count) {
    (void)c; (void)kid; (void)total_ms; (void)count;
    return RPCF_EINVAL;
}
const char* rpcf_kernel_name(int kid) { (void)kid; return NULL; }
int rpcf_num_kernels(void) { return 0; }

/* number of OpenMP threads the oracle uses (reported with the CPU baseline) */
int rpcf_oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
