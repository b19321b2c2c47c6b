// rpcf_hip.hip -- MI355X (gfx950) implementation of the rpcf C ABI (include/rpcf.h):
// context, device memory, launches, and the host side of the reference's step
// orchestration (nextStep, source/solver.cpp:35-101), with the whole step resident on
// the GPU (no per-step host<->device copies, no helper thread).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rpcf.h"
#include "host_prep.hpp"
#include "k_linear.hpp"
#include "k_transform.hpp"

namespace {

enum KernelId {
    KID_VELOCITY = 0, KID_S2P, KID_SPAN_INV, KID_LAMB, KID_SPAN_FWD, KID_P2S_NL, KID_LIN_TV, KID_LINEAR, KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {"k_velocity", "k_cheb_s2p", "k_span_inv", "k_lamb_phys",
                                             "k_span_fwd", "k_p2s_nl",   "k_lin_tv",   "k_lin_solve"};

constexpr int kFftThreads = 256;

struct Pending {
    int kid;
    hipEvent_t a, b;
};

}  // namespace

enum { FIELDS_STALE = 0, FIELDS_FRESH = 1, FIELDS_EXPLICIT = 2 };

struct rpcf_ctx {
    rpcf_params prm{};
    Dims d{};
    LinParams lp{};
    double Ro = 0.0, fac = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;
    int lds_limit = 65536;
    // tiles: sequences per block and LDS row stride (elements) of each transform kernel
    int TB_s2p = 1, TB_nl = 1, TK_inv = 1, TK_fwd = 1, RP = 1;
    int ld_zn = 0, ld_my = 0, ld_mx = 0;
    FftPlan pl_x{}, pl_y{}, pl_zp{}, pl_zn{};
    int cap_x = 0, cap_y = 0, cap_zp = 0, cap_zn = 0;  // fft_capacity of each plan
    // device memory
    double2* F = nullptr;   // [6][P][nz]
    double2* sv = nullptr;  // [P][nz] v-hat (u0 in pencil 0)
    double2* se = nullptr;  // [P][nz] eta-hat (w0 in pencil 0)
    double2* nl[4] = {nullptr, nullptr, nullptr, nullptr};  // N_v, N_v prev, N_eta, N_eta prev
    double2* Y = nullptr;   // [6][hnx][pz][ny] (also Z1 [3][hnx][pz][ny])
    double2* W = nullptr;   // [6][pz][my][nxh]
    double2* Z2 = nullptr;  // [3][pz][my][zld], zld = hnx rounded up to 8 (128-byte rows)
    int zld = 0;
    double2* minv = nullptr;
    double2* Wlin = nullptr;  // [5][P][nz] operator products of the linear right-hand sides
    double *T0 = nullptr, *T2 = nullptr, *T4 = nullptr, *U = nullptr, *dU = nullptr, *ddU = nullptr;
    // transposed tables for k_linear: T0, T2, and the quotients T4*dt*0.5/Re, T2*0.5/Re*dt
    double *T0t = nullptr, *T2t = nullptr, *T4q = nullptr, *T2q = nullptr;
    double* Tf = nullptr;  // [5][RT][S][64] A-operand tiles of k_lin_tv_mfma
    int tv_rt = 0, tv_s = 0;
    double *meanU = nullptr, *qw = nullptr, *tw = nullptr;
    std::vector<double> qw_host;
    double2 q_target = make_double2(0.0, 0.0);
    int have_state = 0;
    // the six spectral fields F relative to the state: FIELDS_STALE = velocity(state), not
    // computed yet; FIELDS_FRESH = computed from the current state; FIELDS_EXPLICIT = set by
    // the caller, or held over a linear step, so not derivable from the state
    int fields = 0;
    // per-kernel timing with hipEvents on the context stream
    int timing = 0;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double t_ms[KID_COUNT] = {0};
    long t_cnt[KID_COUNT] = {0};
    char err[512] = {0};
};

namespace {

int fail(rpcf_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(c, call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return fail((c), RPCF_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int check_sizes(const rpcf_params* prm, Dims* out) {
    if (!prm) return RPCF_EINVAL;
    const int mx = prm->mx, my = prm->my, mz = prm->mz;
    if (mx < 6 || my < 6 || mz < 16 || mx % 2 || my % 2 || mz % 4) return RPCF_EINVAL;
    if (!(prm->Re > 0) || !(prm->dt > 0) || !(prm->n_pi_x > 0) || !(prm->n_pi_y > 0)) return RPCF_EINVAL;
    if (prm->precision != 0) return RPCF_EINVAL;
    Dims d;
    d.mx = mx; d.my = my;
    d.nx = mx / 3 * 2; d.ny = my / 3 * 2;
    d.nz = mz / 4 + 1; d.N = d.nz - 1; d.pz = mz / 2 + 1;
    d.hnx = d.nx / 2 + 1; d.nxh = d.nx / 2; d.P = d.hnx * d.ny;
    d.nslot = d.hnx * (d.ny / 2 + 1);
    if (d.nz > 256 || d.ny < 2 || d.nx < 2) return RPCF_EINVAL;
    int rad[16], nst;
    if (!rpcf_host::factorize(mx, rad, &nst) || !rpcf_host::factorize(my, rad, &nst) ||
        !rpcf_host::factorize(4 * d.N, rad, &nst) || !rpcf_host::factorize(2 * d.N, rad, &nst))
        return RPCF_EINVAL;
    if (out) *out = d;
    return RPCF_OK;
}

hipEvent_t get_event(rpcf_ctx* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Wait for the stream, then fold the recorded kernel times into the totals.
int sync_ctx(rpcf_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->t_ms[p.kid] += ms;
            c->t_cnt[p.kid] += 1;
        }
        c->pool.push_back(p.a);
        c->pool.push_back(p.b);
    }
    c->pending.clear();
    return RPCF_OK;
}

template <typename F>
int launch(rpcf_ctx* c, int kid, F&& fn) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c->timing) {
        a = get_event(c);
        (void)hipEventRecord(a, c->stream);
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, RPCF_EDEVICE, "launch %s: %s", kKernelNames[kid], hipGetErrorString(e));
    if (c->timing) {
        b = get_event(c);
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({kid, a, b});
    }
    return RPCF_OK;
}

template <typename T>
int dalloc(rpcf_ctx* c, T** p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return fail(c, RPCF_ENOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * count, hipGetErrorString(e));
    return RPCF_OK;
}

int make_plan(rpcf_ctx* c, int n, double* tw_dev, FftPlan* pl, int* cap) {
    pl->n = n;
    if (!rpcf_host::factorize(n, pl->rad, &pl->nst)) return fail(c, RPCF_EINVAL, "FFT length %d not 2^a3^b5^c", n);
    *cap = fft_capacity(pl->rad, pl->nst, kFftThreads);
    std::vector<double> tw(2 * (size_t)n);
    rpcf_host::twiddles(n, tw.data());
    HIPCHK(c, hipMemcpy(tw_dev, tw.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    pl->tw = tw_dev;
    return RPCF_OK;
}

// largest tile t <= tmax with t*seq_len elements within the plan's per-block FFT
// capacity and lds(t) bytes within the LDS limit
template <typename L>
int pick_tile(int tmax, int seq_len, int cap, int limit, L lds_bytes) {
    for (int t = tmax; t >= 1; t--)
        if ((long)t * seq_len <= (long)cap && lds_bytes(t) <= (size_t)limit) return t;
    return 0;
}

size_t spec_elems(const rpcf_ctx* c) { return (size_t)c->d.P * c->d.nz; }

int upload_spec(rpcf_ctx* c, double2* dst, const double* src, long stride) {
    if (!src || stride < c->d.nz) return fail(c, RPCF_EINVAL, "bad spectral array or stride");
    const int nz = c->d.nz;
    std::vector<double2> h(spec_elems(c));
    for (size_t p = 0; p < (size_t)c->d.P; p++)
        for (int k = 0; k < nz; k++) {
            const double* s = src + 2 * ((size_t)stride * p + k);
            h[p * nz + k] = make_double2(s[0], s[1]);
        }
    HIPCHK(c, hipMemcpyAsync(dst, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPCF_OK;
}

int download_spec(rpcf_ctx* c, double* dst, const double2* src, long stride) {
    if (!dst || stride < c->d.nz) return fail(c, RPCF_EINVAL, "bad spectral array or stride");
    const int nz = c->d.nz;
    std::vector<double2> h(spec_elems(c));
    int rc = sync_ctx(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(h.data(), src, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
    for (size_t p = 0; p < (size_t)c->d.P; p++)
        for (int k = 0; k < nz; k++) {
            double* s = dst + 2 * ((size_t)stride * p + k);
            s[0] = h[p * nz + k].x;
            s[1] = h[p * nz + k].y;
        }
    return RPCF_OK;
}

// ---- step pieces (asynchronous on c->stream) ------------------------------------------
int enq_velocity(rpcf_ctx* c) {
    const Dims d = c->d;
    const int thr = 64 * ((d.nz + 63) / 64);
    return launch(c, KID_VELOCITY, [&] {
        k_velocity<<<d.P, thr, 0, c->stream>>>(c->sv, c->se, c->F, (long)d.P * d.nz, d, c->lp);
    });
}

int enq_nonlinear(rpcf_ctx* c) {
    const Dims d = c->d;
    int rc;
    const long fS = (long)d.P * d.nz, yS = (long)d.hnx * d.pz * d.ny, wS = (long)d.pz * d.my * d.nxh,
               zS = (long)d.pz * d.my * c->zld;
    const double2* tw4 = (const double2*)c->pl_zp.tw;
    if (c->fields == FIELDS_STALE) {
        if ((rc = enq_velocity(c))) return rc;
        c->fields = FIELDS_FRESH;
    }
    {
        const int TB = c->TB_s2p, ld = c->ld_zn;
        dim3 g((d.ny + TB - 1) / TB, d.hnx, 6);
        size_t lds = sizeof(double2) * TB * ld;
        rc = launch(c, KID_S2P, [&] {
            switch (2 * d.N) {
#define RPCF_CASE(L) case L: k_cheb_s2p<1, L><<<g, kFftThreads, lds, c->stream>>>(c->F, fS, c->Y, yS, d, c->pl_zn, tw4, TB, ld); break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: k_cheb_s2p<1><<<g, kFftThreads, lds, c->stream>>>(c->F, fS, c->Y, yS, d, c->pl_zn, tw4, TB, ld);
            }
        });
        if (rc) return rc;
    }
    {
        const int TK = c->TK_inv, ld = c->ld_my;
        dim3 g((d.nxh + TK - 1) / TK, d.pz, 6);
        size_t lds = sizeof(double2) * TK * ld;
        rc = launch(c, KID_SPAN_INV, [&] {
            switch (d.my) {
#define RPCF_CASE(L) case L: k_span_inv<L><<<g, kFftThreads, lds, c->stream>>>(c->Y, yS, c->W, wS, d, d.pz, c->pl_y, TK, ld); break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: k_span_inv<<<g, kFftThreads, lds, c->stream>>>(c->Y, yS, c->W, wS, d, d.pz, c->pl_y, TK, ld);
            }
        });
        if (rc) return rc;
    }
    {
        const int RP = c->RP, ld = c->ld_mx;
        dim3 g((d.my / 2 + RP - 1) / RP, d.pz);
        size_t lds = sizeof(double2) * 6 * RP * ld;
        rc = launch(c, KID_LAMB, [&] {
            switch (d.mx) {
#define RPCF_CASE(L) case L: k_lamb_phys<L><<<g, kFftThreads, lds, c->stream>>>(c->W, wS, c->Z2, zS, c->zld, c->meanU, c->Ro, c->fac, d, c->pl_x, RP, ld); break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: k_lamb_phys<<<g, kFftThreads, lds, c->stream>>>(c->W, wS, c->Z2, zS, c->zld, c->meanU, c->Ro, c->fac, d,
                                                                         c->pl_x, RP, ld);
            }
        });
        if (rc) return rc;
    }
    {
        const int TK = c->TK_fwd, ld = c->ld_my;
        dim3 g((d.hnx + TK - 1) / TK, d.pz, 3);
        size_t lds = sizeof(double2) * TK * ld;
        rc = launch(c, KID_SPAN_FWD, [&] {
            switch (d.my) {
#define RPCF_CASE(L) case L: k_span_fwd<L><<<g, kFftThreads, lds, c->stream>>>(c->Z2, zS, c->zld, c->Y, yS, d, d.pz, c->pl_y, TK, ld); break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: k_span_fwd<<<g, kFftThreads, lds, c->stream>>>(c->Z2, zS, c->zld, c->Y, yS, d, d.pz, c->pl_y, TK, ld);
            }
        });
        if (rc) return rc;
    }
    // AB2 history swap (source/rhs.cpp:107-108): the new N lands in the old "previous" buffers
    std::swap(c->nl[0], c->nl[1]);
    std::swap(c->nl[2], c->nl[3]);
    {
        const int TB = c->TB_nl, ld = c->ld_zn;
        dim3 g((d.ny + TB - 1) / TB, d.hnx);
        size_t lds = sizeof(double2) * ((size_t)TB * ld + 256);
        rc = launch(c, KID_P2S_NL, [&] {
            switch (2 * d.N) {
#define RPCF_CASE(L) case L: k_p2s_nl<L><<<g, kFftThreads, lds, c->stream>>>(c->Y, yS, c->nl[0], c->nl[2], d, c->pl_zn, tw4, c->lp.aphi, c->lp.beta, TB, ld); break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: k_p2s_nl<<<g, kFftThreads, lds, c->stream>>>(c->Y, yS, c->nl[0], c->nl[2], d, c->pl_zn, tw4,
                                                                      c->lp.aphi, c->lp.beta, TB, ld);
            }
        });
    }
    return rc;
}

constexpr int kTvPencils = 16;  // pencils per block of k_lin_tv

int enq_linear(rpcf_ctx* c) {
    const Dims d = c->d;
    LinState st;
    st.nlv = c->nl[0]; st.nlvp = c->nl[1]; st.nle = c->nl[2]; st.nlep = c->nl[3];
    st.sv = c->sv; st.se = c->se; st.minv = c->minv;
    st.qw_target_re = c->q_target.x; st.qw_target_im = c->q_target.y;
    st.flow_control = c->prm.flow_control;
    int rc = launch(c, KID_LIN_TV, [&] {
        dim3 g((d.P + kTvBlockPencils - 1) / kTvBlockPencils, 5);
        const size_t lds = sizeof(double) * kTvChunk * c->tv_s * 64;
        switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                    \
    case L:                                                                                             \
        k_lin_tv_mfma<(L / 2 + 4) / 4><<<g, 256, lds, c->stream>>>(c->sv, c->se, c->Wlin, d, c->Tf, c->tv_rt); \
        break;
            RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
            default: {
                const int gv = (d.P + kTvPencils - 1) / kTvPencils;
                const size_t lv = sizeof(double2) * 2 * (size_t)d.nz * (kTvPencils + 1);
                k_lin_tv<kTvPencils><<<gv, 384, lv, c->stream>>>(c->sv, c->se, c->Wlin, d, c->T0t, c->T2t, c->T4q,
                                                                 c->T2q);
            }
        }
    });
    if (rc) return rc;
    return launch(c, KID_LINEAR, [&] {
        const int thr = 64 * ((d.nz + 63) / 64);
        k_lin_solve<<<d.nslot, thr, 0, c->stream>>>(st, c->Wlin, d, c->lp, c->T0t, c->U, c->dU, c->ddU, c->qw);
    });
}

void free_ctx(rpcf_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->Wlin, c->F, c->sv, c->se, c->nl[0], c->nl[1], c->nl[2], c->nl[3], c->Y, c->W, c->Z2, c->minv,
                    c->T0, c->T2, c->T4, c->U, c->dU, c->ddU, c->meanU, c->qw, c->tw, c->T0t, c->T2t, c->T4q, c->T2q,
                    c->Tf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Raise a kernel's dynamic-LDS cap to `bytes`; a refusal leaves the default cap, and the
// sticky error it sets is cleared so that the next launch check does not report it.
template <typename K>
void allow_lds(K kernel, int bytes) {
    if (hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        (void)hipGetLastError();
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

int rpcf_sizes(const rpcf_params* prm, int* nx, int* ny, int* nz, int* pz, int* hnx) {
    Dims d;
    int rc = check_sizes(prm, &d);
    if (rc) return rc;
    if (nx) *nx = d.nx;
    if (ny) *ny = d.ny;
    if (nz) *nz = d.nz;
    if (pz) *pz = d.pz;
    if (hnx) *hnx = d.hnx;
    return RPCF_OK;
}

const char* rpcf_last_error(const rpcf_ctx* c) { return c ? c->err : "null context"; }

void rpcf_destroy(rpcf_ctx* c) { free_ctx(c); }

}  // extern "C"

namespace {

// allocation and set-up of a context; on failure the caller frees it
int init_ctx(rpcf_ctx* c, const rpcf_params* prm) {
    const Dims d = c->d;
    int rc;
    // problem(RPCF_Paras): include/data.h:321-322
    c->lp.aphi = prm->n_pi_x / 2;
    c->lp.beta = prm->n_pi_y / 2;
    c->lp.Re = prm->Re;
    c->lp.dt = prm->dt;
    c->Ro = prm->Ro;
    c->fac = 1.0 / d.mx / d.my;  // normalize factor, source/transform.cu:267
    HIPCHK(c, hipGetDevice(&c->device));
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int lim = 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && lim > 0)
        c->lds_limit = lim;

    const size_t P = d.P, nz = d.nz;
    if ((rc = dalloc(c, &c->F, 6 * P * nz))) return rc;
    if ((rc = dalloc(c, &c->sv, P * nz))) return rc;
    if ((rc = dalloc(c, &c->se, P * nz))) return rc;
    for (int i = 0; i < 4; i++)
        if ((rc = dalloc(c, &c->nl[i], P * nz))) return rc;
    if ((rc = dalloc(c, &c->Y, 6 * (size_t)d.hnx * d.pz * d.ny))) return rc;
    if ((rc = dalloc(c, &c->W, 6 * (size_t)d.pz * d.my * d.nxh))) return rc;
    c->zld = (d.hnx + 7) & ~7;
    if ((rc = dalloc(c, &c->Z2, 3 * (size_t)d.pz * d.my * c->zld))) return rc;
    if ((rc = dalloc(c, &c->minv, 2 * (size_t)d.nslot * nz * nz))) return rc;
    if ((rc = dalloc(c, &c->Wlin, 5 * P * nz))) return rc;
    if ((rc = dalloc(c, &c->T0, nz * nz)) || (rc = dalloc(c, &c->T2, nz * nz)) || (rc = dalloc(c, &c->T4, nz * nz)))
        return rc;
    if ((rc = dalloc(c, &c->T0t, nz * nz)) || (rc = dalloc(c, &c->T2t, nz * nz)) || (rc = dalloc(c, &c->T4q, nz * nz)) ||
        (rc = dalloc(c, &c->T2q, nz * nz)))
        return rc;
    if ((rc = dalloc(c, &c->U, nz)) || (rc = dalloc(c, &c->dU, nz)) || (rc = dalloc(c, &c->ddU, nz)) ||
        (rc = dalloc(c, &c->qw, nz)) || (rc = dalloc(c, &c->meanU, (size_t)d.pz)))
        return rc;
    if ((rc = dalloc(c, &c->tw, 2 * ((size_t)d.mx + d.my + 4 * d.N + 2 * d.N)))) return rc;
    for (int i = 0; i < 4; i++) HIPCHK(c, hipMemset(c->nl[i], 0, sizeof(double2) * P * nz));
    HIPCHK(c, hipMemset(c->F, 0, sizeof(double2) * 6 * P * nz));
    HIPCHK(c, hipMemset(c->sv, 0, sizeof(double2) * P * nz));
    HIPCHK(c, hipMemset(c->se, 0, sizeof(double2) * P * nz));

    // operators (initAuxMatrix, source/solver.cpp:265-269)
    std::vector<double> T0(nz * nz), T2(nz * nz), T4(nz * nz), U(nz), dU(nz), ddU(nz), mU(d.pz);
    rpcf_host::t_matrices(d.N, T0.data(), T2.data(), T4.data());
    rpcf_host::base_profile(d.nz, prm->u_profile, U.data(), dU.data(), ddU.data());
    rpcf_host::base_profile_padded(d.pz, prm->u_profile, mU.data());
    c->qw_host.resize(nz);
    for (size_t k = 0; k < nz; k++) c->qw_host[k] = (k % 2 == 0) ? 2.0 / (1.0 - (double)k * (double)k) : 0.0;
    HIPCHK(c, hipMemcpy(c->T0, T0.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->T2, T2.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->T4, T4.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    {
        // transposed T0, and the quotients T2*0.5/Re*dt and T4*dt*0.5/Re of
        // source/linear.cpp:100, 142 evaluated in the reference's order
        std::vector<double> t0(nz * nz), t2(nz * nz), q4(nz * nz), q2(nz * nz);
        const double Re = prm->Re, dt = prm->dt;
        for (size_t i = 0; i < nz; i++)
            for (size_t j = 0; j < nz; j++) {
                t0[j * nz + i] = T0[i * nz + j];
                t2[j * nz + i] = T2[i * nz + j];
                q2[j * nz + i] = T2[i * nz + j] * 0.5 / Re * dt;
                q4[j * nz + i] = T4[i * nz + j] * dt * 0.5 / Re;
            }
        HIPCHK(c, hipMemcpy(c->T0t, t0.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->T2t, t2.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->T4q, q4.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->T2q, q2.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        // the five operators (T0, T2, T4q for v; T0, T2q for eta) as MFMA A-operand tiles
        const int RT = (int)((nz + 15) / 16), S = (int)((nz + 3) / 4);
        c->tv_rt = RT;
        c->tv_s = S;
        const std::vector<double>* ops[5] = {&t0, &t2, &q4, &t0, &q2};  // transposed: op[j*nz + i] = T[i][j]
        std::vector<double> tf((size_t)5 * RT * S * 64, 0.0);
        for (int m = 0; m < 5; m++)
            for (int rt = 0; rt < RT; rt++)
                for (int sk = 0; sk < S; sk++)
                    for (int l = 0; l < 64; l++) {
                        const size_t i = 16 * rt + (l & 15), k = 4 * sk + (l >> 4);
                        if (i < nz && k < nz) tf[(((size_t)m * RT + rt) * S + sk) * 64 + l] = (*ops[m])[k * nz + i];
                    }
        if ((rc = dalloc(c, &c->Tf, tf.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->Tf, tf.data(), sizeof(double) * tf.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(c->U, U.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dU, dU.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->ddU, ddU.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->qw, c->qw_host.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->meanU, mU.data(), sizeof(double) * d.pz, hipMemcpyHostToDevice));

    // FFT plans (the reference's cuFFT plans, source/transform.cu:49-71)
    double* tw = c->tw;
    if ((rc = make_plan(c, d.mx, tw, &c->pl_x, &c->cap_x))) return rc;
    tw += 2 * d.mx;
    if ((rc = make_plan(c, d.my, tw, &c->pl_y, &c->cap_y))) return rc;
    tw += 2 * d.my;
    if ((rc = make_plan(c, 4 * d.N, tw, &c->pl_zp, &c->cap_zp))) return rc;
    tw += 2 * 4 * d.N;
    if ((rc = make_plan(c, 2 * d.N, tw, &c->pl_zn, &c->cap_zn))) return rc;

    // LDS tiles
    lim = c->lds_limit;
    // Chebyshev kernels: tiles sized for three resident blocks per CU
    const int lim3 = lim / 3;
    c->ld_zn = 2 * d.N + 2;  // >= 2N (FFT) and >= 2 nz (L_x, L_y of k_p2s_nl)
    c->ld_my = d.my + 1;
    c->ld_mx = d.mx + 1;
    c->TB_s2p = pick_tile(kChebTile, 2 * d.N, c->cap_zn, lim3, [&](int t) { return sizeof(double2) * (size_t)t * c->ld_zn; });
    c->TB_nl = pick_tile(kChebTile, 2 * d.N, c->cap_zn, lim3, [&](int t) {
        return sizeof(double2) * ((size_t)t * c->ld_zn + 256);
    });
    c->TK_inv = pick_tile(kSpanTile, d.my, c->cap_y, lim3, [&](int t) { return sizeof(double2) * (size_t)t * c->ld_my; });
    c->TK_fwd = c->TK_inv;
    c->RP = pick_tile(kLambPairs, 6 * d.mx, c->cap_x, lim3, [&](int t) { return sizeof(double2) * 6 * (size_t)t * c->ld_mx; });
    if (!c->TB_s2p || !c->TB_nl || !c->TK_inv || !c->RP)
        return fail(c, RPCF_EINVAL, "grid too large for the LDS tiles (limit %d bytes)", lim);
    allow_lds(k_cheb_s2p<1>, lim);
    allow_lds(k_cheb_s2p<0>, lim);
    allow_lds(k_span_inv<0>, lim);
    allow_lds(k_span_fwd<0>, lim);
    allow_lds(k_lamb_phys<0>, lim);
    allow_lds(k_p2s_nl<0>, lim);
#define RPCF_CASE(L)                  \
    allow_lds(k_cheb_s2p<1, L>, lim); \
    allow_lds(k_cheb_s2p<0, L>, lim); \
    allow_lds(k_span_inv<L>, lim);    \
    allow_lds(k_span_fwd<L>, lim);    \
    allow_lds(k_lamb_phys<L>, lim);   \
    allow_lds(k_p2s_nl<L>, lim);      \
    allow_lds(k_lin_tv_mfma<(L / 2 + 4) / 4>, lim);
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    allow_lds(k_cheb_p2s, lim);
    allow_lds(k_x_c2r, lim);
    allow_lds(k_x_r2c, lim);
    allow_lds(k_lin_tv<kTvPencils>, lim);
    (void)hipGetLastError();
    if (getenv("RPCF_VERBOSE"))
        fprintf(stderr, "rpcf: lds_limit=%d TB_s2p=%d TB_nl=%d TK=%d RP=%d nslot=%d\n", lim, c->TB_s2p, c->TB_nl,
                c->TK_inv, c->RP, d.nslot);

    // collocation matrices and inverses, in chunks of slots (initAuxMatrix)
    const size_t per = nz * nz;
    const int chunk =
        (int)std::min<size_t>((size_t)d.nslot, std::max<size_t>(64, ((size_t)1 << 29) / (2 * per * sizeof(double2))));
    double2* work = nullptr;
    int* status = nullptr;
    if ((rc = dalloc(c, &work, 2 * (size_t)chunk * per))) return rc;
    if ((rc = dalloc(c, &status, 1))) { (void)hipFree(work); return rc; }
    (void)hipMemset(status, 0, sizeof(int));
    for (int s0 = 0; s0 < d.nslot && !rc; s0 += chunk) {
        dim3 g(std::min(chunk, d.nslot - s0), 2);
        k_build_inverse<<<g, 256, 0, c->stream>>>(c->minv, work, s0, d, c->lp, c->T0, c->T2, c->T4, c->U, c->ddU,
                                                  status);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, RPCF_EDEVICE, "k_build_inverse: %s", hipGetErrorString(e));
    }
    int hstatus = 0;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "matrix build failed");
    if (!rc && hipMemcpy(&hstatus, status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, RPCF_EDEVICE, "matrix build status copy failed");
    (void)hipFree(work);
    (void)hipFree(status);
    if (!rc && hstatus) rc = fail(c, RPCF_EINVAL, "%d singular collocation matrices", hstatus);
    return rc;
}

}  // namespace

extern "C" {

int rpcf_create(const rpcf_params* prm, rpcf_ctx** out) {
    if (!out) return RPCF_EINVAL;
    *out = nullptr;
    Dims d;
    if (check_sizes(prm, &d)) return RPCF_EINVAL;
    rpcf_ctx* c = new rpcf_ctx();
    c->prm = *prm;
    c->d = d;
    int rc = init_ctx(c, prm);
    if (rc) {
        fprintf(stderr, "rpcf_create: %s\n", c->err);
        free_ctx(c);
        return rc;
    }
    *out = c;
    return RPCF_OK;
}

int rpcf_set_state(rpcf_ctx* c, const double* v_hat, const double* eta_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    int rc = upload_spec(c, c->sv, v_hat, stride);
    if (!rc) rc = upload_spec(c, c->se, eta_hat, stride);
    if (rc) return rc;
    c->have_state = 1;
    // initFlow tail (source/RPCFKernels.cu:413-427); the flow-rate target is the initial bulk of u0
    double2 q = make_double2(0.0, 0.0);
    for (int k = 0; k < c->d.nz; k++) {
        const double* s = v_hat + 2 * k;
        q = make_double2(q.x + s[0] * c->qw_host[k], q.y + s[1] * c->qw_host[k]);
    }
    c->q_target = q;
    c->fields = FIELDS_STALE;
    return RPCF_OK;
}

int rpcf_get_state(rpcf_ctx* c, double* v_hat, double* eta_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    int rc = download_spec(c, v_hat, c->sv, stride);
    return rc ? rc : download_spec(c, eta_hat, c->se, stride);
}

int rpcf_set_field(rpcf_ctx* c, int field, const double* f_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (field < 0 || field > 5) return fail(c, RPCF_EINVAL, "bad field id %d", field);
    if (c->fields == FIELDS_STALE && c->have_state) {
        // the other five fields keep their values: materialise them first
        int rc = enq_velocity(c);
        if (rc) return rc;
    }
    int rc = upload_spec(c, c->F + (size_t)field * spec_elems(c), f_hat, stride);
    if (!rc) {
        c->have_state = 1;
        c->fields = FIELDS_EXPLICIT;
    }
    return rc;
}

int rpcf_get_field(rpcf_ctx* c, int field, double* f_hat, long stride) {
    if (!c) return RPCF_EINVAL;
    if (field < 0 || field > 5) return fail(c, RPCF_EINVAL, "bad field id %d", field);
    if (c->fields == FIELDS_STALE && c->have_state) {
        int rc = enq_velocity(c);
        if (rc) return rc;
        c->fields = FIELDS_FRESH;
    }
    return download_spec(c, f_hat, c->F + (size_t)field * spec_elems(c), stride);
}

int rpcf_get_nonlinear(rpcf_ctx* c, int prev, double* n_v, double* n_eta, long stride) {
    if (!c) return RPCF_EINVAL;
    int rc = download_spec(c, n_v, c->nl[prev ? 1 : 0], stride);
    return rc ? rc : download_spec(c, n_eta, c->nl[prev ? 3 : 2], stride);
}

int rpcf_set_nonlinear(rpcf_ctx* c, int prev, const double* n_v, const double* n_eta, long stride) {
    if (!c) return RPCF_EINVAL;
    int rc = upload_spec(c, c->nl[prev ? 1 : 0], n_v, stride);
    return rc ? rc : upload_spec(c, c->nl[prev ? 3 : 2], n_eta, stride);
}

int rpcf_nonlinear(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return fail(c, RPCF_ESTATE, "rpcf_nonlinear: no state set");
    int rc = enq_nonlinear(c);
    return rc ? rc : sync_ctx(c);
}

int rpcf_linear(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return fail(c, RPCF_ESTATE, "rpcf_linear: no state set");
    // the fields keep the velocity of the state before the solve until rpcf_velocity
    if (c->fields == FIELDS_STALE) {
        int rc = enq_velocity(c);
        if (rc) return rc;
    }
    int rc = enq_linear(c);
    c->fields = FIELDS_EXPLICIT;
    return rc ? rc : sync_ctx(c);
}

int rpcf_velocity(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return fail(c, RPCF_ESTATE, "rpcf_velocity: no state set");
    int rc = enq_velocity(c);
    c->fields = FIELDS_FRESH;
    return rc ? rc : sync_ctx(c);
}

int rpcf_step(rpcf_ctx* c, int nsteps) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return fail(c, RPCF_ESTATE, "rpcf_step: no state set");
    // nextStep + getUVW: the velocity of the new state is formed when it is next needed
    // (by the next nonlinear evaluation, or by rpcf_get_field)
    for (int s = 0; s < nsteps; s++) {
        int rc = enq_nonlinear(c);
        if (!rc) rc = enq_linear(c);
        if (rc) return rc;
        c->fields = FIELDS_STALE;
    }
    return sync_ctx(c);
}

int rpcf_bootstrap(rpcf_ctx* c) {
    if (!c) return RPCF_EINVAL;
    if (!c->have_state) return fail(c, RPCF_ESTATE, "rpcf_bootstrap: no state set");
    int rc = enq_nonlinear(c);
    if (rc) return rc;
    return rpcf_step(c, 1);
}

int rpcf_spec_to_phys(rpcf_ctx* c, const double* spec, double* phys, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!spec || !phys) return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys: null array");
    const Dims d = c->d;
    const int planes = padded ? d.pz : d.nz;
    const FftPlan& plz = c->pl_zn;
    const int L = 2 * d.N, ld = L + 1;
    const int TB = pick_tile(16, L, c->cap_zn, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * ld; });
    const int RP1 = pick_tile(8, d.mx, c->cap_x, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * c->ld_mx; });
    const double2* tw4 = (const double2*)c->pl_zp.tw;
    double2 *f = nullptr, *y = nullptr, *w = nullptr;
    double* ph = nullptr;
    int rc = dalloc(c, &f, spec_elems(c));
    if (!rc) rc = dalloc(c, &y, (size_t)d.hnx * planes * d.ny);
    if (!rc) rc = dalloc(c, &w, (size_t)planes * d.my * d.nxh);
    if (!rc) rc = dalloc(c, &ph, (size_t)planes * d.my * d.mx);
    if (!rc) rc = upload_spec(c, f, spec, stride);
    if (!rc) {
        dim3 g((d.ny + TB - 1) / TB, d.hnx, 1);
        if (padded) k_cheb_s2p<1><<<g, kFftThreads, sizeof(double2) * TB * ld, c->stream>>>(f, 0, y, 0, d, plz, tw4, TB, ld);
        else k_cheb_s2p<0><<<g, kFftThreads, sizeof(double2) * TB * ld, c->stream>>>(f, 0, y, 0, d, plz, tw4, TB, ld);
        dim3 g2((d.nxh + c->TK_inv - 1) / c->TK_inv, planes, 1);
        k_span_inv<0><<<g2, kFftThreads, sizeof(double2) * c->TK_inv * c->ld_my, c->stream>>>(y, 0, w, 0, d, planes,
                                                                                           c->pl_y, c->TK_inv, c->ld_my);
        dim3 g3((d.my / 2 + RP1 - 1) / RP1, planes);
        k_x_c2r<<<g3, kFftThreads, sizeof(double2) * RP1 * c->ld_mx, c->stream>>>(w, ph, d, c->pl_x, RP1, c->ld_mx);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, RPCF_EDEVICE, "spec_to_phys launch: %s", hipGetErrorString(e));
    }
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "spec_to_phys sync");
    if (!rc && hipMemcpy(phys, ph, sizeof(double) * planes * d.my * d.mx, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, RPCF_EDEVICE, "spec_to_phys copy");
    for (void* p : {(void*)f, (void*)y, (void*)w, (void*)ph})
        if (p) (void)hipFree(p);
    return rc;
}

int rpcf_phys_to_spec(rpcf_ctx* c, const double* phys, double* spec, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!spec || !phys) return fail(c, RPCF_EINVAL, "rpcf_phys_to_spec: null array");
    const Dims d = c->d;
    const int planes = padded ? d.pz : d.nz;
    const FftPlan& plz = padded ? c->pl_zp : c->pl_zn;
    const int capz = padded ? c->cap_zp : c->cap_zn;
    const int L = padded ? 4 * d.N : 2 * d.N, ld = L + 1;
    const int TB = pick_tile(16, L, capz, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * ld; });
    const int RP1 = pick_tile(8, d.mx, c->cap_x, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * c->ld_mx; });
    double2 *f = nullptr, *y = nullptr, *z = nullptr;
    double* ph = nullptr;
    int rc = dalloc(c, &f, spec_elems(c));
    if (!rc) rc = dalloc(c, &y, (size_t)d.hnx * planes * d.ny);
    if (!rc) rc = dalloc(c, &z, (size_t)planes * d.my * d.hnx);
    if (!rc) rc = dalloc(c, &ph, (size_t)planes * d.my * d.mx);
    if (!rc && hipMemcpy(ph, phys, sizeof(double) * planes * d.my * d.mx, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, RPCF_EDEVICE, "phys_to_spec copy");
    if (!rc) {
        dim3 g3((d.my / 2 + RP1 - 1) / RP1, planes);
        k_x_r2c<<<g3, kFftThreads, sizeof(double2) * RP1 * c->ld_mx, c->stream>>>(ph, z, d.hnx, c->fac, d, c->pl_x, RP1,
                                                                                  c->ld_mx);
        dim3 g2((d.hnx + c->TK_fwd - 1) / c->TK_fwd, planes, 1);
        k_span_fwd<0><<<g2, kFftThreads, sizeof(double2) * c->TK_fwd * c->ld_my, c->stream>>>(z, 0, d.hnx, y, 0, d, planes,
                                                                                           c->pl_y, c->TK_fwd, c->ld_my);
        dim3 g((d.ny + TB - 1) / TB, d.hnx);
        k_cheb_p2s<<<g, kFftThreads, sizeof(double2) * TB * ld, c->stream>>>(y, f, d, planes, plz, TB, ld);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, RPCF_EDEVICE, "phys_to_spec launch: %s", hipGetErrorString(e));
    }
    if (!rc) rc = download_spec(c, spec, f, stride);
    for (void* p : {(void*)f, (void*)y, (void*)z, (void*)ph})
        if (p) (void)hipFree(p);
    return rc;
}

int rpcf_get_inverse(rpcf_ctx* c, int eq, int pencil, double* out) {
    if (!c || !out || eq < 0 || eq > 1 || pencil < 0 || pencil >= c->d.P) return RPCF_EINVAL;
    const Dims d = c->d;
    const int kx = pencil % d.hnx, ky = pencil / d.hnx;
    const int kyh = ky <= d.ny / 2 ? ky : d.ny - ky;
    const size_t nn = (size_t)d.nz * d.nz;
    std::vector<double2> m(nn);
    int rc = sync_ctx(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(m.data(), c->minv + ((size_t)eq * d.nslot + (size_t)kyh * d.hnx + kx) * nn,
                        sizeof(double2) * nn, hipMemcpyDeviceToHost));
    for (int r = 0; r < d.nz; r++)
        for (int j = 0; j < d.nz; j++) {
            out[2 * ((size_t)r * d.nz + j)] = m[(size_t)j * d.nz + r].x;
            out[2 * ((size_t)r * d.nz + j) + 1] = m[(size_t)j * d.nz + r].y;
        }
    return RPCF_OK;
}

int rpcf_seed_state(const rpcf_params* prm, uint64_t seed, double amplitude, int kmax, int degmax, double kexp,
                    double* v_hat, double* eta_hat, long stride) {
    Dims d;
    if (check_sizes(prm, &d)) return RPCF_EINVAL;
    const int nz = d.nz, ny = d.ny, hnx = d.hnx, nx = d.nx;
    if (!v_hat || !eta_hat || stride < nz || degmax < 0 || degmax + 4 >= nz || kmax < 0) return RPCF_EINVAL;
    const size_t P = (size_t)d.P;
    memset(v_hat, 0, sizeof(double) * 2 * (size_t)stride * P);
    memset(eta_hat, 0, sizeof(double) * 2 * (size_t)stride * P);
    uint64_t s = seed;
    std::vector<double> g(2 * nz), t(2 * nz);
    for (int ky = 0; ky < ny; ky++)
        for (int kx = 0; kx < hnx; kx++) {
            const int kys = ky <= ny / 2 ? ky : ky - ny;
            if (kx == 0 && ky == 0) continue;
            if (kx == nx / 2 || ky == ny / 2) continue;
            if (kx > kmax || std::abs(kys) > kmax) continue;
            if (kx == 0 && kys < 0) continue;
            const double scale = amplitude * pow(1.0 + sqrt((double)kx * kx + (double)kys * kys), -kexp);
            const size_t p = (size_t)ky * hnx + kx;
            for (int field = 0; field < 2; field++) {
                std::fill(g.begin(), g.end(), 0.0);
                for (int k = 0; k <= degmax; k++) {
                    const double re = rpcf_host::gauss(&s);
                    const double im = rpcf_host::gauss(&s);
                    g[2 * k] = re * scale;
                    g[2 * k + 1] = im * scale;
                }
                rpcf_host::mul_one_minus_y2(g.data(), t.data(), nz);
                if (field == 0) rpcf_host::mul_one_minus_y2(t.data(), g.data(), nz);
                const double* res = field == 0 ? g.data() : t.data();
                double* dst = (field == 0 ? v_hat : eta_hat) + 2 * (size_t)stride * p;
                for (int k = 0; k < 2 * nz; k++) dst[k] = res[k];
            }
        }
    for (int ky = 1; ky < ny / 2; ky++) {
        const size_t pp = (size_t)ky * hnx, pm = (size_t)(ny - ky) * hnx;
        for (int f = 0; f < 2; f++) {
            double* a = f ? eta_hat : v_hat;
            for (int k = 0; k < nz; k++) {
                a[2 * ((size_t)stride * pm + k)] = a[2 * ((size_t)stride * pp + k)];
                a[2 * ((size_t)stride * pm + k) + 1] = -a[2 * ((size_t)stride * pp + k) + 1];
            }
        }
    }
    return RPCF_OK;
}

// ---- per-kernel timing (hipEvents on the context's stream) ----------------------------
int rpcf_timing(rpcf_ctx* c, int enable) {
    if (!c) return RPCF_EINVAL;
    int rc = sync_ctx(c);
    if (rc) return rc;
    c->timing = enable ? 1 : 0;
    for (int k = 0; k < KID_COUNT; k++) { c->t_ms[k] = 0.0; c->t_cnt[k] = 0; }
    return RPCF_OK;
}

int rpcf_timing_get(rpcf_ctx* c, int kid, double* total_ms, long* count) {
    if (!c || kid < 0 || kid >= KID_COUNT) return RPCF_EINVAL;
    int rc = sync_ctx(c);
    if (rc) return rc;
    if (total_ms) *total_ms = c->t_ms[kid];
    if (count) *count = c->t_cnt[kid];
    return RPCF_OK;
}

const char* rpcf_kernel_name(int kid) { return (kid >= 0 && kid < KID_COUNT) ? kKernelNames[kid] : nullptr; }

int rpcf_num_kernels(void) { return KID_COUNT; }

}  // extern "C"
