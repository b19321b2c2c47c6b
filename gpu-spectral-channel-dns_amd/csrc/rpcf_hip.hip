// rpcf_hip.hip -- MI355X (gfx950) implementation of the rpcf C ABI (include/rpcf.h,
// include/rpcf_group.h): context, device memory, launches, and the host side of the
// reference's step orchestration (nextStep, source/solver.cpp:35-101), with the whole step
// resident on the GPU (no per-step host<->device copies, no helper thread).
//
// A context holds `capacity` simulations ("slots") of one grid shape.  Every per-slot
// array is one allocation [capacity][...]; a launch covers a list of slots (its member
// coordinate j runs slot ids[j]), so one kernel launch advances all listed simulations.
// The single-simulation ABI of rpcf.h is slot 0 of a context of capacity 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rpcf.h"
#include "../../include/rpcf_group.h"
#include "../../include/rpcf_static.h"
#include "host_prep.hpp"
#include "k_linear.hpp"
#include "k_transform.hpp"
#include "k_cheb_mfma.hpp"
#include "k_cheb_dmfma.hpp"

// Static device scratch of the entry points that allocate nothing (include/rpcf_static.h):
// each GPU holds its own copy, RPCF_STATIC_SCRATCH_BYTES, until the process ends.
static __device__ __align__(256) unsigned char scratch[RPCF_STATIC_SCRATCH_BYTES];

namespace {

enum KernelId {
    KID_VELOCITY = 0, KID_S2P, KID_SPAN_INV, KID_LAMB, KID_SPAN_FWD, KID_P2S_NL, KID_LIN_TV, KID_LINEAR, KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {"k_velocity", "k_cheb_s2p", "k_span_inv", "k_lamb_phys",
                                             "k_span_fwd", "k_p2s_nl",   "k_lin_tv",   "k_lin_solve"};

constexpr int kFftThreads = 256;
constexpr int kTvPencils = 16;  // pencils per block of the fallback k_lin_tv

struct Pending {
    int kid;
    hipEvent_t a, b;
};

// the six spectral fields F of a slot relative to its state: STALE = velocity(state), not
// computed yet; FRESH = computed from the current state; EXPLICIT = set by the caller, or
// held over a linear step, so not derivable from the state
enum { FIELDS_STALE = 0, FIELDS_FRESH = 1, FIELDS_EXPLICIT = 2 };

struct SlotHost {
    int configured = 0;
    int have_state = 0;
    int fields = FIELDS_STALE;
    int phase = 0;  // mirror of the device history phase
    rpcf_params prm{};
    MemberParams mp{};
};

}  // namespace

struct rpcf_ctx {
    rpcf_params shape{};
    Dims d{};
    int capacity = 1;
    double fac = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;
    // the operator products of a step run on a second stream, beside the nonlinear transforms
    // (they depend on the state only): fork after the state is final, join before the solve
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int lds_limit = 65536;
    int ncu = 256;                   // compute units of the device
    long matrix_cache_hits = 0;      // slots whose inverses were copied from another slot
    // tiles: sequences per block and LDS row stride (elements) of each transform kernel
    int TB_s2p = 1, TB_nl = 1, TB_nl2 = 0, TB_vel = 1, TK_inv = 1, TK_fwd = 1, RP = 1;
    int ld_zn = 0, ld_my = 0, ld_mx = 0, zld = 0;
    int mixed = 0;                   // 1: transform intermediates and x/span FFTs in FP32
    FftPlan pl_x{}, pl_y{}, pl_zp{}, pl_zn{};
    FftPlan pl_xf{}, pl_yf{};        // FP32 twiddles of the x and span plans (mixed form)
    float* twf = nullptr;
    int cap_x = 0, cap_y = 0, cap_zp = 0, cap_zn = 0;  // fft_capacity of each plan
    int tv_rt = 0, tv_s = 0;                           // MFMA row tiles and k-steps of the operator products
    // mixed form: Chebyshev transforms as split-FP16 products (k_cheb_mfma.hpp)
    int use_mfma = 0;
    _Float16* MT = nullptr;  // s2p fragments, then p2s fragments
    long mt_p2s = 0;         // offset (halves) of the p2s fragments
    int mt_s0 = 0, mt_s1 = 0;
    // FP64 form: the padded p2s with the nonlinear combination as FP64 products (k_cheb_dmfma.hpp)
    int use_dmfma = 0;
    double* DT = nullptr;  // fragments of the folded p2s tables
    // per-slot arrays ([capacity][...]) and their slot strides in elements
    double2* F = nullptr;     // [6][P][nz] u, v, w, ox, oy, oz
    double2* SV = nullptr;    // [P][nz] v-hat (u0 in pencil 0)
    double2* SE = nullptr;    // [P][nz] eta-hat (w0 in pencil 0)
    double2* NL = nullptr;    // [4][P][nz] N_v (2 buffers), N_eta (2 buffers); current = phase
    double2* Y = nullptr;     // [6][hnx][pz][ny] (also Z1 [3][hnx][pz][ny])
    double2* W = nullptr;     // [6][pz][my][nxh]
    double2* Z2 = nullptr;    // [3][pz][my][zld], zld = hnx rounded up to 8 (128-byte rows; FP32: 16)
    double2* WL = nullptr;    // [5][P][nz] operator products of the linear right-hand sides
    double2* MINV = nullptr;  // [2][nslot][nz*nz] inverses, column-major
    double2* MINVH = nullptr; // [2][nslot][np][nz] parity blocks of the inverses (even base profiles;
                              // the mixed form keeps the v blocks only, [1][nslot][np][nz])
    unsigned* MINVE = nullptr; // [nslot][np][nz] BF16 (re, im): the v blocks' remainders (FP64 form)
                               // or the eta blocks' low parts (mixed form)
    float2* MINVF = nullptr;   // [nslot][np][nz] the eta blocks' FP32 high parts (mixed form)
    int4* PAIRS = nullptr;    // [2][np] row pairs (i, i', sign, 0) of the v and eta systems
    int npair = 0;            // np = N/2 + 1
    double* TF = nullptr;     // [5][RT][S][64] MFMA A-operand tiles of T0, T2, T4q, T0, T2q
    double* T4Q = nullptr;    // [nz*nz] T4*dt*0.5/Re, transposed (fallback operator products)
    double* T2Q = nullptr;    // [nz*nz] T2*0.5/Re*dt, transposed
    double *U = nullptr, *DU = nullptr, *DDU = nullptr;  // [nz] base profile at the collocation points
    double* MEANU = nullptr;                             // [pz] base profile at the padded points
    MemberParams* MP = nullptr;                          // [capacity]
    int* PH = nullptr;                                   // [capacity] history phase
    long sF = 0, sS = 0, sNL = 0, sY = 0, sW = 0, sZ = 0, sWl = 0, sM = 0, sMH = 0, sME = 0, sTf = 0, sQ = 0, sU = 0, sPz = 0;
    // shape tables shared by all slots
    double *T0 = nullptr, *T2 = nullptr, *T4 = nullptr, *T0t = nullptr, *T2t = nullptr, *qw = nullptr, *tw = nullptr;
    std::vector<double> T0h, T2h, T4h, qw_host;
    // slot-id lists on the device: ids_zero = {0}; ids = the list of the last grouped launch
    int* ids_dev = nullptr;
    int* ids_zero = nullptr;
    std::vector<int> ids_cur;
    std::vector<SlotHost> slots;
    // graph of one grouped step of g_ids
    int use_graphs = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<int> g_ids;
    // per-kernel timing with hipEvents on the context stream
    int timing = 0;
    unsigned timing_mask = ~0u;      // kernels (bit kid) bracketed by events while timing
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
    if (prm->precision < 0 || prm->precision > 2) return RPCF_EINVAL;
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

// per-slot element counts of every per-slot array (= the slot strides)
struct SlotSizes {
    long F, S, NL, Y, W, Z, Wl, M, MH, ME, MF, Tf, Q, U, Pz;
    size_t bytes() const {
        return sizeof(double2) * (size_t)(F + 2 * S + NL + Y + W + Z + Wl + M + MH) + sizeof(unsigned) * (size_t)ME +
               sizeof(float2) * (size_t)MF +
               sizeof(double) * (size_t)(Tf + 2 * Q + 3 * U + Pz) + sizeof(MemberParams) + 2 * sizeof(int);
    }
};

// Inverses: the full ones (the mean flow's slot, odd base profiles, and the source of the
// rest), their parity blocks (even base profiles), and in the FP64 form the BF16 remainders
// of the v blocks, which restore the full inverses' results (the blocks alone agree with them
// to ~1e-11 after a step, enough for the mixed form only; the eta blocks need none); in the
// mixed form the eta blocks also as FP32 high and BF16 low parts, which its solves read.
SlotSizes slot_sizes(const Dims& d, int precision) {
    SlotSizes s;
    const long P = d.P, nz = d.nz, zld = (d.hnx + 7) & ~7;
    const long RT = (d.N / 2 + 16) / 16, S = (d.N / 2 + 4) / 4;  // rows and columns per parity: N/2 + 1
    s.F = 6 * P * nz;
    s.S = P * nz;
    s.NL = 4 * P * nz;
    s.Y = 6L * d.hnx * d.pz * d.ny;
    s.W = 6L * d.pz * d.my * d.nxh;
    s.Z = 3L * d.pz * d.my * zld;
    s.Wl = 5 * P * nz;
    s.M = 2L * d.nslot * nz * nz;
    s.MH = (precision == 1 ? 1L : 2L) * d.nslot * nz * (d.N / 2 + 1);  // nz rows x (N/2 + 1) pair columns; mixed: v
    s.ME = precision <= 1 ? (long)d.nslot * nz * (d.N / 2 + 1) : 0;  // v remainders / eta low parts
    s.MF = precision == 1 ? s.ME : 0;                                  // eta high parts
    s.Tf = 5 * 2 * RT * S * 64;
    s.Q = nz * nz;
    s.U = nz;
    s.Pz = d.pz;
    return s;
}

// a transform length with a compile-time FFT plan
bool compiled_len(int n) {
    bool ok = false;
#define RPCF_CASE(L) ok |= n == L;
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    return ok;
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

// Wait for the stream, then fold the recorded kernel times into the totals.  The recorded
// events go back to the pool whatever the outcome, so that a failed wait leaves no stale
// times for a later call to add.
int sync_ctx(rpcf_ctx* c) {
    const hipError_t e = hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (e == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->t_ms[p.kid] += ms;
            c->t_cnt[p.kid] += 1;
        }
        c->pool.push_back(p.a);
        c->pool.push_back(p.b);
    }
    c->pending.clear();
    if (e != hipSuccess) return fail(c, RPCF_EDEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e));
    return RPCF_OK;
}

// Launch with optional HIP-event timing (never while a graph is being captured).
template <typename F>
int launch(rpcf_ctx* c, int kid, F&& fn, bool capturing = false) {
    hipEvent_t a = nullptr, b = nullptr;
    const bool timed = c->timing && ((c->timing_mask >> kid) & 1u) && !capturing;
    if (timed) {
        a = get_event(c);
        (void)hipEventRecord(a, c->stream);
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        if (a) c->pool.push_back(a);
        return fail(c, RPCF_EDEVICE, "launch %s: %s", kKernelNames[kid], hipGetErrorString(e));
    }
    if (timed) {
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

#define RCCHK(expr)                \
    do {                           \
        const int rc_ = (expr);    \
        if (rc_) return rc_;       \
    } while (0)

// Copies and fills on the context's stream, waited for: nothing touches the legacy null
// stream, so a context can be used while another thread captures a graph on its own stream.
int copy_sync(rpcf_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, kind, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPCF_OK;
}

int zero_sync(rpcf_ctx* c, void* dst, size_t bytes) {
    HIPCHK(c, hipMemsetAsync(dst, 0, bytes, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPCF_OK;
}

// a plan's twiddle table: exp(-2 pi i m/n), m < n, then (two-pass lengths) the stage-1
// table, the same values laid out [r - 1][k] (ct_tw1_count, fft_lds.hpp); 2 ct_tw_len(n) doubles
std::vector<double> plan_twiddles(int n) {
    std::vector<double> tw(2 * (size_t)ct_tw_len(n));
    rpcf_host::twiddles(n, tw.data());
    for (int i = 0; i < ct_tw1_count(n); i++) {
        const int m = ct_tw1_index(n, i);
        tw[2 * (n + i)] = tw[2 * m];
        tw[2 * (n + i) + 1] = tw[2 * m + 1];
    }
    return tw;
}

int make_plan(rpcf_ctx* c, int n, double* tw_dev, FftPlan* pl, int* cap) {
    pl->n = n;
    if (!rpcf_host::factorize(n, pl->rad, &pl->nst)) return fail(c, RPCF_EINVAL, "FFT length %d not 2^a3^b5^c", n);
    *cap = fft_capacity(pl->rad, pl->nst, kFftThreads);
    std::vector<double> tw = plan_twiddles(n);
    RCCHK(copy_sync(c, tw_dev, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
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
    RCCHK(copy_sync(c, h.data(), src, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
    for (size_t p = 0; p < (size_t)c->d.P; p++)
        for (int k = 0; k < nz; k++) {
            double* s = dst + 2 * ((size_t)stride * p + k);
            s[0] = h[p * nz + k].x;
            s[1] = h[p * nz + k].y;
        }
    return RPCF_OK;
}

// The six fields of the mixed form's matrix-core path are stored in FP32 (complex float,
// same element indexing): these move one field across the boundary with the conversion.
int upload_spec_f32(rpcf_ctx* c, float2* dst, const double* src, long stride) {
    if (!src || stride < c->d.nz) return fail(c, RPCF_EINVAL, "bad spectral array or stride");
    const int nz = c->d.nz;
    std::vector<float2> h(spec_elems(c));
    for (size_t p = 0; p < (size_t)c->d.P; p++)
        for (int k = 0; k < nz; k++) {
            const double* s = src + 2 * ((size_t)stride * p + k);
            h[p * nz + k] = make_float2((float)s[0], (float)s[1]);
        }
    HIPCHK(c, hipMemcpyAsync(dst, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPCF_OK;
}

int download_spec_f32(rpcf_ctx* c, double* dst, const float2* src, long stride) {
    if (!dst || stride < c->d.nz) return fail(c, RPCF_EINVAL, "bad spectral array or stride");
    const int nz = c->d.nz;
    std::vector<float2> h(spec_elems(c));
    int rc = sync_ctx(c);
    if (rc) return rc;
    RCCHK(copy_sync(c, h.data(), src, sizeof(float2) * h.size(), hipMemcpyDeviceToHost));
    for (size_t p = 0; p < (size_t)c->d.P; p++)
        for (int k = 0; k < nz; k++) {
            double* s = dst + 2 * ((size_t)stride * p + k);
            s[0] = h[p * nz + k].x;
            s[1] = h[p * nz + k].y;
        }
    return RPCF_OK;
}

// history buffer of a slot: which = 0 N_v, 1 N_eta; prev selects N^{n-1}
double2* hist(rpcf_ctx* c, int slot, int which, int prev) {
    const int ph = c->slots[slot].phase;
    const int b = (prev ? 1 - ph : ph) + 2 * which;
    return c->NL + slot * c->sNL + (long)b * c->sS;
}

int check_slot(rpcf_ctx* c, int slot) {
    if (!c) return RPCF_EINVAL;
    if (slot < 0 || slot >= c->capacity) return fail(c, RPCF_EINVAL, "slot %d outside 0..%d", slot, c->capacity - 1);
    if (!c->slots[slot].configured) return fail(c, RPCF_ESTATE, "slot %d has no member", slot);
    return RPCF_OK;
}

// Make `ids` the device slot list of the following grouped launches.
int set_ids(rpcf_ctx* c, const std::vector<int>& ids) {
    if (ids == c->ids_cur) return RPCF_OK;
    HIPCHK(c, hipMemcpyAsync(c->ids_dev, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ids_cur = ids;
    return RPCF_OK;
}

Members members(rpcf_ctx* c, const int* ids) {
    Members m;
    m.ids = ids;
    m.mp = c->MP;
    m.phase = c->PH;
    return m;
}

// ---- step pieces over a device slot list (asynchronous on c->stream) ---------------------
int enq_velocity(rpcf_ctx* c, const int* ids, int n, bool cap = false) {
    const Dims d = c->d;
    const long fS = (long)d.P * d.nz;
    return launch(
        c, KID_VELOCITY,
        [&] {
            const int TB = c->TB_vel, ld = c->ld_zn;
            dim3 g((d.ny + TB - 1) / TB, d.hnx, n);
            const size_t lds = sizeof(double2) * ((size_t)TB * ld + 256);
            switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                          \
    case L:                                                                                                   \
        if (c->use_mfma)                                                                                      \
            k_velocity_tile<L, float2><<<g, kFftThreads, lds, c->stream>>>(c->SV, c->SE, c->sS, (float2*)c->F, fS, \
                                                                           c->sF, d, members(c, ids), TB, ld);   \
        else                                                                                                  \
            k_velocity_tile<L><<<g, kFftThreads, lds, c->stream>>>(c->SV, c->SE, c->sS, c->F, fS, c->sF, d,        \
                                                                   members(c, ids), TB, ld);                    \
        break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: {
                    const int thr = 64 * ((d.nz + 63) / 64);
                    k_velocity<<<dim3(d.P, n), thr, 0, c->stream>>>(c->SV, c->SE, c->sS, c->F, fS, c->sF, d,
                                                                    members(c, ids));
                }
            }
        },
        cap);
}

// Backward transforms, Lamb vector, forward transforms and rhsNonlinear of the listed
// slots with the intermediates Y, W, Z2 and the x/span FFTs in precision R (FP64, or FP32
// in the mixed form); the Chebyshev transforms and everything that reaches the state are
// FP64.
template <typename R>
int enq_transforms(rpcf_ctx* c, const int* ids, int n, bool cap) {
    const Dims d = c->d;
    int rc;
    // rows of Z2 padded to whole 128-byte lines: 8 FP64 or 16 FP32 entries (FP32 rows padded
    // to 8 entries straddle lines, and the span transform fetched 1.75x its bytes); an FP32
    // member keeps its whole FP64-sized slot (twice the stride in FP32 entries), so that the
    // wider padding stays inside it
    const int zld = sizeof(R) == 4 ? (d.hnx + 15) & ~15 : c->zld;
    const long sZ = sizeof(R) == 4 ? 2 * c->sZ : c->sZ;
    const long fS = (long)d.P * d.nz, yS = (long)d.hnx * d.pz * d.ny, wS = (long)d.pz * d.my * d.nxh,
               zS = (long)d.pz * d.my * zld;
    const double2* tw4 = (const double2*)c->pl_zp.tw;
    const Members mb = members(c, ids);
    using C2 = cplx<R>;
    C2* Y = reinterpret_cast<C2*>(c->Y);
    C2* W = reinterpret_cast<C2*>(c->W);
    C2* Z2 = reinterpret_cast<C2*>(c->Z2);
    constexpr bool f64 = sizeof(R) == 8;
    const FftPlan& plx = f64 ? c->pl_x : c->pl_xf;
    const FftPlan& ply = f64 ? c->pl_y : c->pl_yf;
    {
        const int TB = c->TB_s2p, ld = c->ld_zn;
        dim3 g((d.ny + TB - 1) / TB, d.hnx, 6 * n);
        size_t lds = sizeof(double2) * TB * ld;
        rc = launch(
            c, KID_S2P,
            [&] {
                if constexpr (!f64) {
                    if (c->use_mfma) {
                        // persistent blocks: one per compute unit, pairs of items round-robin
                        const int items = n * 6 * d.hnx * ((d.ny + cheb_mfma::kPencils - 1) / cheb_mfma::kPencils);
                        const int grid = std::min((items + 1) / 2, c->ncu);
                        switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                               \
    case L:                                                                                                        \
        if constexpr (L <= 384)                                                                                    \
            k_s2p_mfma<(L <= 384 ? L : 384)><<<grid, S2pMfma<(L <= 384 ? L : 384)>::threads,                     \
                                               S2pMfma<(L <= 384 ? L : 384)>::lds, c->stream>>>(                  \
                (const float2*)c->F, fS, c->sF, Y, yS, c->sY, d, (const h8_t*)c->MT, ids, n);                     \
        break;
                            RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                        }
                        return;
                    }
                }
                switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                             \
    case L:                                                                                                      \
        if constexpr (L >= 64)                                                                                   \
            k_cheb_s2p_h<(L >= 64 ? L : 64), R><<<g, kFftThreads, sizeof(double2) * s2p_h_lds(L, TB), c->stream>>>( \
                c->F, fS, c->sF, Y, yS, c->sY, d, tw4, TB, ids);                                                 \
        else                                                                                                     \
            k_cheb_s2p<1, L, R><<<g, kFftThreads, lds, c->stream>>>(c->F, fS, c->sF, Y, yS, c->sY, d, c->pl_zn, tw4, \
                                                                   TB, ld, ids);                                 \
        break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    default:
                        if constexpr (f64)
                            k_cheb_s2p<1><<<g, kFftThreads, lds, c->stream>>>(c->F, fS, c->sF, Y, yS, c->sY, d,
                                                                              c->pl_zn, tw4, TB, ld, ids);
                }
            },
            cap);
        if (rc) return rc;
    }
    {
        const int TK = c->TK_inv, ld = c->ld_my;
        dim3 g((d.nxh + TK - 1) / TK, d.pz, 6 * n);
        size_t lds = sizeof(C2) * (TK * ld + (compiled_len(d.my) ? ct_tw_lds_len(d.my) : 0));
        rc = launch(
            c, KID_SPAN_INV,
            [&] {
                switch (d.my) {
#define RPCF_CASE(L) case L: k_span_inv<L, R><<<g, kFftThreads, lds, c->stream>>>(Y, yS, c->sY, W, wS, c->sW, d, d.pz, ply, TK, ld, ids); break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    default:
                        if constexpr (f64)
                            k_span_inv<<<g, kFftThreads, lds, c->stream>>>(Y, yS, c->sY, W, wS, c->sW, d, d.pz, ply,
                                                                           TK, ld, ids);
                }
            },
            cap);
        if (rc) return rc;
    }
    {
        const int RP = c->RP, ld = c->ld_mx;
        dim3 g((d.my / 2 + RP - 1) / RP, d.pz, n);
        size_t lds = sizeof(C2) * (6 * RP * ld + (compiled_len(d.mx) ? ct_tw_lds_len(d.mx) : 0));
        rc = launch(
            c, KID_LAMB,
            [&] {
                switch (d.mx) {
#define RPCF_CASE(L) case L: k_lamb_phys<L, R><<<g, kFftThreads, lds, c->stream>>>(W, wS, c->sW, Z2, zS, sZ, zld, c->MEANU, c->sPz, c->fac, d, plx, RP, ld, mb); break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    default:
                        if constexpr (f64)
                            k_lamb_phys<<<g, kFftThreads, lds, c->stream>>>(W, wS, c->sW, Z2, zS, sZ, zld,
                                                                            c->MEANU, c->sPz, c->fac, d, plx, RP, ld,
                                                                            mb);
                }
            },
            cap);
        if (rc) return rc;
    }
    {
        const int TK = c->TK_fwd, ld = c->ld_my;
        dim3 g((d.hnx + TK - 1) / TK, d.pz, 3 * n);
        size_t lds = sizeof(C2) * (TK * ld + (compiled_len(d.my) ? ct_tw_lds_len(d.my) : 0));
        rc = launch(
            c, KID_SPAN_FWD,
            [&] {
                switch (d.my) {
#define RPCF_CASE(L) case L: k_span_fwd<L, R><<<g, kFftThreads, lds, c->stream>>>(Z2, zS, sZ, zld, Y, yS, c->sY, d, d.pz, ply, TK, ld, ids, c->PH); break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    default:
                        if constexpr (f64)
                            k_span_fwd<<<g, kFftThreads, lds, c->stream>>>(Z2, zS, sZ, zld, Y, yS, c->sY, d,
                                                                           d.pz, ply, TK, ld, ids, c->PH);
                }
            },
            cap);
        if (rc) return rc;
    }
    // the AB2 history swap ran inside k_span_fwd: the new N lands in the old "previous" buffers
    if constexpr (!f64) {
        if (c->use_mfma) {
            dim3 g((d.ny + cheb_mfma::kPencils - 1) / cheb_mfma::kPencils, d.hnx, n);
            return launch(
                c, KID_P2S_NL,
                [&] {
                    switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                                 \
    case L:                                                                                                          \
        if constexpr (L <= 384)                                                                                      \
            k_p2s_mfma<(L <= 384 ? L : 384)><<<g, 128 * cheb_mfma::Shape<(L <= 384 ? L : 384)>::RTS,               \
                                             p2s_mfma_lds<(L <= 384 ? L : 384)>(), c->stream>>>(                    \
                Y, yS, c->sY, c->NL, c->sNL, d, (const h8_t*)(c->MT + c->mt_p2s), c->mt_s0, c->mt_s1, mb);          \
        break;
                        RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    }
                },
                cap);
        }
    }
    if constexpr (f64) {
        if (c->use_dmfma) {
            const int nyb = (d.ny + cheb_dmfma::kPencils - 1) / cheb_dmfma::kPencils;
            const int nitems = n * d.hnx * nyb;
            return launch(
                c, KID_P2S_NL,
                [&] {
                    switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                              \
    case L:                                                                                                       \
        if constexpr (L <= 384)                                                                                   \
            k_p2s_dmfma<(L <= 384 ? L : 384)><<<cheb_dmfma::DShape<(L <= 384 ? L : 384)>::blocks(nitems),        \
                                              cheb_dmfma::DShape<(L <= 384 ? L : 384)>::NT,                       \
                                              cheb_dmfma::DShape<(L <= 384 ? L : 384)>::lds_bytes(), c->stream>>>( \
                Y, yS, c->sY, c->NL, c->sNL, d, c->DT, mb, nitems, nyb);                                          \
        break;
                        RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    }
                },
                cap);
        }
    }
    if (c->TB_nl2) {
        const int TB = c->TB_nl2, ld = c->ld_zn;
        dim3 g((d.ny + TB - 1) / TB, d.hnx, n);
        const size_t lds = sizeof(double2) * (2 * (size_t)TB * ld + 256 + ct_tw_lds_len(2 * d.N) + d.nz);
        return launch(
            c, KID_P2S_NL,
            [&] {
                switch (2 * d.N) {
#define RPCF_CASE(L) case L: k_p2s_nl2<L, R><<<g, kFftThreads, lds, c->stream>>>(Y, yS, c->sY, c->NL, c->sNL, d, c->pl_zn, tw4, TB, ld, mb); break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                }
            },
            cap);
    }
    {
        const int TB = c->TB_nl, ld = c->ld_zn;
        dim3 g((d.ny + TB - 1) / TB, d.hnx, n);
        size_t lds = sizeof(double2) * ((size_t)TB * ld + 256);
        rc = launch(
            c, KID_P2S_NL,
            [&] {
                switch (2 * d.N) {
#define RPCF_CASE(L) case L: k_p2s_nl<L, R><<<g, kFftThreads, lds, c->stream>>>(Y, yS, c->sY, c->NL, c->sNL, d, c->pl_zn, tw4, TB, ld, mb); break;
                    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                    default:
                        if constexpr (f64)
                            k_p2s_nl<<<g, kFftThreads, lds, c->stream>>>(Y, yS, c->sY, c->NL, c->sNL, d, c->pl_zn, tw4,
                                                                         TB, ld, mb);
                }
            },
            cap);
    }
    return rc;
}

// getNonlinear for the listed slots; with_velocity: form their fields from the state first
int enq_nonlinear(rpcf_ctx* c, const int* ids, int n, bool with_velocity, bool cap = false) {
    int rc;
    if (with_velocity && (rc = enq_velocity(c, ids, n, cap))) return rc;
    return c->mixed ? enq_transforms<float>(c, ids, n, cap) : enq_transforms<double>(c, ids, n, cap);
}

// The operator products of the linear step on stream s (the context stream, timed when
// asked, or the side stream, untimed).
int enq_lin_tv(rpcf_ctx* c, const int* ids, int n, bool cap, hipStream_t s) {
    const Dims d = c->d;
    auto body = [&] {
            const size_t lds = sizeof(double) * 2 * kTvChunk * c->tv_s * 64;
            dim3 g((d.P + kTvBlockPencils - 1) / kTvBlockPencils, 5 * n);
            // persistent blocks with an operator's whole table in LDS when it fits (2N <= 256):
            // about one block per compute unit (two when the table is small)
            const size_t lds_stream = sizeof(double) * 2 * c->tv_rt * c->tv_s * 64;
            const bool tv_stream = c->tv_s <= 17 && lds_stream <= (size_t)c->lds_limit;
            const int per_cu = lds_stream * 2 <= (size_t)c->lds_limit ? 2 : 1;
            int gs = std::max(1, std::min((d.P + 31) / 32, per_cu * c->ncu / (5 * n)));
            // a multiple of 8: the five operator blocks of one pencil range (linear ids x + gs y)
            // land on the same XCD (id mod 8), so v-hat and eta-hat come from its L2 after the first read
            if (gs >= 16) gs &= ~7;
            switch (2 * d.N) {
#define RPCF_CASE(L)                                                                                             \
    case L:                                                                                                      \
        if (tv_stream)                                                                                           \
            k_lin_tv_stream<(L / 4 + 4) / 4, 8><<<dim3(gs, 5 * n), 512, lds_stream, s>>>(                         \
                c->SV, c->SE, c->sS, c->WL, c->sWl, d, c->TF, c->sTf, c->tv_rt, ids);                            \
        else                                                                                                     \
            k_lin_tv_mfma<(L / 4 + 4) / 4><<<g, 256, lds, s>>>(c->SV, c->SE, c->sS, c->WL, c->sWl, d,            \
                                                               c->TF, c->sTf, c->tv_rt, ids);                   \
        break;
                RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
                default: {
                    dim3 gv((d.P + kTvPencils - 1) / kTvPencils, n);
                    const size_t lv = sizeof(double2) * 2 * (size_t)d.nz * (kTvPencils + 1);
                    k_lin_tv<kTvPencils><<<gv, 384, lv, s>>>(c->SV, c->SE, c->sS, c->WL, c->sWl, d, c->T0t,
                                                             c->T2t, c->T4Q, c->T2Q, c->sQ, ids);
                }
            }
    };
    if (s == c->stream) return launch(c, KID_LIN_TV, body, cap);
    body();
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RPCF_OK : fail(c, RPCF_EDEVICE, "launch k_lin_tv: %s", hipGetErrorString(e));
}

int enq_lin_solve(rpcf_ctx* c, const int* ids, int n, bool cap = false) {
    const Dims d = c->d;
    return launch(
        c, KID_LINEAR,
        [&] {
            // about four column groups of nz threads, at most kSolveThreads; where four groups
            // would take more than 512 threads but three fit in 512 (128 < nz <= 170), 512:
            // three blocks per CU (24 waves) instead of two (18) -- 310 against 321 us at
            // nz = 129.  The group count sets the GEMV's summation order.
            int thr = std::min(kSolveThreads, 64 * ((4 * d.nz + 63) / 64));
            if (thr > 512 && 512 / d.nz >= 3) thr = 512;
            k_lin_solve<<<dim3(d.nslot, n), thr, 0, c->stream>>>(c->SV, c->SE, c->sS, c->NL, c->sNL, c->MINV, c->sM,
                                                                 c->MINVH, c->sMH, c->MINVE, c->sME, c->MINVF, c->PAIRS,
                                                                 c->npair,
                                                                 c->WL, c->sWl, d, members(c, ids), c->T0t, c->U, c->DU,
                                                                 c->DDU, c->sU, c->qw);
        },
        cap);
}

int enq_linear(rpcf_ctx* c, const int* ids, int n, bool cap = false) {
    int rc = enq_lin_tv(c, ids, n, cap, c->stream);
    return rc ? rc : enq_lin_solve(c, ids, n, cap);
}

// One step of the listed slots: the nonlinear evaluation (with the velocity of the state when
// with_velocity), the operator products and the solves.  The operator products read only the
// state, so unless their launches are being timed they run on the side stream, forked after
// the state is final and joined before the solve, beside the nonlinear transforms.
int enq_step(rpcf_ctx* c, const int* ids, int n, bool with_velocity, bool cap = false) {
    const bool side = !(c->timing && ((c->timing_mask >> KID_LIN_TV) & 1u));
    if (!side) {
        int rc = enq_nonlinear(c, ids, n, with_velocity, cap);
        return rc ? rc : enq_linear(c, ids, n, cap);
    }
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    int rc = enq_lin_tv(c, ids, n, cap, c->side);
    // the join is recorded even after a failed launch, so that a capture is not left open
    const hipError_t e1 = hipEventRecord(c->ev_join, c->side);
    if (!rc) rc = enq_nonlinear(c, ids, n, with_velocity, cap);
    const hipError_t e2 = hipStreamWaitEvent(c->stream, c->ev_join, 0);
    if (rc) return rc;
    if (e1 != hipSuccess || e2 != hipSuccess)
        return fail(c, RPCF_EDEVICE, "step fork/join: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    return enq_lin_solve(c, ids, n, cap);
}

// ---- tile autotuner ------------------------------------------------------------------
// The tile of each transform kernel only regroups independent pencils, rows or columns, so
// every candidate gives bit-identical results and the choice is free to follow the clock.
// At context creation the candidates (the largest tile that fits, and half of it) are timed
// with HIP events on slot 0 (zero state, placeholder physics) by coordinate descent over
// the kernels; a candidate replaces the current one when it is 3 % faster.  The choice is
// kept per (mx, my, mz, precision) for the life of the process.  RPCF_AUTOTUNE=0 keeps the
// rule-based tiles; RPCF_AUTOTUNE=half takes the halved candidates untimed (for tests).
struct TileSet {
    int TB_s2p, TK_inv, TK_fwd, RP, TB_nl2, TB_vel;
};
std::mutex g_tune_mu;
std::map<std::array<int, 4>, TileSet> g_tuned;

int* tile_ref(rpcf_ctx* c, int i) {
    int* refs[6] = {&c->TB_s2p, &c->TK_inv, &c->TK_fwd, &c->RP, &c->TB_nl2, &c->TB_vel};
    return refs[i];
}
constexpr int kTileKernel[6] = {KID_S2P, KID_SPAN_INV, KID_SPAN_FWD, KID_LAMB, KID_P2S_NL, KID_VELOCITY};

int autotune_tiles(rpcf_ctx* c) {
    const char* env = getenv("RPCF_AUTOTUNE");
    const std::string mode = env ? env : "1";
    if (mode == "0") return RPCF_OK;
    const Dims d = c->d;
    bool compiled = false;
#define RPCF_CASE(L) compiled |= 2 * d.N == L;
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    // k_velocity_tile and k_p2s_nl2 run for compiled Chebyshev lengths only
    const bool tunable[6] = {!c->use_mfma, true, true, true, c->TB_nl2 > 0 && !c->use_mfma && !c->use_dmfma, compiled};
    if (mode == "half") {
        for (int i = 0; i < 6; i++)
            if (tunable[i] && *tile_ref(c, i) >= 2) *tile_ref(c, i) /= 2;
        return RPCF_OK;
    }
    const std::array<int, 4> key = {d.mx, d.my, d.N, c->mixed};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            const TileSet& t = it->second;
            c->TB_s2p = t.TB_s2p, c->TK_inv = t.TK_inv, c->TK_fwd = t.TK_fwd, c->RP = t.RP;
            c->TB_nl2 = t.TB_nl2, c->TB_vel = t.TB_vel;
            return RPCF_OK;
        }
    }
    // slot 0 with zero fields and placeholder physics (init_slot overwrites all of it)
    MemberParams mp{};
    mp.aphi = mp.beta = 1.0;
    mp.Re = 1000.0;
    mp.dt = 1e-3;
    RCCHK(copy_sync(c, c->MP, &mp, sizeof(MemberParams), hipMemcpyHostToDevice));
    RCCHK(zero_sync(c, c->F, sizeof(double2) * c->sF));
    RCCHK(zero_sync(c, c->SV, sizeof(double2) * c->sS));
    RCCHK(zero_sync(c, c->SE, sizeof(double2) * c->sS));
    RCCHK(zero_sync(c, c->NL, sizeof(double2) * c->sNL));
    auto time_kernel = [&](int kid, double* ms) -> int {
        c->timing = 1;
        c->timing_mask = 1u << kid;
        for (int k = 0; k < KID_COUNT; k++) c->t_ms[k] = 0.0, c->t_cnt[k] = 0;
        int rc = RPCF_OK;
        for (int r = 0; r < 3 && !rc; r++) rc = enq_nonlinear(c, c->ids_zero, 1, true);
        if (!rc) rc = sync_ctx(c);
        c->timing = 0;
        c->timing_mask = ~0u;
        *ms = c->t_cnt[kid] ? c->t_ms[kid] / c->t_cnt[kid] : 0.0;
        return rc;
    };
    for (int i = 0; i < 6; i++) {
        int* t = tile_ref(c, i);
        if (!tunable[i] || *t < 2) continue;
        const int cur = *t;
        double t_cur = 0.0, t_alt = 0.0;
        RCCHK(time_kernel(kTileKernel[i], &t_cur));
        *t = cur / 2;
        RCCHK(time_kernel(kTileKernel[i], &t_alt));
        if (!(t_alt < 0.97 * t_cur)) *t = cur;
    }
    RCCHK(zero_sync(c, c->PH, sizeof(int)));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = TileSet{c->TB_s2p, c->TK_inv, c->TK_fwd, c->RP, c->TB_nl2, c->TB_vel};
    return RPCF_OK;
}

// Device scratch of the entry points that allocate nothing (rpcf_spec_to_phys_static).  In
// host code `scratch` is only a host-side stand-in for the symbol: its device address, one
// per GPU, is looked up once per device while that device is current (init_ctx) and kept
// in g_scratch_dev; kernels and copies take that address, never the name.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int scratch_setup(rpcf_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, RPCF_EDEVICE, "device %d outside 0..%d", c->device, kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[c->device]) return RPCF_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, RPCF_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[c->device] = p;
    return RPCF_OK;
}

void free_ctx(rpcf_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    void* ptrs[] = {c->F,   c->SV, c->SE, c->NL,  c->Y,     c->W,  c->Z2, c->WL, c->MINV, c->TF,  c->T4Q,
                    c->T2Q, c->U,  c->DU, c->DDU, c->MEANU, c->MP, c->PH, c->T0, c->T2,   c->T4,  c->T0t,
                    c->T2t, c->qw, c->tw, c->ids_dev, c->ids_zero, c->twf, c->MINVH, c->MINVE, c->MINVF, c->PAIRS, c->MT, c->DT};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
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

// Host fragments of the split-FP16 Chebyshev products of one length 2N, built once per process.
struct MfmaTables {
    std::vector<_Float16> h;  // s2p fragments, then p2s fragments
    long p2s_offset = 0;
    int s0 = 0, s1 = 0;
};
std::mutex g_mt_mu;
std::map<int, MfmaTables> g_mt;

const MfmaTables& mfma_tables(int N) {
    std::lock_guard<std::mutex> lk(g_mt_mu);
    auto it = g_mt.find(N);
    if (it != g_mt.end()) return it->second;
    MfmaTables t;
    rpcf_host::s2p_tables(N, t.h);
    t.p2s_offset = (long)t.h.size();
    rpcf_host::p2s_tables(N, t.h, &t.s0, &t.s1);
    return g_mt.emplace(N, std::move(t)).first->second;
}

// Host fragments of the FP64 p2s products of one length 2N, built once per process.
std::map<int, std::vector<double>> g_dt;

const std::vector<double>& dmfma_tables(int N) {
    std::lock_guard<std::mutex> lk(g_mt_mu);
    auto it = g_dt.find(N);
    if (it != g_dt.end()) return it->second;
    std::vector<double> t;
    rpcf_host::p2s_tables_f64(N, t);
    return g_dt.emplace(N, std::move(t)).first->second;
}

// allocation and shape set-up of a context of `capacity` slots; on failure the caller frees it
int init_ctx(rpcf_ctx* c) {
    const Dims d = c->d;
    const long C = c->capacity;
    int rc;
    c->fac = 1.0 / d.mx / d.my;  // normalize factor, source/transform.cu:267
    HIPCHK(c, hipGetDevice(&c->device));
    RCCHK(scratch_setup(c));
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    int lim = 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && lim > 0)
        c->lds_limit = lim;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0)
        c->ncu = ncu;

    const SlotSizes ss = slot_sizes(d, c->shape.precision);
    c->sF = ss.F; c->sS = ss.S; c->sNL = ss.NL; c->sY = ss.Y; c->sW = ss.W; c->sZ = ss.Z;
    c->sWl = ss.Wl; c->sM = ss.M; c->sMH = ss.MH; c->sME = ss.ME; c->sTf = ss.Tf; c->sQ = ss.Q; c->sU = ss.U; c->sPz = ss.Pz;
    c->zld = (d.hnx + 7) & ~7;
    c->tv_rt = (d.N / 2 + 16) / 16;
    c->tv_s = (d.N / 2 + 4) / 4;
    if ((rc = dalloc(c, &c->F, C * ss.F)) || (rc = dalloc(c, &c->SV, C * ss.S)) || (rc = dalloc(c, &c->SE, C * ss.S)) ||
        (rc = dalloc(c, &c->NL, C * ss.NL)) || (rc = dalloc(c, &c->Y, C * ss.Y)) || (rc = dalloc(c, &c->W, C * ss.W)) ||
        (rc = dalloc(c, &c->Z2, C * ss.Z)) || (rc = dalloc(c, &c->WL, C * ss.Wl)) ||
        (rc = dalloc(c, &c->MINV, C * ss.M)) || (rc = dalloc(c, &c->MINVH, C * ss.MH)) ||
        (ss.ME && (rc = dalloc(c, &c->MINVE, C * ss.ME))) || (ss.MF && (rc = dalloc(c, &c->MINVF, C * ss.MF))) ||
        (rc = dalloc(c, &c->TF, C * ss.Tf)) ||
        (rc = dalloc(c, &c->T4Q, C * ss.Q)) || (rc = dalloc(c, &c->T2Q, C * ss.Q)) ||
        (rc = dalloc(c, &c->U, C * ss.U)) || (rc = dalloc(c, &c->DU, C * ss.U)) || (rc = dalloc(c, &c->DDU, C * ss.U)) ||
        (rc = dalloc(c, &c->MEANU, C * ss.Pz)) || (rc = dalloc(c, &c->MP, (size_t)C)) ||
        (rc = dalloc(c, &c->PH, (size_t)C)) || (rc = dalloc(c, &c->ids_dev, (size_t)C)) ||
        (rc = dalloc(c, &c->ids_zero, 1)))
        return rc;
    const size_t nz = d.nz;
    if ((rc = dalloc(c, &c->T0, nz * nz)) || (rc = dalloc(c, &c->T2, nz * nz)) || (rc = dalloc(c, &c->T4, nz * nz)) ||
        (rc = dalloc(c, &c->T0t, nz * nz)) || (rc = dalloc(c, &c->T2t, nz * nz)) || (rc = dalloc(c, &c->qw, nz)))
        return rc;
    if ((rc = dalloc(c, &c->tw, 2 * ((size_t)ct_tw_len(d.mx) + ct_tw_len(d.my) + ct_tw_len(4 * d.N) +
                                       ct_tw_len(2 * d.N)))))
        return rc;
    RCCHK(zero_sync(c, c->ids_zero, sizeof(int)));
    RCCHK(zero_sync(c, c->PH, sizeof(int) * C));
    {
        // Row pairs of the collocation systems under the wall reflection y -> -y (the rows of
        // z_p and z_{N-p}; the two wall rows of each condition).  v (4th order): rows 0, 1
        // (value at the walls) and 2, 3 (slope at the walls, which changes sign), rows
        // 4..N at the points p = i - 2; eta (2nd order): rows 0, 1, rows 2..N at p = i - 1.
        // A row paired with itself (the centre point) has i == i'.
        const int N = d.N, np = N / 2 + 1;
        std::vector<int4> pr(2 * (size_t)np);
        int k = 0;
        pr[k++] = make_int4(0, 1, 1, 0);
        pr[k++] = make_int4(2, 3, -1, 0);
        for (int i = 4; i <= N + 4 - i; i++) pr[k++] = make_int4(i, N + 4 - i, 1, 0);
        if (k != np) return fail(c, RPCF_EINVAL, "v row pairs: %d, expected %d", k, np);
        pr[k++] = make_int4(0, 1, 1, 0);
        for (int i = 2; i <= N + 2 - i; i++) pr[k++] = make_int4(i, N + 2 - i, 1, 0);
        if (k != 2 * np) return fail(c, RPCF_EINVAL, "eta row pairs: %d, expected %d", k - np, np);
        c->npair = np;
        if ((rc = dalloc(c, &c->PAIRS, pr.size()))) return rc;
        RCCHK(copy_sync(c, c->PAIRS, pr.data(), sizeof(int4) * pr.size(), hipMemcpyHostToDevice));
    }
    c->slots.assign(C, SlotHost());

    // shape operators (initAuxMatrix, source/solver.cpp:265-269)
    c->T0h.resize(nz * nz);
    c->T2h.resize(nz * nz);
    c->T4h.resize(nz * nz);
    rpcf_host::t_matrices(d.N, c->T0h.data(), c->T2h.data(), c->T4h.data());
    c->qw_host.resize(nz);
    for (size_t k = 0; k < nz; k++) c->qw_host[k] = (k % 2 == 0) ? 2.0 / (1.0 - (double)k * (double)k) : 0.0;
    RCCHK(copy_sync(c, c->T0, c->T0h.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    RCCHK(copy_sync(c, c->T2, c->T2h.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    RCCHK(copy_sync(c, c->T4, c->T4h.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    {
        std::vector<double> t0(nz * nz), t2(nz * nz);
        for (size_t i = 0; i < nz; i++)
            for (size_t j = 0; j < nz; j++) {
                t0[j * nz + i] = c->T0h[i * nz + j];
                t2[j * nz + i] = c->T2h[i * nz + j];
            }
        RCCHK(copy_sync(c, c->T0t, t0.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        RCCHK(copy_sync(c, c->T2t, t2.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
    }
    RCCHK(copy_sync(c, c->qw, c->qw_host.data(), sizeof(double) * nz, hipMemcpyHostToDevice));

    // FFT plans (the reference's cuFFT plans, source/transform.cu:49-71)
    double* tw = c->tw;
    if ((rc = make_plan(c, d.mx, tw, &c->pl_x, &c->cap_x))) return rc;
    tw += 2 * ct_tw_len(d.mx);
    if ((rc = make_plan(c, d.my, tw, &c->pl_y, &c->cap_y))) return rc;
    tw += 2 * ct_tw_len(d.my);
    if (c->mixed) {
        // FP32 copies of the x and span twiddles
        const std::vector<double> t64 = plan_twiddles(d.mx), t64y = plan_twiddles(d.my);
        std::vector<float> t32(t64.size() + t64y.size());
        for (size_t i = 0; i < t64.size(); i++) t32[i] = (float)t64[i];
        for (size_t i = 0; i < t64y.size(); i++) t32[t64.size() + i] = (float)t64y[i];
        if ((rc = dalloc(c, &c->twf, t32.size()))) return rc;
        RCCHK(copy_sync(c, c->twf, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice));
        c->pl_xf = c->pl_x;
        c->pl_xf.tw = c->twf;
        c->pl_yf = c->pl_y;
        c->pl_yf.tw = c->twf + t64.size();
    }
    if ((rc = make_plan(c, 4 * d.N, tw, &c->pl_zp, &c->cap_zp))) return rc;
    tw += 2 * ct_tw_len(4 * d.N);
    if ((rc = make_plan(c, 2 * d.N, tw, &c->pl_zn, &c->cap_zn))) return rc;
    // mixed form: fragments of the split-FP16 Chebyshev products (k_cheb_mfma.hpp) for the
    // compiled lengths up to 2N = 384; RPCF_MIXED_MFMA=0 keeps the FFT-based kernels
    if (c->mixed && 2 * d.N <= 384) {
        const char* e = getenv("RPCF_MIXED_MFMA");
        if (!(e && e[0] == '0')) {
            const MfmaTables& t = mfma_tables(d.N);
            if ((rc = dalloc(c, &c->MT, t.h.size()))) return rc;
            RCCHK(copy_sync(c, c->MT, t.h.data(), sizeof(_Float16) * t.h.size(), hipMemcpyHostToDevice));
            c->mt_p2s = t.p2s_offset;
            c->mt_s0 = t.s0;
            c->mt_s1 = t.s1;
            c->use_mfma = 1;
        }
    }
    // FP64 form: the padded p2s + nonlinear combination as FP64 matrix-core products for the
    // compiled lengths up to 2N = 384; RPCF_P2S_MFMA=0 keeps the FFT-based kernels
    if (!c->mixed && 2 * d.N <= 384) {
        bool compiled = false;
#define RPCF_CASE(L) compiled |= 2 * d.N == L;
        RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
        const char* e = getenv("RPCF_P2S_MFMA");
        if (compiled && !(e && e[0] == '0')) {
            const std::vector<double>& t = dmfma_tables(d.N);
            if ((rc = dalloc(c, &c->DT, t.size()))) return rc;
            RCCHK(copy_sync(c, c->DT, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
            c->use_dmfma = 1;
        }
    }

    // LDS tiles, sized for three resident blocks per CU
    lim = c->lds_limit;
    const int lim3 = lim / 3;
    c->ld_zn = 2 * d.N + 2;  // >= 2N (FFT) and >= 2 nz (L_x, L_y of k_p2s_nl)
    c->ld_my = d.my + 1;
    c->ld_mx = d.mx + 1;
    {
        int tmax = kChebTile;  // the runtime-plan kernel sizes its registers for kChebTile
#define RPCF_CASE(L) if (2 * d.N == L) tmax = s2p_tile_max(L);
        RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
        // room for k_cheb_s2p and for k_cheb_s2p_h (the half-length form, 2N >= 64)
        c->TB_s2p = pick_tile(tmax, 2 * d.N, c->cap_zn, lim3, [&](int t) {
            return sizeof(double2) * std::max((size_t)t * c->ld_zn, s2p_h_lds(2 * d.N, t));
        });
    }
    c->TB_nl = pick_tile(kChebTile, 2 * d.N, c->cap_zn, lim3, [&](int t) {
        return sizeof(double2) * ((size_t)t * c->ld_zn + 256);
    });
    // the batched k_p2s_nl2 (compiled lengths), two resident blocks per CU; 0: not used
    c->TB_nl2 = 0;
#define RPCF_CASE(L)                                                                                     \
    if (2 * d.N == L)                                                                                    \
        c->TB_nl2 = pick_tile(cheb_nl_tile_max(L), 1, cheb_nl_tile_max(L), lim / 2, [&](int t) {             \
            return sizeof(double2) * (2 * (size_t)t * c->ld_zn + 256 + ct_tw_lds_len(2 * d.N) + d.nz);       \
        });
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    // k_velocity_tile (compiled lengths): the tile of the scan kernels
    c->TB_vel = c->TB_nl;
#define RPCF_CASE(L)                                                                              \
    if (2 * d.N == L)                                                                             \
        c->TB_vel = pick_tile(cheb_nl_tile_max(L), 1, cheb_nl_tile_max(L), lim3, [&](int t) {     \
            return sizeof(double2) * ((size_t)t * c->ld_zn + 256);                                \
        });
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    int tk_max = kSpanTile, rp_max = kLambPairs;  // the runtime-plan kernels size registers for these
#define RPCF_CASE(L)                          \
    if (d.my == L) tk_max = span_tile_max(L); \
    if (d.mx == L) rp_max = lamb_pairs_max(L);
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    c->TK_inv = pick_tile(tk_max, d.my, c->cap_y, lim3, [&](int t) {
        return sizeof(double2) * ((size_t)t * c->ld_my + (compiled_len(d.my) ? ct_tw_lds_len(d.my) : 0));
    });
    c->TK_fwd = c->TK_inv;
    c->RP = pick_tile(rp_max, 6 * d.mx, c->cap_x, lim3, [&](int t) {
        return sizeof(double2) * (6 * (size_t)t * c->ld_mx + (compiled_len(d.mx) ? ct_tw_lds_len(d.mx) : 0));
    });
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
    allow_lds(k_lin_tv_mfma<(L / 4 + 4) / 4>, lim);      \
    allow_lds(k_lin_tv_stream<(L / 4 + 4) / 4, 8>, lim); \
    allow_lds(k_cheb_s2p<1, L, float>, lim);              \
    allow_lds(k_span_inv<L, float>, lim);                 \
    allow_lds(k_span_fwd<L, float>, lim);                 \
    allow_lds(k_lamb_phys<L, float>, lim);                \
    allow_lds(k_p2s_nl<L, float>, lim);                   \
    allow_lds(k_p2s_nl2<L>, lim);                         \
    allow_lds(k_cheb_s2p_h<(L >= 64 ? L : 64)>, lim);      \
    allow_lds(k_cheb_s2p_h<(L >= 64 ? L : 64), float>, lim); \
    allow_lds(k_p2s_nl2<L, float>, lim);                  \
    allow_lds(k_velocity_tile<L>, lim);                   \
    allow_lds(k_velocity_tile<L, float2>, lim);           \
    allow_lds(k_p2s_mfma<(L <= 384 ? L : 384)>, lim);    \
    allow_lds(k_s2p_mfma<(L <= 384 ? L : 384)>, lim);    \
    allow_lds(k_p2s_dmfma<(L <= 384 ? L : 384)>, lim);
    RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
    allow_lds(k_cheb_p2s, lim);
    allow_lds(k_x_c2r, lim);
    allow_lds(k_x_r2c, lim);
    allow_lds(k_lin_tv<kTvPencils>, lim);
    (void)hipGetLastError();
    RCCHK(autotune_tiles(c));
    if (getenv("RPCF_VERBOSE"))
        fprintf(stderr, "rpcf: lds_limit=%d TB_s2p=%d TB_nl2=%d TB_vel=%d TK=%d/%d RP=%d nslot=%d capacity=%ld\n", lim,
                c->TB_s2p, c->TB_nl2, c->TB_vel, c->TK_inv, c->TK_fwd, c->RP, d.nslot, C);
    return RPCF_OK;
}

// Physics of one slot: base profile, per-slot operator tables and the inverses of its
// Crank-Nicolson collocation matrices (initAuxMatrix, source/solver.cpp:260-320); the slot's
// state, fields and history are cleared.
// parity blocks of a slot's inverses (k_parity_blocks), on the context stream
int build_parity_blocks(rpcf_ctx* c, int slot) {
    const Dims d = c->d;
    k_parity_blocks<<<dim3(d.nslot, 2), 256, 0, c->stream>>>(c->MINV + slot * c->sM, c->MINVH + slot * c->sMH,
                                                             c->MINVE ? c->MINVE + slot * c->sME : nullptr,
                                                             c->MINVF ? c->MINVF + slot * c->sME : nullptr,
                                                             c->PAIRS, c->npair, d);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RPCF_OK : fail(c, RPCF_EDEVICE, "k_parity_blocks: %s", hipGetErrorString(e));
}

int init_slot(rpcf_ctx* c, int slot, const rpcf_params* prm) {
    const Dims d = c->d;
    const size_t nz = d.nz;
    int rc;
    SlotHost& sh = c->slots[slot];
    sh = SlotHost();
    sh.prm = *prm;
    // problem(RPCF_Paras): include/data.h:321-322
    LinParams lp;
    lp.aphi = prm->n_pi_x / 2;
    lp.beta = prm->n_pi_y / 2;
    lp.Re = prm->Re;
    lp.dt = prm->dt;
    sh.mp.aphi = lp.aphi;
    sh.mp.beta = lp.beta;
    sh.mp.Re = lp.Re;
    sh.mp.dt = lp.dt;
    sh.mp.Ro = prm->Ro;
    sh.mp.q_re = sh.mp.q_im = 0.0;
    sh.mp.flow_control = prm->flow_control;
    // an even base profile U(y) (no linear part) keeps the collocation systems symmetric under
    // y -> -y, so the solves use the parity blocks of the inverses: with their remainders in
    // the FP64 form (the full inverses' values), alone in the others.  RPCF_FULL_INVERSE=1
    // keeps the full inverses.
    const char* fe = getenv("RPCF_FULL_INVERSE");
    const bool full = fe && atoi(fe) != 0;
    sh.mp.parity = full || prm->u_profile[1] != 0.0 ? 0 : c->shape.precision == 0 ? 2 : 1;

    std::vector<double> U(nz), dU(nz), ddU(nz), mU(d.pz);
    rpcf_host::base_profile(d.nz, prm->u_profile, U.data(), dU.data(), ddU.data());
    rpcf_host::base_profile_padded(d.pz, prm->u_profile, mU.data());
    RCCHK(copy_sync(c, c->U + slot * c->sU, U.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    RCCHK(copy_sync(c, c->DU + slot * c->sU, dU.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    RCCHK(copy_sync(c, c->DDU + slot * c->sU, ddU.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    RCCHK(copy_sync(c, c->MEANU + slot * c->sPz, mU.data(), sizeof(double) * d.pz, hipMemcpyHostToDevice));
    {
        // the quotients T2*0.5/Re*dt and T4*dt*0.5/Re of source/linear.cpp:100, 142 in the
        // reference's order (transposed), and the five operators of the right-hand sides as
        // MFMA A-operand tiles (T0, T2, T4q for v; T0, T2q for eta)
        const double Re = prm->Re, dt = prm->dt;
        std::vector<double> t0(nz * nz), t2(nz * nz), q4(nz * nz), q2(nz * nz);
        for (size_t i = 0; i < nz; i++)
            for (size_t j = 0; j < nz; j++) {
                t0[j * nz + i] = c->T0h[i * nz + j];
                t2[j * nz + i] = c->T2h[i * nz + j];
                q2[j * nz + i] = c->T2h[i * nz + j] * 0.5 / Re * dt;
                q4[j * nz + i] = c->T4h[i * nz + j] * dt * 0.5 / Re;
            }
        RCCHK(copy_sync(c, c->T4Q + slot * c->sQ, q4.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        RCCHK(copy_sync(c, c->T2Q + slot * c->sQ, q2.data(), sizeof(double) * nz * nz, hipMemcpyHostToDevice));
        const int RT = c->tv_rt, S = c->tv_s;
        const std::vector<double>* ops[5] = {&t0, &t2, &q4, &t0, &q2};  // op[j*nz + i] = T[i][j]
        std::vector<double> tf((size_t)c->sTf, 0.0);
        // rows i <= N/2 against the even (par 0, j = 2k) and odd (par 1, j = 2k + 1) columns;
        // the kernel mirrors rows N - i by the parity T[N-i][j] = (-1)^j T[i][j]
        const size_t rows = d.N / 2 + 1;
        for (int m = 0; m < 5; m++)
            for (int par = 0; par < 2; par++)
                for (int rt = 0; rt < RT; rt++)
                    for (int sk = 0; sk < S; sk++)
                        for (int l = 0; l < 64; l++) {
                            const size_t i = 16 * rt + (l & 15), j = 2 * (4 * sk + (l >> 4)) + par;
                            if (i < rows && j < nz)
                                tf[((((size_t)m * 2 + par) * RT + rt) * S + sk) * 64 + l] = (*ops[m])[j * nz + i];
                        }
        RCCHK(copy_sync(c, c->TF + slot * c->sTf, tf.data(), sizeof(double) * tf.size(), hipMemcpyHostToDevice));
    }
    RCCHK(copy_sync(c, c->MP + slot, &sh.mp, sizeof(MemberParams), hipMemcpyHostToDevice));
    RCCHK(zero_sync(c, c->PH + slot, sizeof(int)));
    RCCHK(zero_sync(c, c->NL + slot * c->sNL, sizeof(double2) * c->sNL));
    RCCHK(zero_sync(c, c->F + slot * c->sF, sizeof(double2) * c->sF));
    RCCHK(zero_sync(c, c->SV + slot * c->sS, sizeof(double2) * c->sS));
    RCCHK(zero_sync(c, c->SE + slot * c->sS, sizeof(double2) * c->sS));

    // Matrix cache: the inverses depend on (Lx, Lz, Re, dt, U) only -- not on Ro, the flow
    // control or the state -- so a slot of this context already configured with the same
    // operator parameters lends its inverses (a device copy instead of the LU build).
    const size_t per = nz * nz;
    double2* minv = c->MINV + slot * c->sM;
    for (int s2 = 0; s2 < c->capacity; s2++) {
        const SlotHost& o = c->slots[s2];
        if (s2 == slot || !o.configured) continue;
        const rpcf_params& a = o.prm;
        if (a.n_pi_x == prm->n_pi_x && a.n_pi_y == prm->n_pi_y && a.Re == prm->Re && a.dt == prm->dt &&
            a.u_profile[0] == prm->u_profile[0] && a.u_profile[1] == prm->u_profile[1] &&
            a.u_profile[2] == prm->u_profile[2]) {
            HIPCHK(c, hipMemcpyAsync(minv, c->MINV + s2 * c->sM, sizeof(double2) * 2 * (size_t)d.nslot * per,
                                     hipMemcpyDeviceToDevice, c->stream));
            if (sh.mp.parity) RCCHK(build_parity_blocks(c, slot));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            c->matrix_cache_hits++;
            sh.configured = 1;
            return RPCF_OK;
        }
    }
    // collocation matrices and inverses, in chunks of matrix slots (initAuxMatrix)
    const int chunk =
        (int)std::min<size_t>((size_t)d.nslot, std::max<size_t>(64, ((size_t)1 << 29) / (2 * per * sizeof(double2))));
    double2* work = nullptr;
    int* status = nullptr;
    if ((rc = dalloc(c, &work, 2 * (size_t)chunk * per))) return rc;
    if ((rc = dalloc(c, &status, 1))) { (void)hipFree(work); return rc; }
    if ((rc = zero_sync(c, status, sizeof(int)))) { (void)hipFree(work); (void)hipFree(status); return rc; }
    for (int s0 = 0; s0 < d.nslot && !rc; s0 += chunk) {
        dim3 g(std::min(chunk, d.nslot - s0), 2);
        k_build_inverse<<<g, 256, 0, c->stream>>>(minv, work, s0, d, lp, c->T0, c->T2, c->T4, c->U + slot * c->sU,
                                                  c->DDU + slot * c->sU, status);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, RPCF_EDEVICE, "k_build_inverse: %s", hipGetErrorString(e));
    }
    int hstatus = 0;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "matrix build failed");
    if (!rc) rc = copy_sync(c, &hstatus, status, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(work);
    (void)hipFree(status);
    if (!rc && hstatus) rc = fail(c, RPCF_EINVAL, "%d singular collocation matrices", hstatus);
    if (!rc && sh.mp.parity) rc = build_parity_blocks(c, slot);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "parity blocks failed");
    if (!rc) sh.configured = 1;
    return rc;
}

int same_shape(const rpcf_params* a, const rpcf_params* b) { return a->mx == b->mx && a->my == b->my && a->mz == b->mz; }

int set_state_slot(rpcf_ctx* c, int slot, const double* v_hat, const double* eta_hat, long stride) {
    int rc = upload_spec(c, c->SV + slot * c->sS, v_hat, stride);
    if (!rc) rc = upload_spec(c, c->SE + slot * c->sS, eta_hat, stride);
    if (rc) return rc;
    SlotHost& sh = c->slots[slot];
    sh.have_state = 1;
    // initFlow tail (source/RPCFKernels.cu:413-427); the flow-rate target is the initial bulk of u0
    double qre = 0.0, qim = 0.0;
    for (int k = 0; k < c->d.nz; k++) {
        const double* s = v_hat + 2 * k;
        qre = qre + s[0] * c->qw_host[k];
        qim = qim + s[1] * c->qw_host[k];
    }
    sh.mp.q_re = qre;
    sh.mp.q_im = qim;
    RCCHK(copy_sync(c, c->MP + slot, &sh.mp, sizeof(MemberParams), hipMemcpyHostToDevice));
    sh.fields = FIELDS_STALE;
    return RPCF_OK;
}

// nsteps grouped steps of the listed slots: velocity of the state, nonlinear, linear;
// replayed from a captured graph when graphs are enabled (and timing is off)
int group_steps(rpcf_ctx* c, const std::vector<int>& ids, int nsteps) {
    int rc = set_ids(c, ids);
    if (rc) return rc;
    const int n = (int)ids.size();
    const bool graph = c->use_graphs && !c->timing;
    if (graph && (!c->gexec || c->g_ids != ids)) {
        if (c->gexec) {
            (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = enq_step(c, c->ids_dev, n, true, true);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, RPCF_EDEVICE, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->gexec = nullptr;
            return fail(c, RPCF_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
        }
        c->g_ids = ids;
    }
    for (int s = 0; s < nsteps; s++) {
        if (graph) {
            HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        } else {
            rc = enq_step(c, c->ids_dev, n, true);
            if (rc) return rc;
        }
        for (int id : ids) c->slots[id].phase ^= 1;
    }
    for (int id : ids) c->slots[id].fields = FIELDS_STALE;
    return sync_ctx(c);
}

int check_list(rpcf_ctx* c, const int* slots, int n, std::vector<int>* out) {
    if (!c) return RPCF_EINVAL;
    if (!slots || n < 1 || n > c->capacity) return fail(c, RPCF_EINVAL, "bad slot list (n = %d)", n);
    std::vector<int> ids(slots, slots + n);
    std::vector<int> sorted(ids);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(c, RPCF_EINVAL, "slot listed twice");
    for (int s : ids) {
        int rc = check_slot(c, s);
        if (rc) return rc;
        if (!c->slots[s].have_state) return fail(c, RPCF_ESTATE, "slot %d: no state set", s);
    }
    *out = ids;
    return RPCF_OK;
}

int create_ctx(const rpcf_params* shape, int capacity, rpcf_ctx** out) {
    if (!out) return RPCF_EINVAL;
    *out = nullptr;
    Dims d;
    if (check_sizes(shape, &d) || capacity < 1) return RPCF_EINVAL;
    if (shape->precision == 1) {
        // the mixed form runs compiled FFT lengths only
        bool ok_x = false, ok_y = false, ok_z = false;
#define RPCF_CASE(L) ok_x |= d.mx == L; ok_y |= d.my == L; ok_z |= 2 * d.N == L;
        RPCF_FFT_LENGTHS(RPCF_CASE)
#undef RPCF_CASE
        if (!(ok_x && ok_y && ok_z)) return RPCF_EINVAL;
    }
    rpcf_ctx* c = new rpcf_ctx();
    c->shape = *shape;
    c->d = d;
    c->capacity = capacity;
    c->mixed = shape->precision == 1;
    int rc = init_ctx(c);
    if (rc) {
        fprintf(stderr, "rpcf: %s\n", c->err);
        free_ctx(c);
        return rc;
    }
    *out = c;
    return RPCF_OK;
}

}  // namespace

// ======================================================================================
// C ABI: one simulation (slot 0)
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

int rpcf_create(const rpcf_params* prm, rpcf_ctx** out) {
    int rc = create_ctx(prm, 1, out);
    if (rc) return rc;
    rc = init_slot(*out, 0, prm);
    if (rc) {
        fprintf(stderr, "rpcf_create: %s\n", (*out)->err);
        free_ctx(*out);
        *out = nullptr;
    }
    return rc;
}

int rpcf_set_state(rpcf_ctx* c, const double* v_hat, const double* eta_hat, long stride) {
    int rc = check_slot(c, 0);
    return rc ? rc : set_state_slot(c, 0, v_hat, eta_hat, stride);
}

int rpcf_get_state(rpcf_ctx* c, double* v_hat, double* eta_hat, long stride) {
    int rc = check_slot(c, 0);
    if (!rc) rc = download_spec(c, v_hat, c->SV, stride);
    return rc ? rc : download_spec(c, eta_hat, c->SE, stride);
}

int rpcf_set_field(rpcf_ctx* c, int field, const double* f_hat, long stride) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    if (field < 0 || field > 5) return fail(c, RPCF_EINVAL, "bad field id %d", field);
    SlotHost& sh = c->slots[0];
    if (sh.fields == FIELDS_STALE && sh.have_state) {
        // the other five fields keep their values: materialise them first
        if ((rc = enq_velocity(c, c->ids_zero, 1))) return rc;
    }
    rc = c->use_mfma ? upload_spec_f32(c, (float2*)c->F + (size_t)field * spec_elems(c), f_hat, stride)
                     : upload_spec(c, c->F + (size_t)field * spec_elems(c), f_hat, stride);
    if (!rc) {
        sh.have_state = 1;
        sh.fields = FIELDS_EXPLICIT;
    }
    return rc;
}

int rpcf_get_field(rpcf_ctx* c, int field, double* f_hat, long stride) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    if (field < 0 || field > 5) return fail(c, RPCF_EINVAL, "bad field id %d", field);
    SlotHost& sh = c->slots[0];
    if (sh.fields == FIELDS_STALE && sh.have_state) {
        if ((rc = enq_velocity(c, c->ids_zero, 1))) return rc;
        sh.fields = FIELDS_FRESH;
    }
    return c->use_mfma ? download_spec_f32(c, f_hat, (const float2*)c->F + (size_t)field * spec_elems(c), stride)
                       : download_spec(c, f_hat, c->F + (size_t)field * spec_elems(c), stride);
}

int rpcf_get_nonlinear(rpcf_ctx* c, int prev, double* n_v, double* n_eta, long stride) {
    int rc = check_slot(c, 0);
    if (!rc) rc = download_spec(c, n_v, hist(c, 0, 0, prev), stride);
    return rc ? rc : download_spec(c, n_eta, hist(c, 0, 1, prev), stride);
}

int rpcf_set_nonlinear(rpcf_ctx* c, int prev, const double* n_v, const double* n_eta, long stride) {
    int rc = check_slot(c, 0);
    if (!rc) rc = upload_spec(c, hist(c, 0, 0, prev), n_v, stride);
    return rc ? rc : upload_spec(c, hist(c, 0, 1, prev), n_eta, stride);
}

int rpcf_nonlinear(rpcf_ctx* c) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    SlotHost& sh = c->slots[0];
    if (!sh.have_state) return fail(c, RPCF_ESTATE, "rpcf_nonlinear: no state set");
    const bool vel = sh.fields == FIELDS_STALE;
    rc = enq_nonlinear(c, c->ids_zero, 1, vel);
    if (rc) return rc;
    if (vel) sh.fields = FIELDS_FRESH;
    sh.phase ^= 1;
    return sync_ctx(c);
}

int rpcf_linear(rpcf_ctx* c) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    SlotHost& sh = c->slots[0];
    if (!sh.have_state) return fail(c, RPCF_ESTATE, "rpcf_linear: no state set");
    // the fields keep the velocity of the state before the solve until rpcf_velocity
    if (sh.fields == FIELDS_STALE && (rc = enq_velocity(c, c->ids_zero, 1))) return rc;
    rc = enq_linear(c, c->ids_zero, 1);
    sh.fields = FIELDS_EXPLICIT;
    return rc ? rc : sync_ctx(c);
}

int rpcf_velocity(rpcf_ctx* c) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    SlotHost& sh = c->slots[0];
    if (!sh.have_state) return fail(c, RPCF_ESTATE, "rpcf_velocity: no state set");
    rc = enq_velocity(c, c->ids_zero, 1);
    sh.fields = FIELDS_FRESH;
    return rc ? rc : sync_ctx(c);
}

int rpcf_step(rpcf_ctx* c, int nsteps) {
    int rc = check_slot(c, 0);
    if (rc) return rc;
    SlotHost& sh = c->slots[0];
    if (!sh.have_state) return fail(c, RPCF_ESTATE, "rpcf_step: no state set");
    // nextStep + getUVW: the velocity of the new state is formed when it is next needed
    // (by the next nonlinear evaluation, or by rpcf_get_field)
    for (int s = 0; s < nsteps; s++) {
        rc = enq_step(c, c->ids_zero, 1, sh.fields == FIELDS_STALE);
        if (rc) return rc;
        sh.phase ^= 1;
        sh.fields = FIELDS_STALE;
    }
    return sync_ctx(c);
}

int rpcf_bootstrap(rpcf_ctx* c) {
    int rc = rpcf_nonlinear(c);
    return rc ? rc : rpcf_step(c, 1);
}

}  // extern "C"

namespace {
// Pencils of a caller's spectral array (`stride` complex numbers apart) packed to nz apart.
__global__ __launch_bounds__(256) void k_gather_pencils(const double2* __restrict__ src, long stride,
                                                        double2* __restrict__ dst, int nz, long total) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long p = i / nz;
        dst[i] = src[p * stride + (i - p * nz)];
    }
}

// The backward transform of one field on the context's stream: Chebyshev s2p (pencils f,
// nz contiguous) -> planes y -> span C2C -> planes w -> x C2R -> ph; y and w are the
// caller's intermediates of hnx*planes*ny and planes*my*nxh complex values.
int enq_spec_to_phys(rpcf_ctx* c, const double2* f, double2* y, double2* w, double* ph, int padded) {
    const Dims d = c->d;
    const int planes = padded ? d.pz : d.nz;
    const FftPlan& plz = c->pl_zn;
    const int L = 2 * d.N, ld = L + 1;
    const int TB = pick_tile(16, L, c->cap_zn, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * ld; });
    const int RP1 = pick_tile(8, d.mx, c->cap_x, c->lds_limit, [&](int t) { return sizeof(double2) * (size_t)t * c->ld_mx; });
    const double2* tw4 = (const double2*)c->pl_zp.tw;
    dim3 g((d.ny + TB - 1) / TB, d.hnx, 1);
    const size_t lds = sizeof(double2) * TB * ld;
    if (padded) k_cheb_s2p<1><<<g, kFftThreads, lds, c->stream>>>(f, 0, 0, y, 0, 0, d, plz, tw4, TB, ld, c->ids_zero);
    else k_cheb_s2p<0><<<g, kFftThreads, lds, c->stream>>>(f, 0, 0, y, 0, 0, d, plz, tw4, TB, ld, c->ids_zero);
    dim3 g2((d.nxh + c->TK_inv - 1) / c->TK_inv, planes, 1);
    k_span_inv<0><<<g2, kFftThreads, sizeof(double2) * c->TK_inv * c->ld_my, c->stream>>>(
        y, 0, 0, w, 0, 0, d, planes, c->pl_y, c->TK_inv, c->ld_my, c->ids_zero);
    dim3 g3((d.my / 2 + RP1 - 1) / RP1, planes);
    k_x_c2r<<<g3, kFftThreads, sizeof(double2) * RP1 * c->ld_mx, c->stream>>>(w, ph, d, c->pl_x, RP1, c->ld_mx);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RPCF_OK : fail(c, RPCF_EDEVICE, "spec_to_phys launch: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" {

int rpcf_spec_to_phys(rpcf_ctx* c, const double* spec, double* phys, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!spec || !phys) return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys: null array");
    const Dims d = c->d;
    const int planes = padded ? d.pz : d.nz;
    double2 *f = nullptr, *y = nullptr, *w = nullptr;
    double* ph = nullptr;
    int rc = dalloc(c, &f, spec_elems(c));
    if (!rc) rc = dalloc(c, &y, (size_t)d.hnx * planes * d.ny);
    if (!rc) rc = dalloc(c, &w, (size_t)planes * d.my * d.nxh);
    if (!rc) rc = dalloc(c, &ph, (size_t)planes * d.my * d.mx);
    if (!rc) rc = upload_spec(c, f, spec, stride);
    if (!rc) rc = enq_spec_to_phys(c, f, y, w, ph, padded);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "spec_to_phys sync");
    if (!rc) rc = copy_sync(c, phys, ph, sizeof(double) * planes * d.my * d.mx, hipMemcpyDeviceToHost);
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
    if (!rc) rc = copy_sync(c, ph, phys, sizeof(double) * planes * d.my * d.mx, hipMemcpyHostToDevice);
    if (!rc) {
        dim3 g3((d.my / 2 + RP1 - 1) / RP1, planes);
        k_x_r2c<<<g3, kFftThreads, sizeof(double2) * RP1 * c->ld_mx, c->stream>>>(ph, z, d.hnx, c->fac, d, c->pl_x, RP1,
                                                                                  c->ld_mx);
        dim3 g2((d.hnx + c->TK_fwd - 1) / c->TK_fwd, planes, 1);
        k_span_fwd<0><<<g2, kFftThreads, sizeof(double2) * c->TK_fwd * c->ld_my, c->stream>>>(
            z, 0, 0, d.hnx, y, 0, 0, d, planes, c->pl_y, c->TK_fwd, c->ld_my, c->ids_zero);
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

int rpcf_spec_to_phys_static(rpcf_ctx* c, const double* spec_dev, double* phys_dev, long stride, int padded) {
    if (!c) return RPCF_EINVAL;
    if (!spec_dev || !phys_dev) return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys_static: null array");
    if (((uintptr_t)spec_dev & 15) || ((uintptr_t)phys_dev & 7))
        return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys_static: misaligned device array");
    const Dims d = c->d;
    if (stride < d.nz) return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys_static: pencil stride %ld < nz = %d", stride, d.nz);
    const int planes = padded ? d.pz : d.nz;
    auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bf = up256(sizeof(double2) * spec_elems(c));
    const size_t by = up256(sizeof(double2) * (size_t)d.hnx * planes * d.ny);
    const size_t bw = up256(sizeof(double2) * (size_t)planes * d.my * d.nxh);
    if (bf + by + bw > RPCF_STATIC_SCRATCH_BYTES)
        return fail(c, RPCF_EINVAL, "rpcf_spec_to_phys_static: needs %zu bytes of scratch, the maximum is %zu",
                    bf + by + bw, (size_t)RPCF_STATIC_SCRATCH_BYTES);
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = (unsigned char*)g_scratch_dev[c->device];
    }
    if (!base) return fail(c, RPCF_ESTATE, "rpcf_spec_to_phys_static: no scratch on device %d", c->device);
    double2* f = (double2*)base;
    double2* y = (double2*)(base + bf);
    double2* w = (double2*)(base + bf + by);
    int rc = sync_ctx(c);
    if (rc) return rc;
    const long total = (long)spec_elems(c);
    const int blocks = (int)std::min<long>((total + 255) / 256, 4L * c->ncu);
    k_gather_pencils<<<blocks, 256, 0, c->stream>>>((const double2*)spec_dev, stride, f, d.nz, total);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, RPCF_EDEVICE, "rpcf_spec_to_phys_static launch: %s", hipGetErrorString(e));
    rc = enq_spec_to_phys(c, f, y, w, phys_dev, padded);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, RPCF_EDEVICE, "spec_to_phys_static sync");
    return rc;
}

int rpcf_get_inverse(rpcf_ctx* c, int eq, int pencil, double* out) {
    if (!c || !out || eq < 0 || eq > 1 || pencil < 0 || pencil >= c->d.P) return RPCF_EINVAL;
    int rc = check_slot(c, 0);
    if (rc) return rc;
    const Dims d = c->d;
    const int kx = pencil % d.hnx, ky = pencil / d.hnx;
    const int kyh = ky <= d.ny / 2 ? ky : d.ny - ky;
    const size_t nn = (size_t)d.nz * d.nz;
    std::vector<double2> m(nn);
    rc = sync_ctx(c);
    if (rc) return rc;
    RCCHK(copy_sync(c, m.data(), c->MINV + ((size_t)eq * d.nslot + (size_t)kyh * d.hnx + kx) * nn,
                        sizeof(double2) * nn, hipMemcpyDeviceToHost));
    for (int r = 0; r < d.nz; r++)
        for (int j = 0; j < d.nz; j++) {
            out[2 * ((size_t)r * d.nz + j)] = m[(size_t)j * d.nz + r].x;
            out[2 * ((size_t)r * d.nz + j) + 1] = m[(size_t)j * d.nz + r].y;
        }
    return RPCF_OK;
}

// Bound of the peak perturbation velocity of a spectral state: per velocity component (u, v,
// w of getVelocityKernel_sm, source/velocity.cu:90-170) the sum of the magnitudes of its
// coefficients over all pencils, the conjugate pencil of kx > 0 counted (|T_k| <= 1), and
// the largest of the three.  D v by the descending recurrence, b_0 halved.  The oracle
// computes the same sums in the same order, so both libraries seed identical states.
static double seed_velocity_bound(const rpcf_params* prm, const double* v, const double* e, long stride, int nz,
                                  int ny, int hnx) {
    const double aphi = prm->n_pi_x / 2, beta = prm->n_pi_y / 2;
    double su = 0.0, sv = 0.0, sw = 0.0;
    std::vector<double> b(2 * (size_t)nz + 4);
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
            std::fill(b.begin(), b.end(), 0.0);
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
    return std::max(std::max(su, sv), sw);
}

int rpcf_seed_state(const rpcf_params* prm, uint64_t seed, double amplitude, int kmax, int degmax, double kexp,
                    int peak_bound, double* v_hat, double* eta_hat, long stride) {
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
            const double scale = pow(1.0 + sqrt((double)kx * kx + (double)kys * kys), -kexp);
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
    // amplitude: the factor of the Gaussian coefficients, or (peak_bound) the bound of the
    // peak perturbation velocity
    const double bound = peak_bound ? seed_velocity_bound(prm, v_hat, eta_hat, stride, nz, ny, hnx) : 1.0;
    if (bound > 0.0) {
        const double f = peak_bound ? amplitude / bound : amplitude;
        for (size_t p = 0; p < P; p++)
            for (int k = 0; k < 2 * nz; k++) {
                v_hat[2 * (size_t)stride * p + k] *= f;
                eta_hat[2 * (size_t)stride * p + k] *= f;
            }
    }
    return RPCF_OK;
}

// ---- per-kernel timing (hipEvents on the context's stream) ----------------------------
int rpcf_timing(rpcf_ctx* c, int enable) { return rpcf_timing_kernels(c, enable ? ~0u : 0u); }

int rpcf_timing_kernels(rpcf_ctx* c, unsigned mask) {
    if (!c) return RPCF_EINVAL;
    int rc = sync_ctx(c);
    if (rc) return rc;
    c->timing = mask ? 1 : 0;
    c->timing_mask = mask;
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

// ======================================================================================
// C ABI: groups (include/rpcf_group.h)
// ======================================================================================
int rpcf_group_create(const rpcf_params* shape, int capacity, rpcf_ctx** out) {
    return create_ctx(shape, capacity, out);
}

int rpcf_group_capacity(const rpcf_ctx* c) { return c ? c->capacity : 0; }

int rpcf_slot_bytes(const rpcf_params* shape, size_t* bytes) {
    Dims d;
    if (!bytes || check_sizes(shape, &d)) return RPCF_EINVAL;
    *bytes = slot_sizes(d, shape->precision).bytes();
    return RPCF_OK;
}

int rpcf_member_init(rpcf_ctx* c, int slot, const rpcf_params* prm) {
    if (!c || !prm) return RPCF_EINVAL;
    if (slot < 0 || slot >= c->capacity) return fail(c, RPCF_EINVAL, "slot %d outside 0..%d", slot, c->capacity - 1);
    Dims d;
    if (check_sizes(prm, &d) || !same_shape(prm, &c->shape) || prm->precision != c->shape.precision)
        return fail(c, RPCF_EINVAL, "member parameters do not match the group's grid shape and precision");
    int rc = sync_ctx(c);
    if (rc) return rc;
    c->slots[slot].configured = 0;
    return init_slot(c, slot, prm);
}

int rpcf_group_cache_hits(const rpcf_ctx* c, long* hits) {
    if (!c || !hits) return RPCF_EINVAL;
    *hits = c->matrix_cache_hits;
    return RPCF_OK;
}

int rpcf_member_release(rpcf_ctx* c, int slot) {
    int rc = check_slot(c, slot);
    if (rc) return rc;
    rc = sync_ctx(c);
    c->slots[slot] = SlotHost();
    return rc;
}

int rpcf_member_set_state(rpcf_ctx* c, int slot, const double* v_hat, const double* eta_hat, long stride) {
    int rc = check_slot(c, slot);
    return rc ? rc : set_state_slot(c, slot, v_hat, eta_hat, stride);
}

int rpcf_member_get_state(rpcf_ctx* c, int slot, double* v_hat, double* eta_hat, long stride) {
    int rc = check_slot(c, slot);
    if (!rc) rc = download_spec(c, v_hat, c->SV + slot * c->sS, stride);
    return rc ? rc : download_spec(c, eta_hat, c->SE + slot * c->sS, stride);
}

int rpcf_group_bootstrap(rpcf_ctx* c, const int* slots, int n) {
    std::vector<int> ids;
    int rc = check_list(c, slots, n, &ids);
    if (rc) return rc;
    if ((rc = set_ids(c, ids))) return rc;
    rc = enq_nonlinear(c, c->ids_dev, n, true);
    if (rc) return rc;
    for (int id : ids) c->slots[id].phase ^= 1;
    return group_steps(c, ids, 1);
}

int rpcf_group_step(rpcf_ctx* c, const int* slots, int n, int nsteps) {
    std::vector<int> ids;
    int rc = check_list(c, slots, n, &ids);
    if (rc) return rc;
    if (nsteps < 0) return fail(c, RPCF_EINVAL, "negative step count");
    return group_steps(c, ids, nsteps);
}

int rpcf_member_adopt(rpcf_ctx* c, int slot, rpcf_ctx* src, int src_slot) {
    if (!c || !src) return RPCF_EINVAL;
    if (c == src) return fail(c, RPCF_EINVAL, "rpcf_member_adopt: the source is the group itself");
    if (slot < 0 || slot >= c->capacity) return fail(c, RPCF_EINVAL, "slot %d outside 0..%d", slot, c->capacity - 1);
    if (src_slot < 0 || src_slot >= src->capacity)
        return fail(c, RPCF_EINVAL, "source slot %d outside 0..%d", src_slot, src->capacity - 1);
    if (!same_shape(&src->shape, &c->shape) || src->shape.precision != c->shape.precision)
        return fail(c, RPCF_EINVAL, "rpcf_member_adopt: the source group has another grid shape or precision");
    const SlotHost sh = src->slots[src_slot];
    if (!sh.configured || !sh.have_state)
        return fail(c, RPCF_ESTATE, "rpcf_member_adopt: source slot %d holds no member with a state", src_slot);
    int rc = sync_ctx(src);
    if (rc) return fail(c, rc, "rpcf_member_adopt: source: %s", src->err);
    if ((rc = sync_ctx(c))) return rc;
    hipError_t e = hipSuccess;
    auto cp = [&](auto* dst, const auto* from, long stride) {
        if (e == hipSuccess && stride > 0)
            e = hipMemcpyAsync(dst + slot * stride, from + src_slot * stride, sizeof(*dst) * (size_t)stride,
                               hipMemcpyDeviceToDevice, c->stream);
    };
    cp(c->F, src->F, c->sF);
    cp(c->SV, src->SV, c->sS);
    cp(c->SE, src->SE, c->sS);
    cp(c->NL, src->NL, c->sNL);
    cp(c->MINV, src->MINV, c->sM);
    cp(c->MINVH, src->MINVH, c->sMH);
    cp(c->MINVE, src->MINVE, c->sME);
    cp(c->MINVF, src->MINVF, c->MINVF ? c->sME : 0);
    cp(c->TF, src->TF, c->sTf);
    cp(c->T4Q, src->T4Q, c->sQ);
    cp(c->T2Q, src->T2Q, c->sQ);
    cp(c->U, src->U, c->sU);
    cp(c->DU, src->DU, c->sU);
    cp(c->DDU, src->DDU, c->sU);
    cp(c->MEANU, src->MEANU, c->sPz);
    cp(c->MP, src->MP, 1L);
    cp(c->PH, src->PH, 1L);
    if (e != hipSuccess) return fail(c, RPCF_EDEVICE, "rpcf_member_adopt copy: %s", hipGetErrorString(e));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->slots[slot] = sh;
    return RPCF_OK;
}

int rpcf_group_graphs(rpcf_ctx* c, int enable) {
    if (!c) return RPCF_EINVAL;
    int rc = sync_ctx(c);
    c->use_graphs = enable ? 1 : 0;
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    return rc;
}

}  // extern "C"
