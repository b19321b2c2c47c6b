/*
 * rpcf.h -- C ABI of one time step of the Fourier-Chebyshev velocity-vorticity
 * DNS of rotating plane Couette / channel flow.
 *
 * Two libraries export exactly this symbol set:
 *   librpcf_hip.so     the MI355X (gfx950) implementation (HIP kernels, FP64 and mixed)
 *   librpcf_oracle.so  the CPU FP64 restatement used as the parity oracle (tests only)
 *
 * Data conventions (shared by both libraries and by every host array crossing the ABI)
 * ----------------------------------------------------------------------------------
 *  Sizes follow the reference's problem struct (include/data.h:189-334,
 *  source/data.cpp:3-14):  nx = mx/3*2, ny = my/3*2, nz = mz/4+1, pz = mz/2+1,
 *  hnx = nx/2+1.  "x" is streamwise, reference "y" is spanwise, reference "z" is
 *  wall-normal (Chebyshev, points cos(pi*p/(nz-1))).
 *  Spectral arrays are interleaved complex double, one "pencil" per (kx,ky) with
 *  pencil index p = ky*hnx + kx (reference ZXY order, source/linear.cpp:18) and the
 *  Chebyshev index contiguous inside a pencil; consecutive pencils are
 *  `pencil_stride` complex numbers apart (pencil_stride >= nz).  The reference's
 *  tPitch/16 is one valid stride.
 *  The (0,0) pencil of the v array carries the mean streamwise velocity u0 and the
 *  (0,0) pencil of the eta array carries the mean spanwise velocity w0
 *  (source/linear.cpp:43-50, source/velocity.cu:106-124).
 *  Physical arrays are double [plane z][span row y][x], rows of mx values, no pitch.
 *
 * Errors: every entry point returns RPCF_OK (0) or a negative code; the message of
 * the last failure on a context is returned by rpcf_last_error.  Nothing exits the
 * process (the reference's cuCheck exits, source/operation.cpp:86-96).
 * Threading: one context per host thread, or external synchronisation.
 */
#ifndef RPCF_H
#define RPCF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rpcf_ctx rpcf_ctx;

/* Numerical parameters.  mx..dt mirror RPCF_Numerical_Para (include/parameters.h:4-13)
 * and parameter.txt lines 1-3 (source/operation.cpp:115-138): Lx = pi*n_pi_x,
 * Lz(span) = pi*n_pi_y.  The remaining fields extend the reference:
 *   u_profile     base flow U(y) = c0 + c1*y + c2*y^2 replacing the hard-coded
 *                 U = (1+y)/2 of source/coefficient.cu:149-159 and
 *                 source/nonlinear.cu:400-403; {0.5,0.5,0} is the reference.
 *   flow_control  0: none (reference); 1: constant flow rate -- a uniform mean
 *                 pressure-gradient correction keeps the bulk of u0 at its value at set_state.
 *   precision     0: FP64 everywhere (reference REAL_DOUBLE, include/data.h:11); the HIP
 *                 library solves an even base profile (u_profile[1] == 0) with the
 *                 reflection-parity blocks of the inverses (half their bytes) plus the v
 *                 blocks' remainders in BF16 (a sixteenth), which restore the full
 *                 inverses' results: the state agrees with the full-inverse solve to ~1e-13
 *                 after one step (environment RPCF_FULL_INVERSE=1 keeps the full inverses).
 *                 1: mixed (HIP library; the matrix-core form for 2(nz-1) <= 384) -- the
 *                 transform intermediates, the x and span FFTs and the Lamb vector in FP32,
 *                 the wall-normal Chebyshev transforms as split-FP16 matrix-core products
 *                 with FP32 accumulation; velocity recovery, the nonlinear terms, the
 *                 operators and the solves in FP64, an even base profile's solves with
 *                 the parity blocks alone (they agree with the full inverses to the
 *                 inverse's own round-off asymmetry; the eta blocks stored as FP32 high
 *                 and BF16 low parts, ~2^-33 relative).
 *                 2: FP64 with the parity blocks alone (even base profiles; an odd profile
 *                 runs as 0): the state agrees with precision 0 and the reference to about
 *                 1e-11 after one step.  The oracle computes FP64 with full inverses
 *                 whatever this field says. */
typedef struct rpcf_params {
    int mx, my, mz;
    double n_pi_x, n_pi_y;
    double Re, Ro, dt;
    double u_profile[3];
    int flow_control;
    int precision;
} rpcf_params;

enum {
    RPCF_OK = 0,
    RPCF_EINVAL = -1,   /* bad argument or unsupported size */
    RPCF_ENOMEM = -2,   /* host or device allocation failed */
    RPCF_EDEVICE = -3,  /* HIP runtime / kernel launch error */
    RPCF_ESTATE = -4    /* call out of order (e.g. step before set_state) */
};

/* Spectral field ids in the reference's code naming (source/velocity.cu:90-170):
 * u streamwise, v spanwise, w wall-normal velocity; omega_x, omega_y (spanwise),
 * omega_z (wall-normal vorticity).  w and omega_z are the prognostic v and eta. */
enum {
    RPCF_FIELD_U = 0, RPCF_FIELD_V = 1, RPCF_FIELD_W = 2,
    RPCF_FIELD_OX = 3, RPCF_FIELD_OY = 4, RPCF_FIELD_OZ = 5
};

/* Derived sizes of a parameter set; any pointer may be NULL.  Returns RPCF_EINVAL
 * for sizes the implementation cannot run.  Replaces problem::initVars
 * (source/data.cpp:3-31). */
int rpcf_sizes(const rpcf_params* prm, int* nx, int* ny, int* nz, int* pz, int* hnx);

/* Allocate a context: operator T matrices, base profile, per-mode Crank-Nicolson
 * collocation matrices and their inverses, all state buffers.  Replaces
 * getDeviceInfo + allocDeviceMem + initFFT + allocHostMemory + initAuxMatrix
 * (source/main_loop.cpp:22-28; source/RPCFKernels.cu:47-128; source/transform.cu:18-93;
 * source/solver.cpp:216-320). */
int rpcf_create(const rpcf_params* prm, rpcf_ctx** out);
void rpcf_destroy(rpcf_ctx* ctx);
const char* rpcf_last_error(const rpcf_ctx* ctx);

/* Set the prognostic spectral state (v-hat with u0 in pencil (0,0), eta-hat with w0
 * in pencil (0,0)), recover the velocity/vorticity fields and reset rhs_v_p, as the
 * tail of initFlow does (source/RPCFKernels.cu:409-427).  The AB2 history is left
 * as it is (zero after rpcf_create). */
int rpcf_set_state(rpcf_ctx* ctx, const double* v_hat, const double* eta_hat, long pencil_stride);
/* Read the prognostic state (host rhs_v / rhs_omega_y after a step, source/solver.cpp:95-96). */
int rpcf_get_state(rpcf_ctx* ctx, double* v_hat, double* eta_hat, long pencil_stride);

/* Overwrite / read one of the six spectral velocity-vorticity fields consumed by the
 * nonlinear term (device dptr_tu .. dptr_tomega_z).  Coefficients past nz are not read. */
int rpcf_set_field(rpcf_ctx* ctx, int field, const double* f_hat, long pencil_stride);
int rpcf_get_field(rpcf_ctx* ctx, int field, double* f_hat, long pencil_stride);

/* Nonlinear terms N_v and N_eta at the nz collocation points, pencil layout
 * (host nonlinear_v / nonlinear_omega_y, source/rhs.cpp:107-117); prev=1 reads the
 * previous step's (nonlinear_*_p). */
int rpcf_get_nonlinear(rpcf_ctx* ctx, int prev, double* n_v, double* n_eta, long pencil_stride);
/* Restore them (recovery: read_recover_data, source/output.cpp:141-178, which reloads
 * nonlinear_v / nonlinear_omega_y so that the next step's swap makes them the history). */
int rpcf_set_nonlinear(rpcf_ctx* ctx, int prev, const double* n_v, const double* n_eta, long pencil_stride);

/* = getNonlinear + unpadded Chebyshev s2p of N_v, N_eta + AB2 history swap + (0,0)
 * Lamb save (source/nonlinear.cu:25-108; source/rhs.cpp:83-127, 229-236). */
int rpcf_nonlinear(rpcf_ctx* ctx);
/* = get_rhs_v (linear part + AB2) + solveEq(v) + get_rhs_omega + solveEq(eta) +
 * save_0_v_omega_y (source/solver.cpp:43-89; source/rhs.cpp:185-227;
 * source/linear.cpp:7-205). */
int rpcf_linear(rpcf_ctx* ctx);
/* = getUVW after the state copy (source/solver.cpp:95-98; source/velocity.cu:22-170). */
int rpcf_velocity(rpcf_ctx* ctx);
/* nsteps x nextStep (source/solver.cpp:35-101): nonlinear, linear, velocity. */
int rpcf_step(rpcf_ctx* ctx, int nsteps);
/* The reference's AB2 start (source/main_loop.cpp:39-52): an extra step fills the
 * history with N(state0), the state is re-initialised, then the real first step runs,
 * so step 1 uses N^{n-1} = N^n.  After this call one step has been taken. */
int rpcf_bootstrap(rpcf_ctx* ctx);

/* Wall-normal Chebyshev transforms of one field between physical planes and
 * spectral pencils: transform_3d_one (source/transform.cu:95-319).  padded=1 uses the
 * pz dealiasing planes z_q = cos(pi*q/(pz-1)); padded=0 the nz collocation planes.
 * phys is [planes][my][mx] doubles; spec pencils hold nz coefficients (forward keeps
 * k < nz). */
int rpcf_phys_to_spec(rpcf_ctx* ctx, const double* phys, double* spec, long pencil_stride, int padded);
int rpcf_spec_to_phys(rpcf_ctx* ctx, const double* spec, double* phys, long pencil_stride, int padded);

/* Inspection: the inverse collocation matrix of one pencil (eq 0: v equation,
 * 1: eta equation), row-major nz*nz complex; pencil (0,0) holds the u0/w0 matrices
 * (source/solver.cpp:279-290, 297-311). */
int rpcf_get_inverse(rpcf_ctx* ctx, int eq, int pencil, double* out);

/* Seeded wall-satisfying initial field (host only, no device work): complex Gaussian
 * Chebyshev coefficients of degree <= degmax on pencils with kx <= kmax and
 * |signed ky| <= kmax (pencil (0,0) excluded), times (1+|k|)^(-kexp) with
 * |k| = sqrt(kx^2 + ky^2) in index units, multiplied by (1-y^2)^2 for v and (1-y^2)
 * for eta, Hermitian on kx = 0, Nyquist pencils zero.  peak_bound = 0: the coefficients
 * are then multiplied by `amplitude` (the recipe as the benchmark configurations state
 * it); peak_bound = 1: the field is scaled so that `amplitude` bounds the peak
 * perturbation velocity: for each of u, v, w the sum of the magnitudes of its spectral
 * coefficients (which bounds its maximum over the domain) is at most `amplitude` (for
 * recipes whose literal field is too strong for the time step).  Generator: splitmix64 +
 * Box-Muller, draws in pencil order.  No counterpart in the reference (its
 * init_flow_kernel analytic field, source/RPCFKernels.cu:221-356, is replaced). */
int rpcf_seed_state(const rpcf_params* prm, uint64_t seed, double amplitude, int kmax,
                    int degmax, double kexp, int peak_bound, double* v_hat, double* eta_hat,
                    long pencil_stride);

/* Per-kernel timing with HIP events recorded on the context's stream around every
 * launch (no counterpart in the reference, whose step timer is one cudaEvent pair,
 * source/main_loop.cpp:90-102).  rpcf_timing(ctx, 1) clears and enables;
 * rpcf_timing_get returns the summed milliseconds and launch count of kernel `kid`
 * (0 <= kid < rpcf_num_kernels(), named by rpcf_kernel_name).  The oracle implements
 * the same symbols and reports no kernels. */
int rpcf_timing(rpcf_ctx* ctx, int enable);
/* The same with events around the kernels whose bit (1u << kid) is set in `mask` only, so
 * that a timed run pays for the events of the kernels it measures; 0 disables. */
int rpcf_timing_kernels(rpcf_ctx* ctx, unsigned mask);
int rpcf_timing_get(rpcf_ctx* ctx, int kid, double* total_ms, long* count);
const char* rpcf_kernel_name(int kid);
int rpcf_num_kernels(void);

#ifdef __cplusplus
}
#endif

#endif /* RPCF_H */
