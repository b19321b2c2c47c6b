// rpcf_common.hpp -- sizes and small helpers shared by the HIP kernels and the host code.
#pragma once
#include <hip/hip_runtime.h>

// Grid sizes in the reference's naming (include/data.h:189-334, source/data.cpp:3-14):
// mx, my padded x / span lengths; nx = mx/3*2, ny = my/3*2 kept modes; nz = mz/4+1
// Chebyshev coefficients, N = nz-1; pz = mz/2+1 = 2N+1 padded wall-normal points;
// hnx = nx/2+1 kx pencils (the last one is the x-Nyquist, zeroed on the way to physical
// space); nxh = nx/2 kx columns that reach physical space; P = hnx*ny pencils.
struct Dims {
    int mx, my, nx, ny, nz, N, pz, hnx, nxh, P;
    int nslot;  // hnx*(ny/2+1) matrix slots: pencils ky and ny-ky share one matrix
};

// physical span row of pencil ky (source/transpose.cu:238-240)
__host__ __device__ __forceinline__ int span_of(const Dims& d, int ky) {
    return ky > d.ny / 2 ? ky + (d.my - d.ny) : ky;
}
// pencil ky of a physical span row, or -1 if the row is zeroed by the dealiasing mask
// (source/transform.cu:366-368: rows ny/2 .. my-ny/2 are zero on the way back)
__host__ __device__ __forceinline__ int ky_of_span(const Dims& d, int s) {
    if (s < d.ny / 2) return s;
    if (s > d.my - d.ny / 2) return s - (d.my - d.ny);
    return -1;
}

// Scalars of one simulation (slot) of a group: wavenumber scales aphi = Lx/2, beta = Lz/2
// (include/data.h:321-322), Re, dt, rotation number, the flow-rate target (bulk of u0 at
// set_state) and the flow-control switch.
struct MemberParams {
    double aphi, beta, Re, dt, Ro;
    double q_re, q_im;
    int flow_control;
    int parity;  // even base profile: 1, the solves read the parity blocks of the inverses;
                 // 2, the parity blocks and their BF16 remainders (the full inverse's values)
};

// Member addressing of one grouped launch: launch member j runs slot ids[j]; a per-slot
// array of slot s starts at its base pointer + s * (its slot stride, in elements).
struct Members {
    const int* ids;
    const MemberParams* mp;
    const int* phase;  // per slot: N history buffer "current" = phase, "previous" = 1 - phase
};
