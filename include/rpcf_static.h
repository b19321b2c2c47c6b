/*
 * rpcf_static.h -- entry points that allocate no device memory (librpcf_hip.so only).
 *
 * For callers that cannot call hipMalloc / hipFree: code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  The context (rpcf_create) is made
 * beforehand, as for every other entry point; the calls below then allocate nothing.
 * Their intermediates live in one static device array of RPCF_STATIC_SCRATCH_BYTES per GPU
 * (held by each GPU the library runs on until the process ends) that every such call on
 * that GPU shares and that keeps the last call's contents: calls that overlap on one GPU
 * (two host threads, or two contexts of one device used concurrently) race, and must be
 * serialised by the caller.
 */
#ifndef RPCF_STATIC_H
#define RPCF_STATIC_H

#include "rpcf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bytes of the static device scratch array (64 MiB). */
#define RPCF_STATIC_SCRATCH_BYTES (64ul << 20)

/* rpcf_spec_to_phys with device arrays and no allocation.  No counterpart in the
 * reference: it serves the callers that cannot allocate (above).  spec_dev: device
 * array of P pencils of nz interleaved complex doubles, `pencil_stride` complex numbers
 * apart, 16-byte aligned; phys_dev: device array [planes][my][mx] doubles
 * (planes = pz if padded, nz otherwise).  Waits for its work before it returns, as
 * rpcf_spec_to_phys does, and gives bit-identical values.  The intermediates (the
 * pencils, the planes after the Chebyshev transform and after the span transform,
 * 16 * (P*nz + hnx*planes*ny + planes*my*(nx/2)) bytes, each part rounded up to 256 bytes)
 * must fit RPCF_STATIC_SCRATCH_BYTES, else RPCF_EINVAL with a message naming both sizes:
 * e.g. 128x129x128 (mx = my = 192, mz = 512) fits unpadded (60 MB), not padded (102 MB). */
int rpcf_spec_to_phys_static(rpcf_ctx* ctx, const double* spec_dev, double* phys_dev, long pencil_stride,
                             int padded);

#ifdef __cplusplus
}
#endif

#endif /* RPCF_STATIC_H */
