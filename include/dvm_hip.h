/* dvm_hip.h — C-ABI of the MI355X (gfx950) DVM time-step library, libdvm_hip.so.
 *
 * Plain C: opaque handles, plain pointers and sizes, no C++ or torch types.
 * Every entry point returns dvm_status and never throws.  Unless stated
 * otherwise, array arguments are DEVICE pointers and calls are asynchronous
 * on the context's stream; entry points named *_host take host pointers and
 * return after the results are in host memory.
 *
 * Each entry point replaces a piece of the DVMpp reference's C++ interface;
 * the replaced file:line is cited beside it.  The host C++ drop-in classes
 * (headers under dvmpp_amd/host: dvm::VelocityKernel, VortexBlobs, VortexSheet,
 * DVMBase, BackgroundMesh, Random) sit on top of this header.
 */
#ifndef DVM_HIP_H
#define DVM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status
 * Maps onto the reference's exception hierarchy (src/Exceptions.hpp:13-81):
 *   DVM_E_INVALID -> dvm::AlgorithmException   (src/VelocityKernel.cpp:20-27)
 *   DVM_E_SIZE    -> dvm::SizeMismatchException (src/VortexBlobs.cpp:272-285)
 *   DVM_E_NAN     -> dvm::VortexException       (src/VortexSheet.cpp:289-291)
 *   DVM_E_HIP, DVM_E_OOM, DVM_E_STATE -> dvm::DVMException
 */
typedef enum dvm_status {
    DVM_OK = 0,
    DVM_E_INVALID = 1,
    DVM_E_SIZE = 2,
    DVM_E_NAN = 3,
    DVM_E_HIP = 4,
    DVM_E_OOM = 5,
    DVM_E_STATE = 6
} dvm_status;

/* Arithmetic of the pairwise sums.
 *   DVM_F64  : plain FP64, the reference's loop and operation order
 *              (src/VelocityKernel.cpp:30-108); the parity form.
 *   DVM_F32C : FP32 pair terms on spatially sorted, tile-recentred
 *              coordinates with compensated (Kahan) accumulation across
 *              source tiles; results returned in FP64. */
typedef enum dvm_precision { DVM_F64 = 0, DVM_F32C = 1 } dvm_precision;

typedef struct dvm_ctx_s* dvm_ctx_t;

/* ------------------------------------------------------------------ context
 * A context owns (or borrows) one HIP stream, a growable device workspace and
 * the launch-shape cache.  One context per host thread; not thread-safe. */

/* stream == NULL: the context creates and owns a stream on `device`;
 * otherwise it enqueues on the caller's hipStream_t (passed as void*). */
dvm_status  dvm_ctx_create(int device, void* stream, dvm_ctx_t* out);
dvm_status  dvm_ctx_destroy(dvm_ctx_t ctx);
dvm_status  dvm_ctx_sync(dvm_ctx_t ctx);
void*       dvm_ctx_stream(dvm_ctx_t ctx);
const char* dvm_last_error(dvm_ctx_t ctx);
const char* dvm_status_string(dvm_status s);
/* Library version and the gfx target it was built for. */
const char* dvm_version(void);

/* Kernel timing.  While enabled, the library brackets each launch of its
 * dominant kernels with HIP events on the context's stream; dvm_prof_read
 * waits for them and returns the summed elapsed time and launch count of the
 * named kernel (e.g. "bs_f32c", "bs_f64").  No counterpart in the reference
 * (it only prints a wall-clock total, src/main.cpp:89-99). */
dvm_status dvm_prof_enable(dvm_ctx_t ctx, int on);
dvm_status dvm_prof_reset(dvm_ctx_t ctx);
dvm_status dvm_prof_read(dvm_ctx_t ctx, const char* kernel, double* total_ms, int64_t* launches);

/* Device memory helpers for hosts without their own allocator. */
dvm_status dvm_malloc(dvm_ctx_t ctx, size_t bytes, void** dptr);
dvm_status dvm_free(dvm_ctx_t ctx, void* dptr);
dvm_status dvm_memcpy_h2d(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes);
dvm_status dvm_memcpy_d2h(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes);
dvm_status dvm_memcpy_d2d(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------ Biot-Savart
 * Regularised 2-D Biot-Savart direct sum.  For target i and source j:
 *   dx = x_i - x_j, dz = z_i - z_j, r2 = dx^2 + dz^2,
 *   K  = (1 - exp(-r2/s_j^2))/r2 if r2 < 10 s_j^2, else 1/r2,
 *   u_i = -(1/2pi) sum_j K dz circ_j,  w_i = (1/2pi) sum_j K dx circ_j.
 */

/* Blob-on-blob, self excluded by index.
 * Replaces dvm::DirectKernel::computeSelfInduced (src/VelocityKernel.cpp:30-67,
 * decl src/VelocityKernel.hpp:38-40) and the marshal of VortexBlobs::biotsavart
 * (src/VortexBlobs.cpp:229-249).  SoA device arrays of length n. */
dvm_status dvm_bs_self(dvm_ctx_t ctx, int prec, int64_t n, const double* x, const double* z,
                       const double* circ, const double* sigma, double* u, double* w);

/* Sources -> arbitrary targets.  Pairs with r2 <= r2_skip are skipped when
 * guard != 0: guard = 1, r2_skip = 1e-15 gives computeInducedAt
 * (src/VelocityKernel.cpp:69-108, decl src/VelocityKernel.hpp:48-52);
 * guard = 0 gives the unguarded collocation sum of solvevortexsheet
 * (src/VortexSheet.cpp:181-207). */
dvm_status dvm_bs_targets(dvm_ctx_t ctx, int prec, int64_t ns, const double* sx,
                          const double* sz, const double* scirc, const double* ssigma,
                          int64_t nt, const double* tx, const double* tz, int guard,
                          double r2_skip, double* u, double* w);

/* Host-pointer forms of the two sums (AoS {x,z,circ,sigma} particles and AoS
 * {u,w} results, exactly std::vector<dvm::VortexParticle> / VelocityResult
 * storage, src/VelocityKernel.hpp:15-26). */
dvm_status dvm_bs_self_host(dvm_ctx_t ctx, int prec, int64_t n, const double* particles4,
                            double* results2);
dvm_status dvm_bs_targets_host(dvm_ctx_t ctx, int prec, int64_t ns, const double* sources4,
                               int64_t nt, const double* tx, const double* tz,
                               double* results2);

#ifdef __cplusplus
}
#endif
#endif /* DVM_HIP_H */
