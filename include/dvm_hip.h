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
 *              source tiles; results returned in FP64.
 *   DVM_FMM  : blob-on-blob sums by a 2-D fast multipole method in FP64
 *              (the reference's FMMKernel, src/VelocityKernel.hpp:84-110,
 *              whose body there falls back to the direct sum): exact
 *              regularised near field, far field truncated at the context's
 *              expansion order and opening criterion (dvm_set_fmm).  Sums at
 *              separate targets use DVM_F64, as the reference's
 *              FMMKernel::computeInducedAt does.  Not for the batched engine. */
typedef enum dvm_precision { DVM_F64 = 0, DVM_F32C = 1, DVM_FMM = 2 } dvm_precision;

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

/* Allocation-free form of dvm_bs_self; it has no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the sum's
 * scratch (sort keys, tiles, source copies, chunk partials) lives in a device
 * array of the library, dvm_bs_self_noalloc_scratch_bytes() = 64 MiB, one per
 * GPU, held until the process ends; the chunk partials take at most 16 MiB of
 * it.  prec is DVM_F64 or DVM_F32C.  A call that needs more scratch returns
 * DVM_E_SIZE before launching anything (about 1.5M blobs in FP64, 0.5M in
 * FP32C).  The caller passes the inputs and outputs, as for dvm_bs_self; the
 * results agree with it to the form's tolerance (bit for bit unless the
 * partial cap changes the chunking).  Calls that overlap on one GPU share the
 * array and race. */
dvm_status dvm_bs_self_noalloc(dvm_ctx_t ctx, int prec, int64_t n, const double* x,
                               const double* z, const double* circ, const double* sigma,
                               double* u, double* w);
size_t dvm_bs_self_noalloc_scratch_bytes(void);

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

/* DVM_F32C launch variants: source tile (128 | 256 sources), target block
 * (64 | 256 targets, i.e. 1 or 4 waves per source tile visit), sources per
 * far-field reciprocal (1 | 2 | 4) and an occupancy target (resident
 * 256-lane blocks per CU that the source range is split to reach).  Every
 * variant meets the same accuracy bound; their rounding differs, so results
 * are bitwise reproducible only for a fixed variant.
 *   DVM_TUNE_TABLE   (default) a fixed variant per problem size.
 *   DVM_TUNE_MEASURE the first call of each kind (self / targets / guarded
 *                    targets) and size bucket (log2 of target and source
 *                    counts) times every variant on the context's stream,
 *                    synchronising, and the context keeps the fastest.
 *                    Calls made while the stream is being captured use the
 *                    table.  The environment variable DVM_TUNE=measure
 *                    selects this mode at context creation.
 * dvm_set_f32c_variant forces one variant (tile 0 clears it: DVM_E_INVALID
 * for a variant that does not exist); dvm_get_f32c_variant reports the
 * variant of the last F32C call {tile, block, share, occupancy}.  No
 * counterpart in the reference (its kernel has no launch shape). */
enum { DVM_TUNE_TABLE = 0, DVM_TUNE_MEASURE = 1 };
dvm_status dvm_set_tuning(dvm_ctx_t ctx, int mode);
dvm_status dvm_set_f32c_variant(dvm_ctx_t ctx, int tile, int block, int share, int occupancy);
dvm_status dvm_get_f32c_variant(dvm_ctx_t ctx, int32_t* out4);

/* ------------------------------------------------------------------ vortex sheet
 * One body's vortex sheet, resident on the device: panel geometry, the
 * influence matrix A ((Nl+1) x Nl), the least-squares operator
 * (A^T A)^-1 A^T formed once (A is time-invariant), and the surface
 * vorticity gamma. */
typedef struct dvm_sheet_s* dvm_sheet_t;

/* Physical constants of src/VortexSheet.cpp:12-54 (XML constants/, flow/, time/). */
typedef struct dvm_sheet_params {
    double rho;        /* constants/density */
    double nu;         /* constants/nu */
    double dt;         /* time/dt */
    double ux, uz;     /* flow/ux, flow/uz */
    double cutoff_exp; /* constants/cutoff_exp: nascent sigma = ds^cutoff_exp */
    int max_npv;       /* constants/max_NumPanelVort */
} dvm_sheet_params;

/* VortexSheet(const XmlHandler&, stamp) minus file I/O: read_input_coord,
 * form_vortex_sheet and compute_influence_matrix (src/VortexSheet.cpp:12-165).
 * px, pz: HOST arrays of npts body points, closed (last == first), so
 * Nl = npts - 1 panels. */
dvm_status dvm_sheet_create(dvm_ctx_t ctx, int npts, const double* px, const double* pz,
                            const dvm_sheet_params* params, dvm_sheet_t* out);
dvm_status dvm_sheet_destroy(dvm_sheet_t sheet);
/* VortexSheet::size (src/VortexSheet.cpp:516-530). */
dvm_status dvm_sheet_size(dvm_sheet_t sheet, int* nl);
/* Geometry to HOST arrays (each Nl long; any may be NULL); aref = x-extent
 * (src/VortexSheet.cpp:74). */
dvm_status dvm_sheet_geometry(dvm_sheet_t sheet, double* xc, double* zc, double* theta,
                              double* ds, double* enx, double* enz, double* etx, double* etz,
                              double* aref);
/* Influence matrix to a HOST array, row-major (Nl+1) x Nl (src/VortexSheet.cpp:104-165). */
dvm_status dvm_sheet_influence(dvm_sheet_t sheet, double* A);
/* Device pointer to gamma (Nl doubles), valid until dvm_sheet_destroy. */
dvm_status dvm_sheet_gamma_ptr(dvm_sheet_t sheet, double** gamma);
/* gamma to / from HOST memory. */
dvm_status dvm_sheet_get_gamma(dvm_sheet_t sheet, double* gamma);
dvm_status dvm_sheet_set_gamma(dvm_sheet_t sheet, const double* gamma);
/* VortexSheet::solvevortexsheet (src/VortexSheet.cpp:167-215): collocation
 * sum of the n device blobs (unguarded), RHS b = (U+u).n, b_Nl = -sum circ,
 * gamma = (A^T A)^-1 A^T b.  b_host (Nl+1, may be NULL) receives b.  The
 * collocation sum is FP64 whatever prec names (a valid dvm_precision): the
 * loads scale gamma by rho/dt, so its right-hand side keeps FP64. */
dvm_status dvm_sheet_solve(dvm_sheet_t sheet, int prec, int64_t n, const double* x,
                           const double* z, const double* circ, const double* sigma,
                           double* b_host);
/* VortexSheet::vortexsheetbc (src/VortexSheet.cpp:217-273): sheet-induced
 * velocity at n device points, fused (no N x Nl coefficient matrices). */
dvm_status dvm_sheet_on_blobs(dvm_sheet_t sheet, int64_t n, const double* x, const double* z,
                              double* uvs, double* wvs);
/* The same velocities in the form the fast simulations use: blobs farther
 * than 2 body radii from the body's centroid take a 40-term Laurent series of
 * the sheet (its panels are straight vortex segments; truncation about 1e-12
 * of the leading term), the others the panel sum above.  No counterpart in
 * the reference. */
dvm_status dvm_sheet_on_blobs_far(dvm_sheet_t sheet, int64_t n, const double* x, const double* z,
                              double* uvs, double* wvs);
/* VortexSheet::compute_loads (src/VortexSheet.cpp:487-514) and get_forces
 * (:552-555): out4 (HOST) = {fx, fz, cd, cl}. */
dvm_status dvm_sheet_loads(dvm_sheet_t sheet, double ur, double* out4);
/* VortexSheet::probe_velocities (src/VortexSheet.cpp:559-646, decl
 * src/VortexSheet.hpp:126) in the reference's form.  The reference never
 * calls it (DVMBase::probe_velocities is declared, src/DVMBase.hpp:98, but not
 * defined); the isolated-sum probes use dvm_bs_targets instead.  Kept quirks:
 * the blob sum starts at blob 1 and has no self or r^2 guard; the sheet term
 * is that of panel Nl-1 alone (the panel loop assigns instead of adding),
 * paired with gamma_0 with the qy term subtracted, without 1/(2 pi).  Device
 * arrays: n blobs, np probes; u, w (Probe::m_u, m_w) = (1/2pi) blob sum +
 * sheet term + U; uvs, wvs (Probe::m_uvs, m_wvs) = the sheet term. */
dvm_status dvm_sheet_probe_quirk(dvm_sheet_t sheet, int prec, int64_t n, const double* x,
                                 const double* z, const double* circ, const double* sigma,
                                 int64_t np, const double* px, const double* pz, double* u,
                                 double* w, double* uvs, double* wvs);
/* VortexSheet::inside_body (src/VortexSheet.cpp:447-464) for n device points:
 * flags[i] = 1 if inside (crossing number == 1). */
dvm_status dvm_sheet_inside(dvm_sheet_t sheet, int64_t n, const double* x, const double* z,
                            int* flags);
/* VortexSheet::reflect (src/VortexSheet.cpp:345-405), in place on device x, z. */
dvm_status dvm_sheet_reflect(dvm_sheet_t sheet, int64_t n, double* x, double* z);

/* ------------------------------------------------------------------ blob operations
 * Single steps of the time loop on caller-owned device SoA arrays. */

/* Euler convection, src/DVMBase.cpp:179-180: x += ((u + uvs) + ux) dt. */
dvm_status dvm_advect_euler(dvm_ctx_t ctx, int64_t n, double ux, double uz, double dt,
                            const double* u, const double* w, const double* uvs,
                            const double* wvs, double* x, double* z);
/* One low-storage RK3 stage, src/DVMBase.cpp:201-205:
 * q = a q + ((u + uvs) + U) dt; x += b q (q1, q2 zeroed by the caller before
 * stage 0, as src/DVMBase.cpp:188-189 does). */
dvm_status dvm_advect_rk3_stage(dvm_ctx_t ctx, int64_t n, double a, double b, double ux,
                                double uz, double dt, const double* u, const double* w,
                                const double* uvs, const double* wvs, double* q1, double* q2,
                                double* x, double* z);
/* Random-walk diffusion with given increments (VortexBlobs::diffusion_random_walk,
 * src/VortexBlobs.cpp:251-265): x_i += dxdz[2i], z_i += dxdz[2i+1]; the
 * increments are r cos(theta), r sin(theta) of the reference's Box-Muller step. */
dvm_status dvm_diffuse(dvm_ctx_t ctx, int64_t n, const double* dxdz, double* x, double* z);
/* Nascent-vortex release, split around the random draws
 * (VortexSheet::release_nascent_vortices_rw, src/VortexSheet.cpp:275-333):
 * _counts computes the per-panel counts from gamma and returns their total
 * (DVM_E_NAN on non-finite counts); _emit writes the nrv new blobs at index
 * `base` of the given device arrays, blob k offset by rrw[k] (device) along the
 * panel's outward normal; IDs are set to `id_value` (Nold in append_vortices,
 * src/VortexBlobs.cpp:215-218). */
dvm_status dvm_sheet_release_counts(dvm_sheet_t sheet, int64_t* nrv);
dvm_status dvm_sheet_release_emit(dvm_sheet_t sheet, int64_t nrv, const double* rrw,
                                  uint32_t id_value, int64_t base, double* x, double* z,
                                  double* circ, double* sigma, double* u, double* w,
                                  double* uvs, double* wvs, uint32_t* id);
/* VortexSheet::delete_inside (src/VortexSheet.cpp:407-445) in place on device
 * arrays: stable compaction of the blobs outside the body, IDs renumbered
 * 0..M-1 (unchanged when nothing is deleted); *n_out = M. */
dvm_status dvm_sheet_delete_inside(dvm_sheet_t sheet, int64_t n, double* x, double* z,
                                   double* circ, double* sigma, double* u, double* w,
                                   double* uvs, double* wvs, uint32_t* id, int64_t* n_out);

/* ------------------------------------------------------------------ simulation
 * One DVM simulation (one body) with all blob and sheet state resident on the
 * device: DVMBase's time loop (src/DVMBase.cpp:67-209) without its text I/O. */
typedef struct dvm_sim_s* dvm_sim_t;

/* <remesh> block of the XML (src/XmlHandler.cpp:192-201); frequency 0 disables. */
typedef struct dvm_remesh_params {
    int frequency;        /* remesh every `frequency` steps */
    int kernel;           /* 0 = M4prime, 1 = gaussian, 2 = M6 */
    double xmin, xmax, zmin, zmax, dx, dz;
    double sigma;         /* core size of the new blobs */
    double circ_threshold;/* nodes with |circ| <= threshold create no blob */
} dvm_remesh_params;

/* BackgroundMesh::remesh (use site src/DVMBase.cpp:160-165; the reference's
 * BackgroundMesh source is missing, the API is that of
 * tests/unit/test_BackgroundMesh.cpp:14-18,105-111,167-204).  Redistributes n
 * device blobs onto the mesh: nodes with |circ| > circ_threshold in linear
 * order k*nx + i (sigma = params->sigma), then the blobs whose interpolation
 * stencil leaves the mesh, unchanged.  *n_out receives the output count; if it
 * exceeds out_cap nothing is written and DVM_E_SIZE is returned. */
dvm_status dvm_remesh(dvm_ctx_t ctx, const dvm_remesh_params* params, int64_t n,
                      const double* x, const double* z, const double* circ,
                      const double* sigma, int64_t out_cap, double* ox, double* oz,
                      double* ocirc, double* osigma, int64_t* n_out);

typedef enum dvm_scheme { DVM_EULER = 0, DVM_RK3 = 1 } dvm_scheme;
typedef enum dvm_crossing { DVM_DELETE = 0, DVM_ABSORB = 1, DVM_REFLECT = 2 } dvm_crossing;
/* Random stream of the diffusion and release steps.
 *  DVM_RNG_REFERENCE: std::mt19937 + uniform_real_distribution<double> seeded
 *    as src/Random.cpp:3-18, consumed in the reference's order (two draws per
 *    blob for diffusion, one per nascent blob); the host turns the draws into
 *    displacement increments with the C library's libm and the device applies
 *    them, so positions match the reference CPU path bit for bit.
 *  DVM_RNG_DEVICE: counter-based Philox4x32-10 on the device keyed by
 *    (seed, step, blob index): same distributions, no host round trip. */
typedef enum dvm_rng_mode { DVM_RNG_REFERENCE = 0, DVM_RNG_DEVICE = 1 } dvm_rng_mode;

typedef struct dvm_sim_params {
    dvm_sheet_params sheet;
    int scheme;        /* dvm_scheme (time/scheme) */
    int crossing;      /* dvm_crossing (algorithms/surface_crossing) */
    int seed;          /* constants/seed (< 0: std::random_device) */
    int prec;          /* dvm_precision of the pair sums */
    int rng_mode;      /* dvm_rng_mode */
    dvm_remesh_params remesh;
    double merge_threshold; /* algorithms/merge_threshold (src/DVMBase.cpp:49-60) */
    int merge_frequency;    /* algorithms/merge_frequency; merging runs when both are > 0 */
} dvm_sim_params;

/* DVMBase(XmlHandler&, stamp) (src/DVMBase.cpp:8-65): body points px, pz on
 * the HOST (closed contour), zero blobs. */
dvm_status dvm_sim_create(dvm_ctx_t ctx, int npts, const double* px, const double* pz,
                          const dvm_sim_params* params, dvm_sim_t* out);
dvm_status dvm_sim_destroy(dvm_sim_t sim);
/* increment_step + compute_step (src/DVMBase.cpp:70-71, 86-133): convect
 * (Euler or RK3), remesh, diffusion, release + append, crossing, loads, and
 * merge_blobs every merge_frequency steps when configured.
 * DVM_E_NAN on the VortexException path (non-finite release counts). */
dvm_status dvm_sim_step(dvm_sim_t sim);
/* `steps` calls of dvm_sim_step (DVMBase::solve without output). */
dvm_status dvm_sim_run(dvm_sim_t sim, int steps);
/* Sub-steps of compute_step, in its order, for lock-step checking:
 * 0 = increment_step + convect, 1 = remesh, 2 = diffusion,
 * 3 = release + append, 4 = crossing, 5 = loads, 6 = merge (when due). */
dvm_status dvm_sim_substep(dvm_sim_t sim, int which);
/* Merges made by the last merge sub-step (0 when it was not due); the
 * reference prints this count (src/DVMBase.cpp:128-131). */
dvm_status dvm_sim_merged(dvm_sim_t sim, int64_t* merged);
dvm_status dvm_sim_size(dvm_sim_t sim, int64_t* n);
dvm_status dvm_sim_step_index(dvm_sim_t sim, int* step);
/* Uniform draws consumed so far from the reference stream. */
dvm_status dvm_sim_rng_draws(dvm_sim_t sim, uint64_t* draws);
/* Blob state to / from HOST arrays (VortexBlobs fields, src/VortexBlobs.hpp:24-32);
 * any output may be NULL. */
dvm_status dvm_sim_get_state(dvm_sim_t sim, uint32_t* id, double* x, double* z, double* circ,
                             double* sigma, double* u, double* w, double* uvs, double* wvs);
dvm_status dvm_sim_set_state(dvm_sim_t sim, int64_t n, const uint32_t* id, const double* x,
                             const double* z, const double* circ, const double* sigma,
                             const double* u, const double* w, const double* uvs,
                             const double* wvs);
/* gamma (Nl) to / from HOST memory. */
dvm_status dvm_sim_get_gamma(dvm_sim_t sim, double* gamma);
dvm_status dvm_sim_set_gamma(dvm_sim_t sim, const double* gamma);
/* Loads of the last step to HOST: {fx, fz, cd, cl}. */
dvm_status dvm_sim_loads(dvm_sim_t sim, double* out4);
/* Device pointers to the blob SoA (valid until the next call that changes N). */
dvm_status dvm_sim_blobs(dvm_sim_t sim, double** x, double** z, double** circ, double** sigma);

/* Snapshot / restore of a whole simulation: parameters, body points, step
 * index, blobs (all VortexBlobs fields), gamma, the last loads, and both
 * random streams (the mt19937 state and draw count of the reference stream,
 * the counter-based stream's seed).  A restored simulation continues exactly
 * as the snapshotted one would have, bit for bit.  The reference has no
 * checkpoint; its output is per-step text only (src/DVMBase.cpp:151-158).
 * dvm_sim_snapshot_size gives the bytes dvm_sim_snapshot writes into HOST
 * memory; dvm_sim_restore reads such a buffer into a simulation created with
 * the same body and parameters (DVM_E_INVALID otherwise, DVM_E_SIZE for a
 * truncated buffer).  Both synchronise the context's stream. */
dvm_status dvm_sim_snapshot_size(dvm_sim_t sim, size_t* bytes);
dvm_status dvm_sim_snapshot(dvm_sim_t sim, void* buf, size_t bytes);
dvm_status dvm_sim_restore(dvm_sim_t sim, const void* buf, size_t bytes);

/* Output staging without a stall.  dvm_sim_copy_out enqueues, on the
 * context's stream, copies of the current blob positions and circulations
 * (the first min(n, cap)), gamma (Nl) and the last loads {fx, fz, cd, cl}
 * into HOST buffers — pinned (dvm_host_alloc) for the copies to run
 * asynchronously — and returns the blob count n at once; NULL buffers are
 * skipped.  Record an event after it (dvm_event_record) and wait on the event
 * before reading the buffers; the simulation can step meanwhile.  This is
 * the state the reference writes every step (VortexBlobs::write_step,
 * VortexSheet::write_step, src/VortexBlobs.cpp:310-331,
 * src/VortexSheet.cpp:335-343). */
dvm_status dvm_sim_copy_out(dvm_sim_t sim, int64_t cap, double* x, double* z, double* circ,
                            double* gamma, double* loads4, int64_t* n);
/* Probe velocities for the simulation's current state, in the reference's
 * probe_velocities form (dvm_sheet_probe_quirk); device arrays of np probes. */
dvm_status dvm_sim_probe_quirk(dvm_sim_t sim, int prec, int64_t np, const double* px,
                               const double* pz, double* u, double* w, double* uvs, double* wvs);

/* Events and pinned host memory for the staging above.  dvm_event_record
 * marks the context stream's current position; dvm_event_wait blocks until
 * the work before it has finished and may be called from any host thread;
 * dvm_event_query sets *done to 1 once it has. */
typedef struct dvm_event_s* dvm_event_t;
dvm_status dvm_event_create(dvm_ctx_t ctx, dvm_event_t* ev);
dvm_status dvm_event_record(dvm_ctx_t ctx, dvm_event_t ev);
dvm_status dvm_event_wait(dvm_event_t ev);
dvm_status dvm_event_query(dvm_event_t ev, int* done);
dvm_status dvm_event_destroy(dvm_event_t ev);
dvm_status dvm_host_alloc(size_t bytes, void** ptr);
dvm_status dvm_host_free(void* ptr);

/* FMMKernel(expansion_order, theta) (src/VelocityKernel.hpp:84-90; the
 * reference's defaults are 10 and 0.5): the parameters of DVM_FMM sums on
 * this context.  The order (1..32) is the number of expansion terms; theta in
 * (0.2, 1] separates boxes whose centre distance exceeds sqrt(2) * width /
 * theta.  The error falls roughly as theta^order.  Inputs with non-finite
 * values, fewer than two blobs, or too few for a tree of three levels use the
 * direct FP64 sum; dvm_fmm_last_level reports the leaf level of the last
 * DVM_FMM sum (0: it was direct). */
dvm_status dvm_set_fmm(dvm_ctx_t ctx, int expansion_order, double theta);
dvm_status dvm_fmm_last_level(dvm_ctx_t ctx, int* level);

/* VortexBlobs::merge_blobs (src/VortexBlobs.cpp:110-209) on n device blobs,
 * in place: the reference's greedy pass in index order (blob i absorbs each
 * later unabsorbed j with |Gi Gj / (Gi + Gj)| * |xi - xj| < threshold against
 * i's current state; moments conserved; u w uvs wvs of an absorbing blob set
 * to 0), with the same decisions and values.  When anything merged, the
 * survivors are compacted to the front in order and, if `id` is not NULL,
 * their IDs set to 0..m-1.  *merged = n - m (the reference's return value).
 * Synchronizes the context's stream. */
dvm_status dvm_merge_blobs(dvm_ctx_t ctx, double threshold, int64_t n, uint32_t* id, double* x,
                           double* z, double* circ, double* sigma, double* u, double* w,
                           double* uvs, double* wvs, int64_t* merged);
/* Work counts of the context's last merge: blobs that had a partner in the
 * input, and how many of their speculative absorption sequences had to be
 * recomputed because an earlier blob took one of their partners.  No
 * counterpart in the reference. */
dvm_status dvm_merge_stats(dvm_ctx_t ctx, int64_t* candidates, int64_t* recomputed);

/* ------------------------------------------------------------------ batched ensemble
 * Many independent simulations (different bodies, viscosities, time steps,
 * seeds and step counts; ragged blob counts) stepped together: each launch
 * covers every active simulation, the pair sum is one persistent launch over
 * (simulation, target tile) work items, and blob counts, step counters and
 * status words stay on the device so a step is replayed from a hipGraph.
 * Diffusion and release draw from the device Philox stream (DVM_RNG_DEVICE).
 * No counterpart in the reference, which runs one simulation per process
 * (src/main.cpp:20-115); each simulation follows DVMBase::compute_step
 * (src/DVMBase.cpp:86-133). */
typedef struct dvm_batch_s* dvm_batch_t;

typedef struct dvm_batch_config {
    int slots;              /* simulations held at once (<= 1024) */
    int64_t slot_capacity;  /* blob capacity of each slot */
    int max_panels;         /* largest body (panels) a slot accepts (<= 1024) */
    int scheme;             /* dvm_scheme, shared by all slots */
    int crossing;           /* dvm_crossing, shared */
    int prec;               /* dvm_precision of the pair sums */
    int use_graph;          /* 1: replay each step from a captured hipGraph */
    dvm_remesh_params remesh; /* shared remesh settings (frequency 0 = none) */
} dvm_batch_config;

/* Device bytes a batch with this configuration allocates. */
dvm_status dvm_batch_footprint(const dvm_batch_config* cfg, size_t* bytes);
dvm_status dvm_batch_create(dvm_ctx_t ctx, const dvm_batch_config* cfg, dvm_batch_t* out);
dvm_status dvm_batch_destroy(dvm_batch_t batch);
/* Starts a simulation of `steps` steps in a free slot (DVM_E_STATE if none):
 * the body (HOST points, closed) gets its sheet set up as dvm_sheet_create
 * does; zero blobs. */
dvm_status dvm_batch_admit(dvm_batch_t batch, int npts, const double* px, const double* pz,
                           const dvm_sheet_params* params, int seed, int steps, int* slot);
/* dvm_batch_admit for `count` bodies at once (arrays of count entries; px[i],
 * pz[i] HOST points of body i): their set-up (influence matrix, least-squares
 * operator, moments) runs concurrently on several streams, with one host
 * synchronisation in all.  slots[i] receives body i's slot.  No counterpart
 * in the reference (it has one body). */
dvm_status dvm_batch_admit_many(dvm_batch_t batch, int count, const int* npts,
                                const double* const* px, const double* const* pz,
                                const dvm_sheet_params* params, const int* seed,
                                const int* steps, int* slots);
/* Advances every active simulation by `steps` time steps. */
dvm_status dvm_batch_step(dvm_batch_t batch, int steps);
/* Reports up to `max` simulations that stopped since the last poll (finished
 * their steps, or failed: status DVM_E_NAN / DVM_E_SIZE); their slots become
 * free, their state stays readable until the slot is admitted again. */
dvm_status dvm_batch_poll(dvm_batch_t batch, int max, int* slots, int* status, int* count);
dvm_status dvm_batch_slot_info(dvm_batch_t batch, int slot, int64_t* n, int* step, int* active,
                               int* status, double* loads4);
/* A slot's blob fields to HOST arrays (n = dvm_batch_slot_info's n). */
dvm_status dvm_batch_get_state(dvm_batch_t batch, int slot, double* x, double* z, double* circ,
                               double* sigma, double* u, double* w);
/* Replaces an admitted slot's blobs by n blobs from HOST arrays (velocities
 * zero, IDs 0..n-1; the step count is kept), as VortexBlobs' public fields
 * are written directly (src/VortexBlobs.hpp:24-32); n <= slot_capacity.  A
 * one-slot batch so started from a saved wake runs that simulation with
 * graph-replayed steps. */
dvm_status dvm_batch_set_state(dvm_batch_t batch, int slot, int64_t n, const double* x,
                               const double* z, const double* circ, const double* sigma);
/* Diagnostic, no counterpart in the reference: the first n entries of a
 * slot's blob order in its last FP32C pair sum (storage indices sorted by
 * Morton key, stable), n = the slot's count when that sum ran. */
dvm_status dvm_batch_sort_order(dvm_batch_t batch, int slot, int64_t n, uint32_t* order);
/* Slots holding at least min_blobs blobs at the start of a step take that
 * step's blob-on-blob sum by the fast multipole form of DVM_FMM (the
 * reference's FMMKernel, src/VelocityKernel.hpp:84-110, with this expansion
 * order and opening criterion) on their own arrays, before the step's
 * graphs; the others keep the batch's FP32C direct sum.  The choice depends
 * on a slot's own count, so its result is the same alone or in any mix.
 * min_blobs 0 turns it off.  FP32C Euler batches only (DVM_E_INVALID
 * otherwise: an RK3 step sums three times inside one graph).  No counterpart
 * in the reference, which runs one simulation with one velocity kernel. */
dvm_status dvm_batch_set_fmm(dvm_batch_t batch, int64_t min_blobs, int expansion_order,
                             double theta);

#ifdef __cplusplus
}
#endif
#endif /* DVM_HIP_H */
