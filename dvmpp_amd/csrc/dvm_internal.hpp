// Internal declarations shared by the libdvm_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/dvm_hip.h"

struct dvm_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    std::string err;
    // Growable scratch arena: at most one arena per context, reused by every
    // call on the context's stream (calls on one stream run in order).
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // While set (dvm_bs_self_noalloc), ws_get hands out this fixed region
    // instead and never allocates; ws_fixed_partials caps the chunk partials.
    void* ws_fixed = nullptr;
    size_t ws_fixed_bytes = 0;
    size_t ws_fixed_partials = 0;
    // Pinned host staging for small device->host reads.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // F32C pair-sum variant choice: DVM_TUNE_TABLE or DVM_TUNE_MEASURE; a
    // forced variant (tile, block, share, occupancy; tile 0 = none); the
    // variant of the last call; measured choices keyed by (kind, log2 of
    // the target count, log2 of the source count) -> (variant, occupancy).
    int tuning = DVM_TUNE_TABLE;
    int f32c_forced[4] = {0, 0, 0, 0};
    int f32c_last[4] = {0, 0, 0, 0};
    std::map<std::tuple<int, int, int>, std::pair<int, int>> f32c_tune;
    // Last merge: candidate blobs, and speculative sequences recomputed.
    int64_t merge_stats[2] = {0, 0};
    // DVM_FMM parameters (the reference FMMKernel's defaults) and the leaf
    // level of the last DVM_FMM sum (0: direct).
    int fmm_order = 10;
    double fmm_theta = 0.5;
    int fmm_last_level = 0;
    // Kernel timing (dvm_prof_*): pending event pairs and totals per kernel.
    bool prof = false;
    struct ProfEntry {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
        double total_ms = 0.0;
        int64_t count = 0;
    };
    std::map<std::string, ProfEntry> prof_map;
    std::vector<hipEvent_t> event_pool;
};

namespace dvmi {

constexpr double kRpi2 = 1.0 / (2.0 * 3.14159265358979323846);  // src/BaseTypes.hpp:21

struct Error {
    dvm_status st;
    std::string msg;
};

[[noreturn]] inline void fail(dvm_status s, const std::string& m) { throw Error{s, m}; }

#define DVM_HIP_CHECK(expr)                                                            \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::dvmi::fail(e_ == hipErrorOutOfMemory ? DVM_E_OOM : DVM_E_HIP,            \
                         std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define DVM_LAUNCH_CHECK() DVM_HIP_CHECK(hipGetLastError())

template <class F>
dvm_status guard(dvm_ctx_t ctx, F&& f)
{
    try {
        f();
        return DVM_OK;
    } catch (const Error& e) {
        if (ctx) ctx->err = e.msg;
        return e.st;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "host out of memory";
        return DVM_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return DVM_E_STATE;
    }
}

inline void require(bool cond, dvm_status s, const char* msg)
{
    if (!cond) fail(s, msg);
}

// Brackets the launches in its scope with timing events when profiling is on.
class ProfScope {
public:
    ProfScope(dvm_ctx_t ctx, const char* name);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

private:
    dvm_ctx_t ctx_;
    const char* name_;
    hipEvent_t start_ = nullptr;
};

// Scratch arena of at least `bytes`; contents are not preserved on growth.
void* ws_get(dvm_ctx_t ctx, size_t bytes);
// Pinned host buffer of at least `bytes`.
void* pinned_get(dvm_ctx_t ctx, size_t bytes);

// Carves aligned sub-buffers out of one scratch allocation.
struct Carver {
    char* base;
    size_t off = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <class T>
    T* take(size_t n)
    {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += n * sizeof(T);
        return p;
    }
    static size_t need(std::initializer_list<size_t> bytes)
    {
        size_t s = 0;
        for (size_t b : bytes) s = ((s + 255) & ~size_t(255)) + b;
        return s + 256;
    }
};

inline unsigned cdiv(int64_t a, int64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Per-panel constants read by the sheet-on-blob and release kernels, computed
// on the host with the C library's libm at set-up (the values the reference
// recomputes per pair in src/VortexSheet.cpp:225-251 and :325).
struct PanelConst {
    double xc, zc, ds, cs, sn, c2, s2, sig;  // cos/sin(theta), cos/sin(-2 theta), ds^cutoff_exp
};

// Device-resident vortex sheet (one body).
// Bounding box of a body contour for the crossing test's cull (outside_box):
// zlo = min z, zhi = max z, xhi = max x + margin, xlo = min x - margin when
// the contour is closed (last point == first), else -inf.
struct BodyBox {
    double zlo, zhi, xlo, xhi;
};
BodyBox body_box(int npts, const double* px, const double* pz);

// Edges per horizontal slab of a body's z range, for the crossing test: the
// range [z0, z0 + n/inv_h) is cut into n slabs (n = Nl), slab s lists (in
// edge order) every edge whose z extent meets slabs s-1..s+1 of its ends, so
// every edge that can straddle a z in slab s is listed there.  n = 0: none
// (degenerate or non-finite body, or more entries than a caller can hold).
struct EdgeSlabs {
    const int* off = nullptr;   // n + 1 offsets into edge
    const int* edge = nullptr;
    int n = 0;
    int pad = 0;
    double z0 = 0.0, inv_h = 0.0;
};
// Host build: off (n + 1) and edge lists; returns n (0 for none).
int edge_slabs_build(int npts, const double* px, const double* pz, std::vector<int>& off,
                     std::vector<int>& edge, double& z0, double& inv_h);

struct SheetDev {
    int nl = 0;
    double rho = 1, nu = 0, dt = 0, ux = 0, uz = 0, cutoff_exp = 0;
    int max_npv = 1;
    double aref = 0;
    BodyBox box{};                                 // crossing-test cull box
    EdgeSlabs slabs;                               // crossing-test edge lists (device)
    std::vector<int> hslab_off, hslab_edge;        // the same on the host
    std::vector<double> hx, hz;                    // body points (host copy)
    std::vector<double> hxc, hzc, htheta, hds, henx, henz, hetx, hetz;
    char* mem = nullptr;                           // one device allocation
    double *x = nullptr, *z = nullptr;             // nl+1 body points
    double *enx = nullptr, *enz = nullptr, *etx = nullptr, *etz = nullptr, *ds = nullptr;
    double *xc = nullptr, *zc = nullptr, *theta = nullptr;
    PanelConst* panel = nullptr;                   // nl
    double* A = nullptr;                           // (nl+1) x nl row-major
    double* P = nullptr;                           // nl x (nl+1) row-major: (A^T A)^-1 A^T
    double* gamma = nullptr;                       // nl
    double* b = nullptr;                           // nl+1 right-hand side of the last solve
    double* uw = nullptr;                          // 2*nl collocation velocities
    double* scal = nullptr;                        // [0] total circulation, [1..4] fx fz cd cl
    double* red = nullptr;                         // reduction partials
    int* cnt = nullptr;                            // release counts per panel
    int64_t* off = nullptr;                        // release offsets (nl+1), off[nl] = Nrv
    double* cnew = nullptr;                        // circulation per nascent blob per panel
    int* flag = nullptr;                           // [0] NaN-count flag
    // Far-field form of the sheet's velocity (fast sums only): the panels of
    // sheet_pair are straight vortex segments, so outside kSheetFar * mr of
    // (mx, mz) their velocity is a Laurent series with kSheetTerms moments,
    // mA = sum_j integrals of -gamma(s) ((zeta(s) - c) / mr)^k ds over the
    // panel, from the geometric integrals mI0 (weight 1 - s/ds) and mI1
    // (weight s/ds) per panel (complex, interleaved re/im).
    double mx = 0, mz = 0, mr = 1;
    double *mI0 = nullptr, *mI1 = nullptr, *mA = nullptr;
    // Panel groups (sheet_groups) for the batched fast form, which copies
    // them: group centres and per-panel integrals about them.
    int ng = 0;
    double4* gcen = nullptr;
    double *gI0 = nullptr, *gI1 = nullptr;
    // host staging of the set-up's uploads, alive until the caller has
    // synchronised the stream sheet_init ran on
    std::vector<double> hstage, hgcen;
    std::vector<PanelConst> hpanel;
};

constexpr int kSheetTerms = 40;
constexpr double kSheetFar = 2.0;

// Panel groups of the batched fast sheet sum: kSheetGroup consecutive panels,
// with centre c (mean of the segments' midpoints) and radius r (largest
// distance from c to a segment end).  Beyond kGroupFar * r of c a group's
// velocity is the Laurent series above with the group's own moments.
// sheet_groups fills cen (cx, cz, r, 0 per group; r = inf for a degenerate
// group, which is always summed panel by panel).
constexpr int kSheetGroup = 16;
// The groups' series serve the FP32C batch only (1e-5 velocities): they are
// taken from kGroupFar * r on, where 40 terms leave at most 1.5^-40 (1e-7)
// of the group's leading term; the body-wide series keeps kSheetFar.
constexpr double kGroupFar = 1.5;
void sheet_groups(const SheetDev& sh, std::vector<double>& cen);
// Geometric integrals (the layout of mI0, mI1) of the nl panels pc on stream
// st: about (cx, cz) with radius r, or, when gcen is given, about each
// panel's group centre gcen[i / kSheetGroup].
void panel_moments(hipStream_t st, int nl, const PanelConst* pc, double cx, double cz, double r,
                   const double4* gcen, double* i0, double* i1);

// Sets up a body's sheet (geometry, influence matrix, least-squares operator,
// far-field moments) on `st` (default: the context's stream).  With
// sync = false it returns with the work queued: the caller synchronises `st`
// before using or freeing the sheet.
void sheet_init(dvm_ctx_t ctx, SheetDev& sh, int npts, const double* px, const double* pz,
                const dvm_sheet_params* p, hipStream_t st = nullptr, bool sync = true);
void sheet_free(SheetDev& sh);
// solvevortexsheet: collocation sum + RHS + gamma = P b.
// have_uw: the collocation sums are already in sh.uw.
void sheet_solve(dvm_ctx_t ctx, SheetDev& sh, int prec, int64_t n, const double* bx,
                 const double* bz, const double* bc, const double* bs, bool have_uw = false);
// vortexsheetbc
void sheet_on_blobs(dvm_ctx_t ctx, const SheetDev& sh, int prec, int64_t n, const double* bx,
                    const double* bz, double* uvs, double* wvs);
void sheet_loads(dvm_ctx_t ctx, SheetDev& sh, double ur);
// probe_velocities in the reference's (quirk) form.
void sheet_probe_quirk(dvm_ctx_t ctx, const SheetDev& sh, int prec, int64_t n, const double* bx,
                       const double* bz, const double* bc, const double* bs, int64_t np,
                       const double* px, const double* pz, double* u, double* w, double* uvs,
                       double* wvs);
// Release counts and offsets; returns nothing (results in sh.cnt/off/flag).
void sheet_release_count(dvm_ctx_t ctx, SheetDev& sh);
// Writes nrv nascent blobs at index `base` of the blob arrays, using rrw[k].
void sheet_release_emit(dvm_ctx_t ctx, const SheetDev& sh, int64_t nrv, const double* rrw,
                        uint32_t nold, int64_t base, double* x, double* z, double* c, double* s,
                        double* u, double* w, double* uvs, double* wvs, uint32_t* id);
// Crossing: flags[i] = inside_body(x_i, z_i).
void sheet_inside_flags(dvm_ctx_t ctx, const SheetDev& sh, int64_t n, const double* x,
                        const double* z, int* flags);
void sheet_reflect(dvm_ctx_t ctx, const SheetDev& sh, int64_t n, const int* flags, double* x,
                   double* z);
// Deterministic sum of a device vector into *out (device).
void device_sum(dvm_ctx_t ctx, int64_t n, const double* v, double* partials, double* out);

// Blob SoA on the device (VortexBlobs fields, src/VortexBlobs.hpp:24-32).
struct BlobBuf {
    double *x = nullptr, *z = nullptr, *c = nullptr, *s = nullptr;
    double *u = nullptr, *w = nullptr, *uvs = nullptr, *wvs = nullptr;
    uint32_t* id = nullptr;
};

// Remesh of n blobs (BackgroundMesh specification, see remesh.hip).  Phase 1
// bins and gathers onto the mesh and returns the output count (host value;
// synchronizes); phase 2 writes the output into `out` (capacity >= count).
struct RemeshPlan {
    int64_t n_in = 0, nx = 0, nz = 0, n_nodes_out = 0, n_kept = 0;
    int64_t win_i0 = 0, win_k0 = 0, win_w = 0, win_h = 0;  // node window worked on
    char* mem = nullptr;  // grid values, flags, scans: the context's work arena
};
int64_t remesh_count(dvm_ctx_t ctx, const dvm_remesh_params& rp, int64_t n, const BlobBuf& in,
                     RemeshPlan& plan);
void remesh_finish(dvm_ctx_t ctx, const dvm_remesh_params& rp, const BlobBuf& in,
                   RemeshPlan& plan, BlobBuf& out);

// VortexBlobs::merge_blobs (see merge.hip): merges in place in `b`, and when
// anything merged writes the survivors compacted into `out` (IDs 0..m-1);
// returns the number of merges (host value; synchronizes).  `scratch` holds
// merge_scratch_bytes(n) bytes of device memory.
size_t merge_scratch_bytes(int64_t n);
int64_t merge_run(dvm_ctx_t ctx, double threshold, int64_t n, const BlobBuf& b, const BlobBuf& out,
                  void* scratch);

// The FMM expansion order used for a requested one (0: unsupported).
int fmm_terms(int order);

// DVM_FMM blob-on-blob sum (fmm.hip); false when it declined (the caller then
// runs the direct FP64 sum).
// With nt > 0 it also writes the sums at targets (tx, tz) into (tu, tw),
// every source counted (the unguarded direct target sum).
bool fmm_self(dvm_ctx_t ctx, int64_t n, const double* x, const double* z, const double* c,
              const double* s, double* u, double* w, int64_t nt = 0, const double* tx = nullptr,
              const double* tz = nullptr, double* tu = nullptr, double* tw = nullptr);

// Biot-Savart entry points shared with the sheet and simulation code.
void bs_self_launch(dvm_ctx_t ctx, int prec, int64_t n, const double* x, const double* z,
                    const double* c, const double* s, double* u, double* w);
void bs_targets_launch(dvm_ctx_t ctx, int prec, int64_t ns, const double* sx, const double* sz,
                       const double* sc, const double* ss, int64_t nt, const double* tx,
                       const double* tz, int guard, double r2_skip, double* u, double* w);
void f32c_set_variant(dvm_ctx_t ctx, int tile, int block, int share, int occupancy);

}  // namespace dvmi

struct dvm_sheet_s {
    dvm_ctx_t ctx;
    dvmi::SheetDev sh;
};
