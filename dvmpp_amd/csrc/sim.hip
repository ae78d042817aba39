// Device-resident DVM simulation: DVMBase's time loop (src/DVMBase.cpp:67-209)
// over blob SoA buffers that live in HBM for the whole run.
//
// Host work per step is orchestration only, plus — in DVM_RNG_REFERENCE mode —
// the reference's serial random stream: the host draws from std::mt19937 in
// the reference's order and converts the draws to displacement increments
// with the C library's libm (so that diffusion and release positions match the
// reference bit for bit); the increments are generated while the device runs
// the convection stage and are applied by device kernels.
#include "dvm_device.hpp"
#include "dvm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <random>
#include <sstream>
#include <string>

namespace dvmi {

// ----------------------------------------------------------------- kernels

// src/DVMBase.cpp:179-180: x += ((u + uvs) + Ux) dt
__global__ void advect_euler_kernel(int64_t n, double ux, double uz, double dt,
                                    const double* __restrict__ u, const double* __restrict__ w,
                                    const double* __restrict__ uvs, const double* __restrict__ wvs,
                                    double* __restrict__ x, double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += (u[i] + uvs[i] + ux) * dt;
    z[i] += (w[i] + wvs[i] + uz) * dt;
}

// src/DVMBase.cpp:201-205: q = a q + ((u + uvs) + U) dt;  x += b q
__global__ void advect_rk3_kernel(int64_t n, double a, double b, double ux, double uz, double dt,
                                  const double* __restrict__ u, const double* __restrict__ w,
                                  const double* __restrict__ uvs, const double* __restrict__ wvs,
                                  double* __restrict__ q1, double* __restrict__ q2,
                                  double* __restrict__ x, double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double n1 = a * q1[i] + (u[i] + uvs[i] + ux) * dt;
    const double n2 = a * q2[i] + (w[i] + wvs[i] + uz) * dt;
    q1[i] = n1;
    q2[i] = n2;
    x[i] += b * n1;
    z[i] += b * n2;
}

// Diffusion with host-generated increments (tape[2i], tape[2i+1]).
__global__ void diffuse_tape_kernel(int64_t n, const double* __restrict__ tape,
                                    double* __restrict__ x, double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 d = reinterpret_cast<const double2*>(tape)[i];
    x[i] += d.x;
    z[i] += d.y;
}

// Diffusion with device random numbers: the reference's Box-Muller step
// (src/VortexBlobs.cpp:253-264) on Philox draws keyed by (seed, step, blob).
__global__ void diffuse_philox_kernel(int64_t n, uint32_t seed, uint32_t step, double four_nu_dt,
                                      double* __restrict__ x, double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 r = philox4x32(make_uint4((uint32_t)i, (uint32_t)(i >> 32), step, 0u),
                               make_uint2(seed, 0x5eed0001u));
    const double R1 = u01(r.x, r.y);
    const double R2 = u01(r.z, r.w);
    const double rrw = sqrt(four_nu_dt * log(1.0 / R1));
    const double th = 2.0 * 3.14159265358979323846 * R2;
    x[i] += rrw * cos(th);
    z[i] += rrw * sin(th);
}

// Release offsets |sqrt(4 nu dt ln(1/R))| from Philox draws keyed by (seed, step, k).
__global__ void release_philox_kernel(int64_t nrv, uint32_t seed, uint32_t step,
                                      double four_nu_dt, double* __restrict__ rrw)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrv) return;
    const uint4 r = philox4x32(make_uint4((uint32_t)k, (uint32_t)(k >> 32), step, 1u),
                               make_uint2(seed, 0x5eed0002u));
    rrw[k] = fabs(sqrt(four_nu_dt * log(1.0 / u01(r.x, r.y))));
}

// Stable compaction of the blobs whose flag is 0 (src/VortexSheet.cpp:407-445);
// IDs renumbered 0..M-1.  pos = exclusive prefix of keep flags.
__global__ void compact_kernel(int64_t n, const int* __restrict__ inside,
                               const int* __restrict__ pos, BlobBuf in, BlobBuf out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || inside[i]) return;
    const int64_t o = pos[i];
    out.x[o] = in.x[i]; out.z[o] = in.z[i]; out.c[o] = in.c[i]; out.s[o] = in.s[i];
    out.u[o] = in.u[i]; out.w[o] = in.w[i]; out.uvs[o] = in.uvs[i]; out.wvs[o] = in.wvs[i];
    out.id[o] = (uint32_t)o;
}

__global__ void keep_flags_kernel(int64_t n, const int* __restrict__ inside, int* __restrict__ keep)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keep[i] = inside[i] ? 0 : 1;
}

}  // namespace dvmi

using namespace dvmi;

struct dvm_sim_s {
    dvm_ctx_t ctx = nullptr;
    SheetDev sh;
    dvm_sim_params p{};
    double ur = 0.0;
    int step = 0;
    int64_t n = 0;
    int64_t cap = 0;
    char* mem = nullptr;    // blob buffers (two sets), RK3 stores, flags, tape
    BlobBuf b, alt;
    double *q1 = nullptr, *q2 = nullptr, *tape = nullptr;
    int *flags = nullptr, *keep = nullptr, *pos = nullptr;
    double* loads = nullptr;  // last step's fx fz cd cl (device, inside sh.scal)
    int64_t last_merged = 0;  // merges of the last merge sub-step
    // reference random stream (src/Random.hpp:13-15)
    std::uniform_real_distribution<> dist;
    std::mt19937 re;
    uint64_t draws = 0;
    uint32_t philox_seed = 0;
    std::vector<double> htape;
    double* pinned_tape = nullptr;
    size_t pinned_cap = 0;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;

    double rand()
    {
        ++draws;
        return dist(re);
    }
};

namespace {

// Snapshot layout (host byte order): header, mt19937 state as text, body
// points x then z, gamma, ids, then x z circ sigma u w uvs wvs.
constexpr char kSnapMagic[8] = {'D', 'V', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kSnapVersion = 1;
struct SnapHeader {
    char magic[8];
    uint32_t version;
    int32_t npts;
    int32_t nl;
    int32_t step;
    int64_t n;
    uint64_t draws;
    uint32_t philox_seed;
    uint32_t rng_bytes;
    dvm_sim_params p;
    double loads[4];
};

std::string rng_text(const std::mt19937& re)
{
    std::ostringstream o;
    o << re;
    return o.str();
}

size_t snap_bytes(dvm_sim_t s, size_t rng_len)
{
    return sizeof(SnapHeader) + rng_len + sizeof(double) * 2 * s->sh.hx.size() +
           sizeof(double) * s->sh.nl + (sizeof(uint32_t) + 8 * sizeof(double)) * (size_t)s->n;
}

bool params_equal(const dvm_sim_params& a, const dvm_sim_params& b)
{
    const dvm_sheet_params &x = a.sheet, &y = b.sheet;
    const dvm_remesh_params &r = a.remesh, &q = b.remesh;
    return x.rho == y.rho && x.nu == y.nu && x.dt == y.dt && x.ux == y.ux && x.uz == y.uz &&
           x.cutoff_exp == y.cutoff_exp && x.max_npv == y.max_npv && a.scheme == b.scheme &&
           a.crossing == b.crossing && a.seed == b.seed && a.prec == b.prec &&
           a.rng_mode == b.rng_mode && r.frequency == q.frequency && r.kernel == q.kernel &&
           r.xmin == q.xmin && r.xmax == q.xmax && r.zmin == q.zmin && r.zmax == q.zmax &&
           r.dx == q.dx && r.dz == q.dz && r.sigma == q.sigma &&
           r.circ_threshold == q.circ_threshold && a.merge_threshold == b.merge_threshold &&
           a.merge_frequency == b.merge_frequency;
}

void carve_blobs(Carver& cv, BlobBuf& b, int64_t cap)
{
    b.x = cv.take<double>(cap); b.z = cv.take<double>(cap);
    b.c = cv.take<double>(cap); b.s = cv.take<double>(cap);
    b.u = cv.take<double>(cap); b.w = cv.take<double>(cap);
    b.uvs = cv.take<double>(cap); b.wvs = cv.take<double>(cap);
    b.id = cv.take<uint32_t>(cap);
}

void copy_blobs(hipStream_t st, const BlobBuf& from, BlobBuf& to, int64_t n, int64_t src_off = 0,
                int64_t dst_off = 0)
{
    if (n <= 0) return;
    const size_t b = sizeof(double) * n;
    const double* fs[8] = {from.x, from.z, from.c, from.s, from.u, from.w, from.uvs, from.wvs};
    double* ts[8] = {to.x, to.z, to.c, to.s, to.u, to.w, to.uvs, to.wvs};
    for (int k = 0; k < 8; ++k)
        DVM_HIP_CHECK(hipMemcpyAsync(ts[k] + dst_off, fs[k] + src_off, b, hipMemcpyDeviceToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(to.id + dst_off, from.id + src_off, sizeof(uint32_t) * n,
                                 hipMemcpyDeviceToDevice, st));
}

// Grows every per-blob buffer to hold at least `need` blobs, keeping the live ones.
void ensure_cap(dvm_sim_t s, int64_t need)
{
    if (need <= s->cap) return;
    int64_t cap = s->cap * 2;
    if (cap < need) cap = need;
    if (cap < 4096) cap = 4096;
    cap = (cap + 255) & ~int64_t(255);
    const size_t d = sizeof(double) * cap, u = sizeof(uint32_t) * cap, i = sizeof(int) * cap;
    const size_t need_b = Carver::need({d, d, d, d, d, d, d, d, u, d, d, d, d, d, d, d, d, u,
                                        d, d, 2 * d, i, i, i});
    char* mem = nullptr;
    DVM_HIP_CHECK(hipMalloc(&mem, need_b));
    Carver cv(mem);
    BlobBuf nb, na;
    carve_blobs(cv, nb, cap);
    carve_blobs(cv, na, cap);
    double* q1 = cv.take<double>(cap);
    double* q2 = cv.take<double>(cap);
    double* tape = cv.take<double>(2 * cap);
    int* flags = cv.take<int>(cap);
    int* keep = cv.take<int>(cap);
    int* pos = cv.take<int>(cap);
    copy_blobs(s->ctx->stream, s->b, nb, s->n);
    DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    if (s->mem) DVM_HIP_CHECK(hipFree(s->mem));
    s->mem = mem;
    s->cap = cap;
    s->b = nb;
    s->alt = na;
    s->q1 = q1; s->q2 = q2; s->tape = tape;
    s->flags = flags; s->keep = keep; s->pos = pos;
    size_t tmp = 0;
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (int*)nullptr, (int*)nullptr,
                                                   (int)cap, s->ctx->stream));
    if (tmp > s->scan_tmp_bytes) {
        if (s->scan_tmp) DVM_HIP_CHECK(hipFree(s->scan_tmp));
        DVM_HIP_CHECK(hipMalloc(&s->scan_tmp, tmp));
        s->scan_tmp_bytes = tmp;
    }
}

double* pinned_tape(dvm_sim_t s, size_t count)
{
    if (count > s->pinned_cap) {
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        if (s->pinned_tape) DVM_HIP_CHECK(hipHostFree(s->pinned_tape));
        size_t c = count * 2 > 65536 ? count * 2 : 65536;
        DVM_HIP_CHECK(hipHostMalloc((void**)&s->pinned_tape, sizeof(double) * c, hipHostMallocDefault));
        s->pinned_cap = c;
    }
    return s->pinned_tape;
}

// velocities at the current positions: biotsavart, solvevortexsheet, vortexsheetbc
void rhs(dvm_sim_t s)
{
    const int64_t n = s->n;
    bool have_uw = false;
    if (n > 0 && s->p.prec == DVM_FMM) {
        // one tree for the blob velocities and the collocation sums
        have_uw = fmm_self(s->ctx, n, s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w, s->sh.nl,
                           s->sh.xc, s->sh.zc, s->sh.uw, s->sh.uw + s->sh.nl);
        if (!have_uw) bs_self_launch(s->ctx, DVM_F64, n, s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w);
    } else if (n > 0) {
        bs_self_launch(s->ctx, s->p.prec, n, s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w);
    }
    sheet_solve(s->ctx, s->sh, s->p.prec, n, s->b.x, s->b.z, s->b.c, s->b.s, have_uw);
    sheet_on_blobs(s->ctx, s->sh, s->p.prec, n, s->b.x, s->b.z, s->b.uvs, s->b.wvs);
}

void convect(dvm_sim_t s)
{
    const int64_t n = s->n;
    hipStream_t st = s->ctx->stream;
    const double dt = s->p.sheet.dt, ux = s->p.sheet.ux, uz = s->p.sheet.uz;
    if (s->p.scheme == DVM_EULER) {
        rhs(s);
        if (n > 0) {
            advect_euler_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, ux, uz, dt, s->b.u, s->b.w,
                                                              s->b.uvs, s->b.wvs, s->b.x, s->b.z);
            DVM_LAUNCH_CHECK();
        }
    } else {
        const double a[3] = {0, -17. / 32, -32. / 27};
        const double bb[3] = {1. / 4, 8. / 9, 3. / 4};
        if (n > 0) {
            DVM_HIP_CHECK(hipMemsetAsync(s->q1, 0, sizeof(double) * n, st));
            DVM_HIP_CHECK(hipMemsetAsync(s->q2, 0, sizeof(double) * n, st));
        }
        for (int k = 0; k < 3; ++k) {
            rhs(s);
            if (n > 0) {
                advect_rk3_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, a[k], bb[k], ux, uz, dt, s->b.u,
                                                                s->b.w, s->b.uvs, s->b.wvs, s->q1,
                                                                s->q2, s->b.x, s->b.z);
                DVM_LAUNCH_CHECK();
            }
        }
    }
}

void do_remesh(dvm_sim_t s)
{
    const dvm_remesh_params& rp = s->p.remesh;
    if (!(rp.frequency > 0 && s->step > 0 && s->step % rp.frequency == 0)) return;
    RemeshPlan plan;
    const int64_t m = remesh_count(s->ctx, rp, s->n, s->b, plan);
    ensure_cap(s, m > s->n ? m : s->n);
    remesh_finish(s->ctx, rp, s->b, plan, s->alt);
    std::swap(s->b, s->alt);
    s->n = m;
}

// src/VortexBlobs.cpp:251-265
void diffuse(dvm_sim_t s)
{
    const int64_t n = s->n;
    if (n == 0) return;
    hipStream_t st = s->ctx->stream;
    const double nu = s->p.sheet.nu, dt = s->p.sheet.dt;
    if (s->p.rng_mode == DVM_RNG_DEVICE) {
        diffuse_philox_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, s->philox_seed, (uint32_t)s->step,
                                                            4.0 * nu * dt, s->b.x, s->b.z);
        DVM_LAUNCH_CHECK();
        return;
    }
    double* h = pinned_tape(s, 2 * n);
    for (int64_t i = 0; i < n; i++) {
        const double R1 = s->rand();
        const double R2 = s->rand();
        const double rrw = std::sqrt(4.0 * nu * dt * std::log(1.0 / R1));
        const double thetarw = 2.0 * 3.14159265358979323846 * R2;
        h[2 * i] = rrw * std::cos(thetarw);
        h[2 * i + 1] = rrw * std::sin(thetarw);
    }
    DVM_HIP_CHECK(hipMemcpyAsync(s->tape, h, sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
    diffuse_tape_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, s->tape, s->b.x, s->b.z);
    DVM_LAUNCH_CHECK();
    // the pinned tape is reused next step: wait for the copy
    DVM_HIP_CHECK(hipStreamSynchronize(st));
}

// src/VortexSheet.cpp:275-333 + src/VortexBlobs.cpp:211-227
void release(dvm_sim_t s)
{
    hipStream_t st = s->ctx->stream;
    sheet_release_count(s->ctx, s->sh);
    int flag = 0;
    int64_t nrv = 0;
    DVM_HIP_CHECK(hipMemcpyAsync(&flag, s->sh.flag, sizeof(int), hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipMemcpyAsync(&nrv, s->sh.off + s->sh.nl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    if (flag & 1) fail(DVM_E_NAN, "release_nascent_vortices_rw: non-finite PanelNewVort");
    if (flag & 2) fail(DVM_E_SIZE, "release_nascent_vortices_rw: panel count too large");
    const int64_t nold = s->n;
    ensure_cap(s, nold + nrv);
    const double nu = s->p.sheet.nu, dt = s->p.sheet.dt;
    double* rrw = s->tape;  // free after diffusion
    if (s->p.rng_mode == DVM_RNG_DEVICE) {
        release_philox_kernel<<<cdiv(nrv, 256), 256, 0, st>>>(nrv, s->philox_seed,
                                                              (uint32_t)s->step, 4.0 * nu * dt, rrw);
        DVM_LAUNCH_CHECK();
    } else {
        double* h = pinned_tape(s, nrv);
        for (int64_t k = 0; k < nrv; ++k) {
            const double R = s->rand();
            h[k] = std::fabs(std::sqrt(4.0 * nu * dt * std::log(1.0 / R)));
        }
        DVM_HIP_CHECK(hipMemcpyAsync(rrw, h, sizeof(double) * nrv, hipMemcpyHostToDevice, st));
    }
    sheet_release_emit(s->ctx, s->sh, nrv, rrw, (uint32_t)nold, nold, s->b.x, s->b.z, s->b.c,
                       s->b.s, s->b.u, s->b.w, s->b.uvs, s->b.wvs, s->b.id);
    s->n = nold + nrv;
    // the pinned tape of the reference stream is refilled next step: wait for
    // its copy (the device stream has no host tape)
    if (s->p.rng_mode != DVM_RNG_DEVICE) DVM_HIP_CHECK(hipStreamSynchronize(st));
}

// src/DVMBase.cpp:107-118
void crossing(dvm_sim_t s)
{
    const int64_t n = s->n;
    if (n == 0) return;
    hipStream_t st = s->ctx->stream;
    sheet_inside_flags(s->ctx, s->sh, n, s->b.x, s->b.z, s->flags);
    if (s->p.crossing != DVM_DELETE) {  // ABSORB falls back to REFLECT (src/DVMBase.cpp:111-114)
        sheet_reflect(s->ctx, s->sh, n, s->flags, s->b.x, s->b.z);
        return;
    }
    keep_flags_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, s->flags, s->keep);
    DVM_LAUNCH_CHECK();
    size_t tmp = s->scan_tmp_bytes;
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(s->scan_tmp, tmp, s->keep, s->pos, (int)n, st));
    int last_pos = 0, last_keep = 0;
    DVM_HIP_CHECK(hipMemcpyAsync(&last_pos, s->pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipMemcpyAsync(&last_keep, s->keep + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    const int64_t m = (int64_t)last_pos + last_keep;
    if (m == n) return;  // nothing deleted: IDs untouched (src/VortexSheet.cpp:420-422)
    compact_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, s->flags, s->pos, s->b, s->alt);
    DVM_LAUNCH_CHECK();
    std::swap(s->b, s->alt);
    s->n = m;
}

void loads(dvm_sim_t s) { sheet_loads(s->ctx, s->sh, s->ur); }

// src/DVMBase.cpp:126-132
void merge(dvm_sim_t s)
{
    s->last_merged = 0;
    const double thr = s->p.merge_threshold;
    const int f = s->p.merge_frequency;
    if (!(f > 0 && thr > 0 && s->step % f == 0) || s->n < 2) return;
    void* scratch = ws_get(s->ctx, merge_scratch_bytes(s->n));
    const int64_t k = merge_run(s->ctx, thr, s->n, s->b, s->alt, scratch);
    if (k == 0) return;
    std::swap(s->b, s->alt);
    s->n -= k;
    s->last_merged = k;
}

void substep(dvm_sim_t s, int which)
{
    switch (which) {
    case 0:
        s->step++;
        convect(s);
        break;
    case 1: do_remesh(s); break;
    case 2: diffuse(s); break;
    case 3: release(s); break;
    case 4: crossing(s); break;
    case 5: loads(s); break;
    case 6: merge(s); break;
    default: fail(DVM_E_INVALID, "unknown sub-step");
    }
}

}  // namespace

extern "C" {

dvm_status dvm_sim_create(dvm_ctx_t ctx, int npts, const double* px, const double* pz,
                          const dvm_sim_params* params, dvm_sim_t* out)
{
    if (!ctx || !out || !params) return DVM_E_INVALID;
    *out = nullptr;
    dvm_sim_t s = new (std::nothrow) dvm_sim_s();
    if (!s) return DVM_E_OOM;
    s->ctx = ctx;
    dvm_status st = guard(ctx, [&] {
        const dvm_sim_params& p = *params;
        require(p.scheme == DVM_EULER || p.scheme == DVM_RK3, DVM_E_INVALID, "unknown scheme");
        require(p.crossing >= DVM_DELETE && p.crossing <= DVM_REFLECT, DVM_E_INVALID,
                "unknown surface crossing");
        require(p.prec == DVM_F64 || p.prec == DVM_F32C || p.prec == DVM_FMM, DVM_E_INVALID,
                "unknown precision");
        require(p.rng_mode == DVM_RNG_REFERENCE || p.rng_mode == DVM_RNG_DEVICE, DVM_E_INVALID,
                "unknown rng mode");
        if (p.remesh.frequency > 0) {
            require(p.remesh.dx > 0 && p.remesh.dz > 0 && p.remesh.xmax > p.remesh.xmin &&
                        p.remesh.zmax > p.remesh.zmin && p.remesh.sigma > 0,
                    DVM_E_INVALID, "invalid remesh parameters");
            require(p.remesh.kernel >= 0 && p.remesh.kernel <= 2, DVM_E_INVALID,
                    "unknown remesh kernel");
        }
        s->p = p;
        sheet_init(ctx, s->sh, npts, px, pz, &p.sheet);
        s->ur = std::sqrt(p.sheet.ux * p.sheet.ux + p.sheet.uz * p.sheet.uz);
        // src/Random.cpp:3-18: seeded from random_device, reseeded if seed >= 0
        std::random_device rd;
        s->re.seed(rd());
        if (p.seed >= 0) s->re.seed(p.seed);
        s->philox_seed = p.seed >= 0 ? (uint32_t)p.seed : (uint32_t)rd();
        ensure_cap(s, 4096);
    });
    if (st != DVM_OK) {
        sheet_free(s->sh);
        if (s->mem) (void)hipFree(s->mem);
        delete s;
        return st;
    }
    *out = s;
    return DVM_OK;
}

dvm_status dvm_sim_destroy(dvm_sim_t s)
{
    if (!s) return DVM_OK;
    dvm_status st = guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        sheet_free(s->sh);
        if (s->mem) DVM_HIP_CHECK(hipFree(s->mem));
        if (s->scan_tmp) DVM_HIP_CHECK(hipFree(s->scan_tmp));
        if (s->pinned_tape) DVM_HIP_CHECK(hipHostFree(s->pinned_tape));
    });
    delete s;
    return st;
}

dvm_status dvm_sim_step(dvm_sim_t s)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        for (int k = 0; k <= 6; ++k) substep(s, k);
    });
}

dvm_status dvm_sim_run(dvm_sim_t s, int steps)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        for (int j = 0; j < steps; ++j)
            for (int k = 0; k <= 6; ++k) substep(s, k);
    });
}

dvm_status dvm_sim_substep(dvm_sim_t s, int which)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] { substep(s, which); });
}

dvm_status dvm_sim_size(dvm_sim_t s, int64_t* n)
{
    if (!s || !n) return DVM_E_INVALID;
    *n = s->n;
    return DVM_OK;
}

dvm_status dvm_sim_step_index(dvm_sim_t s, int* step)
{
    if (!s || !step) return DVM_E_INVALID;
    *step = s->step;
    return DVM_OK;
}

dvm_status dvm_sim_merged(dvm_sim_t s, int64_t* merged)
{
    if (!s || !merged) return DVM_E_INVALID;
    *merged = s->last_merged;
    return DVM_OK;
}

dvm_status dvm_sim_rng_draws(dvm_sim_t s, uint64_t* draws)
{
    if (!s || !draws) return DVM_E_INVALID;
    *draws = s->draws;
    return DVM_OK;
}

dvm_status dvm_sim_get_state(dvm_sim_t s, uint32_t* id, double* x, double* z, double* circ,
                             double* sigma, double* u, double* w, double* uvs, double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        const int64_t n = s->n;
        if (n == 0) return;
        hipStream_t st = s->ctx->stream;
        const size_t b = sizeof(double) * n;
        double* outs[8] = {x, z, circ, sigma, u, w, uvs, wvs};
        const double* srcs[8] = {s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w, s->b.uvs, s->b.wvs};
        for (int k = 0; k < 8; ++k)
            if (outs[k]) DVM_HIP_CHECK(hipMemcpyAsync(outs[k], srcs[k], b, hipMemcpyDeviceToHost, st));
        if (id) DVM_HIP_CHECK(hipMemcpyAsync(id, s->b.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
    });
}

dvm_status dvm_sim_set_state(dvm_sim_t s, int64_t n, const uint32_t* id, const double* x,
                             const double* z, const double* circ, const double* sigma,
                             const double* u, const double* w, const double* uvs, const double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(n >= 0 && n < (int64_t(1) << 31), DVM_E_SIZE, "blob count out of range");
        require(n == 0 || (id && x && z && circ && sigma && u && w && uvs && wvs), DVM_E_INVALID,
                "null state array");
        ensure_cap(s, n);
        hipStream_t st = s->ctx->stream;
        const size_t b = sizeof(double) * n;
        const double* ins[8] = {x, z, circ, sigma, u, w, uvs, wvs};
        double* dsts[8] = {s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w, s->b.uvs, s->b.wvs};
        if (n > 0) {
            for (int k = 0; k < 8; ++k)
                DVM_HIP_CHECK(hipMemcpyAsync(dsts[k], ins[k], b, hipMemcpyHostToDevice, st));
            DVM_HIP_CHECK(hipMemcpyAsync(s->b.id, id, sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
        }
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        s->n = n;
    });
}

dvm_status dvm_sim_get_gamma(dvm_sim_t s, double* g)
{
    if (!s || !g) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(g, s->sh.gamma, sizeof(double) * s->sh.nl,
                                     hipMemcpyDeviceToHost, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_sim_set_gamma(dvm_sim_t s, const double* g)
{
    if (!s || !g) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(s->sh.gamma, g, sizeof(double) * s->sh.nl,
                                     hipMemcpyHostToDevice, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_sim_loads(dvm_sim_t s, double* out4)
{
    if (!s || !out4) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(out4, s->sh.scal + 1, sizeof(double) * 4,
                                     hipMemcpyDeviceToHost, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_advect_euler(dvm_ctx_t ctx, int64_t n, double ux, double uz, double dt,
                            const double* u, const double* w, const double* uvs,
                            const double* wvs, double* x, double* z)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        if (n == 0) return;
        require(u && w && uvs && wvs && x && z, DVM_E_INVALID, "null array");
        advect_euler_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, ux, uz, dt, u, w, uvs, wvs, x, z);
        DVM_LAUNCH_CHECK();
    });
}

dvm_status dvm_advect_rk3_stage(dvm_ctx_t ctx, int64_t n, double a, double b, double ux,
                                double uz, double dt, const double* u, const double* w,
                                const double* uvs, const double* wvs, double* q1, double* q2,
                                double* x, double* z)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        if (n == 0) return;
        require(u && w && uvs && wvs && q1 && q2 && x && z, DVM_E_INVALID, "null array");
        advect_rk3_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, a, b, ux, uz, dt, u, w, uvs,
                                                                 wvs, q1, q2, x, z);
        DVM_LAUNCH_CHECK();
    });
}

dvm_status dvm_diffuse(dvm_ctx_t ctx, int64_t n, const double* dxdz, double* x, double* z)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        if (n == 0) return;
        require(dxdz && x && z, DVM_E_INVALID, "null array");
        diffuse_tape_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, dxdz, x, z);
        DVM_LAUNCH_CHECK();
    });
}

dvm_status dvm_sheet_release_counts(dvm_sheet_t sheet, int64_t* nrv)
{
    if (!sheet || !nrv) return DVM_E_INVALID;
    dvm_ctx_t ctx = sheet->ctx;
    return guard(ctx, [&] {
        sheet_release_count(ctx, sheet->sh);
        int flag = 0;
        DVM_HIP_CHECK(hipMemcpyAsync(&flag, sheet->sh.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        DVM_HIP_CHECK(hipMemcpyAsync(nrv, sheet->sh.off + sheet->sh.nl, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (flag & 1) fail(DVM_E_NAN, "release_nascent_vortices_rw: non-finite PanelNewVort");
        if (flag & 2) fail(DVM_E_SIZE, "release_nascent_vortices_rw: panel count too large");
    });
}

dvm_status dvm_sheet_release_emit(dvm_sheet_t sheet, int64_t nrv, const double* rrw,
                                  uint32_t id_value, int64_t base, double* x, double* z,
                                  double* circ, double* sigma, double* u, double* w,
                                  double* uvs, double* wvs, uint32_t* id)
{
    if (!sheet) return DVM_E_INVALID;
    dvm_ctx_t ctx = sheet->ctx;
    return guard(ctx, [&] {
        require(nrv >= 0 && base >= 0, DVM_E_SIZE, "negative count");
        if (nrv == 0) return;
        require(rrw && x && z && circ && sigma && u && w && uvs && wvs && id, DVM_E_INVALID,
                "null array");
        sheet_release_emit(ctx, sheet->sh, nrv, rrw, id_value, base, x, z, circ, sigma, u, w, uvs,
                           wvs, id);
    });
}

dvm_status dvm_sheet_delete_inside(dvm_sheet_t sheet, int64_t n, double* x, double* z,
                                   double* circ, double* sigma, double* u, double* w,
                                   double* uvs, double* wvs, uint32_t* id, int64_t* n_out)
{
    if (!sheet || !n_out) return DVM_E_INVALID;
    dvm_ctx_t ctx = sheet->ctx;
    return guard(ctx, [&] {
        require(n >= 0 && n < (int64_t(1) << 31), DVM_E_SIZE, "count out of range");
        *n_out = n;
        if (n == 0) return;
        require(x && z && circ && sigma && u && w && uvs && wvs && id, DVM_E_INVALID, "null array");
        hipStream_t st = ctx->stream;
        size_t tmp = 0;
        DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (int*)nullptr, (int*)nullptr,
                                                       (int)n, st));
        const size_t d = sizeof(double) * n;
        const size_t need = Carver::need({sizeof(int) * n, sizeof(int) * n, sizeof(int) * n, tmp,
                                          d, d, d, d, d, d, d, d, sizeof(uint32_t) * n});
        Carver cv(ws_get(ctx, need));
        int* flags = cv.take<int>(n);
        int* keep = cv.take<int>(n);
        int* pos = cv.take<int>(n);
        void* t = cv.take<char>(tmp);
        BlobBuf in, out;
        in.x = x; in.z = z; in.c = circ; in.s = sigma; in.u = u; in.w = w; in.uvs = uvs;
        in.wvs = wvs; in.id = id;
        carve_blobs(cv, out, n);
        sheet_inside_flags(ctx, sheet->sh, n, x, z, flags);
        keep_flags_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, flags, keep);
        DVM_LAUNCH_CHECK();
        DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t, tmp, keep, pos, (int)n, st));
        int lp = 0, lk = 0;
        DVM_HIP_CHECK(hipMemcpyAsync(&lp, pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipMemcpyAsync(&lk, keep + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        const int64_t m = (int64_t)lp + lk;
        if (m == n) return;
        compact_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, flags, pos, in, out);
        DVM_LAUNCH_CHECK();
        copy_blobs(st, out, in, m);
        *n_out = m;
    });
}

dvm_status dvm_sim_blobs(dvm_sim_t s, double** x, double** z, double** circ, double** sigma)
{
    if (!s) return DVM_E_INVALID;
    if (x) *x = s->b.x;
    if (z) *z = s->b.z;
    if (circ) *circ = s->b.c;
    if (sigma) *sigma = s->b.s;
    return DVM_OK;
}

dvm_status dvm_sim_snapshot_size(dvm_sim_t s, size_t* bytes)
{
    if (!s || !bytes) return DVM_E_INVALID;
    return guard(s->ctx, [&] { *bytes = snap_bytes(s, rng_text(s->re).size()); });
}

dvm_status dvm_sim_snapshot(dvm_sim_t s, void* buf, size_t bytes)
{
    if (!s || !buf) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        const std::string rt = rng_text(s->re);
        require(bytes >= snap_bytes(s, rt.size()), DVM_E_SIZE, "snapshot buffer too small");
        const int64_t n = s->n;
        const int nl = s->sh.nl;
        const size_t npts = s->sh.hx.size();
        SnapHeader h{};
        std::memcpy(h.magic, kSnapMagic, 8);
        h.version = kSnapVersion;
        h.npts = (int32_t)npts;
        h.nl = nl;
        h.step = s->step;
        h.n = n;
        h.draws = s->draws;
        h.philox_seed = s->philox_seed;
        h.rng_bytes = (uint32_t)rt.size();
        h.p = s->p;
        hipStream_t st = s->ctx->stream;
        DVM_HIP_CHECK(hipMemcpyAsync(h.loads, s->sh.scal + 1, sizeof(double) * 4,
                                     hipMemcpyDeviceToHost, st));
        char* o = static_cast<char*>(buf) + sizeof(SnapHeader);
        std::memcpy(o, rt.data(), rt.size());
        o += rt.size();
        std::memcpy(o, s->sh.hx.data(), sizeof(double) * npts);
        o += sizeof(double) * npts;
        std::memcpy(o, s->sh.hz.data(), sizeof(double) * npts);
        o += sizeof(double) * npts;
        DVM_HIP_CHECK(hipMemcpyAsync(o, s->sh.gamma, sizeof(double) * nl, hipMemcpyDeviceToHost, st));
        o += sizeof(double) * nl;
        if (n > 0) {
            DVM_HIP_CHECK(hipMemcpyAsync(o, s->b.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
            o += sizeof(uint32_t) * n;
            const double* srcs[8] = {s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w, s->b.uvs, s->b.wvs};
            for (int k = 0; k < 8; ++k) {
                DVM_HIP_CHECK(hipMemcpyAsync(o, srcs[k], sizeof(double) * n, hipMemcpyDeviceToHost, st));
                o += sizeof(double) * n;
            }
        }
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        std::memcpy(buf, &h, sizeof(SnapHeader));
    });
}

dvm_status dvm_sim_restore(dvm_sim_t s, const void* buf, size_t bytes)
{
    if (!s || !buf) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(bytes >= sizeof(SnapHeader), DVM_E_SIZE, "snapshot truncated");
        SnapHeader h;
        std::memcpy(&h, buf, sizeof(SnapHeader));
        require(std::memcmp(h.magic, kSnapMagic, 8) == 0 && h.version == kSnapVersion,
                DVM_E_INVALID, "not a simulation snapshot of this version");
        const size_t npts = s->sh.hx.size();
        require(h.npts == (int32_t)npts && h.nl == s->sh.nl, DVM_E_INVALID,
                "snapshot of a different body");
        require(h.n >= 0 && h.n < (int64_t(1) << 31), DVM_E_INVALID, "snapshot blob count out of range");
        const int64_t n = h.n;
        const size_t need = sizeof(SnapHeader) + h.rng_bytes + sizeof(double) * 2 * npts +
                            sizeof(double) * h.nl + (sizeof(uint32_t) + 8 * sizeof(double)) * (size_t)n;
        require(bytes >= need, DVM_E_SIZE, "snapshot truncated");
        require(params_equal(h.p, s->p), DVM_E_INVALID, "snapshot of different simulation parameters");
        const char* o = static_cast<const char*>(buf) + sizeof(SnapHeader);
        std::mt19937 re;
        {
            std::istringstream in(std::string(o, h.rng_bytes));
            in >> re;
            require(!in.fail(), DVM_E_INVALID, "corrupt random-stream state in snapshot");
        }
        o += h.rng_bytes;
        require(std::memcmp(o, s->sh.hx.data(), sizeof(double) * npts) == 0 &&
                    std::memcmp(o + sizeof(double) * npts, s->sh.hz.data(), sizeof(double) * npts) == 0,
                DVM_E_INVALID, "snapshot of a different body");
        o += sizeof(double) * 2 * npts;
        ensure_cap(s, n);
        hipStream_t st = s->ctx->stream;
        DVM_HIP_CHECK(hipMemcpyAsync(s->sh.gamma, o, sizeof(double) * h.nl, hipMemcpyHostToDevice, st));
        o += sizeof(double) * h.nl;
        DVM_HIP_CHECK(hipMemcpyAsync(s->sh.scal + 1, h.loads, sizeof(double) * 4, hipMemcpyHostToDevice, st));
        if (n > 0) {
            DVM_HIP_CHECK(hipMemcpyAsync(s->b.id, o, sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
            o += sizeof(uint32_t) * n;
            double* dsts[8] = {s->b.x, s->b.z, s->b.c, s->b.s, s->b.u, s->b.w, s->b.uvs, s->b.wvs};
            for (int k = 0; k < 8; ++k) {
                DVM_HIP_CHECK(hipMemcpyAsync(dsts[k], o, sizeof(double) * n, hipMemcpyHostToDevice, st));
                o += sizeof(double) * n;
            }
        }
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        s->n = n;
        s->step = h.step;
        s->draws = h.draws;
        s->philox_seed = h.philox_seed;
        s->re = re;
        s->dist.reset();
    });
}

dvm_status dvm_sim_copy_out(dvm_sim_t s, int64_t cap, double* x, double* z, double* circ,
                            double* gamma, double* loads4, int64_t* n)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        const int64_t m = cap < 0 ? 0 : (s->n < cap ? s->n : cap);
        hipStream_t st = s->ctx->stream;
        const size_t b = sizeof(double) * m;
        if (m > 0) {
            if (x) DVM_HIP_CHECK(hipMemcpyAsync(x, s->b.x, b, hipMemcpyDeviceToHost, st));
            if (z) DVM_HIP_CHECK(hipMemcpyAsync(z, s->b.z, b, hipMemcpyDeviceToHost, st));
            if (circ) DVM_HIP_CHECK(hipMemcpyAsync(circ, s->b.c, b, hipMemcpyDeviceToHost, st));
        }
        if (gamma)
            DVM_HIP_CHECK(hipMemcpyAsync(gamma, s->sh.gamma, sizeof(double) * s->sh.nl,
                                         hipMemcpyDeviceToHost, st));
        if (loads4)
            DVM_HIP_CHECK(hipMemcpyAsync(loads4, s->sh.scal + 1, sizeof(double) * 4,
                                         hipMemcpyDeviceToHost, st));
        if (n) *n = s->n;
    });
}

dvm_status dvm_sim_probe_quirk(dvm_sim_t s, int prec, int64_t np, const double* px,
                               const double* pz, double* u, double* w, double* uvs, double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(prec == DVM_F64 || prec == DVM_F32C || prec == DVM_FMM, DVM_E_INVALID,
                "unknown precision");
        require(np >= 0, DVM_E_SIZE, "negative probe count");
        require(np == 0 || (px && pz && u && w && uvs && wvs), DVM_E_INVALID, "null probe array");
        sheet_probe_quirk(s->ctx, s->sh, prec, s->n, s->b.x, s->b.z, s->b.c, s->b.s, np, px, pz, u, w,
                          uvs, wvs);
    });
}

}  // extern "C""
