// Remesh (BackgroundMesh::remesh).  The reference's BackgroundMesh source is
// missing (src/DVMBase.hpp:4,51; src/DVMBase.cpp:160-165); this implements the
// specification this build states, with the interpolation kernels of
// tests/unit/test_BackgroundMesh.cpp and the options of src/XmlHandler.cpp:192-201:
//
//   node (i,k) at (xmin + i dx, zmin + k dz), nx = round((xmax-xmin)/dx) + 1;
//   a blob's cell is c = floor((x - xmin)/dx) (same in z); its stencil nodes
//   are c-s+1 .. c+s (s = 2 for M4', 3 for M6 and Gaussian); a blob whose
//   stencil leaves the mesh is kept unchanged;
//   node value = sum over the stencil cells (z-major, ascending) of the blobs
//   of each cell in index order of (circ * wx) * wz (Gaussian weights
//   normalised per axis);
//   output = nodes with |circ| > circ_threshold in linear order k nx + i with
//   sigma = remesh sigma, then kept blobs in index order; IDs 0..M-1.
//
// Device plan: per-blob cell keys, a stable radix sort of (cell, index) —
// which yields every cell's blobs in index order — per-cell ranges, then one
// thread per node gathers its stencil cells (no atomics: the sum order is
// fixed, so values are bit-identical run to run and to the CPU oracle).
#include "dvm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>

namespace dvmi {

constexpr uint32_t kKept = 0xFFFFFFFFu;

__device__ __forceinline__ double remesh_w(int kernel, double r)
{
    const double a = fabs(r);
    if (kernel == 0) {
        if (a < 1.0) return 1.0 - 2.5 * a * a + 1.5 * a * a * a;
        if (a < 2.0) {
            const double t = 2.0 - a;
            return 0.5 * t * t * (1.0 - a);
        }
        return 0.0;
    }
    if (kernel == 2) {
        if (a >= 3.0) return 0.0;
        const double y3 = 3.0 - a, y2 = 2.0 - a, y1 = 1.0 - a;
        const double p3 = y3 > 0.0 ? y3 * y3 * y3 * y3 * y3 : 0.0;
        const double p2 = y2 > 0.0 ? y2 * y2 * y2 * y2 * y2 : 0.0;
        const double p1 = y1 > 0.0 ? y1 * y1 * y1 * y1 * y1 : 0.0;
        return (p3 - 6.0 * p2 + 15.0 * p1) / 120.0;
    }
    if (a >= 3.0) return 0.0;
    return exp(-0.5 * a * a) / sqrt(2.0 * 3.14159265358979323846);
}

// Weight of stencil slot t (node base + t) for fractional position f.
__device__ __forceinline__ double remesh_slot_w(int kernel, int s, double f, int64_t base, int t)
{
    const double w = remesh_w(kernel, f - (double)(base + t));
    if (kernel != 1) return w;
    double S = 0.0;
    for (int q = 0; q < 2 * s; ++q) S += remesh_w(kernel, f - (double)(base + q));
    return w / S;
}

struct MeshGeo {
    double xmin, zmin, dx, dz;
    int64_t nx, nz;
    int s, kernel;
};

// Per blob: its cell key (kKept for a blob whose stencil leaves the mesh)
// and whether it is remeshed; per block: the smallest and largest cell
// column and row of its remeshed blobs in part[4 * block + 0..3] (min cx,
// max cx, min cz, max cz), folded by remesh_rect_kernel.
__global__ __launch_bounds__(256) void remesh_cell_kernel(int64_t n, MeshGeo g,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ z,
                                                          uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ idx,
                                                          int* __restrict__ inside,
                                                          int* __restrict__ part)
{
    __shared__ int red[4][4];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int r0 = INT_MAX, r1 = -1, r2 = INT_MAX, r3 = -1;
    if (b < n) {
        const double fx = (x[b] - g.xmin) / g.dx;
        const double fz = (z[b] - g.zmin) / g.dz;
        bool ok = fx >= -1e15 && fx <= 1e15 && fz >= -1e15 && fz <= 1e15;  // false for NaN
        uint32_t k = kKept;
        if (ok) {
            const int64_t cx = (int64_t)floor(fx), cz = (int64_t)floor(fz);
            ok = (cx - g.s + 1 >= 0) && (cx + g.s <= g.nx - 1) && (cz - g.s + 1 >= 0) &&
                 (cz + g.s <= g.nz - 1);
            if (ok) {
                k = (uint32_t)(cz * g.nx + cx);
                r0 = r1 = (int)cx;
                r2 = r3 = (int)cz;
            }
        }
        key[b] = k;
        idx[b] = (uint32_t)b;
        inside[b] = ok ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        r0 = min(r0, __shfl_xor(r0, o));
        r1 = max(r1, __shfl_xor(r1, o));
        r2 = min(r2, __shfl_xor(r2, o));
        r3 = max(r3, __shfl_xor(r3, o));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = r0; red[wave][1] = r1; red[wave][2] = r2; red[wave][3] = r3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < (int)(blockDim.x >> 6); ++v) {
            r0 = min(r0, red[v][0]); r1 = max(r1, red[v][1]);
            r2 = min(r2, red[v][2]); r3 = max(r3, red[v][3]);
        }
        int* o = part + 4 * (int64_t)blockIdx.x;
        o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
    }
}

// rect[0..3] = the fold of the nb per-block rectangles (one block); with no
// remeshed blob, {INT_MAX, -1, INT_MAX, -1}, which no node lies in.
__global__ __launch_bounds__(256) void remesh_rect_kernel(int nb, const int* __restrict__ part,
                                                          int* __restrict__ rect)
{
    __shared__ int red[4][4];
    int r0 = INT_MAX, r1 = -1, r2 = INT_MAX, r3 = -1;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        r0 = min(r0, part[4 * b]); r1 = max(r1, part[4 * b + 1]);
        r2 = min(r2, part[4 * b + 2]); r3 = max(r3, part[4 * b + 3]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        r0 = min(r0, __shfl_xor(r0, o));
        r1 = max(r1, __shfl_xor(r1, o));
        r2 = min(r2, __shfl_xor(r2, o));
        r3 = max(r3, __shfl_xor(r3, o));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = r0; red[wave][1] = r1; red[wave][2] = r2; red[wave][3] = r3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < (int)(blockDim.x >> 6); ++v) {
            r0 = min(r0, red[v][0]); r1 = max(r1, red[v][1]);
            r2 = min(r2, red[v][2]); r3 = max(r3, red[v][3]);
        }
        rect[0] = r0; rect[1] = r1; rect[2] = r2; rect[3] = r3;
    }
}

// Sorted copies of the remeshed blobs' fractional cell coordinates and
// circulation (position q of the (cell, index) order), so the node gather
// reads them contiguously instead of through the permutation.
__global__ void remesh_sorted_kernel(int64_t n, MeshGeo g, const uint32_t* __restrict__ sidx,
                                     const double* __restrict__ x, const double* __restrict__ z,
                                     const double* __restrict__ c, double* __restrict__ sfx,
                                     double* __restrict__ sfz, double* __restrict__ sc)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t b = sidx[q];
    sfx[q] = (x[b] - g.xmin) / g.dx;
    sfz[q] = (z[b] - g.zmin) / g.dz;
    sc[q] = c[b];
}

__global__ void remesh_range_kernel(int64_t n, const uint32_t* __restrict__ key,
                                    int* __restrict__ start, int* __restrict__ end)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = key[p];
    if (k == kKept) return;
    if (p == 0 || key[p - 1] != k) start[k] = (int)p;
    if (p == n - 1 || key[p + 1] != k) end[k] = (int)p + 1;
}

// One thread per node: the sum over its stencil cells (z-major, ascending)
// of each cell's blobs in index order.  Nodes outside the rectangle the
// remeshed blobs' stencils reach (rect, widened by the stencil) have no
// blobs in reach and take the value 0 without visiting their cells.
__global__ void remesh_gather_kernel(MeshGeo g, const int* __restrict__ rect,
                                     const int* __restrict__ start, const int* __restrict__ end,
                                     const double* __restrict__ sfx,
                                     const double* __restrict__ sfz,
                                     const double* __restrict__ sc, double threshold,
                                     double* __restrict__ G, int* __restrict__ flag)
{
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nx * g.nz) return;
    const int64_t k = node / g.nx;
    const int64_t i = node - k * g.nx;
    const int s = g.s;
    double acc = 0.0;
    if (i >= (int64_t)rect[0] - s + 1 && i <= (int64_t)rect[1] + s &&
        k >= (int64_t)rect[2] - s + 1 && k <= (int64_t)rect[3] + s) {
        for (int64_t cz = k - s; cz <= k + s - 1; ++cz) {
            if (cz < 0 || cz >= g.nz) continue;
            for (int64_t cx = i - s; cx <= i + s - 1; ++cx) {
                if (cx < 0 || cx >= g.nx) continue;
                const int64_t cell = cz * g.nx + cx;
                const int q1 = end[cell];
                for (int q = start[cell]; q < q1; ++q) {
                    const double wx = remesh_slot_w(g.kernel, s, sfx[q], cx - s + 1,
                                                    (int)(i - (cx - s + 1)));
                    const double wz = remesh_slot_w(g.kernel, s, sfz[q], cz - s + 1,
                                                    (int)(k - (cz - s + 1)));
                    acc += (sc[q] * wx) * wz;
                }
            }
        }
    }
    G[node] = acc;
    flag[node] = fabs(acc) > threshold ? 1 : 0;
}

__global__ void remesh_emit_nodes_kernel(MeshGeo g, double sigma, const double* __restrict__ G,
                                         const int* __restrict__ flag, const int* __restrict__ pos,
                                         BlobBuf out)
{
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nx * g.nz || !flag[node]) return;
    const int64_t k = node / g.nx;
    const int64_t i = node - k * g.nx;
    const int64_t o = pos[node];
    out.x[o] = g.xmin + (double)i * g.dx;
    out.z[o] = g.zmin + (double)k * g.dz;
    out.c[o] = G[node];
    out.s[o] = sigma;
    if (out.u) { out.u[o] = 0.0; out.w[o] = 0.0; out.uvs[o] = 0.0; out.wvs[o] = 0.0; }
    if (out.id) out.id[o] = (uint32_t)o;
}

__global__ void remesh_emit_kept_kernel(int64_t n_kept, int64_t first, int64_t base,
                                        const uint32_t* __restrict__ sidx, BlobBuf in, BlobBuf out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_kept) return;
    const uint32_t b = sidx[first + q];
    const int64_t o = base + q;
    out.x[o] = in.x[b]; out.z[o] = in.z[b]; out.c[o] = in.c[b]; out.s[o] = in.s[b];
    if (out.u) {
        out.u[o] = in.u[b]; out.w[o] = in.w[b]; out.uvs[o] = in.uvs[b]; out.wvs[o] = in.wvs[b];
    }
    if (out.id) out.id[o] = (uint32_t)o;
}

__global__ void remesh_tail_kernel(int64_t cells, const int* __restrict__ pos,
                                   const int* __restrict__ flag, int* __restrict__ nin)
{
    if (threadIdx.x != 0) return;
    nin[1] = pos[cells - 1];
    nin[2] = flag[cells - 1];
}

static MeshGeo mesh_geo(const dvm_remesh_params& rp)
{
    MeshGeo g;
    g.xmin = rp.xmin; g.zmin = rp.zmin; g.dx = rp.dx; g.dz = rp.dz;
    g.nx = (int64_t)std::llround((rp.xmax - rp.xmin) / rp.dx) + 1;
    g.nz = (int64_t)std::llround((rp.zmax - rp.zmin) / rp.dz) + 1;
    g.kernel = rp.kernel;
    g.s = rp.kernel == 0 ? 2 : 3;
    return g;
}

// Plan memory layout (offsets recomputed from the sizes on both phases).
struct PlanPtrs {
    uint32_t *key, *key2, *idx, *sidx;
    int *inside, *start, *end, *flag, *pos, *nin, *rect, *rpart;
    double *G, *sfx, *sfz, *sc;
    void* tmp;
    size_t tmp_bytes;
};

static PlanPtrs plan_ptrs(char* mem, int64_t n, int64_t cells, size_t tmp_bytes)
{
    Carver cv(mem);
    PlanPtrs p;
    p.key = cv.take<uint32_t>(n > 0 ? n : 1);
    p.key2 = cv.take<uint32_t>(n > 0 ? n : 1);
    p.idx = cv.take<uint32_t>(n > 0 ? n : 1);
    p.sidx = cv.take<uint32_t>(n > 0 ? n : 1);
    p.inside = cv.take<int>(n > 0 ? n : 1);
    p.start = cv.take<int>(cells);
    p.end = cv.take<int>(cells);
    p.flag = cv.take<int>(cells);
    p.pos = cv.take<int>(cells);
    p.nin = cv.take<int>(4);
    p.rect = cv.take<int>(4);
    p.rpart = cv.take<int>(4 * cdiv(n > 0 ? n : 1, 256));
    p.G = cv.take<double>(cells);
    p.sfx = cv.take<double>(n > 0 ? n : 1);
    p.sfz = cv.take<double>(n > 0 ? n : 1);
    p.sc = cv.take<double>(n > 0 ? n : 1);
    p.tmp = cv.take<char>(tmp_bytes);
    p.tmp_bytes = tmp_bytes;
    return p;
}

static size_t plan_tmp_bytes(int64_t n, int64_t cells, hipStream_t st)
{
    size_t a = 0, b = 0, c = 0;
    const int nn = (int)(n > 0 ? n : 1);
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, a, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                     (uint32_t*)nullptr, (uint32_t*)nullptr, nn, 0, 32, st));
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int*)nullptr, (int*)nullptr, (int)cells, st));
    DVM_HIP_CHECK(hipcub::DeviceReduce::Sum(nullptr, c, (int*)nullptr, (int*)nullptr, nn, st));
    size_t m = a > b ? a : b;
    return m > c ? m : c;
}

int64_t remesh_count(dvm_ctx_t ctx, const dvm_remesh_params& rp, int64_t n, const BlobBuf& in,
                     RemeshPlan& plan)
{
    const MeshGeo g = mesh_geo(rp);
    require(g.nx >= 1 && g.nz >= 1 && g.nx * g.nz < (int64_t)kKept, DVM_E_SIZE,
            "remesh grid too large");
    require(n < (int64_t(1) << 31), DVM_E_SIZE, "too many blobs");
    const int64_t cells = g.nx * g.nz;
    hipStream_t st = ctx->stream;
    const size_t tmp = plan_tmp_bytes(n, cells, st);
    const int64_t nn = n > 0 ? n : 1;
    const size_t bytes = Carver::need({4 * sizeof(uint32_t) * nn, sizeof(int) * nn,
                                       4 * sizeof(int) * cells, sizeof(int) * 4, sizeof(int) * 4,
                                       4 * sizeof(int) * (size_t)cdiv(nn, 256),
                                       sizeof(double) * cells, 3 * sizeof(double) * nn, tmp}) +
                         20 * 256;
    // the context's work arena: nothing else takes it between this count and
    // remesh_finish, and it persists, so repeated remeshes allocate nothing
    plan.mem = static_cast<char*>(ws_get(ctx, bytes));
    PlanPtrs p = plan_ptrs(plan.mem, n, cells, tmp);
    plan.n_in = n;
    plan.nx = g.nx;
    plan.nz = g.nz;
    ProfScope ps(ctx, "remesh");
    DVM_HIP_CHECK(hipMemsetAsync(p.start, 0, sizeof(int) * cells, st));
    DVM_HIP_CHECK(hipMemsetAsync(p.end, 0, sizeof(int) * cells, st));
    DVM_HIP_CHECK(hipMemsetAsync(p.nin, 0, sizeof(int) * 4, st));
    const int nb = (int)cdiv(n > 0 ? n : 1, 256);
    if (n > 0) {
        remesh_cell_kernel<<<nb, 256, 0, st>>>(n, g, in.x, in.z, p.key, p.idx, p.inside, p.rpart);
        DVM_LAUNCH_CHECK();
    }
    remesh_rect_kernel<<<1, 256, 0, st>>>(n > 0 ? nb : 0, p.rpart, p.rect);
    DVM_LAUNCH_CHECK();
    if (n > 0) {
        size_t t = p.tmp_bytes;
        DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(p.tmp, t, p.key, p.key2, p.idx, p.sidx,
                                                         (int)n, 0, 32, st));
        remesh_range_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, p.key2, p.start, p.end);
        DVM_LAUNCH_CHECK();
        remesh_sorted_kernel<<<cdiv(n, 256), 256, 0, st>>>(n, g, p.sidx, in.x, in.z, in.c, p.sfx,
                                                           p.sfz, p.sc);
        DVM_LAUNCH_CHECK();
        t = p.tmp_bytes;
        DVM_HIP_CHECK(hipcub::DeviceReduce::Sum(p.tmp, t, p.inside, p.nin, (int)n, st));
    }
    remesh_gather_kernel<<<cdiv(cells, 256), 256, 0, st>>>(g, p.rect, p.start, p.end, p.sfx,
                                                           p.sfz, p.sc, rp.circ_threshold, p.G,
                                                           p.flag);
    DVM_LAUNCH_CHECK();
    size_t t = p.tmp_bytes;
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(p.tmp, t, p.flag, p.pos, (int)cells, st));
    // one read-back: {remeshed blobs, last node's output offset, its flag}
    remesh_tail_kernel<<<1, 64, 0, st>>>(cells, p.pos, p.flag, p.nin);
    DVM_LAUNCH_CHECK();
    int* hv = static_cast<int*>(pinned_get(ctx, sizeof(int) * 4));
    DVM_HIP_CHECK(hipMemcpyAsync(hv, p.nin, sizeof(int) * 3, hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    plan.n_nodes_out = (int64_t)hv[1] + hv[2];
    plan.n_kept = n - hv[0];
    return plan.n_nodes_out + plan.n_kept;
}

void remesh_finish(dvm_ctx_t ctx, const dvm_remesh_params& rp, const BlobBuf& in,
                   RemeshPlan& plan, BlobBuf& out)
{
    const MeshGeo g = mesh_geo(rp);
    const int64_t cells = g.nx * g.nz;
    hipStream_t st = ctx->stream;
    const int64_t n = plan.n_in;
    const size_t tmp = plan_tmp_bytes(n, cells, st);
    PlanPtrs p = plan_ptrs(plan.mem, n, cells, tmp);
    remesh_emit_nodes_kernel<<<cdiv(cells, 256), 256, 0, st>>>(g, rp.sigma, p.G, p.flag, p.pos, out);
    DVM_LAUNCH_CHECK();
    if (plan.n_kept > 0) {
        remesh_emit_kept_kernel<<<cdiv(plan.n_kept, 256), 256, 0, st>>>(
            plan.n_kept, n - plan.n_kept, plan.n_nodes_out, p.sidx, in, out);
        DVM_LAUNCH_CHECK();
    }
    plan.mem = nullptr;
}

}  // namespace dvmi

using namespace dvmi;

extern "C" {

dvm_status dvm_remesh(dvm_ctx_t ctx, const dvm_remesh_params* rp, int64_t n, const double* x,
                      const double* z, const double* circ, const double* sigma, int64_t out_cap,
                      double* ox, double* oz, double* ocirc, double* osigma, int64_t* n_out)
{
    if (!ctx || !rp || !n_out) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        require(n == 0 || (x && z && circ && sigma), DVM_E_INVALID, "null input array");
        require(rp->dx > 0 && rp->dz > 0 && rp->xmax > rp->xmin && rp->zmax > rp->zmin,
                DVM_E_INVALID, "invalid mesh");
        require(rp->kernel >= 0 && rp->kernel <= 2, DVM_E_INVALID, "unknown remesh kernel");
        // positions, circulation and core size only: no velocity fields or IDs
        BlobBuf in{};
        in.x = const_cast<double*>(x); in.z = const_cast<double*>(z);
        in.c = const_cast<double*>(circ); in.s = const_cast<double*>(sigma);
        RemeshPlan plan;
        const int64_t m = remesh_count(ctx, *rp, n, in, plan);
        *n_out = m;
        if (m > out_cap || (m > 0 && !(ox && oz && ocirc && osigma)))
            fail(DVM_E_SIZE, "remesh output larger than out_cap (n_out holds the size needed)");
        if (m == 0) return;
        BlobBuf out{};
        out.x = ox; out.z = oz; out.c = ocirc; out.s = osigma;
        remesh_finish(ctx, *rp, in, plan, out);
    });
}

}  // extern "C"
