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
// Device plan: per-blob cell keys and the rectangle of the remeshed blobs'
// cells (read back once: it sets the window of nodes and cells worked on, so
// the cost follows the wake's extent, not the mesh's), a stable radix sort of
// (cell, index) — which yields every cell's blobs in index order — the cells'
// start positions, per-blob weight records, then one lane per window node
// gathers its stencil cells from records staged in LDS (no atomics: the sum
// order is fixed, so values are bit-identical run to run and to the CPU
// oracle); nodes outside the window hold no circulation.
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

constexpr int kMaxS = 3;  // stencil half-width: 2 (M4') or 3

// A remeshed blob in (cell, index) order, for stencil half-width S: circ *
// (its weight on each of the 2S stencil columns of its cell) and its weights
// on the 2S stencil rows (remesh_slot_w, evaluated once per blob instead of
// once per node; the product circ * wx rounds as in the node sum).  Its cell
// column is kept beside it (rcx).
template <int S>
struct RemeshRec {
    double cw[2 * S];
    double wz[2 * S];
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

// Records of the remeshed blobs in (cell, index) order (position q), so the
// node gather reads them contiguously instead of through the permutation.
template <int S>
__global__ void remesh_sorted_kernel(int64_t n, MeshGeo g, const uint32_t* __restrict__ sidx,
                                     const uint32_t* __restrict__ key,
                                     const double* __restrict__ x, const double* __restrict__ z,
                                     const double* __restrict__ c, RemeshRec<S>* __restrict__ rec,
                                     int* __restrict__ rcx)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || key[q] == kKept) return;
    const uint32_t b = sidx[q];
    const double fx = (x[b] - g.xmin) / g.dx;
    const double fz = (z[b] - g.zmin) / g.dz;
    const int64_t cx = (int64_t)floor(fx);
    const int64_t bx = cx - S + 1;
    const int64_t bz = (int64_t)floor(fz) - S + 1;
    const double cb = c[b];
    RemeshRec<S> r;
#pragma unroll
    for (int t = 0; t < 2 * S; ++t) {
        r.cw[t] = cb * remesh_slot_w(g.kernel, S, fx, bx, t);
        r.wz[t] = remesh_slot_w(g.kernel, S, fz, bz, t);
    }
    rec[q] = r;
    rcx[q] = (int)cx;
}

// cstart[c] = the first position of the sorted keys holding a key >= c, for
// the window cells c = 0 .. wcells (kKept sorts after every cell): cell c's
// blobs are positions cstart[c] .. cstart[c + 1] - 1, and consecutive cells
// of a row hold consecutive position ranges.
__global__ void remesh_cstart_kernel(int64_t n, int64_t wcells, const uint32_t* __restrict__ key,
                                     int* __restrict__ cstart)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > wcells) return;
    int64_t lo = 0, hi = n;  // first p with key[p] >= c
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)key[mid] < c) lo = mid + 1; else hi = mid;
    }
    cstart[c] = (int)lo;
}

// The part of the mesh a remesh works on: the nodes [i0, i0 + W) x [k0, k0 + H)
// and the cells [cx0, cx0 + CW) x [cz0, cz0 + CH) that hold remeshed blobs;
// cell arrays are indexed (cz - cz0) * CW + (cx - cx0).  Nodes outside the
// window have no remeshed blob in reach and take the value 0.
struct MeshWindow {
    int64_t i0, k0, W, H;
    int64_t cx0, cz0, CW, CH;
};

// Global cell key -> window cell key (kKept stays).
__global__ void remesh_rekey_kernel(int64_t n, MeshGeo g, MeshWindow w,
                                    uint32_t* __restrict__ key)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint32_t k = key[b];
    if (k == kKept) return;
    const int64_t cz = (int64_t)k / g.nx;
    const int64_t cx = (int64_t)k - cz * g.nx;
    key[b] = (uint32_t)((cz - w.cz0) * w.CW + (cx - w.cx0));
}

// A workgroup per 64 x 4 window nodes (64 columns of 4 rows); one lane per
// node sums (circ * wx) * wz over its stencil cells (z-major, ascending) and
// each cell's blobs in index order.  A node's cells in one stencil row hold
// one contiguous range of the sorted records, and the workgroup's ranges in
// that row one contiguous range: the workgroup walks the cell rows its nodes
// reach in ascending order, stages each row's range through LDS 256 records
// at a time, and each lane reads the part of every staged piece that lies in
// its own range (a record is staged about twice, not once per node row).
constexpr int kGatherCols = 64;
constexpr int kGatherRows = 4;

template <int S>
__global__ __launch_bounds__(kGatherCols * kGatherRows) void remesh_gather_kernel(
    MeshGeo g, MeshWindow w, const int* __restrict__ cstart, const RemeshRec<S>* __restrict__ rec,
    const int* __restrict__ rcx, double threshold, double* __restrict__ G, int* __restrict__ flag)
{
    constexpr int kLanes = kGatherCols * kGatherRows;
    __shared__ RemeshRec<S> st[kLanes];
    __shared__ int stc[kLanes];
    const int col = threadIdx.x % kGatherCols;
    const int row = threadIdx.x / kGatherCols;
    const int64_t a = (int64_t)blockIdx.x * kGatherCols + col;  // node column in the window
    const int64_t i = w.i0 + a;
    // the workgroup's node columns and the window's last cell column
    const int64_t ifirst = w.i0 + (int64_t)blockIdx.x * kGatherCols;
    const int64_t ilast = ifirst + kGatherCols - 1 < w.i0 + w.W - 1 ? ifirst + kGatherCols - 1
                                                                  : w.i0 + w.W - 1;
    const int64_t cxe = w.cx0 + w.CW - 1;
    const int64_t ca = ifirst - S > w.cx0 ? ifirst - S : w.cx0;
    const int64_t cb = ilast + S - 1 < cxe ? ilast + S - 1 : cxe;
    for (int64_t kb = (int64_t)blockIdx.y * kGatherRows; kb < w.H;
         kb += (int64_t)gridDim.y * kGatherRows) {  // workgroup-uniform
        const int64_t kk = kb + row;
        const bool valid = a < w.W && kk < w.H;
        const int64_t k = w.k0 + kk;
        const int64_t kl = w.k0 + (kb + kGatherRows - 1 < w.H - 1 ? kb + kGatherRows - 1 : w.H - 1);
        double acc = 0.0;
        for (int64_t cz = w.k0 + kb - S; cz <= kl + S - 1; ++cz) {
            if (cz < w.cz0 || cz >= w.cz0 + w.CH || ca > cb) continue;
            const int64_t rb = (cz - w.cz0) * w.CW - w.cx0;  // cell index = rb + cx
            const int A = cstart[rb + ca], B = cstart[rb + cb + 1];
            const int64_t ta = i - S > w.cx0 ? i - S : w.cx0;
            const int64_t tb = i + S - 1 < cxe ? i + S - 1 : cxe;
            const bool mine = valid && cz >= k - S && cz <= k + S - 1 && ta <= tb;
            const int qa = mine ? cstart[rb + ta] : A, qb = mine ? cstart[rb + tb + 1] : A;
            const int u = (int)(k - (cz - S + 1));  // this node's row slot in the cells' stencils
            for (int q0 = A; q0 < B; q0 += kLanes) {
                __syncthreads();
                if (q0 + (int)threadIdx.x < B) {
                    st[threadIdx.x] = rec[q0 + threadIdx.x];
                    stc[threadIdx.x] = rcx[q0 + threadIdx.x];
                }
                __syncthreads();
                const int lo = qa > q0 ? qa : q0;
                const int hi = qb < q0 + kLanes ? qb : q0 + kLanes;
                for (int q = lo; q < hi; ++q) {
                    const int v = (int)(i - (stc[q - q0] - S + 1));  // column slot
                    acc += st[q - q0].cw[v] * st[q - q0].wz[u];
                }
            }
        }
        if (valid) {
            const int64_t node = kk * w.W + a;
            G[node] = acc;
            flag[node] = fabs(acc) > threshold ? 1 : 0;
        }
    }
}

__global__ void remesh_emit_nodes_kernel(MeshGeo g, MeshWindow w, double sigma,
                                         const double* __restrict__ G,
                                         const int* __restrict__ flag, const int* __restrict__ pos,
                                         BlobBuf out)
{
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= w.W * w.H || !flag[node]) return;
    const int64_t kk = node / w.W;
    const int64_t i = w.i0 + (node - kk * w.W);
    const int64_t k = w.k0 + kk;
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

__global__ void remesh_iota_kernel(int64_t n, uint32_t* __restrict__ sidx)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) sidx[b] = (uint32_t)b;
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
    int *inside, *start, *flag, *pos, *nin, *rect, *rpart;
    double* G;
    void* rec;  // RemeshRec<S> of the remeshed blobs in (cell, index) order
    int* rcx;   // their cell columns
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
    p.start = cv.take<int>(cells + 1);  // cell starts of the window (cstart)
    p.flag = cv.take<int>(cells);
    p.pos = cv.take<int>(cells);
    p.nin = cv.take<int>(4);
    p.rect = cv.take<int>(4);
    p.rpart = cv.take<int>(4 * cdiv(n > 0 ? n : 1, 256));
    p.G = cv.take<double>(cells);
    p.rec = cv.take<RemeshRec<kMaxS>>(n > 0 ? n : 1);
    p.rcx = cv.take<int>(n > 0 ? n : 1);
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
                                       sizeof(double) * cells, sizeof(RemeshRec<kMaxS>) * nn,
                                       sizeof(int) * nn, tmp}) +
                         20 * 256;
    // the context's work arena: nothing else takes it between this count and
    // remesh_finish, and it persists, so repeated remeshes allocate nothing
    plan.mem = static_cast<char*>(ws_get(ctx, bytes));
    PlanPtrs p = plan_ptrs(plan.mem, n, cells, tmp);
    plan.n_in = n;
    plan.nx = g.nx;
    plan.nz = g.nz;
    ProfScope ps(ctx, "remesh");
    DVM_HIP_CHECK(hipMemsetAsync(p.nin, 0, sizeof(int) * 4, st));
    const int nb = (int)cdiv(n > 0 ? n : 1, 256);
    if (n > 0) {
        remesh_cell_kernel<<<nb, 256, 0, st>>>(n, g, in.x, in.z, p.key, p.idx, p.inside, p.rpart);
        DVM_LAUNCH_CHECK();
    }
    remesh_rect_kernel<<<1, 256, 0, st>>>(n > 0 ? nb : 0, p.rpart, p.rect);
    DVM_LAUNCH_CHECK();
    // the rectangle of the remeshed blobs' cells sizes the window
    int* hv = static_cast<int*>(pinned_get(ctx, sizeof(int) * 8));
    DVM_HIP_CHECK(hipMemcpyAsync(hv, p.rect, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    MeshWindow w{};
    const bool any = hv[1] >= hv[0];
    if (any) {
        w.cx0 = hv[0]; w.CW = (int64_t)hv[1] - hv[0] + 1;
        w.cz0 = hv[2]; w.CH = (int64_t)hv[3] - hv[2] + 1;
    }
    if (rp.circ_threshold < 0.0) {  // every node is output: the whole mesh
        w.i0 = 0; w.W = g.nx;
        w.k0 = 0; w.H = g.nz;
    } else if (any) {  // the nodes the remeshed blobs' stencils reach
        w.i0 = w.cx0 - g.s + 1; w.W = w.CW + 2 * g.s - 1;
        w.k0 = w.cz0 - g.s + 1; w.H = w.CH + 2 * g.s - 1;
    }
    const int64_t wcells = w.CW * w.CH, wnodes = w.W * w.H;
    plan.win_i0 = w.i0; plan.win_k0 = w.k0; plan.win_w = w.W; plan.win_h = w.H;
    if (wcells > 0) {
        remesh_rekey_kernel<<<nb, 256, 0, st>>>(n, g, w, p.key);
        DVM_LAUNCH_CHECK();
        // window keys lie below 2^bits - 1, so kKept (all bits set) sorts last
        int bits = 1;
        while ((int64_t(1) << bits) <= wcells) ++bits;
        size_t t = p.tmp_bytes;
        DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(p.tmp, t, p.key, p.key2, p.idx, p.sidx,
                                                         (int)n, 0, bits, st));
        remesh_cstart_kernel<<<cdiv(wcells + 1, 256), 256, 0, st>>>(n, wcells, p.key2, p.start);
        DVM_LAUNCH_CHECK();
        if (g.s == 2)
            remesh_sorted_kernel<2><<<cdiv(n, 256), 256, 0, st>>>(
                n, g, p.sidx, p.key2, in.x, in.z, in.c, static_cast<RemeshRec<2>*>(p.rec), p.rcx);
        else
            remesh_sorted_kernel<3><<<cdiv(n, 256), 256, 0, st>>>(
                n, g, p.sidx, p.key2, in.x, in.z, in.c, static_cast<RemeshRec<3>*>(p.rec), p.rcx);
        DVM_LAUNCH_CHECK();
        t = p.tmp_bytes;
        DVM_HIP_CHECK(hipcub::DeviceReduce::Sum(p.tmp, t, p.inside, p.nin, (int)n, st));
    } else if (n > 0) {
        // nothing remeshed: every blob is kept, in index order
        remesh_iota_kernel<<<nb, 256, 0, st>>>(n, p.sidx);
        DVM_LAUNCH_CHECK();
    }
    if (wnodes > 0) {
        const dim3 grid(cdiv(w.W, kGatherCols),
                        (unsigned)std::min<int64_t>(cdiv(w.H, kGatherRows), 65535));
        if (g.s == 2)
            remesh_gather_kernel<2><<<grid, kGatherCols * kGatherRows, 0, st>>>(
                g, w, p.start, static_cast<const RemeshRec<2>*>(p.rec), p.rcx, rp.circ_threshold,
                p.G, p.flag);
        else
            remesh_gather_kernel<3><<<grid, kGatherCols * kGatherRows, 0, st>>>(
                g, w, p.start, static_cast<const RemeshRec<3>*>(p.rec), p.rcx, rp.circ_threshold,
                p.G, p.flag);
        DVM_LAUNCH_CHECK();
        size_t t = p.tmp_bytes;
        DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(p.tmp, t, p.flag, p.pos, (int)wnodes, st));
        // {remeshed blobs, last node's output offset, its flag}
        remesh_tail_kernel<<<1, 64, 0, st>>>(wnodes, p.pos, p.flag, p.nin);
        DVM_LAUNCH_CHECK();
    }
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
    MeshWindow w{};
    w.i0 = plan.win_i0; w.k0 = plan.win_k0; w.W = plan.win_w; w.H = plan.win_h;
    if (w.W * w.H > 0) {
        remesh_emit_nodes_kernel<<<cdiv(w.W * w.H, 256), 256, 0, st>>>(g, w, rp.sigma, p.G, p.flag,
                                                                       p.pos, out);
        DVM_LAUNCH_CHECK();
    }
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
