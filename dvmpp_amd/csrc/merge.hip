// VortexBlobs::merge_blobs (src/VortexBlobs.cpp:110-209) on the device, with
// the reference's result exactly.
//
// The reference is greedy in index order: blob i (not yet absorbed) tests every
// later unabsorbed blob j against its CURRENT state and absorbs each j whose
// criterion |Gi Gj / (Gi + Gj)| * |xi - xj| is below the threshold, moving its
// own state to the merged one.  Three facts make this parallel without
// changing the result:
//  * when i's turn comes, i and every j > i still hold their input values (a
//    blob changes only on its own turn);
//  * a blob with no partner among the later blobs of the input never absorbs
//    anything (absorptions only remove candidates) -- merge_cand_kernel finds
//    the blobs that have one, all pairs in parallel;
//  * i's sequence of absorptions computed as if nothing had been absorbed
//    before its turn (its "speculative" sequence J) is its true sequence
//    whenever no member of J was absorbed by an earlier blob: at each step the
//    first passing blob is then still available, and no earlier-indexed blob
//    can pass that the speculation did not already see.
// merge_spec_kernel computes every candidate blob's speculative sequence in
// parallel (one wavefront each); merge_commit_kernel then walks the blobs in
// index order in one wavefront, skips the absorbed ones, commits each
// sequence that is still valid, and recomputes the others exactly with the
// absorptions so far.
//
// A "next partner" search returns the lowest passing index above the last
// one, by a cell-list query.  Since |Gi Gj / (Gi + Gj)| >= min(|Gi|,|Gj|)/2,
// a pair can pass only if d < T / min(|Gi|,|Gj|) = max(T/|Gi|, T/|Gj|) with
// T = 2 thr (1 + 1e-9): within the current blob's reach T/|Gi| or within the
// partner's own reach T/|Gj|.  Blobs are binned by reach into three grid
// levels (cell size h, 8h, 64h; a blob goes to the finest level whose cell is
// at least its reach) by a stable sort on (level, cell), so every cell lists
// ascending indices and the cells of one grid row are contiguous; blobs with
// a longer reach form the long-reach list, in index order.  A query visits,
// on each level, the cells within max(T/|Gi|, cell size) and the long-reach
// list up to the best hit; when the current blob's reach covers too many
// cells it scans every later blob instead.  Blobs that can never pass
// (non-finite position or circulation) are in neither.
//
// The criterion and the merged state are evaluated as written, in FP64 with
// no contraction and correctly rounded division and square root, so every
// decision and every value matches the reference's loop.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>

#include "dvm_internal.hpp"

namespace dvmi {

namespace {

constexpr int kCandThreads = 256;
constexpr int kW = 64;              // a wavefront
constexpr int kLevels = 3;          // grid levels, cell sizes h * 8^l
constexpr int kGridSide = 2049;     // cells per axis on the finest level, at most
constexpr int kMaxCells = 65536;    // cells one query may visit (all levels)
constexpr int kCandMax = 16;        // input-state partners recorded per candidate blob
constexpr int kSpecMax = 16;        // absorptions one speculative sequence may hold
constexpr int kSpecBlocks = 2048;   // merge_spec_kernel blocks (4 wavefronts each)
constexpr int kBboxBlocks = 256;

// src/VortexBlobs.cpp:127-142, with a rejection that cannot change the
// outcome: |Gi Gj / (Gi + Gj)| >= min(|Gi|, |Gj|) / 2 for any signs, so when
// min(|Gi|,|Gj|)^2 d^2 > 4 thr^2 (1 + 1e-9) the exact criterion exceeds the
// threshold by far more than its rounding (`rej` = 4 thr^2 (1 + 1e-9)).
__device__ __forceinline__ bool merge_pass(double gi, double xi, double zi, double gj, double xj,
                                           double zj, double thr, double rej)
{
    const double dx = xi - xj;
    const double dz = zi - zj;
    const double d2 = dx * dx + dz * dz;
    const double m = fmin(fabs(gi), fabs(gj));
    if (m * m * d2 > rej) return false;
    const double gs = gi + gj;
    if (fabs(gs) < 1e-15) return false;
    const double dist = sqrt(d2);
    return fabs(gi * gj / gs) * dist < thr;
}

__global__ __launch_bounds__(kCandThreads) void merge_cand_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ z,
    const double* __restrict__ c, double thr, double rej, int* __restrict__ cand,
    int* __restrict__ removed)
{
    __shared__ double sx[kCandThreads], sz[kCandThreads], sc[kCandThreads];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kCandThreads + t;
    const bool valid = i < n;
    double xi = 0.0, zi = 0.0, gi = 0.0;
    if (valid) {
        xi = x[i];
        zi = z[i];
        gi = c[i];
        removed[i] = 0;
    }
    int found = n;
    bool done = !valid;
    for (int j0 = blockIdx.x * kCandThreads; j0 < n; j0 += kCandThreads) {
        // also the barrier before the tile is overwritten
        if (__syncthreads_and(done)) break;
        const int jl = j0 + t;
        if (jl < n) {
            sx[t] = x[jl];
            sz[t] = z[jl];
            sc[t] = c[jl];
        }
        __syncthreads();
        if (!done) {
            const int lim = min(kCandThreads, n - j0);
            for (int k = 0; k < lim; ++k) {
                const int j = j0 + k;
                if (j > i && merge_pass(gi, xi, zi, sc[k], sx[k], sz[k], thr, rej)) {
                    found = j;
                    done = true;
                    break;
                }
            }
        }
    }
    if (valid) cand[i] = found;
}

// Cell lists of the blobs that can pass (see the top of this file).
struct MergeGrid {
    double x0, z0, T;
    double h[kLevels];
    double reach_le[kLevels];  // |G| >= reach_le[l]: reach at most h[l]
    int nx[kLevels], nz[kLevels], off[kLevels];
    int ncell;                     // all levels; ncell = long reach, ncell + 1 = never passes
    const int* __restrict__ sidx;  // blob indices sorted by (level, cell, index)
    const int* __restrict__ cbeg;  // cbeg[k] = first sorted position with key >= k (ncell + 3)
};

__global__ void merge_key_kernel(int n, const double* __restrict__ x, const double* __restrict__ z,
                                 const double* __restrict__ c, MergeGrid g, int* __restrict__ key,
                                 int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], zi = z[i], a = fabs(c[i]);
    int k = g.ncell;  // long reach unless a level takes it
    if (!isfinite(xi) || !isfinite(zi) || !isfinite(a)) {
        k = g.ncell + 1;
    } else {
        for (int l = 0; l < kLevels; ++l) {
            if (a >= g.reach_le[l]) {
                const int cx = min(max((int)floor((xi - g.x0) / g.h[l]), 0), g.nx[l] - 1);
                const int cz = min(max((int)floor((zi - g.z0) / g.h[l]), 0), g.nz[l] - 1);
                k = g.off[l] + cz * g.nx[l] + cx;
                break;
            }
        }
    }
    key[i] = k;
    idx[i] = i;
}

// cbeg[k] = first sorted position with key >= k, k = 0 .. ncell + 2, one
// thread per key (a binary search: empty runs cost nothing extra).
__global__ void merge_cbeg_kernel(int n, const int* __restrict__ key, int ncell, int* __restrict__ cbeg)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > ncell + 2) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    cbeg[k] = lo;
}

__device__ __forceinline__ int wave_min(int v)
{
    for (int o = kW / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

// the lanes of a wavefront see each other's global stores
__device__ __forceinline__ void wave_sync_memory()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// The absorbed flags as the searches and the commit see them: none (the
// speculative pass), the global array, or a bitmap in the commit wavefront's
// LDS (one bit per blob; written back to the global array at the end).  sync()
// makes a wavefront's own flag writes visible to its next reads.
struct TakenNone {
    __device__ bool get(int) const { return false; }
};
struct TakenGlobal {
    int* r;
    __device__ bool get(int j) const { return r[j] != 0; }
    __device__ void set(int j) const { r[j] = 1; }
    __device__ void sync() const { wave_sync_memory(); }
};
struct TakenLds {
    uint32_t* bits;
    __device__ bool get(int j) const { return (bits[j >> 5] >> (j & 31)) & 1u; }
    __device__ void set(int j) const { atomicOr(&bits[j >> 5], 1u << (j & 31)); }
    // LDS operations of one wavefront complete in order: only the compiler
    // must not move them
    __device__ void sync() const
    {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
};
constexpr int kTakenWords = 32768;  // LDS bitmap: up to 1,048,576 blobs

// Cell range [i0, i1] on one axis covering (v - r, v + r); a cell boundary
// within 1e-6 of a cell of the range edge adds that cell (rounding of the
// binning).  Returns false when the range is not finite.
__device__ __forceinline__ bool cell_range(double v, double r, double v0, double h, int nmax, int& i0,
                                           int& i1)
{
    const double lo = (v - r - v0) / h, hi = (v + r - v0) / h;
    if (!(isfinite(lo) && isfinite(hi))) return false;
    double f0 = floor(lo), f1 = floor(hi);
    if (lo - f0 < 1e-6) f0 -= 1.0;
    if (f1 + 1.0 - hi < 1e-6) f1 += 1.0;
    f0 = fmax(f0, 0.0);
    f1 = fmin(f1, (double)(nmax - 1));
    i0 = (int)f0;
    i1 = (int)f1;
    if (f0 > f1) i1 = i0 - 1;  // empty
    return true;
}

// Lowest j >= j0 with !removed[j] that passes against (cg, cx, cz), by a
// forward scan of all later blobs; INT_MAX if none.  Lanes hold consecutive
// indices, so the lowest set ballot bit is the lowest passing index.
template <class T>
__device__ int next_linear(int n, const double* __restrict__ x, const double* __restrict__ z,
                           const double* __restrict__ c, const T& taken, double cg, double cx,
                           double cz, int j0, double thr, double rej)
{
    const int lane = threadIdx.x & (kW - 1);
    for (; j0 < n; j0 += kW) {
        const int j = j0 + lane;
        bool hit = false;
        if (j < n && !taken.get(j)) hit = merge_pass(cg, cx, cz, c[j], x[j], z[j], thr, rej);
        const unsigned long long b = __ballot(hit);
        if (b) return j0 + __ffsll((long long)b) - 1;
    }
    return INT_MAX;
}

// Per-wavefront LDS of the queries.
struct WaveLds {
    int sbeg[kW];  // first sorted entry of each visited grid row
    int soff[kW];  // entry offset of each visited grid row
    int buf[kW];   // hits collected by all_partners (unsorted)
    int out[kW];   // the same, ascending
};

// Visits every blob j >= j0 that might pass against (cg, cx, cz): the grid
// cells within reach on each level (rows in groups of kW, one per lane), then
// the long-reach list from the first index >= j0, in index order, below
// `stop_below` (which the visitor may lower).  Every lane calls
// `visit(j, ordered)` for its entry (j = -1: none; `ordered`: entries come in
// ascending index order); a true return ends the query.  Returns false,
// having visited nothing, if the cells within reach exceed kMaxCells.
template <class Visit>
__device__ bool visit_near(double cg, double cx, double cz, int j0, const MergeGrid& g, WaveLds& L,
                           const int& stop_below, Visit visit)
{
    const int lane = threadIdx.x & (kW - 1);
    const double reach = g.T / fabs(cg);
    int ix0[kLevels], ix1[kLevels], iz0[kLevels], rows[kLevels];
    int nrows = 0;
    double ncells = 0.0;
    for (int l = 0; l < kLevels; ++l) {
        const double rq = fmax(reach, g.h[l]);
        int a0, a1, b0, b1;
        if (!cell_range(cx, rq, g.x0, g.h[l], g.nx[l], a0, a1) ||
            !cell_range(cz, rq, g.z0, g.h[l], g.nz[l], b0, b1))
            return false;
        ix0[l] = a0;
        ix1[l] = a1;
        iz0[l] = b0;
        rows[l] = (a1 >= a0 && b1 >= b0) ? b1 - b0 + 1 : 0;
        nrows += rows[l];
        ncells += (double)rows[l] * (a1 - a0 + 1);
    }
    if (ncells > kMaxCells) return false;
    for (int r0 = 0; r0 < nrows; r0 += kW) {
        // lane r holds visited grid row r0 + r: its run of sorted entries
        int pb = 0, cnt = 0;
        if (r0 + lane < nrows) {
            int l = 0, r = r0 + lane;
            while (r >= rows[l]) r -= rows[l++];
            const int cb = g.off[l] + (iz0[l] + r) * g.nx[l];
            pb = g.cbeg[cb + ix0[l]];
            cnt = g.cbeg[cb + ix1[l] + 1] - pb;
        }
        int incl = cnt;
        for (int o = 1; o < kW; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const int nr = min(kW, nrows - r0);
        const int total = __shfl(incl, nr - 1);
        L.sbeg[lane] = pb;
        L.soff[lane] = incl - cnt;
        __builtin_amdgcn_wave_barrier();
        bool done = false;
        for (int e0 = 0; e0 < total && !done; e0 += kW) {
            const int e = e0 + lane;
            int j = -1;
            if (e < total) {
                int r = 0, rh = nr - 1;  // the row with soff[r] <= e < soff[r + 1]
                while (r < rh) {
                    const int mid = (r + rh + 1) >> 1;
                    if (L.soff[mid] <= e) r = mid;
                    else rh = mid - 1;
                }
                j = g.sidx[L.sbeg[r] + e - L.soff[r]];
                if (j < j0) j = -1;
            }
            done = visit(j, false);
        }
        __builtin_amdgcn_wave_barrier();  // the tables are rewritten next
        if (done) return true;
    }
    // long-reach blobs: first position with index >= j0 by a 64-ary search
    // (entries are in index order), then forward in chunks
    const int L1 = g.cbeg[g.ncell + 1];
    int lo = g.cbeg[g.ncell], hi = L1;  // the answer lies in [lo, hi]
    while (hi - lo > 1) {
        const int step = (hi - lo + kW - 1) / kW;
        const int p = lo + lane * step;
        const unsigned long long b = __ballot(p < hi && g.sidx[p] >= j0);
        const int q = b ? lo + (__ffsll((long long)b) - 1) * step : hi;
        lo = max(lo, q - step);
        hi = q;
        if (step == 1) break;
    }
    for (int p = lo; p < L1; p += kW) {
        const int e = p + lane;
        int j = e < L1 ? g.sidx[e] : -1;
        const bool past = __ballot(j >= stop_below) != 0;  // later entries are all larger
        if (j >= stop_below || j < j0) j = -1;
        if (visit(j, true) || past) return true;
    }
    return true;
}

// Lowest j >= j0 with !removed[j] that passes against (cg, cx, cz), by a
// cell-list query (falls back to next_linear).  TakenNone ignores the
// absorbed flags (the speculative pass, where none are set).
template <class T>
__device__ int next_partner(int n, const double* __restrict__ x, const double* __restrict__ z,
                            const double* __restrict__ c, const T& taken, double cg, double cx,
                            double cz, int j0, double thr, double rej, const MergeGrid& g, WaveLds& L)
{
    int best = INT_MAX;
    auto visit = [&](int j, bool ordered) {
        const bool hit = j >= 0 && j < best && !taken.get(j) &&
                         merge_pass(cg, cx, cz, c[j], x[j], z[j], thr, rej);
        const bool any = __ballot(hit) != 0;
        if (any) best = min(best, wave_min(hit ? j : INT_MAX));
        return ordered && any;  // in index order the first hit is the answer
    };
    if (!visit_near(cg, cx, cz, j0, g, L, best, visit))
        return next_linear(n, x, z, c, taken, cg, cx, cz, j0, thr, rej);
    return best;
}

// Every j >= j0 passing against the state, ascending, into L.out; returns the
// count, or -1 when there are more than kW of them or the query is too wide.
__device__ int all_partners(const double* __restrict__ x, const double* __restrict__ z,
                            const double* __restrict__ c, double cg, double cx, double cz, int j0,
                            double thr, double rej, const MergeGrid& g, WaveLds& L)
{
    const int lane = threadIdx.x & (kW - 1);
    int count = 0;
    const int no_stop = INT_MAX;
    auto collect = [&](int j, bool) {
        const bool hit = j >= 0 && merge_pass(cg, cx, cz, c[j], x[j], z[j], thr, rej);
        const unsigned long long b = __ballot(hit);
        const int at = count + __popcll(b & ((1ull << lane) - 1ull));
        if (hit && at < kW) L.buf[at] = j;
        count += __popcll(b);
        return count > kW;
    };
    if (!visit_near(cg, cx, cz, j0, g, L, no_stop, collect)) return -1;
    if (count > kW) return -1;
    __builtin_amdgcn_wave_barrier();
    if (lane < count) {
        const int v = L.buf[lane];
        int rank = 0;
        for (int m = 0; m < count; ++m) rank += L.buf[m] < v ? 1 : 0;
        L.out[rank] = v;
    }
    __builtin_amdgcn_wave_barrier();
    return count;
}

// src/VortexBlobs.cpp:147-160
__device__ __forceinline__ void absorb(double& cx, double& cz, double& cg, double& cs, double xj,
                                       double zj, double gj, double sj)
{
    const double gi = cg;
    const double gs = gi + gj;
    const double nx = (gi * cx + gj * xj) / gs;
    const double nz = (gi * cz + gj * zj) / gs;
    const double ns = (fabs(gi) * cs + fabs(gj) * sj) / (fabs(gi) + fabs(gj));
    cx = nx;
    cz = nz;
    cg = gs;
    cs = ns;
}

// Candidate blobs in index order.
__global__ void merge_flag_kernel(int n, const int* __restrict__ cand, int* __restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = cand[i] < n ? 1 : 0;
}

__global__ void merge_list_kernel(int n, const int* __restrict__ f, const int* __restrict__ pos,
                                  int* __restrict__ list, int* __restrict__ nlist)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (f[i]) list[pos[i]] = i;
    if (i == n - 1) *nlist = pos[i] + f[i];
}

// Speculative sequences: wavefront w takes list slots w, w + W, ...  For each
// it records up to kSpecMax absorbed indices (count kSpecMax + 1: longer) and
// the final state, and the blob's partners against its input state: the
// kCandMax lowest, ascending, with their count (kCandMax + 1: more exist;
// -1: not known).
__global__ __launch_bounds__(256) void merge_spec_kernel(
    int n, const double* __restrict__ x, const double* __restrict__ z, const double* __restrict__ c,
    const double* __restrict__ s, const int* __restrict__ cand, const int* __restrict__ list,
    const int* __restrict__ nlist, double thr, double rej, MergeGrid g, int* __restrict__ jn,
    int* __restrict__ jl, double* __restrict__ fs, int* __restrict__ cn, int* __restrict__ cl)
{
    __shared__ WaveLds lds[4];
    const int wv = threadIdx.x / kW, lane = threadIdx.x & (kW - 1);
    WaveLds& L = lds[wv];
    const int nl = *nlist;
    for (int slot = blockIdx.x * 4 + wv; slot < nl; slot += gridDim.x * 4) {
        const int ii = list[slot];
        double cx = x[ii], cz = z[ii], cg = c[ii], cs = s[ii];
        const int np = all_partners(x, z, c, cg, cx, cz, cand[ii], thr, rej, g, L);
        if (lane < kCandMax && lane < np) cl[(size_t)slot * kCandMax + lane] = L.out[lane];
        if (lane == 0) cn[slot] = np > kCandMax ? kCandMax + 1 : np;
        int k = 0;
        int jm = cand[ii];  // passes against the input state
        while (jm != INT_MAX) {
            if (k == kSpecMax) {  // longer than recorded: recomputed at commit
                ++k;
                break;
            }
            if (lane == 0) jl[(size_t)slot * kSpecMax + k] = jm;
            ++k;
            absorb(cx, cz, cg, cs, x[jm], z[jm], c[jm], s[jm]);
            jm = next_partner(n, x, z, c, TakenNone{}, cg, cx, cz, jm + 1, thr, rej, g, L);
        }
        if (lane == 0) {
            jn[slot] = k;
            fs[4 * (size_t)slot] = cx;
            fs[4 * (size_t)slot + 1] = cz;
            fs[4 * (size_t)slot + 2] = cg;
            fs[4 * (size_t)slot + 3] = cs;
        }
    }
}

// One wavefront walks the candidate blobs in index order (src/VortexBlobs.cpp:
// 121-171): absorbed ones are skipped, still-valid speculative sequences are
// committed, the others recomputed against the absorptions so far, starting
// from the first input-state partner still available.  Lane 0 writes the
// shared results.  out[0] = merges, out[1] = recomputed sequences.
template <class T>
__device__ void merge_commit(
    const T& taken, int n, double* __restrict__ x, double* __restrict__ z, double* __restrict__ c,
    double* __restrict__ s, double* __restrict__ u, double* __restrict__ w,
    double* __restrict__ uvs, double* __restrict__ wvs, const int* __restrict__ cand,
    const int* __restrict__ list, const int* __restrict__ nlist, const int* __restrict__ jn,
    const int* __restrict__ jl, const double* __restrict__ fs, const int* __restrict__ cn,
    const int* __restrict__ cl, double thr, double rej, const MergeGrid& g, WaveLds& L,
    int* __restrict__ out)
{
    const int lane = threadIdx.x;
    const int nl = *nlist;
    int total = 0, redo = 0;
    for (int slot = 0; slot < nl; ++slot) {
        const int ii = list[slot];
        const int nj = jn[slot];
        const int np = cn[slot];
        const int jv = lane < min(nj, kSpecMax) ? jl[(size_t)slot * kSpecMax + lane] : -1;
        const int pv = lane < min(np, kCandMax) ? cl[(size_t)slot * kCandMax + lane] : -1;
        if (taken.get(ii)) continue;  // absorbed by an earlier blob
        bool valid = nj <= kSpecMax;
        if (valid && __ballot(jv >= 0 && taken.get(jv))) valid = false;
        double cx, cz, cg, cs;
        if (valid) {
            if (jv >= 0) taken.set(jv);
            cx = fs[4 * (size_t)slot];
            cz = fs[4 * (size_t)slot + 1];
            cg = fs[4 * (size_t)slot + 2];
            cs = fs[4 * (size_t)slot + 3];
            total += nj;
        } else {
            cx = x[ii];
            cz = z[ii];
            cg = c[ii];
            cs = s[ii];
            int jm;
            if (np >= 0) {
                // the first still-available partner is the first absorption
                const unsigned long long b = __ballot(pv >= 0 && !taken.get(pv));
                if (b) jm = __shfl(pv, __ffsll((long long)b) - 1);
                else if (np <= kCandMax) continue;  // every partner is taken: no merge
                else jm = next_partner(n, x, z, c, taken, cg, cx, cz,
                                             __shfl(pv, kCandMax - 1) + 1, thr, rej, g, L);
            } else {
                // every j in (ii, cand) fails against the input state
                jm = next_partner(n, x, z, c, taken, cg, cx, cz, cand[ii], thr, rej, g, L);
            }
            ++redo;
            int k = 0;
            while (jm != INT_MAX) {
                absorb(cx, cz, cg, cs, x[jm], z[jm], c[jm], s[jm]);
                if (lane == 0) taken.set(jm);
                ++k;
                taken.sync();
                jm = next_partner(n, x, z, c, taken, cg, cx, cz, jm + 1, thr, rej, g, L);
            }
            total += k;
            if (k == 0) continue;
        }
        if (lane == 0) {
            x[ii] = cx;
            z[ii] = cz;
            c[ii] = cg;
            s[ii] = cs;
            u[ii] = 0.0;
            w[ii] = 0.0;
            uvs[ii] = 0.0;
            wvs[ii] = 0.0;
        }
        taken.sync();
    }
    if (lane == 0) {
        out[0] = total;
        out[1] = redo;
    }
}

__global__ __launch_bounds__(kW) void merge_commit_kernel(
    int n, double* __restrict__ x, double* __restrict__ z, double* __restrict__ c,
    double* __restrict__ s, double* __restrict__ u, double* __restrict__ w,
    double* __restrict__ uvs, double* __restrict__ wvs, const int* __restrict__ cand,
    const int* __restrict__ list, const int* __restrict__ nlist, const int* __restrict__ jn,
    const int* __restrict__ jl, const double* __restrict__ fs, const int* __restrict__ cn,
    const int* __restrict__ cl, int* removed, double thr, double rej, MergeGrid g,
    int* __restrict__ out)
{
    __shared__ WaveLds L;
    merge_commit(TakenGlobal{removed}, n, x, z, c, s, u, w, uvs, wvs, cand, list, nlist, jn, jl,
                 fs, cn, cl, thr, rej, g, L, out);
}

// The same with the absorbed flags in LDS (n <= 32 kTakenWords): no global
// round trip per candidate; the flags go to `removed` at the end.
__global__ __launch_bounds__(kW) void merge_commit_lds_kernel(
    int n, double* __restrict__ x, double* __restrict__ z, double* __restrict__ c,
    double* __restrict__ s, double* __restrict__ u, double* __restrict__ w,
    double* __restrict__ uvs, double* __restrict__ wvs, const int* __restrict__ cand,
    const int* __restrict__ list, const int* __restrict__ nlist, const int* __restrict__ jn,
    const int* __restrict__ jl, const double* __restrict__ fs, const int* __restrict__ cn,
    const int* __restrict__ cl, int* removed, double thr, double rej, MergeGrid g,
    int* __restrict__ out)
{
    __shared__ WaveLds L;
    __shared__ uint32_t bits[kTakenWords];
    const int lane = threadIdx.x;
    const int nw = (n + 31) >> 5;
    for (int k = lane; k < nw; k += kW) bits[k] = 0u;
    __syncthreads();
    const TakenLds taken{bits};
    merge_commit(taken, n, x, z, c, s, u, w, uvs, wvs, cand, list, nlist, jn, jl, fs, cn, cl, thr,
                 rej, g, L, out);
    __syncthreads();
    for (int i = lane; i < n; i += kW) removed[i] = taken.get(i) ? 1 : 0;
}

__global__ void merge_keep_kernel(int n, const int* __restrict__ removed, int* __restrict__ keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keep[i] = removed[i] ? 0 : 1;
}

// src/VortexBlobs.cpp:177-205: survivors in order, IDs 0..m-1.
__global__ void merge_compact_kernel(int n, const int* __restrict__ removed,
                                     const int* __restrict__ pos, BlobBuf in, BlobBuf out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || removed[i]) return;
    const int o = pos[i];
    out.x[o] = in.x[i]; out.z[o] = in.z[i]; out.c[o] = in.c[i]; out.s[o] = in.s[i];
    out.u[o] = in.u[i]; out.w[o] = in.w[i]; out.uvs[o] = in.uvs[i]; out.wvs[o] = in.wvs[i];
    if (out.id) out.id[o] = (uint32_t)o;
}

__global__ void merge_bbox_kernel(int n, const double* __restrict__ x, const double* __restrict__ z,
                                  double* __restrict__ part)
{
    __shared__ double r[4][256];
    double a0 = INFINITY, a1 = -INFINITY, a2 = INFINITY, a3 = -INFINITY;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double xi = x[i], zi = z[i];
        if (isfinite(xi) && isfinite(zi)) {
            a0 = fmin(a0, xi);
            a1 = fmax(a1, xi);
            a2 = fmin(a2, zi);
            a3 = fmax(a3, zi);
        }
    }
    const int t = threadIdx.x;
    r[0][t] = a0; r[1][t] = a1; r[2][t] = a2; r[3][t] = a3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            r[0][t] = fmin(r[0][t], r[0][t + h]);
            r[1][t] = fmax(r[1][t], r[1][t + h]);
            r[2][t] = fmin(r[2][t], r[2][t + h]);
            r[3][t] = fmax(r[3][t], r[3][t + h]);
        }
        __syncthreads();
    }
    if (t < 4) part[blockIdx.x * 4 + t] = r[t][0];
}

int max_cells()
{
    int total = 0, side = kGridSide;
    for (int l = 0; l < kLevels; ++l) {
        total += side * side;
        side = side / 8 + 2;
    }
    return total;
}

struct MergeLayout {
    size_t ib, cb, sort_bytes, scan_bytes, total;
};

MergeLayout merge_layout(int64_t n)
{
    MergeLayout L{};
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    L.ib = al(sizeof(int) * (size_t)n);
    L.cb = al(sizeof(int) * ((size_t)max_cells() + 3));
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, L.scan_bytes, (int*)nullptr, (int*)nullptr,
                                                   (int)n, (hipStream_t)0));
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, L.sort_bytes, (int*)nullptr, (int*)nullptr,
                                                     (int*)nullptr, (int*)nullptr, (int)n, 0, 24,
                                                     (hipStream_t)0));
    // cand removed keep pos key key2 idx sidx list jn cn | jl | cl | fs | cbeg | counts | bbox |
    // scan | sort
    L.total = 11 * L.ib + al(sizeof(int) * kSpecMax * (size_t)n) +
              al(sizeof(int) * kCandMax * (size_t)n) + al(sizeof(double) * 4 * (size_t)n) + L.cb +
              256 + al(sizeof(double) * 4 * kBboxBlocks) + al(L.scan_bytes) + al(L.sort_bytes);
    return L;
}

}  // namespace

size_t merge_scratch_bytes(int64_t n) { return merge_layout(n < 2 ? 2 : n).total; }

int64_t merge_run(dvm_ctx_t ctx, double threshold, int64_t n, const BlobBuf& b, const BlobBuf& out,
                  void* scratch)
{
    if (n < 2) return 0;  // src/VortexBlobs.cpp:112-114
    require(n < INT_MAX / (kSpecMax + kCandMax), DVM_E_SIZE, "merge_blobs: too many blobs");
    hipStream_t st = ctx->stream;
    const MergeLayout L = merge_layout(n);
    char* p = static_cast<char*>(scratch);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) & ~size_t(255);
        return q;
    };
    int* cand = reinterpret_cast<int*>(take(L.ib));
    int* removed = reinterpret_cast<int*>(take(L.ib));
    int* keep = reinterpret_cast<int*>(take(L.ib));
    int* pos = reinterpret_cast<int*>(take(L.ib));
    int* key = reinterpret_cast<int*>(take(L.ib));
    int* key2 = reinterpret_cast<int*>(take(L.ib));
    int* idx = reinterpret_cast<int*>(take(L.ib));
    int* sidx = reinterpret_cast<int*>(take(L.ib));
    int* list = reinterpret_cast<int*>(take(L.ib));
    int* jn = reinterpret_cast<int*>(take(L.ib));
    int* cn = reinterpret_cast<int*>(take(L.ib));
    int* jl = reinterpret_cast<int*>(take(sizeof(int) * kSpecMax * (size_t)n));
    int* cl = reinterpret_cast<int*>(take(sizeof(int) * kCandMax * (size_t)n));
    double* fs = reinterpret_cast<double*>(take(sizeof(double) * 4 * (size_t)n));
    int* cbeg = reinterpret_cast<int*>(take(L.cb));
    int* counts = reinterpret_cast<int*>(take(256));  // [0] list length, [2] merges, [3] recomputed
    double* bbox = reinterpret_cast<double*>(take(sizeof(double) * 4 * kBboxBlocks));
    void* scan_tmp = take(L.scan_bytes);
    void* sort_tmp = take(L.sort_bytes);

    const double rej = 4.0 * threshold * threshold * (1.0 + 1e-9);
    {
        ProfScope ps(ctx, "merge_cand_kernel");
        merge_cand_kernel<<<cdiv(n, kCandThreads), kCandThreads, 0, st>>>(
            (int)n, b.x, b.z, b.c, threshold, rej, cand, removed);
        DVM_LAUNCH_CHECK();
    }
    // candidate list (index order)
    merge_flag_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, cand, keep);
    DVM_LAUNCH_CHECK();
    size_t scb = L.scan_bytes;
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scb, keep, pos, (int)n, st));
    merge_list_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, keep, pos, list, counts);
    DVM_LAUNCH_CHECK();

    // cell lists over the finite positions
    merge_bbox_kernel<<<kBboxBlocks, 256, 0, st>>>((int)n, b.x, b.z, bbox);
    DVM_LAUNCH_CHECK();
    double* hb = static_cast<double*>(pinned_get(ctx, sizeof(double) * 4 * kBboxBlocks));
    DVM_HIP_CHECK(hipMemcpyAsync(hb, bbox, sizeof(double) * 4 * kBboxBlocks, hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    double x0 = INFINITY, x1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY;
    for (int k = 0; k < kBboxBlocks; ++k) {
        x0 = std::fmin(x0, hb[4 * k]);
        x1 = std::fmax(x1, hb[4 * k + 1]);
        z0 = std::fmin(z0, hb[4 * k + 2]);
        z1 = std::fmax(z1, hb[4 * k + 3]);
    }
    if (!(x0 <= x1)) x0 = x1 = z0 = z1 = 0.0;  // no finite position
    MergeGrid g{};
    g.T = 2.0 * threshold * (1.0 + 1e-9);
    g.x0 = x0;
    g.z0 = z0;
    const double W = x1 - x0, H = z1 - z0, ext = std::fmax(W, H);
    // finest cells: about one blob per cell at the mean density of the box
    double h = 1.0;
    if (ext > 0.0) {
        const double area = std::fmax(W, ext / 1024) * std::fmax(H, ext / 1024);
        h = std::fmax(0.75 * std::sqrt(area / (double)n), ext / (kGridSide - 1));
    }
    int off = 0;
    for (int l = 0; l < kLevels; ++l) {
        g.h[l] = h;
        g.reach_le[l] = g.T / h * (1.0 + 1e-9);
        g.nx[l] = std::min((int)std::floor(W / h) + 1, kGridSide);
        g.nz[l] = std::min((int)std::floor(H / h) + 1, kGridSide);
        g.off[l] = off;
        off += g.nx[l] * g.nz[l];
        h *= 8.0;
    }
    g.ncell = off;
    require(g.ncell <= max_cells(), DVM_E_STATE, "merge_blobs: grid larger than its scratch");
    merge_key_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, b.x, b.z, b.c, g, key, idx);
    DVM_LAUNCH_CHECK();
    int bits = 1;
    while ((1 << bits) < g.ncell + 2) ++bits;
    size_t sb = 0;
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, key, key2, idx, sidx, (int)n, 0,
                                                     bits, st));
    require(sb <= L.sort_bytes, DVM_E_STATE, "merge_blobs: sort scratch too small");
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sb, key, key2, idx, sidx, (int)n, 0,
                                                     bits, st));
    merge_cbeg_kernel<<<cdiv(g.ncell + 3, 256), 256, 0, st>>>((int)n, key2, g.ncell, cbeg);
    DVM_LAUNCH_CHECK();
    g.sidx = sidx;
    g.cbeg = cbeg;
    {
        ProfScope ps(ctx, "merge_spec_kernel");
        merge_spec_kernel<<<kSpecBlocks, 256, 0, st>>>((int)n, b.x, b.z, b.c, b.s, cand, list, counts,
                                                      threshold, rej, g, jn, jl, fs, cn, cl);
        DVM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(ctx, "merge_commit_kernel");
        if (n <= 32 * (int64_t)kTakenWords)
            merge_commit_lds_kernel<<<1, kW, 0, st>>>((int)n, b.x, b.z, b.c, b.s, b.u, b.w, b.uvs,
                                                      b.wvs, cand, list, counts, jn, jl, fs, cn, cl,
                                                      removed, threshold, rej, g, counts + 2);
        else
            merge_commit_kernel<<<1, kW, 0, st>>>((int)n, b.x, b.z, b.c, b.s, b.u, b.w, b.uvs, b.wvs,
                                                  cand, list, counts, jn, jl, fs, cn, cl, removed,
                                                  threshold, rej, g, counts + 2);
        DVM_LAUNCH_CHECK();
    }
    int* hc = static_cast<int*>(pinned_get(ctx, sizeof(int) * 4));
    DVM_HIP_CHECK(hipMemcpyAsync(hc, counts, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    const int64_t merged = hc[2];
    ctx->merge_stats[0] = hc[0];
    ctx->merge_stats[1] = hc[3];
    if (merged == 0) return 0;  // src/VortexBlobs.cpp:173-175: nothing rebuilt
    merge_keep_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, removed, keep);
    DVM_LAUNCH_CHECK();
    scb = L.scan_bytes;
    DVM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scb, keep, pos, (int)n, st));
    merge_compact_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, removed, pos, b, out);
    DVM_LAUNCH_CHECK();
    return merged;
}

}  // namespace dvmi

using namespace dvmi;

extern "C" {

dvm_status dvm_merge_blobs(dvm_ctx_t ctx, double threshold, int64_t n, uint32_t* id, double* x,
                           double* z, double* circ, double* sigma, double* u, double* w,
                           double* uvs, double* wvs, int64_t* merged)
{
    if (!ctx || !merged) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative blob count");
        require(n == 0 || (x && z && circ && sigma && u && w && uvs && wvs), DVM_E_INVALID,
                "null blob array");
        *merged = 0;
        if (n < 2) return;
        // scratch: the merge's work arrays, then a compacted copy of the blobs
        const size_t sb = (merge_scratch_bytes(n) + 255) & ~size_t(255);
        const size_t db = (sizeof(double) * (size_t)n + 255) & ~size_t(255);
        const size_t ub = (sizeof(uint32_t) * (size_t)n + 255) & ~size_t(255);
        char* p = static_cast<char*>(ws_get(ctx, sb + 8 * db + ub));
        BlobBuf in{x, z, circ, sigma, u, w, uvs, wvs, id};
        BlobBuf out;
        double** outs[8] = {&out.x, &out.z, &out.c, &out.s, &out.u, &out.w, &out.uvs, &out.wvs};
        for (int k = 0; k < 8; ++k) *outs[k] = reinterpret_cast<double*>(p + sb + k * db);
        out.id = id ? reinterpret_cast<uint32_t*>(p + sb + 8 * db) : nullptr;
        const int64_t k = merge_run(ctx, threshold, n, in, out, p);
        *merged = k;
        if (k == 0) return;
        const int64_t m = n - k;
        hipStream_t st = ctx->stream;
        double* ins[8] = {x, z, circ, sigma, u, w, uvs, wvs};
        for (int q = 0; q < 8; ++q)
            DVM_HIP_CHECK(hipMemcpyAsync(ins[q], *outs[q], sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        if (id)
            DVM_HIP_CHECK(hipMemcpyAsync(id, out.id, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, st));
    });
}

dvm_status dvm_merge_stats(dvm_ctx_t ctx, int64_t* candidates, int64_t* recomputed)
{
    if (!ctx) return DVM_E_INVALID;
    if (candidates) *candidates = ctx->merge_stats[0];
    if (recomputed) *recomputed = ctx->merge_stats[1];
    return DVM_OK;
}

}  // extern "C"
