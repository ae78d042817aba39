// Batched ensemble: many independent simulations (bodies, viscosities, step
// counts, blob counts) advanced together.
//
// Storage: slot s owns blob range [s*cap, (s+1)*cap) of one SoA and panel range
// [s*mp, (s+1)*mp) of the sheet arrays; live counts, step counters and status
// words live on the device, so a whole time step is a fixed sequence of
// launches that reads its sizes from device memory — it is captured once into
// a hipGraph and replayed.  Every launch covers all slots: per-blob kernels
// stride over the concatenated live ranges (slot found by binary search of a
// device prefix), per-slot kernels use one workgroup per slot, and the
// pair sum is one persistent launch whose waves (FP32C) or workgroups (FP64)
// pull work items from a device work counter, largest slots first.  Each
// item's sum order is fixed, so results do not depend on which wave ran it or
// on what else is in the batch.
#include "dvm_device.hpp"
#include "dvm_internal.hpp"


#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dvmi {

constexpr int kBT = 256;       // sources per tile; targets per FP64 item
constexpr int kColChunks = 256;   // most source chunks of the collocation sum per slot
constexpr int kColSources = 256;  // sources per collocation chunk, at least
constexpr int kTB = 64;         // targets per FP32C item (one wave)
constexpr int kMaxChunks = 64;  // most source chunks per slot in the FP32C sum
constexpr int kBaseChunks = 4;  // chunks of every slot large enough for them
constexpr int kSlotItems = 65536;  // FP32C items a slot is cut into, at least
constexpr int kMinChunkTiles = 2;  // a chunk spans at least this many tiles
constexpr int kSortBits = 24;   // Morton key bits sorted (12 per axis)
constexpr int kSortTile = 4096; // keys per sort workgroup (256 lanes x 16)
constexpr int kBoxParts = 32;   // workgroups per slot of the bounding-box pass
constexpr int kXcds = 8;        // accelerator dies (XCDs) of the device
constexpr int kAdmitStreams = 16;  // bodies set up at once by dvm_batch_admit_many
constexpr int kFmmStreams = 8;     // fast multipole sums of large slots at once

struct SlotParam {
    double rho, nu, dt, ux, uz, ur, aref, four_nu_dt;
    int nl, max_npv, steps_target, pad0;
    uint32_t seed, pad1;
    double pad2;
    BodyBox box;  // crossing-test cull box of the body
    int nslab, pad3;              // crossing-test slab lists (EdgeSlabs; 0: none)
    double slab_z0, slab_inv_h;
};

constexpr int kSlabCap = 16;  // slab list entries per panel a slot holds

struct SlotState {
    int64_t n;       // live blobs
    int step;        // completed steps
    int active;      // 1 while stepping
    int status;      // dvm_status of the failure that stopped the slot, else 0
    int nrv;         // nascent blobs of the current step
    double loads[4]; // fx fz cd cl of the last step
    int ext;         // 1: this step's blob velocities were summed before the
                     // convection graph (fast multipole); its pair sum skips the slot
};

// Work lists rebuilt by bt_prefix_kernel at every velocity evaluation.
struct WorkList {
    int total_tiles;
    int work_counter;
    int64_t total_blobs;
    int total_items;  // FP32C work items (slot, chunk, 64-target block)
    int total_stiles; // sort tiles (kSortTile keys of one slot)
    // FP32C item counters, one per XCD (128 bytes apart): the waves of XCD x
    // take items x, x + 8, x + 16, ...
    int xcd_counter[kXcds * 32];
};

struct BatchDev {
    int S = 0;
    int64_t cap = 0;
    int mp = 0;
    int max_tiles = 0;
    SlotParam* prm = nullptr;
    SlotState* st = nullptr;
    WorkList* wl = nullptr;
    int* tile_pre = nullptr;    // S+1, slot order
    int* order = nullptr;       // S, slots sorted by tile count (desc)
    int* otile_pre = nullptr;   // S+1, in `order`
    int64_t* blob_pre = nullptr;  // S+1
    BlobBuf b, alt;
    double *q1 = nullptr, *q2 = nullptr;
    float4* s32 = nullptr;      // sorted order
    float4* near32 = nullptr;   // sorted order
    double2* p64 = nullptr;     // sorted FP64 positions
    TileInfo* tiles = nullptr;  // S * max_tiles (compact, slot order)
    // FP32C sum: per-slot Morton sort and chunked work items
    int* item_pre = nullptr;    // S+1, in `order`
    double4* sbox = nullptr;    // S bounding boxes
    double4* box_part = nullptr;  // S * kBoxParts partial boxes
    uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *perm = nullptr;  // S*cap
    int* stile_pre = nullptr;   // S+1, sort tiles per slot (slot order)
    int* shist = nullptr;       // 256 digit counts per sort tile, S * max_stiles
    int max_stiles = 0;
    TileInfo* tboxes = nullptr; // S * max_tboxes, 64-target boxes
    int max_tboxes = 0;
    double* part = nullptr;     // S * part_slot chunk partials
    int64_t part_slot = 0;      // partials per slot: 2 * max over n <= cap of n * chunks(n)
    int* flags = nullptr;
    // sheet
    double *bx = nullptr, *bz = nullptr;  // S*(mp+1)
    double *enx = nullptr, *enz = nullptr, *etx = nullptr, *etz = nullptr, *ds = nullptr;
    PanelConst* panel = nullptr;
    double* P = nullptr;        // S * mp*(mp+1)
    double* gamma = nullptr;
    double* col = nullptr;      // S * kColChunks * (2*mp + 1) collocation partials
    int* cnt = nullptr;
    int64_t* off = nullptr;     // S*(mp+1)
    double* cnew = nullptr;
    double* lwork = nullptr;    // S * 3*mp
    double* rhs = nullptr;      // S * (mp+1) right-hand sides of the gamma solve
    int* slab_off = nullptr;    // S * (mp+1)
    int* slab_edge = nullptr;   // S * kSlabCap*mp
    // FP32C sheet sum by panel groups (sheet_groups): mg groups per slot
    int mg = 0;
    double4* gcen = nullptr;    // S * mg
    double *gI0 = nullptr, *gI1 = nullptr;  // S * mp * 2*kSheetTerms
    double* gA = nullptr;       // S * mg * 2*kSheetTerms moments of the current gamma
};

// ----------------------------------------------------------------- work lists

// FP32C source chunks of a slot with n live blobs: enough (slot, chunk,
// 64-target block) items that a slot alone keeps the device busy, at least
// kBaseChunks, each chunk at least kMinChunkTiles tiles.  A function of n
// alone, so a slot's sum does not depend on the rest of the batch.
__host__ __device__ __forceinline__ int slot_chunks(int64_t n)
{
    const int ntl = (int)((n + kBT - 1) / kBT);
    const int nblk = (int)((n + kTB - 1) / kTB);
    int c = nblk > 0 ? (kSlotItems + nblk - 1) / nblk : 1;
    c = c > kBaseChunks ? c : kBaseChunks;
    const int cmax = ntl / kMinChunkTiles;
    c = c > cmax ? cmax : c;
    c = c > kMaxChunks ? kMaxChunks : c;
    return c < 1 ? 1 : c;
}

// Collocation source chunks of a slot with n live blobs (a function of n).
__host__ __device__ __forceinline__ int colloc_chunks(int64_t n)
{
    const int64_t c = (n + kColSources - 1) / kColSources;
    return c < 1 ? 1 : (c > kColChunks ? kColChunks : (int)c);
}

// Chunk partials a slot of capacity cap may need: 2 * max n * slot_chunks(n)
// over n <= cap (the chunk count is constant between multiples of kTB).
static int64_t part_slot_size(int64_t cap)
{
    // beyond kSlotItems * kTB blobs the count is kBaseChunks: n * chunks grows
    int64_t m = cap * slot_chunks(cap);
    for (int64_t n = kTB; n < cap && n <= (int64_t)kSlotItems * kTB; n += kTB) {
        const int64_t v = n * slot_chunks(n);
        m = v > m ? v : m;
    }
    return 2 * m;
}

// Exclusive prefix sum of v over the workgroup's lanes (up to 1024 lanes,
// 16 waves): wave scans by shuffles, then the wave totals; total gets the sum.
__device__ __forceinline__ int64_t block_scan_excl(int64_t v, int64_t* wsum, int64_t& total)
{
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int64_t off = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const int64_t t = wsum[k];
        off += k < wv ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// One workgroup, a lane per slot (prefix_lanes(S) lanes): per-slot
// tile and blob counts, their prefixes, the slot order by decreasing tile
// count (ties by index), the FP32C item and sort-tile prefixes, and the
// work-counter reset.
// Lanes of the prefix workgroup: whole waves covering the slots.
static unsigned prefix_lanes(int S) { return (unsigned)((S + 63) / 64 * 64); }

__global__ __launch_bounds__(1024) void bt_prefix_kernel(int S, int64_t cap,
                                                         const SlotState* __restrict__ st,
                                                         WorkList* wl, int* tile_pre, int* order,
                                                         int* otile_pre, int64_t* blob_pre,
                                                         int* item_pre, int* stile_pre)
{
    __shared__ int tiles_s[1024];
    __shared__ int64_t n_s[1024];
    __shared__ int ord_s[1024];
    __shared__ int64_t wsum[16];
    const int t = threadIdx.x;
    const bool slot = t < S;
    const int64_t n = slot && st[t].active ? st[t].n : 0;
    const int tiles = (int)((n + kBT - 1) / kBT);
    tiles_s[t] = tiles;
    n_s[t] = n;
    __syncthreads();
    int r = 0;
    if (slot) {
        for (int o = 0; o < S; ++o)
            r += (tiles_s[o] > tiles) || (tiles_s[o] == tiles && o < t);
        order[r] = t;
        ord_s[r] = t;
    }
    __syncthreads();
    int64_t tot_tiles, tot_blobs, tot_stiles, tot_otiles, tot_items;
    const int64_t tp = block_scan_excl(tiles, wsum, tot_tiles);
    const int64_t bp = block_scan_excl(n, wsum, tot_blobs);
    const int64_t sp = block_scan_excl((n + kSortTile - 1) / kSortTile, wsum, tot_stiles);
    // in slot order: lane t holds the slot of rank t
    const int64_t no = slot ? n_s[ord_s[t]] : 0;
    const int64_t op = block_scan_excl((no + kBT - 1) / kBT, wsum, tot_otiles);
    const int64_t items = no > 0 ? ((no + kTB - 1) / kTB) * slot_chunks(no) : 0;
    const int64_t ip = block_scan_excl(items, wsum, tot_items);
    if (slot) {
        tile_pre[t] = (int)tp;
        blob_pre[t] = bp;
        stile_pre[t] = (int)sp;
        otile_pre[t] = (int)op;
        item_pre[t] = (int)ip;
    }
    if (t == 0) {
        tile_pre[S] = (int)tot_tiles;
        blob_pre[S] = tot_blobs;
        stile_pre[S] = (int)tot_stiles;
        otile_pre[S] = (int)tot_otiles;
        item_pre[S] = (int)tot_items;
        wl->total_tiles = (int)tot_tiles;
        wl->work_counter = 0;
        for (int x = 0; x < kXcds; ++x) wl->xcd_counter[x * 32] = 0;
        wl->total_blobs = tot_blobs;
        wl->total_stiles = (int)tot_stiles;
        wl->total_items = (int)tot_items;
    }
}

template <class T>
__device__ __forceinline__ int find_slot(const T* __restrict__ pre, int S, T g)
{
    int lo = 0, hi = S;  // pre[lo] <= g < pre[lo+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Grid-stride loop over the live blobs of all slots: calls f(slot, blob index).
template <class F>
__device__ __forceinline__ void for_each_blob(const WorkList* wl, const int64_t* blob_pre, int S,
                                              int64_t cap, F&& f)
{
    const int64_t total = wl->total_blobs;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_slot(blob_pre, S, g);
        f(s, (int64_t)s * cap + (g - blob_pre[s]));
    }
}

// ----------------------------------------------------------------- pair sums
//
// FP32C (the form of bs_kernels.hip, per slot): each slot's live blobs are
// Morton-sorted (bounding box, 24-bit keys, a radix sort with several
// workgroups per slot), cut into 256-source tiles stored as FP32 offsets
// from their FP64 centres, and summed by a
// persistent kernel whose waves pull (slot, source chunk, 64-target block)
// items, largest slots first and chunk by chunk, so co-resident waves stream
// the same source slice.  Each item writes FP64 chunk partials; a reduce adds
// them in chunk order and scatters to the storage order.  Chunking and
// order depend on the slot's own count only.  FP64: one 256-target item per
// workgroup over all sources, in index order (the reference's summation
// order).

// Per-slot bounding box of the live blobs: kBoxParts workgroups per slot
// each reduce a contiguous part (min and max are exact, so the box does not
// depend on the split), then one lane per slot combines the parts.
__global__ __launch_bounds__(256) void bt_bbox_kernel(int64_t cap, const SlotState* __restrict__ st,
                                                      BlobBuf b, double4* __restrict__ box_part)
{
    __shared__ double4 red[256];
    const int s = blockIdx.y;
    const int64_t n = st[s].active ? st[s].n : 0;
    const int64_t len = (n + kBoxParts - 1) / kBoxParts;
    const int64_t j0 = (int64_t)blockIdx.x * len;
    const int64_t j1 = j0 + len < n ? j0 + len : n;
    const int64_t base = (int64_t)s * cap;
    double4 m = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int64_t k = j0 + threadIdx.x; k < j1; k += 256) {
        const double x = b.x[base + k], z = b.z[base + k];
        m.x = fmin(m.x, x); m.y = fmax(m.y, x);
        m.z = fmin(m.z, z); m.w = fmax(m.w, z);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double4 o = red[threadIdx.x + w];
            double4& r = red[threadIdx.x];
            r.x = fmin(r.x, o.x); r.y = fmax(r.y, o.y);
            r.z = fmin(r.z, o.z); r.w = fmax(r.w, o.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) box_part[(int64_t)s * kBoxParts + blockIdx.x] = red[0];
}

__global__ void bt_bbox_final_kernel(int S, const double4* __restrict__ box_part,
                                     double4* __restrict__ sbox)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double4 r = box_part[(int64_t)s * kBoxParts];
    for (int p = 1; p < kBoxParts; ++p) {
        const double4 o = box_part[(int64_t)s * kBoxParts + p];
        r.x = fmin(r.x, o.x); r.y = fmax(r.y, o.y);
        r.z = fmin(r.z, o.z); r.w = fmax(r.w, o.w);
    }
    sbox[s] = r;
}

// Morton key (12 bits per axis over the slot's box) and local index.
__global__ void bt_morton_kernel(int S, int64_t cap, const WorkList* wl,
                                 const int64_t* __restrict__ blob_pre,
                                 const double4* __restrict__ sbox, BlobBuf b,
                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int s, int64_t o) {
        const double4 bb = sbox[s];
        const double ext = fmax(fmax(bb.y - bb.x, bb.w - bb.z), 1e-300);
        const double scale = 4095.0 / ext;
        double qx = (b.x[o] - bb.x) * scale;
        double qz = (b.z[o] - bb.z) * scale;
        qx = qx >= 0.0 ? (qx <= 4095.0 ? qx : 4095.0) : 0.0;  // NaN -> 0
        qz = qz >= 0.0 ? (qz <= 4095.0 ? qz : 4095.0) : 0.0;
        keys[o] = spread16((uint32_t)qx) | (spread16((uint32_t)qz) << 1);
        vals[o] = (uint32_t)(o - (int64_t)s * cap);
    });
}

// Stable LSD radix sort of every slot's live keys by 8-bit digits, several
// workgroups per slot: each sort tile (kSortTile consecutive keys of one slot)
// counts its digits; per slot the counts are scanned digit-major over the
// slot's tiles, giving every (tile, digit) its first output rank; each tile
// then ranks its keys stably (wave by wave in key order, equal digits by lane
// ballots) and scatters keys and values into the slot's range.  Counts come
// from the device, so the sort is captured once and replays at any size; its
// output equals any stable sort's.

__global__ __launch_bounds__(256) void bt_sort_hist_kernel(int S, int64_t cap, const WorkList* wl,
                                                           const int* __restrict__ stile_pre,
                                                           const SlotState* __restrict__ st,
                                                           const uint32_t* __restrict__ keys,
                                                           int shift, int* __restrict__ hist)
{
    __shared__ int cnt[256];
    const int total = wl->total_stiles;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        cnt[threadIdx.x] = 0;
        __syncthreads();
        const int s = find_slot(stile_pre, S, g);
        const int64_t n = st[s].n;
        const int64_t i0 = (int64_t)(g - stile_pre[s]) * kSortTile;
        const int64_t i1 = i0 + kSortTile < n ? i0 + kSortTile : n;
        const uint32_t* __restrict__ k = keys + (int64_t)s * cap;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
            atomicAdd(&cnt[(k[i] >> shift) & 255u], 1);
        __syncthreads();
        hist[(int64_t)g * 256 + threadIdx.x] = cnt[threadIdx.x];
        __syncthreads();
    }
}

// One workgroup per slot, one lane per digit.
__global__ __launch_bounds__(256) void bt_sort_scan_kernel(const int* __restrict__ stile_pre,
                                                           int* __restrict__ hist)
{
    __shared__ int tot[256];
    const int s = blockIdx.x;
    const int d = threadIdx.x;
    const int g0 = stile_pre[s], g1 = stile_pre[s + 1];
    if (g0 == g1) return;
    int run = 0;
    for (int g = g0; g < g1; ++g) {
        const int v = hist[(int64_t)g * 256 + d];
        hist[(int64_t)g * 256 + d] = run;
        run += v;
    }
    tot[d] = run;
    __syncthreads();
    for (int w = 1; w < 256; w <<= 1) {  // inclusive scan over digits
        const int v = d >= w ? tot[d - w] : 0;
        __syncthreads();
        tot[d] += v;
        __syncthreads();
    }
    const int base = tot[d] - run;
    for (int g = g0; g < g1; ++g) hist[(int64_t)g * 256 + d] += base;
}

constexpr int kSortPer = kSortTile / 256;  // keys per lane

__global__ __launch_bounds__(256) void bt_sort_scatter_kernel(
    int S, int64_t cap, const WorkList* wl, const int* __restrict__ stile_pre,
    const SlotState* __restrict__ st, const uint32_t* __restrict__ kin,
    const uint32_t* __restrict__ vin, int shift, const int* __restrict__ hist,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout)
{
    __shared__ int wcnt[4][256];  // per wave: digit counts, then offsets
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int total = wl->total_stiles;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        for (int k = threadIdx.x; k < 4 * 256; k += 256) (&wcnt[0][0])[k] = 0;
        __syncthreads();
        const int s = find_slot(stile_pre, S, g);
        const int64_t n = st[s].n;
        const int64_t base = (int64_t)s * cap;
        // wave w ranks keys [w * kSortTile / 4, (w + 1) * kSortTile / 4) of the tile
        const int64_t i0 = (int64_t)(g - stile_pre[s]) * kSortTile + w * (kSortTile / 4);
        uint32_t kk[kSortPer], vv[kSortPer];
        int dig[kSortPer], rk[kSortPer];
#pragma unroll
        for (int j = 0; j < kSortPer; ++j) {
            const int64_t i = i0 + j * 64 + lane;
            const bool valid = i < n;
            const uint32_t key = valid ? kin[base + i] : 0u;
            const int dd = (int)((key >> shift) & 255u);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                const uint64_t m = __ballot((dd >> bit) & 1);
                peers &= ((dd >> bit) & 1) ? m : ~m;
            }
            int r = -1;
            if (valid) r = wcnt[w][dd] + (int)__popcll(peers & below);
            // the lowest lane of each digit group advances the wave's count
            if (valid && (peers & below) == 0) wcnt[w][dd] += (int)__popcll(peers);
            kk[j] = key;
            vv[j] = valid ? vin[base + i] : 0u;
            dig[j] = dd;
            rk[j] = r;
        }
        __syncthreads();
        {
            const int d = threadIdx.x;  // one lane per digit: offsets of the waves
            int a = 0;
            for (int v = 0; v < 4; ++v) {
                const int c = wcnt[v][d];
                wcnt[v][d] = a;
                a += c;
            }
        }
        __syncthreads();
        const int* __restrict__ ho = hist + (int64_t)g * 256;
#pragma unroll
        for (int j = 0; j < kSortPer; ++j) {
            if (rk[j] < 0) continue;
            const int64_t dst = base + ho[dig[j]] + wcnt[w][dig[j]] + rk[j];
            kout[dst] = kk[j];
            vout[dst] = vv[j];
        }
        __syncthreads();
    }
}

// One workgroup per 256-source tile (sorted order): tile record, FP32
// offsets, near-field constants, sorted FP64 target positions, and the four
// 64-target boxes of the tile.
__global__ __launch_bounds__(kBT) void bt_tiles_kernel(int S, int64_t cap, int max_tboxes,
                                                       const WorkList* wl,
                                                       const int* __restrict__ tile_pre,
                                                       const SlotState* __restrict__ st,
                                                       const uint32_t* __restrict__ perm,
                                                       BlobBuf b, float4* __restrict__ s32,
                                                       float4* __restrict__ near32,
                                                       double2* __restrict__ p64,
                                                       TileInfo* __restrict__ tiles,
                                                       TileInfo* __restrict__ tboxes)
{
    __shared__ double red[6][kBT];
    const int total = wl->total_tiles;
    const int lane = threadIdx.x & 63;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int s = find_slot(tile_pre, S, g);
        const int64_t n = st[s].n;
        const int tile = g - tile_pre[s];
        const int64_t k = (int64_t)tile * kBT + threadIdx.x;  // sorted rank
        const bool valid = k < n;
        const int64_t base = (int64_t)s * cap;
        const int64_t src = valid ? base + perm[base + k] : base;
        const double xi = valid ? b.x[src] : 0.0, zi = valid ? b.z[src] : 0.0;
        const double si = valid ? b.s[src] : 0.0;
        // 64-target box of this wave
        double mnx = valid ? xi : INFINITY, mxx = valid ? xi : -INFINITY;
        double mnz = valid ? zi : INFINITY, mxz = valid ? zi : -INFINITY;
        for (int o = 32; o > 0; o >>= 1) {
            mnx = fmin(mnx, __shfl_xor(mnx, o));
            mxx = fmax(mxx, __shfl_xor(mxx, o));
            mnz = fmin(mnz, __shfl_xor(mnz, o));
            mxz = fmax(mxz, __shfl_xor(mxz, o));
        }
        const int64_t w0 = (int64_t)tile * kBT + (threadIdx.x & ~63);
        if (lane == 0 && w0 < n) {
            TileInfo tb;
            tb.cx = 0.5 * (mnx + mxx);
            tb.cz = 0.5 * (mnz + mxz);
            tb.hx = (float)(0.5 * (mxx - mnx)) * 1.000001f + 1e-30f;
            tb.hz = (float)(0.5 * (mxz - mnz)) * 1.000001f + 1e-30f;
            tb.thr_max = 0.0f;
            tb.thr_min = 0.0f;
            tboxes[(int64_t)s * max_tboxes + w0 / kTB] = tb;
        }
        red[0][threadIdx.x] = valid ? xi : INFINITY;
        red[1][threadIdx.x] = valid ? xi : -INFINITY;
        red[2][threadIdx.x] = valid ? zi : INFINITY;
        red[3][threadIdx.x] = valid ? zi : -INFINITY;
        red[4][threadIdx.x] = si * si;
        red[5][threadIdx.x] = valid ? si * si : INFINITY;
        __syncthreads();
        for (int w = kBT / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                const int p = threadIdx.x + w;
                red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][p]);
                red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][p]);
                red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][p]);
                red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][p]);
                red[4][threadIdx.x] = fmax(red[4][threadIdx.x], red[4][p]);
                red[5][threadIdx.x] = fmin(red[5][threadIdx.x], red[5][p]);
            }
            __syncthreads();
        }
        const double cx = 0.5 * (red[0][0] + red[1][0]);
        const double cz = 0.5 * (red[2][0] + red[3][0]);
        if (threadIdx.x == 0) {
            TileInfo ti;
            ti.cx = cx;
            ti.cz = cz;
            ti.hx = (float)(0.5 * (red[1][0] - red[0][0])) * 1.000001f + 1e-30f;
            ti.hz = (float)(0.5 * (red[3][0] - red[2][0])) * 1.000001f + 1e-30f;
            ti.thr_max = (float)(10.0 * red[4][0]);
            ti.thr_min = (float)(10.0 * red[5][0]);
            tiles[g] = ti;
        }
        if (valid) {
            const int64_t o = base + k;
            const double s2 = si * si;
            const float thr = (float)(10.0 * s2);
            const double h = 0.5 * fmax(red[1][0] - red[0][0], red[3][0] - red[2][0]);
            s32[o] = make_float4((float)(xi - cx), (float)(zi - cz), (float)b.c[src],
                                 near_thr_hi(thr, h, 10.0 * red[4][0]));
            const double ox = xi - cx, oz = zi - cz;
            near32[o] = make_float4(thr, (float)(1.0 / s2), (float)(ox - (double)(float)ox),
                                    (float)(oz - (double)(float)oz));
            p64[o] = make_double2(xi, zi);
        }
        __syncthreads();
    }
}

// FP32C pair sum, one wave per (slot, chunk, 64-target block) item.
__global__ __launch_bounds__(256) void bt_bs32_kernel(int S, int64_t cap, int max_tboxes,
                                                      WorkList* wl,
                                                      const int* __restrict__ tile_pre,
                                                      const int* __restrict__ order,
                                                      const int* __restrict__ item_pre,
                                                      const SlotState* __restrict__ st,
                                                      const float4* __restrict__ s32,
                                                      const float4* __restrict__ near32,
                                                      const double2* __restrict__ p64,
                                                      const TileInfo* __restrict__ tiles,
                                                      const TileInfo* __restrict__ tboxes,
                                                      int64_t part_slot, double* __restrict__ part)
{
    const int lane = threadIdx.x & 63;
    const int total = wl->total_items;
    // workgroups are dealt to the XCDs in turn: this one's XCD takes every
    // kXcds-th item, so the XCDs share no counter and stay in the same chunk
    const int xcd = blockIdx.x % kXcds;
    int* ctr = &wl->xcd_counter[xcd * 32];
    for (;;) {
        int k = 0;
        if (lane == 0) k = atomicAdd(ctr, 1);
        // wave-uniform (an SGPR), so the item's slot, tiles and source
        // records are scalar loads, as in the single-simulation kernel
        k = __builtin_amdgcn_readfirstlane(k);
        const int item = xcd + kXcds * k;
        if (item >= total) break;
        const int p = find_slot(item_pre, S, item);
        const int s = order[p];
        if (st[s].ext) continue;  // velocities already summed (dvm_batch_set_fmm)
        const int local = item - item_pre[p];
        const int64_t n = st[s].n;
        const int nblk = (int)((n + kTB - 1) / kTB);
        const int nch = slot_chunks(n);
        const int c = local / nblk;
        const int tbk = local - c * nblk;
        const int ntl = (int)((n + kBT - 1) / kBT);
        const int tpc = (ntl + nch - 1) / nch;
        const int s0 = c * tpc;
        const int s1 = s0 + tpc < ntl ? s0 + tpc : ntl;
        const int64_t base = (int64_t)s * cap;
        const int64_t i = (int64_t)tbk * kTB + lane;  // sorted rank
        const int64_t it = i < n ? i : n - 1;
        const double2 pi = p64[base + it];
        const double xi = pi.x, zi = pi.y;
        const TileInfo tb = tboxes[(int64_t)s * max_tboxes + tbk];
        const TileInfo* __restrict__ tsl = tiles + tile_pre[s];
        static_assert(kBT % kTB == 0, "a target block lies in one source tile");
        const int self_tile = tbk * kTB / kBT;
        float su = 0.0f, cu = 0.0f, sw = 0.0f, cw = 0.0f;
        for (int sti = s0; sti < s1; ++sti) {
            const TileInfo ts = tsl[sti];
            const double Dxd = ts.cx - tb.cx;
            const double Dzd = ts.cz - tb.cz;
            const double gx = fmax(fabs(Dxd) - (double)(tb.hx + ts.hx), 0.0);
            const double gz = fmax(fabs(Dzd) - (double)(tb.hz + ts.hz), 0.0);
            const double g2 = gx * gx + gz * gz;
            const bool near = g2 <= 1.0001 * (double)ts.thr_max + 1e-300;
            const float txd = (float)(xi - ts.cx);
            const float tzd = (float)(zi - ts.cz);
            const int64_t jr = (int64_t)sti * kBT;  // sorted rank of the tile's first source
            const int cnt = n - jr < kBT ? (int)(n - jr) : kBT;
            const float4* __restrict__ sp = s32 + base + jr;
            float au = 0.0f, aw = 0.0f;
            if (!near) {
                const double ex = fabs(Dxd) + (double)(tb.hx + ts.hx);
                const double ez = fabs(Dzd) + (double)(tb.hz + ts.hz);
                if (cnt == kBT && g2 > 1e-4 && ex * ex + ez * ez < 1e4)
                    far_tile_share8<kBT>(sp, txd, tzd, au, aw);
                else if (cnt == kBT && g2 > 1e-8 && ex * ex + ez * ez < 1e8)
                    far_tile_share4<kBT>(sp, txd, tzd, au, aw);
                else if (cnt == kBT)
                    far_tile<kBT>(sp, txd, tzd, au, aw);
                else
                    for (int k = 0; k < cnt; ++k) far_pair(sp[k], txd, tzd, au, aw);
            } else {
                const float txl = (float)((xi - ts.cx) - (double)txd);
                const float tzl = (float)((zi - ts.cz) - (double)tzd);
                // only the tile holding the block's own ranks has the self pair
                if (sti == self_tile) {
                    near_range<true, false>(sp, near32 + base + jr, 0, cnt, txd, tzd, txl, tzl, jr,
                                            i, 0.0f, au, aw);
                } else if (cnt == kBT) {
                    const double ex = fabs(Dxd) + (double)(tb.hx + ts.hx);
                    const double ez = fabs(Dzd) + (double)(tb.hz + ts.hz);
                    const bool share = ts.thr_min > 1e-8f && ex * ex + ez * ez < 1e8;
                    near_tile_share4<kBT>(sp, near32 + base + jr, txd, tzd, txl, tzl, share, au,
                                          aw);
                } else {
                    near_range<false, false>(sp, near32 + base + jr, 0, cnt, txd, tzd, txl, tzl, jr,
                                             i, 0.0f, au, aw);
                }
            }
            kahan_add(su, cu, au);
            kahan_add(sw, cw, aw);
        }
        if (i < n) {
            double* __restrict__ ps = part + (int64_t)s * part_slot + (int64_t)c * 2 * n;
            ps[i] = -((double)su - (double)cu);
            ps[n + i] = (double)sw - (double)cw;
        }
    }
}

// Adds a blob's chunk partials in chunk order, scales by 1/(2 pi) and writes
// the velocity at the blob's storage position.
__global__ void bt_reduce32_kernel(int S, int64_t cap, const WorkList* wl,
                                   const int64_t* __restrict__ blob_pre,
                                   const SlotState* __restrict__ st,
                                   const uint32_t* __restrict__ perm, int64_t part_slot,
                                   const double* __restrict__ part, BlobBuf b)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int s, int64_t o) {
        if (st[s].ext) return;  // velocities already summed (dvm_batch_set_fmm)
        const int64_t n = st[s].n;
        const int nch = slot_chunks(n);
        const double* __restrict__ ps = part + (int64_t)s * part_slot + (o - (int64_t)s * cap);
        double su = 0.0, sw = 0.0;
        for (int c = 0; c < nch; ++c) {
            su += ps[(int64_t)c * 2 * n];
            sw += ps[(int64_t)c * 2 * n + n];
        }
        const int64_t dst = (int64_t)s * cap + perm[o];
        b.u[dst] = su * kRpi2;
        b.w[dst] = sw * kRpi2;
    });
}

// FP64 pair sum: one workgroup per (slot, 256-target tile), sources in index
// order (src/VelocityKernel.cpp:37-61 per slot).
__global__ __launch_bounds__(kBT) void bt_bs64_kernel(int S, int64_t cap, WorkList* wl,
                                                      const int* __restrict__ order,
                                                      const int* __restrict__ otile_pre,
                                                      const SlotState* __restrict__ st, BlobBuf b)
{
    __shared__ int item_sh;
    const int total = wl->total_tiles;
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(&wl->work_counter, 1);
        __syncthreads();
        const int item = item_sh;
        __syncthreads();
        if (item >= total) break;
        const int p = find_slot(otile_pre, S, item);
        const int s = order[p];
        const int tt = item - otile_pre[p];
        const int64_t n = st[s].n;
        const int64_t base = (int64_t)s * cap;
        const int64_t i = (int64_t)tt * kBT + threadIdx.x;
        const int64_t it = i < n ? i : n - 1;
        const double xi = b.x[base + it];
        const double zi = b.z[base + it];
        double ui = 0.0, wi = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            bs_pair_f64(xi, zi, b.x[base + j], b.z[base + j], b.c[base + j], b.s[base + j], ui, wi);
        }
        if (i < n) {
            b.u[base + i] = ui * kRpi2;
            b.w[base + i] = wi * kRpi2;
        }
    }
}

// ----------------------------------------------------------------- sheet

// Collocation sums (unguarded) per (slot, source chunk): u, w partials for the
// slot's Nl panels and the chunk's circulation sum.  The chunk's sources are
// staged 256 at a time in LDS as records {x, z, circ, 10 sigma^2} and read by
// every lane (one panel per lane) in index order; the pair is the FP64 form's
// (src/VelocityKernel.cpp:44-60 with fused accumulation), four sources per
// step with their four 1/r^2 from one Newton-refined reciprocal of the product
// of the four r^2, the near-field correction only for a step with a near pair
// in some lane.
__global__ __launch_bounds__(256) void bt_colloc_kernel(int64_t cap, int mp,
                                                        const SlotParam* __restrict__ prm,
                                                        const SlotState* __restrict__ st, BlobBuf b,
                                                        const PanelConst* __restrict__ panel,
                                                        double* __restrict__ col)
{
    __shared__ double4 srec[256];
    __shared__ double red[256];
    const int s = blockIdx.y;
    const int c = blockIdx.x;
    if (!st[s].active) return;
    const int nl = prm[s].nl;
    const int64_t n = st[s].n;
    const int nch = colloc_chunks(n);
    if (c >= nch) return;
    const int64_t chunk = (n + nch - 1) / nch;
    const int64_t j0 = (int64_t)c * chunk;
    const int64_t j1 = j0 + chunk < n ? j0 + chunk : n;
    const int64_t base = (int64_t)s * cap;
    double* out = col + ((int64_t)s * kColChunks + c) * (2 * mp + 1);
    for (int p0 = 0; p0 < nl; p0 += 256) {  // wave-uniform
        const int pj = p0 + (int)threadIdx.x;
        const bool have = pj < nl;
        double xi = 0.0, zi = 0.0;
        if (have) {
            const PanelConst q = panel[(int64_t)s * mp + pj];
            xi = q.xc;
            zi = q.zc;
        }
        double ui = 0.0, wi = 0.0;
        for (int64_t t0 = j0; t0 < j1; t0 += 256) {
            const int64_t jt = t0 + threadIdx.x;
            __syncthreads();
            if (jt < j1) {
                const double sj = b.s[base + jt];
                srec[threadIdx.x] = make_double4(b.x[base + jt], b.z[base + jt], b.c[base + jt],
                                                 10.0 * (sj * sj));
            } else {
                srec[threadIdx.x] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
            __syncthreads();
            const int cnt = j1 - t0 < 256 ? (int)(j1 - t0) : 256;
            // a wavefront with no panel in this pass only stages sources
            const int kend = __builtin_amdgcn_ballot_w64(have) != 0 ? cnt : 0;
            for (int k = 0; k < kend; k += 4) {
                double dx[4], dz[4], r2[4], K[4];
                bool take[4];
                uint64_t nearm = 0;  // lanes with a near pair in the group
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double4 r = srec[k + q];
                    dx[q] = xi - r.x;
                    dz[q] = zi - r.y;
                    r2[q] = __builtin_fma(dx[q], dx[q], dz[q] * dz[q]);
                    take[q] = have && k + q < cnt;
                    nearm |= __builtin_amdgcn_ballot_w64(take[q] && r2[q] < r.w);
                }
                // the four 1/r2 from one reciprocal of their product, as in the
                // FP64 blob-on-blob sum (bs_f64_kernel)
                const double pab = r2[0] * r2[1], pcd = r2[2] * r2[3], pp = pab * pcd;
                const double inv = recip64(pp);
                const double iab = inv * pcd, icd = inv * pab;
                K[0] = r2[1] * iab;
                K[1] = r2[0] * iab;
                K[2] = r2[3] * icd;
                K[3] = r2[2] * icd;
                const bool shared = fmin(pab, pcd) > 1e-300 && pp > 1e-300 && pp < 1e300;
                if (__builtin_amdgcn_ballot_w64(!shared) != 0) {  // wave-uniform, rare
                    asm volatile("" ::: "memory");  // a branch, not computed ahead
#pragma unroll
                    for (int q = 0; q < 4; ++q) K[q] = shared ? K[q] : recip64(r2[q]);
                }
                if (nearm != 0) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double thr = srec[k + q].w;
                        const bool near = take[q] && r2[q] < thr;
                        const double f = 1.0 - exp(-r2[q] * (10.0 / thr));
                        K[q] = near ? K[q] * f : K[q];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double kc = take[q] ? K[q] * srec[k + q].z : 0.0;
                    ui = __builtin_fma(-kc, dz[q], ui);
                    wi = __builtin_fma(kc, dx[q], wi);
                }
            }
        }
        if (have) {
            out[pj] = ui;
            out[mp + pj] = wi;
        }
    }
    double cs = 0.0;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) cs += b.c[base + j];
    red[threadIdx.x] = cs;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[2 * mp] = red[0];
}

// Per slot: b = (U + u).n, b_Nl = -sum circ (the chunk partials added in
// chunk order), into rhs (mp + 1 per slot).
__global__ __launch_bounds__(64) void bt_rhs_kernel(int mp, const SlotParam* __restrict__ prm,
                                                     const SlotState* __restrict__ st,
                                                     const double* __restrict__ enx,
                                                     const double* __restrict__ enz,
                                                     const double* __restrict__ col,
                                                     double* __restrict__ rhs)
{
    const int s = blockIdx.y;
    if (!st[s].active) return;
    const int nl = prm[s].nl;
    const bool have = st[s].n > 0;
    const double* cs = col + (int64_t)s * kColChunks * (2 * mp + 1);
    const int nch = colloc_chunks(st[s].n);
    const int64_t cstride = 2 * mp + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nl) {
        double v;
        if (i == nl) {
            double tot = 0.0;
#pragma unroll 8
            for (int c = 0; c < nch; ++c) tot += cs[c * cstride + 2 * mp];
            v = have ? -tot : -0.0;
        } else {
            double u = 0.0, w = 0.0;
            if (have) {
#pragma unroll 8
                for (int c = 0; c < nch; ++c) {
                    u += cs[c * cstride + i];
                    w += cs[c * cstride + mp + i];
                }
            }
            const double ex = enx[(int64_t)s * mp + i], ez = enz[(int64_t)s * mp + i];
            v = have ? (prm[s].ux + u * kRpi2) * ex + (prm[s].uz + w * kRpi2) * ez
                     : prm[s].ux * ex + prm[s].uz * ez;
        }
        rhs[(int64_t)s * (mp + 1) + i] = v;
    }
}

// gamma = P b per slot, kGammaBlocks workgroups per slot, one wavefront per
// row: lane l sums columns l, l+64, ... in order, then the lanes' sums are
// added by shuffles — the row arithmetic of gamma_kernel, so a slot's gamma
// equals the single simulation's bit for bit.
constexpr int kGammaBlocks = 16;

__global__ __launch_bounds__(256) void bt_gamma_kernel(int mp, const SlotParam* __restrict__ prm,
                                                       const SlotState* __restrict__ st,
                                                       const double* __restrict__ P,
                                                       const double* __restrict__ rhs,
                                                       double* __restrict__ gamma)
{
    const int s = blockIdx.y;
    if (!st[s].active) return;
    extern __shared__ double sb[];
    const int nl = prm[s].nl;
    const int ncol = nl + 1;
    for (int i = threadIdx.x; i < ncol; i += 256) sb[i] = rhs[(int64_t)s * (mp + 1) + i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double* Ps = P + (int64_t)s * mp * (mp + 1);
    for (int i = blockIdx.x * 4 + wave; i < nl; i += kGammaBlocks * 4) {
        double acc = 0.0;
        for (int c = lane; c < ncol; c += 64) acc += Ps[(int64_t)i * ncol + c] * sb[c];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (lane == 0) gamma[(int64_t)s * mp + i] = acc;
    }
}

__global__ __launch_bounds__(256) void bt_sheet_kernel(int S, int64_t cap, int mp,
                                                       const WorkList* wl,
                                                       const int64_t* __restrict__ blob_pre,
                                                       const SlotParam* __restrict__ prm, BlobBuf b,
                                                       const PanelConst* __restrict__ panel,
                                                       const double* __restrict__ gamma)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int s, int64_t o) {
        const int nl = prm[s].nl;
        const PanelConst* pc = panel + (int64_t)s * mp;
        const double* g = gamma + (int64_t)s * mp;
        const double xi = b.x[o], zi = b.z[o];
        double su = 0.0, sw = 0.0;
        for (int j = 0; j < nl; ++j) sheet_pair(xi, zi, pc[j], g[j], g[j + 1 == nl ? 0 : j + 1], su, sw);
        b.uvs[o] = su * kRpi2;
        b.wvs[o] = sw * kRpi2;
    });
}

// FP32C form of the sheet's velocity by panel groups (sheet_groups): moments
// of the current gamma per (group, term), then per blob sheet_by_groups.
__global__ __launch_bounds__(64) void bt_group_moments_kernel(int mp, int mg,
                                                              const SlotParam* __restrict__ prm,
                                                              const SlotState* __restrict__ st,
                                                              const double* __restrict__ gamma,
                                                              const double* __restrict__ i0,
                                                              const double* __restrict__ i1,
                                                              double* __restrict__ ga)
{
    const int s = blockIdx.y;
    if (!st[s].active) return;
    const int64_t io = (int64_t)s * mp * 2 * kSheetTerms;
    group_moments(blockIdx.x, threadIdx.x, prm[s].nl, gamma + (int64_t)s * mp, i0 + io, i1 + io,
                  ga + (int64_t)s * mg * 2 * kSheetTerms);
}

// Each slot's blobs in its Morton order (the sorted positions p64 and the
// permutation perm of the pair sum's sort), a workgroup per 256-blob tile of
// one slot: a wavefront's blobs sit close together and mostly take the same
// branch per group, and the slot's panels, group centres and moments are
// workgroup-uniform (scalar loads).
__global__ __launch_bounds__(kBT) void bt_sheet_groups_kernel(
    int S, int64_t cap, int mp, int mg, const WorkList* wl, const int* __restrict__ tile_pre,
    const SlotState* __restrict__ st, const SlotParam* __restrict__ prm,
    const uint32_t* __restrict__ perm, const double2* __restrict__ p64,
    const PanelConst* __restrict__ panel, const double* __restrict__ gamma,
    const double4* __restrict__ gcen, const double* __restrict__ ga, BlobBuf b)
{
    const int total = wl->total_tiles;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int s = find_slot(tile_pre, S, g);
        const int64_t k = (int64_t)(g - tile_pre[s]) * kBT + threadIdx.x;  // sorted rank
        if (k >= st[s].n) continue;
        const int64_t o = (int64_t)s * cap + k;
        const double2 p = p64[o];
        double su = 0.0, sw = 0.0;
        sheet_by_groups(p.x, p.y, prm[s].nl, panel + (int64_t)s * mp, gamma + (int64_t)s * mp,
                        gcen + (int64_t)s * mg, ga + (int64_t)s * mg * 2 * kSheetTerms, su, sw);
        const int64_t dst = (int64_t)s * cap + perm[o];
        b.uvs[dst] = su * kRpi2;
        b.wvs[dst] = sw * kRpi2;
    }
}

// ----------------------------------------------------------------- blob updates

__global__ void bt_zero_q_kernel(int S, int64_t cap, const WorkList* wl,
                                 const int64_t* __restrict__ blob_pre, double* q1, double* q2)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int, int64_t o) { q1[o] = 0.0; q2[o] = 0.0; });
}

__global__ void bt_advect_kernel(int S, int64_t cap, const WorkList* wl,
                                 const int64_t* __restrict__ blob_pre,
                                 const SlotParam* __restrict__ prm, int rk, double a, double bb,
                                 BlobBuf b, double* q1, double* q2)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int s, int64_t o) {
        const double dt = prm[s].dt, ux = prm[s].ux, uz = prm[s].uz;
        if (!rk) {
            b.x[o] += (b.u[o] + b.uvs[o] + ux) * dt;
            b.z[o] += (b.w[o] + b.wvs[o] + uz) * dt;
        } else {
            const double n1 = a * q1[o] + (b.u[o] + b.uvs[o] + ux) * dt;
            const double n2 = a * q2[o] + (b.w[o] + b.wvs[o] + uz) * dt;
            q1[o] = n1;
            q2[o] = n2;
            b.x[o] += bb * n1;
            b.z[o] += bb * n2;
        }
    });
}

// Box-Muller random walk on Philox draws keyed by (seed, step, blob index):
// the same stream as the single-simulation device mode.
__global__ void bt_diffuse_kernel(int S, int64_t cap, const WorkList* wl,
                                  const int64_t* __restrict__ blob_pre,
                                  const SlotParam* __restrict__ prm,
                                  const SlotState* __restrict__ st, BlobBuf b)
{
    for_each_blob(wl, blob_pre, S, cap, [&](int s, int64_t o) {
        const int64_t i = o - (int64_t)s * cap;
        const uint32_t step = (uint32_t)(st[s].step + 1);
        const uint4 r = philox4x32(make_uint4((uint32_t)i, (uint32_t)(i >> 32), step, 0u),
                                   make_uint2(prm[s].seed, 0x5eed0001u));
        const double rrw = sqrt(prm[s].four_nu_dt * log(1.0 / u01(r.x, r.y)));
        const double th = 2.0 * 3.14159265358979323846 * u01(r.z, r.w);
        b.x[o] += rrw * cos(th);
        b.z[o] += rrw * sin(th);
    });
}

// One workgroup per slot: release counts, offsets, circulation per nascent
// blob (src/VortexSheet.cpp:282-303); stops the slot on NaN counts or overflow.
__global__ __launch_bounds__(1024) void bt_release_count_kernel(
    int64_t cap, int mp, const SlotParam* __restrict__ prm, SlotState* __restrict__ st,
    const double* __restrict__ gamma, const double* __restrict__ ds, int* __restrict__ cnt,
    int64_t* __restrict__ off, double* __restrict__ cnew)
{
    const int s = blockIdx.x;
    if (!st[s].active) return;
    __shared__ double red[1024];
    __shared__ int64_t pre[1024];
    __shared__ int bad_sh;
    const int t = threadIdx.x;
    const int nl = prm[s].nl;
    const double* g = gamma + (int64_t)s * mp;
    const double* d = ds + (int64_t)s * mp;
    double mx = -INFINITY;
    for (int i = t; i < nl; i += 1024) {
        const double v = g[i] * d[i];
        if (v > mx) mx = v;
    }
    red[t] = mx;
    if (t == 0) bad_sh = 0;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w && red[t + w] > red[t]) red[t] = red[t + w];
        __syncthreads();
    }
    const double m = red[0];
    const int per = (nl + 1023) / 1024;
    const int i0 = t * per;
    const int i1 = i0 + per < nl ? i0 + per : nl;
    int64_t local = 0;
    int bad = 0;
    for (int i = i0; i < i1; ++i) {
        const double pcirc = g[i] * d[i];
        const double c = floor(fabs(pcirc / m * (double)prm[s].max_npv)) + 1.0;
        int ci = 0;
        if (!(c > 1.0e-15)) bad |= 1;
        else if (!(c < 1.0e9)) bad |= 2;
        else ci = (int)c;
        cnt[(int64_t)s * mp + i] = ci;
        cnew[(int64_t)s * mp + i] = pcirc / c;
        local += ci;
    }
    if (bad) atomicOr(&bad_sh, bad);
    pre[t] = local;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {
        const int64_t v = t >= w ? pre[t - w] : 0;
        __syncthreads();
        pre[t] += v;
        __syncthreads();
    }
    int64_t run = pre[t] - local;
    for (int i = i0; i < i1; ++i) {
        off[(int64_t)s * (mp + 1) + i] = run;
        run += cnt[(int64_t)s * mp + i];
    }
    if (t == 1023) {
        const int64_t nrv = pre[1023];
        off[(int64_t)s * (mp + 1) + nl] = nrv;
        int status = 0;
        if (bad_sh & 1) status = DVM_E_NAN;
        else if (bad_sh & 2 || st[s].n + nrv > cap) status = DVM_E_SIZE;
        st[s].nrv = status ? 0 : (int)nrv;
        if (status) {
            st[s].status = status;
            st[s].active = 0;
        }
    }
}

// Nascent blobs of every slot: grid (chunks, S); blob k of slot s goes to
// index n_s + k.
__global__ void bt_release_emit_kernel(int64_t cap, int mp, const SlotParam* __restrict__ prm,
                                       const SlotState* __restrict__ st,
                                       const double* __restrict__ bx, const double* __restrict__ bz,
                                       const double* __restrict__ ds, const double* __restrict__ etx,
                                       const double* __restrict__ etz, const double* __restrict__ enx,
                                       const double* __restrict__ enz, const int* __restrict__ cnt,
                                       const int64_t* __restrict__ off,
                                       const double* __restrict__ cnew,
                                       const PanelConst* __restrict__ panel, BlobBuf b)
{
    const int s = blockIdx.y;
    if (!st[s].active) return;
    const int nrv = st[s].nrv;
    const int nl = prm[s].nl;
    const int64_t pb = (int64_t)s * mp, ob = (int64_t)s * (mp + 1), xb = (int64_t)s * (mp + 1);
    const uint32_t step = (uint32_t)(st[s].step + 1);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nrv; k += gridDim.x * blockDim.x) {
        int lo = 0, hi = nl;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[ob + mid] <= k) lo = mid; else hi = mid;
        }
        const int i = lo;
        const uint4 r = philox4x32(make_uint4((uint32_t)k, 0u, step, 1u),
                                   make_uint2(prm[s].seed, 0x5eed0002u));
        const double rrw = fabs(sqrt(prm[s].four_nu_dt * log(1.0 / u01(r.x, r.y))));
        const double np1 = (double)cnt[pb + i] + 1;
        const double xm = bx[xb + i] + ds[pb + i] * etx[pb + i] / np1;
        const double zm = bz[xb + i] + ds[pb + i] * etz[pb + i] / np1;
        const int64_t o = (int64_t)s * cap + st[s].n + k;
        b.x[o] = xm + rrw * enx[pb + i];
        b.z[o] = zm + rrw * enz[pb + i];
        b.c[o] = cnew[pb + i];
        b.s[o] = panel[pb + i].sig;
        b.u[o] = 0.0; b.w[o] = 0.0; b.uvs[o] = 0.0; b.wvs[o] = 0.0;
        b.id[o] = (uint32_t)st[s].n;
    }
}

__global__ void bt_commit_release_kernel(int S, SlotState* __restrict__ st)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !st[s].active) return;
    st[s].n += st[s].nrv;
    st[s].nrv = 0;
}

// Crossing flags (inside_body_fast with the slot's box and slab lists) and,
// for REFLECT, src/VortexSheet.cpp:345-405 for the inside blobs: each wave
// takes its inside blobs one at a time, all lanes working on one
// (reflect_wave).  A workgroup per 256-blob tile of one slot, so the slot's
// body, box and slab lists are workgroup-uniform (scalar loads) and the loop
// bounds wave-uniform (every lane takes part in the shuffles).
__global__ __launch_bounds__(kBT) void bt_inside_kernel(
    int S, int64_t cap, int mp, const WorkList* wl, const int* __restrict__ tile_pre,
    const SlotState* __restrict__ st, const SlotParam* __restrict__ prm,
    const double* __restrict__ bx, const double* __restrict__ bz, const double* __restrict__ enx,
    const double* __restrict__ enz, const int* __restrict__ slab_off,
    const int* __restrict__ slab_edge, int reflect, BlobBuf b, int* __restrict__ flags)
{
    const int total = wl->total_tiles;
    const int lane = threadIdx.x & 63;
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int s = find_slot(tile_pre, S, g);
        const int64_t k = (int64_t)(g - tile_pre[s]) * kBT + threadIdx.x;
        const int64_t o = (int64_t)s * cap + k;
        const int nl = prm[s].nl;
        const double* sbx = bx + (int64_t)s * (mp + 1);
        const double* sbz = bz + (int64_t)s * (mp + 1);
        bool in = false;
        double x0 = 0.0, z0 = 0.0;
        if (k < st[s].n) {
            x0 = b.x[o];
            z0 = b.z[o];
            EdgeSlabs es;
            es.off = slab_off + (int64_t)s * (mp + 1);
            es.edge = slab_edge + (int64_t)s * kSlabCap * mp;
            es.n = prm[s].nslab;
            es.z0 = prm[s].slab_z0;
            es.inv_h = prm[s].slab_inv_h;
            in = inside_body_fast(nl, sbx, sbz, prm[s].box, es, x0, z0);
            flags[o] = in ? 1 : 0;
        }
        if (!reflect) continue;
        uint64_t mask = __ballot(in);
        while (mask) {
            const int l = __ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1;
            const double px = __shfl(x0, l), pz = __shfl(z0, l);
            const double2 p = reflect_wave(nl, sbx, sbz, enx + (int64_t)s * mp, enz + (int64_t)s * mp,
                                           px, pz);
            if (lane == l) {
                b.x[o] = p.x;
                b.z[o] = p.y;
            }
        }
    }
}

// One workgroup per slot: stable compaction of the blobs outside the body
// (src/VortexSheet.cpp:407-445) through the alternate buffers, IDs renumbered.
__global__ __launch_bounds__(1024) void bt_delete_kernel(int64_t cap, SlotState* __restrict__ st,
                                                         const int* __restrict__ flags, BlobBuf b,
                                                         BlobBuf alt)
{
    const int s = blockIdx.x;
    if (!st[s].active) return;
    __shared__ int scan[1024];
    __shared__ int64_t base_out;
    const int t = threadIdx.x;
    const int64_t n = st[s].n;
    const int64_t sb = (int64_t)s * cap;
    if (t == 0) base_out = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < n; c0 += 1024) {
        const int64_t i = c0 + t;
        const int keep = (i < n && !flags[sb + i]) ? 1 : 0;
        scan[t] = keep;
        __syncthreads();
        for (int w = 1; w < 1024; w <<= 1) {
            const int v = t >= w ? scan[t - w] : 0;
            __syncthreads();
            scan[t] += v;
            __syncthreads();
        }
        if (keep) {
            const int64_t o = sb + base_out + scan[t] - 1;
            const int64_t q = sb + i;
            alt.x[o] = b.x[q]; alt.z[o] = b.z[q]; alt.c[o] = b.c[q]; alt.s[o] = b.s[q];
            alt.u[o] = b.u[q]; alt.w[o] = b.w[q]; alt.uvs[o] = b.uvs[q]; alt.wvs[o] = b.wvs[q];
            alt.id[o] = (uint32_t)(base_out + scan[t] - 1);
        }
        __syncthreads();
        if (t == 1023) base_out += scan[1023];
        __syncthreads();
    }
    const int64_t m = base_out;
    if (m == n) return;
    for (int64_t i = t; i < m; i += 1024) {
        const int64_t q = sb + i;
        b.x[q] = alt.x[q]; b.z[q] = alt.z[q]; b.c[q] = alt.c[q]; b.s[q] = alt.s[q];
        b.u[q] = alt.u[q]; b.w[q] = alt.w[q]; b.uvs[q] = alt.uvs[q]; b.wvs[q] = alt.wvs[q];
        b.id[q] = alt.id[q];
    }
    __syncthreads();
    if (t == 0) st[s].n = m;
}

// One workgroup per slot: the loads of src/VortexSheet.cpp:487-514
// (loads_block; work holds 3 mp doubles per slot), then the step counter; a
// slot that reaches its step count stops.  (Every lane reads `active` before
// the first barrier of loads_block, so lane 0's update below is seen by none.)
__global__ __launch_bounds__(kLoadsThreads) void bt_loads_finish_kernel(
    int S, int mp, const SlotParam* __restrict__ prm, SlotState* __restrict__ st,
    const double* __restrict__ gamma, const double* __restrict__ ds,
    const double* __restrict__ enx, const double* __restrict__ enz, double* __restrict__ work)
{
    const int s = blockIdx.x;
    if (s >= S || !st[s].active) return;
    const int64_t o = (int64_t)s * mp;
    loads_block(prm[s].nl, prm[s].rho, prm[s].dt, prm[s].ur, prm[s].aref, gamma + o, ds + o,
                enx + o, enz + o, work + (int64_t)s * 3 * mp, st[s].loads);
    if (threadIdx.x == 0) {
        st[s].step += 1;
        if (st[s].step >= prm[s].steps_target) st[s].active = 0;
    }
}

}  // namespace dvmi

// ======================================================================= host

using namespace dvmi;

struct dvm_batch_s {
    dvm_ctx_t ctx = nullptr;
    dvm_batch_config cfg{};
    BatchDev d;
    char* mem = nullptr;
    std::vector<int> busy;            // admitted and not yet reported finished
    std::vector<int> host_steps;      // steps completed while busy (mirror for remesh timing)
    std::vector<SlotState> hst;       // last polled states
    hipGraph_t graph_a = nullptr, graph_b = nullptr;
    hipGraphExec_t exec_a = nullptr, exec_b = nullptr;
    std::vector<hipStream_t> admit_streams;  // set-up of several bodies at once
    // dvm_batch_set_fmm: slots of at least fmm_min blobs take the fast
    // multipole sum (fmm_order, fmm_theta) before the convection graph
    int64_t fmm_min = 0;
    int fmm_order = 12;
    double fmm_theta = 0.5;
    SlotState* pst = nullptr;  // pinned staging of the slot states (S)
    // contexts (stream, arenas) of the fast multipole sums: the slots' sums
    // run on them in turn, concurrently; the main stream waits on fmm_ev
    std::vector<dvm_ctx_t> fmm_ctx;
    std::vector<hipEvent_t> fmm_ev;
};

namespace {

void carve_blob_set(Carver& cv, BlobBuf& b, size_t n)
{
    b.x = cv.take<double>(n); b.z = cv.take<double>(n); b.c = cv.take<double>(n);
    b.s = cv.take<double>(n); b.u = cv.take<double>(n); b.w = cv.take<double>(n);
    b.uvs = cv.take<double>(n); b.wvs = cv.take<double>(n); b.id = cv.take<uint32_t>(n);
}

size_t batch_bytes(const dvm_batch_config& c, BatchDev* d, char* mem)
{
    const size_t S = c.slots, cap = c.slot_capacity, mp = c.max_panels;
    const size_t nb = S * cap, np = S * mp, np1 = S * (mp + 1);
    const size_t maxt = (cap + kBT - 1) / kBT;
    Carver cv(mem);
    BatchDev t;
    BatchDev& o = d ? *d : t;
    o.prm = cv.take<SlotParam>(S);
    o.st = cv.take<SlotState>(S);
    o.wl = cv.take<WorkList>(1);
    o.tile_pre = cv.take<int>(S + 1);
    o.order = cv.take<int>(S);
    o.otile_pre = cv.take<int>(S + 1);
    o.blob_pre = cv.take<int64_t>(S + 1);
    carve_blob_set(cv, o.b, nb);
    carve_blob_set(cv, o.alt, nb);
    o.q1 = cv.take<double>(nb);
    o.q2 = cv.take<double>(nb);
    const bool f32c = c.prec == DVM_F32C;
    const size_t nf = f32c ? nb : 0;  // FP32C-only arrays
    const size_t maxtb = (cap + kTB - 1) / kTB;
    o.s32 = cv.take<float4>(nf);
    o.near32 = cv.take<float4>(nf);
    o.p64 = cv.take<double2>(nf);
    o.tiles = cv.take<TileInfo>(S * maxt);
    o.item_pre = cv.take<int>(S + 1);
    o.sbox = cv.take<double4>(S);
    o.keys = cv.take<uint32_t>(nf);
    o.keys2 = cv.take<uint32_t>(nf);
    o.vals = cv.take<uint32_t>(nf);
    o.perm = cv.take<uint32_t>(nf);
    o.box_part = cv.take<double4>(f32c ? S * kBoxParts : 0);
    const size_t maxst = (cap + kSortTile - 1) / kSortTile;
    o.max_stiles = (int)maxst;
    o.stile_pre = cv.take<int>(S + 1);
    o.shist = cv.take<int>(f32c ? S * maxst * 256 : 0);
    o.tboxes = cv.take<TileInfo>(f32c ? S * maxtb : 0);
    o.max_tboxes = (int)maxtb;
    o.part_slot = f32c ? part_slot_size((int64_t)cap) : 0;
    o.part = cv.take<double>(S * (size_t)o.part_slot);
    o.flags = cv.take<int>(nb);
    o.bx = cv.take<double>(np1);
    o.bz = cv.take<double>(np1);
    o.enx = cv.take<double>(np); o.enz = cv.take<double>(np);
    o.etx = cv.take<double>(np); o.etz = cv.take<double>(np);
    o.ds = cv.take<double>(np);
    o.panel = cv.take<PanelConst>(np);
    o.P = cv.take<double>(S * mp * (mp + 1));
    o.gamma = cv.take<double>(np);
    o.col = cv.take<double>(S * kColChunks * (2 * mp + 1));
    o.cnt = cv.take<int>(np);
    o.off = cv.take<int64_t>(np1);
    o.cnew = cv.take<double>(np);
    o.lwork = cv.take<double>(S * 3 * mp);
    o.rhs = cv.take<double>(S * (mp + 1));
    o.slab_off = cv.take<int>(S * (mp + 1));
    o.slab_edge = cv.take<int>(S * kSlabCap * mp);
    const size_t mg = (mp + kSheetGroup - 1) / kSheetGroup;
    o.mg = (int)mg;
    o.gcen = cv.take<double4>(f32c ? S * mg : 0);
    o.gI0 = cv.take<double>(f32c ? np * 2 * kSheetTerms : 0);
    o.gI1 = cv.take<double>(f32c ? np * 2 * kSheetTerms : 0);
    o.gA = cv.take<double>(f32c ? S * mg * 2 * kSheetTerms : 0);
    return cv.off + 256;
}

// The slots' Morton keys (keys, vals) sorted into (keys2, perm): three
// 8-bit passes, (keys, vals) -> (keys2, perm) -> (keys, vals) -> (keys2, perm).
static void record_sort(const BatchDev& d, hipStream_t st)
{
    static_assert(kSortBits == 24, "three 8-bit passes");
    const unsigned grid = (unsigned)std::min<int64_t>((int64_t)d.S * d.max_stiles, 4096);
    uint32_t* kin[3] = {d.keys, d.keys2, d.keys};
    uint32_t* vin[3] = {d.vals, d.perm, d.vals};
    uint32_t* kout[3] = {d.keys2, d.keys, d.keys2};
    uint32_t* vout[3] = {d.perm, d.vals, d.perm};
    for (int p = 0; p < 3; ++p) {
        bt_sort_hist_kernel<<<grid, 256, 0, st>>>(d.S, d.cap, d.wl, d.stile_pre, d.st, kin[p],
                                                  8 * p, d.shist);
        bt_sort_scan_kernel<<<d.S, 256, 0, st>>>(d.stile_pre, d.shist);
        bt_sort_scatter_kernel<<<grid, 256, 0, st>>>(d.S, d.cap, d.wl, d.stile_pre, d.st, kin[p],
                                                     vin[p], 8 * p, d.shist, kout[p], vout[p]);
    }
}

void record_convect(dvm_batch_t B)
{
    BatchDev& d = B->d;
    hipStream_t st = B->ctx->stream;
    const int S = d.S;
    const int64_t cap = d.cap;
    const int mp = d.mp;
    const unsigned gs = 4096;  // grid-stride launches
    const unsigned persistent = (unsigned)B->ctx->num_cus * 6;
    const bool f32c = B->cfg.prec == DVM_F32C;
    const bool rk = B->cfg.scheme == DVM_RK3;
    const double a[3] = {0, -17. / 32, -32. / 27};
    const double bb[3] = {1. / 4, 8. / 9, 3. / 4};
    const int stages = rk ? 3 : 1;
    for (int k = 0; k < stages; ++k) {
        bt_prefix_kernel<<<1, prefix_lanes(S), 0, st>>>(S, cap, d.st, d.wl, d.tile_pre, d.order, d.otile_pre,
                                             d.blob_pre, d.item_pre, d.stile_pre);
        if (rk && k == 0) bt_zero_q_kernel<<<gs, 256, 0, st>>>(S, cap, d.wl, d.blob_pre, d.q1, d.q2);
        if (f32c) {
            bt_bbox_kernel<<<dim3(kBoxParts, S), 256, 0, st>>>(cap, d.st, d.b, d.box_part);
            bt_bbox_final_kernel<<<(S + 63) / 64, 64, 0, st>>>(S, d.box_part, d.sbox);
            bt_morton_kernel<<<gs, 256, 0, st>>>(S, cap, d.wl, d.blob_pre, d.sbox, d.b, d.keys, d.vals);
            record_sort(d, st);
            bt_tiles_kernel<<<gs, kBT, 0, st>>>(S, cap, d.max_tboxes, d.wl, d.tile_pre, d.st, d.perm,
                                                d.b, d.s32, d.near32, d.p64, d.tiles, d.tboxes);
            bt_bs32_kernel<<<persistent, 256, 0, st>>>(S, cap, d.max_tboxes, d.wl, d.tile_pre, d.order,
                                                       d.item_pre, d.st, d.s32, d.near32, d.p64,
                                                       d.tiles, d.tboxes, d.part_slot, d.part);
            bt_reduce32_kernel<<<gs, 256, 0, st>>>(S, cap, d.wl, d.blob_pre, d.st, d.perm,
                                                   d.part_slot, d.part, d.b);
        } else {
            bt_bs64_kernel<<<persistent, kBT, 0, st>>>(S, cap, d.wl, d.order, d.otile_pre, d.st, d.b);
        }
        bt_colloc_kernel<<<dim3(kColChunks, S), 256, 0, st>>>(cap, mp, d.prm, d.st, d.b, d.panel, d.col);
        bt_rhs_kernel<<<dim3((mp + 64) / 64, S), 64, 0, st>>>(mp, d.prm, d.st, d.enx, d.enz, d.col,
                                                              d.rhs);
        bt_gamma_kernel<<<dim3(kGammaBlocks, S), 256, sizeof(double) * (mp + 1), st>>>(
            mp, d.prm, d.st, d.P, d.rhs, d.gamma);
        if (f32c) {
            bt_group_moments_kernel<<<dim3(d.mg, S), 64, 0, st>>>(mp, d.mg, d.prm, d.st, d.gamma,
                                                                  d.gI0, d.gI1, d.gA);
            bt_sheet_groups_kernel<<<gs, kBT, 0, st>>>(S, cap, mp, d.mg, d.wl, d.tile_pre, d.st,
                                                       d.prm, d.perm, d.p64, d.panel, d.gamma,
                                                       d.gcen, d.gA, d.b);
        } else {
            bt_sheet_kernel<<<gs, 256, 0, st>>>(S, cap, mp, d.wl, d.blob_pre, d.prm, d.b, d.panel,
                                                d.gamma);
        }
        bt_advect_kernel<<<gs, 256, 0, st>>>(S, cap, d.wl, d.blob_pre, d.prm, rk ? 1 : 0, a[k], bb[k],
                                             d.b, d.q1, d.q2);
    }
    DVM_LAUNCH_CHECK();
}

void record_rest(dvm_batch_t B)
{
    BatchDev& d = B->d;
    hipStream_t st = B->ctx->stream;
    const int S = d.S;
    const int64_t cap = d.cap;
    const int mp = d.mp;
    const unsigned gs = 4096;
    bt_prefix_kernel<<<1, prefix_lanes(S), 0, st>>>(S, cap, d.st, d.wl, d.tile_pre, d.order, d.otile_pre,
                                         d.blob_pre, d.item_pre, d.stile_pre);
    bt_diffuse_kernel<<<gs, 256, 0, st>>>(S, cap, d.wl, d.blob_pre, d.prm, d.st, d.b);
    bt_release_count_kernel<<<S, 1024, 0, st>>>(cap, mp, d.prm, d.st, d.gamma, d.ds, d.cnt, d.off, d.cnew);
    bt_release_emit_kernel<<<dim3(16, S), 256, 0, st>>>(cap, mp, d.prm, d.st, d.bx, d.bz, d.ds, d.etx,
                                                         d.etz, d.enx, d.enz, d.cnt, d.off, d.cnew,
                                                         d.panel, d.b);
    bt_commit_release_kernel<<<(S + 255) / 256, 256, 0, st>>>(S, d.st);
    bt_prefix_kernel<<<1, prefix_lanes(S), 0, st>>>(S, cap, d.st, d.wl, d.tile_pre, d.order, d.otile_pre,
                                         d.blob_pre, d.item_pre, d.stile_pre);
    const int reflect = B->cfg.crossing != DVM_DELETE ? 1 : 0;
    bt_inside_kernel<<<gs, kBT, 0, st>>>(S, cap, mp, d.wl, d.tile_pre, d.st, d.prm, d.bx, d.bz,
                                         d.enx, d.enz, d.slab_off, d.slab_edge, reflect, d.b,
                                         d.flags);
    if (!reflect) bt_delete_kernel<<<S, 1024, 0, st>>>(cap, d.st, d.flags, d.b, d.alt);
    bt_loads_finish_kernel<<<S, kLoadsThreads, 0, st>>>(S, mp, d.prm, d.st, d.gamma, d.ds, d.enx,
                                                         d.enz, d.lwork);
    DVM_LAUNCH_CHECK();
}

void build_graphs(dvm_batch_t B)
{
    hipStream_t st = B->ctx->stream;
    DVM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    record_convect(B);
    DVM_HIP_CHECK(hipStreamEndCapture(st, &B->graph_a));
    DVM_HIP_CHECK(hipGraphInstantiate(&B->exec_a, B->graph_a, nullptr, nullptr, 0));
    DVM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    record_rest(B);
    DVM_HIP_CHECK(hipStreamEndCapture(st, &B->graph_b));
    DVM_HIP_CHECK(hipGraphInstantiate(&B->exec_b, B->graph_b, nullptr, nullptr, 0));
}

void poll_states(dvm_batch_t B)
{
    DVM_HIP_CHECK(hipMemcpyAsync(B->hst.data(), B->d.st, sizeof(SlotState) * B->d.S,
                                 hipMemcpyDeviceToHost, B->ctx->stream));
    DVM_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
}

// The remeshed blobs of a slot (written to the alternate buffers) back into
// its storage, and its new count.
__global__ void bt_remesh_commit_kernel(int64_t m, BlobBuf from, BlobBuf to, SlotState* st)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        to.x[j] = from.x[j]; to.z[j] = from.z[j]; to.c[j] = from.c[j]; to.s[j] = from.s[j];
        to.u[j] = from.u[j]; to.w[j] = from.w[j]; to.uvs[j] = from.uvs[j]; to.wvs[j] = from.wvs[j];
        to.id[j] = from.id[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->n = m;
}

// Remesh of the slots whose next step is a remesh step (BackgroundMesh
// specification of remesh.hip), one slot at a time between the two graphs.
void remesh_due(dvm_batch_t B)
{
    const dvm_remesh_params& rp = B->cfg.remesh;
    if (rp.frequency <= 0) return;
    bool any = false;
    for (int s = 0; s < B->d.S; ++s)
        if (B->busy[s] && (B->host_steps[s] + 1) % rp.frequency == 0) any = true;
    if (!any) return;
    poll_states(B);
    hipStream_t st = B->ctx->stream;
    for (int s = 0; s < B->d.S; ++s) {
        if (!B->busy[s] || !B->hst[s].active || (B->host_steps[s] + 1) % rp.frequency != 0) continue;
        const int64_t off = (int64_t)s * B->d.cap;
        auto view = [off](const BlobBuf& b) {
            BlobBuf v;
            v.x = b.x + off; v.z = b.z + off; v.c = b.c + off; v.s = b.s + off;
            v.u = b.u + off; v.w = b.w + off; v.uvs = b.uvs + off; v.wvs = b.wvs + off;
            v.id = b.id + off;
            return v;
        };
        BlobBuf in = view(B->d.b), out = view(B->d.alt);
        RemeshPlan plan;
        const int64_t m = remesh_count(B->ctx, rp, B->hst[s].n, in, plan);
        if (m > B->d.cap) {
            SlotState bad = B->hst[s];
            bad.active = 0;
            bad.status = DVM_E_SIZE;
            DVM_HIP_CHECK(hipMemcpyAsync(B->d.st + s, &bad, sizeof(SlotState), hipMemcpyHostToDevice, st));
            DVM_HIP_CHECK(hipStreamSynchronize(st));
            continue;
        }
        remesh_finish(B->ctx, rp, in, plan, out);
        bt_remesh_commit_kernel<<<m > 0 ? (unsigned)std::min<int64_t>(cdiv(m, 256), 1024) : 1, 256, 0,
                                  st>>>(m, out, in, B->d.st + s);
        DVM_LAUNCH_CHECK();
    }
}

// The blob velocities of the slots with at least fmm_min blobs by the fast
// multipole sum (fmm.hip, on the slot's own arrays), enqueued before the
// convection graph on the batch's FMM contexts in turn (their streams run
// concurrently; the main stream waits for them); the ext flags make the
// graph's pair sum and reduce skip those slots.  The choice depends on the
// slot's own count, so a slot's result is the same alone or in any mix.
// Euler batches only (checked when set).
void fmm_due(dvm_batch_t B)
{
    if (B->fmm_min <= 0) return;
    poll_states(B);  // the main stream is idle: positions and counts are final
    const int S = B->d.S;
    const int K = (int)B->fmm_ctx.size();
    std::vector<char> used((size_t)K, 0);
    auto join = [&] {  // the main stream waits for the contexts that ran sums
        for (int k = 0; k < K; ++k) {
            if (!used[(size_t)k]) continue;
            DVM_HIP_CHECK(hipEventRecord(B->fmm_ev[(size_t)k], B->fmm_ctx[(size_t)k]->stream));
            DVM_HIP_CHECK(hipStreamWaitEvent(B->ctx->stream, B->fmm_ev[(size_t)k], 0));
        }
    };
    int j = 0;
    try {
        for (int s = 0; s < S; ++s) {
            SlotState ss = B->hst[s];
            ss.ext = 0;
            if (B->busy[s] && ss.active && ss.n >= B->fmm_min) {
                const int k = j++ % K;
                const int64_t off = (int64_t)s * B->d.cap;
                const BlobBuf& b = B->d.b;
                used[(size_t)k] = 1;
                ss.ext = fmm_self(B->fmm_ctx[(size_t)k], ss.n, b.x + off, b.z + off, b.c + off,
                                  b.s + off, b.u + off, b.w + off) ? 1 : 0;
            }
            B->pst[s] = ss;
        }
    } catch (...) {
        // no slot keeps a flag without this step's sum behind it
        for (int s = 0; s < S; ++s) {
            B->pst[s] = B->hst[s];
            B->pst[s].ext = 0;
        }
        join();
        DVM_HIP_CHECK(hipMemcpyAsync(B->d.st, B->pst, sizeof(SlotState) * S,
                                     hipMemcpyHostToDevice, B->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
        throw;
    }
    join();
    // the staging is rewritten only after the next poll has synchronised
    DVM_HIP_CHECK(hipMemcpyAsync(B->d.st, B->pst, sizeof(SlotState) * S, hipMemcpyHostToDevice,
                                 B->ctx->stream));
}

}  // namespace

// Enqueues on `st` the copies that place a set-up body `sh` in slot s and
// start it (sps, sss: host staging, alive until `st` is synchronised).
static void admit_copy(dvm_batch_t B, int s, const SheetDev& sh, const dvm_sheet_params& pp,
                       int seed, int steps, hipStream_t st, SlotParam& sp, SlotState& ss)
{
    const dvm_sheet_params* p = &pp;
    BatchDev& d = B->d;
    const int nl = sh.nl, mp = d.mp;
    const int npts = nl + 1;
    const size_t db = sizeof(double);
    auto d2d = [&](void* dst, const void* src, size_t bytes) {
        DVM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    };
    d2d(d.bx + (size_t)s * (mp + 1), sh.x, db * npts);
    d2d(d.bz + (size_t)s * (mp + 1), sh.z, db * npts);
    d2d(d.enx + (size_t)s * mp, sh.enx, db * nl);
    d2d(d.enz + (size_t)s * mp, sh.enz, db * nl);
    d2d(d.etx + (size_t)s * mp, sh.etx, db * nl);
    d2d(d.etz + (size_t)s * mp, sh.etz, db * nl);
    d2d(d.ds + (size_t)s * mp, sh.ds, db * nl);
    d2d(d.panel + (size_t)s * mp, sh.panel, sizeof(PanelConst) * nl);
    d2d(d.P + (size_t)s * mp * (mp + 1), sh.P, db * nl * (nl + 1));
    DVM_HIP_CHECK(hipMemsetAsync(d.gamma + (size_t)s * mp, 0, db * nl, st));
    if (B->cfg.prec == DVM_F32C) {
        const size_t mo = (size_t)s * mp * 2 * kSheetTerms;
        d2d(d.gcen + (size_t)s * d.mg, sh.gcen, sizeof(double4) * sh.ng);
        d2d(d.gI0 + mo, sh.gI0, db * 2 * kSheetTerms * nl);
        d2d(d.gI1 + mo, sh.gI1, db * 2 * kSheetTerms * nl);
    }
    sp = SlotParam{};
    sp.rho = p->rho; sp.nu = p->nu; sp.dt = p->dt; sp.ux = p->ux; sp.uz = p->uz;
    sp.ur = std::sqrt(p->ux * p->ux + p->uz * p->uz);
    sp.aref = sh.aref;
    sp.four_nu_dt = 4.0 * p->nu * p->dt;
    sp.nl = nl; sp.max_npv = p->max_npv; sp.steps_target = steps;
    sp.seed = (uint32_t)seed;
    sp.box = sh.box;
    sp.nslab = 0;
    if (sh.slabs.n > 0 && sh.hslab_edge.size() <= (size_t)kSlabCap * mp) {
        DVM_HIP_CHECK(hipMemcpyAsync(d.slab_off + (size_t)s * (mp + 1), sh.hslab_off.data(),
                                     sizeof(int) * sh.hslab_off.size(), hipMemcpyHostToDevice,
                                     st));
        DVM_HIP_CHECK(hipMemcpyAsync(d.slab_edge + (size_t)s * kSlabCap * mp,
                                     sh.hslab_edge.data(), sizeof(int) * sh.hslab_edge.size(),
                                     hipMemcpyHostToDevice, st));
        sp.nslab = sh.slabs.n;
        sp.slab_z0 = sh.slabs.z0;
        sp.slab_inv_h = sh.slabs.inv_h;
    }
    ss = SlotState{};
    ss.active = 1;
    DVM_HIP_CHECK(hipMemcpyAsync(d.prm + s, &sp, sizeof(sp), hipMemcpyHostToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(d.st + s, &ss, sizeof(ss), hipMemcpyHostToDevice, st));
}

extern "C" {

dvm_status dvm_batch_footprint(const dvm_batch_config* cfg, size_t* bytes)
{
    if (!cfg || !bytes) return DVM_E_INVALID;
    *bytes = batch_bytes(*cfg, nullptr, nullptr);
    return DVM_OK;
}

dvm_status dvm_batch_create(dvm_ctx_t ctx, const dvm_batch_config* cfg, dvm_batch_t* out)
{
    if (!ctx || !cfg || !out) return DVM_E_INVALID;
    *out = nullptr;
    dvm_batch_t B = new (std::nothrow) dvm_batch_s();
    if (!B) return DVM_E_OOM;
    B->ctx = ctx;
    dvm_status st = guard(ctx, [&] {
        const dvm_batch_config& c = *cfg;
        require(c.slots >= 1 && c.slots <= 1024, DVM_E_INVALID, "slots must be in [1, 1024]");
        require(c.slot_capacity >= kBT && c.slot_capacity < (int64_t(1) << 31), DVM_E_INVALID,
                "slot capacity out of range");
        require(c.max_panels >= 2 && c.max_panels <= 1024, DVM_E_INVALID, "max_panels must be in [2, 1024]");
        require(c.scheme == DVM_EULER || c.scheme == DVM_RK3, DVM_E_INVALID, "unknown scheme");
        require(c.crossing >= DVM_DELETE && c.crossing <= DVM_REFLECT, DVM_E_INVALID, "unknown crossing");
        require(c.prec == DVM_F64 || c.prec == DVM_F32C, DVM_E_INVALID, "unknown precision");
        B->cfg = c;
        B->d.S = c.slots;
        B->d.cap = c.slot_capacity;
        B->d.mp = c.max_panels;
        B->d.max_tiles = (int)((c.slot_capacity + kBT - 1) / kBT);
        if (c.prec == DVM_F32C)
            require((int64_t)c.slots * c.slot_capacity < (int64_t(1) << 31), DVM_E_INVALID,
                    "slots x slot capacity must stay below 2^31 blobs");
        const size_t bytes = batch_bytes(c, nullptr, nullptr);
        DVM_HIP_CHECK(hipMalloc(&B->mem, bytes));
        batch_bytes(c, &B->d, B->mem);
        DVM_HIP_CHECK(hipMemsetAsync(B->d.st, 0, sizeof(SlotState) * c.slots, ctx->stream));
        DVM_HIP_CHECK(hipMemsetAsync(B->d.prm, 0, sizeof(SlotParam) * c.slots, ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        B->busy.assign(c.slots, 0);
        B->host_steps.assign(c.slots, 0);
        B->hst.resize(c.slots);
        if (c.use_graph) build_graphs(B);
    });
    if (st != DVM_OK) {
        if (B->mem) (void)hipFree(B->mem);
        delete B;
        return st;
    }
    *out = B;
    return DVM_OK;
}

dvm_status dvm_batch_destroy(dvm_batch_t B)
{
    if (!B) return DVM_OK;
    dvm_status st = guard(B->ctx, [&] {
        DVM_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
        if (B->exec_a) DVM_HIP_CHECK(hipGraphExecDestroy(B->exec_a));
        if (B->exec_b) DVM_HIP_CHECK(hipGraphExecDestroy(B->exec_b));
        if (B->graph_a) DVM_HIP_CHECK(hipGraphDestroy(B->graph_a));
        if (B->graph_b) DVM_HIP_CHECK(hipGraphDestroy(B->graph_b));
        if (B->mem) DVM_HIP_CHECK(hipFree(B->mem));
        if (B->pst) DVM_HIP_CHECK(hipHostFree(B->pst));
        for (dvm_ctx_t c : B->fmm_ctx) (void)dvm_ctx_destroy(c);  // waits for its stream
        B->fmm_ctx.clear();
        for (hipEvent_t e : B->fmm_ev) DVM_HIP_CHECK(hipEventDestroy(e));
        B->fmm_ev.clear();
        for (hipStream_t a : B->admit_streams) DVM_HIP_CHECK(hipStreamDestroy(a));
        B->admit_streams.clear();
    });
    delete B;
    return st;
}

dvm_status dvm_batch_admit(dvm_batch_t B, int npts, const double* px, const double* pz,
                           const dvm_sheet_params* p, int seed, int steps, int* slot)
{
    if (!B || !slot) return DVM_E_INVALID;
    return dvm_batch_admit_many(B, 1, &npts, &px, &pz, p, &seed, &steps, slot);
}

dvm_status dvm_batch_admit_many(dvm_batch_t B, int count, const int* npts,
                                const double* const* px, const double* const* pz,
                                const dvm_sheet_params* p, const int* seed, const int* steps,
                                int* slots)
{
    if (!B || count < 0 || (count > 0 && (!npts || !px || !pz || !p || !seed || !steps || !slots)))
        return DVM_E_INVALID;
    dvm_ctx_t ctx = B->ctx;
    return guard(ctx, [&] {
        for (int i = 0; i < count; ++i) slots[i] = -1;
        std::vector<int> free_slots;
        for (int k = 0; k < B->d.S; ++k)
            if (!B->busy[k]) free_slots.push_back(k);
        require((int)free_slots.size() >= count, DVM_E_STATE, "no free slot");
        for (int i = 0; i < count; ++i) {
            require(steps[i] >= 1, DVM_E_INVALID, "steps must be >= 1");
            require(npts[i] - 1 <= B->d.mp, DVM_E_SIZE, "body has more panels than max_panels");
        }
        // Bodies set up on up to kAdmitStreams streams at once: each body's
        // least-squares passes run one workgroup at a time, so concurrent
        // bodies fill the device.  One host synchronisation at the end.
        const int ns = count < kAdmitStreams ? count : kAdmitStreams;
        while ((int)B->admit_streams.size() < ns) {
            hipStream_t st = nullptr;
            DVM_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            B->admit_streams.push_back(st);
        }
        std::vector<SheetDev> sh((size_t)count);
        std::vector<SlotParam> sps((size_t)count);
        std::vector<SlotState> sss((size_t)count);
        // the set-up reads nothing the context's stream still writes
        if (count > 1) DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        auto wait_all = [&]() {
            if (count == 1) {
                DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                return;
            }
            for (int k = 0; k < ns; ++k) DVM_HIP_CHECK(hipStreamSynchronize(B->admit_streams[k]));
        };
        try {
            for (int i = 0; i < count; ++i) {
                hipStream_t st = count == 1 ? ctx->stream : B->admit_streams[i % ns];
                sheet_init(ctx, sh[i], npts[i], px[i], pz[i], &p[i], st, false);
                admit_copy(B, free_slots[i], sh[i], p[i], seed[i], steps[i], st, sps[i], sss[i]);
            }
            wait_all();
        } catch (...) {
            for (int k = 0; k < ns && count > 1; ++k) (void)hipStreamSynchronize(B->admit_streams[k]);
            (void)hipStreamSynchronize(ctx->stream);
            for (auto& x : sh) sheet_free(x);
            throw;
        }
        for (int i = 0; i < count; ++i) {
            sheet_free(sh[i]);
            const int s = free_slots[i];
            B->busy[s] = 1;
            B->host_steps[s] = 0;
            B->hst[s] = sss[i];
            slots[i] = s;
        }
    });
}

dvm_status dvm_batch_step(dvm_batch_t B, int steps)
{
    if (!B) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        for (int k = 0; k < steps; ++k) {
            if (B->cfg.use_graph) {
                ProfScope ps(B->ctx, "batch_step");
                fmm_due(B);
                DVM_HIP_CHECK(hipGraphLaunch(B->exec_a, B->ctx->stream));
                remesh_due(B);
                DVM_HIP_CHECK(hipGraphLaunch(B->exec_b, B->ctx->stream));
            } else {
                ProfScope ps(B->ctx, "batch_step");
                fmm_due(B);
                record_convect(B);
                remesh_due(B);
                record_rest(B);
            }
            for (int s = 0; s < B->d.S; ++s)
                if (B->busy[s]) B->host_steps[s]++;
        }
    });
}

dvm_status dvm_batch_poll(dvm_batch_t B, int max, int* slots, int* status, int* count)
{
    if (!B || !count) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        poll_states(B);
        int c = 0;
        for (int s = 0; s < B->d.S && c < max; ++s) {
            if (B->busy[s] && !B->hst[s].active) {
                if (slots) slots[c] = s;
                if (status) status[c] = B->hst[s].status;
                B->busy[s] = 0;
                ++c;
            }
        }
        *count = c;
    });
}

dvm_status dvm_batch_slot_info(dvm_batch_t B, int slot, int64_t* n, int* step, int* active,
                               int* status, double* loads4)
{
    if (!B || slot < 0 || slot >= B->d.S) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        SlotState ss;
        DVM_HIP_CHECK(hipMemcpyAsync(&ss, B->d.st + slot, sizeof(ss), hipMemcpyDeviceToHost, B->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
        if (n) *n = ss.n;
        if (step) *step = ss.step;
        if (active) *active = ss.active;
        if (status) *status = ss.status;
        if (loads4) std::memcpy(loads4, ss.loads, sizeof(ss.loads));
    });
}

dvm_status dvm_batch_get_state(dvm_batch_t B, int slot, double* x, double* z, double* circ,
                               double* sigma, double* u, double* w)
{
    if (!B || slot < 0 || slot >= B->d.S) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        SlotState ss;
        hipStream_t st = B->ctx->stream;
        DVM_HIP_CHECK(hipMemcpyAsync(&ss, B->d.st + slot, sizeof(ss), hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        const int64_t off = (int64_t)slot * B->d.cap;
        const size_t b = sizeof(double) * ss.n;
        double* outs[6] = {x, z, circ, sigma, u, w};
        const double* srcs[6] = {B->d.b.x, B->d.b.z, B->d.b.c, B->d.b.s, B->d.b.u, B->d.b.w};
        for (int k = 0; k < 6; ++k)
            if (outs[k] && ss.n) DVM_HIP_CHECK(hipMemcpyAsync(outs[k], srcs[k] + off, b, hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
    });
}

dvm_status dvm_batch_sort_order(dvm_batch_t B, int slot, int64_t n, uint32_t* order)
{
    if (!B || slot < 0 || slot >= B->d.S) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        require(B->cfg.prec == DVM_F32C, DVM_E_STATE, "only the FP32C sum sorts");
        require(n >= 0 && n <= B->d.cap, DVM_E_SIZE, "n exceeds the slot capacity");
        require(n == 0 || order, DVM_E_INVALID, "null array");
        hipStream_t st = B->ctx->stream;
        if (n)
            DVM_HIP_CHECK(hipMemcpyAsync(order, B->d.perm + (int64_t)slot * B->d.cap,
                                         sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
    });
}

dvm_status dvm_batch_set_state(dvm_batch_t B, int slot, int64_t n, const double* x,
                               const double* z, const double* circ, const double* sigma)
{
    if (!B || slot < 0 || slot >= B->d.S) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        require(B->busy[slot], DVM_E_STATE, "slot not admitted");
        require(n >= 0 && n <= B->d.cap, DVM_E_SIZE, "blob count exceeds the slot capacity");
        require(n == 0 || (x && z && circ && sigma), DVM_E_INVALID, "null array");
        hipStream_t st = B->ctx->stream;
        const int64_t off = (int64_t)slot * B->d.cap;
        const size_t b = sizeof(double) * n;
        SlotState ss;
        DVM_HIP_CHECK(hipMemcpyAsync(&ss, B->d.st + slot, sizeof(ss), hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        if (n > 0) {
            const double* srcs[4] = {x, z, circ, sigma};
            double* dsts[4] = {B->d.b.x, B->d.b.z, B->d.b.c, B->d.b.s};
            for (int k = 0; k < 4; ++k)
                DVM_HIP_CHECK(hipMemcpyAsync(dsts[k] + off, srcs[k], b, hipMemcpyHostToDevice, st));
            for (double* v : {B->d.b.u, B->d.b.w, B->d.b.uvs, B->d.b.wvs})
                DVM_HIP_CHECK(hipMemsetAsync(v + off, 0, b, st));
            std::vector<uint32_t> ids((size_t)n);
            for (int64_t i = 0; i < n; ++i) ids[(size_t)i] = (uint32_t)i;
            DVM_HIP_CHECK(hipMemcpyAsync(B->d.b.id + off, ids.data(), sizeof(uint32_t) * n,
                                         hipMemcpyHostToDevice, st));
            DVM_HIP_CHECK(hipStreamSynchronize(st));  // ids lives on the host stack
        }
        ss.n = n;
        DVM_HIP_CHECK(hipMemcpyAsync(B->d.st + slot, &ss, sizeof(ss), hipMemcpyHostToDevice, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        B->hst[slot] = ss;
    });
}

dvm_status dvm_batch_set_fmm(dvm_batch_t B, int64_t min_blobs, int expansion_order, double theta)
{
    if (!B) return DVM_E_INVALID;
    return guard(B->ctx, [&] {
        require(min_blobs >= 0, DVM_E_INVALID, "min_blobs must be >= 0");
        if (min_blobs > 0) {
            require(B->cfg.prec == DVM_F32C, DVM_E_INVALID,
                    "the fast multipole slots need an FP32C batch");
            require(B->cfg.scheme == DVM_EULER, DVM_E_INVALID,
                    "the fast multipole slots need an Euler batch (one velocity sum per step)");
            require(expansion_order >= 1 && fmm_terms(expansion_order) > 0, DVM_E_INVALID,
                    "FMM expansion order must be 1..32");
            require(theta > 0.2 && theta <= 1.0, DVM_E_INVALID, "FMM theta must be in (0.2, 1]");
            if (!B->pst)
                DVM_HIP_CHECK(hipHostMalloc((void**)&B->pst, sizeof(SlotState) * B->d.S,
                                            hipHostMallocDefault));
            while ((int)B->fmm_ctx.size() < kFmmStreams) {
                dvm_ctx_t c = nullptr;
                const dvm_status cs = dvm_ctx_create(B->ctx->device, nullptr, &c);
                require(cs == DVM_OK, cs, "cannot create a context for the fast multipole sums");
                B->fmm_ctx.push_back(c);
                hipEvent_t e = nullptr;
                DVM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                B->fmm_ev.push_back(e);
            }
            for (dvm_ctx_t c : B->fmm_ctx) {
                c->fmm_order = expansion_order;
                c->fmm_theta = theta;
            }
        }
        // no slot keeps a flag from an earlier setting
        poll_states(B);
        for (SlotState& ss : B->hst) ss.ext = 0;
        DVM_HIP_CHECK(hipMemcpyAsync(B->d.st, B->hst.data(), sizeof(SlotState) * B->d.S,
                                     hipMemcpyHostToDevice, B->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
        B->fmm_min = min_blobs;
        B->fmm_order = expansion_order;
        B->fmm_theta = theta;
    });
}

}  // extern "C"
