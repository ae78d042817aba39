// Regularised Biot-Savart direct sum for gfx950.
//
// Two arithmetic forms behind one entry point (dvm_bs_self / dvm_bs_targets):
//
//  DVM_F64  (parity form) — one target per lane, sources streamed with
//           wave-uniform addresses so they arrive through the scalar cache
//           into SGPRs; the per-target loop visits sources in index order and
//           evaluates each pair as src/VelocityKernel.cpp:42-61 does, in FP64
//           with the reference's near/far test exactly, 1/r^2 within a few
//           units in the last place (one Newton-refined reciprocal shared by
//           four pairs) and the sums as fused multiply-adds.  When there are
//           too few targets to fill 256 CUs the source range is split into
//           chunks whose partial sums are added in chunk order.
//
//  DVM_F32C (fast form) — sources are Morton-sorted, cut into tiles (128 or
//           256) and stored as FP32 offsets from their tile centre (FP64).
//           For each source tile every target is recentred on that centre in
//           FP64 and rounded once, so a pair's dx is the difference of two
//           small FP32 offsets: the cancellation that plain FP32 would suffer
//           at |x| ~ 15 with 0.01 cores is avoided.  A tile pair whose
//           bounding boxes (source tile, target block of 64 or 256) are
//           farther apart than sqrt(10)*max(sigma) takes the far-field loop
//           (1/r^2 only: 7 VALU ops per pair and one reciprocal per 1, 2,
//           4 or 8 pairs); the others take the full regularised loop (near_k:
//           Taylor series or hardware exponential for 1 - e^-x).  Partial
//           sums within a tile accumulate in FP32 with FMA;
//           tile partials are added with Kahan compensation and the result
//           is returned in FP64.  The variant (tile, block, sources per
//           reciprocal, occupancy target) comes from a size table or is
//           measured on first use (dvm_set_tuning).
#include "dvm_device.hpp"
#include "dvm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace dvmi {

constexpr int kBlock = 256;  // targets per FP64 block (one per lane, 4 waves)

// ----------------------------------------------------------------- FP64 form

// Per-source record of the FP64 form: position, circulation and the
// near-field threshold thr = 10 * sigma^2 exactly as the reference forms it
// per pair (src/VelocityKernel.cpp:48, sigma^2 = sigma * sigma), so a pair is
// near or far exactly when it is in the reference.  Inside the near branch
// 1/sigma^2 is taken as 10 / thr, within an ulp of the reference's 1/sigma^2.
// Records of the FP64 form are padded by this many entries so the prefetch of
// a group past the end of a source range stays inside the allocation.
constexpr int kRec64Pad = 8;

// Writes n records and kRec64Pad zero records after them: a group's unused
// slots then hold finite values, whose terms are multiplied by zero.
__global__ void src_rec64_kernel(int64_t n, const double* __restrict__ x,
                                 const double* __restrict__ z, const double* __restrict__ c,
                                 const double* __restrict__ s, double4* __restrict__ rec)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n + kRec64Pad) return;
    if (j >= n) {
        rec[j] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const double sj = s[j];
    rec[j] = make_double4(x[j], z[j], c[j], 10.0 * (sj * sj));
}

struct Rec64x4 {
    double4 r[4];
};

// One target per lane; sources in index order, four at a time, through
// wave-uniform (scalar) loads with the next group's records in flight while
// the current group is summed.  The pair is that of
// src/VelocityKernel.cpp:44-60 with K = num * (1/r^2) in place of num / r^2,
// the group's four 1/r^2 from one reciprocal, and the two accumulations as
// fused multiply-adds: each pair term is within a few units in the last
// place of the reference's; the far pairs of a group run without branches
// and only a group with a near pair in some lane enters the near-field
// branch.
template <bool SELF, bool GUARD>
__global__ __launch_bounds__(kBlock) void bs_f64_kernel(
    int64_t nt, const double* __restrict__ tx, const double* __restrict__ tz, int64_t ns,
    const double4* __restrict__ rec, int64_t chunk, double r2_skip, double* __restrict__ u,
    double* __restrict__ w, double* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t it = i < nt ? i : nt - 1;
    const double xi = tx[it];
    const double zi = tz[it];
    const int64_t j0 = (int64_t)blockIdx.y * chunk;
    const int64_t j1 = j0 + chunk < ns ? j0 + chunk : ns;
    double ui = 0.0;
    double wi = 0.0;

    // Sum of the pairs with sources g.r[0..cnt-1] (source index jb + q).
    auto group = [&](const Rec64x4& g, int cnt, int64_t jb, bool self_check) {
        double dx[4], dz[4], r2[4], K[4];
        bool take[4];
        // lanes with a near pair in the group (ballots ORed on the scalar unit)
        uint64_t nearm = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            dx[q] = xi - g.r[q].x;
            dz[q] = zi - g.r[q].y;
            r2[q] = __builtin_fma(dx[q], dx[q], dz[q] * dz[q]);
            take[q] = q < cnt;
            if (SELF) take[q] = take[q] && !(self_check && jb + q == i);
            if (GUARD) take[q] = take[q] && (r2[q] > r2_skip);
            nearm |= __builtin_amdgcn_ballot_w64(take[q] && r2[q] < g.r[q].w);
        }
        // 1/r2 of the four from one reciprocal of their product, 1/ra =
        // rb rc rd / (ra rb rc rd), when the partial products are normal
        // numbers (then every factor is, and each 1/r2 is within a few units
        // in the last place); else one reciprocal per pair.
        const double pab = r2[0] * r2[1];
        const double pcd = r2[2] * r2[3];
        const double pp = pab * pcd;
        const double inv = recip64(pp);
        const double iab = inv * pcd;
        const double icd = inv * pab;
        K[0] = r2[1] * iab;
        K[1] = r2[0] * iab;
        K[2] = r2[3] * icd;
        K[3] = r2[2] * icd;
        const bool shared = fmin(pab, pcd) > 1e-300 && pp > 1e-300 && pp < 1e300;
        if (__builtin_amdgcn_ballot_w64(!shared) != 0) {  // wave-uniform, rare
            // kept a branch: the four reciprocals are not computed ahead
            asm volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < 4; ++q) K[q] = shared ? K[q] : recip64(r2[q]);
        }
        if (nearm != 0) {  // wave-uniform: the group's near-field correction
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool near = take[q] && r2[q] < g.r[q].w;
                const double f = 1.0 - exp(-r2[q] * (10.0 / g.r[q].w));
                K[q] = near ? K[q] * f : K[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double kc = take[q] ? K[q] * g.r[q].z : 0.0;
            ui = __builtin_fma(-kc, dz[q], ui);
            wi = __builtin_fma(kc, dx[q], wi);
        }
    };

    // Sources jb .. je-1; the records of the next group are loaded before the
    // current group is summed (they may lie past je: the array is padded).
    auto range = [&](int64_t jb, int64_t je, bool self_check) {
        if (jb >= je) return;
        const Rec64x4* q = reinterpret_cast<const Rec64x4*>(rec + jb);
        Rec64x4 a = q[0];
        int64_t j = jb;
        for (; j + 4 <= je; j += 4) {
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
            __builtin_amdgcn_sched_barrier(0);
            const Rec64x4 b = *reinterpret_cast<const Rec64x4*>(
                rec + __builtin_amdgcn_readfirstlane((int)(j + 4)));
            __builtin_amdgcn_sched_barrier(0);
            group(a, 4, j, self_check);
            __builtin_amdgcn_sched_barrier(0);
            a = b;
        }
        if (j < je) group(a, (int)(je - j), j, self_check);
    };

    if (SELF) {
        // Only the block's own index range can contain j == i.
        const int64_t b0 = (int64_t)blockIdx.x * kBlock;
        const int64_t b1 = b0 + kBlock < nt ? b0 + kBlock : nt;
        const int64_t a_end = j1 < b0 ? j1 : b0;
        range(j0, a_end, false);
        const int64_t m_beg = j0 > b0 ? j0 : b0;
        const int64_t m_end = j1 < b1 ? j1 : b1;
        range(m_beg, m_end, true);
        range(j0 > b1 ? j0 : b1, j1, false);
    } else {
        range(j0, j1, false);
    }

    if (i < nt) {
        if (part == nullptr) {
            u[i] = ui * kRpi2;
            w[i] = wi * kRpi2;
        } else {
            part[((int64_t)blockIdx.y * 2 + 0) * nt + i] = ui;
            part[((int64_t)blockIdx.y * 2 + 1) * nt + i] = wi;
        }
    }
}

// Adds chunk partials in chunk order and applies 1/(2 pi).  If out_index is
// given, result i goes to out_index[i].
__global__ void bs_reduce_kernel(int64_t nt, int nchunks, const double* __restrict__ part,
                                 const uint32_t* __restrict__ out_index, double* __restrict__ u,
                                 double* __restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double su = 0.0, sw = 0.0;
    // Loads of 8 chunks are issued together; the adds stay in chunk order.
    constexpr int kU = 8;
    int c = 0;
    for (; c + kU <= nchunks; c += kU) {
        double a[kU], b[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            a[k] = part[((int64_t)(c + k) * 2 + 0) * nt + i];
            b[k] = part[((int64_t)(c + k) * 2 + 1) * nt + i];
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            su += a[k];
            sw += b[k];
        }
    }
    for (; c < nchunks; ++c) {
        su += part[((int64_t)c * 2 + 0) * nt + i];
        sw += part[((int64_t)c * 2 + 1) * nt + i];
    }
    const int64_t o = out_index ? (int64_t)out_index[i] : i;
    u[o] = su * kRpi2;
    w[o] = sw * kRpi2;
}

// The same sum for few targets (probes, collocation points), where one thread
// per target leaves most CUs idle and each thread walks hundreds of chunks:
// a workgroup takes 64 consecutive targets (lane x, coalesced loads) and 16
// chunk groups (y); thread (x, y) adds chunks y, y + 16, ... in order, and the
// 16 group sums are added in group order through LDS.
constexpr int kRedGroups = 16;
__global__ __launch_bounds__(64 * kRedGroups) void bs_reduce_few_kernel(
    int64_t nt, int nchunks, const double* __restrict__ part, const uint32_t* __restrict__ out_index,
    double* __restrict__ u, double* __restrict__ w)
{
    __shared__ double ru[kRedGroups][64], rw[kRedGroups][64];
    const int x = threadIdx.x, y = threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 64 + x;
    double su = 0.0, sw = 0.0;
    if (i < nt) {
        // loads of 8 chunks issued together, the adds in chunk order
        constexpr int kU = 8;
        int c = y;
        for (; c + (kU - 1) * kRedGroups < nchunks; c += kU * kRedGroups) {
            double a[kU], b[kU];
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                a[k] = part[((int64_t)(c + k * kRedGroups) * 2 + 0) * nt + i];
                b[k] = part[((int64_t)(c + k * kRedGroups) * 2 + 1) * nt + i];
            }
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                su += a[k];
                sw += b[k];
            }
        }
        for (; c < nchunks; c += kRedGroups) {
            su += part[((int64_t)c * 2 + 0) * nt + i];
            sw += part[((int64_t)c * 2 + 1) * nt + i];
        }
    }
    ru[y][x] = su;
    rw[y][x] = sw;
    __syncthreads();
    if (y == 0 && i < nt) {
        double tu = 0.0, tw = 0.0;
        for (int g = 0; g < kRedGroups; ++g) {
            tu += ru[g][x];
            tw += rw[g][x];
        }
        const int64_t o = out_index ? (int64_t)out_index[i] : i;
        u[o] = tu * kRpi2;
        w[o] = tw * kRpi2;
    }
}

// Chunk partials to velocities: one thread per target, or, when that fills
// fewer workgroups than the device has CUs, bs_reduce_few_kernel.
static void launch_reduce(dvm_ctx_t ctx, int64_t nt, int nch, const double* part,
                          const uint32_t* out_index, double* u, double* w)
{
    if (cdiv(nt, 256) < ctx->num_cus && nch >= 2 * kRedGroups)
        bs_reduce_few_kernel<<<(unsigned)cdiv(nt, 64), dim3(64, kRedGroups), 0, ctx->stream>>>(
            nt, nch, part, out_index, u, w);
    else
        bs_reduce_kernel<<<cdiv(nt, 256), 256, 0, ctx->stream>>>(nt, nch, part, out_index, u, w);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- F32C form

__global__ void bbox_partial_kernel(int64_t n, const double* __restrict__ x,
                                    const double* __restrict__ z, double4* __restrict__ part)
{
    __shared__ double4 red[kBlock];
    double4 b = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i], zi = z[i];
        b.x = fmin(b.x, xi); b.y = fmax(b.y, xi);
        b.z = fmin(b.z, zi); b.w = fmax(b.w, zi);
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double4 o = red[threadIdx.x + s];
            double4& m = red[threadIdx.x];
            m.x = fmin(m.x, o.x); m.y = fmax(m.y, o.y);
            m.z = fmin(m.z, o.z); m.w = fmax(m.w, o.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Morton key on a square 65536^2 grid spanning the bounding box.
__global__ void morton_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ z,
                              int nparts, const double4* __restrict__ part,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ idx)
{
    double4 b = part[0];
    for (int p = 1; p < nparts; ++p) {
        const double4 o = part[p];
        b.x = fmin(b.x, o.x); b.y = fmax(b.y, o.y);
        b.z = fmin(b.z, o.z); b.w = fmax(b.w, o.w);
    }
    const double ext = fmax(fmax(b.y - b.x, b.w - b.z), 1e-300);
    const double scale = 65535.0 / ext;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double qx = (x[i] - b.x) * scale;
    double qz = (z[i] - b.z) * scale;
    qx = qx >= 0.0 ? (qx <= 65535.0 ? qx : 65535.0) : 0.0;  // NaN -> 0
    qz = qz >= 0.0 ? (qz <= 65535.0 ? qz : 65535.0) : 0.0;
    keys[i] = spread16((uint32_t)qx) | (spread16((uint32_t)qz) << 1);
    idx[i] = (uint32_t)i;
}

// One block per tile of T consecutive (possibly permuted) points: FP64 centre
// and half extents; for sources also the FP32 offsets from the centre with
// circulation, the FP32 near-field constants and an FP64 sorted copy of the
// positions (the blob-on-blob targets).
template <int T>
__global__ __launch_bounds__(T) void tile_build_kernel(
    int64_t n, const uint32_t* __restrict__ perm, const double* __restrict__ x,
    const double* __restrict__ z, const double* __restrict__ c, const double* __restrict__ s,
    float4* __restrict__ p32, float4* __restrict__ near32, double2* __restrict__ p64,
    TileInfo* __restrict__ tiles)
{
    __shared__ double red[6][T];
    const int64_t k = (int64_t)blockIdx.x * T + threadIdx.x;
    const bool valid = k < n;
    const int64_t src = valid ? (perm ? (int64_t)perm[k] : k) : 0;
    const double xi = valid ? x[src] : 0.0;
    const double zi = valid ? z[src] : 0.0;
    const double si = (valid && s) ? s[src] : 0.0;
    red[0][threadIdx.x] = valid ? xi : INFINITY;
    red[1][threadIdx.x] = valid ? xi : -INFINITY;
    red[2][threadIdx.x] = valid ? zi : INFINITY;
    red[3][threadIdx.x] = valid ? zi : -INFINITY;
    red[4][threadIdx.x] = si * si;
    red[5][threadIdx.x] = valid ? si * si : INFINITY;
    __syncthreads();
    for (int st = T / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][o]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][o]);
            red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][o]);
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][o]);
            red[4][threadIdx.x] = fmax(red[4][threadIdx.x], red[4][o]);
            red[5][threadIdx.x] = fmin(red[5][threadIdx.x], red[5][o]);
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
        tiles[blockIdx.x] = ti;
    }
    if (!valid || !p32) return;
    const double ci = c[src];
    const double s2 = si * si;
    const float thr = (float)(10.0 * s2);
    const double h = 0.5 * fmax(red[1][0] - red[0][0], red[3][0] - red[2][0]);
    p32[k] = make_float4((float)(xi - cx), (float)(zi - cz), (float)ci,
                         near_thr_hi(thr, h, 10.0 * red[4][0]));
    const double ox = xi - cx, oz = zi - cz;
    near32[k] = make_float4(thr, (float)(1.0 / s2), (float)(ox - (double)(float)ox),
                            (float)(oz - (double)(float)oz));
    p64[k] = make_double2(xi, zi);
}

// TILE sources per source tile, BLOCK targets per block (one per lane),
// SHARE sources per far-field reciprocal (SHARE 4 takes 8 where every r^2
// of the tile pair lies in [1e-4, 1e4]).  Targets: SELF -> the sorted FP64
// copy p64 (target i is source i); otherwise tx, tz in the caller's order.
// tblk: bounding box of each target block (the near/far test).  Per source
// tile each target is recentred on the tile centre in FP64 and rounded once,
// so a pair's dx = (x_i - c) - (x_j - c) carries only FP32 rounding relative
// to |x - c| (a few 1e-8 absolute for near pairs, whose tiles span tenths).
template <int TILE, int BLOCK, int SHARE, bool SELF, bool GUARD>
__global__ __launch_bounds__(BLOCK) void bs_f32c_kernel(
    int64_t nt, const double* __restrict__ tx, const double* __restrict__ tz,
    const TileInfo* __restrict__ tblk, int64_t ns, const float4* __restrict__ s32,
    const float4* __restrict__ near32, const double2* __restrict__ p64,
    const TileInfo* __restrict__ stile, int ntiles, int tiles_per_chunk, float r2_skip,
    const uint32_t* __restrict__ out_index, double* __restrict__ u, double* __restrict__ w,
    double* __restrict__ part)
{
    static_assert(TILE % 16 == 0 && BLOCK % 64 == 0, "tile and block shapes");
    static_assert(SHARE == 1 || SHARE == 2 || SHARE == 4, "sources per reciprocal");
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int64_t it = i < nt ? i : nt - 1;
    double xi, zi;
    if (SELF) {
        const double2 p = p64[it];
        xi = p.x;
        zi = p.y;
    } else {
        xi = tx[it];
        zi = tz[it];
    }
    const TileInfo tb = tblk[blockIdx.x];
    float su = 0.0f, cu = 0.0f, sw = 0.0f, cw = 0.0f;
    const int s0 = blockIdx.y * tiles_per_chunk;
    const int s1 = s0 + tiles_per_chunk < ntiles ? s0 + tiles_per_chunk : ntiles;

    for (int st = s0; st < s1; ++st) {
        const TileInfo ts = stile[st];
        const double Dxd = ts.cx - tb.cx;
        const double Dzd = ts.cz - tb.cz;
        const double gx = fmax(fabs(Dxd) - (double)(tb.hx + ts.hx), 0.0);
        const double gz = fmax(fabs(Dzd) - (double)(tb.hz + ts.hz), 0.0);
        // Far tile pair: every pair has r2 > 10 sigma_j^2 (and > r2_skip when
        // guarded), so the 1/r^2 branch holds for all of them.
        const double lim = 1.0001 * (double)ts.thr_max + 1e-300;
        const bool near = gx * gx + gz * gz <= (GUARD ? fmax(lim, 1.0001 * (double)r2_skip) : lim);
        const float txd = (float)(xi - ts.cx);
        const float tzd = (float)(zi - ts.cz);
        const int64_t jb = (int64_t)st * TILE;
        const int cnt = ns - jb < TILE ? (int)(ns - jb) : TILE;
        const float4* __restrict__ sp = s32 + jb;
        float au = 0.0f, aw = 0.0f;
        if (!near) {
            bool shared = false;
            if (SHARE > 1) {
                // r^2 of every pair lies in [gap^2, ext^2]; a product of
                // SHARE of them stays within [1e-34, 1e34].
                const double ex = fabs(Dxd) + (double)(tb.hx + ts.hx);
                const double ez = fabs(Dzd) + (double)(tb.hz + ts.hz);
                const double lo = SHARE == 2 ? 1e-17 : 1e-8;
                const double hi = SHARE == 2 ? 1e17 : 1e8;
                shared = cnt == TILE && gx * gx + gz * gz > lo && ex * ex + ez * ez < hi;
            }
            if (SHARE == 2 && shared) {
                far_tile_share2<TILE>(sp, txd, tzd, au, aw);
            } else if (SHARE == 4 && shared) {
                const double ex = fabs(Dxd) + (double)(tb.hx + ts.hx);
                const double ez = fabs(Dzd) + (double)(tb.hz + ts.hz);
                if (gx * gx + gz * gz > 1e-4 && ex * ex + ez * ez < 1e4)
                    far_tile_share8<TILE>(sp, txd, tzd, au, aw);
                else
                    far_tile_share4<TILE>(sp, txd, tzd, au, aw);
            } else if (cnt == TILE) {
                far_tile<TILE>(sp, txd, tzd, au, aw);
            } else {
                for (int k = 0; k < cnt; ++k) far_pair(sp[k], txd, tzd, au, aw);
            }
        } else {
            const float txl = (float)((xi - ts.cx) - (double)txd);
            const float tzl = (float)((zi - ts.cz) - (double)tzd);
            // only tiles overlapping the block's own indices hold self pairs
            const int64_t b0 = (int64_t)blockIdx.x * BLOCK;
            const bool self_tile = SELF && jb < b0 + BLOCK && jb + TILE > b0;
            if (!self_tile && !GUARD && cnt == TILE) {
                const double ex = fabs(Dxd) + (double)(tb.hx + ts.hx);
                const double ez = fabs(Dzd) + (double)(tb.hz + ts.hz);
                const bool share = ts.thr_min > 1e-8f && ex * ex + ez * ez < 1e8;
                near_tile_share4<TILE>(sp, near32 + jb, txd, tzd, txl, tzl, share, au, aw);
            } else if (self_tile) {
                near_range<SELF, GUARD>(sp, near32 + jb, 0, cnt, txd, tzd, txl, tzl, jb, i,
                                        r2_skip, au, aw);
            } else {
                near_range<false, GUARD>(sp, near32 + jb, 0, cnt, txd, tzd, txl, tzl, jb, i,
                                         r2_skip, au, aw);
            }
        }
        kahan_add(su, cu, au);
        kahan_add(sw, cw, aw);
    }

    if (i < nt) {
        const double U = (double)su - (double)cu;
        const double W = (double)sw - (double)cw;
        if (part == nullptr) {
            const int64_t o = out_index ? (int64_t)out_index[i] : i;
            u[o] = -U * kRpi2;
            w[o] = W * kRpi2;
        } else {
            part[((int64_t)blockIdx.y * 2 + 0) * nt + i] = -U;
            part[((int64_t)blockIdx.y * 2 + 1) * nt + i] = W;
        }
    }
}

// ------------------------------------------------- F32C, matrix-core far field
//
// A wave owns 32 targets (8 quads of 4); a block of 4 waves stages each
// source tile once into LDS (double-buffered, one barrier per tile).  For a
// tile that is far from the wave's target box and well separated from it
// (gap >= kappa * the larger of the tile's half-diagonal and the reach of the
// wave's box from the block's reference point), the tile's pair sums go
// through v_mfma_f32_4x4x1f32 (16 blocks of 4x4, K = 1; each step rounds
// once, like fmaf):
//   (1) r^2 for 64 sources x 4 targets: four accumulating K = 1 steps of
//       r^2 = |t'|^2 + |s'|^2 - 2 t'.s' with t', s' relative to the source
//       tile centre (A: source features, lane = source; B: target features,
//       lane & 3 = target);
//   (2) W = 1/r^2 on the VALU, in place in the result registers;
//   (3) sum_j W * circ_j * (1, x_j - c, z_j - c) with c the block's
//       reference point: four K = 1 steps per 64 sources, A = W (register r
//       holds source 4*(lane/4) + r, lane & 3 = target), B = the staged
//       features (lane & 3 = feature).  Results stay in 16 per-block
//       partials that are folded into FP64 accumulators after each tile and
//       summed over the blocks at the end: u = -(z'' F0 - F2),
//       w = x'' F0 - F1 with x'', z'' the target's offset from c.
// The separation bounds the rounding of both products to a few FP32 units
// relative to the pair term.  Other tiles take the VALU loops of the vector
// form, each lane covering one target and half of the tile's sources.

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kMT = 32;       // targets per wave
constexpr int kMQ = kMT / 4;  // target quads per wave
// Matrix form only for gap^2 >= kKappa2 * reach^2: with |t'| <= r + reach and
// |s'| <= reach the products round to within ~(1 + 2/kappa)^2 FP32 units.
constexpr float kKappa2 = 4.0f;

template <int TILE, bool SELF, bool GUARD>
__global__ __launch_bounds__(256) void bs_f32m_kernel(
    int64_t nt, const double* __restrict__ tx, const double* __restrict__ tz,
    const TileInfo* __restrict__ tgrp, int64_t ns, const float4* __restrict__ s32,
    const float4* __restrict__ near32, const double2* __restrict__ p64,
    const TileInfo* __restrict__ stile, int ntiles, int tiles_per_chunk, float r2_skip,
    float kappa2, const uint32_t* __restrict__ out_index, double* __restrict__ u,
    double* __restrict__ w, double* __restrict__ part)
{
    static_assert(TILE == 128 || TILE == 256, "a tile is 2 or 4 groups of 64 sources");
    __shared__ float4 la[2][TILE];         // (-2x', -2z', |s'|^2, 0) per source
    __shared__ float lf[2][4][TILE + 1];   // circ * (1, x - c, z - c, 0), feature-major
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int64_t grp = (int64_t)blockIdx.x * 4 + (tid >> 6);
    const int64_t t0 = grp * kMT;
    const bool active = t0 < nt;  // wave-uniform; idle waves still stage and synchronise
    const int c4 = lane & 3;
    const int tl = lane & 31;
    const int half = lane >> 5;

    for (int k = tid; k < 2 * (TILE + 1); k += 256) lf[k / (TILE + 1)][3][k % (TILE + 1)] = 0.0f;

    auto target = [&](int64_t idx, double& x, double& z) {
        idx = idx < nt ? idx : nt - 1;
        if (SELF) {
            const double2 p = p64[idx];
            x = p.x;
            z = p.y;
        } else {
            x = tx[idx];
            z = tz[idx];
        }
    };
    const TileInfo tref = tgrp[(int64_t)blockIdx.x * 4];
    const TileInfo tg = tgrp[active ? grp : (int64_t)blockIdx.x * 4];
    const int64_t i = t0 + tl;
    double xi, zi;
    target(i, xi, zi);
    float tqx[kMQ], tqz[kMQ];
#pragma unroll
    for (int q = 0; q < kMQ; ++q) {
        double x, z;
        target(t0 + 4 * q + c4, x, z);
        tqx[q] = (float)(x - tg.cx);
        tqz[q] = (float)(z - tg.cz);
    }
    // reach of this wave's targets from the reference point, squared
    const double ox = fabs(tg.cx - tref.cx) + (double)tg.hx;
    const double oz = fabs(tg.cz - tref.cz) + (double)tg.hz;
    const double reach_t2 = ox * ox + oz * oz;

    double E[kMQ][4];
#pragma unroll
    for (int q = 0; q < kMQ; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) E[q][r] = 0.0;
    float su = 0.0f, cu = 0.0f, sw = 0.0f, cw = 0.0f;
    const int s0 = blockIdx.y * tiles_per_chunk;
    const int s1 = s0 + tiles_per_chunk < ntiles ? s0 + tiles_per_chunk : ntiles;
    const v4f zero4 = {0.0f, 0.0f, 0.0f, 0.0f};

    auto stage_load = [&](int st) {
        const int64_t j = (int64_t)st * TILE + tid;
        return (tid < TILE && j < ns) ? s32[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    float4 pre = stage_load(s0);

    for (int st = s0; st < s1; ++st) {
        const int b = (st - s0) & 1;
        const TileInfo ts = stile[st];
        if (tid < TILE) {
            const float rx = (float)(ts.cx - tref.cx);
            const float rz = (float)(ts.cz - tref.cz);
            la[b][tid] = make_float4(-2.0f * pre.x, -2.0f * pre.y,
                                     __builtin_fmaf(pre.x, pre.x, pre.y * pre.y), 0.0f);
            lf[b][0][tid] = pre.z;
            lf[b][1][tid] = pre.z * (pre.x + rx);
            lf[b][2][tid] = pre.z * (pre.y + rz);
        }
        __syncthreads();
        if (st + 1 < s1) pre = stage_load(st + 1);
        if (!active) continue;

        const double Dxd = ts.cx - tg.cx;
        const double Dzd = ts.cz - tg.cz;
        const double gx = fmax(fabs(Dxd) - (double)(tg.hx + ts.hx), 0.0);
        const double gz = fmax(fabs(Dzd) - (double)(tg.hz + ts.hz), 0.0);
        const double g2 = gx * gx + gz * gz;
        const double lim = 1.0001 * (double)ts.thr_max + 1e-300;
        const bool near = g2 <= (GUARD ? fmax(lim, 1.0001 * (double)r2_skip) : lim);
        const int64_t jb = (int64_t)st * TILE;
        const int cnt = ns - jb < TILE ? (int)(ns - jb) : TILE;
        const double ext_s2 = (double)ts.hx * ts.hx + (double)ts.hz * ts.hz;
        const bool mat = !near && cnt == TILE && g2 > 1e-30 &&
                         g2 >= (double)kappa2 * fmax(ext_s2, reach_t2);
        if (mat) {
            // target box centre relative to the source tile centre
            const float dtx = (float)(tg.cx - ts.cx);
            const float dtz = (float)(tg.cz - ts.cz);
            float bx[kMQ], bz[kMQ], br[kMQ];
#pragma unroll
            for (int q = 0; q < kMQ; ++q) {
                bx[q] = tqx[q] + dtx;
                bz[q] = tqz[q] + dtz;
                br[q] = __builtin_fmaf(bx[q], bx[q], bz[q] * bz[q]);
            }
            v4f D2[kMQ];
#pragma unroll
            for (int q = 0; q < kMQ; ++q) D2[q] = zero4;
#pragma unroll 1
            for (int gg = 0; gg < TILE / 64; ++gg) {
                const float4 a = la[b][gg * 64 + lane];
                float bv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) bv[r] = lf[b][c4][gg * 64 + 4 * (lane >> 2) + r];
#pragma unroll
                for (int q = 0; q < kMQ; ++q) {
                    v4f d = __builtin_amdgcn_mfma_f32_4x4x1f32(1.0f, br[q], zero4, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_4x4x1f32(a.z, 1.0f, d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_4x4x1f32(a.x, bx[q], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_4x4x1f32(a.y, bz[q], d, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) d[r] = __builtin_amdgcn_rcpf(d[r]);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        D2[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(d[r], bv[r], D2[q], 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < kMQ; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) E[q][r] += (double)D2[q][r];
        } else {
            // VALU: this lane's target, sources [k0, k1) of the tile
            const int hs = TILE / 2;
            const int k0 = half * hs;
            const int k1 = cnt < k0 + hs ? cnt : k0 + hs;
            const float4* __restrict__ sp = s32 + jb;
            float au = 0.0f, aw = 0.0f;
            if (!near) {
                const float txd = (float)(xi - ts.cx);
                const float tzd = (float)(zi - ts.cz);
                for (int k = k0; k < k1; ++k) far_pair(sp[k], txd, tzd, au, aw);
            } else {
                const float txd = (float)(xi - ts.cx);
                const float tzd = (float)(zi - ts.cz);
                const float txl = (float)((xi - ts.cx) - (double)txd);
                const float tzl = (float)((zi - ts.cz) - (double)tzd);
                near_range<SELF, GUARD>(sp, near32 + jb, k0, k1, txd, tzd, txl, tzl, jb, i, r2_skip,
                                        au, aw);
            }
            kahan_add(su, cu, au);
            kahan_add(sw, cw, aw);
        }
    }
    if (!active) return;

    // VALU part: add the two source halves of each target.
    double U = (double)su - (double)cu;
    double W = (double)sw - (double)cw;
    U += __shfl_xor(U, 32);
    W += __shfl_xor(W, 32);
    // Matrix part: sum the 16 blocks (lanes of equal lane & 3), then lane t
    // (< 32) collects features 0..2 of its target 4q + r from lanes 0..2.
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
#pragma unroll
    for (int q = 0; q < kMQ; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double e = E[q][r];
            e += __shfl_xor(e, 4);
            e += __shfl_xor(e, 8);
            e += __shfl_xor(e, 16);
            e += __shfl_xor(e, 32);
            const double v0 = __shfl(e, 0);
            const double v1 = __shfl(e, 1);
            const double v2 = __shfl(e, 2);
            if (tl == 4 * q + r) {
                F0 = v0;
                F1 = v1;
                F2 = v2;
            }
        }
    }
    U += (zi - tref.cz) * F0 - F2;
    W += (xi - tref.cx) * F0 - F1;
    if (half == 0 && i < nt) {
        if (part == nullptr) {
            const int64_t o = out_index ? (int64_t)out_index[i] : i;
            u[o] = -U * kRpi2;
            w[o] = W * kRpi2;
        } else {
            part[((int64_t)blockIdx.y * 2 + 0) * nt + i] = -U;
            part[((int64_t)blockIdx.y * 2 + 1) * nt + i] = W;
        }
    }
}

// ----------------------------------------------------------------- launch

static int pick_chunks(dvm_ctx_t ctx, int64_t blocks_x, int64_t max_chunks)
{
    const int64_t want = 4LL * ctx->num_cus;
    int64_t c = (want + blocks_x - 1) / blocks_x;
    if (c > max_chunks) c = max_chunks;
    if (c < 1) c = 1;
    return (int)c;
}

static void launch_f64(dvm_ctx_t ctx, bool self, int64_t ns, const double* sx, const double* sz,
                       const double* sc, const double* ss, int64_t nt, const double* tx,
                       const double* tz, int guard, double r2_skip, double* u, double* w)
{
    const int64_t bx = cdiv(nt, kBlock);
    // Keep chunks of >= 2048 sources (>= 512 when the targets fill at most
    // 16 blocks — collocation points, probes — so that enough blocks run); a
    // single chunk keeps the reference's summation order.
    const int64_t min_chunk = bx <= 16 ? 512 : 2048;
    int64_t max_chunks = (ns + min_chunk - 1) / min_chunk;
    if (ctx->ws_fixed) {
        const int64_t m = (int64_t)(ctx->ws_fixed_partials / (2 * sizeof(double) * (size_t)nt));
        max_chunks = m < max_chunks ? (m < 1 ? 1 : m) : max_chunks;
    }
    const int nch = pick_chunks(ctx, bx, max_chunks);
    const int64_t chunk = (ns + nch - 1) / nch;
    const size_t part_bytes = nch > 1 ? sizeof(double) * 2 * nt * nch : 0;
    Carver cv(ws_get(ctx, Carver::need({sizeof(double4) * (size_t)(ns + kRec64Pad), part_bytes})));
    double4* rec = cv.take<double4>((size_t)(ns + kRec64Pad));
    double* part = nch > 1 ? cv.take<double>(2 * (size_t)nt * nch) : nullptr;
    dim3 grid((unsigned)bx, (unsigned)nch);
    {
        ProfScope ps(ctx, self ? "bs_self_f64" : "bs_targets_f64");
        src_rec64_kernel<<<cdiv(ns + kRec64Pad, 256), 256, 0, ctx->stream>>>(ns, sx, sz, sc, ss,
                                                                              rec);
        DVM_LAUNCH_CHECK();
        if (self) {
            bs_f64_kernel<true, false><<<grid, kBlock, 0, ctx->stream>>>(
                nt, tx, tz, ns, rec, chunk, r2_skip, u, w, part);
        } else if (guard) {
            bs_f64_kernel<false, true><<<grid, kBlock, 0, ctx->stream>>>(
                nt, tx, tz, ns, rec, chunk, r2_skip, u, w, part);
        } else {
            bs_f64_kernel<false, false><<<grid, kBlock, 0, ctx->stream>>>(
                nt, tx, tz, ns, rec, chunk, r2_skip, u, w, part);
        }
        DVM_LAUNCH_CHECK();
    }
    if (part) {
        launch_reduce(ctx, nt, nch, part, nullptr, u, w);
    }
}

// Arguments of one F32C pair-sum launch (host side).
struct F32cArgs {
    int64_t nt;
    const double* tx;
    const double* tz;
    const TileInfo* tblk;
    int64_t ns;
    const float4* s32;
    const float4* near32;
    const double2* p64;
    const TileInfo* stile;
    int ntiles;
    int tpc;
    float r2_skip;
    const uint32_t* out_index;
    double* u;
    double* w;
    double* part;
    float kappa2;  // matrix form only
};

// kind: 0 blob-on-blob, 1 targets, 2 targets with the r2_skip guard.
template <int TILE, int BLOCK, int SHARE>
static void f32c_dispatch(int kind, dim3 grid, hipStream_t st, const F32cArgs& a)
{
#define DVM_F32C_ARGS                                                                        \
    a.nt, a.tx, a.tz, a.tblk, a.ns, a.s32, a.near32, a.p64, a.stile, a.ntiles, a.tpc,        \
        a.r2_skip, a.out_index, a.u, a.w, a.part
    if (kind == 0)
        bs_f32c_kernel<TILE, BLOCK, SHARE, true, false><<<grid, BLOCK, 0, st>>>(DVM_F32C_ARGS);
    else if (kind == 2)
        bs_f32c_kernel<TILE, BLOCK, SHARE, false, true><<<grid, BLOCK, 0, st>>>(DVM_F32C_ARGS);
    else
        bs_f32c_kernel<TILE, BLOCK, SHARE, false, false><<<grid, BLOCK, 0, st>>>(DVM_F32C_ARGS);
#undef DVM_F32C_ARGS
}

// Matrix far-field form: 256-lane blocks of 4 waves x 32 targets.
template <int TILE>
static void f32m_dispatch(int kind, dim3 grid, hipStream_t st, const F32cArgs& a)
{
#define DVM_F32M_ARGS                                                                        \
    a.nt, a.tx, a.tz, a.tblk, a.ns, a.s32, a.near32, a.p64, a.stile, a.ntiles, a.tpc,        \
        a.r2_skip, a.kappa2, a.out_index, a.u, a.w, a.part
    if (kind == 0)
        bs_f32m_kernel<TILE, true, false><<<grid, 256, 0, st>>>(DVM_F32M_ARGS);
    else if (kind == 2)
        bs_f32m_kernel<TILE, false, true><<<grid, 256, 0, st>>>(DVM_F32M_ARGS);
    else
        bs_f32m_kernel<TILE, false, false><<<grid, 256, 0, st>>>(DVM_F32M_ARGS);
#undef DVM_F32M_ARGS
}

template <int T>
static void tile_build(hipStream_t st, int64_t n, const uint32_t* perm, const double* x,
                       const double* z, const double* c, const double* s, float4* p32,
                       float4* near32, double2* p64, TileInfo* tiles)
{
    tile_build_kernel<T><<<(unsigned)cdiv(n, T), T, 0, st>>>(n, perm, x, z, c, s, p32, near32,
                                                             p64, tiles);
}

using F32cFn = void (*)(int, dim3, hipStream_t, const F32cArgs&);
using TileFn = void (*)(hipStream_t, int64_t, const uint32_t*, const double*, const double*,
                        const double*, const double*, float4*, float4*, double2*, TileInfo*);

struct F32cKernel {
    int tile, block, share;
    F32cFn fn;
};

// The compiled variants: source tile x target block x sources per reciprocal.
static const F32cKernel kF32c[] = {
    {128, 64, 1, f32c_dispatch<128, 64, 1>},   {128, 64, 2, f32c_dispatch<128, 64, 2>},
    {128, 64, 4, f32c_dispatch<128, 64, 4>},   {128, 256, 1, f32c_dispatch<128, 256, 1>},
    {128, 256, 2, f32c_dispatch<128, 256, 2>}, {128, 256, 4, f32c_dispatch<128, 256, 4>},
    {256, 64, 1, f32c_dispatch<256, 64, 1>},   {256, 64, 2, f32c_dispatch<256, 64, 2>},
    {256, 64, 4, f32c_dispatch<256, 64, 4>},   {256, 256, 1, f32c_dispatch<256, 256, 1>},
    {256, 256, 2, f32c_dispatch<256, 256, 2>}, {256, 256, 4, f32c_dispatch<256, 256, 4>},
    // share 0: matrix-core far field, 32-target boxes
    {128, 32, 0, f32m_dispatch<128>},          {256, 32, 0, f32m_dispatch<256>},
};
constexpr int kNumF32c = sizeof(kF32c) / sizeof(kF32c[0]);
// Occupancy targets: 256-lane blocks per CU that splitting the source range
// into chunks aims for.  Beyond filling the chip, finer chunks keep the
// source slice that co-resident waves stream small (blocks are dispatched
// chunk by chunk), so it stays in the scalar cache and L2: on MI355X the
// 262,144-blob sum drops from 16.3 ms at 8 to 11.9 ms at 128.
constexpr int kOccupancy[] = {8, 32, 128};
constexpr int kOccupancyFirst = 32;          // used while every variant is timed
constexpr size_t kPartialBytes = 256u << 20;  // cap on the FP64 chunk partials

static TileFn tile_fn(int t)
{
    return t == 32    ? tile_build<32>
           : t == 64  ? tile_build<64>
           : t == 128 ? tile_build<128>
                      : tile_build<256>;
}

static int f32c_find(int tile, int block, int share)
{
    for (int k = 0; k < kNumF32c; ++k)
        if (kF32c[k].tile == tile && kF32c[k].block == block && kF32c[k].share == share) return k;
    return -1;
}

// One complete F32C evaluation with variant v (index into kF32c) and occupancy
// target occ: Morton sort of the sources, tile build, pair sum, chunk reduce.
static void f32c_run(dvm_ctx_t ctx, int kind, int v, int occ, int64_t ns, const double* sx,
                     const double* sz, const double* sc, const double* ss, int64_t nt,
                     const double* tx, const double* tz, double r2_skip, double* u, double* w)
{
    const F32cKernel& K = kF32c[v];
    const bool self = kind == 0;
    const bool matrix = K.share == 0;
    const int ntiles = (int)cdiv(ns, K.tile);
    const int64_t bx = cdiv(nt, K.block);  // target boxes
    // launch blocks: one box each, or 4 boxes of 32 (one per wave) in the
    // matrix form; the occupancy target counts 256-lane blocks per CU
    const int64_t gx = matrix ? cdiv(nt, 4 * kMT) : bx;
    const int nbb = 64;
    const int64_t want = (int64_t)occ * ctx->num_cus * 256 / (matrix ? 256 : K.block);
    int64_t nch64 = (want + gx - 1) / gx;
    const size_t part_cap = ctx->ws_fixed ? ctx->ws_fixed_partials : kPartialBytes;
    const int64_t nch_mem = (int64_t)(part_cap / (2 * sizeof(double) * (size_t)nt));
    nch64 = nch64 > nch_mem ? nch_mem : nch64;
    nch64 = nch64 > ntiles ? ntiles : (nch64 < 1 ? 1 : nch64);
    const int nch = (int)nch64;
    const int tpc = (ntiles + nch - 1) / nch;

    size_t sort_tmp = 0;
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (uint32_t*)nullptr,
                                                     (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                     (uint32_t*)nullptr, (int)ns, 0, 32,
                                                     ctx->stream));
    const size_t need = Carver::need({
        sizeof(uint32_t) * ns * 4, sizeof(double4) * nbb, sort_tmp, sizeof(float4) * ns,
        sizeof(float4) * ns, sizeof(double2) * ns, sizeof(TileInfo) * ntiles,
        sizeof(TileInfo) * bx, nch > 1 ? sizeof(double) * 2 * nt * nch : 0});
    Carver cv(ws_get(ctx, need));
    uint32_t* keys = cv.take<uint32_t>(ns);
    uint32_t* keys2 = cv.take<uint32_t>(ns);
    uint32_t* idx = cv.take<uint32_t>(ns);
    uint32_t* perm = cv.take<uint32_t>(ns);
    double4* bb = cv.take<double4>(nbb);
    void* tmp = cv.take<char>(sort_tmp);
    float4* s32 = cv.take<float4>(ns);
    float4* near32 = cv.take<float4>(ns);
    double2* p64 = cv.take<double2>(ns);
    TileInfo* stile = cv.take<TileInfo>(ntiles);
    TileInfo* tblk = cv.take<TileInfo>(bx);
    double* part = nch > 1 ? cv.take<double>(2 * nt * nch) : nullptr;

    bbox_partial_kernel<<<nbb, kBlock, 0, ctx->stream>>>(ns, sx, sz, bb);
    DVM_LAUNCH_CHECK();
    morton_kernel<<<cdiv(ns, 256), 256, 0, ctx->stream>>>(ns, sx, sz, nbb, bb, keys, idx);
    DVM_LAUNCH_CHECK();
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, sort_tmp, keys, keys2, idx, perm,
                                                     (int)ns, 0, 32, ctx->stream));
    tile_fn(K.tile)(ctx->stream, ns, perm, sx, sz, sc, ss, s32, near32, p64, stile);
    DVM_LAUNCH_CHECK();
    // Target blocks: the sorted sources themselves (self) or the targets in
    // the caller's order; bounding boxes only.
    tile_fn(K.block)(ctx->stream, nt, self ? perm : nullptr, self ? sx : tx, self ? sz : tz,
                     nullptr, nullptr, nullptr, nullptr, nullptr, tblk);
    DVM_LAUNCH_CHECK();
    F32cArgs a{nt,     self ? nullptr : tx, self ? nullptr : tz,  tblk, ns, s32, near32, p64,
               stile,  ntiles, tpc, (float)r2_skip, self ? perm : nullptr, u, w, part, kKappa2};
    dim3 grid((unsigned)gx, (unsigned)nch);
    {
        ProfScope ps(ctx, self ? "bs_self_f32c" : "bs_targets_f32c");
        K.fn(kind, grid, ctx->stream, a);
        DVM_LAUNCH_CHECK();
    }
    if (part) {
        launch_reduce(ctx, nt, nch, part, a.out_index, u, w);
    }
}

static int lg2_bucket(int64_t n)
{
    int b = 0;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

// Fixed choice (DVM_TUNE_TABLE), from sweeps of every variant and
// occupancy target on MI355X (uniform and cylinder-wake blob sets, 4K..1M
// blobs; 257 and 1024 targets over 16K and 262K sources): small sums favour
// few, wide blocks; probe and collocation sums 1-reciprocal tiles of 128 and
// 64-target blocks; large sums 64-target blocks (tight near/far boxes) and 4
// sources per reciprocal; everything gains from a finely split source range.
static void f32c_table(int kind, int64_t nt, int64_t ns, int& v, int& occ)
{
    if ((double)nt * (double)ns < 6.7e7) {
        v = f32c_find(128, 256, 2);
        occ = 8;
    } else if (kind != 0 && nt < 8192) {
        v = f32c_find(128, 64, 1);
        occ = 32;
    } else if (ns >= 196608) {
        v = f32c_find(256, 64, 4);
        occ = 128;
    } else {
        v = f32c_find(128, 64, 4);
        occ = 128;
    }
}

// DVM_TUNE_MEASURE: time every variant at the first occupancy target (after
// an untimed run), then the best one at the other targets; keep the fastest.
static void f32c_measure(dvm_ctx_t ctx, int kind, int64_t ns, const double* sx, const double* sz,
                         const double* sc, const double* ss, int64_t nt, const double* tx,
                         const double* tz, double r2_skip, double* u, double* w, int& v_best,
                         int& occ_best)
{
    hipEvent_t e0, e1;
    DVM_HIP_CHECK(hipEventCreate(&e0));
    DVM_HIP_CHECK(hipEventCreate(&e1));
    const bool prof = ctx->prof;
    ctx->prof = false;
    float best = 3.0e38f;
    auto time_one = [&](int v, int occ) {
        f32c_run(ctx, kind, v, occ, ns, sx, sz, sc, ss, nt, tx, tz, r2_skip, u, w);
        float t = 3.0e38f;
        for (int rep = 0; rep < 2; ++rep) {
            DVM_HIP_CHECK(hipEventRecord(e0, ctx->stream));
            f32c_run(ctx, kind, v, occ, ns, sx, sz, sc, ss, nt, tx, tz, r2_skip, u, w);
            DVM_HIP_CHECK(hipEventRecord(e1, ctx->stream));
            DVM_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DVM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            t = ms < t ? ms : t;
        }
        if (t < best) {
            best = t;
            v_best = v;
            occ_best = occ;
        }
    };
    try {
        for (int v = 0; v < kNumF32c; ++v) time_one(v, kOccupancyFirst);
        const int v0 = v_best;
        for (int occ : kOccupancy)
            if (occ != kOccupancyFirst) time_one(v0, occ);
    } catch (...) {
        ctx->prof = prof;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    ctx->prof = prof;
    DVM_HIP_CHECK(hipEventDestroy(e0));
    DVM_HIP_CHECK(hipEventDestroy(e1));
}

static void launch_f32c(dvm_ctx_t ctx, bool self, int64_t ns, const double* sx, const double* sz,
                        const double* sc, const double* ss, int64_t nt, const double* tx,
                        const double* tz, int guard, double r2_skip, double* u, double* w)
{
    const int kind = self ? 0 : (guard ? 2 : 1);
    int v = -1, occ = 4;
    if (ctx->f32c_forced[0] != 0) {
        v = f32c_find(ctx->f32c_forced[0], ctx->f32c_forced[1], ctx->f32c_forced[2]);
        occ = ctx->f32c_forced[3];
    } else if (ctx->tuning == DVM_TUNE_MEASURE && !ctx->ws_fixed) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        DVM_HIP_CHECK(hipStreamIsCapturing(ctx->stream, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            const auto key = std::make_tuple(kind, lg2_bucket(nt), lg2_bucket(ns));
            auto it = ctx->f32c_tune.find(key);
            if (it == ctx->f32c_tune.end()) {
                f32c_measure(ctx, kind, ns, sx, sz, sc, ss, nt, tx, tz, r2_skip, u, w, v, occ);
                it = ctx->f32c_tune.emplace(key, std::make_pair(v, occ)).first;
            }
            v = it->second.first;
            occ = it->second.second;
        }
    }
    if (v < 0) f32c_table(kind, nt, ns, v, occ);
    const F32cKernel& K = kF32c[v];
    ctx->f32c_last[0] = K.tile;
    ctx->f32c_last[1] = K.block;
    ctx->f32c_last[2] = K.share;
    ctx->f32c_last[3] = occ;
    f32c_run(ctx, kind, v, occ, ns, sx, sz, sc, ss, nt, tx, tz, r2_skip, u, w);
}

void f32c_set_variant(dvm_ctx_t ctx, int tile, int block, int share, int occupancy)
{
    if (tile == 0) {
        ctx->f32c_forced[0] = 0;
        return;
    }
    require(f32c_find(tile, block, share) >= 0, DVM_E_INVALID, "no such F32C variant");
    require(occupancy >= 1 && occupancy <= 1024, DVM_E_INVALID, "occupancy target out of range");
    ctx->f32c_forced[0] = tile;
    ctx->f32c_forced[1] = block;
    ctx->f32c_forced[2] = share;
    ctx->f32c_forced[3] = occupancy;
}

void bs_self_launch(dvm_ctx_t ctx, int prec, int64_t n, const double* x, const double* z,
                    const double* c, const double* s, double* u, double* w)
{
    require(prec == DVM_F64 || prec == DVM_F32C || prec == DVM_FMM, DVM_E_INVALID,
            "unknown precision");
    require(n >= 0 && n < (int64_t(1) << 31), DVM_E_SIZE, "blob count out of range");
    if (n == 0) return;
    require(x && z && c && s && u && w, DVM_E_INVALID, "null array");
    if (prec == DVM_FMM) {
        if (!fmm_self(ctx, n, x, z, c, s, u, w)) launch_f64(ctx, true, n, x, z, c, s, n, x, z, 0, 0.0, u, w);
    } else if (prec == DVM_F64)
        launch_f64(ctx, true, n, x, z, c, s, n, x, z, 0, 0.0, u, w);
    else
        launch_f32c(ctx, true, n, x, z, c, s, n, x, z, 0, 0.0, u, w);
}

void bs_targets_launch(dvm_ctx_t ctx, int prec, int64_t ns, const double* sx, const double* sz,
                       const double* sc, const double* ss, int64_t nt, const double* tx,
                       const double* tz, int guard, double r2_skip, double* u, double* w)
{
    require(prec == DVM_F64 || prec == DVM_F32C || prec == DVM_FMM, DVM_E_INVALID,
            "unknown precision");
    if (prec == DVM_FMM) prec = DVM_F64;  // FMMKernel::computeInducedAt is the direct sum
    require(ns >= 0 && nt >= 0 && ns < (int64_t(1) << 31) && nt < (int64_t(1) << 31), DVM_E_SIZE,
            "count out of range");
    if (nt == 0) return;
    require(tx && tz && u && w, DVM_E_INVALID, "null target array");
    if (ns == 0) {
        DVM_HIP_CHECK(hipMemsetAsync(u, 0, sizeof(double) * nt, ctx->stream));
        DVM_HIP_CHECK(hipMemsetAsync(w, 0, sizeof(double) * nt, ctx->stream));
        return;
    }
    require(sx && sz && sc && ss, DVM_E_INVALID, "null source array");
    if (prec == DVM_F64)
        launch_f64(ctx, false, ns, sx, sz, sc, ss, nt, tx, tz, guard, r2_skip, u, w);
    else
        launch_f32c(ctx, false, ns, sx, sz, sc, ss, nt, tx, tz, guard, r2_skip, u, w);
}

// ------------------------------------------- allocation-free blob-on-blob sum
//
// dvm_bs_self_noalloc runs dvm_bs_self's kernels with the scratch memory
// (sort keys, tiles, source copies, chunk partials, FP64 records) in this
// file-scope device array instead of the context's growable arena, so it
// calls no hipMalloc or hipFree.  Each GPU the library runs on holds its own
// copy until the process ends.  Every byte a call uses is written before it is
// read, so nothing needs clearing between calls; calls that overlap on one GPU
// share the array and race.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchPartialBytes = size_t(16) << 20;  // of which chunk partials
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on `device`, taken once per device with that
// device current (the host-side name is a stand-in, never a device address).
static void* scratch_address(int device)
{
    static std::mutex mu;
    static void* ptr[kScratchMaxDevices] = {};
    require(device >= 0 && device < kScratchMaxDevices, DVM_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        int cur = -1;
        DVM_HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) DVM_HIP_CHECK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) DVM_HIP_CHECK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            fail(DVM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr[device] = p;
    }
    return ptr[device];
}

// Points the context's arena at the scratch array for one call (table tuning,
// no kernel timing) and restores it afterwards.
struct FixedScratch {
    dvm_ctx_t ctx;
    bool prof;
    FixedScratch(dvm_ctx_t c, void* p) : ctx(c), prof(c->prof)
    {
        ctx->ws_fixed = p;
        ctx->ws_fixed_bytes = kScratchBytes;
        ctx->ws_fixed_partials = kScratchPartialBytes;
        ctx->prof = false;
    }
    ~FixedScratch()
    {
        ctx->ws_fixed = nullptr;
        ctx->ws_fixed_bytes = 0;
        ctx->ws_fixed_partials = 0;
        ctx->prof = prof;
    }
};

}  // namespace dvmi

using namespace dvmi;

extern "C" {

dvm_status dvm_bs_self_noalloc(dvm_ctx_t ctx, int prec, int64_t n, const double* x,
                               const double* z, const double* circ, const double* sigma,
                               double* u, double* w)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(prec == DVM_F64 || prec == DVM_F32C, DVM_E_INVALID,
                "dvm_bs_self_noalloc: precision must be DVM_F64 or DVM_F32C");
        void* p = scratch_address(ctx->device);
        FixedScratch fs(ctx, p);
        bs_self_launch(ctx, prec, n, x, z, circ, sigma, u, w);
    });
}

size_t dvm_bs_self_noalloc_scratch_bytes(void) { return kScratchBytes; }

dvm_status dvm_bs_self(dvm_ctx_t ctx, int prec, int64_t n, const double* x, const double* z,
                       const double* circ, const double* sigma, double* u, double* w)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] { bs_self_launch(ctx, prec, n, x, z, circ, sigma, u, w); });
}

dvm_status dvm_bs_targets(dvm_ctx_t ctx, int prec, int64_t ns, const double* sx,
                          const double* sz, const double* scirc, const double* ssigma,
                          int64_t nt, const double* tx, const double* tz, int guard_,
                          double r2_skip, double* u, double* w)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        bs_targets_launch(ctx, prec, ns, sx, sz, scirc, ssigma, nt, tx, tz, guard_, r2_skip, u, w);
    });
}

// Host AoS forms: upload, split AoS -> SoA on device, sum, AoS results back.
__global__ void aos4_to_soa_kernel(int64_t n, const double4* __restrict__ p, double* __restrict__ x,
                                   double* __restrict__ z, double* __restrict__ c,
                                   double* __restrict__ s)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 v = p[i];
    x[i] = v.x; z[i] = v.y; c[i] = v.z; s[i] = v.w;
}

__global__ void soa_to_aos2_kernel(int64_t n, const double* __restrict__ u,
                                   const double* __restrict__ w, double2* __restrict__ o)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    o[i] = make_double2(u[i], w[i]);
}

dvm_status dvm_bs_self_host(dvm_ctx_t ctx, int prec, int64_t n, const double* particles4,
                            double* results2)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        if (n == 0) return;
        require(particles4 && results2, DVM_E_INVALID, "null host array");
        char* buf = nullptr;
        const size_t b = sizeof(double) * n;
        DVM_HIP_CHECK(hipMallocAsync((void**)&buf, b * 11, ctx->stream));
        double4* p = reinterpret_cast<double4*>(buf);
        double* x = reinterpret_cast<double*>(buf + 4 * b);
        double *z = x + n, *c = z + n, *s = c + n, *u = s + n, *w = u + n;
        double2* o = reinterpret_cast<double2*>(buf);
        DVM_HIP_CHECK(hipMemcpyAsync(p, particles4, 4 * b, hipMemcpyHostToDevice, ctx->stream));
        aos4_to_soa_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, p, x, z, c, s);
        DVM_LAUNCH_CHECK();
        bs_self_launch(ctx, prec, n, x, z, c, s, u, w);
        soa_to_aos2_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, u, w, o);
        DVM_LAUNCH_CHECK();
        DVM_HIP_CHECK(hipMemcpyAsync(results2, o, 2 * b, hipMemcpyDeviceToHost, ctx->stream));
        DVM_HIP_CHECK(hipFreeAsync(buf, ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

dvm_status dvm_bs_targets_host(dvm_ctx_t ctx, int prec, int64_t ns, const double* sources4,
                               int64_t nt, const double* tx, const double* tz, double* results2)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(ns >= 0 && nt >= 0, DVM_E_SIZE, "negative count");
        if (nt == 0) return;
        require(tx && tz && results2 && (ns == 0 || sources4), DVM_E_INVALID, "null host array");
        char* buf = nullptr;
        const size_t bs = sizeof(double) * (ns > 0 ? ns : 1);
        const size_t bt = sizeof(double) * nt;
        DVM_HIP_CHECK(hipMallocAsync((void**)&buf, 8 * bs + 6 * bt, ctx->stream));
        double4* p = reinterpret_cast<double4*>(buf);
        double* sx = reinterpret_cast<double*>(buf + 4 * bs);
        double *sz = sx + (bs / 8), *sc = sz + (bs / 8), *ss = sc + (bs / 8);
        double* t = reinterpret_cast<double*>(buf + 8 * bs);
        double *dtx = t, *dtz = t + nt, *u = t + 2 * nt, *w = t + 3 * nt;
        double2* o = reinterpret_cast<double2*>(t + 4 * nt);
        if (ns > 0) {
            DVM_HIP_CHECK(hipMemcpyAsync(p, sources4, 4 * sizeof(double) * ns,
                                         hipMemcpyHostToDevice, ctx->stream));
            aos4_to_soa_kernel<<<cdiv(ns, 256), 256, 0, ctx->stream>>>(ns, p, sx, sz, sc, ss);
            DVM_LAUNCH_CHECK();
        }
        DVM_HIP_CHECK(hipMemcpyAsync(dtx, tx, bt, hipMemcpyHostToDevice, ctx->stream));
        DVM_HIP_CHECK(hipMemcpyAsync(dtz, tz, bt, hipMemcpyHostToDevice, ctx->stream));
        bs_targets_launch(ctx, prec, ns, sx, sz, sc, ss, nt, dtx, dtz, 1, 1e-15, u, w);
        soa_to_aos2_kernel<<<cdiv(nt, 256), 256, 0, ctx->stream>>>(nt, u, w, o);
        DVM_LAUNCH_CHECK();
        DVM_HIP_CHECK(hipMemcpyAsync(results2, o, 2 * bt, hipMemcpyDeviceToHost, ctx->stream));
        DVM_HIP_CHECK(hipFreeAsync(buf, ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

dvm_status dvm_set_tuning(dvm_ctx_t ctx, int mode)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(mode == DVM_TUNE_TABLE || mode == DVM_TUNE_MEASURE, DVM_E_INVALID,
                "unknown tuning mode");
        ctx->tuning = mode;
    });
}

dvm_status dvm_set_f32c_variant(dvm_ctx_t ctx, int tile, int block, int share, int occupancy)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] { f32c_set_variant(ctx, tile, block, share, occupancy); });
}

dvm_status dvm_get_f32c_variant(dvm_ctx_t ctx, int32_t* out4)
{
    if (!ctx || !out4) return DVM_E_INVALID;
    for (int k = 0; k < 4; ++k) out4[k] = ctx->f32c_last[k];
    return DVM_OK;
}

}  // extern "C"
