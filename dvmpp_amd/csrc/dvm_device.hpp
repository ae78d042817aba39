// Device helpers shared by several kernels (inline, one copy per translation unit).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dvm_internal.hpp"

namespace dvmi {

// Tile record of the FP32C sum: FP64 centre, FP32 half extents (rounded up)
// and, for a source tile, the largest and smallest 10 sigma^2 of its sources.
struct TileInfo {
    double cx, cz;
    float hx, hz;
    float thr_max;  // largest 10 sigma^2 of a source tile
    float thr_min;  // smallest 10 sigma^2 of a source tile
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based generator.
__device__ __forceinline__ uint4 philox4x32(uint4 ctr, uint2 key)
{
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, ctr.x);
        const uint32_t lo0 = 0xD2511F53u * ctr.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, ctr.z);
        const uint32_t lo1 = 0xCD9E8D57u * ctr.z;
        ctr = make_uint4(hi1 ^ ctr.y ^ key.x, lo1, hi0 ^ ctr.w ^ key.y, lo0);
        key.x += 0x9E3779B9u;
        key.y += 0xBB67AE85u;
    }
    return ctr;
}

// 53-bit uniform in (0, 1) from two 32-bit words.
__device__ __forceinline__ double u01(uint32_t a, uint32_t b)
{
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// src/VortexSheet.cpp:453-457: 1 when edge i (point i to i+1) straddles zc
// under the half-open rule and crosses z = zc right of xc.
__device__ __forceinline__ int edge_crossing(const double* __restrict__ bxp,
                                             const double* __restrict__ bzp, int i, double xc,
                                             double zc)
{
    const double zi = bzp[i], zi1 = bzp[i + 1];
    if (((zi <= zc) && (zi1 > zc)) || ((zi > zc) && (zi1 <= zc))) {
        const double vt = (zc - zi) / (zi1 - zi);
        if (xc < bxp[i] + vt * (bxp[i + 1] - bxp[i])) return 1;
    }
    return 0;
}

// src/VortexSheet.cpp:447-464: crossing number over the Nl edges; inside iff 1.
__device__ __forceinline__ bool inside_body(int nl, const double* __restrict__ bxp,
                                            const double* __restrict__ bzp, double xc, double zc)
{
    int cn = 0;
    for (int i = 0; i < nl; i++) cn += edge_crossing(bxp, bzp, i, xc, zc);
    return cn == 1;
}

// Points where inside_body is false without visiting the edges: z < zmin or
// z >= zmax (no edge straddles z under the half-open rule), x beyond the
// largest x by more than the rounding of any crossing abscissa, or, for a
// closed contour, below the smallest x by that margin (every straddling edge
// counts and a closed contour has an even number of them, so cn != 1).
__device__ __forceinline__ bool outside_box(const BodyBox& bb, double xc, double zc)
{
    return zc < bb.zlo || zc >= bb.zhi || xc > bb.xhi || xc < bb.xlo;
}

// inside_body with the bounding-box cull and the slab edge lists: only the
// edges listed for zc's slab can straddle zc (EdgeSlabs), and the crossing
// number is a count, so visiting them alone gives the same flag.
__device__ __forceinline__ bool inside_body_fast(int nl, const double* __restrict__ bxp,
                                                 const double* __restrict__ bzp,
                                                 const BodyBox& bb, const EdgeSlabs& es, double xc,
                                                 double zc)
{
    if (outside_box(bb, xc, zc)) return false;
    if (es.n == 0) return inside_body(nl, bxp, bzp, xc, zc);
    if (!(zc == zc)) return false;  // NaN: no edge straddles it
    const double f = floor((zc - es.z0) * es.inv_h);
    const int s = f < 0.0 ? 0 : (f > (double)(es.n - 1) ? es.n - 1 : (int)f);
    int cn = 0;
    for (int q = es.off[s]; q < es.off[s + 1]; ++q) cn += edge_crossing(bxp, bzp, es.edge[q], xc, zc);
    return cn == 1;
}

// src/VortexSheet.cpp:345-405 for the point (px, pz), computed by all 64
// lanes of a wavefront together (the point and the body are the same in
// every lane): lane l searches edges l, l+64, ... keeping the first strict
// minimum of the clamped-projection distance from 1e10 (the reference's
// start), and the wave keeps the smallest distance and, among equal ones,
// the smallest edge — the reference's first strict minimum over all edges
// (no edge closer than 1e10: edge 0).  The moved point's inside test is the
// same crossing count summed over the lanes.  Returns the new position.
__device__ __forceinline__ double2 reflect_wave(int nl, const double* __restrict__ bxp,
                                                const double* __restrict__ bzp,
                                                const double* __restrict__ enx,
                                                const double* __restrict__ enz, double px,
                                                double pz)
{
    const int lane = threadIdx.x & 63;
    double min_dist = 1e10;
    int closest = nl;  // none closer than 1e10 in this lane
    for (int j = lane; j < nl; j += 64) {
        const double x0 = bxp[j], z0 = bzp[j], x1 = bxp[j + 1], z1 = bzp[j + 1];
        const double vx = px - x0, vz = pz - z0;
        const double dx = x1 - x0, dz = z1 - z0;
        const double len_sq = dx * dx + dz * dz;
        double t = (vx * dx + vz * dz) / len_sq;
        const double tm = (t < 1.0) ? t : 1.0;      // std::min(1.0, t)
        t = (0.0 < tm) ? tm : 0.0;                   // std::max(0.0, .)
        const double cx = x0 + t * dx, cz = z0 + t * dz;
        const double dist = sqrt((px - cx) * (px - cx) + (pz - cz) * (pz - cz));
        if (dist < min_dist) { min_dist = dist; closest = j; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double od = __shfl_xor(min_dist, o);
        const int oj = __shfl_xor(closest, o);
        if (od < min_dist || (od == min_dist && oj < closest)) {
            min_dist = od;
            closest = oj;
        }
    }
    if (closest == nl) closest = 0;
    const double offset = 2.0 * min_dist + 1e-6;
    double nx = px + offset * enx[closest];
    double nz = pz + offset * enz[closest];
    int cn = 0;
    for (int j = lane; j < nl; j += 64) cn += edge_crossing(bxp, bzp, j, nx, nz);
    for (int o = 32; o > 0; o >>= 1) cn += __shfl_xor(cn, o);
    if (cn == 1) {
        nx = px + 3.0 * offset * enx[closest];
        nz = pz + 3.0 * offset * enz[closest];
    }
    return make_double2(nx, nz);
}

// 1/a from the hardware reciprocal and two Newton steps on the FMA: below one
// unit in the last place (the correctly rounded quotient the reference's
// divide gives can differ in that last bit).  a = 0 gives NaN, as 0/0 does.
__device__ __forceinline__ double recip64(double a)
{
    double y = __builtin_amdgcn_rcp(a);
    double e = __builtin_fma(-a, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-a, y, 1.0);
    return __builtin_fma(y, e, y);
}

// One (blob, panel) term of src/VortexSheet.cpp:225-272: coefficients
// px, qx, py, qy with the collocation midpoint as panel origin, against
// gamma_j and gamma_{j+1}.
__device__ __forceinline__ void sheet_pair(double xi, double zi, const PanelConst& q, double gj,
                                           double gj1, double& su, double& sw)
{
    const double xj = q.xc, zj = q.zc, dsj = q.ds;
    const double c1 = -(xi - xj) * q.cs - (zi - zj) * q.sn;
    const double c2 = (xi - xj) * (xi - xj) + (zi - zj) * (zi - zj);
    const double c5 = (xi - xj) * q.sn - (zi - zj) * q.cs;
    const double c6 = log(1.0 + dsj * ((dsj + 2 * c1) / c2));
    const double c7 = atan2(c5 * dsj, c2 + c1 * dsj);
    const double c8 = (xi - xj) * q.s2 + (zi - zj) * q.c2;
    const double c9 = (xi - xj) * q.c2 - (zi - zj) * q.s2;
    const double ids = 1.0 / dsj;  // one division for the four quotients by ds
    const double qx = -q.sn + 0.5 * c8 * c6 * ids + (c1 * q.cs + c5 * q.sn) * c7 * ids;
    const double px = -0.5 * c6 * q.sn - c7 * q.cs - qx;
    const double qy = q.cs + 0.5 * c9 * c6 * ids + (c1 * q.sn - c5 * q.cs) * c7 * ids;
    const double py = 0.5 * c6 * q.cs - c7 * q.sn - qy;
    su += px * gj + qx * gj1;
    sw += py * gj + qy * gj1;
}

// Terms of a group's (or the body's) Laurent series at |W|^2 = w2 > 1 (W the
// point's offset in units of the radius): with the moments scaled by the
// radius, term k is at most the leading term times |W|^-k, so the series is
// cut where |W|^-K <= 2^-43 (e = floor(log2 w2) from the exponent bits), at
// most kSheetTerms (all of them below |W| = 2.8).
__device__ __forceinline__ int sheet_terms(double w2)
{
    const int e = (int)((__double_as_longlong(w2) >> 52) & 0x7ff) - 1023;
    const int K = e > 0 ? (86 + e - 1) / e : kSheetTerms;
    return K < kSheetTerms ? K : kSheetTerms;
}

// The sheet's velocity at (x, z) by panel groups (sheet_groups): for each
// group, beyond kGroupFar * r of its centre the Laurent series
// phi = (1/r) sum_k A_k / W^(k+1), W = (z - c) / r (u += Im phi, w += Re phi),
// else sheet_pair over the group's panels.  pc, gm (gamma), gc (centres) and
// ga (moments, 2 kSheetTerms per group) are one body's; the choice depends on
// (x, z) alone.  Adds to su, sw (without the 1/2pi).
__device__ __forceinline__ void sheet_by_groups(double x, double z, int nl, const PanelConst* pc,
                                                const double* gm, const double4* gc,
                                                const double* ga, double& su, double& sw)
{
    const int ng = (nl + kSheetGroup - 1) / kSheetGroup;
    for (int g = 0; g < ng; ++g) {
        const double4 c = gc[g];  // centre, radius, 1 / radius (0: never the series)
        const double wr = (x - c.x) * c.w, wi = (z - c.y) * c.w;
        const double w2 = wr * wr + wi * wi;
        if (w2 > kGroupFar * kGroupFar) {
            const double* a = ga + 2 * g * kSheetTerms;
            const double iw2 = recip64(w2);
            const double tr = wr * iw2, ti = -wi * iw2;  // 1 / W
            const int K = sheet_terms(w2);
            double pr = a[2 * (K - 1)], pim = a[2 * (K - 1) + 1];
            for (int k = K - 2; k >= 0; --k) {
                const double nr = pr * tr - pim * ti + a[2 * k];
                pim = pr * ti + pim * tr + a[2 * k + 1];
                pr = nr;
            }
            su += (pr * ti + pim * tr) * c.w;
            sw += (pr * tr - pim * ti) * c.w;
        } else {
            const int j1 = min((g + 1) * kSheetGroup, nl);
            for (int j = g * kSheetGroup; j < j1; ++j)
                sheet_pair(x, z, pc[j], gm[j], gm[j + 1 == nl ? 0 : j + 1], su, sw);
        }
    }
}

// Group moments A_k = sum over group g's panels, in order, of
// -(gamma_j I0_jk + gamma_{j+1} I1_jk) (k = lane < kSheetTerms); i0, i1 in
// the layout of sheet_groups' integrals, one body's.
__device__ __forceinline__ void group_moments(int g, int k, int nl, const double* gm,
                                              const double* i0, const double* i1, double* ga)
{
    const int j0 = g * kSheetGroup;
    if (j0 >= nl || k >= kSheetTerms) return;
    const int j1 = min(j0 + kSheetGroup, nl);
    double ar = 0.0, ai = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double g0 = gm[j], g1 = gm[j + 1 == nl ? 0 : j + 1];
        const int64_t o = 2 * ((int64_t)j * kSheetTerms + k);
        ar -= g0 * i0[o] + g1 * i1[o];
        ai -= g0 * i0[o + 1] + g1 * i1[o + 1];
    }
    ga[2 * (g * kSheetTerms + k)] = ar;
    ga[2 * (g * kSheetTerms + k) + 1] = ai;
}


// One pair of the reference kernel (src/VelocityKernel.cpp:44-60), FP64.
__device__ __forceinline__ void bs_pair_f64(double xi, double zi, double xj, double zj, double cj,
                                            double sj, double& ui, double& wi)
{
    const double dx = xi - xj;
    const double dz = zi - zj;
    const double r2 = dx * dx + dz * dz;
    const double s2 = sj * sj;
    const double thr = 10.0 * s2;
    double num = 1.0;
    if (r2 < thr) {
        const double rs = 1.0 / s2;
        num = 1.0 - exp(-r2 * rs);
    }
    const double K = num / r2;
    ui -= K * dz * cj;
    wi += K * dx * cj;
}

// Armadillo-style two-accumulator sum (the oracle restates the same order).
__device__ __forceinline__ double pair_sum(const double* a, int n)
{
    double acc1 = 0.0, acc2 = 0.0;
    int j;
    for (j = 1; j < n; j += 2) {
        acc1 += a[j - 1];
        acc2 += a[j];
    }
    if ((j - 1) < n) acc1 += a[j - 1];
    return acc1 + acc2;
}

// src/VortexSheet.cpp:487-514 for one body, by one workgroup of kLoadsThreads
// lanes: out = {fx, fz, cd, cl}.  Every value is formed by the reference's
// expression in the reference's order — the pressure prefix, the maximum and
// the two force sums sequentially by lane 0 — with the independent
// elementwise steps spread over the lanes and the vectors staged in LDS
// (work, 2 nl doubles, global memory, when nl > kLoadsLds).
constexpr int kLoadsThreads = 256;
constexpr int kLoadsLds = 1024;

__device__ __forceinline__ void loads_block(int nl, double rho, double dt, double ur, double aref,
                                            const double* __restrict__ g,
                                            const double* __restrict__ d,
                                            const double* __restrict__ ex,
                                            const double* __restrict__ ez, double* work,
                                            double* __restrict__ out)
{
    __shared__ double la[kLoadsLds], lb[kLoadsLds];
    __shared__ double pmax_s;
    double* A = nl <= kLoadsLds ? la : work;
    double* B = nl <= kLoadsLds ? lb : work + nl;
    const int t = threadIdx.x;
    const double k = -rho / dt;
    for (int i = t; i < nl; i += blockDim.x) A[i] = k * (g[i] * d[i]);
    __syncthreads();
    if (t == 0) {  // P[0] = 0, P[i+1] = P[i] + k (gamma_i ds_i)
        double p = 0.0;
        B[0] = 0.0;
        for (int i = 0; i < nl - 1; ++i) {
            p = p + A[i];
            B[i + 1] = p;
        }
    }
    __syncthreads();
    // P[i] = (P[i] + P[i+1]) / 2 for i < nl - 1 (the old P[i+1]), then the last
    // entry with the new P[0]
    for (int i = t; i < nl - 1; i += blockDim.x) A[i] = 0.5 * (B[i] + B[i + 1]);
    __syncthreads();
    if (t == 0) {
        A[nl - 1] = 0.5 * (B[nl - 1] + A[0]);
        double pm = A[0];
        for (int i = 1; i < nl; ++i) pm = A[i] > pm ? A[i] : pm;
        pmax_s = pm;
    }
    __syncthreads();
    const double shift = (0.5 * rho * ur * ur - pmax_s) * 1.0;
    for (int i = t; i < nl; i += blockDim.x) {
        const double pi = A[i] + shift;
        B[i] = pi * ex[i] * d[i];
        A[i] = pi * ez[i] * d[i];
    }
    __syncthreads();
    if (t == 0) {
        const double fx = -pair_sum(B, nl);
        const double fz = -pair_sum(A, nl);
        out[0] = fx;
        out[1] = fz;
        out[2] = fx / (0.5 * rho * aref * ur * ur);
        out[3] = fz / (0.5 * rho * aref * ur * ur);
    }
}

// Spreads the low 16 bits of v to the even bits (Morton interleave).
__device__ inline uint32_t spread16(uint32_t v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Far-field loops over one source tile: sum of circ_j * (dz, dx) / r^2.
// The shared forms take one reciprocal for 2, 4 or 8 sources:
// 1/ra = rb * (1/(ra rb)); they are used only when every r^2 of the tile pair
// lies in a range where the product stays a normal FP32 number.

__device__ __forceinline__ void far_pair(float4 sj, float txd, float tzd, float& au, float& aw)
{
    const float dx = txd - sj.x;
    const float dz = tzd - sj.y;
    const float r2 = __builtin_fmaf(dx, dx, dz * dz);
    const float kk = __builtin_amdgcn_rcpf(r2) * sj.z;
    au = __builtin_fmaf(kk, dz, au);
    aw = __builtin_fmaf(kk, dx, aw);
}

template <int T>
__device__ __forceinline__ void far_tile(const float4* __restrict__ sp, float txd, float tzd,
                                         float& au, float& aw)
{
#pragma unroll 16
    for (int k = 0; k < T; ++k) far_pair(sp[k], txd, tzd, au, aw);
}

template <int T>
__device__ __forceinline__ void far_tile_share2(const float4* __restrict__ sp, float txd,
                                                float tzd, float& au, float& aw)
{
#pragma unroll 8
    for (int k = 0; k < T; k += 2) {
        const float4 sa = sp[k];
        const float4 sb = sp[k + 1];
        const float dxa = txd - sa.x, dza = tzd - sa.y;
        const float dxb = txd - sb.x, dzb = tzd - sb.y;
        const float ra = __builtin_fmaf(dxa, dxa, dza * dza);
        const float rb = __builtin_fmaf(dxb, dxb, dzb * dzb);
        const float inv = __builtin_amdgcn_rcpf(ra * rb);
        const float ka = (rb * inv) * sa.z;
        const float kb = (ra * inv) * sb.z;
        au = __builtin_fmaf(ka, dza, au);
        aw = __builtin_fmaf(ka, dxa, aw);
        au = __builtin_fmaf(kb, dzb, au);
        aw = __builtin_fmaf(kb, dxb, aw);
    }
}

// s_waitcnt immediate that waits for every outstanding scalar (and LDS) load
// and for nothing else: lgkmcnt = 0, vmcnt = 63 and expcnt = 7 (no wait).
constexpr unsigned kWaitLgkm0 = 0xc07f;

// Eight consecutive source records: one 128-byte scalar load.
struct Src8 {
    float4 v[8];
};

__device__ __forceinline__ void far_quad(const float4 sa, const float4 sb, const float4 sc,
                                         const float4 sd, float txd, float tzd, float& au,
                                         float& aw)
{
    const float dxa = txd - sa.x, dza = tzd - sa.y;
    const float dxb = txd - sb.x, dzb = tzd - sb.y;
    const float dxc = txd - sc.x, dzc = tzd - sc.y;
    const float dxd = txd - sd.x, dzd = tzd - sd.y;
    const float ra = __builtin_fmaf(dxa, dxa, dza * dza);
    const float rb = __builtin_fmaf(dxb, dxb, dzb * dzb);
    const float rc = __builtin_fmaf(dxc, dxc, dzc * dzc);
    const float rd = __builtin_fmaf(dxd, dxd, dzd * dzd);
    const float pab = ra * rb;
    const float pcd = rc * rd;
    const float inv = __builtin_amdgcn_rcpf(pab * pcd);
    const float iab = inv * pcd;
    const float icd = inv * pab;
    const float ka = (rb * iab) * sa.z;
    const float kb = (ra * iab) * sb.z;
    const float kc = (rd * icd) * sc.z;
    const float kd = (rc * icd) * sd.z;
    au = __builtin_fmaf(ka, dza, au);
    aw = __builtin_fmaf(ka, dxa, aw);
    au = __builtin_fmaf(kb, dzb, au);
    aw = __builtin_fmaf(kb, dxb, aw);
    au = __builtin_fmaf(kc, dzc, au);
    aw = __builtin_fmaf(kc, dxc, aw);
    au = __builtin_fmaf(kd, dzd, au);
    aw = __builtin_fmaf(kd, dxd, aw);
}

// Eight far pairs with one reciprocal: the product tree of the eight r^2
// (pairs, quads, all), its reciprocal, and each 1/r^2 as the product of the
// other seven taken back down the tree.  The caller keeps every r^2 within
// [1e-4, 1e4], so every partial product stays within [1e-32, 1e32].
__device__ __forceinline__ void far_oct(const Src8& s, float txd, float tzd, float& au, float& aw)
{
    float dx[8], dz[8], r[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        dx[q] = txd - s.v[q].x;
        dz[q] = tzd - s.v[q].y;
        r[q] = __builtin_fmaf(dx[q], dx[q], dz[q] * dz[q]);
    }
    const float p01 = r[0] * r[1], p23 = r[2] * r[3], p45 = r[4] * r[5], p67 = r[6] * r[7];
    const float p03 = p01 * p23, p47 = p45 * p67;
    const float inv = __builtin_amdgcn_rcpf(p03 * p47);
    const float i03 = inv * p47, i47 = inv * p03;
    const float i01 = i03 * p23, i23 = i03 * p01, i45 = i47 * p67, i67 = i47 * p45;
    const float k[8] = {(r[1] * i01) * s.v[0].z, (r[0] * i01) * s.v[1].z,
                        (r[3] * i23) * s.v[2].z, (r[2] * i23) * s.v[3].z,
                        (r[5] * i45) * s.v[4].z, (r[4] * i45) * s.v[5].z,
                        (r[7] * i67) * s.v[6].z, (r[6] * i67) * s.v[7].z};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        au = __builtin_fmaf(k[q], dz[q], au);
        aw = __builtin_fmaf(k[q], dx[q], aw);
    }
}

// One reciprocal for four sources.  The records go through scalar loads in
// batches of eight with the next batch in flight while the current one is
// summed: an explicit wait for all scalar loads (they complete out of order,
// so no partial wait exists) comes before each batch's loads are issued, and
// scheduling barriers keep the compiler from moving that wait below the
// loads or the arithmetic above them.  The batch index goes through
// readfirstlane so the address stays a base plus immediate offsets.
template <int T>
__device__ __forceinline__ void far_tile_share4(const float4* __restrict__ sp, float txd,
                                                float tzd, float& au, float& aw)
{
    const Src8* __restrict__ q = reinterpret_cast<const Src8*>(sp);
    Src8 a = q[0];
#pragma unroll 1
    for (int b = 0; b < T / 8; b += 2) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        const Src8 c = q[__builtin_amdgcn_readfirstlane(b + 1)];
        __builtin_amdgcn_sched_barrier(0);
        far_quad(a.v[0], a.v[1], a.v[2], a.v[3], txd, tzd, au, aw);
        __builtin_amdgcn_sched_barrier(0);
        far_quad(a.v[4], a.v[5], a.v[6], a.v[7], txd, tzd, au, aw);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        a = q[__builtin_amdgcn_readfirstlane(b + 2 < T / 8 ? b + 2 : b + 1)];
        __builtin_amdgcn_sched_barrier(0);
        far_quad(c.v[0], c.v[1], c.v[2], c.v[3], txd, tzd, au, aw);
        __builtin_amdgcn_sched_barrier(0);
        far_quad(c.v[4], c.v[5], c.v[6], c.v[7], txd, tzd, au, aw);
    }
}

// far_tile_share4 with one reciprocal for each batch of eight (far_oct).
template <int T>
__device__ __forceinline__ void far_tile_share8(const float4* __restrict__ sp, float txd,
                                                float tzd, float& au, float& aw)
{
    const Src8* __restrict__ q = reinterpret_cast<const Src8*>(sp);
    Src8 a = q[0];
#pragma unroll 1
    for (int b = 0; b < T / 8; b += 2) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        const Src8 c = q[__builtin_amdgcn_readfirstlane(b + 1)];
        __builtin_amdgcn_sched_barrier(0);
        far_oct(a, txd, tzd, au, aw);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        a = q[__builtin_amdgcn_readfirstlane(b + 2 < T / 8 ? b + 2 : b + 1)];
        __builtin_amdgcn_sched_barrier(0);
        far_oct(c, txd, tzd, au, aw);
    }
}


// Kernel factor of a near-tile pair (src/VelocityKernel.cpp:48-55):
// K = (1 - exp(-r2/s^2)) / r2 when r2 < 10 s^2, else 1/r2, with
// nj = {10 s^2, 1/s^2}.  With x = r2/s^2, (1 - e^-x)/r2 = f(x)/s^2 where
// f(x) = (1 - e^-x)/x is summed from its Taylor series for x < 0.5
// (truncation below 1.1e-8 relative); above, 1 - e^-x uses the hardware
// exponential (v_exp_f32 on x log2 e; relative error below 1e-6).
// Coincident distinct points (r2 = 0) give NaN, as the reference does.
__device__ __forceinline__ float near_k_inv(float r2, float2 nj, float inv)
{
    const float y = -r2 * nj.y;
    float f = __builtin_fmaf(y, 1.0f / 40320.0f, 1.0f / 5040.0f);
    f = __builtin_fmaf(y, f, 1.0f / 720.0f);
    f = __builtin_fmaf(y, f, 1.0f / 120.0f);
    f = __builtin_fmaf(y, f, 1.0f / 24.0f);
    f = __builtin_fmaf(y, f, 1.0f / 6.0f);
    f = __builtin_fmaf(y, f, 0.5f);
    f = __builtin_fmaf(y, f, 1.0f);
    const float small = f * nj.y;
    const float mid = (1.0f - __builtin_amdgcn_exp2f(y * 1.44269504088896341f)) * inv;
    return r2 < nj.x ? ((y > -0.5f && r2 > 0.0f) ? small : mid) : inv;
}

__device__ __forceinline__ float near_k(float r2, float2 nj)
{
    return near_k_inv(r2, nj, __builtin_amdgcn_rcpf(r2));
}

// Near-tile pairs k0 <= k < k1 of one source tile.  Offsets from the tile
// centre are carried as two FP32 words (rounded value + rounding residual:
// sj.xy + nj.zw for the source, txd + txl for the target), so dx is exact to
// FP32 rounding of dx itself however far the pair sits from the centre.
// nj = {10 sigma^2, 1/sigma^2, x residual, z residual}.  SELF skips source
// j0 + k == i, GUARD skips r2 <= r2_skip.
struct NearPair {
    float dx, dz, r2, c;
    float2 nj;
};

__device__ __forceinline__ NearPair near_prep(float4 sj, float4 nj4, float txd, float tzd,
                                              float txl, float tzl)
{
    NearPair p;
    p.dx = (txd - sj.x) + (txl - nj4.z);
    p.dz = (tzd - sj.y) + (tzl - nj4.w);
    p.r2 = __builtin_fmaf(p.dx, p.dx, p.dz * p.dz);
    p.c = sj.z;
    p.nj = make_float2(nj4.x, nj4.y);
    return p;
}

template <bool SELF, bool GUARD>
__device__ __forceinline__ void near_add(const NearPair& p, float K, int64_t j, int64_t i,
                                         float r2_skip, float& au, float& aw)
{
    bool take = true;
    if (SELF) take = j != i;
    if (GUARD) take = take && (p.r2 > r2_skip);
    const float kk = take ? K * p.c : 0.0f;
    au = __builtin_fmaf(kk, p.dz, au);
    aw = __builtin_fmaf(kk, p.dx, aw);
}

// Four sources at a time: their loads and separations come first, then one
// branch for the group — the regularised factor (near_k) only when one of
// the four pairs of some lane lies inside 10 sigma^2, else 1/r^2 (which
// near_k also returns there) — so the group's scalar loads are in flight
// together instead of one source's at a time behind each pair's branch.
template <bool SELF, bool GUARD>
__device__ __forceinline__ void near_range(const float4* __restrict__ sp,
                                           const float4* __restrict__ np, int k0, int k1,
                                           float txd, float tzd, float txl, float tzl,
                                           int64_t j0, int64_t i, float r2_skip, float& au,
                                           float& aw)
{
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
        const NearPair p0 = near_prep(sp[k], np[k], txd, tzd, txl, tzl);
        const NearPair p1 = near_prep(sp[k + 1], np[k + 1], txd, tzd, txl, tzl);
        const NearPair p2 = near_prep(sp[k + 2], np[k + 2], txd, tzd, txl, tzl);
        const NearPair p3 = near_prep(sp[k + 3], np[k + 3], txd, tzd, txl, tzl);
        float K0 = __builtin_amdgcn_rcpf(p0.r2), K1 = __builtin_amdgcn_rcpf(p1.r2);
        float K2 = __builtin_amdgcn_rcpf(p2.r2), K3 = __builtin_amdgcn_rcpf(p3.r2);
        if ((p0.r2 < p0.nj.x) | (p1.r2 < p1.nj.x) | (p2.r2 < p2.nj.x) | (p3.r2 < p3.nj.x)) {
            K0 = near_k_inv(p0.r2, p0.nj, K0);
            K1 = near_k_inv(p1.r2, p1.nj, K1);
            K2 = near_k_inv(p2.r2, p2.nj, K2);
            K3 = near_k_inv(p3.r2, p3.nj, K3);
        }
        near_add<SELF, GUARD>(p0, K0, j0 + k, i, r2_skip, au, aw);
        near_add<SELF, GUARD>(p1, K1, j0 + k + 1, i, r2_skip, au, aw);
        near_add<SELF, GUARD>(p2, K2, j0 + k + 2, i, r2_skip, au, aw);
        near_add<SELF, GUARD>(p3, K3, j0 + k + 3, i, r2_skip, au, aw);
    }
    for (; k < k1; ++k) {
        const NearPair p = near_prep(sp[k], np[k], txd, tzd, txl, tzl);
        near_add<SELF, GUARD>(p, near_k(p.r2, p.nj), j0 + k, i, r2_skip, au, aw);
    }
}

// Enlarged near-field threshold of a source, kept in the fourth word of its
// FP32 record.  thr = 10 sigma^2 as the near-field constants round it, h the
// larger half extent of the source tile, thr_max the tile's largest thr.  A
// one-word separation (txd - x', tzd - z') differs from the exact one by at
// most 2^-24 (|t - c| + |s - c| + |t - s|) per component; for a pair with
// |t - s| <= 2 sqrt(thr) that is at most 2^-24 (2h + 4 sqrt(thr_max)), and
// the returned value exceeds thr by more than the r^2 such an error (and the
// roundings of both r^2 forms) can move.  So a pair whose one-word r^2 is not
// below it has an exact two-word r^2 not below thr: it is a 1/r^2 pair of the
// reference's kernel (src/VelocityKernel.cpp:48-55).
__device__ __forceinline__ float near_thr_hi(float thr, double h, double thr_max)
{
    const double e = 2.0 * h + 4.0 * sqrt(thr_max);
    const double t = (double)thr;
    return (float)(t * (1.0 + 0x1p-19) + 1e-6 * sqrt(t) * e + 3e-14 * e * e + 1e-37);
}

// One quad of a near tile (sources at np[0..3] of the near-field constants).
// The one-word separations and r^2 come first; if no lane has a pair below
// its source's enlarged threshold (record word w) the quad is four 1/r^2
// pairs, summed as in far_quad (one reciprocal for four when `share`, i.e.
// every r^2 of the tile pair lies in [1e-8, 1e8]); else the whole wave takes
// the exact two-word path of near_range, whose operations it repeats (same
// results; it gives 1/r^2 for the pairs at or beyond 10 sigma^2 too).
__device__ __forceinline__ void near_quad(const float4 sa, const float4 sb, const float4 sc,
                                          const float4 sd, const float4* __restrict__ np,
                                          float txd, float tzd, float txl, float tzl, bool share,
                                          float& au, float& aw)
{
    const float dxa = txd - sa.x, dza = tzd - sa.y;
    const float dxb = txd - sb.x, dzb = tzd - sb.y;
    const float dxc = txd - sc.x, dzc = tzd - sc.y;
    const float dxd = txd - sd.x, dzd = tzd - sd.y;
    const float ra = __builtin_fmaf(dxa, dxa, dza * dza);
    const float rb = __builtin_fmaf(dxb, dxb, dzb * dzb);
    const float rc = __builtin_fmaf(dxc, dxc, dzc * dzc);
    const float rd = __builtin_fmaf(dxd, dxd, dzd * dzd);
    // wave-uniform: compares into lane masks, ORed on the scalar unit
    const uint64_t nearm = __builtin_amdgcn_ballot_w64(ra < sa.w) |
                           __builtin_amdgcn_ballot_w64(rb < sb.w) |
                           __builtin_amdgcn_ballot_w64(rc < sc.w) |
                           __builtin_amdgcn_ballot_w64(rd < sd.w);
    if (nearm != 0) {
        const float4 na = np[0], nb = np[1], nc = np[2], nd = np[3];
        NearPair p[4];
        const float dx[4] = {dxa, dxb, dxc, dxd}, dz[4] = {dza, dzb, dzc, dzd};
        const float4 n4[4] = {na, nb, nc, nd};
        const float cc[4] = {sa.z, sb.z, sc.z, sd.z};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            p[q].dx = dx[q] + (txl - n4[q].z);
            p[q].dz = dz[q] + (tzl - n4[q].w);
            p[q].r2 = __builtin_fmaf(p[q].dx, p[q].dx, p[q].dz * p[q].dz);
            p[q].c = cc[q];
            p[q].nj = make_float2(n4[q].x, n4[q].y);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            near_add<false, false>(p[q], near_k(p[q].r2, p[q].nj), 0, 0, 0.0f, au, aw);
    } else if (share) {
        const float pab = ra * rb;
        const float pcd = rc * rd;
        const float inv = __builtin_amdgcn_rcpf(pab * pcd);
        const float iab = inv * pcd;
        const float icd = inv * pab;
        const float ka = (rb * iab) * sa.z;
        const float kb = (ra * iab) * sb.z;
        const float kc = (rd * icd) * sc.z;
        const float kd = (rc * icd) * sd.z;
        au = __builtin_fmaf(ka, dza, au);
        aw = __builtin_fmaf(ka, dxa, aw);
        au = __builtin_fmaf(kb, dzb, au);
        aw = __builtin_fmaf(kb, dxb, aw);
        au = __builtin_fmaf(kc, dzc, au);
        aw = __builtin_fmaf(kc, dxc, aw);
        au = __builtin_fmaf(kd, dzd, au);
        aw = __builtin_fmaf(kd, dxd, aw);
    } else {
        const float ka = __builtin_amdgcn_rcpf(ra) * sa.z;
        const float kb = __builtin_amdgcn_rcpf(rb) * sb.z;
        const float kc = __builtin_amdgcn_rcpf(rc) * sc.z;
        const float kd = __builtin_amdgcn_rcpf(rd) * sd.z;
        au = __builtin_fmaf(ka, dza, au);
        aw = __builtin_fmaf(ka, dxa, aw);
        au = __builtin_fmaf(kb, dzb, au);
        aw = __builtin_fmaf(kb, dxb, aw);
        au = __builtin_fmaf(kc, dzc, au);
        aw = __builtin_fmaf(kc, dxc, aw);
        au = __builtin_fmaf(kd, dzd, au);
        aw = __builtin_fmaf(kd, dxd, aw);
    }
}

// A full near tile without the self pair or a guard: the source records
// stream as in far_tile_share4 (batches of eight, the next in flight) and
// each quad is classified by near_quad; only quads with a near pair read the
// near-field constants.
template <int T>
__device__ __forceinline__ void near_tile_share4(const float4* __restrict__ sp,
                                                 const float4* __restrict__ np, float txd,
                                                 float tzd, float txl, float tzl, bool share,
                                                 float& au, float& aw)
{
    const Src8* __restrict__ q = reinterpret_cast<const Src8*>(sp);
    Src8 a = q[0];
#pragma unroll 1
    for (int b = 0; b < T / 8; b += 2) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        const int bc = __builtin_amdgcn_readfirstlane(b + 1);
        const Src8 c = q[bc];
        __builtin_amdgcn_sched_barrier(0);
        const float4* __restrict__ na = np + 8 * __builtin_amdgcn_readfirstlane(b);
        near_quad(a.v[0], a.v[1], a.v[2], a.v[3], na, txd, tzd, txl, tzl, share, au, aw);
        near_quad(a.v[4], a.v[5], a.v[6], a.v[7], na + 4, txd, tzd, txl, tzl, share, au, aw);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        a = q[__builtin_amdgcn_readfirstlane(b + 2 < T / 8 ? b + 2 : b + 1)];
        __builtin_amdgcn_sched_barrier(0);
        const float4* __restrict__ nc = np + 8 * bc;
        near_quad(c.v[0], c.v[1], c.v[2], c.v[3], nc, txd, tzd, txl, tzl, share, au, aw);
        near_quad(c.v[4], c.v[5], c.v[6], c.v[7], nc + 4, txd, tzd, txl, tzl, share, au, aw);
    }
}

}  // namespace dvmi
