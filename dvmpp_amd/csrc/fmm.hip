// Fast multipole form of the blob-on-blob sum: the reference's FMMKernel
// (src/VelocityKernel.hpp:84-110, parameters expansion_order and theta),
// whose body there is a stub that calls the direct sum
// (src/VelocityKernel.cpp:113-147).  Here it is a 2-D FMM on the device.
//
// Far field.  Beyond r^2 >= 10 sigma_j^2 the reference kernel is exactly the
// point vortex 1/r^2 (src/VelocityKernel.cpp:47-55), so with
// phi(z) = sum_j G_j / (z - z_j) over the far sources (z = x + i z),
// u = Im(phi) / 2pi and w = Re(phi) / 2pi.  phi is carried by complex
// multipole and local expansions of `p` terms on a uniform quadtree over the
// square bounding box, scaled by each level's box width so every coefficient
// is O(1):
//   P2M  A_k  = sum_j G_j Z_j^k,                      Z_j = (z_j - c) / s
//   M2M  A'_k = sum_{m<=k} C(k,m) 2^-m A_m D^(k-m),   D = (c_child - c) / s
//   M2L  B_l  = (-1)^l sum_k C(k+l,l) A_k T^-(k+l+1), T = (c_tgt - c_src) / s
//   L2L  B'_m = 2^-(m+1) sum_{l>=m} C(l,m) B_l E^(l-m), E = (c_child - c) / s
//   L2P  phi  = (1/s) sum_l B_l W^l,                   W = (z - c) / s
// Two boxes of one level are well separated when their centres are farther
// apart than sqrt(2) s / theta (the half-diagonals sum to theta times the
// distance: theta is the reference's opening criterion); the interaction
// list of a box is the children of its parent's near boxes that are not
// near it.  The truncation error falls roughly as theta^p.
//
// Near field.  Sources in the near leaves are summed directly with the
// reference's regularised kernel.  The leaf width is kept large enough that
// every pair closer than sqrt(10) * max(sigma) lies in near leaves, so the
// only difference from the direct sum is the truncation of the far field.
//
// Layout: particles Morton-sorted by leaf (radix sort); boxes of level l are
// indexed by their Morton code after an offset (4^l - 1) / 3; expansions are
// complex FP64, `P` per box.  Kernels: key, sort, gather, leaf bounds, P2M
// (one thread per leaf), M2M per level, M2L + L2L per level (a wavefront per
// occupied box, its lane groups splitting the interaction list), and L2P + P2P
// (one thread per particle; a wavefront per particle with a dense near field).
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "dvm_internal.hpp"

namespace dvmi {

namespace {

constexpr int kFmmMaxLevel = 10;
constexpr int kBinom = 64;  // binomial table side (2P <= 64)
// Mean blobs per leaf over the square box that sets the depth: wakes fill a
// fraction of the box, so occupied leaves hold several times more.
constexpr double kLeafMean = 16.0;

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx cmul_fma(cplx a, cplx b)
{
    return {__builtin_fma(a.re, b.re, -(a.im * b.im)), __builtin_fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cplx cscale(cplx a, double f) { return {a.re * f, a.im * f}; }

__host__ __device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__host__ __device__ __forceinline__ uint32_t morton(uint32_t ix, uint32_t iz)
{
    return spread_bits(ix) | (spread_bits(iz) << 1);
}
__device__ __forceinline__ uint32_t compact_bits(uint32_t v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

struct FmmTree {
    double x0, z0, S;  // square domain origin and side
    int L;             // leaf level
    int p;             // terms used (<= P)
    const double* binom;  // C(n, k) at [n * kBinom + k]
    const int* il;        // interaction-list offsets (dx, dz) per child position, il_n each
    int il_n[4];
    int il_stride;
    // per (child position, entry, lane l): T^-(l+1) and (-1)^l T^-l, with T the
    // entry's box offset from source to target (box widths; il_cap entries a
    // position, G lanes an entry, four doubles a lane)
    const double* m2l;
    int il_cap;
    const int* nearo;     // near offsets (dx, dz)
    int near_n;
    const int* lbeg;      // leaf particle bounds (4^L + 1)
    int* cnt;             // particles per box (all levels)
    int* need;            // 1: a target lies in the box (its locals are needed)
    cplx* A;              // multipole coefficients per box
    cplx* B;              // local coefficients per box
    int* heavy;           // [0] count, then the heavy targets (fmm_eval_kernel)
};

// Targets whose near field holds more sources than this are summed by a
// wavefront each (fmm_heavy_kernel) instead of one lane: dense clusters (the
// nascent blobs along a body) would otherwise leave a few lanes running long
// after the rest of the launch.
constexpr int kFmmHeavy = 2048;

__host__ __device__ __forceinline__ int64_t level_off(int l) { return ((int64_t(1) << (2 * l)) - 1) / 3; }

__global__ void fmm_binom_kernel(double* binom)
{
    // one thread per row n: C(n, k) for k <= n (Pascal, exact up to 2^53)
    const int n = threadIdx.x;
    if (n >= kBinom) return;
    double v = 1.0;
    for (int k = 0; k < kBinom; ++k) {
        binom[n * kBinom + k] = k <= n ? v : 0.0;
        v = v * (double)(n - k) / (double)(k + 1);
    }
}

__global__ void fmm_key_kernel(int n, const double* __restrict__ x, const double* __restrict__ z,
                               double x0, double z0, double sl, int nb, uint32_t* __restrict__ key,
                               int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = min(max((int)floor((x[i] - x0) / sl), 0), nb - 1);
    const int iz = min(max((int)floor((z[i] - z0) / sl), 0), nb - 1);
    key[i] = morton((uint32_t)ix, (uint32_t)iz);
    idx[i] = i;
}

// Occupancy of the leaves at level lm: one bit per leaf, in Morton order.
__global__ void fmm_occ_kernel(int n, const double* __restrict__ x, const double* __restrict__ z,
                               double x0, double z0, double sl, int nb,
                               unsigned long long* __restrict__ bm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = min(max((int)floor((x[i] - x0) / sl), 0), nb - 1);
    const int iz = min(max((int)floor((z[i] - z0) / sl), 0), nb - 1);
    const uint32_t code = morton((uint32_t)ix, (uint32_t)iz);
    atomicOr(&bm[code >> 6], 1ull << (code & 63));
}

// Occupied boxes of level lm - blockIdx.y from the level-lm bitmap: a box of
// level l covers 4^(lm - l) consecutive leaf codes.
__global__ void fmm_occ_count_kernel(const unsigned long long* __restrict__ bm, int lm,
                                     int* __restrict__ cnt)
{
    const int y = blockIdx.y;
    const int64_t ngroups = int64_t(1) << (2 * (lm - y));
    const int64_t gbits = int64_t(1) << (2 * y);
    int local = 0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups;
         g += (int64_t)gridDim.x * blockDim.x) {
        bool any = false;
        if (gbits <= 64) {
            const int64_t b0 = g * gbits;
            const unsigned long long m =
                gbits == 64 ? ~0ull : ((1ull << gbits) - 1) << (b0 & 63);
            any = (bm[b0 >> 6] & m) != 0;
        } else {
            for (int64_t w = g * gbits / 64; w < (g + 1) * gbits / 64 && !any; ++w) any = bm[w] != 0;
        }
        local += any ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(&cnt[y], local);
}

// Sorted particles as (x, z, circulation, 1 / sigma^2).
__global__ void fmm_gather_kernel(int n, const int* __restrict__ perm, const double* __restrict__ x,
                                  const double* __restrict__ z, const double* __restrict__ c,
                                  const double* __restrict__ s, double4* __restrict__ sp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = perm[k];
    sp[k] = make_double4(x[i], z[i], c[i], 1.0 / (s[i] * s[i]));
}

// lbeg[m] = first sorted position with key >= m, one thread per leaf (a
// binary search, so long runs of empty leaves cost nothing extra).
__global__ void fmm_lbeg_kernel(int n, const uint32_t* __restrict__ key, int nleaf, int* __restrict__ lbeg)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nleaf) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)key[mid] < m) lo = mid + 1;
        else hi = mid;
    }
    lbeg[m] = lo;
}

// e^-x for 0 <= x < 10, the range of the regularised kernel's factor: with
// x = k ln2 + r (k = rint(x / ln2), ln2 split in two parts as in Cody and
// Waite's reduction, so k ln2 is exact to 2^-106), e^-r = sum_{i<=12} (-r)^i / i! for |r| <= ln2 / 2 (truncation
// below 1.7e-16 relative), then scaled by 2^-k.  The library exp also handles
// overflow, underflow and non-finite input, which this range never reaches.
__device__ __forceinline__ double exp_neg_kernel(double x)
{
    constexpr double kLog2e = 1.4426950408889634;      // 1 / ln2
    constexpr double kLn2Hi = 6.93147180369123816490e-01;  // ln2, high part (exact k ln2_hi)
    constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln2 - ln2_hi
    const double kf = __builtin_rint(x * kLog2e);
    const double r = fma(-kf, kLn2Lo, fma(-kf, kLn2Hi, x));  // x - k ln2
    const double q = -r;
    double p = 1.0 / 479001600.0;  // 1/12!
    p = fma(p, q, 1.0 / 39916800.0);
    p = fma(p, q, 1.0 / 3628800.0);
    p = fma(p, q, 1.0 / 362880.0);
    p = fma(p, q, 1.0 / 40320.0);
    p = fma(p, q, 1.0 / 5040.0);
    p = fma(p, q, 1.0 / 720.0);
    p = fma(p, q, 1.0 / 120.0);
    p = fma(p, q, 1.0 / 24.0);
    p = fma(p, q, 1.0 / 6.0);
    p = fma(p, q, 0.5);
    p = fma(p, q, 1.0);
    p = fma(p, q, 1.0);
    return ldexp(p, -(int)kf);
}

// One near pair with the reference's regularised kernel
// (src/VelocityKernel.cpp:44-60): 1/r^2 from the hardware reciprocal and two
// Newton steps (about 1 ulp), the factor 1 - exp(-r^2/sigma^2) inside
// r^2 < 10 sigma^2.  q.w holds 1/sigma^2.
template <bool MASK = false>
__device__ __forceinline__ void near_pair64(double xi, double zi, const double4& q, double& nu,
                                            double& nw, bool self = false)
{
    const double dx = xi - q.x;
    const double dz = zi - q.y;
    // a masked (self) pair takes r^2 = 1 and weight 0: it adds -0 * 0, which
    // leaves both sums as they are
    const double r2 = MASK && self ? 1.0 : fma(dx, dx, dz * dz);
    const double xs = r2 * q.w;
    double inv = __builtin_amdgcn_rcp(r2);
    inv = fma(inv, fma(-r2, inv, 1.0), inv);
    inv = fma(inv, fma(-r2, inv, 1.0), inv);
    const double ci = MASK && self ? 0.0 : q.z * inv;  // before the branch: the record loads whole
    const double g = xs < 10.0 ? (1.0 - exp_neg_kernel(xs)) * ci : ci;
    nu = fma(-g, dz, nu);
    nw = fma(g, dx, nw);
}

__device__ __forceinline__ cplx box_centre(const FmmTree& t, int l, uint32_t m)
{
    const double s = t.S / (double)(1 << l);
    return {t.x0 + (compact_bits(m) + 0.5) * s, t.z0 + (compact_bits(m >> 1) + 0.5) * s};
}

// P2M, one thread per leaf; a leaf of more than kP2mHeavy particles is
// listed in t.heavy instead and done by fmm_p2m_heavy_kernel (a dense leaf
// would otherwise hold one lane long after the rest of the launch).
constexpr int kP2mHeavy = 256;

template <int P>
__device__ __forceinline__ void p2m_accumulate(const FmmTree& t, const double4* __restrict__ sp,
                                               int64_t m, int j0, int j1, int step, cplx* a)
{
    const double s = t.S / (double)(1 << t.L), rs = 1.0 / s;
    const cplx c = box_centre(t, t.L, (uint32_t)m);
    for (int j = j0; j < j1; j += step) {
        const double4 q = sp[j];
        const cplx zj = {(q.x - c.re) * rs, (q.y - c.im) * rs};
        cplx pw = {q.z, 0.0};
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (k < t.p) a[k] = cadd(a[k], pw);
            pw = cmul(pw, zj);
        }
    }
}

template <int P>
__global__ __launch_bounds__(64) void fmm_p2m_kernel(FmmTree t, const double4* __restrict__ sp)
{
    const int64_t nleaf = int64_t(1) << (2 * t.L);
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = m < nleaf;
    const int b = valid ? t.lbeg[m] : 0, e = valid ? t.lbeg[m + 1] : 0;
    const bool heavy = e - b > kP2mHeavy;
    const uint64_t hm = __ballot(heavy);
    if (hm) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)hm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(t.heavy, __popcll(hm));
        base = __shfl(base, leader);
        if (heavy) t.heavy[1 + base + __popcll(hm & ((uint64_t(1) << lane) - 1))] = (int)m;
    }
    if (!valid || heavy) return;
    const int64_t bi = level_off(t.L) + m;
    t.cnt[bi] = e - b;
    cplx a[P];
#pragma unroll
    for (int k = 0; k < P; ++k) a[k] = {0.0, 0.0};
    p2m_accumulate<P>(t, sp, m, b, e, 1, a);
#pragma unroll
    for (int k = 0; k < P; ++k) t.A[bi * P + k] = a[k];
}

// The listed dense leaves, one wavefront each: lane l takes particles l,
// l+64, ..., and the lanes' sums are added by shuffles.
template <int P>
__global__ __launch_bounds__(256) void fmm_p2m_heavy_kernel(FmmTree t, const double4* __restrict__ sp)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    const int cnt = t.heavy[0];
    for (int q = wave; q < cnt; q += nwaves) {
        const int64_t m = t.heavy[1 + q];
        const int64_t bi = level_off(t.L) + m;
        const int b = t.lbeg[m], e = t.lbeg[m + 1];
        cplx a[P];
#pragma unroll
        for (int k = 0; k < P; ++k) a[k] = {0.0, 0.0};
        p2m_accumulate<P>(t, sp, m, b + lane, e, 64, a);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (k >= t.p) continue;
            for (int o = 32; o > 0; o >>= 1) {
                a[k].re += __shfl_xor(a[k].re, o);
                a[k].im += __shfl_xor(a[k].im, o);
            }
        }
        if (lane == 0) {
            t.cnt[bi] = e - b;
#pragma unroll
            for (int k = 0; k < P; ++k) t.A[bi * P + k] = a[k];
        }
    }
}

// M2M into level l from its children at l + 1: one thread per box.
template <int P>
__global__ __launch_bounds__(64) void fmm_m2m_kernel(FmmTree t, int l)
{
    const int64_t nbox = int64_t(1) << (2 * l);
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nbox) return;
    const int64_t bi = level_off(l) + m;
    cplx a[P];
#pragma unroll
    for (int k = 0; k < P; ++k) a[k] = {0.0, 0.0};
    int total = 0;
    t.need[bi] = t.need[level_off(l + 1) + 4 * m] | t.need[level_off(l + 1) + 4 * m + 1] |
                 t.need[level_off(l + 1) + 4 * m + 2] | t.need[level_off(l + 1) + 4 * m + 3];
    for (int ch = 0; ch < 4; ++ch) {
        const int64_t ci = level_off(l + 1) + 4 * m + ch;
        const int cn = t.cnt[ci];
        if (cn == 0) continue;
        total += cn;
        // D = (c_child - c) / s = (+-1/4, +-1/4)
        const cplx d = {(ch & 1) ? 0.25 : -0.25, (ch & 2) ? 0.25 : -0.25};
        cplx h[P];
        double f = 1.0;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            h[k] = cscale(t.A[ci * P + k], f);
            f *= 0.5;
        }
        // a_k += sum_m C(k,m) h_m d^(k-m)
        cplx dp[P];
        dp[0] = {1.0, 0.0};
#pragma unroll
        for (int k = 1; k < P; ++k) dp[k] = cmul(dp[k - 1], d);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (k < t.p) {
                cplx acc = {0.0, 0.0};
#pragma unroll
                for (int mm = 0; mm <= k; ++mm)
                    acc = cadd(acc, cscale(cmul(h[mm], dp[k - mm]), t.binom[k * kBinom + mm]));
                a[k] = cadd(a[k], acc);
            }
        }
    }
    t.cnt[bi] = total;
#pragma unroll
    for (int k = 0; k < P; ++k) t.A[bi * P + k] = a[k];
}

// Group width for P terms: lane l of a group holds coefficient l.
template <int P>
struct Lanes {
    static constexpr int G = P <= 8 ? 8 : P <= 16 ? 16 : 32;
};

__device__ __forceinline__ cplx shfl_c(cplx v, int src, int width)
{
    return {__shfl(v.re, src, width), __shfl(v.im, src, width)};
}

// Lane moves within rows of 16 lanes by DPP (VALU moves, where __shfl takes
// an LDS permute each): CTRL 0x150 + k broadcasts lane k of each row
// (row_newbcast), 0x110 + o shifts up by o (row_shr; the first o lanes of a
// row keep their own value).  A double moves as its two halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xf, 0xf,
                                               false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xf, 0xf,
                                               false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ cplx dpp_c(cplx v)
{
    return {dpp_f64<CTRL>(v.re), dpp_f64<CTRL>(v.im)};
}

// acc += sum_{k < P} C(k + lane, lane) * (lane k's value), lanes in rows of 16.
template <int K, int P>
struct RowSum16 {
    static __device__ __forceinline__ void run(cplx ak, const double* cb, cplx& acc)
    {
        const cplx v = dpp_c<0x150 + K>(ak);
        acc = {__builtin_fma(v.re, cb[K], acc.re), __builtin_fma(v.im, cb[K], acc.im)};
        RowSum16<K + 1, P>::run(ak, cb, acc);
    }
};
template <int P>
struct RowSum16<P, P> {
    static __device__ __forceinline__ void run(cplx, const double*, cplx&) {}
};

// Level l locals: L2L from the parent (l > 2) plus M2L from the interaction
// list, one wavefront per box with particles, made of 64/G groups of G lanes;
// lane l of a group computes B_l.  Group g takes interaction-list entries g,
// g + 64/G, ... (group 0 also the L2L), and the groups' sums are added by
// shuffles, so the list is walked 64/G entries at a time.
//   L2L  B_l  = 2^-(l+1) sum_{m>=l} C(m,l) Bp_m E^(m-l)
//   M2L  B_l += (-1)^l t^l sum_k C(k+l,l) A_k t^(k+1),  t = 1/T
// (A_k reaches lane l by shuffles; t^(l+1) by a prefix product over lanes).
template <int P>
__global__ __launch_bounds__(256) void fmm_down_kernel(FmmTree t, int l)
{
    constexpr int G = Lanes<P>::G;
    constexpr int NG = 64 / G;
    const int64_t nbox = int64_t(1) << (2 * l);
    const int64_t m = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & (G - 1);
    const int grp = (int)((threadIdx.x & 63) / G);
    if (m >= nbox) return;  // whole wavefronts leave together
    const int64_t bi = level_off(l) + m;
    if (t.cnt[bi] == 0 && t.need[bi] == 0) return;
    const int p = t.p;
    const bool act = lane < p;
    cplx b = {0.0, 0.0};
    const int ix = (int)compact_bits((uint32_t)m), iz = (int)compact_bits((uint32_t)m >> 1);
    if (l > 2 && grp == 0) {
        const int64_t pi = level_off(l - 1) + (m >> 2);
        const int ch = (int)(m & 3);
        const cplx e = {(ch & 1) ? 0.25 : -0.25, (ch & 2) ? 0.25 : -0.25};
        const cplx mine = act ? t.B[pi * P + lane] : cplx{0.0, 0.0};
        cplx ep = {1.0, 0.0};  // E^(mm - lane)
        for (int mm = 0; mm < p; ++mm) {
            const cplx bp = shfl_c(mine, mm, G);
            if (act && mm >= lane) {
                b = cadd(b, cscale(cmul(bp, ep), t.binom[mm * kBinom + lane]));
                ep = cmul(ep, e);
            }
        }
        b = cscale(b, ldexp(1.0, -(lane + 1)));
    }
    // binomials C(k + lane, lane) for this lane
    double cb[P];
#pragma unroll
    for (int k = 0; k < P; ++k) cb[k] = t.binom[(k + lane) * kBinom + lane];
    const int side = 1 << l;
    const int pos = (ix & 1) | ((iz & 1) << 1);
    const int* il = t.il + pos * t.il_stride;
    const int nq = t.il_n[pos];
    // Entry q's source box, its count and its lane's coefficient, loaded one
    // entry ahead of the arithmetic (each entry is otherwise three dependent
    // memory round trips); an entry off the grid reads this box's slot.
    auto fetch = [&](int q, int& cn, cplx& a, double4& f) {
        const int sx = ix + il[2 * q], sz = iz + il[2 * q + 1];
        const bool in = sx >= 0 && sz >= 0 && sx < side && sz < side;
        const int64_t si = in ? level_off(l) + morton((uint32_t)sx, (uint32_t)sz) : bi;
        cn = in ? t.cnt[si] : 0;
        a = act ? t.A[si * P + lane] : cplx{0.0, 0.0};
        f = *reinterpret_cast<const double4*>(t.m2l + 4 * ((int64_t)(pos * t.il_cap + q) * G + lane));
    };
    int ncn = 0;
    cplx na = {0.0, 0.0};
    double4 nf = {0.0, 0.0, 0.0, 0.0};
    if (grp < nq) fetch(grp, ncn, na, nf);
    for (int q = grp; q < nq; q += NG) {
        const int cn = ncn;
        const cplx a_src = na;
        const double4 f = nf;
        if (q + NG < nq) fetch(q + NG, ncn, na, nf);
        if (cn == 0) continue;
        // B_l += (-1)^l t^l sum_k C(k+l, l) A_k t^(k+1), t = 1/T: lane k
        // contributes A_k t^(k+1), each lane sums its binomial row
        const cplx ak = act ? cmul_fma(a_src, cplx{f.x, f.y}) : cplx{0.0, 0.0};
        cplx acc = {0.0, 0.0};
        if constexpr (G == 16) {
            RowSum16<0, P>::run(ak, cb, acc);
        } else {
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const cplx v = shfl_c(ak, k, G);
                acc = {__builtin_fma(v.re, cb[k], acc.re), __builtin_fma(v.im, cb[k], acc.im)};
            }
        }
        b = cadd(b, cmul_fma(acc, cplx{f.z, f.w}));
    }
    for (int o = G; o < 64; o <<= 1) {
        b.re += __shfl_xor(b.re, o);
        b.im += __shfl_xor(b.im, o);
    }
    if (grp == 0 && lane < P) t.B[bi * P + lane] = act ? b : cplx{0.0, 0.0};
}

// L2P + P2P: one thread per particle in sorted order.
template <int P>
__global__ void fmm_eval_kernel(FmmTree t, int n, const double4* __restrict__ sp,
                                const uint32_t* __restrict__ skey, const int* __restrict__ perm,
                                double* __restrict__ u, double* __restrict__ w)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n;
    const uint32_t m = valid ? skey[k] : 0u;
    const int L = t.L;
    const int ix = (int)compact_bits(m), iz = (int)compact_bits(m >> 1);
    const int side = 1 << L;
    // near-field size of this target's leaf: heavy ones go to fmm_heavy_kernel
    int work = 0;
    if (valid)
        for (int q = 0; q < t.near_n; ++q) {
            const int sx = ix + t.nearo[2 * q], sz = iz + t.nearo[2 * q + 1];
            if (sx < 0 || sz < 0 || sx >= side || sz >= side) continue;
            const uint32_t lm = morton((uint32_t)sx, (uint32_t)sz);
            work += t.lbeg[lm + 1] - t.lbeg[lm];
        }
    const bool heavy = work > kFmmHeavy;
    const uint64_t hm = __ballot(heavy);
    if (hm) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)hm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(t.heavy, __popcll(hm));
        base = __shfl(base, leader);
        if (heavy) t.heavy[1 + base + __popcll(hm & ((uint64_t(1) << lane) - 1))] = k;
    }
    if (!valid || heavy) return;
    const double4 me = sp[k];
    const double s = t.S / (double)(1 << L);
    double fu = 0.0, fw = 0.0;
    if (L >= 2) {
        const int64_t bi = level_off(L) + m;
        const cplx c = box_centre(t, L, m);
        const cplx wv = {(me.x - c.re) / s, (me.y - c.im) / s};
        cplx phi = {0.0, 0.0};
#pragma unroll
        for (int ll = P - 1; ll >= 0; --ll)
            if (ll < t.p) phi = cadd(cmul(phi, wv), t.B[bi * P + ll]);
        fu = phi.im / s;
        fw = phi.re / s;
    }
    // near field with the reference's regularised kernel
    double nu = 0.0, nw = 0.0;
    for (int q = 0; q < t.near_n; ++q) {
        const int sx = ix + t.nearo[2 * q], sz = iz + t.nearo[2 * q + 1];
        if (sx < 0 || sz < 0 || sx >= side || sz >= side) continue;
        const uint32_t lm = morton((uint32_t)sx, (uint32_t)sz);
        const int b = t.lbeg[lm], e = t.lbeg[lm + 1];
        int j = b;
        for (; j + 3 < e; j += 4) {
            const double4 s0 = sp[j], s1 = sp[j + 1], s2 = sp[j + 2], s3 = sp[j + 3];
            near_pair64<true>(me.x, me.y, s0, nu, nw, j == k);
            near_pair64<true>(me.x, me.y, s1, nu, nw, j + 1 == k);
            near_pair64<true>(me.x, me.y, s2, nu, nw, j + 2 == k);
            near_pair64<true>(me.x, me.y, s3, nu, nw, j + 3 == k);
        }
        for (; j < e; ++j) near_pair64<true>(me.x, me.y, sp[j], nu, nw, j == k);
    }
    const int i = perm[k];
    u[i] = (fu + nu) * kRpi2;
    w[i] = (fw + nw) * kRpi2;
}

// The heavy targets of fmm_eval_kernel, one wavefront each: the lanes split
// every near leaf's sources (lane l takes l, l+64, ...), their sums are added
// by shuffles, and lane 0 adds the local expansion (L2P).  A fixed set of
// wavefronts walks the list.
template <int P>
__global__ __launch_bounds__(256) void fmm_heavy_kernel(FmmTree t, const double4* __restrict__ sp,
                                                        const uint32_t* __restrict__ skey,
                                                        const int* __restrict__ perm,
                                                        double* __restrict__ u,
                                                        double* __restrict__ w)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    const int cnt = t.heavy[0];
    const int L = t.L;
    const int side = 1 << L;
    const double s = t.S / (double)side;
    for (int q = wave; q < cnt; q += nwaves) {
        const int k = t.heavy[1 + q];
        const double4 me = sp[k];
        const uint32_t m = skey[k];
        const int ix = (int)compact_bits(m), iz = (int)compact_bits(m >> 1);
        double nu = 0.0, nw = 0.0;
        for (int qq = 0; qq < t.near_n; ++qq) {
            const int sx = ix + t.nearo[2 * qq], sz = iz + t.nearo[2 * qq + 1];
            if (sx < 0 || sz < 0 || sx >= side || sz >= side) continue;
            const uint32_t lm = morton((uint32_t)sx, (uint32_t)sz);
            const int b = t.lbeg[lm], e = t.lbeg[lm + 1];
            for (int j = b + lane; j < e; j += 64)
                if (j != k) near_pair64(me.x, me.y, sp[j], nu, nw);
        }
        for (int o = 32; o > 0; o >>= 1) {
            nu += __shfl_xor(nu, o);
            nw += __shfl_xor(nw, o);
        }
        if (lane == 0) {
            double fu = 0.0, fw = 0.0;
            if (L >= 2) {
                const int64_t bi = level_off(L) + m;
                const cplx c = box_centre(t, L, m);
                const cplx wv = {(me.x - c.re) / s, (me.y - c.im) / s};
                cplx phi = {0.0, 0.0};
#pragma unroll
                for (int ll = P - 1; ll >= 0; --ll)
                    if (ll < t.p) phi = cadd(cmul(phi, wv), t.B[bi * P + ll]);
                fu = phi.im / s;
                fw = phi.re / s;
            }
            const int i = perm[k];
            u[i] = (fu + nu) * kRpi2;
            w[i] = (fw + nw) * kRpi2;
        }
    }
}

// Leaves holding a target need their local expansions.
__global__ void fmm_mark_kernel(int nt, const double* __restrict__ tx, const double* __restrict__ tz,
                                double x0, double z0, double sl, int L, int* __restrict__ need)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    const int side = 1 << L;
    const int ix = min(max((int)floor((tx[k] - x0) / sl), 0), side - 1);
    const int iz = min(max((int)floor((tz[k] - z0) / sl), 0), side - 1);
    need[level_off(L) + morton((uint32_t)ix, (uint32_t)iz)] = 1;
}

// Targets: L2P at the target's leaf plus every source of the near leaves
// (no pair skipped, as the direct target sum with no guard).  One wavefront
// per target: the lanes share the near-field sources and their partial sums
// are added by shuffles.
template <int P>
__global__ __launch_bounds__(64) void fmm_target_kernel(FmmTree t, int nt, const double* __restrict__ tx,
                                                        const double* __restrict__ tz,
                                                        const double4* __restrict__ sp,
                                                        double* __restrict__ u, double* __restrict__ w)
{
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= nt) return;
    const int L = t.L;
    const int side = 1 << L;
    const double s = t.S / (double)side;
    const double xk = tx[k], zk = tz[k];
    const int ix = min(max((int)floor((xk - t.x0) / s), 0), side - 1);
    const int iz = min(max((int)floor((zk - t.z0) / s), 0), side - 1);
    const uint32_t m = morton((uint32_t)ix, (uint32_t)iz);
    const int64_t bi = level_off(L) + m;
    const cplx c = box_centre(t, L, m);
    const cplx wv = {(xk - c.re) / s, (zk - c.im) / s};
    cplx phi = {0.0, 0.0};
#pragma unroll
    for (int ll = P - 1; ll >= 0; --ll)
        if (ll < t.p) phi = cadd(cmul(phi, wv), t.B[bi * P + ll]);
    double nu = 0.0, nw = 0.0;
    for (int q = 0; q < t.near_n; ++q) {
        const int sx = ix + t.nearo[2 * q], sz = iz + t.nearo[2 * q + 1];
        if (sx < 0 || sz < 0 || sx >= side || sz >= side) continue;
        const uint32_t lm = morton((uint32_t)sx, (uint32_t)sz);
        const int b = t.lbeg[lm], e = t.lbeg[lm + 1];
        for (int j = b + lane; j < e; j += 64) near_pair64(xk, zk, sp[j], nu, nw);
    }
    for (int o = 32; o > 0; o >>= 1) {
        nu += __shfl_xor(nu, o);
        nw += __shfl_xor(nw, o);
    }
    if (lane == 0) {
        u[k] = (phi.im / s + nu) * kRpi2;
        w[k] = (phi.re / s + nw) * kRpi2;
    }
}

template <int P>
void fmm_levels(dvm_ctx_t ctx, const FmmTree& t, int n, const double4* sp, const uint32_t* skey,
                const int* perm, double* u, double* w, int nt, const double* tx, const double* tz,
                double* tu, double* tw)
{
    hipStream_t st = ctx->stream;
    const int64_t nleaf = int64_t(1) << (2 * t.L);
    {
        ProfScope ps(ctx, "fmm_p2m_kernel");
        DVM_HIP_CHECK(hipMemsetAsync(t.heavy, 0, sizeof(int), st));
        fmm_p2m_kernel<P><<<cdiv(nleaf, 64), 64, 0, st>>>(t, sp);
        DVM_LAUNCH_CHECK();
        fmm_p2m_heavy_kernel<P><<<ctx->num_cus * 4, 256, 0, st>>>(t, sp);
        DVM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(ctx, "fmm_m2m_kernel");
        for (int l = t.L - 1; l >= 2; --l) {
            fmm_m2m_kernel<P><<<cdiv(int64_t(1) << (2 * l), 64), 64, 0, st>>>(t, l);
            DVM_LAUNCH_CHECK();
        }
    }
    {
        ProfScope ps(ctx, "fmm_down_kernel");
        for (int l = 2; l <= t.L; ++l) {
            fmm_down_kernel<P><<<cdiv((int64_t(1) << (2 * l)) * 64, 256), 256, 0, st>>>(t, l);
            DVM_LAUNCH_CHECK();
        }
    }
    {
        ProfScope ps(ctx, "fmm_eval_kernel");
        DVM_HIP_CHECK(hipMemsetAsync(t.heavy, 0, sizeof(int), st));
        fmm_eval_kernel<P><<<cdiv(n, 128), 128, 0, st>>>(t, n, sp, skey, perm, u, w);
        DVM_LAUNCH_CHECK();
        fmm_heavy_kernel<P><<<ctx->num_cus * 8, 256, 0, st>>>(t, sp, skey, perm, u, w);
        DVM_LAUNCH_CHECK();
    }
    if (nt > 0) {
        ProfScope ps(ctx, "fmm_target_kernel");
        fmm_target_kernel<P><<<nt, 64, 0, st>>>(t, nt, tx, tz, sp, tu, tw);
        DVM_LAUNCH_CHECK();
    }
}

constexpr int kOrders[] = {8, 12, 16, 20, 24, 32};
constexpr int kRedBlocks = 256;

// Box of the finite positions, largest core size and the count of
// non-finite entries: one partial per block, [xmin, xmax, zmin, zmax, smax, bad].
__global__ void fmm_bbox_kernel(int n, const double* __restrict__ x, const double* __restrict__ z,
                                const double* __restrict__ s, const double* __restrict__ c,
                                double* __restrict__ part)
{
    __shared__ double r[6][256];
    double a0 = INFINITY, a1 = -INFINITY, a2 = INFINITY, a3 = -INFINITY, a4 = 0.0, a5 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double xi = x[i], zi = z[i], si = s[i], ci = c[i];
        if (isfinite(xi) && isfinite(zi) && isfinite(si) && isfinite(ci)) {
            a0 = fmin(a0, xi);
            a1 = fmax(a1, xi);
            a2 = fmin(a2, zi);
            a3 = fmax(a3, zi);
            a4 = fmax(a4, fabs(si));
        } else {
            a5 += 1.0;
        }
    }
    const int t = threadIdx.x;
    r[0][t] = a0; r[1][t] = a1; r[2][t] = a2; r[3][t] = a3; r[4][t] = a4; r[5][t] = a5;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            r[0][t] = fmin(r[0][t], r[0][t + h]);
            r[1][t] = fmax(r[1][t], r[1][t + h]);
            r[2][t] = fmin(r[2][t], r[2][t + h]);
            r[3][t] = fmax(r[3][t], r[3][t + h]);
            r[4][t] = fmax(r[4][t], r[4][t + h]);
            r[5][t] += r[5][t + h];
        }
        __syncthreads();
    }
    if (t < 6) part[blockIdx.x * 6 + t] = r[t][0];
}

}  // namespace

// The expansion order used for a requested one: the next supported order.
int fmm_terms(int order)
{
    for (int p : kOrders)
        if (order <= p) return p;
    return 0;
}

bool fmm_self(dvm_ctx_t ctx, int64_t n, const double* x, const double* z, const double* c,
              const double* s, double* u, double* w, int64_t nt, const double* tx,
              const double* tz, double* tu, double* tw)
{
    const int order = ctx->fmm_order;
    const double theta = ctx->fmm_theta;
    require(n >= 0 && n < (int64_t(1) << 31), DVM_E_SIZE, "blob count out of range");
    const int P = fmm_terms(order);
    require(order >= 1 && P > 0, DVM_E_INVALID, "FMM expansion order must be 1..32");
    require(theta > 0.2 && theta <= 1.0, DVM_E_INVALID, "FMM theta must be in (0.2, 1]");
    require(nt >= 0 && nt < (int64_t(1) << 31), DVM_E_SIZE, "target count out of range");
    require(nt == 0 || (tx && tz && tu && tw), DVM_E_INVALID, "null target array");
    ctx->fmm_last_level = 0;
    if (n < 2) return false;
    hipStream_t st = ctx->stream;

    // near offsets: centre distance <= sqrt(2)/theta leaf widths; interaction
    // list per child position: children of the parent's near boxes not near
    const double r2n = 2.0 / (theta * theta) * (1.0 + 1e-12);
    const int rmax = (int)std::ceil(std::sqrt(r2n)) + 1;
    std::vector<int> nearo;
    for (int dz = -rmax; dz <= rmax; ++dz)
        for (int dx = -rmax; dx <= rmax; ++dx)
            if ((double)(dx * dx + dz * dz) <= r2n) {
                nearo.push_back(dx);
                nearo.push_back(dz);
            }
    std::vector<int> il[4];
    for (int pos = 0; pos < 4; ++pos) {
        const int bx = pos & 1, bz = pos >> 1;
        for (size_t q = 0; q < nearo.size(); q += 2)
            for (int a = 0; a < 2; ++a)
                for (int cc = 0; cc < 2; ++cc) {
                    const int ox = 2 * nearo[q] + a - bx, oz = 2 * nearo[q + 1] + cc - bz;
                    if ((double)(ox * ox + oz * oz) <= r2n) continue;
                    il[pos].push_back(ox);
                    il[pos].push_back(oz);
                }
    }
    // smallest distance between points of two leaves that are not near (leaf widths)
    double dmin = 1e300;
    for (int dz = -2 * rmax; dz <= 2 * rmax; ++dz)
        for (int dx = -2 * rmax; dx <= 2 * rmax; ++dx) {
            if ((double)(dx * dx + dz * dz) <= r2n) continue;
            const double ax = std::max(std::abs(dx) - 1, 0), az = std::max(std::abs(dz) - 1, 0);
            dmin = std::min(dmin, std::sqrt(ax * ax + az * az));
        }

    // pinned host memory of this call: the partials read back, then the
    // interaction tables copied to the device (written after the read-back's
    // synchronisation, so the previous call's copies on this stream are done)
    const size_t il_stride = std::max({il[0].size(), il[1].size(), il[2].size(), il[3].size()}) / 2;
    const size_t rb_bytes = (sizeof(double) * (6 * kRedBlocks + 2 * nt) + 255) & ~size_t(255);
    const size_t tab_ints = 8 * il_stride + nearo.size();
    const int G = P <= 8 ? 8 : P <= 16 ? 16 : 32;  // lanes per coefficient group (Lanes<P>)
    const size_t m2l_doubles = 4 * il_stride * (size_t)G * 4;
    const size_t m2l_off = rb_bytes + ((sizeof(int) * tab_ints + 255) & ~size_t(255));
    char* pin = static_cast<char*>(pinned_get(ctx, m2l_off + sizeof(double) * m2l_doubles));

    // extent, largest core, non-finite entries
    double* red = static_cast<double*>(ws_get(ctx, sizeof(double) * 6 * kRedBlocks));
    fmm_bbox_kernel<<<kRedBlocks, 256, 0, st>>>((int)n, x, z, s, c, red);
    DVM_LAUNCH_CHECK();
    // (targets are few: their coordinates come to the host with the partials)
    double* hr = reinterpret_cast<double*>(pin);
    DVM_HIP_CHECK(hipMemcpyAsync(hr, red, sizeof(double) * 6 * kRedBlocks, hipMemcpyDeviceToHost, st));
    if (nt > 0) {
        DVM_HIP_CHECK(hipMemcpyAsync(hr + 6 * kRedBlocks, tx, sizeof(double) * nt,
                                     hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipMemcpyAsync(hr + 6 * kRedBlocks + nt, tz, sizeof(double) * nt,
                                     hipMemcpyDeviceToHost, st));
    }
    DVM_HIP_CHECK(hipStreamSynchronize(st));
    double x0 = INFINITY, x1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY, smax = 0.0, bad = 0.0;
    for (int k = 0; k < kRedBlocks; ++k) {
        x0 = std::fmin(x0, hr[6 * k]);
        x1 = std::fmax(x1, hr[6 * k + 1]);
        z0 = std::fmin(z0, hr[6 * k + 2]);
        z1 = std::fmax(z1, hr[6 * k + 3]);
        smax = std::fmax(smax, hr[6 * k + 4]);
        bad += hr[6 * k + 5];
    }
    for (int64_t k = 0; k < nt; ++k) {  // the box holds the targets too
        const double a = hr[6 * kRedBlocks + k], b = hr[6 * kRedBlocks + nt + k];
        if (!std::isfinite(a) || !std::isfinite(b)) bad += 1.0;
        x0 = std::fmin(x0, a);
        x1 = std::fmax(x1, a);
        z0 = std::fmin(z0, b);
        z1 = std::fmax(z1, b);
    }
    if (bad > 0.0) return false;  // non-finite input: the direct sum reproduces it
    double S = std::fmax(x1 - x0, z1 - z0);
    if (!(S > 0.0)) return false;
    S *= 1.0 + 1e-9;
    // depth: about kLeafMean blobs per leaf over the box, leaves no narrower than
    // sqrt(10) * max(sigma) / dmin (every regularised pair in the near field)
    int L = 0;
    while (L < kFmmMaxLevel && (double)n / (double)(int64_t(1) << (2 * (L + 1))) >= kLeafMean) ++L;
    const double smin = std::sqrt(10.0) * smax * (1.0 + 1e-9) / dmin;
    while (L > 0 && S / (double)(1 << L) < smin) --L;
    int Ls = L;  // deepest level the core-size rule allows
    double occ_leaves = (double)(int64_t(1) << (2 * L));  // occupied leaves of level L (at most)
    while (Ls < kFmmMaxLevel && S / (double)(1 << (Ls + 1)) >= smin) ++Ls;
    if (Ls > L && L >= 1) {
        // A clustered set fills few of the box's leaves: the mean over the box
        // understates the blobs per occupied leaf.  Count the occupied boxes of
        // the deeper levels the core rule allows and take the deepest whose
        // occupied leaves still hold kLeafMean blobs on average.
        const int64_t words = (int64_t(1) << (2 * Ls)) / 64 > 0 ? (int64_t(1) << (2 * Ls)) / 64 : 1;
        const int ny = Ls - L;  // levels Ls, Ls - 1, ..., L + 1
        char* occ = static_cast<char*>(ws_get(ctx, sizeof(unsigned long long) * words + 256));
        unsigned long long* bm = reinterpret_cast<unsigned long long*>(occ);
        int* ocnt = reinterpret_cast<int*>(occ + sizeof(unsigned long long) * words);
        DVM_HIP_CHECK(hipMemsetAsync(occ, 0, sizeof(unsigned long long) * words + 256, st));
        fmm_occ_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, x, z, x0, z0, S / (double)(1 << Ls),
                                                      1 << Ls, bm);
        DVM_LAUNCH_CHECK();
        const unsigned gx = (unsigned)std::min<int64_t>(cdiv(int64_t(1) << (2 * Ls), 256), 1024);
        fmm_occ_count_kernel<<<dim3(gx, ny), 256, 0, st>>>(bm, Ls, ocnt);
        DVM_LAUNCH_CHECK();
        int* hc = reinterpret_cast<int*>(pin);  // the partials are read already
        DVM_HIP_CHECK(hipMemcpyAsync(hc, ocnt, sizeof(int) * ny, hipMemcpyDeviceToHost, st));
        DVM_HIP_CHECK(hipStreamSynchronize(st));
        for (int y = 0; y < ny; ++y)
            if (hc[y] > 0 && (double)n / (double)hc[y] >= kLeafMean) {
                L = Ls - y;
                occ_leaves = (double)hc[y];
                break;
            }
    }
    // Leaf width within the level: the root side is free above the extent, so
    // the leaves are narrowed to the width at which the occupied ones hold
    // about kLeafMean blobs (the chosen level's count scaled by area), but no
    // narrower than the core-size rule allows.  The power-of-two cut of the
    // extent alone leaves them up to twice that wide, and the near field up to
    // four times as large (a remeshed wake at its core-size limit).
    if (L >= 1) {
        const double wl = S / (double)(1 << L);
        const double wf = std::fmax(smin, wl * std::sqrt(kLeafMean * occ_leaves / (double)n));
        int k = L;
        while (k <= kFmmMaxLevel && wf * (double)(1 << k) < S) ++k;
        if (wf < wl && k <= kFmmMaxLevel) {
            S = wf * (double)(1 << k);
            L = k;
        }
    }
    if (L < 2) return false;  // no well-separated boxes: direct sum
    ctx->fmm_last_level = L;

    const int64_t nbox = level_off(L + 1);
    const int64_t nleaf = int64_t(1) << (2 * L);
    size_t sort_bytes = 0;
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t*)nullptr,
                                                     (uint32_t*)nullptr, (int*)nullptr, (int*)nullptr,
                                                     (int)n, 0, 2 * L, st));
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t need = al(sizeof(double) * kBinom * kBinom) + al(sizeof(int) * 8 * il_stride) +
                        al(sizeof(int) * nearo.size()) + 2 * al(sizeof(uint32_t) * n) +
                        2 * al(sizeof(int) * n) + al(sizeof(double4) * n) + al(sizeof(int) * (nleaf + 1)) +
                        2 * al(sizeof(int) * nbox) + 2 * al(sizeof(cplx) * P * nbox) + al(sort_bytes) +
                        al(sizeof(int) * (n + 1)) + al(sizeof(double) * m2l_doubles);
    char* p = static_cast<char*>(ws_get(ctx, need));
    auto take = [&](size_t bytes) {
        char* q = p;
        p += al(bytes);
        return q;
    };
    double* binom = reinterpret_cast<double*>(take(sizeof(double) * kBinom * kBinom));
    int* ild = reinterpret_cast<int*>(take(sizeof(int) * 8 * il_stride));
    int* neard = reinterpret_cast<int*>(take(sizeof(int) * nearo.size()));
    uint32_t* key = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * n));
    uint32_t* skey = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * n));
    int* idx = reinterpret_cast<int*>(take(sizeof(int) * n));
    int* perm = reinterpret_cast<int*>(take(sizeof(int) * n));
    double4* sp = reinterpret_cast<double4*>(take(sizeof(double4) * n));
    int* lbeg = reinterpret_cast<int*>(take(sizeof(int) * (nleaf + 1)));
    int* cnt = reinterpret_cast<int*>(take(sizeof(int) * nbox));
    int* needb = reinterpret_cast<int*>(take(sizeof(int) * nbox));
    cplx* A = reinterpret_cast<cplx*>(take(sizeof(cplx) * P * nbox));
    cplx* B = reinterpret_cast<cplx*>(take(sizeof(cplx) * P * nbox));
    void* sort_tmp = take(sort_bytes);
    int* heavy = reinterpret_cast<int*>(take(sizeof(int) * (n + 1)));
    double* m2ld = reinterpret_cast<double*>(take(sizeof(double) * m2l_doubles));

    // tables: interaction lists padded to il_stride entries per child position,
    // then the near offsets, staged in the pinned buffer
    int* ilh = reinterpret_cast<int*>(pin + rb_bytes);
    std::fill(ilh, ilh + 8 * il_stride, 0);
    FmmTree t{};
    for (int pos = 0; pos < 4; ++pos) {
        std::copy(il[pos].begin(), il[pos].end(), ilh + pos * 2 * il_stride);
        t.il_n[pos] = (int)(il[pos].size() / 2);
    }
    int* nearh = ilh + 8 * il_stride;
    std::copy(nearo.begin(), nearo.end(), nearh);
    DVM_HIP_CHECK(hipMemcpyAsync(ild, ilh, sizeof(int) * 8 * il_stride, hipMemcpyHostToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(neard, nearh, sizeof(int) * nearo.size(), hipMemcpyHostToDevice, st));
    // M2L factors of each interaction entry: with T = -(dx, dz) the offset from
    // the source box to the target box, t = 1/T, lane l gets t^(l+1) and
    // (-1)^l t^l (the same at every level: offsets are in box widths)
    double* m2lh = reinterpret_cast<double*>(pin + m2l_off);
    std::fill(m2lh, m2lh + m2l_doubles, 0.0);
    for (int pos = 0; pos < 4; ++pos)
        for (size_t q = 0; q < il[pos].size() / 2; ++q) {
            const double tre = -(double)il[pos][2 * q], tim = -(double)il[pos][2 * q + 1];
            const double d = tre * tre + tim * tim;
            const double vre = tre / d, vim = -tim / d;  // t = 1/T
            double pre = vre, pim = vim;                // t^(l+1)
            double fre = 1.0, fim = 0.0;                // t^l
            double* e = m2lh + 4 * ((size_t)(pos * il_stride + q) * G);
            for (int ln = 0; ln < G; ++ln) {
                const double sg = (ln & 1) ? -1.0 : 1.0;
                e[4 * ln] = pre;
                e[4 * ln + 1] = pim;
                e[4 * ln + 2] = sg * fre;
                e[4 * ln + 3] = sg * fim;
                const double nre = pre * vre - pim * vim, nim = pre * vim + pim * vre;
                fre = pre;
                fim = pim;
                pre = nre;
                pim = nim;
            }
        }
    DVM_HIP_CHECK(hipMemcpyAsync(m2ld, m2lh, sizeof(double) * m2l_doubles, hipMemcpyHostToDevice, st));
    fmm_binom_kernel<<<1, kBinom, 0, st>>>(binom);
    DVM_LAUNCH_CHECK();

    const double sl = S / (double)(1 << L);
    fmm_key_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, x, z, x0, z0, sl, 1 << L, key, idx);
    DVM_LAUNCH_CHECK();
    DVM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, key, skey, idx, perm, (int)n,
                                                     0, 2 * L, st));
    fmm_gather_kernel<<<cdiv(n, 256), 256, 0, st>>>((int)n, perm, x, z, c, s, sp);
    DVM_LAUNCH_CHECK();
    fmm_lbeg_kernel<<<cdiv(nleaf + 1, 256), 256, 0, st>>>((int)n, skey, (int)nleaf, lbeg);
    DVM_LAUNCH_CHECK();

    t.x0 = x0;
    t.z0 = z0;
    t.S = S;
    t.L = L;
    t.p = order;
    t.binom = binom;
    t.il = ild;
    t.il_stride = (int)(2 * il_stride);
    t.il_cap = (int)il_stride;
    t.m2l = m2ld;
    t.nearo = neard;
    t.near_n = (int)(nearo.size() / 2);
    t.lbeg = lbeg;
    t.cnt = cnt;
    t.need = needb;
    DVM_HIP_CHECK(hipMemsetAsync(needb, 0, sizeof(int) * nbox, st));
    if (nt > 0) {
        fmm_mark_kernel<<<cdiv(nt, 256), 256, 0, st>>>((int)nt, tx, tz, x0, z0, sl, L, needb);
        DVM_LAUNCH_CHECK();
    }
    t.A = A;
    t.B = B;
    t.heavy = heavy;
    switch (P) {
    case 8: fmm_levels<8>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    case 12: fmm_levels<12>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    case 16: fmm_levels<16>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    case 20: fmm_levels<20>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    case 24: fmm_levels<24>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    default: fmm_levels<32>(ctx, t, (int)n, sp, skey, perm, u, w, (int)nt, tx, tz, tu, tw); break;
    }
    // no wait here: the tables stay in the pinned buffer, which the next call
    // on this context rewrites only after its read-back has synchronised
    return true;
}

}  // namespace dvmi

using namespace dvmi;

extern "C" {

dvm_status dvm_set_fmm(dvm_ctx_t ctx, int expansion_order, double theta)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        require(expansion_order >= 1 && fmm_terms(expansion_order) > 0, DVM_E_INVALID,
                "FMM expansion order must be 1..32");
        require(theta > 0.2 && theta <= 1.0, DVM_E_INVALID, "FMM theta must be in (0.2, 1]");
        ctx->fmm_order = expansion_order;
        ctx->fmm_theta = theta;
    });
}

dvm_status dvm_fmm_last_level(dvm_ctx_t ctx, int* level)
{
    if (!ctx || !level) return DVM_E_INVALID;
    *level = ctx->fmm_last_level;
    return DVM_OK;
}

}  // extern "C"
