// Vortex sheet on the device: influence matrix, least-squares operator,
// per-step gamma solve, sheet-on-blob velocities, loads, nascent release and
// surface crossing (src/VortexSheet.cpp).
//
// Layout: panel data in SoA FP64 arrays plus one 64-byte PanelConst record per
// panel; the per-pair loops stream panels with wave-uniform addresses (scalar
// loads), one blob per lane.
#include "dvm_device.hpp"
#include "dvm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

namespace dvmi {

// ----------------------------------------------------------------- set-up

// One (i, j) entry of A: src/VortexSheet.cpp:117-161.  Panel j contributes its
// start-end coefficient p(i,j) and panel j-1 its finish-end coefficient q.
__device__ void panel_pq(double xci, double zci, double thetai, double xj, double zj,
                         double thetaj, double dsj, double& p, double& q)
{
    const double c1 = -(xci - xj) * cos(thetaj) - (zci - zj) * sin(thetaj);
    const double c2 = (xci - xj) * (xci - xj) + (zci - zj) * (zci - zj);
    const double c3 = sin(thetai - thetaj);
    const double c4 = cos(thetai - thetaj);
    const double c5 = (xci - xj) * sin(thetaj) - (zci - zj) * cos(thetaj);
    const double c6 = log(1.0 + dsj * ((dsj + 2 * c1) / c2));
    const double c7 = atan2((c5 * dsj), (c2 + c1 * dsj));
    const double c9 = (xci - xj) * cos(thetai - 2.0 * thetaj) - (zci - zj) * sin(thetai - 2.0 * thetaj);
    q = c4 + 0.5 * c9 * c6 / dsj - ((c1 * c3 + c4 * c5) * c7) / dsj;
    p = 0.5 * c4 * c6 + c3 * c7 - q;
}

__global__ void influence_kernel(int nl, const double* __restrict__ x, const double* __restrict__ z,
                                 const double* __restrict__ xc, const double* __restrict__ zc,
                                 const double* __restrict__ theta, const double* __restrict__ ds,
                                 double* __restrict__ A)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= nl) return;
    if (i == nl) {
        A[(int64_t)nl * nl + j] = ds[j];
        return;
    }
    const int jm = j == 0 ? nl - 1 : j - 1;
    double p, q, pd, qd;
    if (i == j) {
        p = -1.0;
    } else {
        panel_pq(xc[i], zc[i], theta[i], x[j], z[j], theta[j], ds[j], p, qd);
    }
    if (i == jm) {
        q = 1.0;
    } else {
        panel_pq(xc[i], zc[i], theta[i], x[jm], z[jm], theta[jm], ds[jm], pd, q);
    }
    A[(int64_t)i * nl + j] = -kRpi2 * (p + q);
}

// The least-squares operator P = (A^T A)^-1 A^T (the solve of
// src/VortexSheet.cpp:214 for every right-hand side at once), by
// Gauss-Jordan elimination with partial pivoting on G = [A^T A | A^T]
// (n rows, 2n+1 columns, row-major).  The pivot of column k is the row not
// yet used as a pivot with the largest |G[r][k]| (the first such row on
// ties) — the rows LU with partial pivoting would choose.  Rows are not
// swapped and the pivot row is not scaled; column k is eliminated from every
// other row, so at the end pivot row p of column i holds G[p][i] x_i.

// G = [A^T A | A^T]; A^T A summed over the n+1 rows of A in order.
__global__ void gj_build_kernel(int n, const double* __restrict__ A, double* __restrict__ G)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int W = 2 * n + 1;
    if (j >= W) return;
    if (j < n) {
        double s = 0.0;
        for (int k = 0; k <= n; ++k) s += A[(int64_t)k * n + i] * A[(int64_t)k * n + j];
        G[(int64_t)i * W + j] = s;
    } else {
        G[(int64_t)i * W + j] = A[(int64_t)(j - n) * n + i];
    }
}

constexpr int kGjMaxRows = 4096;   // at most 4096 panels per body
constexpr int kGjThreads = 512;
constexpr int kGjStrip = 17408;    // doubles of the LDS strip: n rows x (KB + 1)

// Blocked form of the elimination, KB pivot columns per pass.  Within a pass
// the steps only touch the strip of columns k0 .. k0+KB-1 and the trailing
// columns; column k is not written after step k, so the multiplier of row r
// at step k is f(r, k) = G[r][k] / G[p_k][k] from the final strip.  The value
// of pivot row p_m when it is used, U_m = T0[p_m] - sum_{c<m} f(p_m, c) U_c,
// and then every row r, T[r] -= f(r, m) U_m for m = 0 .. KB-1 (m with
// p_m != r), give the trailing columns the same operations, in the same order,
// as KB single steps.

// Pivot-candidate order: larger |value| first, then the lower row; n (no
// row) last.
__device__ __forceinline__ void gj_better(double ov, int oi, int n, double& bv, int& bi)
{
    if (oi < n && (bi == n || ov > bv || (ov == bv && oi < bi))) {
        bv = ov;
        bi = oi;
    }
}

// The pass's pivot steps on the strip (in LDS, row stride KB + 1), one
// workgroup, a thread per row (rows t, t + kGjThreads, ...; which of them are
// used as pivots kept in a register mask): each step takes the unused row
// with the largest |G[r][k]| (the first on ties) and eliminates column k from
// the other rows over the strip's later columns, the row segment in
// registers, and offers the row's next-column value as the next pivot
// candidate; one barrier per step (the per-wave candidates double-buffered).
// Then the strip back to G and the pivot rows' trailing columns into U.
template <int KB>
__global__ __launch_bounds__(kGjThreads) void gj_panel_kernel(int n, int k0, double* __restrict__ G,
                                                              int* __restrict__ pivrow,
                                                              double* __restrict__ U)
{
    constexpr int LD = KB + 1;
    constexpr int NW = kGjThreads / 64;
    __shared__ double S[kGjStrip];
    __shared__ unsigned char used[kGjMaxRows];
    __shared__ double wv[2][NW];
    __shared__ int wi[2][NW];
    __shared__ int piv[KB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int W = 2 * n + 1;
    const int kb = min(KB, n - k0);
    for (int e = t; e < n * KB; e += kGjThreads) {
        const int r = e / KB, c = e - r * KB;
        S[r * LD + c] = c < kb ? G[(int64_t)r * W + k0 + c] : 0.0;
    }
    for (int r = t; r < n; r += kGjThreads) used[r] = 0;
    __syncthreads();
    for (int m = t; m < k0; m += kGjThreads) used[pivrow[m]] = 1;
    __syncthreads();
    uint32_t mine = 0;  // bit i: row t + i kGjThreads already a pivot (n <= 32 kGjThreads)
    double best = -1.0;
    int bidx = n;
    for (int r = t, i = 0; r < n; r += kGjThreads, ++i) {
        if (used[r]) {
            mine |= 1u << i;
            continue;
        }
        const double v = fabs(S[r * LD]);
        if (bidx == n || v > best) { best = v; bidx = r; }
    }
    // All KB steps run (unrolled, so the column tests below fold away); in a
    // last, narrower pass the steps c >= kb touch only the strip's padding
    // columns and record nothing.
#pragma unroll
    for (int c = 0; c < KB; ++c) {
        const int buf = c & 1;
        for (int o = 32; o > 0; o >>= 1) gj_better(__shfl_xor(best, o), __shfl_xor(bidx, o), n, best, bidx);
        if (lane == 0) {
            wv[buf][wave] = best;
            wi[buf][wave] = bidx;
        }
        __syncthreads();
        double bv = wv[buf][0];
        int p = wi[buf][0];
        for (int w = 1; w < NW; ++w) gj_better(wv[buf][w], wi[buf][w], n, bv, p);
        if (p >= n) p = 0;  // no unused row left: a padding step, any row reads in bounds
        if (t == 0) {
            piv[c] = p;
            if (c < kb) pivrow[k0 + c] = p;
        }
        if (p % kGjThreads == t) mine |= 1u << (p / kGjThreads);
        const double pk = S[p * LD + c];
        double prow[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) prow[j] = S[p * LD + j];
        best = -1.0;
        bidx = n;
        for (int r = t, i = 0; r < n; r += kGjThreads, ++i) {
            if (r == p) continue;
            const double f = S[r * LD + c] / pk;
            double row[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) row[j] = S[r * LD + j];
            double next = 0.0;
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                double v = row[j];
                v -= f * prow[j];
                next = j == c + 1 ? v : next;
                row[j] = j > c ? v : row[j];  // columns up to k keep their values
            }
#pragma unroll
            for (int j = 0; j < KB; ++j) S[r * LD + j] = row[j];
            if (!(mine >> i & 1u)) {
                const double v = fabs(next);
                if (bidx == n || v > best) { best = v; bidx = r; }
            }
        }
    }
    __syncthreads();
    for (int e = t; e < n * kb; e += kGjThreads) {
        const int r = e / kb, c = e - r * kb;
        G[(int64_t)r * W + k0 + c] = S[r * LD + c];
    }
    // the pivot rows' trailing columns before the update (T0 of U_m)
    const int wt = W - (k0 + kb);
    for (int e = t; e < kb * wt; e += kGjThreads) {
        const int m = e / wt, j = k0 + kb + (e - m * wt);
        U[(int64_t)m * W + j] = G[(int64_t)piv[m] * W + j];
    }
}

// U_m = T0[p_m] - sum_{c<m} f(p_m, c) U_c of 64 trailing columns per
// workgroup (a thread per column, the values so far in LDS), in place in U.
template <int KB>
__global__ __launch_bounds__(64) void gj_pivot_rows_kernel(int n, int k0, const double* __restrict__ G,
                                                           const int* __restrict__ pivrow,
                                                           double* __restrict__ U)
{
    __shared__ double fm[KB][KB];
    __shared__ double us[KB][64];
    const int lane = threadIdx.x;
    const int W = 2 * n + 1;
    const int kb = min(KB, n - k0);
    const int j = k0 + kb + blockIdx.x * 64 + lane;
    for (int e = lane; e < KB * KB; e += 64) {
        const int m = e / KB, c = e - m * KB;
        fm[m][c] = (c < m && m < kb) ? G[(int64_t)pivrow[k0 + m] * W + k0 + c] /
                                           G[(int64_t)pivrow[k0 + c] * W + k0 + c]
                                     : 0.0;
    }
    __syncthreads();
    if (j >= W) return;
    for (int m = 0; m < kb; ++m) {
        double v = U[(int64_t)m * W + j];
#pragma unroll 4
        for (int c = 0; c < m; ++c) v -= fm[m][c] * us[c][lane];
        us[m][lane] = v;
        U[(int64_t)m * W + j] = v;
    }
}

// The pass's trailing update T -= F U on FP64 matrix cores: F[r][m] =
// f(r, m) (0 for p_m = r and past the pass's last column), U the pivot rows'
// values when used (n x KB times KB x trailing columns).  A workgroup of 4
// waves takes 16 rows and 64 trailing columns, a wave one 16 x 16 tile: KB/4
// steps of v_mfma_f64_16x16x4f64 (lane l supplies F[l % 16][k], k = l / 16 of
// the step, and U[k][l % 16]; it receives the product's rows 4v + l / 16,
// v = 0..3, of column l % 16), then the tile is subtracted from T.  Every row
// takes the same terms as KB single elimination steps; the products are
// summed by the matrix unit before the one subtraction, so the trailing
// values differ from the stepwise order in the last bits.
constexpr int kGjTile = 16;  // rows per workgroup of the trailing update (the MFMA tile)

typedef double gj_v4d __attribute__((ext_vector_type(4)));

template <int KB>
__global__ __launch_bounds__(256) void gj_update_kernel(int n, int k0, double* __restrict__ G,
                                                        const int* __restrict__ pivrow,
                                                        const double* __restrict__ U)
{
    constexpr int KP = (KB + 3) / 4 * 4;  // whole MFMA steps (columns past KB are zero)
    __shared__ double f[kGjTile][KP];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int W = 2 * n + 1;
    const int kb = min(KB, n - k0);
    const int r0 = blockIdx.y * kGjTile;
    for (int e = t; e < kGjTile * KP; e += 256) {
        const int rr = e / KP, m = e - rr * KP;
        const int r = r0 + rr;
        double v = 0.0;
        if (r < n && m < kb) {
            const int p = pivrow[k0 + m];
            v = p == r ? 0.0 : G[(int64_t)r * W + k0 + m] / G[(int64_t)p * W + k0 + m];
        }
        f[rr][m] = v;
    }
    __syncthreads();
    const int j = k0 + kb + blockIdx.x * 64 + wave * 16 + (lane & 15);  // this lane's column
    const int kq = lane >> 4;
    gj_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m0 = 0; m0 < KP; m0 += 4) {
        const int m = m0 + kq;
        const double a = f[lane & 15][m];
        const double b = (m < kb && j < W) ? U[(int64_t)m * W + j] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (j >= W) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = r0 + 4 * v + kq;
        if (r < n) G[(int64_t)r * W + j] -= acc[v];
    }
}

template <int KB>
static void gj_passes(hipStream_t st, int n, double* G, int* pivrow, double* U)
{
    const int W = 2 * n + 1;
    for (int k0 = 0; k0 < n; k0 += KB) {
        const int kb = std::min(KB, n - k0);
        gj_panel_kernel<KB><<<1, kGjThreads, 0, st>>>(n, k0, G, pivrow, U);
        DVM_LAUNCH_CHECK();
        const int wt = W - (k0 + kb);
        gj_pivot_rows_kernel<KB><<<cdiv(wt, 64), 64, 0, st>>>(n, k0, G, pivrow, U);
        DVM_LAUNCH_CHECK();
        gj_update_kernel<KB><<<dim3(cdiv(wt, 64), cdiv(n, kGjTile)), 256, 0, st>>>(n, k0, G, pivrow, U);
        DVM_LAUNCH_CHECK();
    }
}

// Widest pass whose strip fits the LDS strip.
static int gj_pass_width(int n)
{
    for (int kb = 32; kb > 2; kb >>= 1)
        if ((int64_t)n * (kb + 1) <= kGjStrip) return kb;
    return 2;
}

// P[i][c] = G[p][n + c] / G[p][i] with p the pivot row of column i.
__global__ void gj_extract_kernel(int n, const double* __restrict__ G, const int* __restrict__ pivrow,
                                  double* __restrict__ P)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (c > n) return;
    const int W = 2 * n + 1;
    const int p = pivrow[i];
    P[(int64_t)i * (n + 1) + c] = G[(int64_t)p * W + n + c] / G[(int64_t)p * W + i];
}

void sheet_free(SheetDev& sh)
{
    if (sh.mem) (void)hipFree(sh.mem);
    sh.mem = nullptr;
}

// Nodes and weights of n-point Gauss-Legendre quadrature on [-1, 1]
// (Newton iteration on the Legendre polynomial from the Chebyshev guess).
static void gauss_legendre(int n, double* x, double* w)
{
    for (int i = 0; i < n; ++i) {
        double t = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0);
            const double dt = p1 / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-16) break;
        }
        double p0 = 1.0, p1 = t;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (t * p1 - p0) / (t * t - 1.0);
        x[i] = t;
        w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
}

struct GlRule24 {
    double x[24], w[24];
};

static const GlRule24& gl_rule24()
{
    static const GlRule24 r = [] {
        GlRule24 g;
        gauss_legendre(24, g.x, g.w);
        return g;
    }();
    return r;
}

// Geometric moments of the panels as sheet_pair places them (panel i is the
// segment from (xc, zc) along theta, length ds) about a centre c and radius r:
// for k < kSheetTerms, i0[i][k] = int (1 - s/ds) W^k ds and
// i1[i][k] = int (s/ds) W^k ds with W = (zeta(s) - c) / r (complex,
// interleaved re/im), by 24-point Gauss-Legendre quadrature, exact for these
// polynomials up to degree 47.  c, r are (cx, cz, r), or with gcen the centre
// of panel i's group (r = inf: a degenerate group, moments zero).  One
// workgroup per panel: lanes q < 24 form the powers at their quadrature point,
// lanes k < kSheetTerms add the weighted powers over the points in order.
__global__ __launch_bounds__(64) void panel_moments_kernel(int nl, const PanelConst* __restrict__ pc,
                                                           double cx, double cz, double r,
                                                           const double4* __restrict__ gcen,
                                                           GlRule24 gl, double* __restrict__ i0,
                                                           double* __restrict__ i1)
{
    __shared__ double pw[24][2 * kSheetTerms];
    __shared__ double wt[24][2];
    const int i = blockIdx.x, t = threadIdx.x;
    if (i >= nl) return;
    if (gcen) {
        const double4 c = gcen[i / kSheetGroup];
        cx = c.x;
        cz = c.y;
        r = c.z;
    }
    const bool ok = r < INFINITY;
    const PanelConst q = pc[i];
    if (t < 24) {
        const double tt = 0.5 * (gl.x[t] + 1.0);  // s / ds
        const double wq = 0.5 * gl.w[t] * q.ds;
        wt[t][0] = wq * (1.0 - tt);
        wt[t][1] = wq * tt;
        const double zr = (q.xc + tt * q.ds * q.cs - cx) / r;
        const double zi = (q.zc + tt * q.ds * q.sn - cz) / r;
        double pr = 1.0, pim = 0.0;
        for (int k = 0; k < kSheetTerms; ++k) {
            pw[t][2 * k] = pr;
            pw[t][2 * k + 1] = pim;
            const double nr = pr * zr - pim * zi;
            pim = pr * zi + pim * zr;
            pr = nr;
        }
    }
    __syncthreads();
    if (t < kSheetTerms) {
        double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
        if (ok) {
            for (int n = 0; n < 24; ++n) {
                a0r += wt[n][0] * pw[n][2 * t];
                a0i += wt[n][0] * pw[n][2 * t + 1];
                a1r += wt[n][1] * pw[n][2 * t];
                a1i += wt[n][1] * pw[n][2 * t + 1];
            }
        }
        const size_t o = 2 * ((size_t)i * kSheetTerms + t);
        i0[o] = a0r;
        i0[o + 1] = a0i;
        i1[o] = a1r;
        i1[o + 1] = a1i;
    }
}

void panel_moments(hipStream_t st, int nl, const PanelConst* pc, double cx, double cz, double r,
                   const double4* gcen, double* i0, double* i1)
{
    if (nl <= 0) return;
    panel_moments_kernel<<<nl, 64, 0, st>>>(nl, pc, cx, cz, r, gcen, gl_rule24(), i0, i1);
    DVM_LAUNCH_CHECK();
}

void sheet_groups(const SheetDev& sh, std::vector<double>& cen)
{
    const int nl = sh.nl;
    const int ng = (nl + kSheetGroup - 1) / kSheetGroup;
    cen.assign(4 * (size_t)ng, 0.0);
    for (int g = 0; g < ng; ++g) {
        const int j0 = g * kSheetGroup, j1 = std::min(j0 + kSheetGroup, nl);
        double cx = 0.0, cz = 0.0;
        for (int j = j0; j < j1; ++j) {
            cx += sh.hxc[j] + 0.5 * sh.hds[j] * std::cos(sh.htheta[j]);
            cz += sh.hzc[j] + 0.5 * sh.hds[j] * std::sin(sh.htheta[j]);
        }
        cx /= (j1 - j0);
        cz /= (j1 - j0);
        double r = 0.0;
        for (int j = j0; j < j1; ++j) {
            const double ax = sh.hxc[j] - cx, az = sh.hzc[j] - cz;
            const double bx = ax + sh.hds[j] * std::cos(sh.htheta[j]);
            const double bz = az + sh.hds[j] * std::sin(sh.htheta[j]);
            r = std::max(r, std::max(std::hypot(ax, az), std::hypot(bx, bz)));
        }
        // a degenerate or non-finite group never takes the series
        const bool ok = r > 0.0 && std::isfinite(r) && std::isfinite(cx) && std::isfinite(cz);
        cen[4 * (size_t)g] = ok ? cx : 0.0;
        cen[4 * (size_t)g + 1] = ok ? cz : 0.0;
        cen[4 * (size_t)g + 2] = ok ? r : INFINITY;
        cen[4 * (size_t)g + 3] = ok ? 1.0 / r : 0.0;
    }
}

BodyBox body_box(int npts, const double* px, const double* pz)
{
    for (int i = 0; i < npts; ++i)
        if (!std::isfinite(px[i]) || !std::isfinite(pz[i]))
            return BodyBox{-INFINITY, INFINITY, -INFINITY, INFINITY};  // no cull
    double x0 = px[0], x1 = px[0], z0 = pz[0], z1 = pz[0];
    for (int i = 1; i < npts; ++i) {
        x0 = px[i] < x0 ? px[i] : x0; x1 = px[i] > x1 ? px[i] : x1;
        z0 = pz[i] < z0 ? pz[i] : z0; z1 = pz[i] > z1 ? pz[i] : z1;
    }
    // a crossing abscissa x_i + vt (x_{i+1} - x_i) rounds to within a few
    // units of |x|; the margin is many orders of magnitude wider
    const double m = 1e-9 * (std::fabs(x0) + std::fabs(x1) + (x1 - x0)) + 1e-300;
    const bool closed = px[npts - 1] == px[0] && pz[npts - 1] == pz[0];
    BodyBox b;
    b.zlo = z0;
    b.zhi = z1;
    b.xhi = x1 + m;
    b.xlo = closed ? x0 - m : -INFINITY;
    return b;
}

int edge_slabs_build(int npts, const double* px, const double* pz, std::vector<int>& off,
                     std::vector<int>& edge, double& z0, double& inv_h)
{
    off.clear();
    edge.clear();
    const int nl = npts - 1;
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < npts; ++i) {
        if (!std::isfinite(px[i]) || !std::isfinite(pz[i])) return 0;
        lo = pz[i] < lo ? pz[i] : lo;
        hi = pz[i] > hi ? pz[i] : hi;
    }
    if (!(hi > lo) || nl < 1) return 0;
    const int n = nl;
    z0 = lo;
    inv_h = (double)n / (hi - lo);
    auto slab = [&](double z) {
        const double f = std::floor((z - z0) * inv_h);
        return f < 0.0 ? 0 : (f > (double)(n - 1) ? n - 1 : (int)f);
    };
    std::vector<int> cnt(n + 1, 0);
    auto span = [&](int i, int& a, int& b) {
        const double za = pz[i], zb = pz[i + 1];
        a = slab(za < zb ? za : zb) - 1;
        b = slab(za < zb ? zb : za) + 1;
        a = a < 0 ? 0 : a;
        b = b > n - 1 ? n - 1 : b;
    };
    for (int i = 0; i < nl; ++i) {
        int a, b;
        span(i, a, b);
        for (int q = a; q <= b; ++q) ++cnt[q + 1];
    }
    off.assign(n + 1, 0);
    for (int q = 0; q < n; ++q) off[q + 1] = off[q] + cnt[q + 1];
    edge.resize(off[n]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int i = 0; i < nl; ++i) {
        int a, b;
        span(i, a, b);
        for (int q = a; q <= b; ++q) edge[fill[q]++] = i;
    }
    return n;
}

void sheet_init(dvm_ctx_t ctx, SheetDev& sh, int npts, const double* px, const double* pz,
                const dvm_sheet_params* p, hipStream_t st, bool sync)
{
    const bool own_stream = st == nullptr || st == ctx->stream;
    if (!st) st = ctx->stream;
    require(npts >= 3, DVM_E_SIZE, "a body needs at least 3 points (2 panels)");
    require(npts <= 4097, DVM_E_SIZE, "at most 4096 panels per body");
    require(px && pz && p, DVM_E_INVALID, "null body or parameters");
    require(p->max_npv > 0 && p->dt > 0.0, DVM_E_INVALID, "max_NumPanelVort and dt must be > 0");
    const int nl = npts - 1;
    sh.nl = nl;
    sh.rho = p->rho; sh.nu = p->nu; sh.dt = p->dt; sh.ux = p->ux; sh.uz = p->uz;
    sh.cutoff_exp = p->cutoff_exp; sh.max_npv = p->max_npv;
    sh.hx.assign(px, px + npts);
    sh.hz.assign(pz, pz + npts);
    sh.box = body_box(npts, px, pz);
    double sz0 = 0.0, sinv = 0.0;
    const int nslab = edge_slabs_build(npts, px, pz, sh.hslab_off, sh.hslab_edge, sz0, sinv);
    // read_input_coord / form_vortex_sheet (src/VortexSheet.cpp:56-102): O(Nl)
    // set-up on the host with the C library's libm.
    double mn = px[0], mx = px[0];
    for (int i = 1; i < npts; ++i) { mn = px[i] < mn ? px[i] : mn; mx = px[i] > mx ? px[i] : mx; }
    sh.aref = mx - mn;
    sh.hxc.resize(nl); sh.hzc.resize(nl); sh.htheta.resize(nl); sh.hds.resize(nl);
    sh.henx.resize(nl); sh.henz.resize(nl); sh.hetx.resize(nl); sh.hetz.resize(nl);
    std::vector<PanelConst>& pc = sh.hpanel;
    pc.resize(nl);
    for (int i = 0; i < nl; ++i) {
        sh.hxc[i] = 0.5 * (px[i] + px[i + 1]);
        sh.hzc[i] = 0.5 * (pz[i] + pz[i + 1]);
        const double dx = px[i + 1] - px[i];
        const double dz = pz[i + 1] - pz[i];
        sh.hds[i] = std::sqrt(dx * dx + dz * dz);
        sh.htheta[i] = std::atan2(dz, dx);
        sh.henx[i] = dz / sh.hds[i];
        sh.henz[i] = -dx / sh.hds[i];
        sh.hetx[i] = -sh.henz[i];
        sh.hetz[i] = sh.henx[i];
        const double th = sh.htheta[i];
        pc[i] = PanelConst{sh.hxc[i], sh.hzc[i], sh.hds[i], std::cos(th), std::sin(th),
                           std::cos(-2.0 * th), std::sin(-2.0 * th),
                           std::pow(sh.hds[i], sh.cutoff_exp)};
    }
    // far-field moments of the panels as sheet_pair places them: segments
    // from the collocation point along theta, length ds (24-point
    // Gauss-Legendre, exact for the polynomial integrands up to degree 47)
    sh.mx = 0.0;
    sh.mz = 0.0;
    for (int i = 0; i < nl; ++i) {
        sh.mx += px[i];
        sh.mz += pz[i];
    }
    sh.mx /= nl;
    sh.mz /= nl;
    sh.mr = 0.0;
    for (int i = 0; i < nl; ++i) {
        const double ax = pc[i].xc - sh.mx, az = pc[i].zc - sh.mz;
        const double bx = ax + pc[i].ds * pc[i].cs, bz = az + pc[i].ds * pc[i].sn;
        sh.mr = std::max(sh.mr, std::max(std::hypot(ax, az), std::hypot(bx, bz)));
    }
    if (!(sh.mr > 0.0)) sh.mr = 1.0;
    const size_t d = sizeof(double);
    const size_t nred = 3 * (size_t)nl > 1024 ? 3 * (size_t)nl : 1024;
    const size_t nm = 2 * (size_t)nl * kSheetTerms;
    std::vector<double>& gcen = sh.hgcen;  // live until the caller's synchronisation
    sheet_groups(sh, gcen);
    const int ng = (int)(gcen.size() / 4);
    const size_t need = Carver::need({d * npts, d * npts, d * nl * 8, sizeof(PanelConst) * nl,
                                      d * (nl + 1) * nl, d * nl * (nl + 1), d * nl, d * (nl + 1),
                                      d * 2 * nl, d * 8, d * nred, sizeof(int) * nl,
                                      sizeof(int64_t) * (nl + 1), d * nl, sizeof(int) * 4, d * nm,
                                      d * nm, d * 2 * kSheetTerms,
                                      sizeof(int) * (sh.hslab_off.size() + 1),
                                      sizeof(int) * (sh.hslab_edge.size() + 1),
                                      sizeof(double4) * ng, d * nm, d * nm});
    DVM_HIP_CHECK(hipMalloc(&sh.mem, need));
    Carver cv(sh.mem);
    sh.x = cv.take<double>(npts);
    sh.z = cv.take<double>(npts);
    double* geo = cv.take<double>((size_t)nl * 8);
    sh.xc = geo; sh.zc = geo + nl; sh.theta = geo + 2 * nl; sh.ds = geo + 3 * nl;
    sh.enx = geo + 4 * nl; sh.enz = geo + 5 * nl; sh.etx = geo + 6 * nl; sh.etz = geo + 7 * nl;
    sh.panel = cv.take<PanelConst>(nl);
    sh.A = cv.take<double>((size_t)(nl + 1) * nl);
    sh.P = cv.take<double>((size_t)nl * (nl + 1));
    sh.gamma = cv.take<double>(nl);
    sh.b = cv.take<double>(nl + 1);
    sh.uw = cv.take<double>(2 * nl);
    sh.scal = cv.take<double>(8);
    sh.red = cv.take<double>(nred);
    sh.cnt = cv.take<int>(nl);
    sh.off = cv.take<int64_t>(nl + 1);
    sh.cnew = cv.take<double>(nl);
    sh.flag = cv.take<int>(4);
    sh.mI0 = cv.take<double>(nm);
    sh.mI1 = cv.take<double>(nm);
    sh.mA = cv.take<double>(2 * kSheetTerms);
    int* slab_off = cv.take<int>(sh.hslab_off.size() + 1);
    int* slab_edge = cv.take<int>(sh.hslab_edge.size() + 1);
    sh.ng = ng;
    sh.gcen = cv.take<double4>(ng);
    sh.gI0 = cv.take<double>(nm);
    sh.gI1 = cv.take<double>(nm);
    if (nslab > 0) {
        DVM_HIP_CHECK(hipMemcpyAsync(slab_off, sh.hslab_off.data(),
                                     sizeof(int) * sh.hslab_off.size(), hipMemcpyHostToDevice, st));
        DVM_HIP_CHECK(hipMemcpyAsync(slab_edge, sh.hslab_edge.data(),
                                     sizeof(int) * sh.hslab_edge.size(), hipMemcpyHostToDevice,
                                     st));
        sh.slabs.off = slab_off;
        sh.slabs.edge = slab_edge;
        sh.slabs.n = nslab;
        sh.slabs.z0 = sz0;
        sh.slabs.inv_h = sinv;
    } else {
        sh.slabs = EdgeSlabs{};
    }

    std::vector<double>& g = sh.hstage;
    g.resize((size_t)nl * 8);
    std::memcpy(&g[0], sh.hxc.data(), d * nl);
    std::memcpy(&g[nl], sh.hzc.data(), d * nl);
    std::memcpy(&g[2 * nl], sh.htheta.data(), d * nl);
    std::memcpy(&g[3 * nl], sh.hds.data(), d * nl);
    std::memcpy(&g[4 * nl], sh.henx.data(), d * nl);
    std::memcpy(&g[5 * nl], sh.henz.data(), d * nl);
    std::memcpy(&g[6 * nl], sh.hetx.data(), d * nl);
    std::memcpy(&g[7 * nl], sh.hetz.data(), d * nl);
    DVM_HIP_CHECK(hipMemcpyAsync(sh.x, sh.hx.data(), d * npts, hipMemcpyHostToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(sh.z, sh.hz.data(), d * npts, hipMemcpyHostToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(geo, g.data(), d * nl * 8, hipMemcpyHostToDevice, st));
    DVM_HIP_CHECK(hipMemcpyAsync(sh.panel, pc.data(), sizeof(PanelConst) * nl,
                                 hipMemcpyHostToDevice, st));
    panel_moments(st, nl, sh.panel, sh.mx, sh.mz, sh.mr, nullptr, sh.mI0, sh.mI1);
    DVM_HIP_CHECK(hipMemcpyAsync(sh.gcen, gcen.data(), d * gcen.size(), hipMemcpyHostToDevice, st));
    panel_moments(st, nl, sh.panel, 0.0, 0.0, 1.0, sh.gcen, sh.gI0, sh.gI1);
    DVM_HIP_CHECK(hipMemsetAsync(sh.gamma, 0, d * nl, st));
    DVM_HIP_CHECK(hipMemsetAsync(sh.scal, 0, d * 8, st));

    // compute_influence_matrix, then the least-squares operator (A^T A)^-1 A^T.
    const int W = 2 * nl + 1;
    const int kb = gj_pass_width(nl);
    double* G = nullptr;
    DVM_HIP_CHECK(hipMallocAsync((void**)&G, d * (size_t)(nl + kb) * W + sizeof(int) * nl, st));
    double* U = G + (size_t)nl * W;
    int* pivrow = reinterpret_cast<int*>(U + (size_t)kb * W);
    {
        ProfScope ps(own_stream ? ctx : nullptr, "panel_influence");
        influence_kernel<<<dim3(cdiv(nl, 128), nl + 1), 128, 0, st>>>(nl, sh.x, sh.z, sh.xc, sh.zc,
                                                                      sh.theta, sh.ds, sh.A);
        DVM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(own_stream ? ctx : nullptr, "sheet_operator");
        gj_build_kernel<<<dim3(cdiv(W, 128), nl), 128, 0, st>>>(nl, sh.A, G);
        DVM_LAUNCH_CHECK();
        switch (kb) {
            case 32: gj_passes<32>(st, nl, G, pivrow, U); break;
            case 16: gj_passes<16>(st, nl, G, pivrow, U); break;
            case 8: gj_passes<8>(st, nl, G, pivrow, U); break;
            case 4: gj_passes<4>(st, nl, G, pivrow, U); break;
            default: gj_passes<2>(st, nl, G, pivrow, U); break;
        }
        gj_extract_kernel<<<dim3(cdiv(nl + 1, 128), nl), 128, 0, st>>>(nl, G, pivrow, sh.P);
        DVM_LAUNCH_CHECK();
    }
    DVM_HIP_CHECK(hipFreeAsync(G, st));
    if (sync) DVM_HIP_CHECK(hipStreamSynchronize(st));
}

// ----------------------------------------------------------------- reductions

__global__ void sum_partial_kernel(int64_t n, const double* __restrict__ v, double* __restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sum_final_kernel(int np, const double* __restrict__ part, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += part[i];
        *out = s;
    }
}

void device_sum(dvm_ctx_t ctx, int64_t n, const double* v, double* partials, double* out)
{
    const int nb = 256;
    sum_partial_kernel<<<nb, 256, 0, ctx->stream>>>(n, v, partials);
    DVM_LAUNCH_CHECK();
    sum_final_kernel<<<1, 64, 0, ctx->stream>>>(nb, partials, out);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- solve

// b_i = (U + u_i).n_i, b_Nl = -sum circ (src/VortexSheet.cpp:173-175, 209-210);
// gamma = P b with one wave per row, lanes over columns.
__global__ __launch_bounds__(256) void gamma_kernel(int nl, const double* __restrict__ P,
                                                    const double* __restrict__ enx,
                                                    const double* __restrict__ enz, double ux,
                                                    double uz, const double* __restrict__ uw,
                                                    int have_blobs, const double* __restrict__ scal,
                                                    double* __restrict__ b,
                                                    double* __restrict__ gamma)
{
    extern __shared__ double sb[];
    for (int i = threadIdx.x; i <= nl; i += 256) {
        double v;
        if (i == nl) {
            v = have_blobs ? -scal[0] : -0.0;
        } else if (have_blobs) {
            v = (ux + uw[i]) * enx[i] + (uz + uw[nl + i]) * enz[i];
        } else {
            v = ux * enx[i] + uz * enz[i];
        }
        sb[i] = v;
        if (blockIdx.x == 0) b[i] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ncol = nl + 1;
    for (int i = blockIdx.x * 4 + wave; i < nl; i += gridDim.x * 4) {
        double s = 0.0;
        for (int c = lane; c < ncol; c += 64) s += P[(int64_t)i * ncol + c] * sb[c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) gamma[i] = s;
    }
}

void sheet_solve(dvm_ctx_t ctx, SheetDev& sh, int prec, int64_t n, const double* bx,
                 const double* bz, const double* bc, const double* bs, bool have_uw)
{
    const int nl = sh.nl;
    if (n > 0) {
        // The collocation sums stay FP64 in every form: the loads take
        // dp = -rho/dt gamma ds, so an FP32 rounding of the right-hand side
        // reaches Cd, Cl amplified (about 4e-6 at 96 panels, dt 0.005), and
        // Nl targets are a small share of a step's pairs.
        if (!have_uw)
            bs_targets_launch(ctx, DVM_F64, n, bx, bz, bc, bs, nl, sh.xc, sh.zc, 0, 0.0, sh.uw,
                              sh.uw + nl);
        device_sum(ctx, n, bc, sh.red, sh.scal);
    }
    const unsigned blocks = cdiv(nl, 4) < 64 ? cdiv(nl, 4) : 64;
    ProfScope ps(ctx, "sheet_gamma");
    gamma_kernel<<<blocks, 256, sizeof(double) * (nl + 1), ctx->stream>>>(
        nl, sh.P, sh.enx, sh.enz, sh.ux, sh.uz, sh.uw, n > 0 ? 1 : 0, sh.scal, sh.b, sh.gamma);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- sheet on blobs

// src/VortexSheet.cpp:225-272 fused: per (blob, panel) the coefficients
// px, qx, py, qy with the collocation midpoint as panel origin, summed against
// gamma_j and gamma_{j+1} (cyclic).
__global__ __launch_bounds__(256) void sheet_on_blobs_kernel(
    int64_t n, int nl, const double* __restrict__ bx, const double* __restrict__ bz,
    const PanelConst* __restrict__ pc, const double* __restrict__ gamma, double* __restrict__ uvs,
    double* __restrict__ wvs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = bx[i];
    const double zi = bz[i];
    double su = 0.0, sw = 0.0;
    for (int j = 0; j < nl; ++j)
        sheet_pair(xi, zi, pc[j], gamma[j], gamma[j + 1 == nl ? 0 : j + 1], su, sw);
    uvs[i] = su * kRpi2;
    wvs[i] = sw * kRpi2;
}

// Moments of the current gamma: mA_k = sum_j -(gamma_j I0_jk + gamma_{j+1} I1_jk)
// (gamma(s) = gamma_j (1 - s/ds) + gamma_{j+1} s/ds on panel j, cyclic): one
// workgroup per moment, lane t summing panels t, t+256, ... and the lanes'
// sums added by a fixed tree.  Block 0 also clears the near-blob counter of
// sheet_on_blobs_far_kernel.
__global__ __launch_bounds__(256) void sheet_moments_kernel(int nl, const double* __restrict__ gamma,
                                                            const double* __restrict__ i0,
                                                            const double* __restrict__ i1,
                                                            double* __restrict__ ma,
                                                            int* __restrict__ near_count)
{
    __shared__ double rr[256], ri[256];
    const int k = blockIdx.x, t = threadIdx.x;
    if (k == 0 && t == 0) *near_count = 0;
    double ar = 0.0, ai = 0.0;
    for (int j = t; j < nl; j += 256) {
        const double g0 = gamma[j], g1 = gamma[j + 1 == nl ? 0 : j + 1];
        const size_t o = 2 * ((size_t)j * kSheetTerms + k);
        ar -= g0 * i0[o] + g1 * i1[o];
        ai -= g0 * i0[o + 1] + g1 * i1[o + 1];
    }
    rr[t] = ar;
    ri[t] = ai;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            rr[t] += rr[t + h];
            ri[t] += ri[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        ma[2 * k] = rr[0];
        ma[2 * k + 1] = ri[0];
    }
}

// Fast form: outside kSheetFar * mr of the body centre the Laurent series
// phi = (1/mr) sum_k A_k / W^(k+1), W = (z - c) / mr, gives uvs = Im(phi) / 2pi,
// wvs = Re(phi) / 2pi (truncation about kSheetFar^-kSheetTerms = 1e-12 of the
// leading term).  Blobs inside it are listed (near_list, near_count) for
// sheet_on_blobs_near_kernel.
__global__ __launch_bounds__(256) void sheet_on_blobs_far_kernel(
    int64_t n, const double* __restrict__ bx, const double* __restrict__ bz,
    const double* __restrict__ ma, double cx, double cz, double rr, double* __restrict__ uvs,
    double* __restrict__ wvs, int* __restrict__ near_list, int* __restrict__ near_count)
{
    __shared__ double a[2 * kSheetTerms];
    for (int k = threadIdx.x; k < 2 * kSheetTerms; k += blockDim.x) a[k] = ma[k];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double xi = i < n ? bx[i] : 0.0;
    const double zi = i < n ? bz[i] : 0.0;
    const double wr = (xi - cx) / rr, wi = (zi - cz) / rr;
    const double w2 = wr * wr + wi * wi;
    // near blobs: one counter update per wavefront
    const bool is_near = i < n && !(w2 > kSheetFar * kSheetFar);
    const uint64_t m = __ballot(is_near);
    if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(near_count, __popcll(m));
        base = __shfl(base, leader);
        if (is_near) near_list[base + __popcll(m & ((uint64_t(1) << lane) - 1))] = (int)i;
    }
    if (i >= n || is_near) return;
    {
        // t = 1 / W; phi * mr = t (A_0 + t (A_1 + ...))
        const double tr = wr / w2, ti = -wi / w2;
        const int K = sheet_terms(w2);
        double pr = a[2 * (K - 1)], pim = a[2 * (K - 1) + 1];
        for (int k = K - 2; k >= 0; --k) {
            const double nr = pr * tr - pim * ti + a[2 * k];
            pim = pr * ti + pim * tr + a[2 * k + 1];
            pr = nr;
        }
        const double fr = (pr * tr - pim * ti) / rr, fi = (pr * ti + pim * tr) / rr;
        uvs[i] = fi * kRpi2;
        wvs[i] = fr * kRpi2;
    }
}

// The panel sum (sheet_pair) for the listed near blobs, one wavefront per
// blob: lane l sums panels l, l+64, ..., and the lanes' sums are added by
// shuffles.  A fixed set of wavefronts walks the list (its length is read on
// the device).
__global__ __launch_bounds__(256) void sheet_on_blobs_near_kernel(
    int nl, const double* __restrict__ bx, const double* __restrict__ bz,
    const PanelConst* __restrict__ pc, const double* __restrict__ gamma,
    const int* __restrict__ near_list, const int* __restrict__ near_count,
    double* __restrict__ uvs, double* __restrict__ wvs)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    const int cnt = *near_count;
    for (int q = wave; q < cnt; q += nwaves) {
        const int i = near_list[q];
        const double xi = bx[i], zi = bz[i];
        double su = 0.0, sw = 0.0;
        for (int j = lane; j < nl; j += 64)
            sheet_pair(xi, zi, pc[j], gamma[j], gamma[j + 1 == nl ? 0 : j + 1], su, sw);
        for (int o = 32; o > 0; o >>= 1) {
            su += __shfl_xor(su, o);
            sw += __shfl_xor(sw, o);
        }
        if (lane == 0) {
            uvs[i] = su * kRpi2;
            wvs[i] = sw * kRpi2;
        }
    }
}

void sheet_on_blobs(dvm_ctx_t ctx, const SheetDev& sh, int prec, int64_t n, const double* bx,
                    const double* bz, double* uvs, double* wvs)
{
    if (n <= 0) return;
    ProfScope ps(ctx, "sheet_on_blobs");
    if (prec == DVM_F64) {
        sheet_on_blobs_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, sh.nl, bx, bz, sh.panel,
                                                                     sh.gamma, uvs, wvs);
        DVM_LAUNCH_CHECK();
        return;
    }
    require(n < (int64_t(1) << 31), DVM_E_SIZE, "too many blobs");
    int* near_list = static_cast<int*>(ws_get(ctx, sizeof(int) * (size_t)n));
    int* near_count = sh.flag + 1;
    sheet_moments_kernel<<<kSheetTerms, 256, 0, ctx->stream>>>(sh.nl, sh.gamma, sh.mI0, sh.mI1,
                                                                sh.mA, near_count);
    DVM_LAUNCH_CHECK();
    sheet_on_blobs_far_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(
        n, bx, bz, sh.mA, sh.mx, sh.mz, sh.mr, uvs, wvs, near_list, near_count);
    DVM_LAUNCH_CHECK();
    sheet_on_blobs_near_kernel<<<ctx->num_cus * 8, 256, 0, ctx->stream>>>(
        sh.nl, bx, bz, sh.panel, sh.gamma, near_list, near_count, uvs, wvs);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- probes (quirk form)

// Sheet part of VortexSheet::probe_velocities (src/VortexSheet.cpp:597-638):
// the panel loop runs j = 1 .. Nl-1 and assigns (does not accumulate), so
// only panel Nl-1 survives, paired with gamma_0 and with the qy term
// subtracted (:630-634); its coefficients differ from vortexsheetbc in c9,
// qy and py (:612-626) and carry no 1/(2 pi).  u, w arrive holding the blob
// part already scaled by 1/(2 pi); the result is (u + uvs) + Ux (:641-644).
__global__ void probe_quirk_kernel(int64_t np, const double* __restrict__ px,
                                   const double* __restrict__ pz, int nl,
                                   const PanelConst* __restrict__ panel,
                                   const double* __restrict__ gamma, double ux, double uz,
                                   double* __restrict__ u, double* __restrict__ w,
                                   double* __restrict__ uvs, double* __restrict__ wvs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double su = 0.0, sw = 0.0;
    if (nl >= 2) {
        const PanelConst q = panel[nl - 1];
        const double dx = px[i] - q.xc;
        const double dz = pz[i] - q.zc;
        const double dsj = q.ds;
        const double c1 = -dx * q.cs - dz * q.sn;
        const double c2 = dx * dx + dz * dz;
        const double c5 = dx * q.sn - dz * q.cs;
        const double c6 = log(1.0 + dsj * ((dsj + 2 * c1) / c2));
        const double c7 = atan2(c5 * dsj, c2 + c1 * dsj);
        const double c8 = dx * q.s2 + dz * q.c2;
        const double c9 = dx * q.c2 + dz * q.s2;
        const double qx = -q.sn + 0.5 * c8 * c6 / dsj + (c1 * q.cs + c5 * q.sn) * c7 / dsj;
        const double pxj = -0.5 * c6 * q.sn - c7 * q.cs - qx;
        const double qy = q.cs + 0.5 * c9 * c6 / dsj + (c1 * q.cs - c5 * q.sn) * c7 / dsj;
        const double pyj = -0.5 * c6 * q.cs - c7 * q.sn - qy;
        su = pxj * gamma[nl - 1] + qx * gamma[0];
        sw = pyj * gamma[nl - 1] - qy * gamma[0];
    }
    uvs[i] = su;
    wvs[i] = sw;
    u[i] = u[i] + su + ux;
    w[i] = w[i] + sw + uz;
}

void sheet_probe_quirk(dvm_ctx_t ctx, const SheetDev& sh, int prec, int64_t n, const double* bx,
                       const double* bz, const double* bc, const double* bs, int64_t np,
                       const double* px, const double* pz, double* u, double* w, double* uvs,
                       double* wvs)
{
    if (np <= 0) return;
    // Blob part: blobs 1 .. n-1 (the reference loop starts at j = 1,
    // src/VortexSheet.cpp:573), the A1 pair with no self or r^2 guard.
    if (n >= 2) {
        bs_targets_launch(ctx, prec, n - 1, bx + 1, bz + 1, bc + 1, bs + 1, np, px, pz, 0, 0.0,
                          u, w);
    } else {
        DVM_HIP_CHECK(hipMemsetAsync(u, 0, sizeof(double) * np, ctx->stream));
        DVM_HIP_CHECK(hipMemsetAsync(w, 0, sizeof(double) * np, ctx->stream));
    }
    probe_quirk_kernel<<<cdiv(np, 256), 256, 0, ctx->stream>>>(np, px, pz, sh.nl, sh.panel,
                                                              sh.gamma, sh.ux, sh.uz, u, w,
                                                              uvs, wvs);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- loads

// src/VortexSheet.cpp:487-514 (loads_block; one workgroup).  work holds
// max(1024, 3 nl) doubles.
__global__ __launch_bounds__(kLoadsThreads) void loads_kernel(
    int nl, double rho, double dt, double ur, double aref, const double* __restrict__ gamma,
    const double* __restrict__ ds, const double* __restrict__ enx, const double* __restrict__ enz,
    double* __restrict__ work, double* __restrict__ out)
{
    loads_block(nl, rho, dt, ur, aref, gamma, ds, enx, enz, work, out);
}

void sheet_loads(dvm_ctx_t ctx, SheetDev& sh, double ur)
{
    // sh.red holds max(1024, 3 nl) doubles: the loads work vectors
    loads_kernel<<<1, kLoadsThreads, 0, ctx->stream>>>(sh.nl, sh.rho, sh.dt, ur, sh.aref,
                                                       sh.gamma, sh.ds, sh.enx, sh.enz, sh.red,
                                                       sh.scal + 1);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- release

// src/VortexSheet.cpp:282-303: counts, NaN check, circulation per nascent blob
// and the exclusive prefix of counts (one workgroup).
__global__ __launch_bounds__(1024) void release_count_kernel(
    int nl, int max_npv, const double* __restrict__ gamma, const double* __restrict__ ds,
    int* __restrict__ cnt, int64_t* __restrict__ off, double* __restrict__ cnew,
    int* __restrict__ flag)
{
    __shared__ double red[1024];
    __shared__ int64_t pre[1024];
    const int t = threadIdx.x;
    // arma::max: running maximum that only moves on a strictly greater value
    // (NaN entries never win)
    double mx = -INFINITY;
    for (int i = t; i < nl; i += 1024) {
        const double v = gamma[i] * ds[i];
        if (v > mx) mx = v;
    }
    red[t] = mx;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            const double o = red[t + s];
            if (o > red[t]) red[t] = o;
        }
        __syncthreads();
    }
    const double m = red[0];
    __syncthreads();
    if (t == 0) flag[0] = 0;
    __syncthreads();
    // each thread owns a contiguous run of panels for the prefix
    const int per = (nl + 1023) / 1024;
    const int i0 = t * per;
    const int i1 = i0 + per < nl ? i0 + per : nl;
    int64_t local = 0;
    int bad = 0;
    for (int i = i0; i < i1; ++i) {
        const double pcirc = gamma[i] * ds[i];
        const double c = floor(fabs(pcirc / m * (double)max_npv)) + 1.0;
        int ci = 0;
        if (!(c > 1.0e-15)) {
            bad |= 1;  // NaN count: the reference throws VortexException
        } else if (!(c < 1.0e9)) {
            bad |= 2;  // count too large to emit
        } else {
            ci = (int)c;
        }
        cnt[i] = ci;
        cnew[i] = pcirc / c;
        local += ci;
    }
    if (bad) atomicOr(flag, bad);
    pre[t] = local;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {
        const int64_t v = t >= s ? pre[t - s] : 0;
        __syncthreads();
        pre[t] += v;
        __syncthreads();
    }
    int64_t run = pre[t] - local;
    for (int i = i0; i < i1; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (t == 1023) off[nl] = pre[1023];
}

void sheet_release_count(dvm_ctx_t ctx, SheetDev& sh)
{
    release_count_kernel<<<1, 1024, 0, ctx->stream>>>(sh.nl, sh.max_npv, sh.gamma, sh.ds, sh.cnt,
                                                      sh.off, sh.cnew, sh.flag);
    DVM_LAUNCH_CHECK();
}

// src/VortexSheet.cpp:305-331: nascent blob k of panel i sits on the outward
// normal through xm = x_i + ds_i et_i / (n_i + 1), at distance rrw[k].
__global__ void release_emit_kernel(int64_t nrv, int nl, const int64_t* __restrict__ off,
                                    const int* __restrict__ cnt, const double* __restrict__ bxp,
                                    const double* __restrict__ bzp, const double* __restrict__ ds,
                                    const double* __restrict__ etx, const double* __restrict__ etz,
                                    const double* __restrict__ enx, const double* __restrict__ enz,
                                    const double* __restrict__ cnew,
                                    const PanelConst* __restrict__ pc,
                                    const double* __restrict__ rrw, uint32_t nold, int64_t base,
                                    double* __restrict__ x, double* __restrict__ z,
                                    double* __restrict__ c, double* __restrict__ s,
                                    double* __restrict__ u, double* __restrict__ w,
                                    double* __restrict__ uvs, double* __restrict__ wvs,
                                    uint32_t* __restrict__ id)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrv) return;
    int lo = 0, hi = nl;  // find i with off[i] <= k < off[i+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= k) lo = mid; else hi = mid;
    }
    const int i = lo;
    const double np1 = (double)cnt[i] + 1;
    const double xm = bxp[i] + ds[i] * etx[i] / np1;
    const double zm = bzp[i] + ds[i] * etz[i] / np1;
    const double r = rrw[k];
    const int64_t o = base + k;
    x[o] = xm + r * enx[i];
    z[o] = zm + r * enz[i];
    c[o] = cnew[i];
    s[o] = pc[i].sig;
    u[o] = 0.0; w[o] = 0.0; uvs[o] = 0.0; wvs[o] = 0.0;
    id[o] = nold;
}

void sheet_release_emit(dvm_ctx_t ctx, const SheetDev& sh, int64_t nrv, const double* rrw,
                        uint32_t nold, int64_t base, double* x, double* z, double* c, double* s,
                        double* u, double* w, double* uvs, double* wvs, uint32_t* id)
{
    if (nrv <= 0) return;
    release_emit_kernel<<<cdiv(nrv, 256), 256, 0, ctx->stream>>>(
        nrv, sh.nl, sh.off, sh.cnt, sh.x, sh.z, sh.ds, sh.etx, sh.etz, sh.enx, sh.enz, sh.cnew,
        sh.panel, rrw, nold, base, x, z, c, s, u, w, uvs, wvs, id);
    DVM_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- crossing

__global__ __launch_bounds__(256) void inside_kernel(int64_t n, int nl, const double* __restrict__ bxp,
                                                     const double* __restrict__ bzp, BodyBox bb,
                                                     EdgeSlabs es, const double* __restrict__ x,
                                                     const double* __restrict__ z,
                                                     int* __restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = inside_body_fast(nl, bxp, bzp, bb, es, x[i], z[i]) ? 1 : 0;
}

void sheet_inside_flags(dvm_ctx_t ctx, const SheetDev& sh, int64_t n, const double* x,
                        const double* z, int* flags)
{
    if (n <= 0) return;
    ProfScope ps(ctx, "crossing_flags");
    inside_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, sh.nl, sh.x, sh.z, sh.box, sh.slabs,
                                                         x, z, flags);
    DVM_LAUNCH_CHECK();
}

// src/VortexSheet.cpp:345-405 for the flagged (inside) blobs: each
// wavefront takes the flagged blobs among its 64 one at a time, all lanes
// working on one (reflect_wave).  No lane leaves early: every lane takes
// part in the shuffles.
__global__ __launch_bounds__(256) void reflect_kernel(int64_t n, int nl,
                                                      const double* __restrict__ bxp,
                                                      const double* __restrict__ bzp,
                                                      const double* __restrict__ enx,
                                                      const double* __restrict__ enz,
                                                      const int* __restrict__ flags,
                                                      double* __restrict__ x,
                                                      double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = i < n && flags[i] != 0;
    const double xi = mine ? x[i] : 0.0, zi = mine ? z[i] : 0.0;
    uint64_t mask = __ballot(mine);
    while (mask) {
        const int l = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        const double px = __shfl(xi, l), pz = __shfl(zi, l);
        const double2 p = reflect_wave(nl, bxp, bzp, enx, enz, px, pz);
        if ((int)(threadIdx.x & 63) == l) {
            x[i] = p.x;
            z[i] = p.y;
        }
    }
}

void sheet_reflect(dvm_ctx_t ctx, const SheetDev& sh, int64_t n, const int* flags, double* x,
                   double* z)
{
    if (n <= 0) return;
    reflect_kernel<<<cdiv(n, 256), 256, 0, ctx->stream>>>(n, sh.nl, sh.x, sh.z, sh.enx, sh.enz,
                                                          flags, x, z);
    DVM_LAUNCH_CHECK();
}

}  // namespace dvmi

using namespace dvmi;

extern "C" {

dvm_status dvm_sheet_create(dvm_ctx_t ctx, int npts, const double* px, const double* pz,
                            const dvm_sheet_params* params, dvm_sheet_t* out)
{
    if (!ctx || !out) return DVM_E_INVALID;
    *out = nullptr;
    dvm_sheet_t s = new (std::nothrow) dvm_sheet_s();
    if (!s) return DVM_E_OOM;
    s->ctx = ctx;
    dvm_status st = guard(ctx, [&] { sheet_init(ctx, s->sh, npts, px, pz, params); });
    if (st != DVM_OK) {
        sheet_free(s->sh);
        delete s;
        return st;
    }
    *out = s;
    return DVM_OK;
}

dvm_status dvm_sheet_destroy(dvm_sheet_t s)
{
    if (!s) return DVM_OK;
    dvm_status st = guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        sheet_free(s->sh);
    });
    delete s;
    return st;
}

dvm_status dvm_sheet_size(dvm_sheet_t s, int* nl)
{
    if (!s || !nl) return DVM_E_INVALID;
    *nl = s->sh.nl;
    return DVM_OK;
}

dvm_status dvm_sheet_geometry(dvm_sheet_t s, double* xc, double* zc, double* theta, double* ds,
                              double* enx, double* enz, double* etx, double* etz, double* aref)
{
    if (!s) return DVM_E_INVALID;
    const SheetDev& h = s->sh;
    const size_t b = sizeof(double) * h.nl;
    if (xc) std::memcpy(xc, h.hxc.data(), b);
    if (zc) std::memcpy(zc, h.hzc.data(), b);
    if (theta) std::memcpy(theta, h.htheta.data(), b);
    if (ds) std::memcpy(ds, h.hds.data(), b);
    if (enx) std::memcpy(enx, h.henx.data(), b);
    if (enz) std::memcpy(enz, h.henz.data(), b);
    if (etx) std::memcpy(etx, h.hetx.data(), b);
    if (etz) std::memcpy(etz, h.hetz.data(), b);
    if (aref) *aref = h.aref;
    return DVM_OK;
}

dvm_status dvm_sheet_influence(dvm_sheet_t s, double* A)
{
    if (!s || !A) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(A, s->sh.A, sizeof(double) * (s->sh.nl + 1) * s->sh.nl,
                                     hipMemcpyDeviceToHost, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_sheet_gamma_ptr(dvm_sheet_t s, double** g)
{
    if (!s || !g) return DVM_E_INVALID;
    *g = s->sh.gamma;
    return DVM_OK;
}

dvm_status dvm_sheet_get_gamma(dvm_sheet_t s, double* g)
{
    if (!s || !g) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(g, s->sh.gamma, sizeof(double) * s->sh.nl,
                                     hipMemcpyDeviceToHost, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_sheet_set_gamma(dvm_sheet_t s, const double* g)
{
    if (!s || !g) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        DVM_HIP_CHECK(hipMemcpyAsync(s->sh.gamma, g, sizeof(double) * s->sh.nl,
                                     hipMemcpyHostToDevice, s->ctx->stream));
        DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}

dvm_status dvm_sheet_solve(dvm_sheet_t s, int prec, int64_t n, const double* x, const double* z,
                           const double* circ, const double* sigma, double* b_host)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(prec == DVM_F64 || prec == DVM_F32C || prec == DVM_FMM, DVM_E_INVALID,
                "unknown precision");
        require(n >= 0, DVM_E_SIZE, "negative blob count");
        require(n == 0 || (x && z && circ && sigma), DVM_E_INVALID, "null blob array");
        sheet_solve(s->ctx, s->sh, prec, n, x, z, circ, sigma);
        if (b_host) {
            DVM_HIP_CHECK(hipMemcpyAsync(b_host, s->sh.b, sizeof(double) * (s->sh.nl + 1),
                                         hipMemcpyDeviceToHost, s->ctx->stream));
            DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        }
    });
}

dvm_status dvm_sheet_on_blobs(dvm_sheet_t s, int64_t n, const double* x, const double* z,
                              double* uvs, double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative blob count");
        require(n == 0 || (x && z && uvs && wvs), DVM_E_INVALID, "null array");
        sheet_on_blobs(s->ctx, s->sh, DVM_F64, n, x, z, uvs, wvs);
    });
}

dvm_status dvm_sheet_on_blobs_far(dvm_sheet_t s, int64_t n, const double* x, const double* z,
                                  double* uvs, double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative blob count");
        require(n == 0 || (x && z && uvs && wvs), DVM_E_INVALID, "null array");
        sheet_on_blobs(s->ctx, s->sh, DVM_F32C, n, x, z, uvs, wvs);
    });
}

dvm_status dvm_sheet_loads(dvm_sheet_t s, double ur, double* out4)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        sheet_loads(s->ctx, s->sh, ur);
        if (out4) {
            DVM_HIP_CHECK(hipMemcpyAsync(out4, s->sh.scal + 1, sizeof(double) * 4,
                                         hipMemcpyDeviceToHost, s->ctx->stream));
            DVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        }
    });
}

dvm_status dvm_sheet_probe_quirk(dvm_sheet_t s, int prec, int64_t n, const double* x,
                                 const double* z, const double* circ, const double* sigma,
                                 int64_t np, const double* px, const double* pz, double* u,
                                 double* w, double* uvs, double* wvs)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(prec == DVM_F64 || prec == DVM_F32C || prec == DVM_FMM, DVM_E_INVALID,
                "unknown precision");
        require(n >= 0 && np >= 0, DVM_E_SIZE, "negative count");
        require(n < 2 || (x && z && circ && sigma), DVM_E_INVALID, "null blob array");
        require(np == 0 || (px && pz && u && w && uvs && wvs), DVM_E_INVALID,
                "null probe array");
        sheet_probe_quirk(s->ctx, s->sh, prec, n, x, z, circ, sigma, np, px, pz, u, w, uvs, wvs);
    });
}

dvm_status dvm_sheet_inside(dvm_sheet_t s, int64_t n, const double* x, const double* z, int* flags)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        require(n == 0 || (x && z && flags), DVM_E_INVALID, "null array");
        sheet_inside_flags(s->ctx, s->sh, n, x, z, flags);
    });
}

dvm_status dvm_sheet_reflect(dvm_sheet_t s, int64_t n, double* x, double* z)
{
    if (!s) return DVM_E_INVALID;
    return guard(s->ctx, [&] {
        require(n >= 0, DVM_E_SIZE, "negative count");
        if (n == 0) return;
        require(x && z, DVM_E_INVALID, "null array");
        int* flags = static_cast<int*>(ws_get(s->ctx, sizeof(int) * n));
        sheet_inside_flags(s->ctx, s->sh, n, x, z, flags);
        sheet_reflect(s->ctx, s->sh, n, flags, x, z);
    });
}

}  // extern "C"
