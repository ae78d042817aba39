// Test infrastructure only: the oracle (see dvm_oracle.h).
//
// CPU restatement of the DVMpp reference time-step path, in the reference's
// loop and operation order, FP64, with the reference's four OpenMP loops kept
// parallel and every other loop serial.  Compiled with -ffp-contract=off so
// that each expression rounds exactly as written (the reference build's
// `-O3 -march=native` contraction choices are not reproducible without the
// reference binary; see the pinning notes in tests/test_oracle_*.py).
//
// Citations are to the reference tree (src/..., tests/...).
#include "dvm_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

namespace orc {

// src/BaseTypes.hpp:18-21
constexpr double pi = 3.14159265358979323846;
constexpr double rpi2 = 1.0 / (2.0 * 3.14159265358979323846);

using Vec = std::vector<double>;

// Armadillo's accumulate for dense vectors (arma::sum / arma::accu): two
// interleaved partial sums, odd tail into the first, then acc1 + acc2.
// Armadillo is unpinned and absent here, so this order is parity-unpinned;
// every consumer of it is compared to a tolerance, never bit-for-bit.
static double arma_sum(const double* a, size_t n)
{
    double acc1 = 0.0, acc2 = 0.0;
    size_t j;
    for (j = 1; j < n; j += 2) {
        acc1 += a[j - 1];
        acc2 += a[j];
    }
    if ((j - 1) < n) acc1 += a[j - 1];
    return acc1 + acc2;
}

// ---------------------------------------------------------------- Random
// src/Random.hpp:10-27, src/Random.cpp:3-18: mt19937 + uniform_real_distribution<double>
struct Rng {
    std::uniform_real_distribution<> dist;
    std::mt19937 re;
    uint64_t draws = 0;
    Rng() { std::random_device rd; re.seed(rd()); }
    double rand() { ++draws; return dist(re); }
    void seed(int s) { if (s >= 0) re.seed(s); }
};

// The library's device random stream (not the reference's): Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11) with the published multipliers and Weyl increments, 10 rounds.
void philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53 random bits from two words, centred in their interval: (m + 0.5) 2^-53.
double u01(uint32_t a, uint32_t b)
{
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Draws of the device stream: index i of step `step`, stream tag 0 (diffusion:
// two uniforms per blob) or 1 (release: one per nascent blob).
struct DeviceDraw {
    uint32_t seed;
    uint32_t step;
    void two(uint64_t i, uint32_t tag, double& r1, double& r2) const
    {
        const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), step, tag};
        const uint32_t key[2] = {seed, tag == 0 ? 0x5eed0001u : 0x5eed0002u};
        uint32_t o[4];
        philox(ctr, key, o);
        r1 = u01(o[0], o[1]);
        r2 = u01(o[2], o[3]);
    }
};

// ---------------------------------------------------------------- Blobs
// src/VortexBlobs.hpp:24-32 (public SoA fields)
struct Blobs {
    std::vector<uint32_t> id;
    Vec x, z, circ, sigma, u, w, uvs, wvs;
    size_t size() const { return id.size(); }
    void resize(size_t n)
    {
        id.assign(n, 0u);
        for (Vec* v : {&x, &z, &circ, &sigma, &u, &w, &uvs, &wvs}) v->assign(n, 0.0);
    }
    // src/VortexBlobs.cpp:211-227 (append_vortices): IDs pushed as Nold + ID_new
    void append(const Blobs& o)
    {
        const uint32_t nold = static_cast<uint32_t>(size());
        for (size_t i = 0; i < o.size(); ++i) id.push_back(nold + o.id[i]);
        x.insert(x.end(), o.x.begin(), o.x.end());
        z.insert(z.end(), o.z.begin(), o.z.end());
        circ.insert(circ.end(), o.circ.begin(), o.circ.end());
        sigma.insert(sigma.end(), o.sigma.begin(), o.sigma.end());
        u.insert(u.end(), o.u.begin(), o.u.end());
        w.insert(w.end(), o.w.begin(), o.w.end());
        uvs.insert(uvs.end(), o.uvs.begin(), o.uvs.end());
        wvs.insert(wvs.end(), o.wvs.begin(), o.wvs.end());
    }
    // src/VortexBlobs.cpp:267-270
    double totalcirc() const { return arma_sum(circ.data(), circ.size()); }

    // src/VortexBlobs.cpp:110-209 (merge_blobs): greedy in index order; blob i
    // absorbs every later unmerged j whose criterion against i's current state
    // is below the threshold; survivors keep their order, IDs become 0..m-1.
    size_t merge(double threshold)
    {
        const size_t n = size();
        if (n < 2) return 0;
        std::vector<bool> merged(n, false);
        size_t count = 0;
        for (size_t i = 0; i < n; ++i) {
            if (merged[i]) continue;
            for (size_t j = i + 1; j < n; ++j) {
                if (merged[j]) continue;
                const double gi = circ[i];
                const double gj = circ[j];
                const double gs = gi + gj;
                if (std::abs(gs) < 1e-15) continue;
                const double dx = x[i] - x[j];
                const double dz = z[i] - z[j];
                const double dist = std::sqrt(dx * dx + dz * dz);
                const double crit = std::abs(gi * gj / gs) * dist;
                if (crit < threshold) {
                    const double nx = (gi * x[i] + gj * x[j]) / gs;
                    const double nz = (gi * z[i] + gj * z[j]) / gs;
                    const double ns = (std::abs(gi) * sigma[i] + std::abs(gj) * sigma[j]) /
                                      (std::abs(gi) + std::abs(gj));
                    x[i] = nx;
                    z[i] = nz;
                    circ[i] = gs;
                    sigma[i] = ns;
                    u[i] = 0.0;
                    w[i] = 0.0;
                    uvs[i] = 0.0;
                    wvs[i] = 0.0;
                    merged[j] = true;
                    ++count;
                }
            }
        }
        if (count == 0) return 0;
        size_t o = 0;
        for (size_t i = 0; i < n; ++i) {
            if (merged[i]) continue;
            id[o] = (uint32_t)o;
            x[o] = x[i]; z[o] = z[i]; circ[o] = circ[i]; sigma[o] = sigma[i];
            u[o] = u[i]; w[o] = w[i]; uvs[o] = uvs[i]; wvs[o] = wvs[i];
            ++o;
        }
        id.resize(o);
        for (Vec* v : {&x, &z, &circ, &sigma, &u, &w, &uvs, &wvs}) v->resize(o);
        return count;
    }
};

// ---------------------------------------------------------------- Biot-Savart
// src/VelocityKernel.cpp:30-67 (DirectKernel::computeSelfInduced)
static void bs_self(int64_t N, const double* px, const double* pz, const double* pc,
                    const double* ps, double* U, double* W)
{
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) {
        double u_i = 0.0;
        double w_i = 0.0;
        for (int64_t j = 0; j < N; ++j) {
            if (i != j) {
                const double dx_ij = px[i] - px[j];
                const double dz_ij = pz[i] - pz[j];
                const double dr_ij2 = dx_ij * dx_ij + dz_ij * dz_ij;
                const double sigma_j_sq = ps[j] * ps[j];
                const double threshold = 10.0 * sigma_j_sq;
                const double rsigmasqr = 1.0 / sigma_j_sq;
                double dK_ij;
                if (dr_ij2 < threshold) {
                    dK_ij = (1.0 - std::exp(-dr_ij2 * rsigmasqr)) / dr_ij2;
                } else {
                    dK_ij = 1.0 / dr_ij2;
                }
                u_i -= dK_ij * dz_ij * pc[j];
                w_i += dK_ij * dx_ij * pc[j];
            }
        }
        U[i] = u_i * rpi2;
        W[i] = w_i * rpi2;
    }
}

// src/VelocityKernel.cpp:69-108 (DirectKernel::computeInducedAt): r^2 <= 1e-15 skipped
static void bs_targets(int64_t Ns, const double* sx, const double* sz, const double* sc,
                       const double* ss, int64_t Nt, const double* tx, const double* tz,
                       double* U, double* W)
{
#pragma omp parallel for
    for (int64_t i = 0; i < Nt; ++i) {
        double u_i = 0.0;
        double w_i = 0.0;
        for (int64_t j = 0; j < Ns; ++j) {
            const double dx_ij = tx[i] - sx[j];
            const double dz_ij = tz[i] - sz[j];
            const double dr_ij2 = dx_ij * dx_ij + dz_ij * dz_ij;
            if (dr_ij2 > 1e-15) {
                const double sigma_j_sq = ss[j] * ss[j];
                const double threshold = 10.0 * sigma_j_sq;
                const double rsigmasqr = 1.0 / sigma_j_sq;
                double dK_ij;
                if (dr_ij2 < threshold) {
                    dK_ij = (1.0 - std::exp(-dr_ij2 * rsigmasqr)) / dr_ij2;
                } else {
                    dK_ij = 1.0 / dr_ij2;
                }
                u_i -= dK_ij * dz_ij * sc[j];
                w_i += dK_ij * dx_ij * sc[j];
            }
        }
        U[i] = u_i * rpi2;
        W[i] = w_i * rpi2;
    }
}

// Long-double version of bs_self: same branch and exclusion, wider arithmetic.
static void bs_self_ld(int64_t N, const double* px, const double* pz, const double* pc,
                       const double* ps, double* U, double* W)
{
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) {
        long double u_i = 0.0L, w_i = 0.0L;
        for (int64_t j = 0; j < N; ++j) {
            if (i == j) continue;
            const long double dx = (long double)px[i] - (long double)px[j];
            const long double dz = (long double)pz[i] - (long double)pz[j];
            const long double r2 = dx * dx + dz * dz;
            const long double s2 = (long double)ps[j] * (long double)ps[j];
            long double K;
            if ((double)r2 < 10.0 * ((double)ps[j] * (double)ps[j]))
                K = -std::expm1(-r2 / s2) / r2;
            else
                K = 1.0L / r2;
            u_i -= K * dz * (long double)pc[j];
            w_i += K * dx * (long double)pc[j];
        }
        const long double r = 1.0L / (2.0L * 3.14159265358979323846264338327950288L);
        U[i] = (double)(u_i * r);
        W[i] = (double)(w_i * r);
    }
}

// ---------------------------------------------------------------- Sheet
struct Sheet {
    Vec x, z;                       // body points (closed: last == first)
    Vec xc, zc, theta, ds, enx, enz, etx, etz, gamma;
    double A = 0.0;                 // reference length (x-extent)
    Vec infM;                       // (Nl+1) x Nl, row-major
    double rho = 1, nu = 0, dt = 0, Ux = 0, Uz = 0, cutoff_exp = 0;
    unsigned maxNPV = 1;
    double fx = 0, fz = 0, cd = 0, cl = 0;
    int Nl = 0;

    // src/VortexSheet.cpp:56-76 (read_input_coord): A = max(x) - min(x)
    // src/VortexSheet.cpp:78-102 (form_vortex_sheet)
    void form(int npts, const double* px, const double* pz)
    {
        x.assign(px, px + npts);
        z.assign(pz, pz + npts);
        A = *std::max_element(x.begin(), x.end()) - *std::min_element(x.begin(), x.end());
        Nl = npts - 1;
        xc.resize(Nl); zc.resize(Nl); theta.resize(Nl); ds.resize(Nl);
        enx.resize(Nl); enz.resize(Nl); etx.resize(Nl); etz.resize(Nl);
        gamma.assign(Nl, 0.0);
        for (int i = 0; i < Nl; ++i) {
            xc[i] = 0.5 * (x[i] + x[i + 1]);
            zc[i] = 0.5 * (z[i] + z[i + 1]);
            const double dx = x[i + 1] - x[i];
            const double dz = z[i + 1] - z[i];
            // arma::sqrt(arma::pow(dx, 2) + arma::pow(dz, 2)); pow(d, 2) taken as d*d
            ds[i] = std::sqrt(dx * dx + dz * dz);
            theta[i] = std::atan2(dz, dx);
            enx[i] = dz / ds[i];
            enz[i] = -dx / ds[i];
            etx[i] = -enz[i];
            etz[i] = enx[i];
        }
    }

    // src/VortexSheet.cpp:104-165 (compute_influence_matrix), panel START nodes
    void influence()
    {
        const int n = Nl;
        Vec p((size_t)n * n), q((size_t)n * n);
        for (int i = 0; i < n; i++) {
            const double xci = xc[i], zci = zc[i], thetai = theta[i];
            for (int j = 0; j < n; j++) {
                if (i == j) {
                    p[(size_t)i * n + j] = -1.0;
                    q[(size_t)i * n + j] = 1.0;
                } else {
                    const double xj = x[j], zj = z[j], thetaj = theta[j], dsj = ds[j];
                    const double c1 = -(xci - xj) * std::cos(thetaj) - (zci - zj) * std::sin(thetaj);
                    const double c2 = (xci - xj) * (xci - xj) + (zci - zj) * (zci - zj);
                    const double c3 = std::sin(thetai - thetaj);
                    const double c4 = std::cos(thetai - thetaj);
                    const double c5 = (xci - xj) * std::sin(thetaj) - (zci - zj) * std::cos(thetaj);
                    const double c6 = std::log(1.0 + dsj * ((dsj + 2 * c1) / c2));
                    const double c7 = std::atan2((c5 * dsj), (c2 + c1 * dsj));
                    const double c9 = (xci - xj) * std::cos(thetai - 2.0 * thetaj)
                                      - (zci - zj) * std::sin(thetai - 2.0 * thetaj);
                    const double qq = c4 + 0.5 * c9 * c6 / dsj - ((c1 * c3 + c4 * c5) * c7) / dsj;
                    q[(size_t)i * n + j] = qq;
                    p[(size_t)i * n + j] = 0.5 * c4 * c6 + c3 * c7 - qq;
                }
            }
        }
        infM.assign((size_t)(n + 1) * n, 0.0);
        for (int i = 0; i < n; i++) {
            infM[(size_t)i * n + 0] = -rpi2 * (p[(size_t)i * n + 0] + q[(size_t)i * n + n - 1]);
            for (int j = 1; j < n; j++)
                infM[(size_t)i * n + j] = -rpi2 * (p[(size_t)i * n + j] + q[(size_t)i * n + j - 1]);
        }
        for (int j = 0; j < n; j++) infM[(size_t)n * n + j] = ds[j];
    }

    // solve(A^T A, A^T b): A^T A and A^T b by plain sequential dot products,
    // then LU with partial pivoting (LAPACK dgesv semantics).  Armadillo may
    // pick Cholesky for a symmetric positive definite matrix; both agree to
    // cond(A^T A)*eps, which is what the tolerance on gamma covers.
    void lsq_solve(const Vec& b)
    {
        const int n = Nl, m = Nl + 1;
        Vec M((size_t)n * n), r(n);
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int k = 0; k < m; ++k) s += infM[(size_t)k * n + i] * infM[(size_t)k * n + j];
                M[(size_t)i * n + j] = s;
            }
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += infM[(size_t)k * n + i] * b[k];
            r[i] = s;
        }
        std::vector<int> piv(n);
        for (int k = 0; k < n; ++k) {
            int pk = k;
            double best = std::fabs(M[(size_t)k * n + k]);
            for (int i = k + 1; i < n; ++i) {
                const double v = std::fabs(M[(size_t)i * n + k]);
                if (v > best) { best = v; pk = i; }
            }
            piv[k] = pk;
            if (pk != k) {
                for (int j = 0; j < n; ++j) std::swap(M[(size_t)k * n + j], M[(size_t)pk * n + j]);
                std::swap(r[k], r[pk]);
            }
            const double d = M[(size_t)k * n + k];
            for (int i = k + 1; i < n; ++i) {
                const double l = M[(size_t)i * n + k] / d;
                M[(size_t)i * n + k] = l;
                for (int j = k + 1; j < n; ++j) M[(size_t)i * n + j] -= l * M[(size_t)k * n + j];
                r[i] -= l * r[k];
            }
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = r[i];
            for (int j = i + 1; j < n; ++j) s -= M[(size_t)i * n + j] * gamma[j];
            gamma[i] = s / M[(size_t)i * n + i];
        }
    }

    // src/VortexSheet.cpp:167-215 (solvevortexsheet)
    void solve(const Blobs& bl, Vec* b_out = nullptr)
    {
        const int n = Nl;
        Vec brhs(n + 1);
        const int64_t N = (int64_t)bl.size();
        if (N == 0) {
            for (int i = 0; i < n; ++i) brhs[i] = Ux * enx[i] + Uz * enz[i];
            brhs[n] = -bl.totalcirc();
        } else {
            Vec u(n), w(n);
#pragma omp parallel for
            for (int i = 0; i < n; i++) {
                u[i] = 0.0;
                w[i] = 0.0;
                for (int64_t j = 0; j < N; j++) {
                    const double dx_ij = xc[i] - bl.x[j];
                    const double dz_ij = zc[i] - bl.z[j];
                    const double dr_ij2 = dx_ij * dx_ij + dz_ij * dz_ij;
                    const double threshold = 10 * (bl.sigma[j] * bl.sigma[j]);
                    const double rsigmasqr = 1.0 / (bl.sigma[j] * bl.sigma[j]);
                    double dK_ij;
                    if (dr_ij2 < threshold) {
                        dK_ij = (1.0 - std::exp(-dr_ij2 * rsigmasqr)) / dr_ij2;
                    } else {
                        dK_ij = 1.0 / dr_ij2;
                    }
                    u[i] -= dK_ij * dz_ij * bl.circ[j];
                    w[i] += dK_ij * dx_ij * bl.circ[j];
                }
            }
            for (int i = 0; i < n; ++i) { u[i] *= rpi2; w[i] *= rpi2; }
            for (int i = 0; i < n; ++i) brhs[i] = (Ux + u[i]) * enx[i] + (Uz + w[i]) * enz[i];
            brhs[n] = -bl.totalcirc();
        }
        if (b_out) *b_out = brhs;
        lsq_solve(brhs);
    }

    // src/VortexSheet.cpp:217-273 (vortexsheetbc): panel origin at the collocation MIDPOINT
    void bc(int64_t N, const double* bx, const double* bz, double* uvs, double* wvs) const
    {
        const int n = Nl;
        const int last = n - 1;
#pragma omp parallel for
        for (int64_t i = 0; i < N; i++) {
            const double xi = bx[i], zi = bz[i];
            double su = 0.0, sw = 0.0;
            for (int j = 0; j < n; j++) {
                const double xj = xc[j], zj = zc[j], thetaj = theta[j], dsj = ds[j];
                const double c1 = -(xi - xj) * std::cos(thetaj) - (zi - zj) * std::sin(thetaj);
                const double c2 = (xi - xj) * (xi - xj) + (zi - zj) * (zi - zj);
                const double c5 = (xi - xj) * std::sin(thetaj) - (zi - zj) * std::cos(thetaj);
                const double c6 = std::log(1.0 + dsj * ((dsj + 2 * c1) / c2));
                const double c7 = std::atan2(c5 * dsj, c2 + c1 * dsj);
                const double c8 = (xi - xj) * std::sin(-2.0 * thetaj) + (zi - zj) * std::cos(-2.0 * thetaj);
                const double c9 = (xi - xj) * std::cos(-2.0 * thetaj) - (zi - zj) * std::sin(-2.0 * thetaj);
                const double qx = -std::sin(thetaj) + 0.5 * c8 * c6 / dsj
                                  + (c1 * std::cos(thetaj) + c5 * std::sin(thetaj)) * c7 / dsj;
                const double px = -0.5 * c6 * std::sin(thetaj) - c7 * std::cos(thetaj) - qx;
                const double qy = std::cos(thetaj) + 0.5 * c9 * c6 / dsj
                                  + (c1 * std::sin(thetaj) - c5 * std::cos(thetaj)) * c7 / dsj;
                const double py = 0.5 * c6 * std::cos(thetaj) - c7 * std::sin(thetaj) - qy;
                const double gj = gamma[j];
                const double gj1 = (j == last) ? gamma[0] : gamma[j + 1];
                su += px * gj + qx * gj1;
                sw += py * gj + qy * gj1;
            }
            uvs[i] = su * rpi2;
            wvs[i] = sw * rpi2;
        }
    }

    // src/VortexSheet.cpp:447-464 (inside_body): crossing number, inside iff cn == 1
    bool inside(double xcoor, double zcoor) const
    {
        int cn = 0;
        for (int i = 0; i < Nl; i++) {
            if (((z[i] <= zcoor) && (z[i + 1] > zcoor)) || ((z[i] > zcoor) && (z[i + 1] <= zcoor))) {
                const double vt = (zcoor - z[i]) / (z[i + 1] - z[i]);
                if (xcoor < x[i] + vt * (x[i + 1] - x[i])) ++cn;
            }
        }
        return cn == 1;
    }

    // src/VortexSheet.cpp:275-333 (release_nascent_vortices_rw)
    // Returns false on the VortexException path (NaN counts).
    bool release(Rng& rng, Blobs& out) const
    {
        return release_with([&rng](unsigned) { return rng.rand(); }, out);
    }

    // The same with the k-th nascent blob's uniform from draw(k).
    template <class Draw>
    bool release_with(Draw draw, Blobs& out) const
    {
        const int n = Nl;
        Vec pc(n), cnt(n), cnew(n);
        for (int i = 0; i < n; ++i) pc[i] = gamma[i] * ds[i];
        const double mx = *std::max_element(pc.begin(), pc.end());
        for (int i = 0; i < n; ++i)
            cnt[i] = std::floor(std::fabs(pc[i] / mx * (double)maxNPV)) + 1.0;
        for (int i = 0; i < n; ++i)
            if (!(cnt[i] > 1.0e-15)) return false;
        for (int i = 0; i < n; ++i) cnew[i] = pc[i] / cnt[i];
        const unsigned Nrv = static_cast<unsigned>(arma_sum(cnt.data(), cnt.size()));
        out.resize(Nrv);
        unsigned counter = 0;
        for (int i = 0; i < n; i++) {
            unsigned j = 0;
            while (j < static_cast<unsigned>(cnt[i])) {
                const double xm = x[i] + ds[i] * etx[i] / (cnt[i] + 1);
                const double zm = z[i] + ds[i] * etz[i] / (cnt[i] + 1);
                const double R = draw(counter);
                const double rrw = std::fabs(std::sqrt(4.0 * nu * dt * std::log(1.0 / R)));
                out.x[counter] = xm + rrw * enx[i];
                out.z[counter] = zm + rrw * enz[i];
                out.circ[counter] = cnew[i];
                out.sigma[counter] = std::pow(ds[i], cutoff_exp);
                counter++;
                j++;
            }
        }
        return true;
    }

    // src/VortexSheet.cpp:345-405 (reflect)
    void reflect(Blobs& v) const
    {
        for (size_t i = 0; i < v.size(); i++) {
            if (!inside(v.x[i], v.z[i])) continue;
            const double px = v.x[i], pz = v.z[i];
            unsigned closest_panel = 0;
            double min_dist = 1e10;
            for (int j = 0; j < Nl; j++) {
                const double x0 = x[j], z0 = z[j], x1 = x[j + 1], z1 = z[j + 1];
                const double vx = px - x0, vz = pz - z0;
                const double dx = x1 - x0, dz = z1 - z0;
                const double len_sq = dx * dx + dz * dz;
                double t = (vx * dx + vz * dz) / len_sq;
                t = std::max(0.0, std::min(1.0, t));
                const double cx = x0 + t * dx, cz = z0 + t * dz;
                const double dist = std::sqrt((px - cx) * (px - cx) + (pz - cz) * (pz - cz));
                if (dist < min_dist) { min_dist = dist; closest_panel = j; }
            }
            const double offset = 2.0 * min_dist + 1e-6;
            v.x[i] = px + offset * enx[closest_panel];
            v.z[i] = pz + offset * enz[closest_panel];
            if (inside(v.x[i], v.z[i])) {
                v.x[i] = px + 3.0 * offset * enx[closest_panel];
                v.z[i] = pz + 3.0 * offset * enz[closest_panel];
            }
        }
    }

    // src/VortexSheet.cpp:407-445 (delete_inside): stable, IDs renumbered 0..M-1
    void delete_inside(Blobs& v) const
    {
        std::vector<size_t> keep;
        keep.reserve(v.size());
        for (size_t i = 0; i < v.size(); ++i)
            if (!inside(v.x[i], v.z[i])) keep.push_back(i);
        if (keep.size() == v.size()) return;
        Blobs nv;
        nv.resize(keep.size());
        for (size_t i = 0; i < keep.size(); ++i) {
            const size_t s = keep[i];
            nv.id[i] = (uint32_t)i;
            nv.x[i] = v.x[s]; nv.z[i] = v.z[s]; nv.circ[i] = v.circ[s]; nv.sigma[i] = v.sigma[s];
            nv.u[i] = v.u[s]; nv.w[i] = v.w[s]; nv.uvs[i] = v.uvs[s]; nv.wvs[i] = v.wvs[s];
        }
        v = std::move(nv);
    }

    // src/VortexSheet.cpp:487-514 (compute_loads)
    void loads(double Ur)
    {
        const int n = Nl;
        Vec P(n, 0.0), Dp(n);
        const double k = -rho / dt;
        for (int i = 0; i < n; ++i) Dp[i] = k * (gamma[i] * ds[i]);
        for (int i = 0; i < n - 1; i++) P[i + 1] = P[i] + Dp[i];
        for (int i = 0; i < n - 1; i++) P[i] = 0.5 * (P[i] + P[i + 1]);
        P[n - 1] = 0.5 * (P[n - 1] + P[0]);
        const double pmax = *std::max_element(P.begin(), P.end());
        const double pref = 0.5 * rho * Ur * Ur;
        for (int i = 0; i < n; ++i) P[i] += (pref - pmax) * 1.0;
        Vec tx(n), tz(n);
        for (int i = 0; i < n; ++i) { tx[i] = P[i] * enx[i] * ds[i]; tz[i] = P[i] * enz[i] * ds[i]; }
        fx = -arma_sum(tx.data(), n);
        fz = -arma_sum(tz.data(), n);
        cd = fx / (0.5 * rho * A * Ur * Ur);
        cl = fz / (0.5 * rho * A * Ur * Ur);
    }

    // src/VortexSheet.cpp:559-646 (probe_velocities), quirks included: loops
    // start at j = 1, the sheet term is assigned (not accumulated), its signs
    // differ from vortexsheetbc and it carries no 1/(2 pi).
    void probe_quirk(const Blobs& bl, int64_t np, const double* prx, const double* prz,
                     double* pu, double* pw, double* puvs, double* pwvs) const
    {
        const int n = Nl;
        for (int64_t i = 0; i < np; i++) { pu[i] = 0.0; pw[i] = 0.0; }
        for (int64_t i = 0; i < np; i++) {
            const double x_i = prx[i], z_i = prz[i];
            double u_i = 0.0, w_i = 0.0;
            for (size_t j = 1; j < bl.size(); j++) {
                const double dx_ij = x_i - bl.x[j];
                const double dz_ij = z_i - bl.z[j];
                const double dr_ij2 = dx_ij * dx_ij + dz_ij * dz_ij;
                const double threshold = 10 * (bl.sigma[j] * bl.sigma[j]);
                const double rsigmasqr = 1.0 / (bl.sigma[j] * bl.sigma[j]);
                double dK_ij;
                if (dr_ij2 < threshold) dK_ij = (1.0 - std::exp(-dr_ij2 * rsigmasqr)) / dr_ij2;
                else dK_ij = 1.0 / dr_ij2;
                const double xkernel = dK_ij * dz_ij;
                const double zkernel = dK_ij * dx_ij;
                u_i = -xkernel * bl.circ[j];
                w_i = +zkernel * bl.circ[j];
                pu[i] = pu[i] + u_i;
                pw[i] = pw[i] + w_i;
            }
        }
        for (int64_t i = 0; i < np; i++) { puvs[i] = 0.0; pwvs[i] = 0.0; }
        for (int64_t i = 0; i < np; i++) {
            for (int j = 1; j < n; j++) {
                const double dxp = prx[i] - xc[j], dzp = prz[i] - zc[j];
                const double th = theta[j], dsj = ds[j];
                const double c1 = -dxp * std::cos(th) - dzp * std::sin(th);
                const double c2 = dxp * dxp + dzp * dzp;
                const double c5 = dxp * std::sin(th) - dzp * std::cos(th);
                const double c6 = std::log(1.0 + dsj * ((dsj + 2 * c1) / c2));
                const double c7 = std::atan2((c5 * dsj), (c2 + c1 * dsj));
                const double c8 = dxp * std::sin(-2.0 * th) + dzp * std::cos(-2.0 * th);
                const double c9 = dxp * std::cos(-2.0 * th) + dzp * std::sin(-2.0 * th);
                const double qx = -std::sin(th) + 0.5 * c8 * c6 / dsj
                                  + (c1 * std::cos(th) + c5 * std::sin(th)) * c7 / dsj;
                const double px = -0.5 * c6 * std::sin(th) - c7 * std::cos(th) - qx;
                const double qy = std::cos(th) + 0.5 * c9 * c6 / dsj
                                  + (c1 * std::cos(th) - c5 * std::sin(th)) * c7 / dsj;
                const double py = -0.5 * c6 * std::cos(th) - c7 * std::sin(th) - qy;
                if (j == n - 1) {
                    puvs[i] = (px * gamma[n - 1] + qx * gamma[0]);
                    pwvs[i] = (py * gamma[n - 1] - qy * gamma[0]);
                } else {
                    puvs[i] = (px * gamma[j] + qx * gamma[j + 1]);
                    pwvs[i] = (py * gamma[j] + qy * gamma[j + 1]);
                }
            }
        }
        for (int64_t i = 0; i < np; i++) {
            pu[i] = rpi2 * pu[i] + puvs[i] + Ux;
            pw[i] = rpi2 * pw[i] + pwvs[i] + Uz;
        }
    }
};

// ---------------------------------------------------------------- Remesh
// BackgroundMesh source is missing from the reference (src/DVMBase.hpp:4,51;
// src/DVMBase.cpp:160-165).  This is the specification the build states,
// with kernels from tests/unit/test_BackgroundMesh.cpp:53-102,149-157,207-263
// and options from src/XmlHandler.cpp:192-201:
//   node (i,k) at (xmin + i*dx, zmin + k*dz), nx = round((xmax-xmin)/dx)+1;
//   blob cell c = floor((x-xmin)/dx); stencil nodes c-s+1 .. c+s (s = 2 for
//   M4', 3 for M6 and Gaussian); a blob whose stencil leaves the mesh is kept
//   unchanged; node value = sum over stencil cells (z-major, ascending) of
//   blobs in index order of (circ * wx) * wz; Gaussian weights are
//   normalised per axis; output = nodes with |circ| > threshold in linear
//   order k*nx + i with sigma = remesh sigma, then kept blobs in order; IDs
//   renumbered 0..M-1.
static double remesh_w(int kernel, double r)
{
    const double a = std::fabs(r);
    if (kernel == 0) {  // M4'
        if (a < 1.0) return 1.0 - 2.5 * a * a + 1.5 * a * a * a;
        if (a < 2.0) { const double t = 2.0 - a; return 0.5 * t * t * (1.0 - a); }
        return 0.0;
    }
    if (kernel == 2) {  // M6: quintic B-spline
        auto p5 = [](double y) { return y > 0.0 ? y * y * y * y * y : 0.0; };
        if (a >= 3.0) return 0.0;
        return (p5(3.0 - a) - 6.0 * p5(2.0 - a) + 15.0 * p5(1.0 - a)) / 120.0;
    }
    // Gaussian, truncated at |r| >= 3
    if (a >= 3.0) return 0.0;
    return std::exp(-0.5 * a * a) / std::sqrt(2.0 * pi);
}

static int remesh_support(int kernel) { return kernel == 0 ? 2 : 3; }

struct RemeshOut { Vec x, z, circ, sigma; std::vector<int64_t> kept; };

static void remesh(const orc_params& p, int64_t n, const double* bx, const double* bz,
                   const double* bc, const double* bs, RemeshOut& out)
{
    const int64_t nx = (int64_t)std::llround((p.remesh_xmax - p.remesh_xmin) / p.remesh_dx) + 1;
    const int64_t nz = (int64_t)std::llround((p.remesh_zmax - p.remesh_zmin) / p.remesh_dz) + 1;
    const int s = remesh_support(p.remesh_kernel);
    std::vector<int64_t> cell(n);
    std::vector<char> inside(n, 0);
    for (int64_t b = 0; b < n; ++b) {
        const double fx = (bx[b] - p.remesh_xmin) / p.remesh_dx;
        const double fz = (bz[b] - p.remesh_zmin) / p.remesh_dz;
        bool ok = std::isfinite(fx) && std::isfinite(fz) && fx >= -1e15 && fx <= 1e15 &&
                  fz >= -1e15 && fz <= 1e15;
        if (ok) {
            const int64_t cx = (int64_t)std::floor(fx), cz = (int64_t)std::floor(fz);
            ok = (cx - s + 1 >= 0) && (cx + s <= nx - 1) && (cz - s + 1 >= 0) && (cz + s <= nz - 1);
            if (ok) cell[b] = cz * nx + cx;
        }
        inside[b] = ok;
    }
    // bin inside blobs by cell, index order within a cell
    std::vector<int64_t> start(nx * nz + 1, 0), order;
    for (int64_t b = 0; b < n; ++b) if (inside[b]) start[cell[b] + 1]++;
    for (int64_t c = 0; c < nx * nz; ++c) start[c + 1] += start[c];
    order.resize(start[nx * nz]);
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t b = 0; b < n; ++b) if (inside[b]) order[fill[cell[b]]++] = b;
    }
    const double xmin = p.remesh_xmin, zmin = p.remesh_zmin, dx = p.remesh_dx, dz = p.remesh_dz;
    auto wts = [&](double f, int64_t base, double* w) {
        // w[t] for stencil node base + t, t in [0, 2s)
        for (int t = 0; t < 2 * s; ++t) w[t] = remesh_w(p.remesh_kernel, f - (double)(base + t));
        if (p.remesh_kernel == 1) {
            double S = 0.0;
            for (int t = 0; t < 2 * s; ++t) S += w[t];
            for (int t = 0; t < 2 * s; ++t) w[t] = w[t] / S;
        }
    };
    out.x.clear(); out.z.clear(); out.circ.clear(); out.sigma.clear(); out.kept.clear();
    for (int64_t k = 0; k < nz; ++k) {
        for (int64_t i = 0; i < nx; ++i) {
            double G = 0.0;
            for (int64_t cz = k - s; cz <= k + s - 1; ++cz) {
                if (cz < 0 || cz >= nz) continue;
                for (int64_t cx = i - s; cx <= i + s - 1; ++cx) {
                    if (cx < 0 || cx >= nx) continue;
                    const int64_t c = cz * nx + cx;
                    for (int64_t q = start[c]; q < start[c + 1]; ++q) {
                        const int64_t b = order[q];
                        const double fx = (bx[b] - xmin) / dx;
                        const double fz = (bz[b] - zmin) / dz;
                        double wx[6], wz[6];
                        wts(fx, cx - s + 1, wx);
                        wts(fz, cz - s + 1, wz);
                        const int tx = (int)(i - (cx - s + 1));
                        const int tz = (int)(k - (cz - s + 1));
                        G += (bc[b] * wx[tx]) * wz[tz];
                    }
                }
            }
            if (std::fabs(G) > p.remesh_circ_threshold) {
                out.x.push_back(xmin + (double)i * dx);
                out.z.push_back(zmin + (double)k * dz);
                out.circ.push_back(G);
                out.sigma.push_back(p.remesh_sigma);
            }
        }
    }
    for (int64_t b = 0; b < n; ++b) {
        if (!inside[b]) {
            out.x.push_back(bx[b]); out.z.push_back(bz[b]);
            out.circ.push_back(bc[b]); out.sigma.push_back(bs[b]);
            out.kept.push_back(b);
        }
    }
}

// ---------------------------------------------------------------- Sim
// src/DVMBase.cpp:8-65 (ctor), 67-133 (solve/compute_step), 160-209 (remesh/convect)
struct Sim {
    orc_params p{};
    Sheet sheet;
    Blobs v;
    Rng rng;
    int step = 0;
    double Ur = 0.0;

    void convect()
    {
        const int64_t N = (int64_t)v.size();
        auto rhs = [&]() {
            if (N > 0) bs_self(N, v.x.data(), v.z.data(), v.circ.data(), v.sigma.data(),
                               v.u.data(), v.w.data());
            sheet.solve(v);
            sheet.bc(N, v.x.data(), v.z.data(), v.uvs.data(), v.wvs.data());
        };
        if (p.scheme == 0) {
            rhs();
            for (int64_t i = 0; i < N; ++i) {
                v.x[i] += (v.u[i] + v.uvs[i] + p.ux) * p.dt;
                v.z[i] += (v.w[i] + v.wvs[i] + p.uz) * p.dt;
            }
        } else {
            const double a[3] = {0, -17. / 32, -32. / 27};
            const double b[3] = {1. / 4, 8. / 9, 3. / 4};
            Vec q1(N, 0.0), q2(N, 0.0);
            for (int s = 0; s < 3; ++s) {
                rhs();
                for (int64_t i = 0; i < N; ++i) {
                    q1[i] = a[s] * q1[i] + (v.u[i] + v.uvs[i] + p.ux) * p.dt;
                    q2[i] = a[s] * q2[i] + (v.w[i] + v.wvs[i] + p.uz) * p.dt;
                }
                for (int64_t i = 0; i < N; ++i) {
                    v.x[i] += b[s] * q1[i];
                    v.z[i] += b[s] * q2[i];
                }
            }
        }
    }

    void do_remesh()
    {
        if (!(p.remesh_frequency > 0 && step > 0 && step % p.remesh_frequency == 0)) return;
        RemeshOut o;
        remesh(p, (int64_t)v.size(), v.x.data(), v.z.data(), v.circ.data(), v.sigma.data(), o);
        const size_t ng = o.x.size() - o.kept.size();
        Blobs nv;
        nv.resize(o.x.size());
        for (size_t i = 0; i < o.x.size(); ++i) {
            nv.id[i] = (uint32_t)i;
            nv.x[i] = o.x[i]; nv.z[i] = o.z[i]; nv.circ[i] = o.circ[i]; nv.sigma[i] = o.sigma[i];
            if (i >= ng) {
                const int64_t b = o.kept[i - ng];
                nv.u[i] = v.u[b]; nv.w[i] = v.w[b]; nv.uvs[i] = v.uvs[b]; nv.wvs[i] = v.wvs[b];
            }
        }
        v = std::move(nv);
    }

    // src/VortexBlobs.cpp:251-265 (diffusion_random_walk)
    void diffuse()
    {
        const DeviceDraw dd{(uint32_t)p.seed, (uint32_t)step};
        for (size_t i = 0; i < v.size(); i++) {
            double R1, R2;
            if (p.rng_mode == 1) {
                dd.two(i, 0, R1, R2);
            } else {
                R1 = rng.rand();
                R2 = rng.rand();
            }
            const double rrw = std::sqrt(4.0 * p.nu * p.dt * std::log(1.0 / R1));
            const double thetarw = 2.0 * pi * R2;
            v.x[i] += rrw * std::cos(thetarw);
            v.z[i] += rrw * std::sin(thetarw);
        }
    }

    bool release()
    {
        Blobs nb;
        if (p.rng_mode == 1) {
            const DeviceDraw dd{(uint32_t)p.seed, (uint32_t)step};
            auto draw = [&dd](unsigned k) {
                double r1, r2;
                dd.two(k, 1, r1, r2);
                return r1;
            };
            if (!sheet.release_with(draw, nb)) return false;
        } else if (!sheet.release(rng, nb)) {
            return false;
        }
        v.append(nb);
        return true;
    }

    void crossing()
    {
        switch (p.crossing) {
        case 0: sheet.delete_inside(v); break;
        case 1: sheet.reflect(v); break;  // ABSORB falls back to reflect (src/DVMBase.cpp:111-114)
        default: sheet.reflect(v); break;
        }
    }

    int do_step()
    {
        step++;
        convect();
        do_remesh();
        diffuse();
        if (!release()) return 3;
        crossing();
        sheet.loads(Ur);
        do_merge();
        return 0;
    }

    // src/DVMBase.cpp:126-132
    size_t do_merge()
    {
        if (p.merge_frequency > 0 && p.merge_threshold > 0 && step % p.merge_frequency == 0)
            return v.merge(p.merge_threshold);
        return 0;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

void* orc_rng_create(int seed)
{
    Rng* r = new Rng();
    r->seed(seed);
    return r;
}
void orc_rng_destroy(void* r) { delete static_cast<Rng*>(r); }
double orc_rng_rand(void* r) { return static_cast<Rng*>(r)->rand(); }
void orc_rng_fill(void* r, int64_t n, double* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = static_cast<Rng*>(r)->rand();
}
uint32_t orc_rng_raw(void* r) { return static_cast<Rng*>(r)->re(); }

int orc_num_threads(void)
{
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}

void orc_bs_self(int64_t n, const double* x, const double* z, const double* circ,
                 const double* sigma, double* u, double* w)
{
    bs_self(n, x, z, circ, sigma, u, w);
}
void orc_bs_targets(int64_t ns, const double* sx, const double* sz, const double* scirc,
                    const double* ssigma, int64_t nt, const double* tx, const double* tz,
                    double* u, double* w)
{
    bs_targets(ns, sx, sz, scirc, ssigma, nt, tx, tz, u, w);
}
void orc_bs_self_ld(int64_t n, const double* x, const double* z, const double* circ,
                    const double* sigma, double* u, double* w)
{
    bs_self_ld(n, x, z, circ, sigma, u, w);
}

static void sheet_init(Sheet& s, int npts, const double* px, const double* pz, const orc_params* p)
{
    s.rho = p->rho; s.nu = p->nu; s.dt = p->dt; s.Ux = p->ux; s.Uz = p->uz;
    s.maxNPV = (unsigned)p->max_npv; s.cutoff_exp = p->cutoff_exp;
    s.form(npts, px, pz);
    s.influence();
}

void* orc_sheet_create(int npts, const double* px, const double* pz, const orc_params* p)
{
    Sheet* s = new Sheet();
    sheet_init(*s, npts, px, pz, p);
    return s;
}
void orc_sheet_destroy(void* sh) { delete static_cast<Sheet*>(sh); }
int orc_sheet_size(void* sh) { return static_cast<Sheet*>(sh)->Nl; }
void orc_sheet_geometry(void* sh, double* xc, double* zc, double* theta, double* ds,
                        double* enx, double* enz, double* etx, double* etz, double* aref)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    const size_t n = s.Nl * sizeof(double);
    std::memcpy(xc, s.xc.data(), n); std::memcpy(zc, s.zc.data(), n);
    std::memcpy(theta, s.theta.data(), n); std::memcpy(ds, s.ds.data(), n);
    std::memcpy(enx, s.enx.data(), n); std::memcpy(enz, s.enz.data(), n);
    std::memcpy(etx, s.etx.data(), n); std::memcpy(etz, s.etz.data(), n);
    *aref = s.A;
}
void orc_sheet_influence(void* sh, double* A)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    std::memcpy(A, s.infM.data(), s.infM.size() * sizeof(double));
}
void orc_sheet_get_gamma(void* sh, double* g)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    std::memcpy(g, s.gamma.data(), s.Nl * sizeof(double));
}
void orc_sheet_set_gamma(void* sh, const double* g)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    std::memcpy(s.gamma.data(), g, s.Nl * sizeof(double));
}
static Blobs make_blobs(int64_t n, const double* x, const double* z, const double* c, const double* s)
{
    Blobs b;
    b.resize(n);
    for (int64_t i = 0; i < n; ++i) { b.x[i] = x[i]; b.z[i] = z[i]; b.circ[i] = c[i]; b.sigma[i] = s[i]; }
    return b;
}
void orc_sheet_solve(void* sh, int64_t n, const double* bx, const double* bz,
                     const double* bcirc, const double* bsigma, double* b_out, double* gamma_out)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    Blobs b = make_blobs(n, bx, bz, bcirc, bsigma);
    Vec rhs;
    s.solve(b, &rhs);
    if (b_out) std::memcpy(b_out, rhs.data(), rhs.size() * sizeof(double));
    if (gamma_out) std::memcpy(gamma_out, s.gamma.data(), s.Nl * sizeof(double));
}
void orc_sheet_bc(void* sh, int64_t n, const double* bx, const double* bz, double* uvs, double* wvs)
{
    static_cast<Sheet*>(sh)->bc(n, bx, bz, uvs, wvs);
}
int orc_sheet_inside(void* sh, double x, double z) { return static_cast<Sheet*>(sh)->inside(x, z) ? 1 : 0; }
void orc_sheet_loads(void* sh, double ur, double* fx, double* fz, double* cd, double* cl)
{
    Sheet& s = *static_cast<Sheet*>(sh);
    s.loads(ur);
    *fx = s.fx; *fz = s.fz; *cd = s.cd; *cl = s.cl;
}
void orc_sheet_probe_quirk(void* sh, int64_t n, const double* bx, const double* bz,
                           const double* bcirc, const double* bsigma, int64_t np,
                           const double* px, const double* pz, double* u, double* w,
                           double* uvs, double* wvs)
{
    Blobs b = make_blobs(n, bx, bz, bcirc, bsigma);
    static_cast<Sheet*>(sh)->probe_quirk(b, np, px, pz, u, w, uvs, wvs);
}

void* orc_sim_create(int npts, const double* px, const double* pz, const orc_params* p)
{
    Sim* s = new Sim();
    s->p = *p;
    sheet_init(s->sheet, npts, px, pz, p);
    s->rng.seed(p->seed);
    s->Ur = std::sqrt(p->ux * p->ux + p->uz * p->uz);
    return s;
}
void orc_sim_destroy(void* sim) { delete static_cast<Sim*>(sim); }
int64_t orc_sim_size(void* sim) { return (int64_t)static_cast<Sim*>(sim)->v.size(); }
int orc_sim_step(void* sim) { return static_cast<Sim*>(sim)->do_step(); }
void orc_sim_get_state(void* sim, uint32_t* id, double* x, double* z, double* circ,
                       double* sigma, double* u, double* w, double* uvs, double* wvs)
{
    const Blobs& v = static_cast<Sim*>(sim)->v;
    const size_t n = v.size() * sizeof(double);
    if (id) std::memcpy(id, v.id.data(), v.size() * sizeof(uint32_t));
    if (x) std::memcpy(x, v.x.data(), n);
    if (z) std::memcpy(z, v.z.data(), n);
    if (circ) std::memcpy(circ, v.circ.data(), n);
    if (sigma) std::memcpy(sigma, v.sigma.data(), n);
    if (u) std::memcpy(u, v.u.data(), n);
    if (w) std::memcpy(w, v.w.data(), n);
    if (uvs) std::memcpy(uvs, v.uvs.data(), n);
    if (wvs) std::memcpy(wvs, v.wvs.data(), n);
}
void orc_sim_set_state(void* sim, int64_t n, const uint32_t* id, const double* x, const double* z,
                       const double* circ, const double* sigma, const double* u, const double* w,
                       const double* uvs, const double* wvs)
{
    Blobs& v = static_cast<Sim*>(sim)->v;
    v.resize(n);
    const size_t b = n * sizeof(double);
    std::memcpy(v.id.data(), id, n * sizeof(uint32_t));
    std::memcpy(v.x.data(), x, b); std::memcpy(v.z.data(), z, b);
    std::memcpy(v.circ.data(), circ, b); std::memcpy(v.sigma.data(), sigma, b);
    std::memcpy(v.u.data(), u, b); std::memcpy(v.w.data(), w, b);
    std::memcpy(v.uvs.data(), uvs, b); std::memcpy(v.wvs.data(), wvs, b);
}
void orc_sim_get_gamma(void* sim, double* g)
{
    const Sheet& s = static_cast<Sim*>(sim)->sheet;
    std::memcpy(g, s.gamma.data(), s.Nl * sizeof(double));
}
void orc_sim_set_gamma(void* sim, const double* g)
{
    Sheet& s = static_cast<Sim*>(sim)->sheet;
    std::memcpy(s.gamma.data(), g, s.Nl * sizeof(double));
}
void orc_sim_get_loads(void* sim, double* fx, double* fz, double* cd, double* cl)
{
    const Sheet& s = static_cast<Sim*>(sim)->sheet;
    *fx = s.fx; *fz = s.fz; *cd = s.cd; *cl = s.cl;
}
int orc_sim_step_index(void* sim) { return static_cast<Sim*>(sim)->step; }
uint64_t orc_sim_rng_draws(void* sim) { return static_cast<Sim*>(sim)->rng.draws; }
void orc_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox(ctr, key, out); }
double orc_u01(uint32_t a, uint32_t b) { return u01(a, b); }
void orc_sim_set_clock(void* sim, int step, uint64_t draws)
{
    // libstdc++ generate_canonical<double, 53> over mt19937 takes 2 words per
    // draw (the stream that src/Random.cpp:3-18 produces)
    Sim* s = static_cast<Sim*>(sim);
    s->step = step;
    s->rng.seed(s->p.seed);
    s->rng.re.discard(2 * draws);
    s->rng.draws = draws;
}
void orc_sim_convect(void* sim) { Sim* s = static_cast<Sim*>(sim); s->step++; s->convect(); }
void orc_sim_remesh(void* sim) { static_cast<Sim*>(sim)->do_remesh(); }
void orc_sim_diffuse(void* sim) { static_cast<Sim*>(sim)->diffuse(); }
int orc_sim_release(void* sim) { return static_cast<Sim*>(sim)->release() ? 0 : 3; }
void orc_sim_crossing(void* sim) { static_cast<Sim*>(sim)->crossing(); }
void orc_sim_loads(void* sim) { Sim* s = static_cast<Sim*>(sim); s->sheet.loads(s->Ur); }
int64_t orc_sim_merge(void* sim) { return (int64_t)static_cast<Sim*>(sim)->do_merge(); }

int64_t orc_merge_blobs(double threshold, int64_t n, uint32_t* id, double* x, double* z,
                        double* circ, double* sigma, double* u, double* w, double* uvs,
                        double* wvs)
{
    Blobs b;
    b.id.assign(id, id + n);
    b.x.assign(x, x + n); b.z.assign(z, z + n); b.circ.assign(circ, circ + n);
    b.sigma.assign(sigma, sigma + n); b.u.assign(u, u + n); b.w.assign(w, w + n);
    b.uvs.assign(uvs, uvs + n); b.wvs.assign(wvs, wvs + n);
    const size_t merged = b.merge(threshold);
    const size_t m = b.size();
    std::copy(b.id.begin(), b.id.end(), id);
    double* outs[8] = {x, z, circ, sigma, u, w, uvs, wvs};
    const Vec* ins[8] = {&b.x, &b.z, &b.circ, &b.sigma, &b.u, &b.w, &b.uvs, &b.wvs};
    for (int k = 0; k < 8; ++k) std::copy(ins[k]->begin(), ins[k]->begin() + m, outs[k]);
    return (int64_t)merged;
}

int64_t orc_remesh(const orc_params* p, int64_t n, const double* x, const double* z,
                   const double* circ, const double* sigma, double* ox, double* oz,
                   double* ocirc, double* osigma)
{
    RemeshOut o;
    remesh(*p, n, x, z, circ, sigma, o);
    if (ox) {
        const size_t b = o.x.size() * sizeof(double);
        std::memcpy(ox, o.x.data(), b); std::memcpy(oz, o.z.data(), b);
        std::memcpy(ocirc, o.circ.data(), b); std::memcpy(osigma, o.sigma.data(), b);
    }
    return (int64_t)o.x.size();
}
double orc_remesh_weight(int kernel, double r) { return remesh_w(kernel, r); }

}  // extern "C"
