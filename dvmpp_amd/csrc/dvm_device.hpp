// Device helpers shared by several kernels (inline, one copy per translation unit).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dvm_internal.hpp"

namespace dvmi {

// Tile record of the FP32C sum: FP64 centre, FP32 half extents (rounded up)
// and 10 * max(sigma)^2 over the tile's sources.
struct TileInfo {
    double cx, cz;
    float hx, hz;
    float thr_max;
    float pad;
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

// src/VortexSheet.cpp:447-464: crossing number over the Nl edges; inside iff 1.
__device__ __forceinline__ bool inside_body(int nl, const double* __restrict__ bxp,
                                            const double* __restrict__ bzp, double xc, double zc)
{
    int cn = 0;
    for (int i = 0; i < nl; i++) {
        const double zi = bzp[i], zi1 = bzp[i + 1];
        if (((zi <= zc) && (zi1 > zc)) || ((zi > zc) && (zi1 <= zc))) {
            const double vt = (zc - zi) / (zi1 - zi);
            if (xc < bxp[i] + vt * (bxp[i + 1] - bxp[i])) ++cn;
        }
    }
    return cn == 1;
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
    const double qx = -q.sn + 0.5 * c8 * c6 / dsj + (c1 * q.cs + c5 * q.sn) * c7 / dsj;
    const double px = -0.5 * c6 * q.sn - c7 * q.cs - qx;
    const double qy = q.cs + 0.5 * c9 * c6 / dsj + (c1 * q.sn - c5 * q.cs) * c7 / dsj;
    const double py = 0.5 * c6 * q.cs - c7 * q.sn - qy;
    su += px * gj + qx * gj1;
    sw += py * gj + qy * gj1;
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

}  // namespace dvmi
