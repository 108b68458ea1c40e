// Device-side types and small 3x3 helpers shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace pk
{

// Per-group CG state, owned by the device. The last workgroup of each
// reduction updates it; every kernel of the solve reads `done` first.
struct GroupState
{
    double delta;     // r.z of the current residual (deltaNew)
    double target;    // tol^2 * delta_0
    double alpha, beta;
    double rr;        // r.r of the current residual
    double residual;  // reported residual norm
    double tol;
    int maxIt;
    int iters;
    int done;         // the CG loop of this group is over
    int converged;
    int failed;       // the solve is unusable: do not integrate
    int nonfinite;    // integration produced a non-finite value
    int hasRows;
    int jacobi;
    unsigned cntRhs, cntSpmv, cntUpd;
    int pad;
};

// One workgroup of a row-tiled kernel: rows [rowBegin, rowEnd) of group `group`.
struct RowTile
{
    int group, rowBegin, rowEnd, pad;
};

struct Partial
{
    double a, b;
    int flag, pad;
};

__device__ inline bool fin(double v) { return __builtin_isfinite(v); }

// 3x3 row-major helpers. Column j of m is (m[j], m[3+j], m[6+j]).
__device__ inline double det3(const double* m)
{
    // c0 . (c1 x c2), Mat3.h:52-55
    const double c1x = m[1], c1y = m[4], c1z = m[7];
    const double c2x = m[2], c2y = m[5], c2z = m[8];
    const double rx = c1y * c2z - c1z * c2y;
    const double ry = c1z * c2x - c1x * c2z;
    const double rz = c1x * c2y - c1y * c2x;
    return m[0] * rx + m[3] * ry + m[6] * rz;
}

// Mat3::Inverse (Mat3.h:57-74): rows of the inverse are the cross products of
// the columns divided by the determinant; |det| <= 1e-6f gives Identity.
__device__ inline void inv3(const double* m, double* o)
{
    const double c0x = m[0], c0y = m[3], c0z = m[6];
    const double c1x = m[1], c1y = m[4], c1z = m[7];
    const double c2x = m[2], c2y = m[5], c2z = m[8];
    const double r0x = c1y * c2z - c1z * c2y, r0y = c1z * c2x - c1x * c2z, r0z = c1x * c2y - c1y * c2x;
    const double r1x = c2y * c0z - c2z * c0y, r1y = c2z * c0x - c2x * c0z, r1z = c2x * c0y - c2y * c0x;
    const double r2x = c0y * c1z - c0z * c1y, r2y = c0z * c1x - c0x * c1z, r2z = c0x * c1y - c0y * c1x;
    const double d = c0x * r0x + c0y * r0y + c0z * r0z;
    if (fabs(d) <= static_cast<double>(0.000001f))
    {
        o[0] = 1.0; o[1] = 0.0; o[2] = 0.0;
        o[3] = 0.0; o[4] = 1.0; o[5] = 0.0;
        o[6] = 0.0; o[7] = 0.0; o[8] = 1.0;
        return;
    }
    const double id = 1.0 / d;
    o[0] = r0x * id; o[1] = r0y * id; o[2] = r0z * id;
    o[3] = r1x * id; o[4] = r1y * id; o[5] = r1z * id;
    o[6] = r2x * id; o[7] = r2y * id; o[8] = r2z * id;
}

__device__ inline bool fin9(const double* m)
{
    bool ok = true;
#pragma unroll
    for (int e = 0; e < 9; ++e)
        ok = ok && fin(m[e]);
    return ok;
}

}  // namespace pk
