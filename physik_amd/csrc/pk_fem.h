// Per-tet finite-element helpers shared by the forward and reverse kernels.
// 3x3 matrices are row-major double[9].
#pragma once

#include "pk_device.h"

namespace pk
{

// FEMModel.cpp:115-163 with the reference thresholds (float literals).
// `kind` (optional) reports how R was obtained: 0 Newton limit, 1 Newton
// limit with column 2 negated, 2 Identity fallback.
__device__ inline void polar(const double* F, double* R, int* kind = nullptr)
{
    bool ident = !fin9(F) || fabs(det3(F)) <= static_cast<double>(1.0e-8f);
    double c[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        c[e] = F[e];
    if (!ident)
    {
        for (int it = 0; it < 12; ++it)
        {
            double t[9], ti[9];
            t[0] = c[0]; t[1] = c[3]; t[2] = c[6];
            t[3] = c[1]; t[4] = c[4]; t[5] = c[7];
            t[6] = c[2]; t[7] = c[5]; t[8] = c[8];
            inv3(t, ti);
            if (!fin9(ti) || fabs(det3(ti)) <= static_cast<double>(1.0e-8f))
            {
                ident = true;
                break;
            }
            double nx[9];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                nx[e] = (c[e] + ti[e]) * 0.5;
            if (!fin9(nx))
            {
                ident = true;
                break;
            }
            double diff = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                const double dx = nx[j] - c[j], dy = nx[3 + j] - c[3 + j], dz = nx[6 + j] - c[6 + j];
                diff += dx * dx + dy * dy + dz * dz;
            }
#pragma unroll
            for (int e = 0; e < 9; ++e)
                c[e] = nx[e];
            if (diff <= static_cast<double>(1.0e-10f))
                break;
        }
        if (!ident && !fin9(c))
            ident = true;
    }
    if (ident)
    {
#pragma unroll
        for (int e = 0; e < 9; ++e)
            R[e] = (e % 4 == 0) ? 1.0 : 0.0;
        if (kind)
            *kind = 2;
        return;
    }
    if (kind)
        *kind = 0;
    if (det3(c) < 0.0)
    {
        c[2] = -c[2];
        c[5] = -c[5];
        c[8] = -c[8];
        if (kind)
            *kind = 1;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e)
        R[e] = c[e];
}

// Local pair p of the 10 stored pairs a <= b: (0,0),(0,1),(0,2),(0,3),(1,1),
// (1,2),(1,3),(2,2),(2,3),(3,3).
__device__ inline void pair_of(int p, int& a, int& b)
{
    constexpr int A[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    constexpr int B[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    a = A[p];
    b = B[p];
}

// Shape gradients: g_a = row a-1 of Dm^-1 (a = 1..3), g_0 = -(g_1 + g_2 + g_3).
__device__ inline void shape_gradients(const double* Dinv, double g[4][3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        g[1][i] = Dinv[i];
        g[2][i] = Dinv[3 + i];
        g[3][i] = Dinv[6 + i];
        g[0][i] = -((g[1][i] + g[2][i]) + g[3][i]);
    }
}

__device__ inline double gdot(const double* ga, const double* gb) { return ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]; }

// R K_ab R^T from the premultiplied factors vl = V lam, vm = V mu and
// vmg = V mu (g_a . g_b) (the tet's stiffness record); RRt is full 3x3.
// The sum vl rga_i rgb_j + vm rgb_i rga_j + vmg RRt_ij is written as explicit
// fused multiply-adds: left to the compiler, which products it fuses depends
// on the code around each inlined copy, and every kernel that forms these
// blocks (the gathers, the inspection records) must round them alike.
__device__ inline void rotated_block_scaled(const double* rga, const double* rgb, const double* RRt, double vl,
                                            double vm, double vmg, double* K)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double sym = (vm * rgb[i]) * rga[j];
            K[3 * i + j] = __builtin_fma(vmg, RRt[3 * i + j], __builtin_fma(vl * rga[i], rgb[j], sym));
        }
}

// R K_aa R^T of one corner, symmetric bit for bit: with w = R g_a the block is
// V lambda w w^T + V mu w w^T + V mu (g_a . g_a) R R^T, and each product
// w_i w_j is formed once for (i, j) and (j, i) (RRt is symmetric too), so the
// diagonal blocks of A, their sums and their block-Jacobi inverses are
// symmetric and the inverse can be kept as its upper triangle.
__device__ inline void rotated_diag_block_scaled(const double* w, const double* RRt, double vl, double vm, double vmg,
                                                 double* K)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double t = w[i < j ? i : j] * w[i < j ? j : i];
            K[3 * i + j] = __builtin_fma(vmg, RRt[3 * i + j], __builtin_fma(vl, t, vm * t));
        }
}


}  // namespace pk
