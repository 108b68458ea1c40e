// MI355X (gfx950) implicit-Euler step: HIP kernels and the pk_system runtime.
//
// One substep runs, on one stream:
//   k_reset      group CG state
//   k_element    one lane per tet: F = Ds Dm^-1, Newton polar R, corotated
//                forces and the 10 blocks R K_ab R^T, K v per corner
//                (FEMModel.cpp:614-679, :377-424, :336-375, :115-163)
//   k_assemble   one lane per upper BSR slot: gather the slot's element blocks
//                in tet order (the precomputed element->slot map; no atomics),
//                A = M + dt^2 K written to both triangles, 3x3 block-Jacobi
//                inverse on diagonal slots (SolverData.cpp:308-465, :260-306)
//   k_rhs        one lane per dynamic row: b = dt (f + m g) - dt^2 K v, CG start
//                x = 0, r = b, d = M^-1 r, delta_0 = r.d (ConjugateGradientSolver.cpp:151-176)
//   CG loop      k_spmv (q = A d, d.q -> alpha), k_update (x, r, z = M^-1 r,
//                r.z -> beta, stop test), k_direction (d = z + beta d)
//                (ConjugateGradientSolver.cpp:178-241)
//   k_integrate  check then commit v += dv, x += dt v, all-or-nothing per
//                group (World.cpp:644-695)
//
// Reductions are deterministic: every workgroup writes a partial, and the
// last workgroup of a group (counted with one atomic) sums the group's
// partials in a fixed order and updates that group's scalars. Nothing else
// uses atomics, so two runs are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "physik_amd.h"
#include "pk_kernels.h"
#include "pk_topology.h"

namespace pk
{

thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

#define PK_HIP(expr)                                                                                   \
    do                                                                                                 \
    {                                                                                                  \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
        {                                                                                              \
            set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));                      \
            return PK_ERR_HIP;                                                                         \
        }                                                                                              \
    } while (0)

constexpr int NT = 256;  // threads per workgroup for every kernel

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup; the result is valid in thread 0.
__device__ inline double block_sum(double v, double* sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return s;
}

// Publishes this workgroup's partial and tells whether it is the last of its
// group. The last workgroup then reads every partial of the group.
__device__ inline bool publish_partial(Partial* part, int wg, double a, double b, int flag, unsigned* counter,
                                       int groupWgs)
{
    __shared__ int last;
    if (threadIdx.x == 0)
    {
        Partial p;
        p.a = a;
        p.b = b;
        p.flag = flag;
        p.pad = 0;
        part[wg] = p;
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = (prev == static_cast<unsigned>(groupWgs) - 1u) ? 1 : 0;
    }
    __syncthreads();
    if (last)
        __threadfence();
    return last != 0;
}

// Fixed-order sum of the group's partials by the last workgroup.
__device__ inline void gather_partials(const Partial* part, int begin, int count, double* sh, double* a, double* b,
                                       int* flag)
{
    double sa = 0.0, sb = 0.0;
    int f = 0;
    for (int i = threadIdx.x; i < count; i += NT)
    {
        const Partial p = part[begin + i];
        sa += p.a;
        sb += p.b;
        f |= p.flag;
    }
    *a = block_sum(sa, sh);
    *b = block_sum(sb, sh);
    *flag = __syncthreads_or(f);
}

__device__ inline double norm_of(double rr) { return (fin(rr) && rr >= 0.0) ? sqrt(rr) : INFINITY; }

__device__ inline void finish_group(GroupState& s, int* active, bool converged)
{
    s.done = 1;
    s.converged = converged ? 1 : 0;
    atomicSub(active, 1);
}

// ---------------------------------------------------------------------------
// k_reset
// ---------------------------------------------------------------------------
__global__ void k_reset(GroupState* gs, int G, int* active)
{
    __shared__ int cnt;
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
    {
        GroupState& s = gs[g];
        s.delta = s.target = s.alpha = s.beta = s.rr = s.residual = 0.0;
        s.iters = 0;
        s.converged = 0;
        s.failed = 0;
        s.nonfinite = 0;
        s.cntRhs = s.cntSpmv = s.cntUpd = 0;
        s.done = s.hasRows ? 0 : 1;
        if (s.hasRows)
            atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *active = cnt;
}

// ---------------------------------------------------------------------------
// k_element
// ---------------------------------------------------------------------------
// FEMModel.cpp:115-163 with the reference thresholds (float literals).
__device__ inline void polar(const double* F, double* R)
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
        return;
    }
    if (det3(c) < 0.0)
    {
        c[2] = -c[2];
        c[5] = -c[5];
        c[8] = -c[8];
    }
#pragma unroll
    for (int e = 0; e < 9; ++e)
        R[e] = c[e];
}

__global__ __launch_bounds__(NT) void k_element(int T, const int4* __restrict__ tets,
                                                const unsigned char* __restrict__ live,
                                                const double* __restrict__ prm, const double* __restrict__ x,
                                                const double* __restrict__ v, double* __restrict__ elem)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= T || !live[t])
        return;
    const int4 n4 = tets[t];
    const int ids[4] = {n4.x, n4.y, n4.z, n4.w};
    double xs[4][3], vs[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[a][i] = x[3 * ids[a] + i];
            vs[a][i] = v[3 * ids[a] + i];
        }
    const size_t TT = static_cast<size_t>(T);
    double Dinv[9], Erest[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
    {
        Dinv[k] = prm[(P_DMINV + k) * TT + t];
        Erest[k] = prm[(P_EDGE + k) * TT + t];
    }
    const double vol = prm[P_VOL * TT + t];
    const double lam = prm[P_LAMBDA * TT + t];
    const double mu = prm[P_MU * TT + t];
    const double damp = prm[P_DAMP * TT + t];
    double* out = elem + TT * 0 + static_cast<size_t>(t) * E_STRIDE;

    double ds[3][3];  // ds[k] = x_{k+1} - x_0
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            ds[k][i] = xs[k + 1][i] - xs[0][i];
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = ds[0][i] * Dinv[j] + ds[1][i] * Dinv[3 + j] + ds[2][i] * Dinv[6 + j];
    if (!fin9(F))
    {
        // the reference skips the tet: it contributes nothing this substep
        for (int k = 0; k < E_STRIDE; ++k)
            out[k] = 0.0;
        return;
    }
    double R[9];
    polar(F, R);

    // shape gradients: g_a = row a-1 of Dm^-1 (a = 1..3), g_0 = -(g_1 + g_2 + g_3)
    double g[4][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        g[1][i] = Dinv[i];
        g[2][i] = Dinv[3 + i];
        g[3][i] = Dinv[6 + i];
        g[0][i] = -((g[1][i] + g[2][i]) + g[3][i]);
    }
    // u_a = R^T (x_a - x_0) - (X_a - X_0); u_0 = 0
    double u[4][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        u[0][i] = 0.0;
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            u[a][i] = (R[i] * ds[a - 1][0] + R[3 + i] * ds[a - 1][1] + R[6 + i] * ds[a - 1][2]) -
                      Erest[3 * (a - 1) + i];
    // engineering strain and stress (FEMModel.cpp:174-225, 386-389)
    double exx = 0, eyy = 0, ezz = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
    for (int a = 1; a < 4; ++a)
    {
        exx += g[a][0] * u[a][0];
        eyy += g[a][1] * u[a][1];
        ezz += g[a][2] * u[a][2];
        gxy += g[a][1] * u[a][0] + g[a][0] * u[a][1];
        gyz += g[a][2] * u[a][1] + g[a][1] * u[a][2];
        gxz += g[a][2] * u[a][0] + g[a][0] * u[a][2];
    }
    const double tr = exx + eyy + ezz;
    const double sxx = lam * tr + 2.0 * mu * exx;
    const double syy = lam * tr + 2.0 * mu * eyy;
    const double szz = lam * tr + 2.0 * mu * ezz;
    const double txy = mu * gxy, tyz = mu * gyz, txz = mu * gxz;
    // f_a = R (-V B_a^T sigma) - c v_a
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
        const double fx = -vol * (sxx * g[a][0] + txy * g[a][1] + txz * g[a][2]);
        const double fy = -vol * (txy * g[a][0] + syy * g[a][1] + tyz * g[a][2]);
        const double fz = -vol * (txz * g[a][0] + tyz * g[a][1] + szz * g[a][2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[E_FORCE + 3 * a + i] = (R[3 * i] * fx + R[3 * i + 1] * fy + R[3 * i + 2] * fz) - damp * vs[a][i];
    }
    // R K_ab R^T = V (lam (R g_a)(R g_b)^T + mu (R g_b)(R g_a)^T + mu (g_a.g_b) R R^T)
    double rg[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            rg[a][i] = R[3 * i] * g[a][0] + R[3 * i + 1] * g[a][1] + R[3 * i + 2] * g[a][2];
    double RRt[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            RRt[3 * i + j] = R[3 * i] * R[3 * j] + R[3 * i + 1] * R[3 * j + 1] + R[3 * i + 2] * R[3 * j + 2];
    double kv[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            kv[a][i] = 0.0;
    int p = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b, ++p)
        {
            const double gab = mu * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
            double K[9];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    K[3 * i + j] = vol * (lam * rg[a][i] * rg[b][j] + mu * rg[b][i] * rg[a][j] + gab * RRt[3 * i + j]);
#pragma unroll
            for (int e = 0; e < 9; ++e)
                out[E_BLOCKS + 9 * p + e] = K[e];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                kv[a][i] += K[3 * i] * vs[b][0] + K[3 * i + 1] * vs[b][1] + K[3 * i + 2] * vs[b][2];
            if (a != b)
            {
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    kv[b][i] += K[i] * vs[a][0] + K[3 + i] * vs[a][1] + K[6 + i] * vs[a][2];
            }
        }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[E_KV + 3 * a + i] = kv[a][i];
}

// ---------------------------------------------------------------------------
// k_assemble
// ---------------------------------------------------------------------------
template <class M>
__global__ __launch_bounds__(NT) void k_assemble(int nnzu, const int* __restrict__ cstart,
                                                 const int* __restrict__ clist, const int* __restrict__ uFull,
                                                 const int* __restrict__ uMirror, const int* __restrict__ slotRow,
                                                 const int* __restrict__ rowGroup, const double* __restrict__ elem,
                                                 const double* __restrict__ mrow, double dt2, M* __restrict__ vals,
                                                 long long nnzf, double* __restrict__ minv, int nd, GroupState* gs)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= nnzu)
        return;
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] = 0.0;
    const int k1 = cstart[s + 1];
    for (int k = cstart[s]; k < k1; ++k)
    {
        const int c = clist[k];
        const double* b = elem + static_cast<size_t>(c >> 5) * E_STRIDE + 9 * ((c >> 1) & 15);
        if (c & 1)
        {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    a[3 * i + j] += b[3 * j + i];
        }
        else
        {
#pragma unroll
            for (int e = 0; e < 9; ++e)
                a[e] += b[e];
        }
    }
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] *= dt2;
    const int mirror = uMirror[s];
    const int r = slotRow[s];
    if (mirror < 0)
    {
        const double m = mrow[r];
        a[0] = m + a[0];
        a[4] = m + a[4];
        a[8] = m + a[8];
    }
    const size_t F = static_cast<size_t>(nnzf);
    const int full = uFull[s];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        vals[e * F + full] = static_cast<M>(a[e]);
    if (mirror >= 0)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                vals[(3 * i + j) * F + mirror] = static_cast<M>(a[3 * j + i]);
    }
    const bool ok = fin9(a);
    const int g = rowGroup[r];
    if (!ok)
        gs[g].failed = 1;  // a non-finite stiffness block rejects the assembly
    if (mirror < 0)
    {
        double mi[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (gs[g].jacobi)
        {
            bool pok = ok;
            if (pok)
            {
                const double d = det3(a);
                pok = fin(d) && fabs(d) > static_cast<double>(1.0e-8f);
            }
            if (pok)
            {
                inv3(a, mi);
                pok = fin9(mi);
            }
            if (!pok)
                gs[g].failed = 1;  // BuildInversePreconditioner fails the solve
        }
#pragma unroll
        for (int e = 0; e < 9; ++e)
            minv[static_cast<size_t>(e) * nd + r] = mi[e];
    }
}

// ---------------------------------------------------------------------------
// k_rhs: right-hand side and CG start
// ---------------------------------------------------------------------------
template <class M>
__global__ __launch_bounds__(NT) void k_rhs(const RowTile* __restrict__ tiles, const int* __restrict__ wgBegin,
                                            const int* __restrict__ wgCount, const int* __restrict__ nstart,
                                            const int* __restrict__ nlist, const double* __restrict__ elem,
                                            const double* __restrict__ mrow, double gx, double gy, double gz,
                                            double dt, double dt2, const double* __restrict__ minv, int nd,
                                            double* __restrict__ xs, double* __restrict__ rv,
                                            double* __restrict__ zv, M* __restrict__ dv, Partial* part,
                                            GroupState* gs, int* active)
{
    __shared__ double sh[4];
    const RowTile tl = tiles[blockIdx.x];
    const int g = tl.group;
    const int row = tl.rowBegin + threadIdx.x;
    double pa = 0.0, pb = 0.0;
    int flag = 0;
    if (row < tl.rowEnd)
    {
        double f[3] = {0, 0, 0}, kv[3] = {0, 0, 0};
        const int k1 = nstart[row + 1];
        for (int k = nstart[row]; k < k1; ++k)
        {
            const int c = nlist[k];
            const double* e = elem + static_cast<size_t>(c >> 2) * E_STRIDE;
            const int a = c & 3;
            const double f0 = e[E_FORCE + 3 * a], f1 = e[E_FORCE + 3 * a + 1], f2 = e[E_FORCE + 3 * a + 2];
            if (fin(f0) && fin(f1) && fin(f2))  // non-finite pushes are skipped (SolverData.cpp:145-154)
            {
                f[0] += f0;
                f[1] += f1;
                f[2] += f2;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
                kv[i] += e[E_KV + 3 * a + i];
        }
        const double m = mrow[row];
        const double gv[3] = {gx, gy, gz};
        double b[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            b[i] = (f[i] + gv[i] * m) * dt - kv[i] * dt2;
        if (!(fin(b[0]) && fin(b[1]) && fin(b[2])))
            flag |= 1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            z[i] = minv[(3 * i) * static_cast<size_t>(nd) + row] * b[0] +
                   minv[(3 * i + 1) * static_cast<size_t>(nd) + row] * b[1] +
                   minv[(3 * i + 2) * static_cast<size_t>(nd) + row] * b[2];
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag |= 2;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[3 * row + i] = 0.0;
            rv[3 * row + i] = b[i];
            zv[3 * row + i] = z[i];
            dv[3 * row + i] = static_cast<M>(z[i]);
        }
        pa = b[0] * z[0] + b[1] * z[1] + b[2] * z[2];
        pb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    }
    pa = block_sum(pa, sh);
    pb = block_sum(pb, sh);
    flag = __syncthreads_or(flag);
    GroupState& s = gs[g];
    if (!publish_partial(part, blockIdx.x, pa, pb, flag, &s.cntRhs, wgCount[g]))
        return;
    double delta0, rr;
    int flags;
    gather_partials(part, wgBegin[g], wgCount[g], sh, &delta0, &rr, &flags);
    if (threadIdx.x != 0)
        return;
    s.cntRhs = 0;
    s.rr = rr;
    s.delta = delta0;
    if (s.failed || (flags & 1) || s.maxIt <= 0 || !fin(s.tol) || !(s.tol > 0.0))
    {
        // rejected: the solution is cleared and nothing is integrated
        s.failed = 1;
        finish_group(s, active, false);
        return;
    }
    if (flags & 2)
    {
        finish_group(s, active, false);  // preconditioner application failed: dv = 0
        return;
    }
    if (!fin(delta0) || delta0 < 0.0)
    {
        finish_group(s, active, false);
        return;
    }
    const double target = s.tol * s.tol * delta0;
    s.target = target;
    if (!fin(target))
    {
        finish_group(s, active, false);
        return;
    }
    if (delta0 <= target)
    {
        s.residual = norm_of(rr);
        finish_group(s, active, true);
    }
}

// ---------------------------------------------------------------------------
// k_spmv: q = A d and d.q -> alpha
// ---------------------------------------------------------------------------
template <class M, int LPR>
__global__ __launch_bounds__(NT) void k_spmv(const RowTile* __restrict__ tiles, const int* __restrict__ wgBegin,
                                             const int* __restrict__ wgCount, const int* __restrict__ fstart,
                                             const int* __restrict__ fcol, const M* __restrict__ vals,
                                             long long nnzf, const M* __restrict__ d, M* __restrict__ q,
                                             Partial* part, GroupState* gs, int* active)
{
    __shared__ double sh[4];
    const RowTile tl = tiles[blockIdx.x];
    const int g = tl.group;
    GroupState& s = gs[g];
    if (s.done)
        return;
    const int lane = threadIdx.x % LPR;
    const int sub = threadIdx.x / LPR;
    constexpr int RPP = NT / LPR;
    const size_t F = static_cast<size_t>(nnzf);
    double pd = 0.0;
    for (int row = tl.rowBegin + sub; row < tl.rowEnd; row += RPP)
    {
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        const int k1 = fstart[row + 1];
        for (int k = fstart[row] + lane; k < k1; k += LPR)
        {
            const int c = fcol[k];
            const double d0 = d[3 * c], d1 = d[3 * c + 1], d2 = d[3 * c + 2];
            y0 += static_cast<double>(vals[0 * F + k]) * d0 + static_cast<double>(vals[1 * F + k]) * d1 +
                  static_cast<double>(vals[2 * F + k]) * d2;
            y1 += static_cast<double>(vals[3 * F + k]) * d0 + static_cast<double>(vals[4 * F + k]) * d1 +
                  static_cast<double>(vals[5 * F + k]) * d2;
            y2 += static_cast<double>(vals[6 * F + k]) * d0 + static_cast<double>(vals[7 * F + k]) * d1 +
                  static_cast<double>(vals[8 * F + k]) * d2;
        }
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1)
        {
            y0 += __shfl_xor(y0, o, LPR);
            y1 += __shfl_xor(y1, o, LPR);
            y2 += __shfl_xor(y2, o, LPR);
        }
        if (lane == 0)
        {
            const M q0 = static_cast<M>(y0), q1 = static_cast<M>(y1), q2 = static_cast<M>(y2);
            q[3 * row] = q0;
            q[3 * row + 1] = q1;
            q[3 * row + 2] = q2;
            pd += static_cast<double>(d[3 * row]) * q0 + static_cast<double>(d[3 * row + 1]) * q1 +
                  static_cast<double>(d[3 * row + 2]) * q2;
        }
    }
    pd = block_sum(pd, sh);
    if (!publish_partial(part, blockIdx.x, pd, 0.0, 0, &s.cntSpmv, wgCount[g]))
        return;
    double dq, unused;
    int fl;
    gather_partials(part, wgBegin[g], wgCount[g], sh, &dq, &unused, &fl);
    if (threadIdx.x != 0)
        return;
    s.cntSpmv = 0;
    const double alpha = s.delta / dq;
    if (!fin(dq) || dq <= 0.0 || !fin(alpha))
    {
        s.residual = norm_of(s.rr);
        finish_group(s, active, s.delta <= s.target);
        return;
    }
    s.alpha = alpha;
}

// ---------------------------------------------------------------------------
// k_update: x += alpha d, r -= alpha q, z = M^-1 r, r.z and r.r -> beta, stop
// ---------------------------------------------------------------------------
template <class M>
__global__ __launch_bounds__(NT) void k_update(const RowTile* __restrict__ tiles, const int* __restrict__ wgBegin,
                                               const int* __restrict__ wgCount, const double* __restrict__ minv,
                                               int nd, const M* __restrict__ d, const M* __restrict__ q,
                                               double* __restrict__ xs, double* __restrict__ rv,
                                               double* __restrict__ zv, Partial* part, GroupState* gs,
                                               int* active)
{
    __shared__ double sh[4];
    const RowTile tl = tiles[blockIdx.x];
    const int g = tl.group;
    GroupState& s = gs[g];
    if (s.done)
        return;
    const double alpha = s.alpha;
    const int row = tl.rowBegin + threadIdx.x;
    double pa = 0.0, pb = 0.0;
    int flag = 0;
    if (row < tl.rowEnd)
    {
        double r[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[3 * row + i] = xs[3 * row + i] + static_cast<double>(d[3 * row + i]) * alpha;
            r[i] = rv[3 * row + i] + static_cast<double>(q[3 * row + i]) * (-alpha);
            rv[3 * row + i] = r[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            z[i] = minv[(3 * i) * static_cast<size_t>(nd) + row] * r[0] +
                   minv[(3 * i + 1) * static_cast<size_t>(nd) + row] * r[1] +
                   minv[(3 * i + 2) * static_cast<size_t>(nd) + row] * r[2];
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag = 1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            zv[3 * row + i] = z[i];
        pa = r[0] * z[0] + r[1] * z[1] + r[2] * z[2];
        pb = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    }
    pa = block_sum(pa, sh);
    pb = block_sum(pb, sh);
    flag = __syncthreads_or(flag);
    if (!publish_partial(part, blockIdx.x, pa, pb, flag, &s.cntUpd, wgCount[g]))
        return;
    double rz, rr;
    int flags;
    gather_partials(part, wgBegin[g], wgCount[g], sh, &rz, &rr, &flags);
    if (threadIdx.x != 0)
        return;
    s.cntUpd = 0;
    s.rr = rr;
    if (flags)
    {
        s.residual = norm_of(rr);
        finish_group(s, active, s.delta <= s.target);
        return;
    }
    const double old = s.delta;
    s.delta = rz;
    const double beta = rz / old;
    if (!fin(rz) || !fin(old) || old <= 0.0 || !fin(beta))
    {
        s.residual = norm_of(rr);
        finish_group(s, active, rz <= s.target);
        return;
    }
    s.beta = beta;
    s.iters += 1;
    if (rz <= s.target || s.iters >= s.maxIt)
    {
        s.residual = norm_of(rr);
        finish_group(s, active, rz <= s.target);
    }
}

// ---------------------------------------------------------------------------
// k_direction: d = z + beta d
// ---------------------------------------------------------------------------
template <class M>
__global__ __launch_bounds__(NT) void k_direction(const RowTile* __restrict__ tiles, const double* __restrict__ zv,
                                                  M* __restrict__ d, const GroupState* gs)
{
    const RowTile tl = tiles[blockIdx.x];
    const GroupState& s = gs[tl.group];
    if (s.done)
        return;
    const double beta = s.beta;
    const int row = tl.rowBegin + threadIdx.x;
    if (row >= tl.rowEnd)
        return;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        d[3 * row + i] = static_cast<M>(zv[3 * row + i] + static_cast<double>(d[3 * row + i]) * beta);
}

// ---------------------------------------------------------------------------
// integration
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_integrate(const RowTile* __restrict__ tiles, const int* __restrict__ rowNode,
                                                  const double* __restrict__ dvel, double* __restrict__ x,
                                                  double* __restrict__ v, double dt, GroupState* gs, int commit)
{
    const RowTile tl = tiles[blockIdx.x];
    GroupState& s = gs[tl.group];
    if (s.failed || (commit && s.nonfinite))
        return;
    const int row = tl.rowBegin + threadIdx.x;
    if (row >= tl.rowEnd)
        return;
    const int n = rowNode[row];
    double nv[3], nx[3];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const double xo = x[3 * n + i], vo = v[3 * n + i], dd = dvel[3 * row + i];
        nv[i] = vo + dd;
        nx[i] = xo + nv[i] * dt;
        ok = ok && fin(xo) && fin(vo) && fin(dd) && fin(nv[i]) && fin(nx[i]);
    }
    if (!commit)
    {
        if (!ok)
            s.nonfinite = 1;
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        v[3 * n + i] = nv[i];
        x[3 * n + i] = nx[i];
    }
}

// Explicit Euler on dynamic nodes (World.cpp:697-725).
__global__ __launch_bounds__(NT) void k_explicit(const RowTile* __restrict__ tiles, const int* __restrict__ rowNode,
                                                 const int* __restrict__ nstart, const int* __restrict__ nlist,
                                                 const double* __restrict__ elem, const double* __restrict__ mrow,
                                                 double gx, double gy, double gz, double dt,
                                                 double* __restrict__ x, double* __restrict__ v)
{
    const RowTile tl = tiles[blockIdx.x];
    const int row = tl.rowBegin + threadIdx.x;
    if (row >= tl.rowEnd)
        return;
    double f[3] = {0, 0, 0};
    for (int k = nstart[row]; k < nstart[row + 1]; ++k)
    {
        const int c = nlist[k];
        const double* e = elem + static_cast<size_t>(c >> 2) * E_STRIDE + E_FORCE + 3 * (c & 3);
        if (fin(e[0]) && fin(e[1]) && fin(e[2]))
        {
            f[0] += e[0];
            f[1] += e[1];
            f[2] += e[2];
        }
    }
    const double m = mrow[row];
    const double gv[3] = {gx, gy, gz};
    const int n = rowNode[row];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const double acc = (f[i] + gv[i] * m) / m;
        v[3 * n + i] = v[3 * n + i] + acc * dt;
        x[3 * n + i] = x[3 * n + i] + v[3 * n + i] * dt;
    }
}

// y = A x on the full BSR (inspection hook).
template <class M>
__global__ __launch_bounds__(NT) void k_spmv_plain(int nd, const int* __restrict__ fstart,
                                                   const int* __restrict__ fcol, const M* __restrict__ vals,
                                                   long long nnzf, const double* __restrict__ xin,
                                                   double* __restrict__ y)
{
    const int row = blockIdx.x * NT + threadIdx.x;
    if (row >= nd)
        return;
    const size_t F = static_cast<size_t>(nnzf);
    double acc[3] = {0, 0, 0};
    for (int k = fstart[row]; k < fstart[row + 1]; ++k)
    {
        const int c = fcol[k];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            acc[i] += static_cast<double>(vals[(3 * i) * F + k]) * xin[3 * c] +
                      static_cast<double>(vals[(3 * i + 1) * F + k]) * xin[3 * c + 1] +
                      static_cast<double>(vals[(3 * i + 2) * F + k]) * xin[3 * c + 2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        y[3 * row + i] = acc[i];
}

// ---------------------------------------------------------------------------
// host runtime
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
    int alloc(size_t count)
    {
        if (p)
        {
            (void)hipFree(p);
            p = nullptr;
        }
        n = count;
        if (count == 0)
            return PK_OK;
        PK_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        return PK_OK;
    }
    int upload(const std::vector<T>& h, hipStream_t st)
    {
        int rc = alloc(h.size());
        if (rc != PK_OK)
            return rc;
        if (!h.empty())
            PK_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        return PK_OK;
    }
};

struct TileTable
{
    DevBuf<RowTile> tiles;
    DevBuf<int> wgBegin, wgCount;
    DevBuf<Partial> part;
    int count = 0;
};

struct Profile
{
    bool on = false;
    long long launches[PK_K_COUNT] = {};
    double ms[PK_K_COUNT] = {};
    std::vector<hipEvent_t> pool;
    int used = 0;
    struct Rec
    {
        int kernel, iter;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
};

}  // namespace pk

struct pk_system
{
    pk::Topology topo;
    int precision = PK_FP64;
    int lpr = 16;
    int spmvRows = 64;
    hipStream_t stream = nullptr;
    int device = 0;

    pk::DevBuf<int4> tets;
    pk::DevBuf<unsigned char> live;
    pk::DevBuf<double> prm, x, v, elem, mrow, minv, xs, r, z;
    pk::DevBuf<int> cstart, clist, uFull, uMirror, slotRow, rowGroup, fstart, fcol, nstart, nlist, rowNode;
    pk::DevBuf<char> vals, d, q;  // precision-dependent element type
    pk::DevBuf<pk::GroupState> gs;
    pk::DevBuf<int> active;
    pk::TileTable rowTiles, spmvTiles;
    int* pinned = nullptr;
    hipEvent_t pollEv[2] = {nullptr, nullptr};

    double gravity[3] = {0, 0, 0};
    double tol = 1e-4;
    int maxIt = 128;
    int jacobi = 1;
    int implicit = 1;
    int lastIters = 0;
    std::vector<pk::GroupState> hostGs;
    pk::Profile prof;

    ~pk_system()
    {
        for (auto e : prof.pool)
            (void)hipEventDestroy(e);
        for (auto e : pollEv)
            if (e)
                (void)hipEventDestroy(e);
        if (pinned)
            (void)hipHostFree(pinned);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace pk
{

static int build_tiles(pk_system* s, TileTable& tt, int rowsPerWg)
{
    const Topology& t = s->topo;
    std::vector<RowTile> tiles;
    std::vector<int> wb(static_cast<size_t>(t.G), 0), wc(static_cast<size_t>(t.G), 0);
    for (int g = 0; g < t.G; ++g)
    {
        wb[g] = static_cast<int>(tiles.size());
        for (int r = t.groupRowBegin[g]; r < t.groupRowEnd[g]; r += rowsPerWg)
            tiles.push_back(RowTile{g, r, std::min(r + rowsPerWg, t.groupRowEnd[g]), 0});
        wc[g] = static_cast<int>(tiles.size()) - wb[g];
    }
    tt.count = static_cast<int>(tiles.size());
    int rc;
    if ((rc = tt.tiles.upload(tiles, s->stream)) != PK_OK) return rc;
    if ((rc = tt.wgBegin.upload(wb, s->stream)) != PK_OK) return rc;
    if ((rc = tt.wgCount.upload(wc, s->stream)) != PK_OK) return rc;
    if ((rc = tt.part.alloc(std::max(1, tt.count))) != PK_OK) return rc;
    return PK_OK;
}

static int upload_state_groups(pk_system* s)
{
    const Topology& t = s->topo;
    s->hostGs.assign(static_cast<size_t>(t.G), GroupState{});
    for (int g = 0; g < t.G; ++g)
    {
        GroupState& st = s->hostGs[g];
        std::memset(&st, 0, sizeof(st));
        st.hasRows = t.groupRowEnd[g] > t.groupRowBegin[g] ? 1 : 0;
        st.tol = s->tol;
        st.maxIt = s->maxIt;
        st.jacobi = s->jacobi;
        st.done = 1;
    }
    int rc = s->gs.upload(s->hostGs, s->stream);
    if (rc != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

template <class M>
static size_t esize()
{
    return sizeof(M);
}

// Event-bracketed launch when profiling.
struct Bracket
{
    pk_system* s;
    int kernel, iter;
    hipEvent_t a = nullptr;
    Bracket(pk_system* sys, int k, int it = -1) : s(sys), kernel(k), iter(it)
    {
        if (!s->prof.on)
            return;
        a = take();
        (void)hipEventRecord(a, s->stream);
    }
    hipEvent_t take()
    {
        Profile& p = s->prof;
        if (p.used == static_cast<int>(p.pool.size()))
        {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            p.pool.push_back(e);
        }
        return p.pool[p.used++];
    }
    ~Bracket()
    {
        if (!s->prof.on)
            return;
        hipEvent_t b = take();
        (void)hipEventRecord(b, s->stream);
        s->prof.recs.push_back(Profile::Rec{kernel, iter, a, b});
    }
};

static void harvest_profile(pk_system* s, int realIters)
{
    Profile& p = s->prof;
    if (!p.on)
        return;
    (void)hipStreamSynchronize(s->stream);
    for (const auto& rec : p.recs)
    {
        // CG launches after every group finished did no work
        if (rec.iter >= 0)
        {
            const int limit = rec.kernel == PK_K_DIRECTION ? realIters - 1 : realIters;
            if (rec.iter >= limit)
                continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, rec.a, rec.b) == hipSuccess)
        {
            p.ms[rec.kernel] += ms;
            p.launches[rec.kernel] += 1;
        }
    }
    p.recs.clear();
    p.used = 0;
}

template <class M>
static int launch_spmv(pk_system* s, int it)
{
    Bracket br(s, PK_K_SPMV, it);
    const dim3 grid(s->spmvTiles.count), block(NT);
    auto* vals = reinterpret_cast<const M*>(s->vals.p);
    auto* d = reinterpret_cast<const M*>(s->d.p);
    auto* q = reinterpret_cast<M*>(s->q.p);
    const long long F = s->topo.nnzf();
    switch (s->lpr)
    {
    case 4:
        hipLaunchKernelGGL((k_spmv<M, 4>), grid, block, 0, s->stream, s->spmvTiles.tiles.p, s->spmvTiles.wgBegin.p,
                           s->spmvTiles.wgCount.p, s->fstart.p, s->fcol.p, vals, F, d, q, s->spmvTiles.part.p,
                           s->gs.p, s->active.p);
        break;
    case 8:
        hipLaunchKernelGGL((k_spmv<M, 8>), grid, block, 0, s->stream, s->spmvTiles.tiles.p, s->spmvTiles.wgBegin.p,
                           s->spmvTiles.wgCount.p, s->fstart.p, s->fcol.p, vals, F, d, q, s->spmvTiles.part.p,
                           s->gs.p, s->active.p);
        break;
    case 32:
        hipLaunchKernelGGL((k_spmv<M, 32>), grid, block, 0, s->stream, s->spmvTiles.tiles.p, s->spmvTiles.wgBegin.p,
                           s->spmvTiles.wgCount.p, s->fstart.p, s->fcol.p, vals, F, d, q, s->spmvTiles.part.p,
                           s->gs.p, s->active.p);
        break;
    default:
        hipLaunchKernelGGL((k_spmv<M, 16>), grid, block, 0, s->stream, s->spmvTiles.tiles.p, s->spmvTiles.wgBegin.p,
                           s->spmvTiles.wgCount.p, s->fstart.p, s->fcol.p, vals, F, d, q, s->spmvTiles.part.p,
                           s->gs.p, s->active.p);
        break;
    }
    return PK_OK;
}

template <class M>
static int run_assembly(pk_system* s, double dt)
{
    const Topology& t = s->topo;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(NT), 0, s->stream, s->gs.p, t.G, s->active.p);
    if (t.T > 0)
    {
        Bracket br(s, PK_K_ELEMENT);
        hipLaunchKernelGGL(k_element, dim3((t.T + NT - 1) / NT), dim3(NT), 0, s->stream, t.T, s->tets.p, s->live.p,
                           s->prm.p, s->x.p, s->v.p, s->elem.p);
    }
    if (t.nd == 0)
        return PK_OK;
    const int nnzu = static_cast<int>(t.nnzu());
    {
        Bracket br(s, PK_K_ASSEMBLE);
        hipLaunchKernelGGL((k_assemble<M>), dim3((nnzu + NT - 1) / NT), dim3(NT), 0, s->stream, nnzu, s->cstart.p,
                           s->clist.p, s->uFull.p, s->uMirror.p, s->slotRow.p, s->rowGroup.p, s->elem.p, s->mrow.p,
                           dt * dt, reinterpret_cast<M*>(s->vals.p), t.nnzf(), s->minv.p, t.nd, s->gs.p);
    }
    {
        Bracket br(s, PK_K_RHS);
        hipLaunchKernelGGL((k_rhs<M>), dim3(s->rowTiles.count), dim3(NT), 0, s->stream, s->rowTiles.tiles.p,
                           s->rowTiles.wgBegin.p, s->rowTiles.wgCount.p, s->nstart.p, s->nlist.p, s->elem.p,
                           s->mrow.p, s->gravity[0], s->gravity[1], s->gravity[2], dt, dt * dt, s->minv.p, t.nd,
                           s->xs.p, s->r.p, s->z.p, reinterpret_cast<M*>(s->d.p), s->rowTiles.part.p, s->gs.p,
                           s->active.p);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

template <class M>
static int run_cg(pk_system* s)
{
    const int maxIt = std::max(0, s->maxIt);
    int launched = 0, poll = 0;
    // the first chunk reaches close to the previous solve's count; later
    // chunks are short so few launches run after convergence
    int chunk = std::max(4, s->lastIters - 4);
    while (launched < maxIt)
    {
        const int n = std::min(chunk, maxIt - launched);
        for (int i = 0; i < n; ++i)
        {
            const int it = launched + i;
            int rc = launch_spmv<M>(s, it);
            if (rc != PK_OK)
                return rc;
            {
                Bracket br(s, PK_K_UPDATE, it);
                hipLaunchKernelGGL((k_update<M>), dim3(s->rowTiles.count), dim3(NT), 0, s->stream,
                                   s->rowTiles.tiles.p, s->rowTiles.wgBegin.p, s->rowTiles.wgCount.p, s->minv.p,
                                   s->topo.nd, reinterpret_cast<const M*>(s->d.p), reinterpret_cast<const M*>(s->q.p),
                                   s->xs.p, s->r.p, s->z.p, s->rowTiles.part.p, s->gs.p, s->active.p);
            }
            {
                Bracket br(s, PK_K_DIRECTION, it);
                hipLaunchKernelGGL((k_direction<M>), dim3(s->rowTiles.count), dim3(NT), 0, s->stream,
                                   s->rowTiles.tiles.p, s->z.p, reinterpret_cast<M*>(s->d.p), s->gs.p);
            }
        }
        launched += n;
        chunk = 4;
        const int slot = poll & 1;
        PK_HIP(hipMemcpyAsync(s->pinned + slot, s->active.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        PK_HIP(hipEventRecord(s->pollEv[slot], s->stream));
        if (poll > 0)
        {
            PK_HIP(hipEventSynchronize(s->pollEv[slot ^ 1]));
            if (s->pinned[slot ^ 1] == 0)
                break;
        }
        ++poll;
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

template <class M>
static int run_substep(pk_system* s, double dt)
{
    int rc = run_assembly<M>(s, dt);
    if (rc != PK_OK || s->topo.nd == 0)
        return rc;
    if (s->implicit)
    {
        if ((rc = run_cg<M>(s)) != PK_OK)
            return rc;
        Bracket br(s, PK_K_INTEGRATE);
        hipLaunchKernelGGL(k_integrate, dim3(s->rowTiles.count), dim3(NT), 0, s->stream, s->rowTiles.tiles.p,
                           s->rowNode.p, s->xs.p, s->x.p, s->v.p, dt, s->gs.p, 0);
        hipLaunchKernelGGL(k_integrate, dim3(s->rowTiles.count), dim3(NT), 0, s->stream, s->rowTiles.tiles.p,
                           s->rowNode.p, s->xs.p, s->x.p, s->v.p, dt, s->gs.p, 1);
    }
    else
    {
        Bracket br(s, PK_K_INTEGRATE);
        hipLaunchKernelGGL(k_explicit, dim3(s->rowTiles.count), dim3(NT), 0, s->stream, s->rowTiles.tiles.p,
                           s->rowNode.p, s->nstart.p, s->nlist.p, s->elem.p, s->mrow.p, s->gravity[0],
                           s->gravity[1], s->gravity[2], dt, s->x.p, s->v.p);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

static int read_groups(pk_system* s)
{
    if (s->topo.G == 0)
        return PK_OK;
    PK_HIP(hipMemcpyAsync(s->hostGs.data(), s->gs.p, s->hostGs.size() * sizeof(GroupState), hipMemcpyDeviceToHost,
                          s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    int mx = 0;
    for (const auto& g : s->hostGs)
        mx = std::max(mx, g.iters);
    s->lastIters = mx;
    return PK_OK;
}

static int choose_lpr(const Topology& t)
{
    if (t.nd == 0)
        return 16;
    const double avg = static_cast<double>(t.nnzf()) / t.nd;
    if (avg <= 6.0)
        return 8;
    if (avg <= 18.0)
        return 16;
    return 32;
}

static int setup_device(pk_system* s)
{
    const Topology& t = s->topo;
    hipStream_t st = s->stream;
    int rc;
    std::vector<int4> tv(static_cast<size_t>(t.T));
    for (int i = 0; i < t.T; ++i)
        tv[i] = make_int4(t.tets[4 * i], t.tets[4 * i + 1], t.tets[4 * i + 2], t.tets[4 * i + 3]);
    if ((rc = s->tets.upload(tv, st)) != PK_OK) return rc;
    std::vector<unsigned char> live(t.live.begin(), t.live.end());
    if ((rc = s->live.upload(live, st)) != PK_OK) return rc;
    if ((rc = s->prm.upload(t.prm, st)) != PK_OK) return rc;
    std::vector<double> mrow(static_cast<size_t>(t.nd));
    for (int r = 0; r < t.nd; ++r)
        mrow[r] = t.mass[t.rowNode[r]];
    if ((rc = s->mrow.upload(mrow, st)) != PK_OK) return rc;
    if ((rc = s->cstart.upload(t.cstart, st)) != PK_OK) return rc;
    if ((rc = s->clist.upload(t.clist, st)) != PK_OK) return rc;
    if ((rc = s->uFull.upload(t.uFull, st)) != PK_OK) return rc;
    if ((rc = s->uMirror.upload(t.uMirror, st)) != PK_OK) return rc;
    std::vector<int> slotRow(static_cast<size_t>(t.nnzu())), rowGroup(static_cast<size_t>(t.nd));
    for (int r = 0; r < t.nd; ++r)
        for (int k = t.ustart[r]; k < t.ustart[r + 1]; ++k)
            slotRow[k] = r;
    for (int g = 0; g < t.G; ++g)
        for (int r = t.groupRowBegin[g]; r < t.groupRowEnd[g]; ++r)
            rowGroup[r] = g;
    if ((rc = s->slotRow.upload(slotRow, st)) != PK_OK) return rc;
    if ((rc = s->rowGroup.upload(rowGroup, st)) != PK_OK) return rc;
    if ((rc = s->fstart.upload(t.fstart, st)) != PK_OK) return rc;
    if ((rc = s->fcol.upload(t.fcol, st)) != PK_OK) return rc;
    if ((rc = s->nstart.upload(t.nstart, st)) != PK_OK) return rc;
    if ((rc = s->nlist.upload(t.nlist, st)) != PK_OK) return rc;
    if ((rc = s->rowNode.upload(t.rowNode, st)) != PK_OK) return rc;
    const size_t es = s->precision == PK_FP32_STORAGE ? sizeof(float) : sizeof(double);
    if ((rc = s->elem.alloc(static_cast<size_t>(t.T) * E_STRIDE)) != PK_OK) return rc;
    if ((rc = s->vals.alloc(static_cast<size_t>(t.nnzf()) * 9 * es)) != PK_OK) return rc;
    if ((rc = s->d.alloc(static_cast<size_t>(t.nd) * 3 * es)) != PK_OK) return rc;
    if ((rc = s->q.alloc(static_cast<size_t>(t.nd) * 3 * es)) != PK_OK) return rc;
    if ((rc = s->minv.alloc(static_cast<size_t>(t.nd) * 9)) != PK_OK) return rc;
    if ((rc = s->xs.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    if ((rc = s->r.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    if ((rc = s->z.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    if ((rc = s->active.alloc(1)) != PK_OK) return rc;
    // initial state: rest positions, zero velocity
    std::vector<double> zero(static_cast<size_t>(t.N) * 3, 0.0);
    if ((rc = s->v.upload(zero, st)) != PK_OK) return rc;
    if ((rc = s->x.alloc(static_cast<size_t>(t.N) * 3)) != PK_OK) return rc;
    s->lpr = choose_lpr(t);
    s->spmvRows = (NT / s->lpr) * 4;
    if ((rc = build_tiles(s, s->rowTiles, NT)) != PK_OK) return rc;
    if ((rc = build_tiles(s, s->spmvTiles, s->spmvRows)) != PK_OK) return rc;
    if ((rc = upload_state_groups(s)) != PK_OK) return rc;
    return PK_OK;
}

}  // namespace pk

using namespace pk;

extern "C" {

const char* pk_last_error(void) { return g_error.c_str(); }

int pk_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

int pk_system_create(const pk_mesh_desc* desc, int precision, pk_system** out)
{
    if (!desc || !out || (precision != PK_FP64 && precision != PK_FP32_STORAGE))
    {
        set_error("pk_system_create: bad argument");
        return PK_ERR_ARG;
    }
    *out = nullptr;
    if (pk_device_count() <= 0)
    {
        set_error("pk_system_create: no GPU visible");
        return PK_ERR_NO_DEVICE;
    }
    std::unique_ptr<pk_system> s(new pk_system());
    s->precision = precision;
    std::string err;
    int rc = build_topology(*desc, s->topo, err, false);
    if (rc != PK_OK)
    {
        set_error(err);
        return rc;
    }
    PK_HIP(hipGetDevice(&s->device));
    PK_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    PK_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->pinned), 2 * sizeof(int), hipHostMallocDefault));
    PK_HIP(hipEventCreateWithFlags(&s->pollEv[0], hipEventDisableTiming));
    PK_HIP(hipEventCreateWithFlags(&s->pollEv[1], hipEventDisableTiming));
    if ((rc = setup_device(s.get())) != PK_OK)
        return rc;
    std::vector<double> rest(desc->rest, desc->rest + 3 * static_cast<size_t>(desc->node_count));
    PK_HIP(hipMemcpyAsync(s->x.p, rest.data(), rest.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    *out = s.release();
    return PK_OK;
}

void pk_system_destroy(pk_system* sys) { delete sys; }

int pk_system_sizes(const pk_system* s, int* nodes, int* tets, int* groups, int* rows, long long* nnzu,
                    long long* nnzf)
{
    if (!s)
        return PK_ERR_ARG;
    if (nodes) *nodes = s->topo.N;
    if (tets) *tets = s->topo.T;
    if (groups) *groups = s->topo.G;
    if (rows) *rows = s->topo.nd;
    if (nnzu) *nnzu = s->topo.nnzu();
    if (nnzf) *nnzf = s->topo.nnzf();
    return PK_OK;
}

int pk_system_set_state(pk_system* s, const double* x, const double* v)
{
    if (!s)
        return PK_ERR_ARG;
    const size_t bytes = static_cast<size_t>(s->topo.N) * 3 * sizeof(double);
    if (x && bytes)
        PK_HIP(hipMemcpyAsync(s->x.p, x, bytes, hipMemcpyHostToDevice, s->stream));
    if (v && bytes)
        PK_HIP(hipMemcpyAsync(s->v.p, v, bytes, hipMemcpyHostToDevice, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_get_state(pk_system* s, double* x, double* v)
{
    if (!s)
        return PK_ERR_ARG;
    const size_t bytes = static_cast<size_t>(s->topo.N) * 3 * sizeof(double);
    if (x && bytes)
        PK_HIP(hipMemcpyAsync(x, s->x.p, bytes, hipMemcpyDeviceToHost, s->stream));
    if (v && bytes)
        PK_HIP(hipMemcpyAsync(v, s->v.p, bytes, hipMemcpyDeviceToHost, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_set_gravity(pk_system* s, double gx, double gy, double gz)
{
    if (!s)
        return PK_ERR_ARG;
    s->gravity[0] = gx;
    s->gravity[1] = gy;
    s->gravity[2] = gz;
    return PK_OK;
}

int pk_system_set_solver(pk_system* s, double tolerance, int max_iterations, int use_jacobi)
{
    if (!s)
        return PK_ERR_ARG;
    s->tol = tolerance;
    s->maxIt = max_iterations;
    s->jacobi = use_jacobi ? 1 : 0;
    return upload_state_groups(s);
}

int pk_system_set_mode(pk_system* s, int implicit)
{
    if (!s)
        return PK_ERR_ARG;
    s->implicit = implicit ? 1 : 0;
    return PK_OK;
}

int pk_system_set_young(pk_system* s, const double* young)
{
    if (!s || !young)
        return PK_ERR_ARG;
    Topology& t = s->topo;
    if (t.T == 0)
        return PK_OK;
    const size_t T = static_cast<size_t>(t.T);
    for (size_t i = 0; i < T; ++i)
        lame(young[i], t.poisson[i], &t.prm[P_LAMBDA * T + i], &t.prm[P_MU * T + i]);
    // lambda and mu are adjacent rows of the parameter block
    PK_HIP(hipMemcpyAsync(s->prm.p + P_LAMBDA * T, t.prm.data() + P_LAMBDA * T, 2 * T * sizeof(double),
                          hipMemcpyHostToDevice, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_step(pk_system* s, double dt, int substeps, int frames)
{
    if (!s || substeps < 1 || frames < 0)
    {
        set_error("pk_system_step: bad argument");
        return PK_ERR_ARG;
    }
    if (s->topo.massErrorNode >= 0)
    {
        set_error("Non-fixed node has no physical mass contribution. nodeIndex=" +
                  std::to_string(s->topo.massErrorNode));
        return PK_ERR_MASS;
    }
    if (!(dt > 0.0))
        return PK_OK;  // World::Step does nothing for dt <= 0
    const double h = dt / substeps;
    for (int f = 0; f < frames; ++f)
        for (int k = 0; k < substeps; ++k)
        {
            const int rc = s->precision == PK_FP32_STORAGE ? run_substep<float>(s, h) : run_substep<double>(s, h);
            if (rc != PK_OK)
                return rc;
            // read the CG outcome once per substep so the next solve's first
            // chunk is sized from this one
            if (s->implicit && s->topo.nd > 0)
            {
                const int rc2 = read_groups(s);
                if (rc2 != PK_OK)
                    return rc2;
                harvest_profile(s, s->lastIters + 1);
            }
        }
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_group_results(pk_system* s, int* iterations, double* residual, int* converged, int* status)
{
    if (!s)
        return PK_ERR_ARG;
    for (int g = 0; g < s->topo.G; ++g)
    {
        const GroupState& st = s->hostGs[g];
        if (iterations) iterations[g] = st.iters;
        if (residual) residual[g] = st.residual;
        if (converged) converged[g] = st.converged;
        if (status)
            status[g] = !st.hasRows ? PK_GROUP_NO_DYNAMIC
                        : st.failed ? PK_GROUP_SOLVE_FAILED
                        : st.nonfinite ? PK_GROUP_NONFINITE
                                       : PK_GROUP_INTEGRATED;
    }
    return PK_OK;
}

int pk_system_profile(pk_system* s, int enable)
{
    if (!s)
        return PK_ERR_ARG;
    s->prof.on = enable != 0;
    for (int k = 0; k < PK_K_COUNT; ++k)
    {
        s->prof.launches[k] = 0;
        s->prof.ms[k] = 0.0;
    }
    s->prof.recs.clear();
    s->prof.used = 0;
    return PK_OK;
}

int pk_system_kernel_time(pk_system* s, int kernel, long long* launches, double* total_ms)
{
    if (!s || kernel < 0 || kernel >= PK_K_COUNT)
        return PK_ERR_ARG;
    if (launches) *launches = s->prof.launches[kernel];
    if (total_ms) *total_ms = s->prof.ms[kernel];
    return PK_OK;
}

int pk_system_assemble(pk_system* s, double dt)
{
    if (!s)
        return PK_ERR_ARG;
    const int rc = s->precision == PK_FP32_STORAGE ? run_assembly<float>(s, dt) : run_assembly<double>(s, dt);
    if (rc != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    return read_groups(s);
}

int pk_system_get_matrix(pk_system* s, double* full_vals, double* rhs, double* minv)
{
    if (!s)
        return PK_ERR_ARG;
    const Topology& t = s->topo;
    const size_t F = static_cast<size_t>(t.nnzf());
    if (full_vals && F)
    {
        const size_t es = s->precision == PK_FP32_STORAGE ? 4 : 8;
        std::vector<char> raw(F * 9 * es);
        PK_HIP(hipMemcpy(raw.data(), s->vals.p, raw.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < F; ++k)
            for (int e = 0; e < 9; ++e)
                full_vals[9 * k + e] = es == 4 ? reinterpret_cast<const float*>(raw.data())[e * F + k]
                                               : reinterpret_cast<const double*>(raw.data())[e * F + k];
    }
    if (rhs && t.nd)
        PK_HIP(hipMemcpy(rhs, s->r.p, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (minv && t.nd)
    {
        std::vector<double> raw(static_cast<size_t>(t.nd) * 9);
        PK_HIP(hipMemcpy(raw.data(), s->minv.p, raw.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int r = 0; r < t.nd; ++r)
            for (int e = 0; e < 9; ++e)
                minv[9 * r + e] = raw[static_cast<size_t>(e) * t.nd + r];
    }
    return PK_OK;
}

int pk_system_get_elements(pk_system* s, double* blocks, double* forces, double* kv)
{
    if (!s)
        return PK_ERR_ARG;
    const Topology& t = s->topo;
    std::vector<double> raw(static_cast<size_t>(t.T) * E_STRIDE);
    if (!raw.empty())
        PK_HIP(hipMemcpy(raw.data(), s->elem.p, raw.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < t.T; ++i)
    {
        const double* e = raw.data() + static_cast<size_t>(i) * E_STRIDE;
        const bool lv = t.live[i] != 0;
        for (int k = 0; k < 90; ++k)
            if (blocks) blocks[90 * static_cast<size_t>(i) + k] = lv ? e[E_BLOCKS + k] : 0.0;
        for (int k = 0; k < 12; ++k)
        {
            if (forces) forces[12 * static_cast<size_t>(i) + k] = lv ? e[E_FORCE + k] : 0.0;
            if (kv) kv[12 * static_cast<size_t>(i) + k] = lv ? e[E_KV + k] : 0.0;
        }
    }
    return PK_OK;
}

int pk_system_spmv(pk_system* s, const double* x, double* y)
{
    if (!s)
        return PK_ERR_ARG;
    const Topology& t = s->topo;
    if (t.nd == 0)
        return PK_OK;
    DevBuf<double> dx, dy;
    int rc;
    if ((rc = dx.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    if ((rc = dy.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    PK_HIP(hipMemcpy(dx.p, x, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyHostToDevice));
    if (s->precision == PK_FP32_STORAGE)
        hipLaunchKernelGGL((k_spmv_plain<float>), dim3((t.nd + NT - 1) / NT), dim3(NT), 0, s->stream, t.nd,
                           s->fstart.p, s->fcol.p, reinterpret_cast<const float*>(s->vals.p), t.nnzf(), dx.p, dy.p);
    else
        hipLaunchKernelGGL((k_spmv_plain<double>), dim3((t.nd + NT - 1) / NT), dim3(NT), 0, s->stream, t.nd,
                           s->fstart.p, s->fcol.p, reinterpret_cast<const double*>(s->vals.p), t.nnzf(), dx.p,
                           dy.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(s->stream));
    PK_HIP(hipMemcpy(y, dy.p, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return PK_OK;
}

int pk_system_get_solution(pk_system* s, double* dv)
{
    if (!s || !dv)
        return PK_ERR_ARG;
    if (s->topo.nd)
        PK_HIP(hipMemcpy(dv, s->xs.p, static_cast<size_t>(s->topo.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return PK_OK;
}

}  // extern "C"
