// Block-Jacobi PCG kernels for gfx950 (ConjugateGradientSolver.cpp:114-241).
//
// Iteration k is two launches:
//   k_spmv(k)    every workgroup first sums the (r.z, r.r) partials the
//                previous launch left for its body, takes the reference's
//                stop decisions for the end of iteration k-1 and forms
//                beta = delta_k / delta_{k-1}; it then applies
//                d_k = z_k + beta d_{k-1} on the fly while computing
//                q = A d_k on the full BSR (both triangles), stores d_k for its
//                own rows and leaves a d.q partial.
//   k_update(k)  every workgroup sums the body's d.q partials, alpha =
//                delta_k / d.q (or stops on d.q <= 0), updates x += alpha d,
//                r -= alpha q, z = M^-1 r and leaves (r.z, r.r) partials.
// Totals are recomputed by every workgroup from the same partials in the same
// order, so no workgroup ever waits on another and no device-scope fence is
// needed inside a launch. One designated workgroup per body records the
// decisions (delta, stop, iterations, residual) for later launches and the
// host; the host polls the count of bodies still iterating.
#include "pk_device.h"

namespace pk
{

// Records the end of a body's CG loop (designated workgroup only).
__device__ inline void record_stop(GroupState& s, int* active, int k, int inUpdate, int iters, bool converged,
                                   double residual, bool failed)
{
    s.stopAt = k;
    s.stopInUpdate = inUpdate;
    s.iters = iters;
    s.converged = converged ? 1 : 0;
    s.residual = residual;
    if (failed)
        s.failed = 1;
    atomicSub(active, 1);
}

// Reads a group field that another workgroup of the same launch may be
// writing, once, so the whole workgroup takes the same branch.
__device__ inline int read_uniform(const int* p)
{
    __shared__ int v;
    if (threadIdx.x == 0)
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return v;
}

template <class M>
__device__ inline M direction(M z, M dprev, double beta)
{
    return static_cast<M>(__builtin_fma(beta, static_cast<double>(dprev), static_cast<double>(z)));
}

template <class M, int LPR>
__global__ __launch_bounds__(NT) void k_spmv(int k, TileView st, TileView vt, const int* __restrict__ fstart,
                                             const int* __restrict__ fcol, const M* __restrict__ vals,
                                             long long nnzf, const M* __restrict__ dprev, M* __restrict__ dcur,
                                             const M* __restrict__ z, M* __restrict__ q, double* __restrict__ pa,
                                             const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    const int t = tile_of(st);
    if (t < 0)
        return;
    const RowTile tl = st.tiles[t];
    const int g = tl.group;
    GroupState& s = gs[g];
    if (read_uniform(&s.stopAt) >= 0)
        return;
    double rz, rr;
    int flags;
    sum_partials(pbPrev, vt.groupBegin[g], vt.groupCount[g], rz, rr, flags);
    const bool designated = t == st.groupBegin[g] && threadIdx.x == 0;

    // Stop decisions for the end of iteration k-1 (k = 0: the CG start).
    double beta = 0.0;
    bool stop = false, conv = false, failed = false;
    double resid = 0.0;
    int iters = k;
    if (k == 0)
    {
        double target = 0.0;
        if (s.failed || (flags & 1) || s.maxIt <= 0 || !fin(s.tol) || !(s.tol > 0.0))
            stop = failed = true;  // rejected: dv cleared, nothing integrated
        else if ((flags & 2) || !fin(rz) || rz < 0.0)
            stop = true;
        else
        {
            target = s.tol * s.tol * rz;
            if (!fin(target))
                stop = true;
            else if (rz <= target)
            {
                stop = conv = true;
                resid = norm_of(rr);
            }
        }
        if (designated && !stop)
        {
            s.delta[0] = rz;
            s.target = target;
        }
    }
    else
    {
        const double old = s.delta[(k - 1) & 1];
        resid = norm_of(rr);
        if (flags)
        {
            stop = true;  // M^-1 r was not finite: break before delta changes
            iters = k - 1;
            conv = old <= s.target;
        }
        else
        {
            beta = rz / old;
            if (!fin(rz) || !fin(old) || old <= 0.0 || !fin(beta))
            {
                stop = true;
                iters = k - 1;
                conv = rz <= s.target;
            }
            else if (rz <= s.target)
                stop = conv = true;
            else if (k >= s.maxIt)
                stop = true;
        }
        if (designated && !stop)
            s.delta[k & 1] = rz;
    }
    if (stop)
    {
        if (designated)
            record_stop(s, active, k, 0, iters, conv, resid, failed);
        return;
    }

    const int lane = threadIdx.x % LPR;
    const int sub = threadIdx.x / LPR;
    constexpr int RPP = NT / LPR;
    const size_t F = static_cast<size_t>(nnzf);
    double pd = 0.0;
    for (int row = tl.rowBegin + sub; row < tl.rowEnd; row += RPP)
    {
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        const int k1 = fstart[row + 1];
        for (int kk = fstart[row] + lane; kk < k1; kk += LPR)
        {
            const int c = fcol[kk];
            M m0, m1, m2;
            if (k == 0)
            {
                m0 = dcur[3 * c];
                m1 = dcur[3 * c + 1];
                m2 = dcur[3 * c + 2];
            }
            else
            {
                m0 = direction(z[3 * c], dprev[3 * c], beta);
                m1 = direction(z[3 * c + 1], dprev[3 * c + 1], beta);
                m2 = direction(z[3 * c + 2], dprev[3 * c + 2], beta);
            }
            const double d0 = m0, d1 = m1, d2 = m2;
            const M* a = vals + kk;
            y0 += static_cast<double>(__builtin_nontemporal_load(a)) * d0 +
                  static_cast<double>(__builtin_nontemporal_load(a + F)) * d1 +
                  static_cast<double>(__builtin_nontemporal_load(a + 2 * F)) * d2;
            y1 += static_cast<double>(__builtin_nontemporal_load(a + 3 * F)) * d0 +
                  static_cast<double>(__builtin_nontemporal_load(a + 4 * F)) * d1 +
                  static_cast<double>(__builtin_nontemporal_load(a + 5 * F)) * d2;
            y2 += static_cast<double>(__builtin_nontemporal_load(a + 6 * F)) * d0 +
                  static_cast<double>(__builtin_nontemporal_load(a + 7 * F)) * d1 +
                  static_cast<double>(__builtin_nontemporal_load(a + 8 * F)) * d2;
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
            M r0, r1, r2;
            if (k == 0)
            {
                r0 = dcur[3 * row];
                r1 = dcur[3 * row + 1];
                r2 = dcur[3 * row + 2];
            }
            else
            {
                r0 = direction(z[3 * row], dprev[3 * row], beta);
                r1 = direction(z[3 * row + 1], dprev[3 * row + 1], beta);
                r2 = direction(z[3 * row + 2], dprev[3 * row + 2], beta);
                dcur[3 * row] = r0;
                dcur[3 * row + 1] = r1;
                dcur[3 * row + 2] = r2;
            }
            const M q0 = static_cast<M>(y0), q1 = static_cast<M>(y1), q2 = static_cast<M>(y2);
            q[3 * row] = q0;
            q[3 * row + 1] = q1;
            q[3 * row + 2] = q2;
            pd += static_cast<double>(r0) * static_cast<double>(q0) + static_cast<double>(r1) * static_cast<double>(q1) +
                  static_cast<double>(r2) * static_cast<double>(q2);
        }
    }
    double u1 = 0.0, u2 = 0.0;
    block_sum3(pd, u1, u2);
    if (threadIdx.x == 0)
        pa[t] = pd;
}

template <class M>
__global__ __launch_bounds__(NT) void k_update(int k, TileView vt, TileView st, const double* __restrict__ minv,
                                               int nd, const M* __restrict__ dcur, const M* __restrict__ q,
                                               double* __restrict__ xs, double* __restrict__ rv, M* __restrict__ zv,
                                               const double* __restrict__ pa, Partial* __restrict__ pbCur,
                                               const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    const int t = tile_of(vt);
    if (t < 0)
        return;
    const RowTile tl = vt.tiles[t];
    const int g = tl.group;
    GroupState& s = gs[g];
    const int stopAt = read_uniform(&s.stopAt);
    if (stopAt >= 0 && stopAt <= k)
        return;
    const double dq = sum_doubles(pa, st.groupBegin[g], st.groupCount[g]);
    const double delta = s.delta[k & 1];
    const double alpha = delta / dq;
    if (!fin(dq) || dq <= 0.0 || !fin(alpha))
    {
        // break before any update (ConjugateGradientSolver.cpp:192-202)
        double rz, rr;
        int fl;
        sum_partials(pbPrev, vt.groupBegin[g], vt.groupCount[g], rz, rr, fl);
        if (t == vt.groupBegin[g] && threadIdx.x == 0)
            record_stop(s, active, k, 1, k, delta <= s.target, norm_of(rr), false);
        return;
    }
    const size_t N = static_cast<size_t>(nd);
    double pz = 0.0, pr = 0.0;
    int flag = 0;
    for (int row = tl.rowBegin + threadIdx.x; row < tl.rowEnd; row += NT)
    {
        double r[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[3 * row + i] = xs[3 * row + i] + static_cast<double>(dcur[3 * row + i]) * alpha;
            r[i] = rv[3 * row + i] + static_cast<double>(q[3 * row + i]) * (-alpha);
            rv[3 * row + i] = r[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            z[i] = minv[(3 * i) * N + row] * r[0] + minv[(3 * i + 1) * N + row] * r[1] +
                   minv[(3 * i + 2) * N + row] * r[2];
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag = 1;
        M zs[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            zs[i] = static_cast<M>(z[i]);
            zv[3 * row + i] = zs[i];
        }
        pz += r[0] * static_cast<double>(zs[0]) + r[1] * static_cast<double>(zs[1]) + r[2] * static_cast<double>(zs[2]);
        pr += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    }
    double u = 0.0;
    block_sum3(pz, pr, u);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
        pbCur[t] = Partial{pz, pr, flag, 0};
}

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

template <class M>
void launch_spmv(const DeviceView& v, int k, int lanesPerRow)
{
    if (v.spmv.count == 0)
        return;
    const dim3 grid(8u * static_cast<unsigned>(v.spmv.per)), block(NT);
    const M* dprev = static_cast<const M*>(v.d[(k + 1) & 1]);
    M* dcur = static_cast<M*>(v.d[k & 1]);
    const M* z = static_cast<const M*>(v.z);
    M* q = static_cast<M*>(v.q);
    const M* vals = static_cast<const M*>(v.vals);
    const Partial* pbPrev = v.pb[(k + 1) & 1];
#define PK_SPMV(L)                                                                                              \
    hipLaunchKernelGGL((k_spmv<M, L>), grid, block, 0, v.stream, k, v.spmv, v.vec, v.fstart, v.fcol, vals,     \
                       v.nnzf, dprev, dcur, z, q, v.pa, pbPrev, v.gs, v.active)
    switch (lanesPerRow)
    {
    case 4: PK_SPMV(4); break;
    case 8: PK_SPMV(8); break;
    case 32: PK_SPMV(32); break;
    default: PK_SPMV(16); break;
    }
#undef PK_SPMV
}

template <class M>
void launch_update(const DeviceView& v, int k)
{
    if (v.vec.count == 0)
        return;
    hipLaunchKernelGGL((k_update<M>), dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, k, v.vec,
                       v.spmv, v.minv, v.nd, static_cast<const M*>(v.d[k & 1]), static_cast<const M*>(v.q), v.xs,
                       v.r, static_cast<M*>(v.z), v.pa, v.pb[k & 1], v.pb[(k + 1) & 1], v.gs, v.active);
}

template <class M>
void launch_spmv_plain(const DeviceView& v, const double* x, double* y)
{
    if (v.nd > 0)
        hipLaunchKernelGGL((k_spmv_plain<M>), dim3((v.nd + NT - 1) / NT), dim3(NT), 0, v.stream, v.nd, v.fstart,
                           v.fcol, static_cast<const M*>(v.vals), v.nnzf, x, y);
}

template void launch_spmv<double>(const DeviceView&, int, int);
template void launch_spmv<float>(const DeviceView&, int, int);
template void launch_update<double>(const DeviceView&, int);
template void launch_update<float>(const DeviceView&, int);
template void launch_spmv_plain<double>(const DeviceView&, const double*, double*);
template void launch_spmv_plain<float>(const DeviceView&, const double*, double*);

}  // namespace pk
