// Block-Jacobi PCG kernels for gfx950 (ConjugateGradientSolver.cpp:114-241).
//
// Iteration k is three launches:
//   k_dir(k)     vector tiles. Applies x += alpha_{k-1} d_{k-1} for its rows;
//                every workgroup sums the (r.z, r.r) partials the previous
//                launch left for its body, takes the reference's stop
//                decisions for the end of iteration k-1 (k = 0: the CG
//                start) and forms beta = delta_k / delta_{k-1}; it then writes
//                d_k = z_k + beta d_{k-1} for its rows.
//   k_spmv(k)    BSR row tiles, no reduction prologue: q = A d_k on the full
//                BSR (both triangles) and a d.q partial per tile.
//   k_update(k)  vector tiles. Every workgroup sums the body's d.q partials,
//                alpha = delta_k / d.q (or stops on d.q <= 0), updates
//                r -= alpha q, z = M^-1 r and leaves (r.z, r.r) partials; the
//                matching x += alpha d_k waits for k_dir(k + 1), which reads
//                d_k anyway.
// Totals are recomputed by every workgroup of the two vector kernels from the
// same partials in the same order, so no workgroup waits on another and no
// device-scope fence is needed inside a launch; the vector kernels have few
// workgroups, so the redundant sums are cheap, and the SpMV (thousands of
// workgroups) carries none. One designated workgroup per body records the
// decisions (delta, stop, iterations, residual) for later launches and the
// host; the host polls the count of bodies still iterating.
//
// SpMV layout: block values are stored as 9 planes (A_ij of every block in
// one contiguous plane), so LPR lanes of a row read consecutive blocks with
// fully coalesced 8-byte (or 4-byte) loads; U rows per lane group are loaded
// before any is reduced, to keep more bytes in flight per wave.
#include "pk_device.h"

#include <climits>
#include <type_traits>

namespace pk
{

// Records the end of a body's CG loop (designated workgroup only).
__device__ inline void record_stop(GroupState& s, int* active, int k, int inUpdate, int iters, bool converged,
                                   double residual, bool failed)
{
    s.stopAt = k;
    s.stopInUpdate = inUpdate;
    if (!inUpdate)
        s.dirStop = k;
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

// Wall-clock ticks at the start and end of each workgroup, one slot pair per
// workgroup (plain stores, no contention), when profiling.
__device__ inline void stamp_begin(unsigned long long* stamp)
{
    if (stamp && threadIdx.x == 0)
        stamp[2 * blockIdx.x] = static_cast<unsigned long long>(wall_clock64());
}
__device__ inline void stamp_end(unsigned long long* stamp)
{
    if (stamp && threadIdx.x == 0)
        stamp[2 * blockIdx.x + 1] = static_cast<unsigned long long>(wall_clock64());
}

// Iteration index and stamp slots of a CG launch. Issued on its own
// (kb == nullptr): k0 itself, and the host has already pointed `stamps` at
// this launch's slots. Replayed from a captured chunk: *kb + k0 (the chunk's
// base lives on the device and advances after every replay), and the slots
// of launch k are found from the buffer base. The first form adds no load to
// the start of a kernel, which matters for the short vector kernels.
struct LaunchIndex
{
    int k;
    unsigned long long* stamp;
};

__device__ inline LaunchIndex launch_index(const int* kb, int k0, unsigned long long* stamps)
{
    if (kb == nullptr)
        return LaunchIndex{k0, stamps};
    const int k = __builtin_amdgcn_readfirstlane(*kb + k0);
    return LaunchIndex{k, stamps ? stamps + 2ull * gridDim.x * static_cast<unsigned long long>(k) : nullptr};
}

template <class M>
__device__ inline M direction(M z, M dprev, double beta)
{
    return static_cast<M>(__builtin_fma(beta, static_cast<double>(dprev), static_cast<double>(z)));
}

// The reference's stop rules for the end of iteration k-1 (k = 0: the CG
// start, ConjugateGradientSolver.cpp:151-176; k > 0: :206-231 and the loop
// test at :180). Identical in every workgroup of the body. Returns true when
// the loop is over; otherwise beta is set for k > 0.
__device__ inline bool stop_rules(int k, GroupState& s, double rz, double rr, int flags, bool designated, int* active,
                                  double& beta)
{
    bool stop = false, conv = false, failed = false;
    double resid = 0.0;
    int iters = k;
    beta = 0.0;
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
    if (stop && designated)
        record_stop(s, active, k, 0, iters, conv, resid, failed);
    return stop;
}

// The end of iteration k-1 and the start of iteration k for the tile's rows.
// The solution's updates are applied two iterations at a time: at even k,
// x += alpha_{k-2} d_{k-2}, then x += alpha_{k-1} d_{k-1} (two fused
// multiply-adds in iteration order, so x is bit for bit what one update per
// iteration gives, ConjugateGradientSolver.cpp:204); at odd k the update of
// iteration k-1 waits, unless the loop ends here. Then the stop decisions for
// the end of iteration k-1 and, when the loop goes on, d_k = z_k +
// beta d_{k-1}. At k = 0 the right-hand side already wrote x = 0 and d_0. A
// body whose loop an earlier launch ended is skipped (a stop that this
// launch's designated workgroup records, stopAt == k, is reached by every
// workgroup, after its share of x).
template <class M, bool EVEN>
__device__ inline void dir_body(int k, const TileView& vt, const M* __restrict__ z, const M* __restrict__ dprev,
                                const M* __restrict__ dprev2, M* __restrict__ dcur, double* __restrict__ xs,
                                const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(vt, t, tl))
        return;
    // the thread's first kPre values of z, d_{k-1} (and at even k d_{k-2}
    // and x; a 512-row tile whole) are in flight while the stop flag and the
    // partials are read. x (like r in k_update) is stored non-temporally: it
    // is read again two iterations later, after the SpMV has streamed the
    // matrix through the L2s, so it need not stay dirty in an L2 that the
    // end of the launch has to write back (q, z and d are read by the very
    // next launch on the same rows and keep the L2)
    constexpr int kPre = 6;
    const int i0 = 3 * tl.rowBegin + static_cast<int>(threadIdx.x), iEnd = 3 * tl.rowEnd;
    const bool two = EVEN && k > 0;
    M zp[kPre], dp[kPre], ep[kPre];
    double xp[kPre];
    if (k > 0 && tl.rowEnd > tl.rowBegin)
    {
#pragma unroll
        for (int j = 0; j < kPre; ++j)
        {
            const int i = i0 + j * NT < iEnd ? i0 + j * NT : i0 - static_cast<int>(threadIdx.x);
            zp[j] = z[i];
            dp[j] = dprev[i];
            if (EVEN)
            {
                ep[j] = dprev2[i];
                xp[j] = xs[i];
            }
        }
    }
    const int g = tl.group;
    GroupState& s = gs[g];
    const Partial p0 = first_partial(pbPrev, tl.vBegin, tl.vCount);
    const int stopAt = read_uniform(&s.stopAt);
    if (stopAt >= 0 && stopAt < k)
        return;
    const double a1 = k > 0 ? s.alpha[(k - 1) & 1] : 0.0;
    const double a2 = two ? s.alpha[k & 1] : 0.0;  // alpha_{k-2}
    double rz, rr;
    int flags;
    sum_partials_from(p0, pbPrev, tl.vBegin, tl.vCount, rz, rr, flags);
    double beta;
    const bool stop = stop_rules(k, s, rz, rr, flags, t == tl.vBegin && threadIdx.x == 0, active, beta);
    if (k == 0)
        return;
#pragma unroll
    for (int j = 0; j < kPre; ++j)
    {
        const int i = i0 + j * NT;
        if (i < iEnd)
        {
            if (EVEN)
                __builtin_nontemporal_store(
                    __builtin_fma(static_cast<double>(dp[j]), a1, __builtin_fma(static_cast<double>(ep[j]), a2, xp[j])),
                    &xs[i]);
            else if (stop)
                __builtin_nontemporal_store(__builtin_fma(static_cast<double>(dp[j]), a1, xs[i]), &xs[i]);
            if (!stop)
                dcur[i] = direction(zp[j], dp[j], beta);
        }
    }
    for (int i = i0 + kPre * NT; i < iEnd; i += NT)
    {
        const M dv = dprev[i];
        if (EVEN)
            __builtin_nontemporal_store(
                __builtin_fma(static_cast<double>(dv), a1, __builtin_fma(static_cast<double>(dprev2[i]), a2, xs[i])),
                &xs[i]);
        else if (stop)
            __builtin_nontemporal_store(__builtin_fma(static_cast<double>(dv), a1, xs[i]), &xs[i]);
        if (!stop)
            dcur[i] = direction(z[i], dv, beta);
    }
}

template <class M, bool EVEN>
__global__ __launch_bounds__(NT) void k_dir(const int* kb, int k0, TileView vt, const M* __restrict__ z,
                                            const M* __restrict__ dprev, const M* __restrict__ dprev2,
                                            M* __restrict__ dcur, double* __restrict__ xs,
                                            const Partial* __restrict__ pbPrev, GroupState* gs, int* active,
                                            unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    dir_body<M, EVEN>(k, vt, z, dprev, dprev2, dcur, xs, pbPrev, gs, active);
    stamp_end(stamp);
}

// Row i of a 3x3 block times (d0, d1, d2) with the rounding pinned by explicit
// fused multiply-adds: a_i0 d0, then + a_i1 d1, then + a_i2 d2. Every SpMV
// form (full or upper-triangle storage, any launch shape) forms its products
// through this, so they agree bit for bit whatever the compiler would
// otherwise contract.
template <class M>
__device__ inline double brow(const M* a, int i, double d0, double d1, double d2)
{
    return __builtin_fma(static_cast<double>(a[3 * i + 2]), d2,
                         __builtin_fma(static_cast<double>(a[3 * i + 1]), d1, static_cast<double>(a[3 * i]) * d0));
}

// Where an SpMV reads the direction vector d: global memory, or the
// segments of it that the workgroup staged in LDS.
template <class M>
struct GlobalDir
{
    const M* __restrict__ d;
    __device__ double at(int c, int j) const { return static_cast<double>(d[3 * c + j]); }
};

// LDS row of column c: c + sh0, plus (SEG: segments) the deltas of every
// later segment whose start c has reached. (Written as masked deltas, not a
// chain of selects, which the compiler turns into a lookup in scratch
// memory.) Without SEG the tile's staging is one contiguous window.
template <class M, bool SEG>
struct StagedDir
{
    const M* lds;
    int lo1, lo2, lo3;  // starts of segments 1..3 (unused: INT_MAX)
    int sh0;            // -lo0: LDS row of column c in segment 0 is c + sh0
    int dl1, dl2, dl3;  // shift of segment s minus that of segment s-1 (unused: 0)
    __device__ int at_row(int c) const
    {
        if (!SEG)
            return c + sh0;
        return c + sh0 + (c >= lo1 ? dl1 : 0) + (c >= lo2 ? dl2 : 0) + (c >= lo3 ? dl3 : 0);
    }
    __device__ double at(int c, int j) const { return static_cast<double>(lds[3 * at_row(c) + j]); }
};

// d_row . q_row for a tile's partial, pinned the same way.
template <class M, class Dir>
__device__ inline double dot_dq(const Dir& d, int row, M q0, M q1, M q2)
{
    return __builtin_fma(d.at(row, 2), static_cast<double>(q2),
                         __builtin_fma(d.at(row, 1), static_cast<double>(q1), d.at(row, 0) * static_cast<double>(q0)));
}

// Dynamic LDS of the SpMV launches: the staged segments of d.
extern __shared__ __align__(16) unsigned char pk_stage_lds[];

// Copies the tile's segments of d into LDS with coalesced loads and waits
// for the workgroup; returns the staged view. LDS element i belongs to
// segment s with 3 end[s-1] <= i < 3 end[s] and holds d element i - 3 shift[s].
template <class M, bool SEG>
__device__ inline StagedDir<M, SEG> stage_segments(const M* __restrict__ d, const StageSegs& seg)
{
    M* lds = reinterpret_cast<M*>(pk_stage_lds);
    const int4 lo = seg.lo, end = seg.end;
    // shifts: segment s's column lo_s sits at LDS row end_{s-1}
    const int sh0 = -lo.x;
    const int sh1 = lo.y == INT_MAX ? sh0 : end.x - lo.y;
    const int sh2 = lo.z == INT_MAX ? sh1 : end.y - lo.z;
    const int sh3 = lo.w == INT_MAX ? sh2 : end.z - lo.w;
    const StagedDir<M, SEG> v{lds, lo.y, lo.z, lo.w, sh0, sh1 - sh0, sh2 - sh1, sh3 - sh2};
    const int n = 3 * end.w, e0 = 3 * end.x, e1 = 3 * end.y, e2 = 3 * end.z;
    for (int i0 = threadIdx.x; i0 < n; i0 += 4 * NT)
    {
        M x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int i = i0 + j * NT;
            const int sh = SEG ? v.sh0 + (i >= e0 ? v.dl1 : 0) + (i >= e1 ? v.dl2 : 0) + (i >= e2 ? v.dl3 : 0)
                               : v.sh0;
            x[j] = i < n ? d[i - 3 * sh] : M(0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i0 + j * NT < n)
                lds[i0 + j * NT] = x[j];
    }
    __syncthreads();
    return v;
}

// Block k of the full BSR: its 9 values (row-major) and its column.
__device__ inline void load_block(const double* __restrict__ vals, size_t F, const int* __restrict__ fcol, int k,
                                  double* a, int& col)
{
    col = fcol[k];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] = vals[e * F + k];
}

__device__ inline void load_block(const float* __restrict__ vals, size_t, const int* __restrict__, int k, float* a,
                                  int& col)
{
    const float2* p = reinterpret_cast<const float2*>(vals) + 5 * static_cast<size_t>(k);
    const float2 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
    a[0] = w0.x; a[1] = w0.y; a[2] = w1.x; a[3] = w1.y; a[4] = w2.x;
    a[5] = w2.y; a[6] = w3.x; a[7] = w3.y; a[8] = w4.x;
    col = __float_as_int(w4.y);
}

// The start of an SpMV workgroup: k_dir(k) has decided, and the body is
// skipped when its loop ended at or before k.
__device__ inline bool spmv_live(int k, const RowTile& tl, GroupState* gs)
{
    const int stopAt = read_uniform(&gs[tl.group].stopAt);
    return !(stopAt >= 0 && stopAt <= k);
}

// The rows of one tile (full storage) with one lane per row: a row's sum is
// (its blocks below the diagonal, columns descending) + (the diagonal and
// the blocks above it, columns ascending), the order of sliced upper-only
// storage; fdiag holds each row's full index of its diagonal block.
template <class M, int U, class Dir>
__device__ __attribute__((always_inline)) inline double full_rows1(const RowTile& tl, const int* __restrict__ fstart,
                                                                   const int* __restrict__ fcol,
                                                                   const int* __restrict__ fdiag,
                                                                   const M* __restrict__ vals, size_t F, Dir d,
                                                                   M* __restrict__ q)
{
    double pd = 0.0;
    const int nrows = tl.rowEnd - tl.rowBegin;
    for (int i0 = static_cast<int>(threadIdx.x); i0 < nrows; i0 += NT * U)
    {
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int i = i0 + u * NT;
            if (i >= nrows)
                break;
            const int row = tl.rowBegin + i;
            const int f0 = fstart[row], f1 = fstart[row + 1], dp = fdiag[row];
            double l[3] = {0.0, 0.0, 0.0}, h[3] = {0.0, 0.0, 0.0};
            for (int p = dp - 1; p >= f0; --p)
            {
                M a[9];
                int c;
                load_block(vals, F, fcol, p, a, c);
                const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    l[r] += brow(a, r, d0, d1, d2);
            }
            for (int p = dp; p < f1; ++p)
            {
                M a[9];
                int c;
                load_block(vals, F, fcol, p, a, c);
                const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    h[r] += brow(a, r, d0, d1, d2);
            }
            const M q0 = static_cast<M>(l[0] + h[0]), q1 = static_cast<M>(l[1] + h[1]),
                    q2 = static_cast<M>(l[2] + h[2]);
            q[3 * row] = q0;
            q[3 * row + 1] = q1;
            q[3 * row + 2] = q2;
            pd += dot_dq(d, row, q0, q1, q2);
        }
    }
    return pd;
}

// The rows of one tile (full storage); returns the lane's d.q partial.
// Lane l of a row's LPR lanes sums the row's blocks l, l + LPR, ... in that
// order, then the lanes' sums are combined by a butterfly (LPR = 1:
// full_rows1).
template <class M, int LPR, int U, class Dir>
__device__ __attribute__((always_inline)) inline double full_rows(const RowTile& tl, const int* __restrict__ fstart,
                                                                  const int* __restrict__ fcol,
                                                                  const int* __restrict__ fdiag,
                                                                  const M* __restrict__ vals, long long nnzf, Dir d,
                                                                  M* __restrict__ q)
{
    if constexpr (LPR == 1)
        return full_rows1<M, U>(tl, fstart, fcol, fdiag, vals, static_cast<size_t>(nnzf), d, q);
    const int lane = threadIdx.x % LPR;
    const int sub = threadIdx.x / LPR;
    constexpr int RPP = NT / LPR;
    const size_t F = static_cast<size_t>(nnzf);
    double pd = 0.0;
    const int nrows = tl.rowEnd - tl.rowBegin;
    for (int i0 = sub; i0 < nrows; i0 += RPP * U)
    {
        int kb[U], ke[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int i = i0 + u * RPP;
            const int row = tl.rowBegin + i;
            const bool live = i < nrows;
            kb[u] = live ? fstart[row] + lane : 0;
            ke[u] = live ? fstart[row + 1] : 0;
        }
        // first block of every row: all loads issued before any use
        int col[U];
        M a[U][9];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int idx = kb[u] < ke[u] ? kb[u] : 0;
            load_block(vals, F, fcol, idx, a[u], col[u]);
        }
        double y[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const bool has = kb[u] < ke[u];
            // a lane with no block reads row tl.rowBegin's column, which every staging holds
            const int c = has ? col[u] : tl.rowBegin;
            const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
#pragma unroll
            for (int i = 0; i < 3; ++i)
                y[u][i] = has ? brow(a[u], i, d0, d1, d2) : 0.0;
        }
        // rows longer than LPR blocks
#pragma unroll
        for (int u = 0; u < U; ++u)
            for (int kk = kb[u] + LPR; kk < ke[u]; kk += LPR)
            {
                M b[9];
                int c;
                load_block(vals, F, fcol, kk, b, c);
                const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    y[u][i] += brow(b, i, d0, d1, d2);
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1)
            {
                y[u][0] += __shfl_xor(y[u][0], o, LPR);
                y[u][1] += __shfl_xor(y[u][1], o, LPR);
                y[u][2] += __shfl_xor(y[u][2], o, LPR);
            }
            const int i = i0 + u * RPP;
            const int row = tl.rowBegin + i;
            if (lane == 0 && i < nrows)
            {
                const M q0 = static_cast<M>(y[u][0]), q1 = static_cast<M>(y[u][1]), q2 = static_cast<M>(y[u][2]);
                q[3 * row] = q0;
                q[3 * row + 1] = q1;
                q[3 * row + 2] = q2;
                pd += dot_dq(d, row, q0, q1, q2);
            }
        }
    }
    return pd;
}

// q = A d_k for the tile's rows and the tile's d.q partial. A body whose loop
// ended at or before k (decided by k_dir(k) or earlier) is skipped. When the
// tile's segments win[t] (up to four runs of rows covering every column its
// rows reference, their own rows included: on a grid, the layers below, at
// and above the tile) fit the launch's staging capacity, the workgroup first
// copies them from d into LDS with coalesced loads and reads every direction
// block from there; otherwise it gathers from global memory. Products and
// sums are the same either way.
template <class M, int LPR, int U, bool SEG>
__device__ inline void spmv_body(int k, const TileView& st, const int* __restrict__ fstart,
                                 const int* __restrict__ fcol, const int* __restrict__ fdiag,
                                 const M* __restrict__ vals, long long nnzf,
                                 const M* __restrict__ d, M* __restrict__ q, double* __restrict__ pa,
                                 GroupState* gs, const StageSegs* __restrict__ win, int winCap)
{
    int t;
    RowTile tl;
    if (!tile_for(st, t, tl))
        return;
    StageSegs sg{make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
    if (winCap > 0)
        sg = win[t];  // in flight with the stop flag
    if (!spmv_live(k, tl, gs))
        return;
    double pd;
    if (winCap > 0 && sg.end.w <= winCap)  // uniform: one tile per workgroup
        pd = full_rows<M, LPR, U>(tl, fstart, fcol, fdiag, vals, nnzf, stage_segments<M, SEG>(d, sg), q);
    else
        pd = full_rows<M, LPR, U>(tl, fstart, fcol, fdiag, vals, nnzf, GlobalDir<M>{d}, q);
    double u1 = 0.0, u2 = 0.0;
    block_sum3(pd, u1, u2);
    if (threadIdx.x == 0)
        pa[t] = pd;
}

// Element e of sliced upper-only storage: its 9 values (row-major), as two
// 16-byte loads and one 4-byte load (FP32) or four 16-byte loads and one
// 8-byte load (FP64), each from its own plane.
__device__ inline void load_element(const float* __restrict__ rec, size_t nE, int e, float* a)
{
    const float4 g0 = reinterpret_cast<const float4*>(rec)[e];
    const float4 g1 = reinterpret_cast<const float4*>(rec + 4 * nE)[e];
    a[0] = g0.x; a[1] = g0.y; a[2] = g0.z; a[3] = g0.w;
    a[4] = g1.x; a[5] = g1.y; a[6] = g1.z; a[7] = g1.w;
    a[8] = rec[8 * nE + e];
}

__device__ inline void load_element(const double* __restrict__ rec, size_t nE, int e, double* a)
{
#pragma unroll
    for (int g = 0; g < 4; ++g)
    {
        const double2 w = reinterpret_cast<const double2*>(rec + 2 * g * nE)[e];
        a[2 * g] = w.x;
        a[2 * g + 1] = w.y;
    }
    a[8] = rec[8 * nE + e];
}

__device__ inline bool lane_bit(unsigned long long m, int lane) { return ((m >> lane) & 1ull) != 0; }

// Number of lanes of the wave below this one whose bit is set in m.
__device__ inline int lanes_below(unsigned long long m)
{
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0u)));
}

// One slice (64 rows, one lane per row) of sliced upper-only storage. Round
// i multiplies the row's upper block at position i (i = 0: the diagonal)
// and its lower block at position i - 1. The lower ELL runs down from the
// diagonal, so the transpose of the mirror at offset -o sits at the position
// the upper block at offset +o takes; on a regular grid (aligned slices) the
// two reads of one record -- by its own row and, transposed, by its mirror's
// row -- then fall in the same round, in this wave or in the workgroup a
// layer of rows earlier, and the second finds the record in L2. In an
// aligned slice a block's column is the row plus the position's offset;
// otherwise it comes from the column planes. The loop is software-pipelined:
// round i's masks, mirror elements (and columns) are loaded during round
// i - 1, so each round waits on one load of values (and of d) only. The
// row's sum is (lower blocks, columns descending) + (upper blocks, columns
// ascending), each through brow: full storage at one lane per row sums the
// same way, so the two storages agree bit for bit. Every load of a wave
// reads consecutive entries of one plane, or in aligned slices the mirror
// records of consecutive rows. Returns the lane's d.q contribution.
template <class M, bool ALIGNED, bool COMPACT, int U, class Dir>
__device__ __attribute__((always_inline)) inline double slice_body(const SliceView& sv, const int4& ds, int row,
                                                                   int lane, bool live, const M* __restrict__ rec,
                                                                   size_t nE, const Dir& d, M* __restrict__ q)
{
    const int wu = ds.z & 0xffff, wl = (ds.z >> 16) & 0x3fff;
    const int rounds = wu > wl + 1 ? wu : wl + 1;
    const int* __restrict__ off = sv.off + ds.w;
    const unsigned long long* __restrict__ mask = sv.mask + ds.w;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
    // compact slices: first entries of the next rounds' upper and lower
    // positions (a position's entries follow those of the positions before
    // it); otherwise position k of lane l is entry base + 64 k + l
    int ub = ds.x, lb = ds.y;
    // what round i needs before its values: whether the lane has an upper
    // and a lower block, its upper element, the mirror element of its lower
    // block and (packed slices) their columns
    auto fetch = [&](int i, bool& h_u, bool& h_l, int& e_u, int& e_m, int& c_u, int& c_l) {
        int li;
        if constexpr (COMPACT)
        {
            const unsigned long long mu = i < wu ? mask[i] : 0ull;
            const unsigned long long ml = (i >= 1 && i <= wl) ? mask[wu + i - 1] : 0ull;
            h_u = lane_bit(mu, lane);
            h_l = lane_bit(ml, lane);
            e_u = ub + lanes_below(mu);
            li = lb + lanes_below(ml);
            ub += __popcll(mu);
            lb += __popcll(ml);
        }
        else
        {
            // the two ways of testing a lane's bit give the same flags but
            // schedule differently: measured, the short-circuit form is
            // faster with FP64 values and the select form with FP32
            if constexpr (sizeof(M) == 8)
            {
                h_u = i < wu && lane_bit(mask[i], lane);
                h_l = i >= 1 && i <= wl && lane_bit(mask[wu + i - 1], lane);
            }
            else
            {
                h_u = lane_bit(i < wu ? mask[i] : 0ull, lane);
                h_l = lane_bit((i >= 1 && i <= wl) ? mask[wu + i - 1] : 0ull, lane);
            }
            e_u = ds.x + kSliceRows * i + lane;
            li = ds.y + kSliceRows * (i - 1) + lane;
        }
        e_m = 0;
        c_u = 0;
        c_l = 0;
        if (h_l)
            e_m = sv.mirE[li];
        if (!ALIGNED && h_u)
            c_u = sv.ecol[e_u];
        if (!ALIGNED && h_l)
            c_l = sv.mirC[li];
    };
    bool hu[U], hl[U];
    int eu[U], em[U], cu[U], cl[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        fetch(u, hu[u], hl[u], eu[u], em[u], cu[u], cl[u]);
    for (int i0 = 0; i0 < rounds; i0 += U)
    {
        bool hu1[U], hl1[U];
        int eu1[U], em1[U], cu1[U], cl1[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            fetch(i0 + U + u, hu1[u], hl1[u], eu1[u], em1[u], cu1[u], cl1[u]);
        M a[U][9], b[U][9];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            if (hu[u])
                load_element(rec, nE, COMPACT ? eu[u] : ds.x + kSliceRows * (i0 + u) + lane, a[u]);
            if (hl[u])
                load_element(rec, nE, em[u], b[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int i = i0 + u;
            if (hu[u])
            {
                const int c = ALIGNED ? row + off[i] : cu[u];
                const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
                u0 += brow(a[u], 0, d0, d1, d2);
                u1 += brow(a[u], 1, d0, d1, d2);
                u2 += brow(a[u], 2, d0, d1, d2);
            }
            if (hl[u])
            {
                const int c = ALIGNED ? row - off[wu + i - 1] : cl[u];
                const double d0 = d.at(c, 0), d1 = d.at(c, 1), d2 = d.at(c, 2);
                M t[9];
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
                        t[3 * ii + jj] = b[u][3 * jj + ii];
                l0 += brow(t, 0, d0, d1, d2);
                l1 += brow(t, 1, d0, d1, d2);
                l2 += brow(t, 2, d0, d1, d2);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            hu[u] = hu1[u];
            hl[u] = hl1[u];
            eu[u] = eu1[u];
            em[u] = em1[u];
            cu[u] = cu1[u];
            cl[u] = cl1[u];
        }
    }
    if (!live)
        return 0.0;
    const M q0 = static_cast<M>(l0 + u0), q1 = static_cast<M>(l1 + u1), q2 = static_cast<M>(l2 + u2);
    q[3 * row] = q0;
    q[3 * row + 1] = q1;
    q[3 * row + 2] = q2;
    return dot_dq(d, row, q0, q1, q2);
}

// The rows of one tile in sliced upper-only storage: wave w of the
// workgroup takes the tile's slices w, w + NT / 64, ... (slice_body, U
// rounds loaded at a time).
template <class M, int U, class Dir>
__device__ __attribute__((always_inline)) inline double slice_rows(const RowTile& tl, int slice0,
                                                                   const SliceView& sv, const M* __restrict__ rec,
                                                                   Dir d, M* __restrict__ q)
{
    constexpr int W = kSliceRows, WAVES = NT / kSliceRows;
    const int wave = static_cast<int>(threadIdx.x) / W;
    const int lane = static_cast<int>(threadIdx.x) % W;
    const int nsl = (tl.rowEnd - tl.rowBegin + W - 1) / W;
    const size_t nE = static_cast<size_t>(sv.nE);
    double pd = 0.0;
    for (int j = wave; j < nsl; j += WAVES)
    {
        const int sl = __builtin_amdgcn_readfirstlane(slice0 + j);
        const int4 ds = sv.slices[sl];
        const int row = tl.rowBegin + W * j + lane;
        const bool live = row < tl.rowEnd;
        const bool compact = ((ds.z >> 30) & 1) != 0;  // uniform, like aligned
        if (ds.z < 0)
        {
            if (compact)
                pd += slice_body<M, true, true, U>(sv, ds, row, lane, live, rec, nE, d, q);
            else
                pd += slice_body<M, true, false, U>(sv, ds, row, lane, live, rec, nE, d, q);
        }
        else if (compact)
            pd += slice_body<M, false, true, U>(sv, ds, row, lane, live, rec, nE, d, q);
        else
            pd += slice_body<M, false, false, U>(sv, ds, row, lane, live, rec, nE, d, q);
    }
    return pd;
}

// spmv_body over sliced upper-only storage.
template <class M, int U, bool SEG>
__device__ inline void spmv_body_upper(int k, const TileView& st, const SliceView& sv, const M* __restrict__ rec,
                                       const M* __restrict__ d, M* __restrict__ q, double* __restrict__ pa,
                                       GroupState* gs, const StageSegs* __restrict__ win, int winCap)
{
    int t;
    RowTile tl;
    if (!tile_for(st, t, tl))
        return;
    StageSegs sg{make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
    if (winCap > 0)
        sg = win[t];  // in flight with the stop flag
    const int s0 = sv.tileSlice[t];
    if (!spmv_live(k, tl, gs))
        return;
    double pd;
    if (winCap > 0 && sg.end.w <= winCap)  // uniform: one tile per workgroup
        pd = slice_rows<M, U>(tl, s0, sv, rec, stage_segments<M, SEG>(d, sg), q);
    else
        pd = slice_rows<M, U>(tl, s0, sv, rec, GlobalDir<M>{d}, q);
    double u1 = 0.0, u2 = 0.0;
    block_sum3(pd, u1, u2);
    if (threadIdx.x == 0)
        pa[t] = pd;
}

template <class M, int U, bool SEG>
__global__ __launch_bounds__(NT) void k_spmv_upper(const int* kb, int k0, TileView st, SliceView sv,
                                                   const M* __restrict__ rec, const M* __restrict__ d,
                                                   M* __restrict__ q, double* __restrict__ pa, GroupState* gs,
                                                   const StageSegs* __restrict__ win, int winCap,
                                                   unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    spmv_body_upper<M, U, SEG>(k, st, sv, rec, d, q, pa, gs, win, winCap);
    stamp_end(stamp);
}

template <class M, int LPR, int U, bool SEG>
__global__ __launch_bounds__(NT) void k_spmv(const int* kb, int k0, TileView st, const int* __restrict__ fstart,
                                             const int* __restrict__ fcol, const int* __restrict__ fdiag,
                                             const M* __restrict__ vals,
                                             long long nnzf, const M* __restrict__ d, M* __restrict__ q,
                                             double* __restrict__ pa, GroupState* gs,
                                             const StageSegs* __restrict__ win, int winCap,
                                             unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    spmv_body<M, LPR, U, SEG>(k, st, fstart, fcol, fdiag, vals, nnzf, d, q, pa, gs, win, winCap);
    stamp_end(stamp);
}

template <class M, bool LONG>
__device__ inline void update_body(int k, const TileView& vt, const TileView& st, const M* __restrict__ minv,
                                   int nd, const M* __restrict__ q, double* __restrict__ rv, M* __restrict__ zv,
                                   const M* __restrict__ dprev, double* __restrict__ xs,
                                   const double* __restrict__ pa, Partial* __restrict__ pbCur,
                                   const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(vt, t, tl))
        return;
    // the thread's first kPre rows (q, r, M^-1) are in flight while the stop
    // flag and the d.q partials are read
    constexpr int kPre = 2;
    const size_t N = static_cast<size_t>(nd);
    const int row0 = tl.rowBegin + static_cast<int>(threadIdx.x);
    M qpre[kPre][3], mpre[kPre][kMinvPlanes];
    double rpre[kPre][3];
#pragma unroll
    for (int j = 0; j < kPre; ++j)
    {
        const int row = row0 + j * NT;
        const int rr = row < tl.rowEnd ? row : tl.rowBegin;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            qpre[j][i] = q[3 * rr + i];
            rpre[j][i] = rv[3 * rr + i];
        }
#pragma unroll
        for (int e = 0; e < kMinvPlanes; ++e)
            mpre[j][e] = minv[e * N + rr];
    }
    const int g = tl.group;
    GroupState& s = gs[g];
    const double pa0 = LONG ? 0.0 : first_double(pa, tl.sBegin, tl.sCount);
    // ended by an earlier launch (k_dir(k) included: dirStop is from a
    // previous launch); a stop that this launch's designated workgroup
    // records (stopAt == k) is reached by every workgroup below
    const int stopAt = read_uniform(&s.stopAt);
    if (stopAt >= 0 && (stopAt < k || s.dirStop == k))
        return;
    const double dq = LONG ? sum_doubles_long(pa, tl.sBegin, tl.sCount) : sum_doubles_from(pa0, pa, tl.sBegin, tl.sCount);
    const double delta = s.delta[k & 1];
    const double alpha = delta / dq;
    const bool designated = t == tl.vBegin && threadIdx.x == 0;
    if (!fin(dq) || dq <= 0.0 || !fin(alpha))
    {
        // break before any update (ConjugateGradientSolver.cpp:192-202); at
        // odd k the solution's update of iteration k-1 is still pending
        if (k & 1)
        {
            const double a1 = s.alpha[(k - 1) & 1];
            for (int i = 3 * tl.rowBegin + static_cast<int>(threadIdx.x); i < 3 * tl.rowEnd; i += NT)
                __builtin_nontemporal_store(__builtin_fma(static_cast<double>(dprev[i]), a1, xs[i]), &xs[i]);
        }
        double rz, rr;
        int fl;
        sum_partials(pbPrev, tl.vBegin, tl.vCount, rz, rr, fl);
        if (designated)
            record_stop(s, active, k, 1, k, delta <= s.target, norm_of(rr), false);
        return;
    }
    // x += alpha d_k is left to k_dir, which reads d_k anyway
    if (designated)
        s.alpha[k & 1] = alpha;
    double pz = 0.0, pr = 0.0;
    int flag = 0;
    // one row: r -= alpha q, z = M^-1 r (upper triangle m of the inverse),
    // and the partials
    auto one = [&](int row, const M* qq, const double* rq, const M* m) {
        const M mm[9] = {m[0], m[1], m[2], m[1], m[3], m[4], m[2], m[4], m[5]};
        double r[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            r[i] = rq[i] + static_cast<double>(qq[i]) * (-alpha);
            __builtin_nontemporal_store(r[i], &rv[3 * row + i]);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            z[i] = static_cast<double>(mm[3 * i]) * r[0] + static_cast<double>(mm[3 * i + 1]) * r[1] +
                   static_cast<double>(mm[3 * i + 2]) * r[2];
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
    };
#pragma unroll
    for (int j = 0; j < kPre; ++j)
        if (row0 + j * NT < tl.rowEnd)
            one(row0 + j * NT, qpre[j], rpre[j], mpre[j]);
    for (int row = row0 + kPre * NT; row < tl.rowEnd; row += NT)
    {
        M qq[3], mm[kMinvPlanes];
        double rq[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            qq[i] = q[3 * row + i];
            rq[i] = rv[3 * row + i];
        }
#pragma unroll
        for (int e = 0; e < kMinvPlanes; ++e)
            mm[e] = minv[e * N + row];
        one(row, qq, rq, mm);
    }
    double u = 0.0;
    block_sum3(pz, pr, u);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
        pbCur[t] = Partial{pz, pr, flag, 0};
}

template <class M, bool LONG>
__global__ __launch_bounds__(NT) void k_update(const int* kb, int k0, TileView vt, TileView st,
                                               const M* __restrict__ minv, int nd, const M* __restrict__ q,
                                               double* __restrict__ rv, M* __restrict__ zv,
                                               const M* __restrict__ dprev, double* __restrict__ xs,
                                               const double* __restrict__ pa, Partial* __restrict__ pbCur,
                                               const Partial* __restrict__ pbPrev, GroupState* gs, int* active,
                                               unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    update_body<M, LONG>(k, vt, st, minv, nd, q, rv, zv, dprev, xs, pa, pbCur, pbPrev, gs, active);
    stamp_end(stamp);
}

// Span of each launch: earliest workgroup start to latest workgroup end.
// Workgroup l reduces launch l (grid slot pairs at raw + 2 * grid * l).
__global__ __launch_bounds__(NT) void k_stamp_reduce(const unsigned long long* __restrict__ raw, int grid,
                                                     unsigned long long* __restrict__ span)
{
    const unsigned long long* p = raw + 2 * static_cast<size_t>(grid) * blockIdx.x;
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int i = threadIdx.x; i < grid; i += NT)
    {
        // a zero slot belongs to a workgroup the launch did not have (its
        // grid was cut to the tiles still iterating)
        const unsigned long long b = p[2 * i];
        if (b != 0ull)
        {
            lo = min(lo, b);
            hi = max(hi, p[2 * i + 1]);
        }
    }
    __shared__ unsigned long long sl[NT], sh[NT];
    sl[threadIdx.x] = lo;
    sh[threadIdx.x] = hi;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1)
    {
        if (static_cast<int>(threadIdx.x) < o)
        {
            sl[threadIdx.x] = min(sl[threadIdx.x], sl[threadIdx.x + o]);
            sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        span[2 * blockIdx.x] = sl[0];
        span[2 * blockIdx.x + 1] = sh[0];
    }
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
        M a[9];
        int c;
        load_block(vals, F, fcol, k, a, c);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            acc[i] += brow(a, i, xin[3 * c], xin[3 * c + 1], xin[3 * c + 2]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        y[3 * row + i] = acc[i];
}

// y = A x over sliced upper-only storage, one thread per row, blocks in
// full-row order (the same sums as k_spmv_plain over full storage), columns
// from the column planes.
template <class M>
__global__ __launch_bounds__(NT) void k_spmv_plain_upper(int nd, SliceView sv, const M* __restrict__ rec,
                                                         const double* __restrict__ xin, double* __restrict__ y)
{
    const int row = blockIdx.x * NT + threadIdx.x;
    if (row >= nd)
        return;
    const int2 rs = sv.rowSlice[row];  // (slice, lane)
    const int4 ds = sv.slices[rs.x];
    const int wu = ds.z & 0xffff, wl = (ds.z >> 16) & 0x3fff, l = rs.y;
    const bool compact = ((ds.z >> 30) & 1) != 0;
    const unsigned long long below = (1ull << l) - 1ull;
    const size_t n = static_cast<size_t>(sv.nE);
    double acc[3] = {0, 0, 0};
    int lb = ds.y;  // first entry of lower position k
    for (int k = 0; k < wl; ++k)
        lb += compact ? __popcll(sv.mask[ds.w + wu + k]) : kSliceRows;
    for (int k = wl - 1; k >= 0; --k)  // the lower ELL runs from the diagonal down
    {
        const unsigned long long m = sv.mask[ds.w + wu + k];
        lb -= compact ? __popcll(m) : kSliceRows;
        if (lane_bit(m, l))
        {
            const int idx = lb + (compact ? __popcll(m & below) : l);
            M a[9], t[9];
            load_element(rec, n, sv.mirE[idx], a);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    t[3 * i + j] = a[3 * j + i];
            const int c = sv.mirC[idx];
            for (int i = 0; i < 3; ++i)
                acc[i] += brow(t, i, xin[3 * c], xin[3 * c + 1], xin[3 * c + 2]);
        }
    }
    int ub = ds.x;  // first entry of upper position k
    for (int k = 0; k < wu; ++k)
    {
        const unsigned long long m = sv.mask[ds.w + k];
        if (lane_bit(m, l))
        {
            const int e = ub + (compact ? __popcll(m & below) : l);
            M a[9];
            load_element(rec, n, e, a);
            const int c = sv.ecol[e];
            for (int i = 0; i < 3; ++i)
                acc[i] += brow(a, i, xin[3 * c], xin[3 * c + 1], xin[3 * c + 2]);
        }
        ub += compact ? __popcll(m) : kSliceRows;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        y[3 * row + i] = acc[i];
}

template <class M, int U>
static void launch_spmv_upper(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    const size_t lds = 3 * sizeof(M) * static_cast<size_t>(v.winCap);
    auto kernel = v.stageSegs ? k_spmv_upper<M, U, true> : k_spmv_upper<M, U, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.spmv.per)), dim3(NT), lds, v.stream, kb, k0,
                       v.spmv, v.slice, static_cast<const M*>(v.vals), static_cast<const M*>(v.d[k0 & 3]),
                       static_cast<M*>(v.q), v.pa, v.gs, v.spmvWin, v.winCap, stamp);
}

// The iteration of a launch is k0 (kb == nullptr) or *kb + k0 with *kb even,
// so k0 alone picks the ping-pong buffers.
template <class M, int LPR, int U>
static void launch_spmv_shape(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    const size_t lds = 3 * sizeof(M) * static_cast<size_t>(v.winCap);
    auto kernel = v.stageSegs ? k_spmv<M, LPR, U, true> : k_spmv<M, LPR, U, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.spmv.per)), dim3(NT), lds, v.stream, kb, k0,
                       v.spmv, v.fstart, v.fcol, v.fdiag, static_cast<const M*>(v.vals), v.nnzf,
                       static_cast<const M*>(v.d[k0 & 3]), static_cast<M*>(v.q), v.pa, v.gs, v.spmvWin, v.winCap,
                       stamp);
}

template <class M>
void launch_dir(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    if (v.vec.count == 0)
        return;
    // d_k lives in d[k & 3]: a captured chunk starts at a multiple of 4, so
    // the slot of its launch k0 is k0 & 3 as well
    auto kernel = (k0 & 1) ? k_dir<M, false> : k_dir<M, true>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, kb, k0, v.vec,
                       static_cast<const M*>(v.z), static_cast<const M*>(v.d[(k0 + 3) & 3]),
                       static_cast<const M*>(v.d[(k0 + 2) & 3]), static_cast<M*>(v.d[k0 & 3]), v.xs,
                       v.pb[(k0 + 1) & 1], v.gs, v.active, stamp);
}

// Upper-only storage runs one lane per row (lanesPerRow 1) and one round
// per pipeline step (two rounds per step ran slower on MI355X: the extra
// registers cost more waves than the deeper pipeline gained), whatever
// `unroll`; full storage any listed shape.
template <class M>
void launch_spmv(const DeviceView& v, const int* kb, int k0, int lanesPerRow, int unroll, unsigned long long* stamp)
{
    if (v.spmv.count == 0)
        return;
    if (v.upper)
    {
        launch_spmv_upper<M, 1>(v, kb, k0, stamp);
        return;
    }
#define PK_SHAPE(L, U)                        \
    if (lanesPerRow == L && unroll == U)      \
    {                                         \
        launch_spmv_shape<M, L, U>(v, kb, k0, stamp); \
        return;                               \
    }
    PK_SHAPE(1, 1) PK_SHAPE(1, 2) PK_SHAPE(1, 4)
    PK_SHAPE(4, 1) PK_SHAPE(4, 2) PK_SHAPE(4, 4)
    PK_SHAPE(8, 1) PK_SHAPE(8, 2) PK_SHAPE(8, 4)
    PK_SHAPE(16, 1) PK_SHAPE(16, 2) PK_SHAPE(16, 4)
    PK_SHAPE(32, 1) PK_SHAPE(32, 2)
#undef PK_SHAPE
    launch_spmv_shape<M, 16, 1>(v, kb, k0, stamp);
}

// LDS one SpMV workgroup can stage d in, for the kernel a launch shape
// selects (both staging forms), without fewer workgroups fitting on a CU
// than its registers allow: the CU's LDS shared among that many in 2 KB
// units, less the kernel's own LDS. 20 KB when the device cannot be asked.
template <class M, int LPR, int U>
static size_t stage_share(bool upper, int ldsPerCU)
{
    const void* kernels[2] = {upper ? reinterpret_cast<const void*>(k_spmv_upper<M, 1, true>)
                                    : reinterpret_cast<const void*>(k_spmv<M, LPR, U, true>),
                              upper ? reinterpret_cast<const void*>(k_spmv_upper<M, 1, false>)
                                    : reinterpret_cast<const void*>(k_spmv<M, LPR, U, false>)};
    size_t share = SIZE_MAX;
    for (const void* k : kernels)
    {
        int blocks = 0;
        hipFuncAttributes attr{};
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, NT, 0) != hipSuccess || blocks <= 0 ||
            hipFuncGetAttributes(&attr, k) != hipSuccess)
            return 20 * 1024;
        const size_t each = (static_cast<size_t>(ldsPerCU) / static_cast<size_t>(blocks)) / 2048 * 2048;
        const size_t own = attr.sharedSizeBytes;
        share = std::min(share, each > own ? each - own : 0);
    }
    return share;
}

template <class M>
size_t spmv_stage_bytes(bool upper, int lanesPerRow, int unroll)
{
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || lds <= 0)
        return 20 * 1024;
    if (upper)
        return stage_share<M, 1, 1>(true, lds);
#define PK_SHAPE(L, U)                   \
    if (lanesPerRow == L && unroll == U) \
        return stage_share<M, L, U>(false, lds);
    PK_SHAPE(1, 1) PK_SHAPE(1, 2) PK_SHAPE(1, 4)
    PK_SHAPE(4, 1) PK_SHAPE(4, 2) PK_SHAPE(4, 4)
    PK_SHAPE(8, 1) PK_SHAPE(8, 2) PK_SHAPE(8, 4)
    PK_SHAPE(16, 1) PK_SHAPE(16, 2) PK_SHAPE(16, 4)
    PK_SHAPE(32, 1) PK_SHAPE(32, 2)
#undef PK_SHAPE
    return stage_share<M, 16, 1>(false, lds);
}

template <class M>
void launch_update(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    if (v.vec.count == 0)
        return;
    auto kernel = v.longSums ? k_update<M, true> : k_update<M, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, kb, k0,
                       v.vec, v.spmv, static_cast<const M*>(v.minv), v.nd, static_cast<const M*>(v.q), v.r,
                       static_cast<M*>(v.z), static_cast<const M*>(v.d[(k0 + 3) & 3]), v.xs, v.pa, v.pb[k0 & 1],
                       v.pb[(k0 + 1) & 1], v.gs, v.active, stamp);
}

__global__ void k_advance(int* kb, int by)
{
    if (threadIdx.x == 0)
        *kb += by;
}

void launch_advance(const DeviceView& v, int* kb, int by)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, v.stream, kb, by);
}

// Builds the workgroup-to-tile maps of a compacted batched CG launch (see
// TileView) for the vector tiling and the SpMV tiling: the tiles of the
// bodies whose loop is still running form a list in increasing order, entry i
// of which goes to workgroup (i % per) * 8 + i / per with per = ceil(n / 8),
// the XCD-contiguous order of an uncompacted launch over n tiles; every other
// workgroup of the full grid gets group -1. Every workgroup forms the
// per-group list offsets in LDS (an exclusive scan over the groups, -1 for a
// stopped body); thread j of workgroups [0, vecBlocks) then places vector
// tile j and settles workgroup slot j of the vector map, the rest the same
// for the SpMV tiling.
__global__ __launch_bounds__(NT) void k_compact(int G, const GroupState* __restrict__ gs, TileView vt, TileView st,
                                                int vecBlocks, RowTile* __restrict__ vecPerm,
                                                RowTile* __restrict__ spmvPerm, int* __restrict__ counts)
{
    __shared__ int offV[kCompactMaxGroups], offS[kCompactMaxGroups];
    __shared__ int scanV[NT], scanS[NT];
    __shared__ int carryV, carryS;
    const int tid = static_cast<int>(threadIdx.x);
    if (tid == 0)
        carryV = carryS = 0;
    __syncthreads();
    for (int g0 = 0; g0 < G; g0 += NT)
    {
        const int g = g0 + tid;
        const bool run = g < G && gs[g].stopAt < 0;
        const int cv = run ? vt.groupCount[g] : 0;
        const int cs = run ? st.groupCount[g] : 0;
        scanV[tid] = cv;
        scanS[tid] = cs;
        __syncthreads();
        for (int o = 1; o < NT; o <<= 1)  // inclusive Hillis-Steele scan
        {
            const int av = tid >= o ? scanV[tid - o] : 0;
            const int as = tid >= o ? scanS[tid - o] : 0;
            __syncthreads();
            scanV[tid] += av;
            scanS[tid] += as;
            __syncthreads();
        }
        if (g < G)
        {
            offV[g] = run ? carryV + scanV[tid] - cv : -1;
            offS[g] = run ? carryS + scanS[tid] - cs : -1;
        }
        __syncthreads();
        if (tid == NT - 1)
        {
            carryV += scanV[NT - 1];
            carryS += scanS[NT - 1];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && tid == 0)
    {
        // tiles still iterating: later launches need no more than 8 *
        // ceil(count / 8) workgroups
        counts[0] = carryV;
        counts[1] = carryS;
    }
    const bool vec = static_cast<int>(blockIdx.x) < vecBlocks;
    const TileView& tv = vec ? vt : st;
    RowTile* perm = vec ? vecPerm : spmvPerm;
    const int* off = vec ? offV : offS;
    const int n = vec ? carryV : carryS;
    const int per = (n + 7) >> 3;
    const int j = (vec ? static_cast<int>(blockIdx.x) : static_cast<int>(blockIdx.x) - vecBlocks) * NT + tid;
    if (j < tv.count)
    {
        const int g = tv.tiles[j].group;
        const int o = off[g];
        if (o >= 0)
        {
            const int i = o + j - tv.groupBegin[g];
            RowTile tl = tv.tiles[j];
            tl.pad = j;
            perm[(i % per) * 8 + i / per] = tl;
        }
    }
    if (j < 8 * tv.per && ((j >> 3) >= per || (j & 7) * per + (j >> 3) >= n))
        perm[j] = RowTile{-1, 0, 0, -1};
}

void launch_compact(const DeviceView& v, RowTile* vecPerm, RowTile* spmvPerm)
{
    const int vb = (8 * v.vec.per + NT - 1) / NT;
    const int sb = (8 * v.spmv.per + NT - 1) / NT;
    if (v.G > 0 && v.G <= kCompactMaxGroups && vb + sb > 0)
        hipLaunchKernelGGL(k_compact, dim3(static_cast<unsigned>(vb + sb)), dim3(NT), 0, v.stream, v.G, v.gs, v.vec,
                           v.spmv, vb, vecPerm, spmvPerm, v.active + 1);
}

int spmv_grid(const DeviceView& v) { return 8 * v.spmv.per; }
int update_grid(const DeviceView& v) { return 8 * v.vec.per; }

void launch_stamp_reduce(hipStream_t stream, const unsigned long long* raw, int grid, int launches,
                         unsigned long long* span)
{
    if (launches > 0 && grid > 0)
        hipLaunchKernelGGL(k_stamp_reduce, dim3(static_cast<unsigned>(launches)), dim3(NT), 0, stream, raw, grid,
                           span);
}

__global__ __launch_bounds__(NT) void k_zero_u64(unsigned long long* __restrict__ p, long long n)
{
    for (long long i = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * NT)
        p[i] = 0ull;
}

void launch_zero_stamps(hipStream_t stream, unsigned long long* raw, long long n)
{
    if (raw && n > 0)
        hipLaunchKernelGGL(k_zero_u64, dim3(static_cast<unsigned>(std::min<long long>((n + NT - 1) / NT, 4096))),
                           dim3(NT), 0, stream, raw, n);
}

template <class M>
void launch_spmv_plain(const DeviceView& v, const double* x, double* y)
{
    if (v.upper)
    {
        if (v.nd > 0)
            hipLaunchKernelGGL(k_spmv_plain_upper<M>, dim3((v.nd + NT - 1) / NT), dim3(NT), 0, v.stream, v.nd,
                               v.slice, static_cast<const M*>(v.vals), x, y);
        return;
    }
    if (v.nd > 0)
        hipLaunchKernelGGL((k_spmv_plain<M>), dim3((v.nd + NT - 1) / NT), dim3(NT), 0, v.stream, v.nd, v.fstart,
                           v.fcol, static_cast<const M*>(v.vals), v.nnzf, x, y);
}

template void launch_dir<double>(const DeviceView&, const int*, int, unsigned long long*);
template void launch_dir<float>(const DeviceView&, const int*, int, unsigned long long*);
template void launch_spmv<double>(const DeviceView&, const int*, int, int, int, unsigned long long*);
template size_t spmv_stage_bytes<double>(bool, int, int);
template size_t spmv_stage_bytes<float>(bool, int, int);
template void launch_spmv<float>(const DeviceView&, const int*, int, int, int, unsigned long long*);
template void launch_update<double>(const DeviceView&, const int*, int, unsigned long long*);
template void launch_update<float>(const DeviceView&, const int*, int, unsigned long long*);
__global__ __launch_bounds__(NT) void k_record_columns(long long F, const int* __restrict__ fcol,
                                                       float* __restrict__ vals)
{
    const long long k = static_cast<long long>(blockIdx.x) * NT + threadIdx.x;
    if (k < F)
        vals[BsrStore<float>::at(static_cast<size_t>(F), static_cast<size_t>(k), 9)] = __int_as_float(fcol[k]);
}

// The column word of every FP32 full-storage record (upper-only storage
// keeps its columns in a plane of their own, set at creation).
void launch_record_columns(const DeviceView& v, bool f32)
{
    const long long n = v.nnzf;
    if (n <= 0 || v.upper || !f32)
        return;
    const dim3 grid(static_cast<unsigned>((n + NT - 1) / NT));
    hipLaunchKernelGGL(k_record_columns, grid, dim3(NT), 0, v.stream, n, v.fcol, static_cast<float*>(v.vals));
}

template void launch_spmv_plain<double>(const DeviceView&, const double*, double*);
template void launch_spmv_plain<float>(const DeviceView&, const double*, double*);

}  // namespace pk
