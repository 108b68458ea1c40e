// Block-Jacobi PCG kernels for gfx950 (ConjugateGradientSolver.cpp:114-241).
//
// Iteration k is three launches:
//   k_dir(k)     vector tiles. Every workgroup sums the (r.z, r.r) partials
//                the previous launch left for its body, takes the reference's
//                stop decisions for the end of iteration k-1 (k = 0: the CG
//                start) and forms beta = delta_k / delta_{k-1}; it then writes
//                d_k = z_k + beta d_{k-1} for its rows.
//   k_spmv(k)    BSR row tiles, no reduction prologue: q = A d_k on the full
//                BSR (both triangles) and a d.q partial per tile.
//   k_update(k)  vector tiles. Every workgroup sums the body's d.q partials,
//                alpha = delta_k / d.q (or stops on d.q <= 0), updates
//                x += alpha d, r -= alpha q, z = M^-1 r and leaves (r.z, r.r)
//                partials.
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

// d_k = z_k + beta d_{k-1} for the tile's rows, after the stop decisions for
// the end of iteration k-1. At k = 0 the right-hand side already wrote d_0.
template <class M>
__device__ inline void dir_body(int k, const TileView& vt, const M* __restrict__ z, const M* __restrict__ dprev,
                                M* __restrict__ dcur, const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(vt, t, tl))
        return;
    const int g = tl.group;
    GroupState& s = gs[g];
    if (read_uniform(&s.stopAt) >= 0)
        return;
    double rz, rr;
    int flags;
    sum_partials(pbPrev, vt.groupBegin[g], vt.groupCount[g], rz, rr, flags);
    double beta;
    if (stop_rules(k, s, rz, rr, flags, t == vt.groupBegin[g] && threadIdx.x == 0, active, beta))
        return;
    if (k == 0)
        return;
    for (int i = 3 * tl.rowBegin + static_cast<int>(threadIdx.x); i < 3 * tl.rowEnd; i += NT)
        dcur[i] = direction(z[i], dprev[i], beta);
}

template <class M>
__global__ __launch_bounds__(NT) void k_dir(const int* kb, int k0, TileView vt, const M* __restrict__ z,
                                            const M* __restrict__ dprev, M* __restrict__ dcur,
                                            const Partial* __restrict__ pbPrev, GroupState* gs, int* active,
                                            unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    dir_body<M>(k, vt, z, dprev, dcur, pbPrev, gs, active);
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
// What a staging copy reads for element gi of d: d itself (mode 0), or, when
// the SpMV also does the direction update of iteration k > 0 (mode 1),
// z_k + beta d_{k-1}, formed exactly as k_dir forms it and stored to d_k for
// the tile's own rows [own0, own1) (each row has one owner tile).
template <class M>
struct DirSource
{
    const M* __restrict__ d;
    const M* __restrict__ z;
    const M* __restrict__ dprev;
    M* dcur;
    double beta;
    int mode;
    int own0, own1;
    __device__ M get(int gi) const
    {
        if (mode == 0)
            return d[gi];
        const M v = direction(z[gi], dprev[gi], beta);
        if (gi >= own0 && gi < own1)
            dcur[gi] = v;
        return v;
    }
};

template <class M, bool SEG>
__device__ inline StagedDir<M, SEG> stage_segments(const DirSource<M>& src, const StageSegs& seg)
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
            x[j] = i < n ? src.get(i - 3 * sh) : M(0);
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

// The start of an SpMV workgroup, with `fuse` (every tile's staging fits and
// there is no direction launch): the stop rules for the end of iteration k-1
// and beta, taken as k_dir takes them (each workgroup from the same partials;
// the body's first tile records them), then the direction source of the
// staging copy. Without `fuse`: k_dir(k) has decided, and the body is skipped
// when its loop ended at or before k. Returns false when the workgroup has
// nothing to do.
template <class M>
__device__ inline bool spmv_start(int k, int t, const RowTile& tl, const TileView& st, const TileView& vt,
                                  int fuse, const M* __restrict__ d, const M* __restrict__ z,
                                  const M* __restrict__ dprev, M* dcur, const Partial* __restrict__ pbPrev,
                                  GroupState* gs, int* active, DirSource<M>& src)
{
    GroupState& s = gs[tl.group];
    double beta = 0.0;
    if (fuse)
    {
        if (read_uniform(&s.stopAt) >= 0)
            return false;
        double rz, rr;
        int flags;
        sum_partials(pbPrev, vt.groupBegin[tl.group], vt.groupCount[tl.group], rz, rr, flags);
        if (stop_rules(k, s, rz, rr, flags, t == st.groupBegin[tl.group] && threadIdx.x == 0, active, beta))
            return false;
    }
    else
    {
        const int stopAt = read_uniform(&s.stopAt);
        if (stopAt >= 0 && stopAt <= k)
            return false;
    }
    // at k = 0 the right-hand side kernel wrote d_0
    src = DirSource<M>{fuse && k > 0 ? dcur : d, z, dprev, dcur, beta, fuse && k > 0 ? 1 : 0,
                       3 * tl.rowBegin, 3 * tl.rowEnd};
    return true;
}

// The rows of one tile (full storage); returns the lane's d.q partial.
template <class M, int LPR, int U, class Dir>
__device__ __attribute__((always_inline)) inline double full_rows(const RowTile& tl, int2 c2,
                                                                  const int* __restrict__ fstart,
                                                                  const int* __restrict__ fcol,
                                                                  const M* __restrict__ vals, long long nnzf, Dir d,
                                                                  M* __restrict__ q)
{
    const int lane = threadIdx.x % LPR;
    const int sub = threadIdx.x / LPR;
    constexpr int RPP = NT / LPR;
    const size_t F = static_cast<size_t>(nnzf);
    double pd = 0.0;
    // position i of the tile: row rowBegin + i of its first run, then
    // c2.x + (i - len1) of its second (pencil tiles)
    const int len1 = tl.rowEnd - tl.rowBegin, nrows = len1 + c2.y;
    for (int i0 = sub; i0 < nrows; i0 += RPP * U)
    {
        int kb[U], ke[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int i = i0 + u * RPP;
            const int row = i < len1 ? tl.rowBegin + i : c2.x + (i - len1);
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
            const int row = i < len1 ? tl.rowBegin + i : c2.x + (i - len1);
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
// ended at or before k (decided by k_dir(k) or earlier) is skipped. The
// direction blocks come from the tile's segments of d staged in LDS when
// they fit the launch's capacity (as in spmv_body_upper), else from global
// memory.
template <class M, int LPR, int U, bool SEG>
__device__ inline void spmv_body(int k, const TileView& st, const int* __restrict__ fstart,
                                 const int* __restrict__ fcol, const M* __restrict__ vals, long long nnzf,
                                 const M* __restrict__ d, M* __restrict__ q, double* __restrict__ pa,
                                 GroupState* gs, const StageSegs* __restrict__ win, int winCap,
                                 const int2* __restrict__ chunk2, int fuse,
                                 const TileView& vt, const M* __restrict__ z, const M* __restrict__ dprev, M* dcur,
                                 const Partial* __restrict__ pbPrev, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(st, t, tl))
        return;
    StageSegs sg{make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
    if (winCap > 0)
        sg = win[t];  // in flight with the stop flag
    const int2 c2 = chunk2 ? chunk2[t] : make_int2(0, 0);
    DirSource<M> src;
    if (!spmv_start(k, t, tl, st, vt, fuse, d, z, dprev, dcur, pbPrev, gs, active, src))
        return;
    double pd;
    if (winCap > 0 && sg.end.w <= winCap)  // uniform: one tile per workgroup
        pd = full_rows<M, LPR, U>(tl, c2, fstart, fcol, vals, nnzf, stage_segments<M, SEG>(src, sg), q);
    else
        pd = full_rows<M, LPR, U>(tl, c2, fstart, fcol, vals, nnzf, GlobalDir<M>{d}, q);
    double u1 = 0.0, u2 = 0.0;
    block_sum3(pd, u1, u2);
    if (threadIdx.x == 0)
        pa[t] = pd;
}

// Block at position p of a row's full-order block list in upper-only
// storage: a lower position (p < nl) is the transpose of the upper record
// lower[lb + p].x, with column lower[lb + p].y; an upper position is record
// ub + p - nl, with its column in word 9. The transpose is taken by selecting
// elements, so the products formed from `a` are exactly those of full storage.
__device__ inline void load_record(const float* __restrict__ rec, int slot, float* w, int& col)
{
    const float2* r2 = reinterpret_cast<const float2*>(rec) + 5 * static_cast<size_t>(slot);
    const float2 w0 = r2[0], w1 = r2[1], w2 = r2[2], w3 = r2[3], w4 = r2[4];
    w[0] = w0.x; w[1] = w0.y; w[2] = w1.x; w[3] = w1.y; w[4] = w2.x;
    w[5] = w2.y; w[6] = w3.x; w[7] = w3.y; w[8] = w4.x;
    col = __float_as_int(w4.y);
}

__device__ inline void load_record(const double* __restrict__ rec, int slot, double* w, int& col)
{
    const double2* r2 = reinterpret_cast<const double2*>(rec) + 5 * static_cast<size_t>(slot);
    const double2 w0 = r2[0], w1 = r2[1], w2 = r2[2], w3 = r2[3], w4 = r2[4];
    w[0] = w0.x; w[1] = w0.y; w[2] = w1.x; w[3] = w1.y; w[4] = w2.x;
    w[5] = w2.y; w[6] = w3.x; w[7] = w3.y; w[8] = w4.x;
    col = static_cast<int>(__double_as_longlong(w4.y));
}

template <class M>
__device__ inline void load_upper_block(const M* __restrict__ rec, const int2* __restrict__ lower, int lb, int ub,
                                        int nl, int p, M* a, int& col)
{
    const bool tr = p < nl;
    int2 e = make_int2(0, 0);
    if (tr)
        e = lower[lb + p];
    const int slot = tr ? e.x : ub + p - nl;
    M w[9];
    int rc;
    load_record(rec, slot, w, rc);
    col = tr ? e.y : rc;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            a[3 * i + j] = tr ? w[3 * j + i] : w[3 * i + j];
}

// spmv_body over upper-only storage: the same lanes, positions, products and
// reduction order as the full-storage body, each upper record read by its own
// row and again, transposed, by the row of its mirror (at most a layer of
// rows later on the same XCD). With tiles of several passes about half of the
// mirror reads miss L2 (the two reads land in different passes of their
// tiles): the launch moves ~1.5x its upper-storage bytes. Tiles of one pass,
// or every position of a row loaded up front, keep them in L2 (1.0x) but ran
// slower on MI355X: this kernel is bound by its chain of dependent loads and
// by the address unit, not by HBM bytes, and those variants lose occupancy.
// The rows of one tile (upper-only storage); returns the lane's d.q partial.
template <class M, int LPR, int U, class Dir>
__device__ __attribute__((always_inline)) inline double upper_rows(const RowTile& tl, int2 c2,
                                                                   const int2* __restrict__ rowptr,
                                                                   const int2* __restrict__ lower,
                                                                   const M* __restrict__ rec, Dir d,
                                                                   M* __restrict__ q)
{
    const int lane = threadIdx.x % LPR;
    const int sub = threadIdx.x / LPR;
    constexpr int RPP = NT / LPR;
    double pd = 0.0;
    // position i of the tile: row rowBegin + i of its first run, then
    // c2.x + (i - len1) of its second (pencil tiles)
    const int len1 = tl.rowEnd - tl.rowBegin, nrows = len1 + c2.y;
    for (int i0 = sub; i0 < nrows; i0 += RPP * U)
    {
        int kb[U], ke[U], lb[U], ub[U], nl[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int i = i0 + u * RPP;
            const int row = i < len1 ? tl.rowBegin + i : c2.x + (i - len1);
            const bool live = i < nrows;
            const int2 r0 = live ? rowptr[row] : make_int2(0, 0);
            const int2 r1 = live ? rowptr[row + 1] : make_int2(0, 0);
            lb[u] = r0.x;
            ub[u] = r0.y;
            nl[u] = r1.x - r0.x;
            kb[u] = lane;
            ke[u] = (r1.x - r0.x) + (r1.y - r0.y);
        }
        int col[U];
        M a[U][9];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const bool has = kb[u] < ke[u];
            load_upper_block(rec, lower, has ? lb[u] : 0, has ? ub[u] : 0, has ? nl[u] : 0, has ? kb[u] : 0, a[u],
                             col[u]);
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
#pragma unroll
        for (int u = 0; u < U; ++u)
            for (int kk = kb[u] + LPR; kk < ke[u]; kk += LPR)
            {
                M b[9];
                int c;
                load_upper_block(rec, lower, lb[u], ub[u], nl[u], kk, b, c);
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
            const int row = i < len1 ? tl.rowBegin + i : c2.x + (i - len1);
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

// q = A d_k for one tile of upper-only storage and its d.q partial. When the
// tile's segments win[t] (up to four runs of rows covering every column its
// rows reference, their own rows included: on a grid, the layers below, at
// and above the tile) fit the launch's staging capacity, the workgroup first
// copies them from d into LDS with coalesced loads and reads every direction
// block from there: the scattered gathers leave the address unit and the
// chain of dependent global loads. Otherwise it gathers from global memory.
// Products and sums are the same either way.
template <class M, int LPR, int U, bool SEG>
__device__ inline void spmv_body_upper(int k, const TileView& st, const int2* __restrict__ rowptr,
                                       const int2* __restrict__ lower, const M* __restrict__ rec,
                                       const M* __restrict__ d, M* __restrict__ q, double* __restrict__ pa,
                                       GroupState* gs, const StageSegs* __restrict__ win, int winCap,
                                 const int2* __restrict__ chunk2, int fuse,
                                       const TileView& vt, const M* __restrict__ z, const M* __restrict__ dprev,
                                       M* dcur, const Partial* __restrict__ pbPrev, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(st, t, tl))
        return;
    StageSegs sg{make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
    if (winCap > 0)
        sg = win[t];  // in flight with the stop flag
    const int2 c2 = chunk2 ? chunk2[t] : make_int2(0, 0);
    DirSource<M> src;
    if (!spmv_start(k, t, tl, st, vt, fuse, d, z, dprev, dcur, pbPrev, gs, active, src))
        return;
    double pd;
    if (winCap > 0 && sg.end.w <= winCap)  // uniform: one tile per workgroup
        pd = upper_rows<M, LPR, U>(tl, c2, rowptr, lower, rec, stage_segments<M, SEG>(src, sg), q);
    else
        pd = upper_rows<M, LPR, U>(tl, c2, rowptr, lower, rec, GlobalDir<M>{d}, q);
    double u1 = 0.0, u2 = 0.0;
    block_sum3(pd, u1, u2);
    if (threadIdx.x == 0)
        pa[t] = pd;
}

template <class M, int LPR, int U, bool SEG>
__global__ __launch_bounds__(NT) void k_spmv_upper(const int* kb, int k0, TileView st, const int2* __restrict__ rowptr,
                                                   const int2* __restrict__ lower, const M* __restrict__ rec,
                                                   const M* __restrict__ d, M* __restrict__ q,
                                                   double* __restrict__ pa, GroupState* gs,
                                                   const StageSegs* __restrict__ win, int winCap,
                                                   const int2* __restrict__ chunk2, int fuse,
                                                   TileView vt, const M* __restrict__ z,
                                                   const M* __restrict__ dprev, M* dcur,
                                                   const Partial* __restrict__ pbPrev, int* active,
                                                   unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    spmv_body_upper<M, LPR, U, SEG>(k, st, rowptr, lower, rec, d, q, pa, gs, win, winCap, chunk2, fuse, vt, z, dprev,
                                    dcur,
                                    pbPrev, active);
    stamp_end(stamp);
}

template <class M, int LPR, int U, bool SEG>
__global__ __launch_bounds__(NT) void k_spmv(const int* kb, int k0, TileView st, const int* __restrict__ fstart,
                                             const int* __restrict__ fcol, const M* __restrict__ vals,
                                             long long nnzf, const M* __restrict__ d, M* __restrict__ q,
                                             double* __restrict__ pa, GroupState* gs,
                                             const StageSegs* __restrict__ win, int winCap,
                                             const int2* __restrict__ chunk2, int fuse, TileView vt,
                                             const M* __restrict__ z, const M* __restrict__ dprev, M* dcur,
                                             const Partial* __restrict__ pbPrev, int* active,
                                             unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    spmv_body<M, LPR, U, SEG>(k, st, fstart, fcol, vals, nnzf, d, q, pa, gs, win, winCap, chunk2, fuse, vt, z, dprev,
                              dcur,
                              pbPrev, active);
    stamp_end(stamp);
}

template <class M, bool LONG>
__device__ inline void update_body(int k, const TileView& vt, const TileView& st, const M* __restrict__ minv,
                                   int nd, const M* __restrict__ dcur, const M* __restrict__ q,
                                   double* __restrict__ xs, double* __restrict__ rv, M* __restrict__ zv,
                                   const double* __restrict__ pa, Partial* __restrict__ pbCur,
                                   const Partial* __restrict__ pbPrev, GroupState* gs, int* active)
{
    int t;
    RowTile tl;
    if (!tile_for(vt, t, tl))
        return;
    // the thread's first kPre rows (d, q, x, r, M^-1) are in flight while the
    // stop flag and the d.q partials are read
    constexpr int kPre = 2;
    const size_t N = static_cast<size_t>(nd);
    const int row0 = tl.rowBegin + static_cast<int>(threadIdx.x);
    M dpre[kPre][3], qpre[kPre][3], mpre[kPre][9];
    double xpre[kPre][3], rpre[kPre][3];
#pragma unroll
    for (int j = 0; j < kPre; ++j)
    {
        const int row = row0 + j * NT;
        const int rr = row < tl.rowEnd ? row : tl.rowBegin;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            dpre[j][i] = dcur[3 * rr + i];
            qpre[j][i] = q[3 * rr + i];
            xpre[j][i] = xs[3 * rr + i];
            rpre[j][i] = rv[3 * rr + i];
        }
#pragma unroll
        for (int e = 0; e < 9; ++e)
            mpre[j][e] = minv[e * N + rr];
    }
    const int g = tl.group;
    GroupState& s = gs[g];
    const int stopAt = read_uniform(&s.stopAt);
    if (stopAt >= 0 && stopAt <= k)
        return;
    const double dq = LONG ? sum_doubles_long(pa, st.groupBegin[g], st.groupCount[g])
                           : sum_doubles(pa, st.groupBegin[g], st.groupCount[g]);
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
    double pz = 0.0, pr = 0.0;
    int flag = 0;
    // one row: x += alpha d, r -= alpha q, z = M^-1 r, and the partials
    auto one = [&](int row, const M* dd, const M* qq, const double* xx, const double* rq, const M* mm) {
        double r[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[3 * row + i] = xx[i] + static_cast<double>(dd[i]) * alpha;
            r[i] = rq[i] + static_cast<double>(qq[i]) * (-alpha);
            rv[3 * row + i] = r[i];
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
            one(row0 + j * NT, dpre[j], qpre[j], xpre[j], rpre[j], mpre[j]);
    for (int row = row0 + kPre * NT; row < tl.rowEnd; row += NT)
    {
        M dd[3], qq[3], mm[9];
        double xx[3], rq[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            dd[i] = dcur[3 * row + i];
            qq[i] = q[3 * row + i];
            xx[i] = xs[3 * row + i];
            rq[i] = rv[3 * row + i];
        }
#pragma unroll
        for (int e = 0; e < 9; ++e)
            mm[e] = minv[e * N + row];
        one(row, dd, qq, xx, rq, mm);
    }
    double u = 0.0;
    block_sum3(pz, pr, u);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
        pbCur[t] = Partial{pz, pr, flag, 0};
}

template <class M, bool LONG>
__global__ __launch_bounds__(NT) void k_update(const int* kb, int k0, TileView vt, TileView st,
                                               const M* __restrict__ minv, int nd, const M* __restrict__ dcur,
                                               const M* __restrict__ q, double* __restrict__ xs,
                                               double* __restrict__ rv, M* __restrict__ zv,
                                               const double* __restrict__ pa, Partial* __restrict__ pbCur,
                                               const Partial* __restrict__ pbPrev, GroupState* gs, int* active,
                                               unsigned long long* stamps)
{
    const LaunchIndex li = launch_index(kb, k0, stamps);
    const int k = li.k;
    unsigned long long* stamp = li.stamp;
    stamp_begin(stamp);
    update_body<M, LONG>(k, vt, st, minv, nd, dcur, q, xs, rv, zv, pa, pbCur, pbPrev, gs, active);
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
        lo = min(lo, p[2 * i]);
        hi = max(hi, p[2 * i + 1]);
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

// y = A x over upper-only FP32 storage, one thread per row, blocks in
// full-row order (the same sums as k_spmv_plain over full storage).
template <class M>
__global__ __launch_bounds__(NT) void k_spmv_plain_upper(int nd, const int2* __restrict__ rowptr,
                                                         const int2* __restrict__ lower, const M* __restrict__ rec,
                                                         const double* __restrict__ xin, double* __restrict__ y)
{
    const int row = blockIdx.x * NT + threadIdx.x;
    if (row >= nd)
        return;
    const int2 r0 = rowptr[row], r1 = rowptr[row + 1];
    const int nl = r1.x - r0.x, n = nl + (r1.y - r0.y);
    double acc[3] = {0, 0, 0};
    for (int p = 0; p < n; ++p)
    {
        M a[9];
        int c;
        load_upper_block(rec, lower, r0.x, r0.y, nl, p, a, c);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            acc[i] += brow(a, i, xin[3 * c], xin[3 * c + 1], xin[3 * c + 2]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        y[3 * row + i] = acc[i];
}

template <class M, int LPR, int U>
static void launch_spmv_upper_shape(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    const size_t lds = 3 * sizeof(M) * static_cast<size_t>(v.winCap);
    auto kernel = v.stageSegs ? k_spmv_upper<M, LPR, U, true> : k_spmv_upper<M, LPR, U, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.spmv.per)), dim3(NT), lds, v.stream, kb, k0,
                       v.spmv, v.rowptr, v.lower, static_cast<const M*>(v.vals), static_cast<const M*>(v.d[k0 & 1]),
                       static_cast<M*>(v.q), v.pa, v.gs, v.spmvWin, v.winCap, v.spmvChunk2, v.fuseDir, v.vec,
                       static_cast<const M*>(v.z), static_cast<const M*>(v.d[(k0 + 1) & 1]),
                       static_cast<M*>(v.d[k0 & 1]), v.pb[(k0 + 1) & 1], v.active, stamp);
}

// The iteration of a launch is k0 (kb == nullptr) or *kb + k0 with *kb even,
// so k0 alone picks the ping-pong buffers.
template <class M, int LPR, int U>
static void launch_spmv_shape(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    const size_t lds = 3 * sizeof(M) * static_cast<size_t>(v.winCap);
    auto kernel = v.stageSegs ? k_spmv<M, LPR, U, true> : k_spmv<M, LPR, U, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.spmv.per)), dim3(NT), lds, v.stream, kb, k0,
                       v.spmv, v.fstart, v.fcol, static_cast<const M*>(v.vals), v.nnzf,
                       static_cast<const M*>(v.d[k0 & 1]), static_cast<M*>(v.q), v.pa, v.gs, v.spmvWin, v.winCap,
                       v.spmvChunk2, v.fuseDir, v.vec, static_cast<const M*>(v.z), static_cast<const M*>(v.d[(k0 + 1) & 1]),
                       static_cast<M*>(v.d[k0 & 1]), v.pb[(k0 + 1) & 1], v.active, stamp);
}

template <class M>
void launch_dir(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    if (v.vec.count == 0)
        return;
    hipLaunchKernelGGL((k_dir<M>), dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, kb, k0,
                       v.vec, static_cast<const M*>(v.z), static_cast<const M*>(v.d[(k0 + 1) & 1]),
                       static_cast<M*>(v.d[k0 & 1]), v.pb[(k0 + 1) & 1], v.gs, v.active, stamp);
}

template <class M>
void launch_spmv(const DeviceView& v, const int* kb, int k0, int lanesPerRow, int unroll, unsigned long long* stamp)
{
    if (v.spmv.count == 0)
        return;
    if (v.upper)
    {
#define PK_SHAPE(L, U)                                      \
    if (lanesPerRow == L && unroll == U)                    \
    {                                                       \
        launch_spmv_upper_shape<M, L, U>(v, kb, k0, stamp); \
        return;                                             \
    }
        PK_SHAPE(4, 1) PK_SHAPE(4, 2) PK_SHAPE(4, 4)
        PK_SHAPE(8, 1) PK_SHAPE(8, 2) PK_SHAPE(8, 4)
        PK_SHAPE(16, 1) PK_SHAPE(16, 2) PK_SHAPE(16, 4)
        PK_SHAPE(32, 1) PK_SHAPE(32, 2)
#undef PK_SHAPE
        launch_spmv_upper_shape<M, 16, 1>(v, kb, k0, stamp);
        return;
    }
#define PK_SHAPE(L, U)                        \
    if (lanesPerRow == L && unroll == U)      \
    {                                         \
        launch_spmv_shape<M, L, U>(v, kb, k0, stamp); \
        return;                               \
    }
    PK_SHAPE(4, 1) PK_SHAPE(4, 2) PK_SHAPE(4, 4)
    PK_SHAPE(8, 1) PK_SHAPE(8, 2) PK_SHAPE(8, 4)
    PK_SHAPE(16, 1) PK_SHAPE(16, 2) PK_SHAPE(16, 4)
    PK_SHAPE(32, 1) PK_SHAPE(32, 2)
#undef PK_SHAPE
    launch_spmv_shape<M, 16, 1>(v, kb, k0, stamp);
}

template <class M>
void launch_update(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp)
{
    if (v.vec.count == 0)
        return;
    auto kernel = v.longSums ? k_update<M, true> : k_update<M, false>;
    hipLaunchKernelGGL(kernel, dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, kb, k0,
                       v.vec, v.spmv, static_cast<const M*>(v.minv), v.nd, static_cast<const M*>(v.d[k0 & 1]),
                       static_cast<const M*>(v.q), v.xs, v.r, static_cast<M*>(v.z), v.pa, v.pb[k0 & 1],
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
                                                RowTile* __restrict__ spmvPerm)
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
                           v.spmv, vb, vecPerm, spmvPerm);
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

template <class M>
void launch_spmv_plain(const DeviceView& v, const double* x, double* y)
{
    if (v.upper)
    {
        if (v.nd > 0)
            hipLaunchKernelGGL(k_spmv_plain_upper<M>, dim3((v.nd + NT - 1) / NT), dim3(NT), 0, v.stream, v.nd,
                               v.rowptr, v.lower, static_cast<const M*>(v.vals), x, y);
        return;
    }
    if (v.nd > 0)
        hipLaunchKernelGGL((k_spmv_plain<M>), dim3((v.nd + NT - 1) / NT), dim3(NT), 0, v.stream, v.nd, v.fstart,
                           v.fcol, static_cast<const M*>(v.vals), v.nnzf, x, y);
}

template void launch_dir<double>(const DeviceView&, const int*, int, unsigned long long*);
template void launch_dir<float>(const DeviceView&, const int*, int, unsigned long long*);
template void launch_spmv<double>(const DeviceView&, const int*, int, int, int, unsigned long long*);
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

__global__ __launch_bounds__(NT) void k_record_columns64(long long U, const int* __restrict__ ucol,
                                                         double* __restrict__ vals)
{
    const long long k = static_cast<long long>(blockIdx.x) * NT + threadIdx.x;
    if (k < U)
        vals[upper_at(static_cast<size_t>(k), 9)] = __longlong_as_double(static_cast<long long>(ucol[k]));
}

void launch_record_columns(const DeviceView& v, bool f32)
{
    const long long n = v.upper ? v.nnzu : v.nnzf;
    if (n <= 0)
        return;
    const dim3 grid(static_cast<unsigned>((n + NT - 1) / NT));
    if (f32)
        hipLaunchKernelGGL(k_record_columns, grid, dim3(NT), 0, v.stream, n, v.upper ? v.ucol : v.fcol,
                           static_cast<float*>(v.vals));
    else if (v.upper)
        hipLaunchKernelGGL(k_record_columns64, grid, dim3(NT), 0, v.stream, n, v.ucol, static_cast<double*>(v.vals));
}

template void launch_spmv_plain<double>(const DeviceView&, const double*, double*);
template void launch_spmv_plain<float>(const DeviceView&, const double*, double*);

}  // namespace pk
