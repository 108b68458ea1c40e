// Device-side types and helpers shared by the HIP kernels, and the launch
// interface between the host runtime and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace pk
{

constexpr int NT = 256;  // threads per workgroup for every kernel

// FEM model of a tet (the reference's FemModel, FEMModel.h:40-45).
enum FemModelId
{
    FEM_LINEAR = 0,        // K1 with R = I (FEMModel.cpp:559-612)
    FEM_COROTATIONAL = 1,  // the default
    FEM_NONE = 2           // NeoHookean: FEMModel::ComputeForces produces nothing for it
};

// Per-group CG state. No workgroup reads a field that another workgroup of
// the same launch may write: delta is double-buffered by iteration parity,
// and stop decisions are taken identically by every workgroup of the group
// (each sums the same partials in the same order) before one designated
// workgroup records them for later launches and the host.
struct GroupState
{
    double delta[2];  // r.z entering iteration k, at index k & 1
    double alpha[2];  // step length of iteration k at index k & 1 (k_update(k) writes)
    double target;    // tol^2 * delta_0
    double residual;  // reported residual norm
    double tol;
    int maxIt;
    int iters;
    int stopAt;       // launch index that ended the loop (-1 while running)
    int stopInUpdate; // 1 when the update launch of stopAt ended it, 0 for the SpMV launch
    int dirStop;      // launch index whose k_dir ended the loop (-1 otherwise)
    int converged;
    int failed;       // the solve is unusable: do not integrate
    int nonfinite;    // integration produced a non-finite value
    int hasRows;
    int jacobi;
    int integrated;   // the substep's state update of this body is done
};

// One workgroup of a row-tiled kernel: rows [rowBegin, rowEnd) of group
// `group`, and where the group's partials of each tiling lie (vector tiles
// vBegin .. vBegin + vCount - 1, SpMV tiles sBegin .. sBegin + sCount - 1),
// so a CG workgroup finds them in the same load as its rows.
struct RowTile
{
    int group, rowBegin, rowEnd, pad;
    int vBegin, vCount, sBegin, sCount;
};

// Per-workgroup partial sums of a reduction.
struct Partial
{
    double a, b;
    int flag, pad;
};

// The rows of d an SpMV tile references, as up to four ascending,
// disjoint segments of rows [lo[s], lo[s] + len[s]) packed one after another
// in LDS: segment s starts at LDS row end[s-1] (0 for s = 0), and end[3] is
// the rows staged in all. An unused segment has lo = INT_MAX and length 0.
struct StageSegs
{
    int4 lo;
    int4 end;
};

// A tiled launch: tiles of consecutive rows, grouped by body. Workgroup b runs
// tile (b % 8) * per + b / 8 so that each of the 8 XCDs (workgroups are
// dispatched round-robin over them) walks one contiguous eighth of the rows
// and keeps its share of the gathered vectors in its own L2.
//
// A compacted view (perm != nullptr, CG launches of a batch) instead gives
// each workgroup of the full grid its tile directly: perm[b] is the tile of
// workgroup b (its id in `pad`) or has group -1. The compaction kernel fills
// it with the tiles whose body is still iterating, in the same XCD-contiguous
// order over that shorter list, so the bodies left at the end of a batched
// solve are spread over all XCDs instead of the one or two that own their
// rows, and the workgroups past the list return at once.
struct TileView
{
    const RowTile* tiles;
    const int* groupBegin;  // first tile of each group
    const int* groupCount;  // tiles of each group
    int count;              // tiles
    int per;                // ceil(count / 8)
    const RowTile* perm = nullptr;  // compacted: tile of each workgroup (8 * per entries)
};

// Limits of one chunk of the assembly gather by (row, tet) pairs: corner-
// list pairs (one lane each), off-diagonal contributions, rows and upper
// slots; its LDS (kPairLds bytes, 53.0 KB: three workgroups per CU; 54 KB
// admits two) holds the blocks (9 doubles each), the rows' diagonal blocks,
// masses and bodies, and the slots' run descriptors, elements and mirror
// elements. A tet gives 4 pairs and 6 off-diagonal contributions, so 256
// pairs go with 384.
constexpr int kPairChunk = NT;
constexpr int kOffChunk = 384;
constexpr int kRowChunk = 48;
constexpr int kSlotChunk = 224;
constexpr size_t kPairLds = 8 * 9 * static_cast<size_t>(kPairChunk + kOffChunk + kRowChunk) +
                            12 * static_cast<size_t>(kSlotChunk) + 16 * static_cast<size_t>(kRowChunk);
// three workgroups per CU (160 KB of LDS, allocated in blocks of up to 4 KB)
static_assert(kPairLds <= 53248, "pair gather LDS beyond a third of a CU");
// Run descriptor of an upper slot in its chunk: where its contributions lie
// in the chunk's block planes (the diagonal ones: the row's pairs), how many,
// the row within the chunk and whether the slot is the row's diagonal.
__host__ __device__ inline int pack_run(int begin, int len, int row, bool diag)
{
    return begin | (len << 10) | (row << 20) | (diag ? 1 << 26 : 0);
}

// Largest group count the compaction handles (its per-group offsets live in
// LDS); larger batches run uncompacted.
constexpr int kCompactMaxGroups = 4096;
// CG iterations between two compactions of a batched solve.
constexpr int kCompactEvery = 4;

// Value storage of the full BSR. FP64: 9 planes, value e of block k at
// e * F + k (the lanes of a row read 8-byte words of one plane). FP32: one
// record of 10 words per block -- the 9 values, then the block's column as
// int bits -- so a lane reads a block with five 8-byte loads instead of ten
// 4-byte ones.
template <class M> struct BsrStore;
template <> struct BsrStore<double>
{
    static constexpr int kWords = 9;
    __host__ __device__ static size_t at(size_t F, size_t k, int e) { return static_cast<size_t>(e) * F + k; }
};
template <> struct BsrStore<float>
{
    static constexpr int kWords = 10;
    __host__ __device__ static size_t at(size_t, size_t k, int e) { return 10 * k + static_cast<size_t>(e); }
};

// Upper-only storage, both precisions, sliced for one lane per row. The
// dynamic rows of each body are cut into slices of kSliceRows consecutive
// rows (one wavefront). A slice keeps its rows' upper blocks as an ELL block
// of width wu: position k holds the k-th upper block of every row that has
// one (the diagonal, then the columns above it ascending). Full slices keep
// the block of the row at lane l at element ubase + kSliceRows * k + l; a
// slice whose positions are far from full is stored compactly instead: the
// entries of position k follow those of positions 0 .. k - 1 from element
// ubase on, in lane order, so lane l's block is element ubase + (blocks at
// positions before k) + (lanes below l with a block at k), and lanes that
// skip a position leave no gap in the run their wave reads. An element's 9
// values (row-major) are kept in planes of nE elements, so that the lanes
// of a wave read one plane of consecutive elements as one contiguous run:
//   FP32: values 0-3 (16 bytes) | values 4-7 (16 bytes) | value 8
//   FP64: values 0-1 | 2-3 | 4-5 | 6-7 (16 bytes each) | value 8
// and its column is entry e of a separate int32 plane (the pattern's, set
// once). A row's lower blocks (columns c < r, descending) are a second ELL
// block of width wl, laid out the same way from entry lbase on, each entry
// the element of its upper mirror (c, r) (one int32 plane) and c (another).
// Slice descriptor: (ubase, lbase, wu | wl << 16 | compact << 30 |
// aligned << 31, mbase); bit l of mask[mbase + k] says lane l has an upper
// position k, bit l of mask[mbase + wu + k] a lower position k. In an
// aligned slice position k means one stencil offset for every row -- the
// upper block at column r + off[mbase + k], the lower block at column
// r - off[mbase + wu + k] (offsets ascending; a row without that neighbour
// has its bit clear) -- so columns need no load, and on a regular grid the
// transposed mirror of a lower position sits at the same position of the
// mirror row's slice. Slices whose offsets do not line up keep each row's
// blocks packed from position 0 and read their columns from the planes.
// Entries of no block (in full slices, and to start every slice on a
// multiple of 8 entries) are never read.
constexpr int kSliceRows = 64;
struct SliceView
{
    const int4* slices;
    const unsigned long long* mask;
    const int* off;       // stencil offsets of aligned slices, indexed like mask
    const int* mirE;      // lower ELL: element of the upper mirror
    const int* mirC;      // lower ELL: its column
    const int* ecol;      // column of every element
    const int* tileSlice; // first slice of every SpMV tile
    const int2* rowSlice; // (slice, lane) of every row
    long long nE;
};

template <class M>
struct SliceStore
{
    static constexpr size_t kGroup = 16 / sizeof(M);  // values per 16-byte plane entry
    static constexpr size_t kWords = 9;               // values per element
    // index into the value array of value w (0..8) of element e
    __host__ __device__ static size_t word(size_t nE, size_t e, int w)
    {
        if (w < 8)
            return kGroup * ((static_cast<size_t>(w) / kGroup) * nE + e) + static_cast<size_t>(w) % kGroup;
        return 8 * nE + e;
    }
    // the 9 values of element e, each 16-byte plane entry in one store
    __device__ static void put(M* vals, size_t nE, size_t e, const double* v)
    {
        if constexpr (sizeof(M) == 4)
        {
            *reinterpret_cast<float4*>(vals + 4 * e) = make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                                                                   static_cast<float>(v[2]), static_cast<float>(v[3]));
            *reinterpret_cast<float4*>(vals + 4 * (nE + e)) = make_float4(
                static_cast<float>(v[4]), static_cast<float>(v[5]), static_cast<float>(v[6]), static_cast<float>(v[7]));
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *reinterpret_cast<double2*>(vals + 2 * (q * nE + e)) = make_double2(v[2 * q], v[2 * q + 1]);
        }
        vals[8 * nE + e] = static_cast<M>(v[8]);
    }
};

__device__ inline int tile_of(const TileView& tv)
{
    const int b = static_cast<int>(blockIdx.x);
    const int t = (b & 7) * tv.per + (b >> 3);
    return t < tv.count ? t : -1;
}

// Workgroup order for a flat launch over consecutive blocks of work: with a
// grid of 8 * ceil(blocks / 8) workgroups, workgroup b takes block
// (b % 8) * per + b / 8, so each XCD walks one contiguous eighth of the work
// and neighbouring blocks share its L2 (blocks >= the real count find no work).
__device__ inline int xcd_block()
{
    const int per = static_cast<int>(gridDim.x) >> 3;
    const int b = static_cast<int>(blockIdx.x);
    return (b & 7) * per + (b >> 3);
}
inline unsigned xcd_grid(long long blocks) { return 8u * static_cast<unsigned>((blocks + 7) / 8); }

// The tile (id t, rows tl) of this workgroup in a view that may be compacted;
// false when the workgroup has none.
__device__ inline bool tile_for(const TileView& tv, int& t, RowTile& tl)
{
    if (tv.perm != nullptr)
    {
        tl = tv.perm[blockIdx.x];
        t = tl.pad;
        return tl.group >= 0;
    }
    t = tile_of(tv);
    if (t < 0)
        return false;
    tl = tv.tiles[t];
    return true;
}

__device__ inline bool fin(double v) { return __builtin_isfinite(v); }

__device__ inline double norm_of(double rr) { return (fin(rr) && rr >= 0.0) ? sqrt(rr) : INFINITY; }

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup sums of three values, returned to every thread (fixed order).
__device__ inline void block_sum3(double& a, double& b, double& c)
{
    __shared__ double sh[12];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
    {
        sh[w] = a;
        sh[4 + w] = b;
        sh[8 + w] = c;
    }
    __syncthreads();
    a = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    b = (sh[4] + sh[5]) + (sh[6] + sh[7]);
    c = (sh[8] + sh[9]) + (sh[10] + sh[11]);
    __syncthreads();
}

// Sums partials [begin, begin + count) in a fixed order; every workgroup that
// calls this on the same range gets bit-identical totals.
__device__ inline void sum_partials(const Partial* p, int begin, int count, double& a, double& b, int& flag)
{
    double x = 0.0, y = 0.0, z = 0.0;
    int f = 0;
    for (int i = threadIdx.x; i < count; i += NT)
    {
        const Partial q = p[begin + i];
        x += q.a;
        y += q.b;
        f |= q.flag;
    }
    block_sum3(x, y, z);
    a = x;
    b = y;
    flag = __syncthreads_or(f);
}

// The first partial a thread adds in sum_partials (or sum_doubles), loaded
// ahead of the workgroup's other work; the *_from forms continue from it in
// the same order, so their totals are bit-identical to the plain forms'.
__device__ inline Partial first_partial(const Partial* p, int begin, int count)
{
    const int i = static_cast<int>(threadIdx.x);
    return i < count ? p[begin + i] : Partial{0.0, 0.0, 0, 0};
}

__device__ inline double first_double(const double* p, int begin, int count)
{
    const int i = static_cast<int>(threadIdx.x);
    return i < count ? p[begin + i] : 0.0;
}

__device__ inline void sum_partials_from(const Partial& q0, const Partial* p, int begin, int count, double& a,
                                         double& b, int& flag)
{
    double x = 0.0 + q0.a, y = 0.0 + q0.b, z = 0.0;
    int f = q0.flag;
    for (int i = static_cast<int>(threadIdx.x) + NT; i < count; i += NT)
    {
        const Partial q = p[begin + i];
        x += q.a;
        y += q.b;
        f |= q.flag;
    }
    block_sum3(x, y, z);
    a = x;
    b = y;
    flag = __syncthreads_or(f);
}

// sum_doubles for long lists (a large body's SpMV partials): eight loads in
// flight per thread, the adds in the same order, so the total is
// bit-identical to sum_doubles'.
__device__ inline double sum_doubles_long(const double* p, int begin, int count)
{
    double x = 0.0, y = 0.0, z = 0.0;
    int i = threadIdx.x;
    for (; i + 7 * NT < count; i += 8 * NT)
    {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            v[j] = p[begin + i + j * NT];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            x += v[j];
    }
    for (; i < count; i += NT)
        x += p[begin + i];
    block_sum3(x, y, z);
    return x;
}

__device__ inline double sum_doubles(const double* p, int begin, int count)
{
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = threadIdx.x; i < count; i += NT)
        x += p[begin + i];
    block_sum3(x, y, z);
    return x;
}

__device__ inline double sum_doubles_from(double v0, const double* p, int begin, int count)
{
    double x = 0.0 + v0, y = 0.0, z = 0.0;
    for (int i = static_cast<int>(threadIdx.x) + NT; i < count; i += NT)
        x += p[begin + i];
    block_sum3(x, y, z);
    return x;
}

// 3x3 row-major helpers. Column j of m is (m[j], m[3+j], m[6+j]).
__device__ inline double det3(const double* m)
{
    // c0 . (c1 x c2), Mat3.h:52-55
    const double rx = m[4] * m[8] - m[7] * m[5];
    const double ry = m[7] * m[2] - m[1] * m[8];
    const double rz = m[1] * m[5] - m[4] * m[2];
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

// Inverse of a symmetric 3x3 block from its upper triangle, symmetric bit for
// bit (each mirrored entry is a copy): the cofactors as explicit fused
// multiply-adds, so every kernel it is inlined into rounds alike, and the
// Identity below |det| <= 1e-6 as inv3 (Mat3.h:57-70).
__device__ inline void inv3_sym(const double* m, double* o)
{
    const double a00 = m[0], a01 = m[1], a02 = m[2], a11 = m[4], a12 = m[5], a22 = m[8];
    const double c00 = __builtin_fma(a11, a22, -(a12 * a12));
    const double c01 = __builtin_fma(a12, a02, -(a01 * a22));
    const double c02 = __builtin_fma(a01, a12, -(a11 * a02));
    const double c11 = __builtin_fma(a22, a00, -(a02 * a02));
    const double c12 = __builtin_fma(a02, a01, -(a12 * a00));
    const double c22 = __builtin_fma(a00, a11, -(a01 * a01));
    const double d = __builtin_fma(a00, c00, __builtin_fma(a01, c01, a02 * c02));
    if (fabs(d) <= static_cast<double>(0.000001f))
    {
        o[0] = 1.0; o[1] = 0.0; o[2] = 0.0;
        o[3] = 0.0; o[4] = 1.0; o[5] = 0.0;
        o[6] = 0.0; o[7] = 0.0; o[8] = 1.0;
        return;
    }
    const double id = 1.0 / d;
    o[0] = c00 * id;
    o[1] = o[3] = c01 * id;
    o[2] = o[6] = c02 * id;
    o[4] = c11 * id;
    o[5] = o[7] = c12 * id;
    o[8] = c22 * id;
}

// The block-Jacobi inverse of the CG kept as its upper triangle: 6 planes of
// nd values, entries (0,0), (0,1), (0,2), (1,1), (1,2), (2,2).
constexpr int kMinvPlanes = 6;

// z = M^-1 r for row `row` of the 6-plane inverse.
template <class M>
__device__ inline void minv_apply(const M* __restrict__ minv, size_t nd, int row, const double* r, double* z)
{
    double m[kMinvPlanes];
#pragma unroll
    for (int p = 0; p < kMinvPlanes; ++p)
        m[p] = static_cast<double>(minv[p * nd + row]);
    z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
    z[1] = m[1] * r[0] + m[3] * r[1] + m[4] * r[2];
    z[2] = m[2] * r[0] + m[4] * r[1] + m[5] * r[2];
}

// Stores the block-Jacobi inverse mi (row-major, symmetric) of row `row`.
template <class M>
__device__ inline void minv_store(M* __restrict__ minv, size_t nd, int row, const double* mi)
{
    minv[0 * nd + row] = static_cast<M>(mi[0]);
    minv[1 * nd + row] = static_cast<M>(mi[1]);
    minv[2 * nd + row] = static_cast<M>(mi[2]);
    minv[3 * nd + row] = static_cast<M>(mi[4]);
    minv[4 * nd + row] = static_cast<M>(mi[5]);
    minv[5 * nd + row] = static_cast<M>(mi[8]);
}

__device__ inline bool fin9(const double* m)
{
    bool ok = true;
#pragma unroll
    for (int e = 0; e < 9; ++e)
        ok = ok && fin(m[e]);
    return ok;
}

// Device pointers and sizes of one system, passed to the launchers.
struct DeviceView
{
    hipStream_t stream;
    int N, T, nd, G;
    long long nnzu, nnzf;
    // topology
    const int4* tets;
    const unsigned char* live;
    const unsigned char* fem;  // per-tet FEM model (FemModel values), or nullptr: all corotational
    const double* prm;
    double* kdat;  // per-tet stiffness records (KD_COUNT doubles each)
    const int* cstart;
    const int* clist;
    const int* uFull;
    const int* uMirror;
    const int* slotRow;
    const int* diagSlot;  // nd: the upper slot of each row's diagonal block
    const int* rowGroup;
    const int* fstart;
    const int* fcol;
    const int* fdiag;  // full storage: each row's full index of its diagonal block
    // upper-only storage (upper != 0), sliced (SliceStore, SliceView), and
    // the element of every upper slot
    int upper;
    long long nE;
    SliceView slice;
    const int* slotEll;
    const int* ucol;
    const int* nstart;
    const int* nlist;
    // the assembly gather by (row, tet) pairs: each chunk's rows, pairs and
    // slots (nchunks 0: the gather by rows, k_assemble), each upper slot's run in
    // its chunk, and per nlist pair the chunk-relative places of its
    // off-diagonal contributions
    const int4* chunks;  // nchunks: first row, first pair, first upper slot, rows | slots << 8 | pairs << 16
    int nchunks;
    const int* slotRun;  // nnzu run descriptors (pack_run)
    const int* pairOff;
    const int* rowNode;
    const int* astart;     // node -> (tet, corner) over live tets
    const int* alist;
    const int* nodeRow;    // node -> dynamic row or -1
    const int* nodeGroup;  // node -> body
    const double* mrow;
    // state and work arrays
    const int* cpos;     // 4 T: where the element kernel writes each dynamic corner
    double* x;
    double* v;
    double* basis;       // per-tet R g_a and R R^T planes
    double* corner;      // forces and K v in row order
    double* adj;         // reverse-mode corner terms, N_ADJ per alist entry
    const int* apos;     // 4 T: alist index of corner a of tet t, or -1
    void* minv;   // M: block-Jacobi inverse, kMinvPlanes planes of nd (upper triangle)
    double* xs;   // CG solution dv
    double* r;
    void* vals;   // M
    void* z;      // M
    void* d[4];   // M, d_k at index k & 3
    void* q;      // M
    GroupState* gs;
    int* active;  // bodies still iterating; active[1], active[2]: vector and SpMV tiles of the last compaction
    int* kbase;        // base iteration of a replayed CG chunk (zeroed by the reset)
    // point connections (none when nconn == 0)
    int nconn;
    const int4* connNodes;
    const double* connData;  // 9 per connection: weights (4), target (3), k, c
    double* connEval;        // 7 per connection: force (3), point velocity (3), ok
    int nconnRows;           // dynamic rows the connections touch
    const int* connRows;
    const int* connRowStart; // nconnRows + 1 offsets into connEntries
    const int2* connEntries; // (connection, corner a) in connection order
    double* crhs;            // nd x 3: the connections' right-hand side (or force) per row
    const int* connNodeStart;  // N + 1 offsets into connNodeList
    const int2* connNodeList;  // node -> (connection, corner a), w_a != 0, in connection order
    double* connAdj;           // 6 per connection: reverse-mode terms before the corner weight
    const double* cslot;     // per upper slot: sum of k w_a w_b (null without connections)
    // collision spheres (none when nsph == 0): per substep, each candidate
    // tet's contact terms, then their sums per upper slot and per row on top
    // of the point connections'
    int nsph;
    const double* sph;       // 6 per sphere: centre (3), radius, k, c
    const unsigned char* tact;  // T: the tet's active flag
    int ncand;               // candidate tets: active, four active nodes
    const int* cand;
    double* ck;              // per candidate: sum of k
    double* cg;              // 3 per candidate: rhs term (or force) before the 1/4 weight
    const int* csStart;      // nnzu + 1 offsets into csList
    const int* csList;       // upper slot -> candidates, once per ordered pair landing there
    const int* crStart;      // nd + 1 offsets into crList
    const int* crList;       // dynamic row -> candidates, once per corner on that row
    double* sslot;           // per upper slot: cslot + the contacts' k/16
    double* srhs;            // nd x 3: crhs + the contacts' terms / 4
    const int* candNodeStart;  // N + 1 offsets into candNodeList
    const int* candNodeList;   // node -> candidate tets, once per corner, in candidate order
    double* candAdj;           // 6 per candidate: reverse-mode terms each corner receives
    // what the assembly adds per upper slot and the right-hand side (or the
    // explicit force) per row: sslot/srhs with spheres, else cslot/crhs with
    // connections, else null
    const double* slotAdd;
    const double* rowAdd;
    // SpMV staging: per SpMV tile the segments of d its blocks reference; a
    // tile whose segments hold at most winCap rows in all reads d from LDS
    // (winCap 0: never)
    const StageSegs* spmvWin;
    int winCap;
    int stageSegs;  // 0: every tile's staging is one window (segment 0 only)
    // reductions; longSums: some body has 4 * NT SpMV tiles or more (its
    // update workgroups sum long partial lists)
    int longSums;
    TileView vec, spmv;
    double* pa;        // per SpMV tile: d.q
    Partial* pb[2];    // per vector tile: (r.z, r.r, flags), by iteration parity
};

// Launchers (pk_assembly.hip, pk_cg.hip). M is the storage type of A and of
// the vectors the SpMV reads (double, or float for FP32 storage).
void launch_reset(const DeviceView& v, double tol, int maxIt);
// g_a . g_b planes of the tet parameters (after every full upload of them).
void launch_gdot(const DeviceView& v, double* prm);
// Every tet's element outputs at the current state as 114-double records
// (10 blocks, 4 forces, 4 K v), for inspection.
void launch_element_records(const DeviceView& v, double* out);
void launch_stiff_records(const DeviceView& v);
void launch_element(const DeviceView& v, double dt, int implicit);
template <class M> void launch_assemble(const DeviceView& v, double dt);
template <class M> void launch_rhs(const DeviceView& v, double gx, double gy, double gz, double dt);
void launch_integrate(const DeviceView& v, double dt, int commit);
void launch_mark_integrated(const DeviceView& v);
void launch_explicit(const DeviceView& v, double gx, double gy, double gz, double dt);
// Point connections: per-connection force and point velocity at the current
// state, then each touched row's right-hand side terms (implicit) or force
// (explicit) into crhs.
void launch_connections(const DeviceView& v, double dt, int implicit);
// Collision spheres: each candidate tet's contacts at the current state, then
// sslot and srhs (no-op without spheres). After launch_connections.
void launch_spheres(const DeviceView& v, double dt, int implicit);
// Zeros the element outputs of the n tets listed (switched off in place).
void launch_zero_tets(const DeviceView& v, int n, const int* ids);
// live[ids[i]] = liveVal[i] and tact[ids[i]] = actVal[i] for i < n.
void launch_tet_flags(hipStream_t stream, int n, const int* ids, const unsigned char* liveVal,
                      const unsigned char* actVal, unsigned char* live, unsigned char* tact);
// cslot[idx[i]] = val ? val[i] : 0 for i < n.
void launch_slot_scatter(hipStream_t stream, int n, const int* idx, const double* val, double* cslot);
// crhs[3 r .. 3 r + 2] = 0 for the n rows listed.
void launch_zero_rows(hipStream_t stream, int n, const int* rows, double* crhs);
// out[i] = (float) x[i] and out[n3 + i] = (float) v[i] for i < n3 (a null
// source is skipped).
void launch_export_frame(hipStream_t stream, long long n3, const double* x, const double* v, float* out);
template <class M> void launch_dir(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp);
// LDS bytes an SpMV workgroup of this shape may stage d in (pk_cg.hip).
template <class M>
size_t spmv_stage_bytes(bool upper, int lanesPerRow, int unroll);
template <class M>
void launch_spmv(const DeviceView& v, const int* kb, int k0, int lanesPerRow, int unroll,
                 unsigned long long* stamp);
template <class M> void launch_update(const DeviceView& v, const int* kb, int k0, unsigned long long* stamp);
// *kb += by (the base of a replayed CG chunk).
void launch_advance(const DeviceView& v, int* kb, int by);
// Profiling stamps: a CG launch writes (start, end) wall-clock ticks of each
// of its grid workgroups into 2 * grid slots: the slots passed (kb ==
// nullptr), or block k of the buffer passed (replayed chunks, iteration k);
// the reduce turns `launches` consecutive such blocks into one (start, end)
// pair each.
int spmv_grid(const DeviceView& v);
int update_grid(const DeviceView& v);
// Writes the compacted workgroup-to-tile maps of the bodies still iterating
// (see TileView), 8 * per entries for each of the two tilings.
void launch_compact(const DeviceView& v, RowTile* vecPerm, RowTile* spmvPerm);
void launch_stamp_reduce(hipStream_t stream, const unsigned long long* raw, int grid, int launches,
                         unsigned long long* span);
// Zeroes n stamp words (a workgroup that does not run leaves its slots zero).
void launch_zero_stamps(hipStream_t stream, unsigned long long* raw, long long n);
template <class M> void launch_spmv_plain(const DeviceView& v, const double* x, double* y);
// Record storage (FP32, or upper-only in either precision): writes each
// block's column into word 9 of its record (one record per full block, or
// per upper slot for upper-only storage).
void launch_record_columns(const DeviceView& v, bool f32);

// Reverse mode (pk_adjoint.hip). `ok` holds one flag per body: the forward
// substep integrated. The CG of A lam = u runs through launch_spmv/update.
void launch_record_ok(const DeviceView& v, int* ok);
template <class M>
void launch_adj_rhs(const DeviceView& v, const int* ok, const double* xbar, const double* vbar, double dt);
// Saves / restores the per-body assembly rejection flags (G ints).
void launch_save_failed(const DeviceView& v, int* out);
void launch_restore_failed(const DeviceView& v, const int* in);
void launch_adj_element(const DeviceView& v, const int* ok, const double* vnext, double dt, double* gradLogE);
void launch_adj_nodes(const DeviceView& v, const int* ok, double dt, double* xbar, double* vbar);
void launch_adj_contacts(const DeviceView& v, double dt);

}  // namespace pk
