// pk_system runtime: device buffers, the substep driver and the pk_* C ABI.
//
// One substep (World.cpp:48-70) is issued on the system's own stream:
//   reset -> element -> assemble -> rhs -> CG launches -> integrate (check,
//   commit), or the explicit integrator in explicit mode.
// The CG loop is launched in chunks; after each chunk the host queues a copy
// of the count of bodies still iterating and checks the previous chunk's copy,
// so the GPU never waits on the host and at most two chunks run past the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <map>
#include <set>
#include <array>
#include <vector>

#include "physik_amd.h"
#include "pk_device.h"
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

template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
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
    // alloc, unless the buffer already holds exactly `count` elements
    int ensure(size_t count) { return (p && n == count) ? PK_OK : alloc(count); }
    void free()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int upload(const std::vector<T>& h, hipStream_t st)
    {
        const int rc = alloc(h.size());
        if (rc != PK_OK)
            return rc;
        if (!h.empty())
            PK_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        return PK_OK;
    }
};

struct Tiles
{
    DevBuf<RowTile> tiles;
    DevBuf<int> groupBegin, groupCount;
    int count = 0;
    int per = 0;
    int longest = 0;             // most tiles of one group
    std::vector<int> hostBegin;  // first tile of each group

    TileView view() const { return TileView{tiles.p, groupBegin.p, groupCount.p, count, per}; }
};

struct Profile
{
    bool on = false;
    long long launches[PK_K_COUNT] = {};
    double ms[PK_K_COUNT] = {};
    long long spmvBlocks = 0, spmvUpper = 0, spmvRows = 0, spmvBodyLaunches = 0;  // work of the counted SpMV launches
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
    int lanesPerRow = 16;
    int spmvRows = 128;
    int vecRows = 1024;
    int unroll = 2;
    hipStream_t stream = nullptr;

    pk::DevBuf<int4> tets;
    pk::DevBuf<unsigned char> live, tact;
    pk::DevBuf<unsigned char> fem;  // per-tet FEM model; not allocated while every tet is corotational
    pk::DevBuf<double> prm, kdat, x, v, basis, corner, adj, mrow, xs, r, pa;
    pk::DevBuf<char> minv;  // storage type, like the CG vectors
    pk::DevBuf<int> cstart, clist, uFull, uMirror, slotRow, rowGroup, fstart, fcol, nstart, nlist, rowNode, active;
    pk::DevBuf<int> slotRun, pairOff;  // the assembly gather by (row, tet) pairs
    pk::DevBuf<int4> gatherChunks;
    int nchunks = 0;
    pk::DevBuf<int> astart, alist, apos, nodeRow, nodeGroup, cpos, diagSlot;
    // sliced upper-only storage: slice descriptors, lane masks, lower ELL
    // entries, the element of each upper slot, each row's (slice, lane) and
    // each SpMV tile's first slice
    pk::DevBuf<int> fdiag;  // full storage: full index of each row's diagonal block
    pk::DevBuf<int4> slices;
    pk::DevBuf<unsigned long long> smask;
    pk::DevBuf<int2> rowSlice;
    pk::DevBuf<int> slotEll, tileSlice, ecol, soff, mirE, mirC;
    std::vector<int> groupSlice;  // first slice of each body
    long long nE = 0;
    int nSlices = 0, nCompact = 0;  // slices of upper-only storage, compact ones
    pk::DevBuf<int> ucol;
    pk::DevBuf<char> vals, z, d0, d1, d2, d3, q;  // element type set by the precision
    pk::DevBuf<pk::Partial> pb0, pb1;
    // batched solves: the tile of each CG workgroup, over the bodies still
    // iterating (vector and SpMV tilings), rebuilt every few iterations
    pk::DevBuf<pk::RowTile> vecPerm, spmvPerm;
    pk::DevBuf<pk::GroupState> gs;
    pk::DevBuf<int> kbase;
    // profiling: per-workgroup (start, end) ticks of every CG launch, and
    // the per-launch spans reduced from them
    pk::DevBuf<unsigned long long> rawSpmv, rawUpd, rawDir, stampSpmv, stampUpd, stampDir;
    double wallClockKHz = 0.0;
    pk::Tiles vecTiles, spmvTiles;
    // SpMV staging (upper-only storage): the column window of every SpMV tile
    // and the rows of d a workgroup stages in LDS (0: no staging)
    pk::DevBuf<pk::StageSegs> spmvWin;
    // host copy of the SpMV tiling for the staging set-up
    std::vector<pk::RowTile> spmvHostTiles;
    int winCap = 0;
    int stageCap = 0;  // rows the staging capacity holds (0: staging off)
    int stageSegs = 0;
    int stagingMode = 1;  // 0 off, 1 staging
    int* pinned = nullptr;
    hipEvent_t pollEv[2] = {nullptr, nullptr};

    // frame export (pk_system_export_begin): float copies of x and v, made on
    // the device and copied into pinned host memory; the copy completes at
    // frameEv. stateVer changes with every write of x or v, so an export made
    // before it is stale.
    float* framePinned = nullptr;
    pk::DevBuf<float> frameDev;
    hipEvent_t frameEv = nullptr;
    int frameWhat = 0;
    unsigned long long stateVer = 0, frameVer = ~0ull;

    // reverse mode: checkpoints (x, v) every `tapeEvery` substeps of a
    // recorded run and after its last substep, the per-body integrate flags
    // of each segment's last substep, the states of the segment being
    // reversed, per-body integrate flags of that segment, the adjoints and
    // the gradient
    pk::DevBuf<double> tapeX, tapeV, segX, segV, xbar, vbar, gradE, xEnd, vEnd;
    pk::DevBuf<int> tapeOk, segOk;
    // the matrices of the segment being reversed (A and M^-1 of every
    // substep, and the assembly's rejection flags), kept from its replay when
    // they fit storeCap bytes, so the reverse step need not assemble again
    pk::DevBuf<char> segVals, segMinv;
    pk::DevBuf<int> segFailed;
    long long storeCap = 8LL << 30;
    int segColumns = -1;  // epoch whose FP32 record columns segVals holds
    char* valsView = nullptr;  // when set, the matrix the kernels use
    char* minvView = nullptr;
    double tapeDt = 0.0;
    int tapeSteps = 0, tapeEvery = 0;
    bool tapeValid = false;
    long long tapeId = 0;  // counts the recorded runs of this system
    int lastItersAdj = 0;
    int firstStopAdj = 0;  // the previous adjoint solve's earliest stop among its bodies

    double gravity[3] = {0, 0, 0};
    double tol = 1e-4;
    int maxIt = 128;
    int jacobi = 1;
    int implicit = 1;
    int lastIters = 0;
    int firstStop = 0;  // the previous forward solve's earliest stop among its bodies
    std::vector<pk::GroupState> hostGs;
    std::vector<long long> groupBlocks, groupUpper, groupRows;  // full and upper BSR blocks, dynamic rows per group
    pk::Profile prof;

    // point connections in force (pk_system_set_connections): their data, the
    // rows and upper slots they touch (kept to clear them when replaced)
    int nconn = 0, nconnRows = 0, nslots = 0;
    pk::DevBuf<int4> connNodes;
    pk::DevBuf<double> connData, connEval, crhs, cslot, slotVal;
    pk::DevBuf<int> connRows, connRowStart, slotIdx;
    pk::DevBuf<int2> connEntries;

    // collision spheres in force (pk_system_set_spheres); the candidate tets
    // and their slot and row lists are built on first use (ncand < 0 before)
    int nsph = 0, ncand = -1;
    pk::DevBuf<double> sph, ck, cg, sslot, srhs;
    pk::DevBuf<int> cand, csStart, csList, crStart, crList;
    // reverse mode over contacts: per node its connections and candidate
    // tets, and their per-substep terms
    pk::DevBuf<int> connNodeStart, candNodeStart, candNodeList;
    pk::DevBuf<int2> connNodeList;
    pk::DevBuf<double> connAdj, candAdj;

    // CG chunks replayed from captured HIP graphs: the first chunk of a solve
    // (full tiles) and every later one (compaction first, for a batch). A
    // capture bakes in buffer addresses and tile shapes: `epoch` changes when
    // they do, and the stamp buffers in use are part of the key.
    bool useGraphs = false;
    int graphMode = 0;     // pk_system_set_graphs: 0 off, 1 CG chunks, 2 whole forward substeps
    int epoch = 0;
    long long cfgVer = 0;  // changes with every setting a captured substep bakes in
    struct ChunkGraphs
    {
        hipGraphExec_t first = nullptr, next = nullptr;
        int epoch = -1;
        const void* stamps[3] = {nullptr, nullptr, nullptr};
        const void* matrix[2] = {nullptr, nullptr};
    } chunks;
    // one implicit forward substep captured whole: reset, assembly,
    // right-hand side, `chunks` CG chunks and the integration of the bodies
    // whose solve ended within them
    struct StepGraph
    {
        hipGraphExec_t exec = nullptr;
        int epoch = -1;
        long long cfgVer = -1;
        double dt = 0.0;
        int chunks = 0;
        const void* matrix[2] = {nullptr, nullptr};
    } stepGraph;

    void drop_graphs()
    {
        for (hipGraphExec_t* g : {&chunks.first, &chunks.next, &stepGraph.exec})
            if (*g)
            {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
        chunks.epoch = -1;
        stepGraph.epoch = -1;
    }

    ~pk_system()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto e : prof.pool)
            (void)hipEventDestroy(e);
        for (auto e : pollEv)
            if (e)
                (void)hipEventDestroy(e);
        if (pinned)
            (void)hipHostFree(pinned);
        if (framePinned)
            (void)hipHostFree(framePinned);
        if (frameEv)
            (void)hipEventDestroy(frameEv);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    pk::DeviceView view() const
    {
        pk::DeviceView dv{};
        dv.stream = stream;
        dv.N = topo.N;
        dv.T = topo.T;
        dv.nd = topo.nd;
        dv.G = topo.G;
        dv.nnzu = topo.nnzu();
        dv.nnzf = topo.nnzf();
        dv.tets = tets.p;
        dv.live = live.p;
        dv.fem = fem.p;
        dv.tact = tact.p;
        dv.prm = prm.p;
        dv.kdat = kdat.p;
        dv.cstart = cstart.p;
        dv.clist = clist.p;
        dv.uFull = uFull.p;
        dv.uMirror = uMirror.p;
        dv.slotRow = slotRow.p;
        dv.diagSlot = diagSlot.p;
        dv.rowGroup = rowGroup.p;
        dv.fstart = fstart.p;
        dv.fdiag = fdiag.p;
        dv.upper = (precision == PK_FP32_UPPER || precision == PK_FP64_UPPER) ? 1 : 0;
        dv.nE = nE;
        dv.slice = pk::SliceView{slices.p, smask.p, soff.p, mirE.p, mirC.p, ecol.p, tileSlice.p, rowSlice.p, nE};
        dv.slotEll = slotEll.p;
        dv.ucol = ucol.p;
        dv.fcol = fcol.p;
        dv.nstart = nstart.p;
        dv.nlist = nlist.p;
        dv.chunks = gatherChunks.p;
        dv.nchunks = nchunks;
        dv.slotRun = slotRun.p;
        dv.pairOff = pairOff.p;
        dv.rowNode = rowNode.p;
        dv.astart = astart.p;
        dv.alist = alist.p;
        dv.apos = apos.p;
        dv.nodeRow = nodeRow.p;
        dv.nodeGroup = nodeGroup.p;
        dv.mrow = mrow.p;
        dv.x = x.p;
        dv.v = v.p;
        dv.basis = basis.p;
        dv.corner = corner.p;
        dv.adj = adj.p;
        dv.cpos = cpos.p;
        dv.minv = minvView ? minvView : minv.p;
        dv.xs = xs.p;
        dv.r = r.p;
        dv.vals = valsView ? valsView : vals.p;
        dv.z = z.p;
        dv.d[0] = d0.p;
        dv.d[1] = d1.p;
        dv.d[2] = d2.p;
        dv.d[3] = d3.p;
        dv.q = q.p;
        dv.gs = gs.p;
        dv.active = active.p;
        dv.kbase = kbase.p;
        dv.nconn = nconn;
        dv.connNodes = connNodes.p;
        dv.connData = connData.p;
        dv.connEval = connEval.p;
        dv.nconnRows = nconnRows;
        dv.connRows = connRows.p;
        dv.connRowStart = connRowStart.p;
        dv.connNodeStart = connNodeStart.p;
        dv.connNodeList = connNodeList.p;
        dv.connAdj = connAdj.p;
        dv.candNodeStart = candNodeStart.p;
        dv.candNodeList = candNodeList.p;
        dv.candAdj = candAdj.p;
        dv.connEntries = connEntries.p;
        dv.crhs = crhs.p;
        dv.cslot = cslot.p;
        dv.nsph = nsph;
        dv.sph = sph.p;
        dv.ncand = std::max(0, ncand);
        dv.cand = cand.p;
        dv.ck = ck.p;
        dv.cg = cg.p;
        dv.csStart = csStart.p;
        dv.csList = csList.p;
        dv.crStart = crStart.p;
        dv.crList = crList.p;
        dv.sslot = sslot.p;
        dv.srhs = srhs.p;
        dv.slotAdd = nsph > 0 ? sslot.p : (nconn > 0 ? cslot.p : nullptr);
        dv.rowAdd = nsph > 0 ? srhs.p : (nconn > 0 ? crhs.p : nullptr);
        dv.vec = vecTiles.view();
        dv.spmv = spmvTiles.view();
        dv.longSums = spmvTiles.longest >= 4 * pk::NT ? 1 : 0;
        dv.spmvWin = spmvWin.p;
        dv.winCap = winCap;
        dv.stageSegs = stageSegs;
        dv.pa = pa.p;
        dv.pb[0] = pb0.p;
        dv.pb[1] = pb1.p;
        return dv;
    }
};

namespace pk
{

// One tiling of the dynamic rows: runs of `rows` consecutive rows of one
// body, with the first tile and the tile count of each body.
struct HostTiling
{
    std::vector<RowTile> tiles;
    std::vector<int> gb, gc;
};

static HostTiling tile_rows(const Topology& t, int rows)
{
    HostTiling h;
    h.gb.assign(static_cast<size_t>(t.G), 0);
    h.gc.assign(static_cast<size_t>(t.G), 0);
    for (int g = 0; g < t.G; ++g)
    {
        h.gb[g] = static_cast<int>(h.tiles.size());
        for (int r = t.groupRowBegin[g]; r < t.groupRowEnd[g]; r += rows)
            h.tiles.push_back(RowTile{g, r, std::min(r + rows, t.groupRowEnd[g]), 0, 0, 0, 0, 0});
        h.gc[g] = static_cast<int>(h.tiles.size()) - h.gb[g];
    }
    return h;
}

static int upload_tiling(pk_system* s, Tiles& tt, const HostTiling& h)
{
    tt.count = static_cast<int>(h.tiles.size());
    tt.per = (tt.count + 7) / 8;
    tt.longest = h.gc.empty() ? 0 : *std::max_element(h.gc.begin(), h.gc.end());
    tt.hostBegin = h.gb;
    int rc;
    if ((rc = tt.tiles.upload(h.tiles, s->stream)) != PK_OK) return rc;
    if ((rc = tt.groupBegin.upload(h.gb, s->stream)) != PK_OK) return rc;
    return tt.groupCount.upload(h.gc, s->stream);
}

static bool is_upper(const pk_system* s);

// The vector tiling (vecRows rows per tile) and the SpMV tiling (spmvRows;
// with sliced upper-only storage a multiple of kSliceRows, so every tile is
// whole slices: tileSlice holds its first). Every tile of either records
// its body's partial ranges in both.
static int build_tilings(pk_system* s)
{
    const Topology& t = s->topo;
    HostTiling hv = tile_rows(t, s->vecRows), hs = tile_rows(t, s->spmvRows);
    for (HostTiling* h : {&hv, &hs})
        for (RowTile& tl : h->tiles)
        {
            tl.vBegin = hv.gb[tl.group];
            tl.vCount = hv.gc[tl.group];
            tl.sBegin = hs.gb[tl.group];
            tl.sCount = hs.gc[tl.group];
        }
    int rc;
    if (is_upper(s))
    {
        std::vector<int> first;
        first.reserve(hs.tiles.size());
        for (const RowTile& tl : hs.tiles)
            first.push_back(s->groupSlice[tl.group] + (tl.rowBegin - t.groupRowBegin[tl.group]) / kSliceRows);
        if ((rc = s->tileSlice.upload(first, s->stream)) != PK_OK)
            return rc;
    }
    if ((rc = upload_tiling(s, s->vecTiles, hv)) != PK_OK) return rc;
    s->spmvHostTiles = hs.tiles;
    return upload_tiling(s, s->spmvTiles, hs);
}

static int upload_groups(pk_system* s)
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
        st.stopAt = 0;
    }
    const int rc = s->gs.upload(s->hostGs, s->stream);
    if (rc != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
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

// Accumulates the recorded launches; CG launches issued after every body had
// stopped did no work and are left out.
static void harvest_profile(pk_system* s)
{
    Profile& p = s->prof;
    if (!p.on)
        return;
    (void)hipStreamSynchronize(s->stream);
    int spmvWork = 0, updWork = 0;
    for (const auto& g : s->hostGs)
        if (g.hasRows)
        {
            spmvWork = std::max(spmvWork, g.stopAt + g.stopInUpdate);
            updWork = std::max(updWork, g.stopAt);
        }
    // CG kernels: span between the first workgroup's start and the last one's
    // end, from the wall clock read inside the kernels
    if (s->stampSpmv.p && s->wallClockKHz > 0.0)
    {
        // launch k of the direction kernel ran for every k <= the last stop;
        // SpMV and update launches k < spmvWork (updWork) did work
        int dirWork = 0;
        for (const auto& g : s->hostGs)
            if (g.hasRows)
                dirWork = std::max(dirWork, g.stopAt + 1);
        const int n = std::max(dirWork, std::max(spmvWork, updWork));
        const size_t bytes = 2 * static_cast<size_t>(n) * sizeof(unsigned long long);
        std::vector<unsigned long long> a(2 * static_cast<size_t>(n)), b(a.size()), c(a.size());
        if (n > 0 && hipMemcpy(a.data(), s->stampSpmv.p, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(b.data(), s->stampUpd.p, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(c.data(), s->stampDir.p, bytes, hipMemcpyDeviceToHost) == hipSuccess)
        {
            auto add = [&](int kernel, const std::vector<unsigned long long>& v, int k) {
                if (v[2 * k + 1] > v[2 * k])
                {
                    p.ms[kernel] += static_cast<double>(v[2 * k + 1] - v[2 * k]) / s->wallClockKHz;
                    p.launches[kernel] += 1;
                }
            };
            for (size_t g = 0; g < s->hostGs.size(); ++g)
                if (s->hostGs[g].hasRows)
                {
                    // SpMV launches k < stopAt (+1 when the loop ended in the
                    // update) multiplied this body's rows
                    const long long c = s->hostGs[g].stopAt + s->hostGs[g].stopInUpdate;
                    p.spmvBlocks += c * s->groupBlocks[g];
                    p.spmvUpper += c * s->groupUpper[g];
                    p.spmvRows += c * s->groupRows[g];
                    p.spmvBodyLaunches += c;
                }
            for (int k = 0; k < n; ++k)
            {
                if (k < spmvWork)
                    add(PK_K_SPMV, a, k);
                if (k < updWork)
                    add(PK_K_UPDATE, b, k);
                if (k < dirWork)
                    add(PK_K_DIRECTION, c, k);
            }
        }
    }
    for (const auto& rec : p.recs)
    {
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

// reset -> element (-> contacts) -> off-diagonal slots -> rows. `mode`:
// kAsmExplicit the element forces only (explicit integration);
// kAsmImplicit the matrix, the right-hand side and the CG start;
// kAsmMatrix the matrix only (a reverse step brings its own right-hand side).
enum AssemblyMode
{
    kAsmExplicit = 0,
    kAsmImplicit = 1,
    kAsmMatrix = 2
};

template <class M>
static int run_assembly(pk_system* s, double dt, double tol, int maxIt, int mode)
{
    const DeviceView dv = s->view();
    launch_reset(dv, tol, maxIt);
    {
        Bracket br(s, PK_K_ELEMENT);
        launch_element(dv, dt, mode != kAsmExplicit);
    }
    if (s->nconn > 0 || s->nsph > 0)
    {
        Bracket br(s, PK_K_CONTACT);
        launch_connections(dv, dt, mode != kAsmExplicit);
        launch_spheres(dv, dt, mode != kAsmExplicit);
    }
    if (s->topo.nd == 0 || mode == kAsmExplicit)
        return PK_OK;
    {
        Bracket br(s, PK_K_ASSEMBLE);
        launch_assemble<M>(dv, dt);
    }
    if (mode == kAsmImplicit)
    {
        Bracket br(s, PK_K_RHS);
        launch_rhs<M>(dv, s->gravity[0], s->gravity[1], s->gravity[2], dt);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

// Captures one CG chunk: kChunk iterations (direction, SpMV, update each) at
// iterations *kbase + 0 .. kChunk-1, then *kbase += kChunk; a batch's later
// chunks first rebuild the compacted tile maps.
constexpr int kChunk = kCompactEvery;

template <class M>
static int capture_chunk(pk_system* s, const DeviceView& dv, bool compacted, unsigned long long* stS,
                         unsigned long long* stU, unsigned long long* stD, hipGraphExec_t& exec)
{
    PK_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    DeviceView cv = dv;
    if (compacted)
    {
        launch_compact(dv, s->vecPerm.p, s->spmvPerm.p);
        cv.vec.perm = s->vecPerm.p;
        cv.spmv.perm = s->spmvPerm.p;
    }
    for (int o = 0; o < kChunk; ++o)
    {
        launch_dir<M>(cv, dv.kbase, o, stD);
        launch_spmv<M>(cv, dv.kbase, o, s->lanesPerRow, s->unroll, stS);
        launch_update<M>(cv, dv.kbase, o, stU);
    }
    launch_advance(dv, dv.kbase, kChunk);
    const hipError_t launchErr = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t endErr = hipStreamEndCapture(s->stream, &g);
    if (launchErr != hipSuccess || endErr != hipSuccess)
    {
        if (g)
            (void)hipGraphDestroy(g);
        set_error(std::string("CG chunk capture failed: ") +
                  hipGetErrorString(launchErr != hipSuccess ? launchErr : endErr));
        return PK_ERR_HIP;
    }
    const hipError_t instErr = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PK_HIP(instErr);
    return PK_OK;
}

// The CG launches up to maxIt; lastIters (the previous solve's count of the
// same kind) sizes the first run of launches before the host starts polling,
// and firstStop (its earliest stop among the bodies) tells when compactions
// can start to pay.
// Launched one by one, or (useGraphs) replayed kChunk iterations at a time
// from captured graphs: the same kernels in the same order either way, so the
// results are identical.
template <class M>
static int run_cg(pk_system* s, int maxItArg, int lastIters, int firstStop)
{
    const DeviceView dv = s->view();
    const int maxIt = std::max(0, maxItArg);
    const int gS = spmv_grid(dv), gU = update_grid(dv);
    // launches k = 0 .. maxIt, rounded up to whole chunks
    const int launches = (maxIt + kChunk) / kChunk * kChunk;
    unsigned long long* stS = nullptr;
    unsigned long long* stU = nullptr;
    unsigned long long* stD = nullptr;
    if (s->prof.on)
    {
        const size_t n = static_cast<size_t>(launches);
        if (s->rawSpmv.n < n * 2 * gS || s->rawUpd.n < n * 2 * gU || s->rawDir.n < n * 2 * gU ||
            s->stampSpmv.n < 2 * n)
        {
            int rc;
            if ((rc = s->rawSpmv.alloc(n * 2 * gS)) != PK_OK) return rc;
            if ((rc = s->rawUpd.alloc(n * 2 * gU)) != PK_OK) return rc;
            if ((rc = s->rawDir.alloc(n * 2 * gU)) != PK_OK) return rc;
            if ((rc = s->stampSpmv.alloc(2 * n)) != PK_OK) return rc;
            if ((rc = s->stampUpd.alloc(2 * n)) != PK_OK) return rc;
            if ((rc = s->stampDir.alloc(2 * n)) != PK_OK) return rc;
            // slots a launch does not write stay zero (launches cut to the
            // tiles still iterating have fewer workgroups)
            launch_zero_stamps(s->stream, s->rawSpmv.p, static_cast<long long>(s->rawSpmv.n));
            launch_zero_stamps(s->stream, s->rawUpd.p, static_cast<long long>(s->rawUpd.n));
            launch_zero_stamps(s->stream, s->rawDir.p, static_cast<long long>(s->rawDir.n));
        }
        stS = s->rawSpmv.p;
        stU = s->rawUpd.p;
        stD = s->rawDir.p;
    }
    const bool compact = s->spmvPerm.p != nullptr;
    if (s->useGraphs)
    {
        auto& cg = s->chunks;
        if (cg.epoch != s->epoch || cg.stamps[0] != stS || cg.stamps[1] != stU || cg.stamps[2] != stD ||
            cg.matrix[0] != dv.vals || cg.matrix[1] != dv.minv)
        {
            s->drop_graphs();
            int rc;
            if ((rc = capture_chunk<M>(s, dv, false, stS, stU, stD, cg.first)) != PK_OK)
                return rc;
            if (compact && (rc = capture_chunk<M>(s, dv, true, stS, stU, stD, cg.next)) != PK_OK)
                return rc;
            cg.epoch = s->epoch;
            cg.stamps[0] = stS;
            cg.stamps[1] = stU;
            cg.stamps[2] = stD;
            cg.matrix[0] = dv.vals;
            cg.matrix[1] = dv.minv;
        }
    }
    // launch k runs the direction, SpMV and update of iteration k (the
    // direction kernel first decides the end of k-1); k = maxIt only takes
    // that decision, and launches past it find every body stopped; a batch
    // switches to the compacted tile maps at the first compaction (every
    // kChunk iterations)
    // after a compaction, a poll also brings the count of tiles still
    // iterating, and later launches cover only those (the count only falls,
    // so a count one chunk old is enough)
    // A compaction (every kCompactEvery launches) only pays once a body has
    // stopped: it runs when a poll has shown fewer bodies iterating than the
    // last map holds or, before the first poll, from a few launches before
    // the previous solve's earliest stop. (It only changes which workgroups
    // hold which tiles, never a result.)
    DeviceView cv = dv;
    int k = 0, poll = 0;
    bool compacted = false, pollCompacted[2] = {false, false};
    int knownActive = -1, mapActive = INT_MAX;  // last polled count (-1: none yet); bodies in the last map
    int chunk = std::max(kChunk, lastIters - 4);
    while (k <= maxIt)
    {
        const int end = std::min(maxIt + 1, k + chunk);
        if (s->useGraphs)
        {
            for (; k < end; k += kChunk)
                PK_HIP(hipGraphLaunch(k == 0 || !compact ? s->chunks.first : s->chunks.next, s->stream));
        }
        else
        {
            for (; k < end; ++k)
            {
                const bool due = knownActive >= 0 ? knownActive < mapActive : k + kCompactEvery >= firstStop;
                if (compact && k > 0 && k % kCompactEvery == 0 && due)
                {
                    launch_compact(dv, s->vecPerm.p, s->spmvPerm.p);
                    cv.vec.perm = s->vecPerm.p;
                    cv.spmv.perm = s->spmvPerm.p;
                    compacted = true;
                    if (knownActive >= 0)
                        mapActive = knownActive;
                }
                launch_dir<M>(cv, nullptr, k, stD ? stD + 2 * static_cast<size_t>(gU) * k : nullptr);
                if (k < maxIt)
                    launch_spmv<M>(cv, nullptr, k, s->lanesPerRow, s->unroll,
                                   stS ? stS + 2 * static_cast<size_t>(gS) * k : nullptr);
                if (k < maxIt)
                    launch_update<M>(cv, nullptr, k, stU ? stU + 2 * static_cast<size_t>(gU) * k : nullptr);
            }
        }
        chunk = kChunk;
        const int slot = poll & 1;
        PK_HIP(hipMemcpyAsync(s->pinned + 4 * slot, s->active.p, 3 * sizeof(int), hipMemcpyDeviceToHost,
                              s->stream));
        pollCompacted[slot] = compacted;
        PK_HIP(hipEventRecord(s->pollEv[slot], s->stream));
        if (poll > 0)
        {
            PK_HIP(hipEventSynchronize(s->pollEv[slot ^ 1]));
            const int* pv = s->pinned + 4 * (slot ^ 1);
            if (pv[0] == 0)
                break;
            knownActive = pv[0];
            if (pollCompacted[slot ^ 1] && !s->useGraphs)
            {
                cv.vec.per = std::max(1, std::min(cv.vec.per, (pv[1] + 7) / 8));
                cv.spmv.per = std::max(1, std::min(cv.spmv.per, (pv[2] + 7) / 8));
            }
        }
        ++poll;
    }
    if (stS)
    {
        // graph chunks ran the SpMV and update of every launch they covered
        const int nS = s->useGraphs ? k : std::min(k, maxIt);
        launch_stamp_reduce(s->stream, stD, gU, k, s->stampDir.p);
        launch_stamp_reduce(s->stream, stS, gS, nS, s->stampSpmv.p);
        launch_stamp_reduce(s->stream, stU, gU, nS, s->stampUpd.p);
        const long long used = 2ll * k;
        launch_zero_stamps(s->stream, stS, used * gS);
        launch_zero_stamps(s->stream, stU, used * gU);
        launch_zero_stamps(s->stream, stD, used * gU);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

// One substep; `failedOut` (reverse-mode replays) receives the assembly's
// per-body rejection flags.
// Captures one implicit forward substep: the assembly's launches, `chunks`
// CG chunks (as capture_chunk records them, one after another, the device
// iteration base advancing) and the integration of every body whose solve
// ended within them.
template <class M>
static int capture_substep(pk_system* s, const DeviceView& dv, double dt, int chunks, hipGraphExec_t& exec)
{
    PK_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    launch_reset(dv, s->tol, s->maxIt);
    launch_element(dv, dt, 1);
    if (s->nconn > 0 || s->nsph > 0)
    {
        launch_connections(dv, dt, 1);
        launch_spheres(dv, dt, 1);
    }
    launch_assemble<M>(dv, dt);
    launch_rhs<M>(dv, s->gravity[0], s->gravity[1], s->gravity[2], dt);
    DeviceView cv = dv;
    const bool compact = s->spmvPerm.p != nullptr;
    for (int c = 0; c < chunks; ++c)
    {
        if (compact && c > 0)
        {
            launch_compact(dv, s->vecPerm.p, s->spmvPerm.p);
            cv.vec.perm = s->vecPerm.p;
            cv.spmv.perm = s->spmvPerm.p;
        }
        for (int o = 0; o < kChunk; ++o)
        {
            launch_dir<M>(cv, dv.kbase, o, nullptr);
            launch_spmv<M>(cv, dv.kbase, o, s->lanesPerRow, s->unroll, nullptr);
            launch_update<M>(cv, dv.kbase, o, nullptr);
        }
        launch_advance(dv, dv.kbase, kChunk);
    }
    launch_integrate(dv, dt, 0);
    launch_integrate(dv, dt, 1);
    launch_mark_integrated(dv);
    const hipError_t launchErr = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t endErr = hipStreamEndCapture(s->stream, &g);
    if (launchErr != hipSuccess || endErr != hipSuccess)
    {
        if (g)
            (void)hipGraphDestroy(g);
        set_error(std::string("substep capture failed: ") +
                  hipGetErrorString(launchErr != hipSuccess ? launchErr : endErr));
        return PK_ERR_HIP;
    }
    const hipError_t instErr = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PK_HIP(instErr);
    return PK_OK;
}

// One implicit forward substep from a captured graph (pk_system_set_graphs
// mode 2). The graph holds enough CG chunks for the previous solve's
// iterations and one more; it is captured again when the settings, the
// buffers or that count change. A body still iterating when the graph ends
// is finished by launches from the host (the same kernels at the same
// iterations), then integrated: the results are those of run_cg's launches.
template <class M>
static int run_substep_graph(pk_system* s, double dt)
{
    const DeviceView dv = s->view();
    const int launches = std::max(0, s->maxIt) + 1;
    const int most = (launches + kChunk - 1) / kChunk;
    const int want = std::min(most, (std::min(std::max(s->lastIters, 0) + 1, launches) + kChunk - 1) / kChunk + 1);
    auto& g = s->stepGraph;
    const bool same = g.exec && g.epoch == s->epoch && g.cfgVer == s->cfgVer && g.dt == dt &&
                      g.matrix[0] == dv.vals && g.matrix[1] == dv.minv;
    if (!same || g.chunks < want || g.chunks > want + 4)
    {
        if (g.exec)
        {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
        int rc;
        if ((rc = capture_substep<M>(s, dv, dt, want, g.exec)) != PK_OK)
            return rc;
        g.epoch = s->epoch;
        g.cfgVer = s->cfgVer;
        g.dt = dt;
        g.chunks = want;
        g.matrix[0] = dv.vals;
        g.matrix[1] = dv.minv;
    }
    PK_HIP(hipGraphLaunch(g.exec, s->stream));
    // bodies still iterating: the remaining launches, checked every chunk
    const bool compact = s->spmvPerm.p != nullptr;
    DeviceView cv = dv;
    if (compact && g.chunks > 1)
    {
        cv.vec.perm = s->vecPerm.p;
        cv.spmv.perm = s->spmvPerm.p;
    }
    int k = g.chunks * kChunk;
    bool tail = false;
    while (true)
    {
        PK_HIP(hipMemcpyAsync(s->pinned, s->active.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        PK_HIP(hipEventRecord(s->pollEv[0], s->stream));
        PK_HIP(hipEventSynchronize(s->pollEv[0]));
        if (s->pinned[0] == 0 || k > s->maxIt)
            break;
        tail = true;
        for (const int end = std::min(s->maxIt + 1, k + kChunk); k < end; ++k)
        {
            if (compact && k % kCompactEvery == 0)
            {
                launch_compact(dv, s->vecPerm.p, s->spmvPerm.p);
                cv.vec.perm = s->vecPerm.p;
                cv.spmv.perm = s->spmvPerm.p;
            }
            launch_dir<M>(cv, nullptr, k, nullptr);
            if (k < s->maxIt)
            {
                launch_spmv<M>(cv, nullptr, k, s->lanesPerRow, s->unroll, nullptr);
                launch_update<M>(cv, nullptr, k, nullptr);
            }
        }
    }
    if (tail)
    {
        launch_integrate(dv, dt, 0);
        launch_integrate(dv, dt, 1);
        launch_mark_integrated(dv);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

template <class M>
static int run_substep(pk_system* s, double dt, int* failedOut = nullptr)
{
    if (s->graphMode == 2 && s->implicit && !failedOut && !s->prof.on && s->topo.nd > 0)
        return run_substep_graph<M>(s, dt);
    int rc = run_assembly<M>(s, dt, s->tol, s->maxIt, s->implicit ? kAsmImplicit : kAsmExplicit);
    if (rc != PK_OK || s->topo.nd == 0)
        return rc;
    const DeviceView dv = s->view();
    if (failedOut)
        launch_save_failed(dv, failedOut);
    if (s->implicit)
    {
        if ((rc = run_cg<M>(s, s->maxIt, s->lastIters, s->firstStop)) != PK_OK)
            return rc;
        Bracket br(s, PK_K_INTEGRATE);
        launch_integrate(dv, dt, 0);
        launch_integrate(dv, dt, 1);
        launch_mark_integrated(dv);
    }
    else
    {
        Bracket br(s, PK_K_INTEGRATE);
        launch_explicit(dv, s->gravity[0], s->gravity[1], s->gravity[2], dt);
    }
    PK_HIP(hipGetLastError());
    return PK_OK;
}

// Copies the per-body CG state to the host; `lastIters` receives the largest
// iteration count (it sizes the next solve of the same kind) and `firstStop`
// the smallest among the bodies with rows (when the next solve's compactions
// start to pay).
static int read_groups(pk_system* s, int& lastIters, int& firstStop)
{
    if (s->hostGs.empty())
        return PK_OK;
    PK_HIP(hipMemcpyAsync(s->hostGs.data(), s->gs.p, s->hostGs.size() * sizeof(GroupState), hipMemcpyDeviceToHost,
                          s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    int mx = 0, mn = INT_MAX;
    for (const auto& g : s->hostGs)
    {
        mx = std::max(mx, g.iters);
        if (g.hasRows)
            mn = std::min(mn, g.iters);
    }
    lastIters = mx;
    firstStop = mn == INT_MAX ? 0 : mn;
    return PK_OK;
}

static int choose_lanes(const Topology& t)
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

// A, M^-1 and the SpMV vectors in FP32 (full or upper-only storage).
static bool is_f32(const pk_system* s) { return s->precision == PK_FP32_STORAGE || s->precision == PK_FP32_UPPER; }
// A stored as its upper triangle, one record per slot.
static bool is_upper(const pk_system* s) { return s->precision == PK_FP32_UPPER || s->precision == PK_FP64_UPPER; }

// Sliced upper-only storage (SliceStore, SliceView): per body, slices of
// kSliceRows rows; per slice its upper and lower ELL blocks (aligned to the
// slice's stencil offsets where those line up, else packed), their lane
// masks and offsets; the element and column of every upper slot; and for
// every lower position the element of its upper mirror and the mirror's row.
static int build_slices(pk_system* s, const std::vector<int>& slotRow)
{
    const Topology& t = s->topo;
    constexpr int W = kSliceRows;
    // a slice aligns when its distinct offsets need at most this many more
    // positions than its longest row
    constexpr int kAlignSlack = 2;
    // a slice is stored compactly when its blocks fill less than this share
    // of its 64 (wu + wl) positions; every slice starts on a multiple of
    // kSliceAlign entries
    constexpr double kCompactBelow = 0.97;
    constexpr long long kSliceAlign = 8;
    std::vector<int4> desc;
    std::vector<unsigned long long> masks;
    std::vector<int> offs;
    std::vector<int2> rowSlice(static_cast<size_t>(std::max(t.nd, 1)), make_int2(0, 0));
    std::vector<int> slotEll(static_cast<size_t>(std::max<long long>(t.nnzu(), 1)), -1);
    // lower ELL position of each lower block, by its full index
    std::vector<int> lowPos(static_cast<size_t>(std::max<long long>(t.nnzf(), 1)), -1);
    s->groupSlice.assign(static_cast<size_t>(t.G), 0);
    s->nCompact = 0;
    long long nE = 0, nL = 0;
    std::vector<int> up, lo;
    for (int g = 0; g < t.G; ++g)
    {
        s->groupSlice[g] = static_cast<int>(desc.size());
        for (int r0 = t.groupRowBegin[g]; r0 < t.groupRowEnd[g]; r0 += W)
        {
            const int n = std::min(W, t.groupRowEnd[g] - r0);
            int mu = 0, ml = 0;  // longest row's upper and lower counts
            up.clear();
            lo.clear();
            for (int l = 0; l < n; ++l)
            {
                const int r = r0 + l;
                const int f0 = t.fstart[r], f1 = t.fstart[r + 1];
                const int nu = t.ustart[r + 1] - t.ustart[r], nlr = (f1 - f0) - nu;
                mu = std::max(mu, nu);
                ml = std::max(ml, nlr);
                for (int k = f0; k < f1; ++k)
                    (t.fcol[k] >= r ? up : lo).push_back(t.fcol[k] >= r ? t.fcol[k] - r : r - t.fcol[k]);
            }
            std::sort(up.begin(), up.end());
            up.erase(std::unique(up.begin(), up.end()), up.end());
            std::sort(lo.begin(), lo.end());
            lo.erase(std::unique(lo.begin(), lo.end()), lo.end());
            const bool aligned = static_cast<int>(up.size()) <= mu + kAlignSlack &&
                                 static_cast<int>(lo.size()) <= ml + kAlignSlack;
            const int wu = aligned ? static_cast<int>(up.size()) : mu;
            const int wl = aligned ? static_cast<int>(lo.size()) : ml;
            nE = (nE + kSliceAlign - 1) / kSliceAlign * kSliceAlign;
            nL = (nL + kSliceAlign - 1) / kSliceAlign * kSliceAlign;
            if (wu > 0xffff || wl > 0x3fff || nE + static_cast<long long>(W) * wu > INT_MAX ||
                nL + static_cast<long long>(W) * wl > INT_MAX || masks.size() + wu + wl > static_cast<size_t>(INT_MAX))
            {
                set_error("upper-only storage: the pattern exceeds the sliced layout's limits");
                return PK_ERR_ARG;
            }
            const int sl = static_cast<int>(desc.size());
            const int mbase = static_cast<int>(masks.size());
            desc.push_back(make_int4(static_cast<int>(nE), static_cast<int>(nL),
                                     wu | (wl << 16) | (aligned ? static_cast<int>(0x80000000u) : 0), mbase));
            masks.resize(masks.size() + wu + wl, 0ull);
            offs.resize(masks.size(), 0);
            if (aligned)
            {
                std::copy(up.begin(), up.end(), offs.begin() + mbase);
                std::copy(lo.begin(), lo.end(), offs.begin() + mbase + wu);
            }
            // positions first, then entries: in a compact slice the entries
            // of one position follow those of the positions before it, lane
            // l's at its position's base plus the number of lanes below l
            // that hold a block there; otherwise position k of lane l is
            // entry base + 64 k + l
            for (int l = 0; l < n; ++l)
            {
                const int r = r0 + l;
                rowSlice[r] = make_int2(sl, l);
                const int f0 = t.fstart[r], f1 = t.fstart[r + 1];
                const int nlr = (f1 - f0) - (t.ustart[r + 1] - t.ustart[r]);
                // upper slots ustart[r] + k (columns ascending)
                for (int k = 0; k < t.ustart[r + 1] - t.ustart[r]; ++k)
                {
                    const int u = t.ustart[r] + k;
                    const int pos = aligned ? static_cast<int>(std::lower_bound(up.begin(), up.end(), t.ucol[u] - r) -
                                                               up.begin())
                                            : k;
                    masks[mbase + pos] |= 1ull << l;
                    slotEll[u] = pos;
                }
                // lower blocks at full indices f0 .. f0 + nlr - 1 (columns ascending);
                // packed: the k-th below the diagonal at position k
                for (int m = f0; m < f0 + nlr; ++m)
                {
                    const int pos = aligned ? static_cast<int>(std::lower_bound(lo.begin(), lo.end(), r - t.fcol[m]) -
                                                               lo.begin())
                                            : (f0 + nlr - 1 - m);
                    masks[mbase + wu + pos] |= 1ull << l;
                    lowPos[m] = pos;
                }
            }
            long long used = 0;
            for (int k = 0; k < wu + wl; ++k)
                used += __builtin_popcountll(masks[mbase + k]);
            const bool compact = static_cast<double>(used) < kCompactBelow * W * (wu + wl);
            if (compact)
            {
                desc.back().z |= 1 << 30;
                ++s->nCompact;
            }
            std::vector<long long> ubase(static_cast<size_t>(wu) + 1, nE), lbase(static_cast<size_t>(wl) + 1, nL);
            for (int k = 0; k < wu; ++k)
                ubase[k + 1] = ubase[k] + (compact ? __builtin_popcountll(masks[mbase + k]) : W);
            for (int k = 0; k < wl; ++k)
                lbase[k + 1] = lbase[k] + (compact ? __builtin_popcountll(masks[mbase + wu + k]) : W);
            for (int l = 0; l < n; ++l)
            {
                const int r = r0 + l;
                const unsigned long long below = (1ull << l) - 1ull;
                for (int u = t.ustart[r]; u < t.ustart[r + 1]; ++u)
                {
                    const int pos = slotEll[u];
                    slotEll[u] = static_cast<int>(ubase[pos] +
                                                  (compact ? __builtin_popcountll(masks[mbase + pos] & below) : l));
                }
                const int f0 = t.fstart[r];
                const int nlr = (t.fstart[r + 1] - f0) - (t.ustart[r + 1] - t.ustart[r]);
                for (int m = f0; m < f0 + nlr; ++m)
                {
                    const int pos = lowPos[m];
                    lowPos[m] = static_cast<int>(
                        lbase[pos] + (compact ? __builtin_popcountll(masks[mbase + wu + pos] & below) : l));
                }
            }
            nE = ubase[wu];
            nL = lbase[wl];
        }
    }
    // the lower block at full index m = uMirror[u] is the transposed mirror of
    // upper slot u
    std::vector<int> mirE(static_cast<size_t>(std::max<long long>(nL, 1)), -1),
        mirC(static_cast<size_t>(std::max<long long>(nL, 1)), 0);
    long long placed = 0;
    for (long long u = 0; u < t.nnzu(); ++u)
    {
        const int m = t.uMirror[u];
        if (m < 0)
            continue;
        const int idx = m < t.nnzf() ? lowPos[m] : -1;
        if (idx < 0 || mirE[idx] != -1 || slotEll[u] < 0)
        {
            set_error("upper-only storage: inconsistent mirror map");
            return PK_ERR_ARG;
        }
        mirE[idx] = slotEll[u];
        mirC[idx] = slotRow[u];
        ++placed;
    }
    for (long long u = 0; u < t.nnzu(); ++u)
        if (slotEll[u] < 0)
        {
            set_error("upper-only storage: an upper slot outside every body's rows");
            return PK_ERR_ARG;
        }
    if (placed != t.nnzf() - t.nnzu())
    {
        set_error("upper-only storage: a lower block has no upper mirror");
        return PK_ERR_ARG;
    }
    s->nE = nE;
    s->nSlices = static_cast<int>(desc.size());
    std::vector<int> ecol(static_cast<size_t>(std::max<long long>(nE, 1)), 0);
    for (long long u = 0; u < t.nnzu(); ++u)
        ecol[slotEll[u]] = t.ucol[u];
    for (auto& v : mirE)
        v = std::max(v, 0);  // padding entries (never read)
    const hipStream_t st = s->stream;
    int rc;
    if ((rc = s->ecol.upload(ecol, st)) != PK_OK) return rc;
    if ((rc = s->slices.upload(desc, st)) != PK_OK) return rc;
    if ((rc = s->smask.upload(masks, st)) != PK_OK) return rc;
    if ((rc = s->soff.upload(offs, st)) != PK_OK) return rc;
    if ((rc = s->mirE.upload(mirE, st)) != PK_OK) return rc;
    if ((rc = s->mirC.upload(mirC, st)) != PK_OK) return rc;
    if ((rc = s->rowSlice.upload(rowSlice, st)) != PK_OK) return rc;
    return s->slotEll.upload(slotEll, st);
}

static int alloc_work(pk_system* s)
{
    const Topology& t = s->topo;
    const size_t es = is_f32(s) ? sizeof(float) : sizeof(double);
    const size_t nd3 = static_cast<size_t>(t.nd) * 3;
    int rc;
    if ((rc = s->basis.alloc(static_cast<size_t>(t.T) * EB_COUNT)) != PK_OK) return rc;
    if ((rc = s->corner.alloc(t.nlist.size() * N_CORNER)) != PK_OK) return rc;
    const size_t words = is_upper(s) ? SliceStore<double>::kWords : is_f32(s) ? 10 : BsrStore<double>::kWords;
    const size_t blocks = static_cast<size_t>(is_upper(s) ? s->nE : t.nnzf());
    if ((rc = s->vals.alloc(blocks * words * es)) != PK_OK) return rc;
    if (is_upper(s) && blocks)  // the padding elements of the slices stay zero (never read)
        PK_HIP(hipMemsetAsync(s->vals.p, 0, blocks * words * es, s->stream));
    if ((rc = s->z.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->d0.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->d1.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->d2.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->d3.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->q.alloc(nd3 * es)) != PK_OK) return rc;
    if ((rc = s->minv.alloc(static_cast<size_t>(t.nd) * pk::kMinvPlanes * es)) != PK_OK) return rc;
    if ((rc = s->xs.alloc(nd3)) != PK_OK) return rc;
    if ((rc = s->r.alloc(nd3)) != PK_OK) return rc;
    if ((rc = s->active.alloc(4)) != PK_OK) return rc;
    if ((rc = s->kbase.alloc(1)) != PK_OK) return rc;
    ++s->epoch;
    return PK_OK;
}

static bool is_f32(const pk_system* s);

// Staging of the SpMV tiles (same order as build_tilings). Each tile's
// columns span a window [first, last]; when every window fits the LDS share of
// the SpMV kernel the system's shape selects (spmv_stage_bytes: as much as
// leaves every workgroup its registers allow on a CU) the
// tiles stage their windows (one segment, the cheapest lookup). Otherwise
// each tile stages its distinct columns cut at the three widest gaps between
// them into up to four segments (on a grid: the layers below, at and above
// the tile), and tiles whose segments still exceed the capacity gather from
// global memory. winCap is the widest staging kept, at most the capacity.
static int setup_windows(pk_system* s)
{
    const Topology& t = s->topo;
    s->winCap = 0;
    s->stageSegs = 0;
    s->stageCap = 0;
    if (s->stagingMode == 0 || t.nd == 0)
        return PK_OK;
    const size_t bytesPerRow = 3 * (is_f32(s) ? sizeof(float) : sizeof(double));
    const size_t share = is_f32(s) ? pk::spmv_stage_bytes<float>(is_upper(s), s->lanesPerRow, s->unroll)
                                   : pk::spmv_stage_bytes<double>(is_upper(s), s->lanesPerRow, s->unroll);
    const int cap = static_cast<int>(share / bytesPerRow);
    s->stageCap = cap;
    std::vector<pk::StageSegs> wins, segs;
    std::vector<int> cols;
    std::vector<std::pair<int, int>> gaps;  // (gap rows, index of the column before it)
    int widestWin = 0, widestSeg = 0;
    for (size_t ti = 0; ti < s->spmvHostTiles.size(); ++ti)
        {
            const pk::RowTile& tl = s->spmvHostTiles[ti];
            cols.assign(t.fcol.begin() + t.fstart[tl.rowBegin], t.fcol.begin() + t.fstart[tl.rowEnd]);
            for (int r = tl.rowBegin; r < tl.rowEnd; ++r)
                cols.push_back(r);
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            const int len = cols.back() - cols.front() + 1;
            wins.push_back(pk::StageSegs{make_int4(cols.front(), INT_MAX, INT_MAX, INT_MAX),
                                         make_int4(len, len, len, len)});
            widestWin = std::max(widestWin, len);
            gaps.clear();
            for (size_t i = 0; i + 1 < cols.size(); ++i)
                if (cols[i + 1] - cols[i] > 1)
                    gaps.push_back({cols[i + 1] - cols[i] - 1, static_cast<int>(i)});
            const size_t cuts = std::min<size_t>(3, gaps.size());
            std::partial_sort(gaps.begin(), gaps.begin() + cuts, gaps.end(),
                              [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                                  return a.first != b.first ? a.first > b.first : a.second < b.second;
                              });
            std::vector<int> at;  // index of the last column of each segment
            for (size_t c = 0; c < cuts; ++c)
                at.push_back(gaps[c].second);
            std::sort(at.begin(), at.end());
            at.push_back(static_cast<int>(cols.size()) - 1);
            int lo[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX}, end[4] = {0, 0, 0, 0};
            int first = 0, total = 0;
            for (size_t k = 0; k < 4; ++k)
            {
                if (k < at.size())
                {
                    lo[k] = cols[first];
                    total += cols[at[k]] - cols[first] + 1;
                    first = at[k] + 1;
                }
                end[k] = total;
            }
            segs.push_back(pk::StageSegs{make_int4(lo[0], lo[1], lo[2], lo[3]),
                                         make_int4(end[0], end[1], end[2], end[3])});
            widestSeg = std::max(widestSeg, total);
        }
    if (widestWin <= cap)
    {
        s->winCap = widestWin;
        return s->spmvWin.upload(wins, s->stream);
    }
    s->stageSegs = 1;
    s->winCap = std::min(widestSeg, cap);
    return s->spmvWin.upload(segs, s->stream);
}

static int setup_tiles(pk_system* s)
{
    int rc;
    ++s->epoch;
    if ((rc = build_tilings(s)) != PK_OK) return rc;
    if ((rc = setup_windows(s)) != PK_OK) return rc;
    if ((rc = s->pa.alloc(std::max(1, s->spmvTiles.count))) != PK_OK) return rc;
    if ((rc = s->pb0.alloc(std::max(1, s->vecTiles.count))) != PK_OK) return rc;
    if ((rc = s->pb1.alloc(std::max(1, s->vecTiles.count))) != PK_OK) return rc;
    if (s->topo.G > 1 && s->topo.G <= kCompactMaxGroups)
    {
        if ((rc = s->vecPerm.alloc(std::max(8, 8 * s->vecTiles.per))) != PK_OK) return rc;
        if ((rc = s->spmvPerm.alloc(std::max(8, 8 * s->spmvTiles.per))) != PK_OK) return rc;
    }
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

// The assembly gather by (row, tet) pairs (k_assemble_pairs): consecutive
// rows grouped into chunks of at most kPairChunk corner-list pairs,
// kOffChunk off-diagonal contributions, kRowChunk rows and kSlotChunk upper
// slots; each upper slot's run in its chunk (pack_run: the diagonal slot's
// contributions are the row's pairs, an off-diagonal slot's lie in the
// chunk's run of the contribution lists with the diagonal slots left out);
// and per pair the chunk-relative places of its contributions to the row's
// upper off-diagonal slots, one 10-bit field per other corner of the tet in
// increasing order (0x3ff: none). A row too large for a chunk leaves
// nchunks 0, and the assembly gathers by rows (k_assemble).
static int build_pair_gather(pk_system* s, const std::vector<int>& diagSlot)
{
    const Topology& t = s->topo;
    const int nd = t.nd;
    const long long nnzu = t.nnzu();
    s->nchunks = 0;
    if (nd == 0 || nnzu == 0)
        return PK_OK;
    // off-diagonal contributions before each upper slot
    std::vector<int> ostart(static_cast<size_t>(nnzu) + 1, 0);
    for (int r = 0; r < nd; ++r)
        for (int q = t.ustart[r]; q < t.ustart[r + 1]; ++q)
            ostart[q + 1] = ostart[q] + (q == diagSlot[r] ? 0 : t.cstart[q + 1] - t.cstart[q]);
    std::vector<int> chunkRow(1, 0);
    int cp = 0, co = 0, cr = 0, cs = 0;
    for (int r = 0; r < nd; ++r)
    {
        const int np = t.nstart[r + 1] - t.nstart[r];
        const int no = ostart[t.ustart[r + 1]] - ostart[t.ustart[r]];
        const int ns = t.ustart[r + 1] - t.ustart[r];
        if (np > kPairChunk || no > kOffChunk || ns > kSlotChunk)
            return PK_OK;  // a row too large: gather by rows
        if (cp + np > kPairChunk || co + no > kOffChunk || cr + 1 > kRowChunk || cs + ns > kSlotChunk)
        {
            chunkRow.push_back(r);
            cp = co = cr = cs = 0;
        }
        cp += np;
        co += no;
        ++cr;
        cs += ns;
    }
    chunkRow.push_back(nd);
    std::vector<int> pairOff(std::max<size_t>(1, t.nlist.size()), 0x3fffffff);
    std::vector<int> slotRun(static_cast<size_t>(nnzu), 0);
    for (size_t c = 0; c + 1 < chunkRow.size(); ++c)
    {
        const int r0 = chunkRow[c], r1 = chunkRow[c + 1];
        const int oBase = ostart[t.ustart[r0]], pBase = t.nstart[r0];
        for (int r = r0; r < r1; ++r)
        {
            const int* nb = t.nlist.data() + t.nstart[r];
            const int* ne = t.nlist.data() + t.nstart[r + 1];
            for (int q = t.ustart[r]; q < t.ustart[r + 1]; ++q)
            {
                if (q == diagSlot[r])
                {
                    slotRun[q] = pack_run(t.nstart[r] - pBase, t.nstart[r + 1] - t.nstart[r], r - r0, true);
                    continue;
                }
                slotRun[q] = pack_run(ostart[q] - oBase, ostart[q + 1] - ostart[q], r - r0, false);
                const int cc = t.ucol[q];
                for (int k = t.cstart[q]; k < t.cstart[q + 1]; ++k)
                {
                    const int tet = t.clist[k] >> 5;
                    // the row's pair of this tet (the row list is in tet order)
                    const int* it = std::lower_bound(nb, ne, tet << 2);
                    if (it == ne || (*it >> 2) != tet)
                    {
                        set_error("pk_system_create: the contribution lists disagree with the row lists");
                        return PK_ERR_ARG;
                    }
                    const int a = *it & 3;
                    int b = -1;
                    for (int x = 0; x < 4; ++x)
                        if (t.dyn[t.tets[4 * static_cast<size_t>(tet) + x]] == cc)
                            b = x;
                    if (b < 0 || b == a)
                    {
                        set_error("pk_system_create: a contribution without its corner");
                        return PK_ERR_ARG;
                    }
                    const int j = b < a ? b : b - 1;
                    const int o = ostart[q] + (k - t.cstart[q]) - oBase;
                    int& w = pairOff[static_cast<size_t>(it - t.nlist.data())];
                    w = (w & ~(0x3ff << (10 * j))) | (o << (10 * j));
                }
            }
        }
    }
    // each chunk: first row, first pair, first upper slot, rows | slots << 8 | pairs << 16
    std::vector<int4> chunks(chunkRow.size() - 1);
    for (size_t c = 0; c + 1 < chunkRow.size(); ++c)
    {
        const int r0 = chunkRow[c], r1 = chunkRow[c + 1];
        const int np = t.nstart[r1] - t.nstart[r0], ns = t.ustart[r1] - t.ustart[r0];
        chunks[c] = make_int4(r0, t.nstart[r0], t.ustart[r0], (r1 - r0) | (ns << 8) | (np << 16));
    }
    int rc;
    if ((rc = s->slotRun.upload(slotRun, s->stream)) != PK_OK) return rc;
    if ((rc = s->pairOff.upload(pairOff, s->stream)) != PK_OK) return rc;
    if ((rc = s->gatherChunks.upload(chunks, s->stream)) != PK_OK) return rc;
    s->nchunks = static_cast<int>(chunks.size());
    return PK_OK;
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
    std::vector<unsigned char> tact(t.tetActive.begin(), t.tetActive.end());
    if ((rc = s->tact.upload(tact, st)) != PK_OK) return rc;
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
    s->groupBlocks.assign(static_cast<size_t>(t.G), 0);
    s->groupUpper.assign(static_cast<size_t>(t.G), 0);
    s->groupRows.assign(static_cast<size_t>(t.G), 0);
    for (int g = 0; g < t.G; ++g)
    {
        for (int r = t.groupRowBegin[g]; r < t.groupRowEnd[g]; ++r)
            rowGroup[r] = g;
        s->groupRows[g] = t.groupRowEnd[g] - t.groupRowBegin[g];
        if (s->groupRows[g] > 0)
        {
            s->groupBlocks[g] = t.fstart[t.groupRowEnd[g]] - t.fstart[t.groupRowBegin[g]];
            s->groupUpper[g] = t.ustart[t.groupRowEnd[g]] - t.ustart[t.groupRowBegin[g]];
        }
    }
    if ((rc = s->slotRow.upload(slotRow, st)) != PK_OK) return rc;
    // the diagonal slot of each row: the row's first upper slot (its column
    // is the row itself, the smallest column >= the row)
    std::vector<int> diagSlot(static_cast<size_t>(t.nd));
    for (int r = 0; r < t.nd; ++r)
    {
        if (t.ustart[r] >= t.ustart[r + 1] || t.ucol[t.ustart[r]] != r)
        {
            set_error("pk_system_create: a dynamic row has no diagonal block in its pattern");
            return PK_ERR_ARG;
        }
        diagSlot[r] = t.ustart[r];
    }
    if ((rc = s->diagSlot.upload(diagSlot, st)) != PK_OK) return rc;
    if ((rc = build_pair_gather(s, diagSlot)) != PK_OK) return rc;
    if (!is_upper(s))
    {
        std::vector<int> fdiag(static_cast<size_t>(t.nd));
        for (int r = 0; r < t.nd; ++r)
            fdiag[r] = t.uFull[diagSlot[r]];
        if ((rc = s->fdiag.upload(fdiag, st)) != PK_OK) return rc;
    }
    if ((rc = s->rowGroup.upload(rowGroup, st)) != PK_OK) return rc;
    if ((rc = s->fstart.upload(t.fstart, st)) != PK_OK) return rc;
    if ((rc = s->fcol.upload(t.fcol, st)) != PK_OK) return rc;
    if ((rc = s->nstart.upload(t.nstart, st)) != PK_OK) return rc;
    if ((rc = s->nlist.upload(t.nlist, st)) != PK_OK) return rc;
    if ((rc = s->rowNode.upload(t.rowNode, st)) != PK_OK) return rc;
    if ((rc = s->astart.upload(t.astart, st)) != PK_OK) return rc;
    if ((rc = s->alist.upload(t.alist, st)) != PK_OK) return rc;
    if ((rc = s->apos.upload(t.apos, st)) != PK_OK) return rc;
    if ((rc = s->nodeRow.upload(t.dyn, st)) != PK_OK) return rc;
    if ((rc = s->nodeGroup.upload(t.nodeGroup, st)) != PK_OK) return rc;
    if ((rc = s->cpos.upload(t.cpos, st)) != PK_OK) return rc;
    if (is_upper(s))
    {
        if ((rc = build_slices(s, slotRow)) != PK_OK) return rc;
        if ((rc = s->ucol.upload(t.ucol, st)) != PK_OK) return rc;
    }
    if ((rc = alloc_work(s)) != PK_OK) return rc;
    launch_gdot(s->view(), s->prm.p);
    if ((rc = s->kdat.alloc(static_cast<size_t>(t.T) * KD_COUNT)) != PK_OK) return rc;
    launch_stiff_records(s->view());
    if (is_f32(s) && !is_upper(s))
        launch_record_columns(s->view(), true);
    std::vector<double> zero(static_cast<size_t>(t.N) * 3, 0.0);
    if ((rc = s->v.upload(zero, st)) != PK_OK) return rc;
    if ((rc = s->x.alloc(static_cast<size_t>(t.N) * 3)) != PK_OK) return rc;
    s->lanesPerRow = is_upper(s) ? 1 : choose_lanes(t);
    if (is_upper(s))
        s->spmvRows = 8 * kSliceRows;  // two slices per wave
    if ((rc = setup_tiles(s)) != PK_OK) return rc;
    return upload_groups(s);
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
    if (!desc || !out || precision < PK_FP64 || precision > PK_FP64_UPPER)
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
    PK_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    {
        int dev = 0, khz = 0;
        PK_HIP(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess)
            s->wallClockKHz = khz;
    }
    PK_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->pinned), 8 * sizeof(int), hipHostMallocDefault));
    PK_HIP(hipEventCreateWithFlags(&s->pollEv[0], hipEventDisableTiming));
    PK_HIP(hipEventCreateWithFlags(&s->pollEv[1], hipEventDisableTiming));
    if ((rc = setup_device(s.get())) != PK_OK)
        return rc;
    if (desc->node_count > 0)
        PK_HIP(hipMemcpyAsync(s->x.p, desc->rest, 3 * static_cast<size_t>(desc->node_count) * sizeof(double),
                              hipMemcpyHostToDevice, s->stream));
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

int pk_system_tune(pk_system* s, int lanes_per_row, int spmv_rows, int vector_rows, int unroll)
{
    const bool lanesOk = lanes_per_row == 1 || lanes_per_row == 4 || lanes_per_row == 8 || lanes_per_row == 16 ||
                         lanes_per_row == 32;
    const bool unrollOk = unroll == 1 || unroll == 2 || (unroll == 4 && lanes_per_row <= 16);
    if (!s || !lanesOk || !unrollOk || spmv_rows < 1 || vector_rows < 1)
    {
        set_error("pk_system_tune: lanes_per_row in {1, 4, 8, 16, 32}, unroll in {1, 2, 4} (4 only up to 16 lanes), "
                  "positive row counts");
        return PK_ERR_ARG;
    }
    if (is_upper(s) && (lanes_per_row != 1 || spmv_rows % kSliceRows != 0))
    {
        set_error("pk_system_tune: upper-only storage runs one lane per row (lanes_per_row 1) over whole slices "
                  "(spmv_rows a multiple of 64)");
        return PK_ERR_ARG;
    }
    s->tapeValid = false;  // a recorded run no longer replays bit-identically
    ++s->cfgVer;  // a captured substep bakes the old setting in
    s->lanesPerRow = lanes_per_row;
    s->spmvRows = spmv_rows;
    s->vecRows = vector_rows;
    s->unroll = unroll;
    return setup_tiles(s);
}

int pk_system_set_spmv_staging(pk_system* s, int mode)
{
    if (!s || mode < 0 || mode > 1)
    {
        set_error("pk_system_set_spmv_staging: mode 0 (off) or 1 (staging)");
        return PK_ERR_ARG;
    }
    s->tapeValid = false;  // a recorded run no longer replays the same launches
    ++s->cfgVer;  // a captured substep bakes the old setting in
    s->stagingMode = mode;
    ++s->epoch;
    const int rc = setup_windows(s);
    if (rc != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_get_spmv_staging(const pk_system* s, int* mode, int* rows, int* segmented)
{
    if (!s)
        return PK_ERR_ARG;
    if (mode) *mode = s->stagingMode;
    if (rows) *rows = s->winCap;
    if (segmented) *segmented = s->stageSegs;
    return PK_OK;
}

int pk_system_get_spmv_stage_capacity(const pk_system* s, int* rows)
{
    if (!s || !rows)
        return PK_ERR_ARG;
    *rows = s->stageCap;
    return PK_OK;
}

int pk_system_get_slice_layout(const pk_system* s, int* slices, int* compact_slices, long long* elements)
{
    if (!s)
        return PK_ERR_ARG;
    const bool up = is_upper(s);
    if (slices) *slices = up ? s->nSlices : 0;
    if (compact_slices) *compact_slices = up ? s->nCompact : 0;
    if (elements) *elements = up ? s->nE : 0;
    return PK_OK;
}

int pk_system_get_tune(const pk_system* s, int* lanes_per_row, int* spmv_rows, int* vector_rows, int* unroll)
{
    if (!s)
        return PK_ERR_ARG;
    if (lanes_per_row) *lanes_per_row = s->lanesPerRow;
    if (spmv_rows) *spmv_rows = s->spmvRows;
    if (vector_rows) *vector_rows = s->vecRows;
    if (unroll) *unroll = s->unroll;
    return PK_OK;
}

}  // extern "C"

namespace pk
{
// x and v in or out, host or device pointers (kind H2D/D2H or D2D).
static int copy_state_in(pk_system* s, const double* x, const double* v, hipMemcpyKind kind)
{
    ++s->stateVer;
    const size_t bytes = static_cast<size_t>(s->topo.N) * 3 * sizeof(double);
    if (x && bytes)
        PK_HIP(hipMemcpyAsync(s->x.p, x, bytes, kind, s->stream));
    if (v && bytes)
        PK_HIP(hipMemcpyAsync(s->v.p, v, bytes, kind, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

static int copy_state_out(pk_system* s, double* x, double* v, hipMemcpyKind kind)
{
    const size_t bytes = static_cast<size_t>(s->topo.N) * 3 * sizeof(double);
    if (x && bytes)
        PK_HIP(hipMemcpyAsync(x, s->x.p, bytes, kind, s->stream));
    if (v && bytes)
        PK_HIP(hipMemcpyAsync(v, s->v.p, bytes, kind, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}
}  // namespace pk

extern "C" {

int pk_system_set_state(pk_system* s, const double* x, const double* v)
{
    return s ? copy_state_in(s, x, v, hipMemcpyHostToDevice) : PK_ERR_ARG;
}

int pk_system_get_state(pk_system* s, double* x, double* v)
{
    return s ? copy_state_out(s, x, v, hipMemcpyDeviceToHost) : PK_ERR_ARG;
}

int pk_system_set_state_device(pk_system* s, const double* x, const double* v)
{
    return s ? copy_state_in(s, x, v, hipMemcpyDeviceToDevice) : PK_ERR_ARG;
}

int pk_system_get_state_device(pk_system* s, double* x, double* v)
{
    return s ? copy_state_out(s, x, v, hipMemcpyDeviceToDevice) : PK_ERR_ARG;
}

int pk_system_tape_id(const pk_system* s, long long* id)
{
    if (!s || !id)
        return PK_ERR_ARG;
    *id = s->tapeValid ? s->tapeId : 0;
    return PK_OK;
}

int pk_system_export_begin(pk_system* s, int what)
{
    if (!s || what < 1 || what > (PK_EXPORT_POSITIONS | PK_EXPORT_VELOCITIES))
    {
        set_error("pk_system_export_begin: what is PK_EXPORT_POSITIONS, PK_EXPORT_VELOCITIES or both");
        return PK_ERR_ARG;
    }
    const size_t n3 = static_cast<size_t>(s->topo.N) * 3;
    s->frameVer = ~0ull;
    s->frameWhat = 0;
    if (n3 == 0)
        return PK_OK;
    if (!s->framePinned)
        PK_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->framePinned), 2 * n3 * sizeof(float),
                             hipHostMallocDefault));
    if (!s->frameEv)
        PK_HIP(hipEventCreateWithFlags(&s->frameEv, hipEventDisableTiming));
    int rc;
    if ((rc = s->frameDev.ensure(2 * n3)) != PK_OK)
        return rc;
    const bool px = (what & PK_EXPORT_POSITIONS) != 0, pv = (what & PK_EXPORT_VELOCITIES) != 0;
    pk::launch_export_frame(s->stream, static_cast<long long>(n3), px ? s->x.p : nullptr, pv ? s->v.p : nullptr,
                            s->frameDev.p);
    PK_HIP(hipGetLastError());
    if (px && pv)
        PK_HIP(hipMemcpyAsync(s->framePinned, s->frameDev.p, 2 * n3 * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream));
    else
    {
        const size_t off = px ? 0 : n3;
        PK_HIP(hipMemcpyAsync(s->framePinned + off, s->frameDev.p + off, n3 * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream));
    }
    PK_HIP(hipEventRecord(s->frameEv, s->stream));
    s->frameWhat = what;
    s->frameVer = s->stateVer;
    return PK_OK;
}

int pk_system_export_view(pk_system* s, const float** positions, const float** velocities)
{
    if (!s)
        return PK_ERR_ARG;
    if (positions)
        *positions = nullptr;
    if (velocities)
        *velocities = nullptr;
    const bool wantX = positions != nullptr, wantV = velocities != nullptr;
    if (s->frameVer != s->stateVer || (wantX && !(s->frameWhat & PK_EXPORT_POSITIONS)) ||
        (wantV && !(s->frameWhat & PK_EXPORT_VELOCITIES)))
    {
        set_error("pk_system_export_view: no export of that kind since the state last changed");
        return PK_ERR_STATE;
    }
    const size_t n3 = static_cast<size_t>(s->topo.N) * 3;
    if (n3 > 0)
        PK_HIP(hipEventSynchronize(s->frameEv));
    if (wantX)
        *positions = s->framePinned;
    if (wantV)
        *velocities = n3 > 0 ? s->framePinned + n3 : s->framePinned;
    return PK_OK;
}

int pk_system_export_read(pk_system* s, int first, int count, float* positions, float* velocities)
{
    if (!s || first < 0 || count < 0 || static_cast<long long>(first) + count > s->topo.N)
    {
        set_error("pk_system_export_read: node range outside the system");
        return PK_ERR_ARG;
    }
    const float *px = nullptr, *pv = nullptr;
    const int rc = pk_system_export_view(s, positions ? &px : nullptr, velocities ? &pv : nullptr);
    if (rc != PK_OK)
        return rc;
    const size_t off = 3 * static_cast<size_t>(first), len = 3 * static_cast<size_t>(count) * sizeof(float);
    if (positions && len)
        std::memcpy(positions, px + off, len);
    if (velocities && len)
        std::memcpy(velocities, pv + off, len);
    return PK_OK;
}

int pk_system_set_gravity(pk_system* s, double gx, double gy, double gz)
{
    if (!s)
        return PK_ERR_ARG;
    s->tapeValid = false;  // a recorded run no longer replays the same map
    ++s->cfgVer;  // a captured substep bakes the old setting in
    s->gravity[0] = gx;
    s->gravity[1] = gy;
    s->gravity[2] = gz;
    return PK_OK;
}

int pk_system_set_solver(pk_system* s, double tolerance, int max_iterations, int use_jacobi)
{
    if (!s)
        return PK_ERR_ARG;
    s->tapeValid = false;  // a recorded run no longer replays the same map
    ++s->cfgVer;  // a captured substep bakes the old setting in
    s->tol = tolerance;
    s->maxIt = max_iterations;
    s->jacobi = use_jacobi ? 1 : 0;
    return upload_groups(s);
}

int pk_system_set_mode(pk_system* s, int implicit)
{
    if (!s)
        return PK_ERR_ARG;
    s->tapeValid = false;  // a recorded run no longer replays the same map
    ++s->cfgVer;  // a captured substep bakes the old setting in
    s->implicit = implicit ? 1 : 0;
    return PK_OK;
}

int pk_system_set_fem_models(pk_system* s, const unsigned char* models)
{
    if (!s || (!models && s->topo.T > 0))
        return PK_ERR_ARG;
    const size_t T = static_cast<size_t>(s->topo.T);
    bool allCorot = true;
    for (size_t t = 0; t < T; ++t)
    {
        if (models[t] > FEM_NONE)
        {
            set_error("pk_system_set_fem_models: a model is not 0 (linear), 1 (corotational) or 2 (none)");
            return PK_ERR_ARG;
        }
        allCorot = allCorot && models[t] == FEM_COROTATIONAL;
    }
    s->tapeValid = false;  // a recorded run no longer replays the same map
    ++s->cfgVer;  // a captured substep bakes the old setting in
    ++s->epoch;            // captured graphs hold the element launch's arguments
    if (allCorot)
    {
        s->fem.free();
        return PK_OK;
    }
    int rc;
    if ((rc = s->fem.ensure(T)) != PK_OK)
        return rc;
    PK_HIP(hipMemcpyAsync(s->fem.p, models, T, hipMemcpyHostToDevice, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_set_young(pk_system* s, const double* young)
{
    if (!s || !young)
        return PK_ERR_ARG;
    s->tapeValid = false;  // a recorded run no longer replays the same map
    ++s->cfgVer;  // a captured substep bakes the old setting in
    Topology& t = s->topo;
    if (t.T == 0)
        return PK_OK;
    const size_t T = static_cast<size_t>(t.T);
    for (size_t i = 0; i < T; ++i)
        lame(young[i], t.poisson[i], &t.prm[P_LAMBDA * T + i], &t.prm[P_MU * T + i]);
    // lambda and mu are adjacent planes of the parameter block
    PK_HIP(hipMemcpyAsync(s->prm.p + P_LAMBDA * T, t.prm.data() + P_LAMBDA * T, 2 * T * sizeof(double),
                          hipMemcpyHostToDevice, s->stream));
    launch_stiff_records(s->view());
    PK_HIP(hipGetLastError());
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
    ++s->stateVer;
    const double h = dt / substeps;
    for (int f = 0; f < frames; ++f)
        for (int k = 0; k < substeps; ++k)
        {
            const int rc = is_f32(s) ? run_substep<float>(s, h) : run_substep<double>(s, h);
            if (rc != PK_OK)
                return rc;
            if (s->implicit && s->topo.nd > 0)
            {
                // the outcome sizes the next solve's first chunk
                const int rc2 = read_groups(s, s->lastIters, s->firstStop);
                if (rc2 != PK_OK)
                    return rc2;
                harvest_profile(s);
            }
        }
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

}  // extern "C"

namespace pk
{

// ---- reverse mode ----

// Device copy of n doubles on the system's stream.
static int dcopy(pk_system* s, double* dst, const double* src, size_t n)
{
    if (n)
        PK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    return PK_OK;
}

static int icopy(pk_system* s, int* dst, const int* src, size_t n)
{
    if (n)
        PK_HIP(hipMemcpyAsync(dst, src, n * sizeof(int), hipMemcpyDeviceToDevice, s->stream));
    return PK_OK;
}

// One forward substep of a replay: the plain substep, then the per-body
// integrate flags into `ok`.
template <class M>
static int replay_substep(pk_system* s, double dt, int* ok, int* failedOut = nullptr)
{
    int rc = run_substep<M>(s, dt, failedOut);
    if (rc != PK_OK)
        return rc;
    launch_record_ok(s->view(), ok);
    PK_HIP(hipGetLastError());
    if ((rc = read_groups(s, s->lastIters, s->firstStop)) != PK_OK)
        return rc;
    harvest_profile(s);
    return PK_OK;
}

// Reverses substep j of the current segment: (segX[j], segV[j]) -> segV[j+1].
template <class M>
static int reverse_substep(pk_system* s, double dt, int j, double adjTol, int adjMaxIt, int* iters)
{
    const size_t n3 = static_cast<size_t>(s->topo.N) * 3;
    const int* ok = s->segOk.p + static_cast<size_t>(j) * s->topo.G;
    int rc;
    if ((rc = dcopy(s, s->x.p, s->segX.p + j * n3, n3)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->v.p, s->segV.p + j * n3, n3)) != PK_OK) return rc;
    if (s->valsView)
    {
        // A(x_j) and its rejection flags as the replay left them
        const DeviceView sv = s->view();
        launch_reset(sv, adjTol, adjMaxIt);
        launch_restore_failed(sv, s->segFailed.p + static_cast<size_t>(j) * s->topo.G);
    }
    else if ((rc = run_assembly<M>(s, dt, adjTol, adjMaxIt, kAsmMatrix)) != PK_OK)
        return rc;
    const DeviceView dv = s->view();
    launch_adj_rhs<M>(dv, ok, s->xbar.p, s->vbar.p, dt);
    if ((rc = run_cg<M>(s, adjMaxIt, s->lastItersAdj, s->firstStopAdj)) != PK_OK)
        return rc;
    launch_adj_element(dv, ok, s->segV.p + (j + 1) * n3, dt, s->gradE.p);
    launch_adj_contacts(dv, dt);
    launch_adj_nodes(dv, ok, dt, s->xbar.p, s->vbar.p);
    PK_HIP(hipGetLastError());
    if ((rc = read_groups(s, s->lastItersAdj, s->firstStopAdj)) != PK_OK)
        return rc;
    harvest_profile(s);
    if (iters)
        *iters = s->lastItersAdj;
    return PK_OK;
}

template <class M>
static int record_run(pk_system* s, double dt, int steps, int every)
{
    const size_t n3 = static_cast<size_t>(s->topo.N) * 3;
    const size_t G = static_cast<size_t>(std::max(1, s->topo.G));
    const int nseg = (steps + every - 1) / every;
    int rc;
    s->tapeValid = false;
    // checkpoint c: the state before substep c * every (c < nseg), and the
    // end state (c = nseg)
    if ((rc = s->tapeX.ensure((nseg + 1) * n3)) != PK_OK) return rc;
    if ((rc = s->tapeV.ensure((nseg + 1) * n3)) != PK_OK) return rc;
    if ((rc = s->tapeOk.ensure(nseg * G)) != PK_OK) return rc;
    if ((rc = s->segOk.ensure(static_cast<size_t>(every) * G)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->tapeX.p, s->x.p, n3)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->tapeV.p, s->v.p, n3)) != PK_OK) return rc;
    for (int n = 0; n < steps; ++n)
    {
        if ((rc = replay_substep<M>(s, dt, s->segOk.p)) != PK_OK)
            return rc;
        if ((n + 1) % every == 0 || n + 1 == steps)
        {
            // a segment's last substep: its integrate flags, and the state
            // after it (the next segment's start or the end state)
            const size_t c = static_cast<size_t>(n / every);
            if ((rc = icopy(s, s->tapeOk.p + c * G, s->segOk.p, G)) != PK_OK) return rc;
            if ((rc = dcopy(s, s->tapeX.p + (c + 1) * n3, s->x.p, n3)) != PK_OK) return rc;
            if ((rc = dcopy(s, s->tapeV.p + (c + 1) * n3, s->v.p, n3)) != PK_OK) return rc;
        }
    }
    PK_HIP(hipStreamSynchronize(s->stream));
    s->tapeDt = dt;
    s->tapeSteps = steps;
    s->tapeEvery = every;
    s->tapeValid = true;
    ++s->tapeId;
    return PK_OK;
}

template <class M>
static int backward_run(pk_system* s, double adjTol, int adjMaxIt, int* adjIters)
{
    const Topology& t = s->topo;
    const size_t n3 = static_cast<size_t>(t.N) * 3;
    const int steps = s->tapeSteps, every = s->tapeEvery;
    const double dt = s->tapeDt;
    int rc;
    if ((rc = s->segX.alloc((every + 1) * n3)) != PK_OK) return rc;
    if ((rc = s->segV.alloc((every + 1) * n3)) != PK_OK) return rc;
    if ((rc = s->xEnd.alloc(n3)) != PK_OK) return rc;
    if ((rc = s->vEnd.alloc(n3)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->xEnd.p, s->x.p, n3)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->vEnd.p, s->v.p, n3)) != PK_OK) return rc;
    // keep the segment's matrices from the replay when they fit
    const size_t valsBytes = s->vals.n, minvBytes = s->minv.n;
    bool store = t.nd > 0 && s->storeCap > 0 &&
                 static_cast<double>(every) * static_cast<double>(valsBytes + minvBytes) <=
                     static_cast<double>(s->storeCap);
    if (store)
    {
        if (!(s->segVals.p && s->segVals.n == every * valsBytes))
            s->segColumns = -1;  // new memory: its record columns are not written yet
        if (s->segVals.ensure(every * valsBytes) != PK_OK || s->segMinv.ensure(every * minvBytes) != PK_OK ||
            s->segFailed.ensure(static_cast<size_t>(every) * std::max(1, t.G)) != PK_OK)
        {
            // no room: assemble again in every reverse step instead
            (void)hipGetLastError();
            s->segVals.free();
            s->segMinv.free();
            s->segFailed.free();
            store = false;
        }
    }
    if (store)
    {
        if (is_f32(s) && !is_upper(s) && s->segColumns != s->epoch)
        {
            for (int j = 0; j < every; ++j)
            {
                s->valsView = s->segVals.p + j * valsBytes;
                launch_record_columns(s->view(), is_f32(s));
            }
            s->valsView = nullptr;
            s->segColumns = s->epoch;
        }
    }
    auto slot = [&](int j) {
        s->valsView = store ? s->segVals.p + j * valsBytes : nullptr;
        s->minvView = store ? s->segMinv.p + j * minvBytes : nullptr;
    };
    const int nseg = (steps + every - 1) / every;
    for (int seg = nseg - 1; seg >= 0; --seg)
    {
        const int n0 = seg * every, len = std::min(every, steps - n0);
        if ((rc = dcopy(s, s->x.p, s->tapeX.p + seg * n3, n3)) != PK_OK) return rc;
        if ((rc = dcopy(s, s->v.p, s->tapeV.p + seg * n3, n3)) != PK_OK) return rc;
        if ((rc = dcopy(s, s->segX.p, s->x.p, n3)) != PK_OK) return rc;
        if ((rc = dcopy(s, s->segV.p, s->v.p, n3)) != PK_OK) return rc;
        for (int j = 0; j + 1 < len; ++j)
        {
            slot(j);
            rc = replay_substep<M>(s, dt, s->segOk.p + static_cast<size_t>(j) * t.G,
                                   store ? s->segFailed.p + static_cast<size_t>(j) * t.G : nullptr);
            s->valsView = s->minvView = nullptr;
            if (rc != PK_OK)
                return rc;
            if ((rc = dcopy(s, s->segX.p + (j + 1) * n3, s->x.p, n3)) != PK_OK) return rc;
            if ((rc = dcopy(s, s->segV.p + (j + 1) * n3, s->v.p, n3)) != PK_OK) return rc;
        }
        {
            // the segment's last substep is not solved again: the record run
            // kept the state after it (the next checkpoint) and its integrate
            // flags; only its matrix is needed, assembled here at its start
            // state when the segment's matrices are kept (otherwise the
            // reverse step assembles it)
            const int j = len - 1;
            const size_t G = static_cast<size_t>(std::max(1, t.G));
            if ((rc = icopy(s, s->segOk.p + static_cast<size_t>(j) * G, s->tapeOk.p + seg * G, G)) != PK_OK)
                return rc;
            if ((rc = dcopy(s, s->segX.p + (j + 1) * n3, s->tapeX.p + (seg + 1) * n3, n3)) != PK_OK) return rc;
            if ((rc = dcopy(s, s->segV.p + (j + 1) * n3, s->tapeV.p + (seg + 1) * n3, n3)) != PK_OK) return rc;
            if (store)
            {
                if ((rc = dcopy(s, s->x.p, s->segX.p + j * n3, n3)) != PK_OK) return rc;
                if ((rc = dcopy(s, s->v.p, s->segV.p + j * n3, n3)) != PK_OK) return rc;
                slot(j);
                rc = run_assembly<M>(s, dt, s->tol, s->maxIt, kAsmMatrix);
                if (rc == PK_OK && t.nd > 0)
                    launch_save_failed(s->view(), s->segFailed.p + static_cast<size_t>(j) * t.G);
                s->valsView = s->minvView = nullptr;
                if (rc != PK_OK)
                    return rc;
            }
        }
        for (int j = len - 1; j >= 0; --j)
        {
            slot(j);
            rc = reverse_substep<M>(s, dt, j, adjTol, adjMaxIt, adjIters ? adjIters + n0 + j : nullptr);
            s->valsView = s->minvView = nullptr;
            if (rc != PK_OK)
                return rc;
        }
    }
    if ((rc = dcopy(s, s->x.p, s->xEnd.p, n3)) != PK_OK) return rc;
    if ((rc = dcopy(s, s->v.p, s->vEnd.p, n3)) != PK_OK) return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

}  // namespace pk

extern "C" {

int pk_system_record(pk_system* s, double dt, int steps, int every)
{
    if (!s || !(dt > 0.0) || steps < 1 || every < 1)
    {
        set_error("pk_system_record: need dt > 0, steps >= 1, every >= 1");
        return PK_ERR_ARG;
    }
    if (!s->implicit)
    {
        set_error("pk_system_record: the reverse mode covers the implicit integrator only");
        return PK_ERR_ARG;
    }
    if (s->fem.p)
    {
        set_error("pk_system_record: the reverse mode covers the corotational FEM model only");
        return PK_ERR_ARG;
    }
    if (s->topo.massErrorNode >= 0)
    {
        set_error("Non-fixed node has no physical mass contribution. nodeIndex=" +
                  std::to_string(s->topo.massErrorNode));
        return PK_ERR_MASS;
    }
    every = std::min(every, steps);
    if (s->topo.nd == 0)
    {
        s->tapeValid = false;
        set_error("pk_system_record: no dynamic node");
        return PK_ERR_ARG;
    }
    ++s->stateVer;
    return is_f32(s) ? record_run<float>(s, dt, steps, every)
                                           : record_run<double>(s, dt, steps, every);
}

}  // extern "C"

namespace pk
{
static int backward_entry(pk_system* s, const double* xbar, const double* vbar, double adj_tol, int adj_max_iter,
                          double* grad_log_young, double* grad_x0, double* grad_v0, int* adj_iterations, bool device)
{
    if (!s)
        return PK_ERR_ARG;
    if (!s->tapeValid)
    {
        set_error("pk_system_backward: no recorded run (call pk_system_record after the last change)");
        return PK_ERR_ARG;
    }
    const hipMemcpyKind in = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    ++s->stateVer;  // the replays write x and v (the end state is restored, bit for bit)
    const Topology& t = s->topo;
    const size_t n3 = static_cast<size_t>(t.N) * 3;
    int rc;
    if ((rc = s->xbar.ensure(n3)) != PK_OK) return rc;
    if ((rc = s->vbar.ensure(n3)) != PK_OK) return rc;
    if ((rc = s->gradE.ensure(static_cast<size_t>(t.T))) != PK_OK) return rc;
    if (s->adj.n < std::max<size_t>(1, t.alist.size()) * N_ADJ)
    {
        if ((rc = s->adj.alloc(std::max<size_t>(1, t.alist.size()) * N_ADJ)) != PK_OK) return rc;
        // tets switched off in place are still in the node gather lists and
        // the reverse element kernel skips them: their entries must read 0
        PK_HIP(hipMemsetAsync(s->adj.p, 0, s->adj.n * sizeof(double), s->stream));
    }
    if (xbar)
        PK_HIP(hipMemcpyAsync(s->xbar.p, xbar, n3 * sizeof(double), in, s->stream));
    else
        PK_HIP(hipMemsetAsync(s->xbar.p, 0, n3 * sizeof(double), s->stream));
    if (vbar)
        PK_HIP(hipMemcpyAsync(s->vbar.p, vbar, n3 * sizeof(double), in, s->stream));
    else
        PK_HIP(hipMemsetAsync(s->vbar.p, 0, n3 * sizeof(double), s->stream));
    if (t.T)
        PK_HIP(hipMemsetAsync(s->gradE.p, 0, static_cast<size_t>(t.T) * sizeof(double), s->stream));
    rc = is_f32(s) ? backward_run<float>(s, adj_tol, adj_max_iter, adj_iterations)
                   : backward_run<double>(s, adj_tol, adj_max_iter, adj_iterations);
    if (rc != PK_OK)
        return rc;
    if (grad_log_young && t.T)
        PK_HIP(hipMemcpyAsync(grad_log_young, s->gradE.p, static_cast<size_t>(t.T) * sizeof(double), out,
                              s->stream));
    if (grad_x0 && n3)
        PK_HIP(hipMemcpyAsync(grad_x0, s->xbar.p, n3 * sizeof(double), out, s->stream));
    if (grad_v0 && n3)
        PK_HIP(hipMemcpyAsync(grad_v0, s->vbar.p, n3 * sizeof(double), out, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}
}  // namespace pk

extern "C" {

int pk_system_backward(pk_system* s, const double* xbar, const double* vbar, double adj_tol, int adj_max_iter,
                       double* grad_log_young, double* grad_x0, double* grad_v0, int* adj_iterations)
{
    return backward_entry(s, xbar, vbar, adj_tol, adj_max_iter, grad_log_young, grad_x0, grad_v0, adj_iterations,
                          false);
}

int pk_system_backward_device(pk_system* s, const double* xbar, const double* vbar, double adj_tol,
                              int adj_max_iter, double* grad_log_young, double* grad_x0, double* grad_v0,
                              int* adj_iterations)
{
    return backward_entry(s, xbar, vbar, adj_tol, adj_max_iter, grad_log_young, grad_x0, grad_v0, adj_iterations,
                          true);
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
    s->prof.spmvBlocks = s->prof.spmvUpper = s->prof.spmvRows = s->prof.spmvBodyLaunches = 0;
    s->prof.recs.clear();
    s->prof.used = 0;
    return PK_OK;
}

int pk_system_set_connections(pk_system* s, int count, const int* nodes, const double* weights,
                              const double* targets, const double* stiffness, const double* damping)
{
    if (!s || count < 0 || (count > 0 && (!nodes || !weights || !targets || !stiffness || !damping)))
    {
        set_error("pk_system_set_connections: bad argument");
        return PK_ERR_ARG;
    }
    const Topology& t = s->topo;
    // the connections the reference keeps (World.cpp:126-131, 748-760;
    // PointConnection.cpp:58-66, 81-84)
    std::vector<int4> hn;
    std::vector<double> hd;
    for (int i = 0; i < count; ++i)
    {
        bool ok = std::isfinite(stiffness[i]) && std::isfinite(damping[i]) && stiffness[i] >= 0.0 &&
                  damping[i] >= 0.0;
        for (int a = 0; a < 4; ++a)
        {
            const int n = nodes[4 * i + a];
            ok = ok && n >= 0 && n < t.N && t.nodeActive[n];
        }
        for (int j = 0; j < 3; ++j)
            ok = ok && std::isfinite(targets[3 * i + j]);
        if (!ok)
            continue;
        hn.push_back(make_int4(nodes[4 * i], nodes[4 * i + 1], nodes[4 * i + 2], nodes[4 * i + 3]));
        for (int a = 0; a < 4; ++a)
            hd.push_back(weights[4 * i + a]);
        for (int j = 0; j < 3; ++j)
            hd.push_back(targets[3 * i + j]);
        hd.push_back(stiffness[i]);
        hd.push_back(damping[i]);
    }
    const int nc = static_cast<int>(hn.size());
    // the upper slots their k w_a w_b I blocks land in (each dynamic pair
    // once, as the reference skips a lower block whose mirror is present,
    // SolverData.cpp:440-446) and the rows their terms reach
    std::map<int, double> slots;
    std::vector<std::array<int, 3>> ent;  // (row, connection, a)
    auto find_slot = [&](int r, int c) {
        const auto b = t.ucol.begin() + t.ustart[r], e = t.ucol.begin() + t.ustart[r + 1];
        const auto it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? static_cast<int>(it - t.ucol.begin()) : -1;
    };
    for (int ci = 0; ci < nc; ++ci)
    {
        const int n4[4] = {hn[ci].x, hn[ci].y, hn[ci].z, hn[ci].w};
        const double* d = hd.data() + 9 * static_cast<size_t>(ci);
        for (int a = 0; a < 4; ++a)
            if (d[a] != 0.0 && t.dyn[n4[a]] >= 0)
                ent.push_back({t.dyn[n4[a]], ci, a});
        if (d[7] <= 0.0)
            continue;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
            {
                if (d[a] == 0.0 || d[b] == 0.0)
                    continue;
                const int ra = t.dyn[n4[a]], rb = t.dyn[n4[b]];
                if (ra < 0 || rb < 0 || ra > rb)
                    continue;
                const int sl = find_slot(ra, rb);
                if (sl < 0)
                {
                    set_error("pk_system_set_connections: nodes " + std::to_string(n4[a]) + " and " +
                              std::to_string(n4[b]) + " are not coupled in the BSR pattern (add the pair to "
                              "pk_mesh_desc.extra_pairs)");
                    return PK_ERR_ARG;
                }
                slots[sl] += d[7] * d[a] * d[b];
            }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const auto& p, const auto& q) { return p[0] < q[0]; });
    std::vector<int> rows, rstart(1, 0);
    std::vector<int2> entries;
    for (size_t k = 0; k < ent.size(); ++k)
    {
        if (rows.empty() || rows.back() != ent[k][0])
        {
            if (!rows.empty())
                rstart.push_back(static_cast<int>(entries.size()));
            rows.push_back(ent[k][0]);
        }
        entries.push_back(make_int2(ent[k][1], ent[k][2]));
    }
    if (!rows.empty())
        rstart.push_back(static_cast<int>(entries.size()));
    std::vector<int> sidx;
    std::vector<double> sval;
    for (const auto& kv : slots)
    {
        sidx.push_back(kv.first);
        sval.push_back(kv.second);
    }

    hipStream_t st = s->stream;
    int rc;
    // clear what the previous connections set
    if (s->nslots > 0)
        launch_slot_scatter(st, s->nslots, s->slotIdx.p, nullptr, s->cslot.p);
    if (s->nconnRows > 0)
        launch_zero_rows(st, s->nconnRows, s->connRows.p, s->crhs.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(st));
    s->nconn = s->nconnRows = s->nslots = 0;
    if (nc == 0)
        return PK_OK;
    if (!s->cslot.p && t.nnzu() > 0)
    {
        if ((rc = s->cslot.alloc(static_cast<size_t>(t.nnzu()))) != PK_OK) return rc;
        PK_HIP(hipMemsetAsync(s->cslot.p, 0, s->cslot.n * sizeof(double), st));
    }
    if (!s->crhs.p && t.nd > 0)
    {
        if ((rc = s->crhs.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
        PK_HIP(hipMemsetAsync(s->crhs.p, 0, s->crhs.n * sizeof(double), st));
    }
    if ((rc = s->connNodes.upload(hn, st)) != PK_OK) return rc;
    if ((rc = s->connData.upload(hd, st)) != PK_OK) return rc;
    if ((rc = s->connEval.alloc(7 * static_cast<size_t>(nc))) != PK_OK) return rc;
    if ((rc = s->connRows.upload(rows, st)) != PK_OK) return rc;
    if ((rc = s->connRowStart.upload(rstart, st)) != PK_OK) return rc;
    if ((rc = s->connEntries.upload(entries, st)) != PK_OK) return rc;
    {
        // every node's (connection, corner) entries with w_a != 0, fixed nodes
        // included, in connection order
        std::vector<int> nstart(static_cast<size_t>(t.N) + 1, 0);
        for (int ci = 0; ci < nc; ++ci)
        {
            const int n4[4] = {hn[ci].x, hn[ci].y, hn[ci].z, hn[ci].w};
            for (int a = 0; a < 4; ++a)
                if (hd[9 * static_cast<size_t>(ci) + a] != 0.0)
                    ++nstart[n4[a] + 1];
        }
        for (int n = 0; n < t.N; ++n)
            nstart[n + 1] += nstart[n];
        std::vector<int2> nlistv(static_cast<size_t>(std::max(1, nstart[t.N])));
        std::vector<int> fill(nstart.begin(), nstart.end() - 1);
        for (int ci = 0; ci < nc; ++ci)
        {
            const int n4[4] = {hn[ci].x, hn[ci].y, hn[ci].z, hn[ci].w};
            for (int a = 0; a < 4; ++a)
                if (hd[9 * static_cast<size_t>(ci) + a] != 0.0)
                    nlistv[fill[n4[a]]++] = make_int2(ci, a);
        }
        if ((rc = s->connNodeStart.upload(nstart, st)) != PK_OK) return rc;
        if ((rc = s->connNodeList.upload(nlistv, st)) != PK_OK) return rc;
        if ((rc = s->connAdj.alloc(6 * static_cast<size_t>(nc))) != PK_OK) return rc;
    }
    if ((rc = s->slotIdx.upload(sidx, st)) != PK_OK) return rc;
    if ((rc = s->slotVal.upload(sval, st)) != PK_OK) return rc;
    launch_slot_scatter(st, static_cast<int>(sidx.size()), s->slotIdx.p, s->slotVal.p, s->cslot.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(st));
    s->nconn = nc;
    s->nconnRows = static_cast<int>(rows.size());
    s->nslots = static_cast<int>(sidx.size());
    s->tapeValid = false;
    ++s->cfgVer;  // a captured substep bakes the old setting in
    return PK_OK;
}

// The tets collision spheres can touch (four active nodes: the contacts the
// reference keeps, CollisionSphereComponent.cpp:141-162, World.cpp:126-131;
// the contact kernel skips those switched off) and, for each, the upper
// slots of its ordered corner pairs (r_a <= r_b, both dynamic) and the rows
// of its dynamic corners, as per-slot and per-row lists in candidate order.
static int build_contact_lists(pk_system* s)
{
    const Topology& t = s->topo;
    std::vector<int> cand;
    for (int tt = 0; tt < t.T; ++tt)
    {
        bool ok = true;
        for (int a = 0; a < 4; ++a)
        {
            const int n = t.tets[4 * static_cast<size_t>(tt) + a];
            ok = ok && n >= 0 && n < t.N && t.nodeActive[n];
        }
        if (ok)
            cand.push_back(tt);
    }
    const long long nnzu = t.nnzu();
    if (static_cast<long long>(cand.size()) * 16 >= (1LL << 31))
    {
        set_error("pk_system_set_spheres: too many tets for the contact lists");
        return PK_ERR_ARG;
    }
    auto find_slot = [&](int r, int c) {
        const auto b = t.ucol.begin() + t.ustart[r], e = t.ucol.begin() + t.ustart[r + 1];
        const auto it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? static_cast<int>(it - t.ucol.begin()) : -1;
    };
    std::vector<int> sstart(static_cast<size_t>(nnzu) + 1, 0), rstart(static_cast<size_t>(t.nd) + 1, 0);
    std::vector<int> slotOf;  // per (candidate, ordered pair) in fill order
    for (size_t i = 0; i < cand.size(); ++i)
    {
        const int* n4 = t.tets.data() + 4 * static_cast<size_t>(cand[i]);
        for (int a = 0; a < 4; ++a)
        {
            const int ra = t.dyn[n4[a]];
            if (ra < 0)
                continue;
            ++rstart[ra + 1];
            for (int b = 0; b < 4; ++b)
            {
                const int rb = t.dyn[n4[b]];
                if (rb < 0 || ra > rb)
                    continue;
                const int sl = find_slot(ra, rb);
                if (sl < 0)
                {
                    set_error("pk_system_set_spheres: a tet's corner pair is missing from the BSR pattern");
                    return PK_ERR_ARG;
                }
                slotOf.push_back(sl);
                ++sstart[sl + 1];
            }
        }
    }
    for (long long k = 0; k < nnzu; ++k)
        sstart[k + 1] += sstart[k];
    for (int r = 0; r < t.nd; ++r)
        rstart[r + 1] += rstart[r];
    std::vector<int> slist(static_cast<size_t>(sstart[nnzu])), rlist(static_cast<size_t>(rstart[t.nd]));
    {
        std::vector<int> sfill(sstart.begin(), sstart.end() - 1), rfill(rstart.begin(), rstart.end() - 1);
        size_t q = 0;
        for (size_t i = 0; i < cand.size(); ++i)
        {
            const int* n4 = t.tets.data() + 4 * static_cast<size_t>(cand[i]);
            for (int a = 0; a < 4; ++a)
            {
                const int ra = t.dyn[n4[a]];
                if (ra < 0)
                    continue;
                rlist[rfill[ra]++] = static_cast<int>(i);
                for (int b = 0; b < 4; ++b)
                {
                    const int rb = t.dyn[n4[b]];
                    if (rb < 0 || ra > rb)
                        continue;
                    slist[sfill[slotOf[q++]]++] = static_cast<int>(i);
                }
            }
        }
    }
    hipStream_t st = s->stream;
    int rc;
    if ((rc = s->cand.upload(cand, st)) != PK_OK) return rc;
    if ((rc = s->csStart.upload(sstart, st)) != PK_OK) return rc;
    if ((rc = s->csList.upload(slist, st)) != PK_OK) return rc;
    if ((rc = s->crStart.upload(rstart, st)) != PK_OK) return rc;
    if ((rc = s->crList.upload(rlist, st)) != PK_OK) return rc;
    {
        // every node's candidate tets (once per corner) in candidate order
        std::vector<int> nstart(static_cast<size_t>(t.N) + 1, 0);
        for (size_t i = 0; i < cand.size(); ++i)
            for (int a = 0; a < 4; ++a)
                ++nstart[t.tets[4 * static_cast<size_t>(cand[i]) + a] + 1];
        for (int n = 0; n < t.N; ++n)
            nstart[n + 1] += nstart[n];
        std::vector<int> nlistv(static_cast<size_t>(std::max(1, nstart[t.N])));
        std::vector<int> fill(nstart.begin(), nstart.end() - 1);
        for (size_t i = 0; i < cand.size(); ++i)
            for (int a = 0; a < 4; ++a)
                nlistv[fill[t.tets[4 * static_cast<size_t>(cand[i]) + a]]++] = static_cast<int>(i);
        if ((rc = s->candNodeStart.upload(nstart, st)) != PK_OK) return rc;
        if ((rc = s->candNodeList.upload(nlistv, st)) != PK_OK) return rc;
        if ((rc = s->candAdj.alloc(6 * std::max<size_t>(1, cand.size()))) != PK_OK) return rc;
    }
    if ((rc = s->ck.alloc(std::max<size_t>(1, cand.size()))) != PK_OK) return rc;
    if ((rc = s->cg.alloc(3 * std::max<size_t>(1, cand.size()))) != PK_OK) return rc;
    if ((rc = s->sslot.alloc(static_cast<size_t>(std::max<long long>(1, nnzu)))) != PK_OK) return rc;
    if ((rc = s->srhs.alloc(3 * static_cast<size_t>(std::max(1, t.nd)))) != PK_OK) return rc;
    PK_HIP(hipStreamSynchronize(st));
    s->ncand = static_cast<int>(cand.size());
    return PK_OK;
}

int pk_system_set_spheres(pk_system* s, int count, const double* centers, const double* radii,
                          const double* stiffness, const double* damping)
{
    if (!s || count < 0 || (count > 0 && (!centers || !radii || !stiffness || !damping)))
    {
        set_error("pk_system_set_spheres: bad argument");
        return PK_ERR_ARG;
    }
    std::vector<double> hs;
    for (int i = 0; i < count; ++i)
    {
        const double v[6] = {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2], radii[i], stiffness[i],
                             damping[i]};
        for (int j = 0; j < 6; ++j)
            if (!std::isfinite(v[j]) || (j >= 3 && v[j] < 0.0))
            {
                set_error("pk_system_set_spheres: sphere " + std::to_string(i) +
                          " needs a finite centre and finite radius, stiffness and damping >= 0");
                return PK_ERR_ARG;
            }
        hs.insert(hs.end(), v, v + 6);
    }
    int rc;
    if (count > 0 && s->ncand < 0 && (rc = build_contact_lists(s)) != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    s->nsph = 0;
    if (count == 0)
        return PK_OK;
    // kept when the count is unchanged (spheres that move every frame)
    if ((rc = s->sph.ensure(hs.size())) != PK_OK)
        return rc;
    PK_HIP(hipMemcpyAsync(s->sph.p, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    s->nsph = count;
    s->tapeValid = false;
    ++s->cfgVer;  // a captured substep bakes the old setting in
    return PK_OK;
}

int pk_system_set_tets_active(pk_system* s, int count, const int* tets, const unsigned char* active)
{
    if (!s || count < 0 || (count > 0 && (!tets || !active)))
    {
        set_error("pk_system_set_tets_active: bad argument");
        return PK_ERR_ARG;
    }
    Topology& t = s->topo;
    std::map<int, uint8_t> want;  // the last entry for a tet wins
    for (int i = 0; i < count; ++i)
    {
        if (tets[i] < 0 || tets[i] >= t.T)
        {
            set_error("pk_system_set_tets_active: tet " + std::to_string(tets[i]) + " out of range");
            return PK_ERR_ARG;
        }
        want[tets[i]] = active[i] ? 1 : 0;
    }
    const double* vol = t.prm.data() + static_cast<size_t>(P_VOL) * t.T;
    std::vector<int> ids;
    std::vector<unsigned char> liveVal, actVal;
    std::vector<int> off;  // listed tets switched off: their outputs become zeros
    std::set<int> nodes;
    for (const auto& kv : want)
    {
        const int id = kv.first;
        if (t.tetActive[id] == kv.second)
            continue;
        const uint8_t live = (kv.second && vol[id] > 0.0) ? 1 : 0;
        if (live && !t.listed[id])
        {
            set_error("pk_system_set_tets_active: tet " + std::to_string(id) +
                      " was not live when the system was created; create it again");
            return PK_ERR_REBUILD;
        }
        ids.push_back(id);
        liveVal.push_back(live);
        actVal.push_back(kv.second);
        if (!live && t.listed[id])
            off.push_back(id);
        for (int a = 0; a < 4; ++a)
            nodes.insert(t.tets[4 * static_cast<size_t>(id) + a]);
    }
    if (ids.empty())
        return PK_OK;
    // lumped masses of the touched nodes as the creation sums them (tet order
    // over the tets live now, TetMeshPhysicsComponent.cpp:127-152); a node
    // that would join or leave the dynamic set changes the system's rows
    std::vector<uint8_t> liveNow(t.live);
    for (size_t i = 0; i < ids.size(); ++i)
        liveNow[ids[i]] = liveVal[i];
    std::vector<std::pair<int, double>> masses;
    for (const int n : nodes)
    {
        if (n < 0 || n >= t.N || t.nodeFixed[n])
            continue;
        double m = 0.0;
        for (int k = t.astart[n]; k < t.astart[n + 1]; ++k)
        {
            const int tt = t.alist[k] >> 2;
            if (liveNow[tt])
                m += t.tetShare[tt];
        }
        const bool dyn = t.nodeActive[n] && std::isfinite(m) && m > 0.0;
        if (dyn != (t.dyn[n] >= 0))
        {
            set_error("pk_system_set_tets_active: node " + std::to_string(n) +
                      (dyn ? " joins" : " leaves") + " the dynamic set; create the system again");
            return PK_ERR_REBUILD;
        }
        masses.push_back({n, m});
    }
    std::vector<int> rowIdx;
    std::vector<double> rowMass;
    for (const auto& nm : masses)
    {
        t.mass[nm.first] = nm.second;
        if (t.dyn[nm.first] >= 0)  // a node outside the dynamic set has no row
        {
            rowIdx.push_back(t.dyn[nm.first]);
            rowMass.push_back(nm.second);
        }
    }
    for (size_t i = 0; i < ids.size(); ++i)
    {
        t.live[ids[i]] = liveVal[i];
        t.tetActive[ids[i]] = actVal[i];
    }
    hipStream_t st = s->stream;
    int rc;
    pk::DevBuf<int> dIds, dOff, dRows;
    pk::DevBuf<unsigned char> dLive, dAct;
    pk::DevBuf<double> dMass;
    if ((rc = dIds.upload(ids, st)) != PK_OK) return rc;
    if ((rc = dLive.upload(liveVal, st)) != PK_OK) return rc;
    if ((rc = dAct.upload(actVal, st)) != PK_OK) return rc;
    if ((rc = dOff.upload(off, st)) != PK_OK) return rc;
    if ((rc = dRows.upload(rowIdx, st)) != PK_OK) return rc;
    if ((rc = dMass.upload(rowMass, st)) != PK_OK) return rc;
    launch_tet_flags(st, static_cast<int>(ids.size()), dIds.p, dLive.p, dAct.p, s->live.p, s->tact.p);
    launch_slot_scatter(st, static_cast<int>(rowIdx.size()), dRows.p, dMass.p, s->mrow.p);
    launch_zero_tets(s->view(), static_cast<int>(off.size()), dOff.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(st));
    s->tapeValid = false;
    ++s->cfgVer;  // a captured substep bakes the old setting in
    return PK_OK;
}

int pk_system_set_matrix_store(pk_system* s, long long max_bytes)
{
    if (!s || max_bytes < 0)
        return PK_ERR_ARG;
    s->storeCap = max_bytes;
    if (max_bytes == 0)
    {
        s->segVals.free();
        s->segMinv.free();
        s->segFailed.free();
        s->segColumns = -1;
    }
    return PK_OK;
}

int pk_system_set_graphs(pk_system* s, int mode)
{
    if (!s || mode < 0 || mode > 2)
    {
        set_error("pk_system_set_graphs: mode 0 (launches), 1 (CG chunks) or 2 (whole substeps)");
        return PK_ERR_ARG;
    }
    s->graphMode = mode;
    s->useGraphs = mode != 0;
    s->drop_graphs();
    return PK_OK;
}

int pk_system_spmv_work(pk_system* s, long long* blocks, long long* upper_blocks, long long* rows,
                        long long* body_launches)
{
    if (!s)
        return PK_ERR_ARG;
    if (blocks) *blocks = s->prof.spmvBlocks;
    if (upper_blocks) *upper_blocks = s->prof.spmvUpper;
    if (rows) *rows = s->prof.spmvRows;
    if (body_launches) *body_launches = s->prof.spmvBodyLaunches;
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
    const int rc = is_f32(s) ? run_assembly<float>(s, dt, s->tol, s->maxIt, kAsmImplicit)
                             : run_assembly<double>(s, dt, s->tol, s->maxIt, kAsmImplicit);
    if (rc != PK_OK)
        return rc;
    PK_HIP(hipStreamSynchronize(s->stream));
    int unused = 0, unusedFirst = 0;
    const int rc2 = read_groups(s, unused, unusedFirst);
    s->prof.recs.clear();
    s->prof.used = 0;
    return rc2;
}

int pk_system_get_matrix(pk_system* s, double* full_vals, double* rhs, double* minv)
{
    if (!s)
        return PK_ERR_ARG;
    const Topology& t = s->topo;
    const size_t F = static_cast<size_t>(t.nnzf());
    PK_HIP(hipStreamSynchronize(s->stream));
    if (full_vals && F && is_upper(s))
    {
        // one element per upper slot: the block at its full index, its
        // transpose at the mirror
        const size_t U = static_cast<size_t>(t.nnzu()), nE = static_cast<size_t>(s->nE);
        const bool f32 = is_f32(s);
        std::vector<char> bytes(nE * SliceStore<double>::kWords * (f32 ? 4 : 8));
        PK_HIP(hipMemcpy(bytes.data(), s->vals.p, bytes.size(), hipMemcpyDeviceToHost));
        std::vector<int> ell(U);
        PK_HIP(hipMemcpy(ell.data(), s->slotEll.p, U * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t u = 0; u < U; ++u)
        {
            double w[9];
            for (int e = 0; e < 9; ++e)
                w[e] = f32 ? reinterpret_cast<const float*>(bytes.data())[SliceStore<float>::word(nE, ell[u], e)]
                           : reinterpret_cast<const double*>(bytes.data())[SliceStore<double>::word(nE, ell[u], e)];
            const size_t full = static_cast<size_t>(t.uFull[u]);
            for (int e = 0; e < 9; ++e)
                full_vals[9 * full + e] = w[e];
            if (t.uMirror[u] >= 0)
            {
                const size_t m = static_cast<size_t>(t.uMirror[u]);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        full_vals[9 * m + 3 * i + j] = w[3 * j + i];
            }
        }
    }
    else if (full_vals && F)
    {
        const bool f32 = is_f32(s);
        const size_t es = f32 ? 4 : 8;
        std::vector<char> raw(F * (f32 ? BsrStore<float>::kWords : BsrStore<double>::kWords) * es);
        PK_HIP(hipMemcpy(raw.data(), s->vals.p, raw.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < F; ++k)
            for (int e = 0; e < 9; ++e)
                full_vals[9 * k + e] = f32 ? reinterpret_cast<const float*>(raw.data())[BsrStore<float>::at(F, k, e)]
                                           : reinterpret_cast<const double*>(raw.data())[BsrStore<double>::at(F, k, e)];
    }
    if (rhs && t.nd)
        PK_HIP(hipMemcpy(rhs, s->r.p, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (minv && t.nd)
    {
        const bool f32 = is_f32(s);
        std::vector<char> raw(static_cast<size_t>(t.nd) * pk::kMinvPlanes * (f32 ? 4 : 8));
        PK_HIP(hipMemcpy(raw.data(), s->minv.p, raw.size(), hipMemcpyDeviceToHost));
        // the device keeps the upper triangle (the blocks are symmetric)
        constexpr int plane[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int r = 0; r < t.nd; ++r)
            for (int e = 0; e < 9; ++e)
            {
                const size_t i = static_cast<size_t>(plane[e]) * t.nd + r;
                minv[9 * r + e] = f32 ? reinterpret_cast<const float*>(raw.data())[i]
                                      : reinterpret_cast<const double*>(raw.data())[i];
            }
    }
    return PK_OK;
}

namespace
{
// Runs the inspection kernel and returns every tet's EREC_COUNT record.
int element_records(pk_system* s, std::vector<double>& raw)
{
    const size_t T = static_cast<size_t>(s->topo.T);
    raw.clear();
    PK_HIP(hipStreamSynchronize(s->stream));
    if (T == 0)
        return PK_OK;
    DevBuf<double> tmp;
    int rc;
    if ((rc = tmp.alloc(T * EREC_COUNT)) != PK_OK)
        return rc;
    launch_element_records(s->view(), tmp.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(s->stream));
    raw.resize(T * EREC_COUNT);
    PK_HIP(hipMemcpy(raw.data(), tmp.p, raw.size() * sizeof(double), hipMemcpyDeviceToHost));
    return PK_OK;
}
}  // namespace

int pk_system_get_elements(pk_system* s, double* blocks, double* forces, double* kv)
{
    if (!s)
        return PK_ERR_ARG;
    std::vector<double> raw;
    const int rc = element_records(s, raw);
    if (rc != PK_OK)
        return rc;
    const size_t T = static_cast<size_t>(s->topo.T);
    for (size_t i = 0; i < T; ++i)
    {
        const double* r = raw.data() + EREC_COUNT * i;
        for (int k = 0; k < 90; ++k)
            if (blocks) blocks[90 * i + k] = r[EREC_K + k];
        for (int k = 0; k < 12; ++k)
        {
            if (forces) forces[12 * i + k] = r[EREC_F + k];
            if (kv) kv[12 * i + k] = r[EREC_KV + k];
        }
    }
    return PK_OK;
}

int pk_system_get_rotations(pk_system* s, double* rotations, int* kind)
{
    if (!s)
        return PK_ERR_ARG;
    std::vector<double> raw;
    const int rc = element_records(s, raw);
    if (rc != PK_OK)
        return rc;
    const size_t T = static_cast<size_t>(s->topo.T);
    for (size_t i = 0; i < T; ++i)
    {
        const double* r = raw.data() + EREC_COUNT * i;
        for (int k = 0; k < 9; ++k)
            if (rotations) rotations[9 * i + k] = r[EREC_R + k];
        if (kind) kind[i] = static_cast<int>(r[EREC_KIND]);
    }
    return PK_OK;
}

int pk_system_spmv(pk_system* s, const double* x, double* y)
{
    if (!s || !x || !y)
        return PK_ERR_ARG;
    const Topology& t = s->topo;
    if (t.nd == 0)
        return PK_OK;
    DevBuf<double> dx, dy;
    int rc;
    if ((rc = dx.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    if ((rc = dy.alloc(static_cast<size_t>(t.nd) * 3)) != PK_OK) return rc;
    PK_HIP(hipMemcpyAsync(dx.p, x, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyHostToDevice, s->stream));
    const DeviceView dv = s->view();
    if (is_f32(s))
        launch_spmv_plain<float>(dv, dx.p, dy.p);
    else
        launch_spmv_plain<double>(dv, dx.p, dy.p);
    PK_HIP(hipGetLastError());
    PK_HIP(hipMemcpyAsync(y, dy.p, static_cast<size_t>(t.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PK_HIP(hipStreamSynchronize(s->stream));
    return PK_OK;
}

int pk_system_get_solution(pk_system* s, double* dv)
{
    if (!s || !dv)
        return PK_ERR_ARG;
    PK_HIP(hipStreamSynchronize(s->stream));
    if (s->topo.nd)
        PK_HIP(hipMemcpy(dv, s->xs.p, static_cast<size_t>(s->topo.nd) * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return PK_OK;
}

}  // extern "C"
