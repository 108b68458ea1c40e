// Per-substep assembly kernels for gfx950: element, slot gather, right-hand
// side with the CG start, the integrators, contact forces and frame export.
//
//   k_element        one lane per tet: F = Ds Dm^-1, Newton polar R (or R = I
//                    for the linear model), the tet's basis record (R g_a,
//                    R R^T) and per dynamic corner one entry dt f_a -
//                    dt^2 (K v)_a (or f_a explicit), written in row-list order
//                    (FEMModel.cpp:614-679, :377-424, :559-612, :115-163)
//   k_stiff_records  one lane per tet: V lambda, V mu, V mu g_a.g_b
//   k_assemble_pairs the slot gather: resident workgroups walking chunks of
//                    consecutive rows, one lane per (row, tet) pair forming
//                    the pair's blocks from the basis and stiffness records
//                    into LDS, then one thread per slot adding its
//                    contributions in tet order (the reference's order, no
//                    atomics); A = M + dt^2 K into full or sliced upper
//                    storage and the block-Jacobi inverse (SolverData.cpp:
//                    308-465, :260-306)
//   k_assemble       the same by 16 lanes per dynamic row (the diagonal block
//                    over the row's corner list, then its off-diagonal upper
//                    slots four at a time, blocks added in tet order by
//                    shuffles), for a system with a row too large for a chunk
//   k_rhs            b = sum of the row's corner entries + dt m g (+ contacts)
//                    per dynamic row, then the CG start x = 0, r = b,
//                    z = d = M^-1 r with the (r.z, r.r) partials
//                    (ConjugateGradientSolver.cpp:151-164)
//   k_integrate      all-or-nothing v += dv, x += dt v per group, for the
//                    groups whose solve has ended (World.cpp:644-695)
//   k_explicit       explicit Euler (World.cpp:697-725)
#include "pk_device.h"
#include "pk_fem.h"
#include "pk_topology.h"

namespace pk
{

__global__ void k_reset(GroupState* gs, int G, int* active, int* kbase, double tol, int maxIt)
{
    __shared__ int cnt;
    if (threadIdx.x == 0)
    {
        cnt = 0;
        *kbase = 0;
    }
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
    {
        GroupState& s = gs[g];
        s.delta[0] = s.delta[1] = s.alpha[0] = s.alpha[1] = s.target = s.residual = 0.0;
        s.tol = tol;
        s.maxIt = maxIt;
        s.iters = 0;
        s.stopAt = s.hasRows ? -1 : 0;
        s.stopInUpdate = 0;
        s.dirStop = -1;
        s.converged = 0;
        s.failed = 0;
        s.nonfinite = 0;
        s.integrated = 0;
        if (s.hasRows)
            atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *active = cnt;
}

// Everything one tet contributes to a substep (FEMModel.cpp:614-679 with
// :377-424, :336-375 and :115-163), streamed to a sink as it is formed so that
// no output waits in registers for the others:
//   sink.rotation(R, kind)   R (row-major) and how it was found: 0 the Newton
//                            polar limit, 1 with column 2 negated, 2 Identity
//                            fallback, 3 the linear model (R = I)
//   sink.basis(k, value)     the 18 basis values: R g_a (k = 3a + i), then
//                            R R^T (xx, xy, xz, yy, yz, zz)
//   sink.corner(a, f, kv)    corner a's element force (elastic + damping) and
//                            (K v)_a = sum_b R K_ab R^T v_b
// Both corner terms have the form V R M g_a with a 3x3 M: the stress
// sigma(R^T (x_a - x_0) - (X_a - X_0)) for the force, and
// S(H) = lambda tr(H) I + mu (H + H^T) with H = R^T sum_b v_b g_b^T for K v,
// which is sum_b R Ke_ab R^T v_b written out (Ke_ab = V (lambda g_a g_b^T +
// mu g_b g_a^T + mu (g_a . g_b) I)) for any R. Returns false when the tet
// contributes nothing this substep (the sink is then not called).
template <class Sink>
__device__ __attribute__((always_inline)) inline bool element_eval(int t, size_t TT, const int4* __restrict__ tets,
                                                                   const double* __restrict__ prm,
                                                                   const double* __restrict__ x,
                                                                   const double* __restrict__ v, int model,
                                                                   Sink& sink)
{
    if (model == FEM_NONE)
        return false;
    const int4 n4 = tets[t];
    const int ids[4] = {n4.x, n4.y, n4.z, n4.w};
    double ds[3][3];  // ds[k] = x_{k+1} - x_0
    {
        const double x0[3] = {x[3 * ids[0]], x[3 * ids[0] + 1], x[3 * ids[0] + 2]};
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                ds[k][i] = x[3 * ids[k + 1] + i] - x0[i];
    }
    double Dinv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        Dinv[k] = prm[(P_DMINV + k) * TT + t];
    double R[9];
    int kind;
    if (model == FEM_LINEAR)
    {
        // FEMModel::ComputeElasticForces: no deformation-gradient check and
        // no rotation; u_a = (x_a - x_0) - (X_a - X_0) below is the
        // displacement x_a - X_a relative to corner 0 (B maps both to the
        // same strain)
#pragma unroll
        for (int e = 0; e < 9; ++e)
            R[e] = (e % 4 == 0) ? 1.0 : 0.0;
        kind = 3;
    }
    else
    {
        double F[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = ds[0][i] * Dinv[j] + ds[1][i] * Dinv[3 + j] + ds[2][i] * Dinv[6 + j];
        if (!fin9(F))
            return false;  // the reference skips the tet this substep
        polar(F, R, &kind);
    }
    sink.rotation(R, kind);
    double g[4][3];
    shape_gradients(Dinv, g);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            sink.basis(EB_RG + 3 * a + i, R[3 * i] * g[a][0] + R[3 * i + 1] * g[a][1] + R[3 * i + 2] * g[a][2]);
    sink.basis(EB_RRT + 0, R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    sink.basis(EB_RRT + 1, R[0] * R[3] + R[1] * R[4] + R[2] * R[5]);
    sink.basis(EB_RRT + 2, R[0] * R[6] + R[1] * R[7] + R[2] * R[8]);
    sink.basis(EB_RRT + 3, R[3] * R[3] + R[4] * R[4] + R[5] * R[5]);
    sink.basis(EB_RRT + 4, R[3] * R[6] + R[4] * R[7] + R[5] * R[8]);
    sink.basis(EB_RRT + 5, R[6] * R[6] + R[7] * R[7] + R[8] * R[8]);
    const double lam = prm[P_LAMBDA * TT + t];
    const double mu = prm[P_MU * TT + t];
    // engineering strain of u_a = R^T ds_a - (X_a - X_0), u_0 = 0, and the
    // stress (FEMModel.cpp:174-225, 386-389)
    double Pf[9];
    {
        double exx = 0, eyy = 0, ezz = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
        for (int a = 1; a < 4; ++a)
        {
            double u[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                u[i] = (R[i] * ds[a - 1][0] + R[3 + i] * ds[a - 1][1] + R[6 + i] * ds[a - 1][2]) -
                       prm[(P_EDGE + 3 * (a - 1) + i) * TT + t];
            exx += g[a][0] * u[0];
            eyy += g[a][1] * u[1];
            ezz += g[a][2] * u[2];
            gxy += g[a][1] * u[0] + g[a][0] * u[1];
            gyz += g[a][2] * u[1] + g[a][1] * u[2];
            gxz += g[a][2] * u[0] + g[a][0] * u[2];
        }
        const double tr = exx + eyy + ezz;
        const double sg[9] = {lam * tr + 2.0 * mu * exx, mu * gxy, mu * gxz,
                              mu * gxy, lam * tr + 2.0 * mu * eyy, mu * gyz,
                              mu * gxz, mu * gyz, lam * tr + 2.0 * mu * ezz};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Pf[3 * i + j] = R[3 * i] * sg[j] + R[3 * i + 1] * sg[3 + j] + R[3 * i + 2] * sg[6 + j];
    }
    double vs[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vs[a][i] = v[3 * ids[a] + i];
    double Pk[9];
    {
        double W[9];  // sum_b v_b g_b^T
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                W[3 * i + j] = ((vs[0][i] * g[0][j] + vs[1][i] * g[1][j]) + vs[2][i] * g[2][j]) + vs[3][i] * g[3][j];
        double H[9];  // R^T W
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                H[3 * i + j] = R[i] * W[j] + R[3 + i] * W[3 + j] + R[6 + i] * W[6 + j];
        const double lt = lam * (H[0] + H[4] + H[8]);
        const double S[9] = {lt + 2.0 * mu * H[0], mu * (H[1] + H[3]), mu * (H[2] + H[6]),
                             mu * (H[1] + H[3]), lt + 2.0 * mu * H[4], mu * (H[5] + H[7]),
                             mu * (H[2] + H[6]), mu * (H[5] + H[7]), lt + 2.0 * mu * H[8]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Pk[3 * i + j] = R[3 * i] * S[j] + R[3 * i + 1] * S[3 + j] + R[3 * i + 2] * S[6 + j];
    }
    const double vol = prm[P_VOL * TT + t];
    const double damp = prm[P_DAMP * TT + t];
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
        double f[3], kv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            f[i] = -vol * (Pf[3 * i] * g[a][0] + Pf[3 * i + 1] * g[a][1] + Pf[3 * i + 2] * g[a][2]) - damp * vs[a][i];
            kv[i] = vol * (Pk[3 * i] * g[a][0] + Pk[3 * i + 1] * g[a][1] + Pk[3 * i + 2] * g[a][2]);
        }
        sink.corner(a, f, kv);
    }
    return true;
}

// The sink of the step's element kernel: basis values into the thread's LDS
// record, and per dynamic corner one entry at its place in the row lists:
// implicit, dt f - dt^2 (K v) (the force part dropped when it is not finite,
// as SolverData::PrecomputeImplicitSolve skips a non-finite push,
// SolverData.cpp:145-154); explicit, the force f.
struct StepSink
{
    double* rec;  // the thread's 18 LDS slots
    const int* cpos;
    size_t TT;
    int t;
    double* entries;
    double dt, dt2;
    int implicit;

    __device__ void rotation(const double*, int) {}
    __device__ void basis(int k, double val) { rec[k] = val; }
    __device__ void corner_at(int a, const double* f, const double* kv)
    {
        const int pos = cpos[a * TT + t];
        if (pos < 0)
            return;
        double* out = entries + static_cast<size_t>(pos) * N_CORNER;
        if (!implicit)
        {
            out[0] = f[0];
            out[1] = f[1];
            out[2] = f[2];
            return;
        }
        const bool ff = fin(f[0]) && fin(f[1]) && fin(f[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[i] = (ff ? f[i] * dt : 0.0) - kv[i] * dt2;
    }
    __device__ void corner(int a, const double* f, const double* kv) { corner_at(a, f, kv); }
};

// Writes the tet's basis record and each dynamic corner's entry.
__global__ __launch_bounds__(NT) void k_element(int T, const int4* __restrict__ tets,
                                                const unsigned char* __restrict__ live,
                                                const unsigned char* __restrict__ fem,
                                                const double* __restrict__ prm, const double* __restrict__ x,
                                                const double* __restrict__ v, const int* __restrict__ cpos,
                                                double* __restrict__ basis, double* __restrict__ corner, double dt,
                                                int implicit)
{
    const int t0 = xcd_block() * NT;
    const int t = t0 + static_cast<int>(threadIdx.x);
    const size_t TT = static_cast<size_t>(T);
    // the workgroup's basis records go out as one contiguous run: each thread
    // puts its record in LDS, then the workgroup copies the run with
    // consecutive 16-byte stores (a tet that contributes nothing stores
    // zeros, which is what it holds)
    __shared__ double2 stage[NT * EB_COUNT / 2];
    double* rec = reinterpret_cast<double*>(stage) + threadIdx.x * EB_COUNT;
    const bool work = t < T && live[t];
    StepSink sink{rec, cpos, TT, t, corner, dt, dt * dt, implicit};
    const bool ok = work && element_eval(t, TT, tets, prm, x, v, fem ? fem[t] : FEM_COROTATIONAL, sink);
    if (!ok)
    {
#pragma unroll
        for (int k = 0; k < EB_COUNT; ++k)
            rec[k] = 0.0;
        if (work)
        {
            const double z[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int a = 0; a < 4; ++a)
                sink.corner_at(a, z, z);
        }
    }
    __syncthreads();
    const int nrec = min(NT, T - t0);
    double2* dst = reinterpret_cast<double2*>(basis + static_cast<size_t>(t0) * EB_COUNT);
    // whole lines, read only by the next kernel: stored past the L2, which
    // the end of the launch would otherwise have to write back
    for (int i = threadIdx.x; i < nrec * (EB_COUNT / 2); i += NT)
    {
        const double2 u = stage[i];
        __builtin_nontemporal_store(u.x, &dst[i].x);
        __builtin_nontemporal_store(u.y, &dst[i].y);
    }
}

// g_a . g_b of the 10 local pairs into the parameter planes (rest data, once
// per mesh), with the arithmetic the element kernel uses.
__global__ __launch_bounds__(NT) void k_gdot(int T, double* __restrict__ prm)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= T)
        return;
    const size_t TT = static_cast<size_t>(T);
    double Dinv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        Dinv[k] = prm[(P_DMINV + k) * TT + t];
    double g[4][3];
    shape_gradients(Dinv, g);
#pragma unroll
    for (int p = 0; p < 10; ++p)
    {
        int a, b;
        pair_of(p, a, b);
        prm[(P_GDOT + p) * TT + t] = gdot(g[a], g[b]);
    }
}

// The stiffness record of each tet from the parameter planes (at creation,
// after k_gdot, and whenever lambda and mu change).
__global__ __launch_bounds__(NT) void k_stiff_records(int T, const double* __restrict__ prm, double* __restrict__ kdat)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= T)
        return;
    const size_t TT = static_cast<size_t>(T);
    const double vol = prm[P_VOL * TT + t];
    const double vm = vol * prm[P_MU * TT + t];
    double* k = kdat + static_cast<size_t>(t) * KD_COUNT;
    k[KD_VLAM] = vol * prm[P_LAMBDA * TT + t];
    k[KD_VMU] = vm;
#pragma unroll
    for (int p = 0; p < 10; ++p)
        k[KD_VMUG + p] = vm * prm[(P_GDOT + p) * TT + t];
}

// Block R K_ab R^T of local pair p of a tet from its basis record rec and
// its stiffness record kd.
__device__ inline void element_block_from(const double* __restrict__ rec, const double* __restrict__ kd, int p,
                                          double* K)
{
    int a, b;
    pair_of(p, a, b);
    double rga[3], rgb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        rga[i] = rec[EB_RG + 3 * a + i];
        rgb[i] = rec[EB_RG + 3 * b + i];
    }
    // R R^T: 48 bytes on a 16-byte boundary (the record is 144 bytes)
    const double2* rr = reinterpret_cast<const double2*>(rec + EB_RRT);
    const double2 rr0 = rr[0], rr1 = rr[1], rr2 = rr[2];
    const double r00 = rr0.x, r01 = rr0.y, r02 = rr1.x, r11 = rr1.y, r12 = rr2.x, r22 = rr2.y;
    const double RRt[9] = {r00, r01, r02, r01, r11, r12, r02, r12, r22};
    const double2 f = *reinterpret_cast<const double2*>(kd + KD_VLAM);
    if (a == b)
        rotated_diag_block_scaled(rga, RRt, f.x, f.y, kd[KD_VMUG + p], K);
    else
        rotated_block_scaled(rga, rgb, RRt, f.x, f.y, kd[KD_VMUG + p], K);
}

// Block R K_ab R^T of local pair p of tet t from the basis and stiffness records.
__device__ inline void element_block(const double* __restrict__ basis, const double* __restrict__ kdat, int t, int p,
                                     double* K)
{
    element_block_from(basis + static_cast<size_t>(t) * EB_COUNT, kdat + static_cast<size_t>(t) * KD_COUNT, p, K);
}

// The inspection sink: everything into registers.
struct RecordSink
{
    double basisv[EB_COUNT];
    double f[4][3], kv[4][3];
    double R[9];
    int kind;

    __device__ void rotation(const double* r, int k)
    {
#pragma unroll
        for (int e = 0; e < 9; ++e)
            R[e] = r[e];
        kind = k;
    }
    __device__ void basis(int k, double val) { basisv[k] = val; }
    __device__ void corner(int a, const double* ff, const double* kk)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            f[a][i] = ff[i];
            kv[a][i] = kk[i];
        }
    }
};

// The element outputs of every tet as one record (10 blocks, 4 forces, 4 K v,
// the rotation R and how polar() found it: EREC_COUNT doubles), for
// inspection; zero (and kind -1) for tets that contribute nothing. The blocks
// are formed from the basis values exactly as the slot gather forms them.
__global__ __launch_bounds__(NT) void k_element_records(int T, const int4* __restrict__ tets,
                                                        const unsigned char* __restrict__ live,
                                                        const unsigned char* __restrict__ fem,
                                                        const double* __restrict__ prm,
                                                        const double* __restrict__ kdat, const double* __restrict__ x,
                                                        const double* __restrict__ v, double* __restrict__ out)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= T)
        return;
    const size_t TT = static_cast<size_t>(T);
    RecordSink o;
    const bool ok = live[t] && element_eval(t, TT, tets, prm, x, v, fem ? fem[t] : FEM_COROTATIONAL, o);
    double* r = out + EREC_COUNT * static_cast<size_t>(t);
    for (int p = 0; p < 10; ++p)
    {
        double K[9];
        if (ok)
            element_block_from(o.basisv, kdat + static_cast<size_t>(t) * KD_COUNT, p, K);
        for (int e = 0; e < 9; ++e)
            r[EREC_K + 9 * p + e] = ok ? K[e] : 0.0;
    }
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i)
        {
            r[EREC_F + 3 * a + i] = ok ? o.f[a][i] : 0.0;
            r[EREC_KV + 3 * a + i] = ok ? o.kv[a][i] : 0.0;
        }
    for (int e = 0; e < 9; ++e)
        r[EREC_R + e] = ok ? o.R[e] : 0.0;
    r[EREC_KIND] = ok ? static_cast<double>(o.kind) : -1.0;
}

// Lanes per upper slot in the assembly gather: a node's star has ~20
// contributions and an edge ~5, so one lane per slot leaves most lanes of a
// wave idle while its diagonal slots finish.
constexpr int kSlotLanes = 4;

// Index of the stored local pair (a, b), a <= b (pair_of's inverse).
__device__ constexpr int kPairOf[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};

// Lanes per dynamic row in the assembly gather: a node's star holds ~24
// tets, each of the row's ~7 off-diagonal upper slots ~5.
constexpr int kRowGatherLanes = 16;

// Adds the blocks b of a group of `lanes` lanes to a, in lane order, for
// the first n lanes: every lane of the group gets the same sum, added in
// list order.
template <int lanes>
__device__ inline void add_in_lane_order(double* a, const double* b, int n)
{
    for (int l = 0; l < n; ++l)
#pragma unroll
        for (int e = 0; e < 9; ++e)
            a[e] += __shfl(b[e], l, lanes);
}

// A = dt^2 K (+ contacts) of off-diagonal upper slot s from its contribution
// list: the 4 lanes of a sub-group form 4 blocks at a time, added in list
// order (tet order) by shuffles; the sub-group's first lane writes the block
// to both triangles (SolverData.cpp:422-450).
template <class M>
__device__ inline void gather_slot(int s, int lane4, const int* __restrict__ cstart, const int* __restrict__ clist,
                                   const int* __restrict__ uFull, const int* __restrict__ uMirror,
                                   const double* __restrict__ basis, const double* __restrict__ kdat,
                                   double dt2, M* __restrict__ vals, size_t F, int upperOnly,
                                   const double* __restrict__ cslot, int& failed)
{
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] = 0.0;
    const int k1 = cstart[s + 1];
    for (int k0 = cstart[s]; k0 < k1; k0 += kSlotLanes)
    {
        double b[9] = {};
        if (k0 + lane4 < k1)
        {
            const int c = clist[k0 + lane4];
            double blk[9];
            element_block(basis, kdat, c >> 5, (c >> 1) & 15, blk);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    b[3 * r + q] = (c & 1) ? blk[3 * q + r] : blk[3 * r + q];
        }
        add_in_lane_order<kSlotLanes>(a, b, k1 - k0 < kSlotLanes ? k1 - k0 : kSlotLanes);
    }
    if (lane4 != 0)
        return;
    if (cslot)
    {
        const double cs = cslot[s];
        a[0] += cs;
        a[4] += cs;
        a[8] += cs;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] *= dt2;
    // full storage: the block at its full index and its transpose at the
    // mirror; upper-only storage: the slot's element
    const int full = uFull[s];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        vals[upperOnly ? SliceStore<M>::word(F, full, e) : BsrStore<M>::at(F, full, e)] = static_cast<M>(a[e]);
    if (!upperOnly)
    {
        const int mirror = uMirror[s];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                vals[BsrStore<M>::at(F, mirror, 3 * i + j)] = static_cast<M>(a[3 * j + i]);
    }
    if (!fin9(a))
        failed = 1;  // a non-finite stiffness block rejects the assembly
}

// The assembly gather, atomic-free: kRowGatherLanes lanes per dynamic row r.
// First the diagonal block over the row's corner list (all lanes, in tet
// order): A_rr = m I + dt^2 (sum_a R K_aa R^T + contacts) and its block-Jacobi
// inverse (SolverData.cpp:407-420, 260-306); then the row's off-diagonal
// upper slots, four at a time, each by a sub-group of 4 lanes. A row's blocks
// all come from the tets around node r, so they are read while in L2. A
// non-finite block or a refused inverse marks the body's solve failed.
// uFull maps each upper slot to its full index (full storage, nnzf blocks)
// or to its element (sliced upper-only storage, nnzf then holds nE).
template <class M>
__global__ __launch_bounds__(NT) void k_assemble(int nnzu, int nd, const int* __restrict__ cstart,
                                                 const int* __restrict__ clist, const int* __restrict__ nstart,
                                                 const int* __restrict__ nlist, const int* __restrict__ diagSlot,
                                                 const int* __restrict__ uFull, const int* __restrict__ uMirror,
                                                 const int* __restrict__ rowGroup, const double* __restrict__ basis,
                                                 const double* __restrict__ kdat,
                                                 const double* __restrict__ mrow, double dt2, M* __restrict__ vals,
                                                 long long nnzf, M* __restrict__ minv, GroupState* gs, int upperOnly,
                                                 const double* __restrict__ cslot)
{
    const int row = (xcd_block() * NT + static_cast<int>(threadIdx.x)) / kRowGatherLanes;
    const int lane = threadIdx.x % kRowGatherLanes;
    if (row >= nd)
        return;  // whole lane groups
    const size_t F = static_cast<size_t>(nnzf);
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        a[e] = 0.0;
    const int k1 = nstart[row + 1];
    for (int k0 = nstart[row]; k0 < k1; k0 += kRowGatherLanes)
    {
        double b[9] = {};
        if (k0 + lane < k1)
        {
            const int c = nlist[k0 + lane];
            element_block(basis, kdat, c >> 2, kPairOf[c & 3][c & 3], b);
        }
        add_in_lane_order<kRowGatherLanes>(a, b, k1 - k0 < kRowGatherLanes ? k1 - k0 : kRowGatherLanes);
    }
    const int sd = diagSlot[row];
    const int se = row + 1 < nd ? diagSlot[row + 1] : nnzu;
    const int g = rowGroup[row];
    int failed = 0;
    if (lane == 0)
    {
        if (cslot)
        {
            // point connections' and contacts' k w_a w_b I (PointConnection.cpp:110-128)
            const double cs = cslot[sd];
            a[0] += cs;
            a[4] += cs;
            a[8] += cs;
        }
#pragma unroll
        for (int e = 0; e < 9; ++e)
            a[e] *= dt2;
        const double m = mrow[row];
        a[0] = m + a[0];
        a[4] = m + a[4];
        a[8] = m + a[8];
        const int full = uFull[sd];
#pragma unroll
        for (int e = 0; e < 9; ++e)
            vals[upperOnly ? SliceStore<M>::word(F, full, e) : BsrStore<M>::at(F, full, e)] = static_cast<M>(a[e]);
        const bool ok = fin9(a);
        if (!ok)
            failed = 1;
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
                inv3_sym(a, mi);  // the block is symmetric bit for bit
                pok = fin9(mi);
            }
            if (!pok)
                failed = 1;  // BuildInversePreconditioner fails the solve
        }
        minv_store(minv, static_cast<size_t>(nd), row, mi);
    }
    // the row's off-diagonal slots (sd + 1 .. se - 1), sub-group lane / 4 on
    // slots sd + 1 + sub, sd + 5 + sub, ...
    const int sub = lane / kSlotLanes;
    for (int s = sd + 1 + sub; s < se; s += kRowGatherLanes / kSlotLanes)
        gather_slot(s, lane % kSlotLanes, cstart, clist, uFull, uMirror, basis, kdat, dt2, vals, F, upperOnly,
                    cslot, failed);
    if (failed)
        gs[g].failed = 1;
}

// The assembly gather by (row, tet) pairs, atomic-free. The host cuts the
// dynamic rows into chunks of consecutive rows (at most kPairChunk pairs of
// the rows' corner lists, kOffChunk off-diagonal contributions, kRowChunk
// rows and kSlotChunk upper slots); workgroup w takes chunks w, w + G,
// w + 2G, ... (G workgroups, as many as are resident at once). Per chunk:
// Phase 1, one lane per pair (row r, tet t, corner a of r): the lane forms
// from t's basis and stiffness records the blocks the pair contributes --
// R K_aa R^T to r's diagonal slot, R K_ab R^T to the upper slot (r, row of b)
// for every other corner b whose row is greater -- with the arithmetic of
// gather_slot (the block of the local pair (min, max), transposed when
// a > b), into LDS planes: diagonal blocks at the pair's place, off-diagonal
// ones at their place in the chunk's run of the off-diagonal lists (pairOff,
// three 10-bit fields, 0x3ff for none); the other lookups go to LDS as well:
// each slot's run descriptor (pack_run) and element, each row's mass, body
// and preconditioner switch. Phase 2, one thread per slot: its 9 components'
// contributions summed in list order, which is tet order (the order the
// reference adds them, SolverData.cpp:407-450), then the contacts, dt^2 and
// the mass as in k_assemble, written to full or sliced upper storage. Phase
// 3, one thread per row: the block-Jacobi inverse (SolverData.cpp:260-306),
// kept as its upper triangle. A non-finite block or a refused inverse marks
// the body's solve failed.
// The loads run ahead of the arithmetic, so that a chunk's three phases
// overlap the dependent loads (descriptor -> pair -> records) of the next
// ones: chunk descriptors three chunks ahead (scalar loads), each lane's pair
// and lookups two ahead, the pair's records and the rows' preconditioner
// switches one ahead. None of these loads is conditional (a lane past a
// chunk's pairs, slots or rows reads the chunk's first one; past the last
// chunk the workgroup reads its last one again), and nothing reads what they
// return before the next chunk: the compiler can then count the loads in
// flight, and a chunk waits only for loads issued during the chunk before.

// One lane's pair of a chunk: its corner-list entry (tet << 2 | corner) and
// its three off-diagonal offsets.
struct PairLane
{
    int cc, offs;
};

// One lane's lookups of a chunk: the run descriptor, element and mirror
// element (full storage) of slot lane, the body and mass of row lane.
struct PairLookups
{
    int run, elem, mirror, body;
    double mass;
};

// The records of a pair (t, a): t's basis record (R g_0 .. R g_3, R R^T),
// V lambda, V mu and V mu g_a.g_b for b = 0 .. 3.
struct PairRecords
{
    double w[12], rr[6], vl, vm, vg[4];
};

__device__ inline PairLane load_pair_lane(const int4& ch, int tid, const int* __restrict__ nlist,
                                          const int* __restrict__ pairOff)
{
    const int k = ch.y + (tid < (ch.w >> 16) ? tid : 0);
    return PairLane{nlist[k], pairOff[k]};
}

__device__ inline PairLookups load_pair_lookups(const int4& ch, int tid, const int* __restrict__ slotRun,
                                                const int* __restrict__ uElem, const int* __restrict__ mirror,
                                                const int* __restrict__ rowGroup, const double* __restrict__ mrow)
{
    const int s = ch.z + (tid < ((ch.w >> 8) & 0xff) ? tid : 0);
    const int r = ch.x + (tid < (ch.w & 0xff) ? tid : 0);
    return PairLookups{slotRun[s], uElem[s], mirror[s], rowGroup[r], mrow[r]};
}

// V mu g_a.g_b of local pair (a, b) sits at KD_VMUG + 4 lo - lo (lo - 1) / 2
// + hi - lo (lo, hi = min, max), the stored pair order.
__device__ inline void load_pair_records(const PairLane& pl, const double* __restrict__ basis,
                                         const double* __restrict__ kdat, PairRecords& R)
{
    const int t = pl.cc >> 2, a = pl.cc & 3;
    const double2* rec2 = reinterpret_cast<const double2*>(basis + static_cast<size_t>(t) * EB_COUNT);
    const double* kd = kdat + static_cast<size_t>(t) * KD_COUNT;
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        const double2 u = rec2[i];
        R.w[2 * i] = u.x;
        R.w[2 * i + 1] = u.y;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const double2 u = rec2[EB_RRT / 2 + i];
        R.rr[2 * i] = u.x;
        R.rr[2 * i + 1] = u.y;
    }
    const double2 lm = *reinterpret_cast<const double2*>(kd + KD_VLAM);
    R.vl = lm.x;
    R.vm = lm.y;
#pragma unroll
    for (int b = 0; b < 4; ++b)
    {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        R.vg[b] = kd[KD_VMUG + 4 * lo - (lo * (lo - 1)) / 2 + hi - lo];
    }
}

// v[k] for a lane-varying k in 0..3, as selects (no indexed registers).
__device__ inline double pick4(int k, double v0, double v1, double v2, double v3)
{
    return k == 0 ? v0 : (k == 1 ? v1 : (k == 2 ? v2 : v3));
}

// Phase 1's blocks of one pair into the LDS planes P (diagonal, at the pair's
// place tid) and O (off-diagonal, at the offsets).
__device__ inline void form_pair_blocks(const PairLane& pl, const PairRecords& R, int tid, double* P, double* O)
{
    const int a = pl.cc & 3;
    const double RRt[9] = {R.rr[0], R.rr[1], R.rr[2], R.rr[1], R.rr[3], R.rr[4], R.rr[2], R.rr[4], R.rr[5]};
    double wa[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        wa[i] = pick4(a, R.w[i], R.w[3 + i], R.w[6 + i], R.w[9 + i]);
    double K[9];
    rotated_diag_block_scaled(wa, RRt, R.vl, R.vm, pick4(a, R.vg[0], R.vg[1], R.vg[2], R.vg[3]), K);
#pragma unroll
    for (int e = 0; e < 9; ++e)
        P[e * kPairChunk + tid] = K[e];
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        const int o = (pl.offs >> (10 * j)) & 0x3ff;
        if (o == 0x3ff)
            continue;
        const int b = j < a ? j : j + 1;  // the other corners in increasing order
        double wb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            wb[i] = pick4(b, R.w[i], R.w[3 + i], R.w[6 + i], R.w[9 + i]);
        // the stored block of the local pair (min, max), transposed when the
        // row's corner is the larger one
        const bool flip = a > b;
        double lo[3], hi[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            lo[i] = flip ? wb[i] : wa[i];
            hi[i] = flip ? wa[i] : wb[i];
        }
        rotated_block_scaled(lo, hi, RRt, R.vl, R.vm, pick4(b, R.vg[0], R.vg[1], R.vg[2], R.vg[3]), K);
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y)
                O[(3 * x + y) * kOffChunk + o] = flip ? K[3 * y + x] : K[3 * x + y];
    }
}

static_assert(kSlotChunk <= NT && kRowChunk <= NT, "one lane per slot and per row of a chunk");

// The 9 sums of a slot's run in the LDS block planes (component e of term j
// at src[e * stride + j]), each added in list order (tet order) from zero;
// two terms of all 9 components in flight at a time.
__device__ inline void sum_run(const double* src, int stride, int n, double* v)
{
#pragma unroll
    for (int e = 0; e < 9; ++e)
        v[e] = 0.0;
    int j = 0;
    for (; j + 2 <= n; j += 2)
    {
        double a[9], b[9];
#pragma unroll
        for (int e = 0; e < 9; ++e)
        {
            a[e] = src[e * stride + j];
            b[e] = src[e * stride + j + 1];
        }
#pragma unroll
        for (int e = 0; e < 9; ++e)
            v[e] = (v[e] + a[e]) + b[e];
    }
    if (j < n)
    {
#pragma unroll
        for (int e = 0; e < 9; ++e)
            v[e] += src[e * stride + j];
    }
}

template <class M>
__global__ __launch_bounds__(NT) void k_assemble_pairs(
    int nchunks, const int4* __restrict__ chunks, int nd, const int* __restrict__ nlist,
    const int* __restrict__ pairOff, const int* __restrict__ slotRun, const int* __restrict__ uElem,
    const int* __restrict__ uMirror,
    const int* __restrict__ rowGroup, const double* __restrict__ basis, const double* __restrict__ kdat,
    const double* __restrict__ mrow, double dt2, M* __restrict__ vals, long long nnzf, M* __restrict__ minv,
    GroupState* gs, int upperOnly, const double* __restrict__ cslot)
{
    extern __shared__ __align__(16) double pk_pair_lds[];
    double* P = pk_pair_lds;                 // 9 planes of kPairChunk: diagonal blocks
    double* O = P + 9 * kPairChunk;          // 9 planes of kOffChunk: off-diagonal blocks
    double* D = O + 9 * kOffChunk;           // kRowChunk diagonal blocks (row-major)
    double* RM = D + 9 * kRowChunk;          // kRowChunk row masses
    int* RG = reinterpret_cast<int*>(RM + kRowChunk);  // kRowChunk bodies (-1 - body: no preconditioner)
    int* SR = RG + kRowChunk;                // kSlotChunk run descriptors
    int* SE = SR + kSlotChunk;               // kSlotChunk elements
    int* SM = SE + kSlotChunk;               // kSlotChunk mirror elements (full storage)
    const int G = static_cast<int>(gridDim.x);
    const int first = xcd_block();
    if (first >= nchunks)
        return;
    const int niter = (nchunks - first + G - 1) / G;  // chunks first + i G, i < niter
    const int tid = static_cast<int>(threadIdx.x);
    const size_t F = static_cast<size_t>(nnzf);
    // upper-only storage has no mirrors: read the elements again instead
    const int* mirrorOf = upperOnly ? uElem : uMirror;

    // the descriptor of the workgroup's chunk i (its last chunk past the end)
    auto desc = [&](int i) { return chunks[first + (i < niter ? i : niter - 1) * G]; };

    // chunk 0's pair, lookups, records and switches; chunk 1's pair and
    // lookups; the descriptors of chunks 0 .. 2
    int4 c0 = desc(0), c1 = desc(1), c2 = desc(2);
    PairLane p0 = load_pair_lane(c0, tid, nlist, pairOff);
    PairLookups l0 = load_pair_lookups(c0, tid, slotRun, uElem, mirrorOf, rowGroup, mrow);
    PairLane p1 = load_pair_lane(c1, tid, nlist, pairOff);
    PairLookups l1 = load_pair_lookups(c1, tid, slotRun, uElem, mirrorOf, rowGroup, mrow);
    PairRecords r0;
    load_pair_records(p0, basis, kdat, r0);
    int j0 = gs[l0.body].jacobi;

    for (int i = 0; i < niter; ++i)
    {
        const int4 ch = c0;
        const int r0w = ch.x, sBase = ch.z;
        const int nR = ch.w & 0xff, nS = (ch.w >> 8) & 0xff, nP = ch.w >> 16;

        // ahead: chunk i + 3's descriptor, chunk i + 1's records and
        // switches, chunk i + 2's pair and lookups
        const int4 c3 = desc(i + 3);
        PairRecords r1;
        load_pair_records(p1, basis, kdat, r1);
        const int j1 = gs[l1.body].jacobi;
        const PairLane p2 = load_pair_lane(c2, tid, nlist, pairOff);
        const PairLookups l2 = load_pair_lookups(c2, tid, slotRun, uElem, mirrorOf, rowGroup, mrow);

        // phase 1
        if (tid < nS)
        {
            SR[tid] = l0.run;
            SE[tid] = l0.elem;
            SM[tid] = l0.mirror;
        }
        if (tid < nR)
        {
            RM[tid] = l0.mass;
            RG[tid] = j0 ? l0.body : -1 - l0.body;
        }
        if (tid < nP)
            form_pair_blocks(p0, r0, tid, P, O);
        __syncthreads();

        // phase 2: one thread per slot, its 9 components at once
        if (tid < nS)
        {
            const int run = SR[tid];
            const int begin = run & 0x3ff, n = (run >> 10) & 0x3ff, row = (run >> 20) & 0x3f;
            const bool diag = (run >> 26) & 1;
            double v[9];
            sum_run(diag ? P + begin : O + begin, diag ? kPairChunk : kOffChunk, n, v);
            if (cslot)
            {
                // point connections' and contacts' k w_a w_b I (PointConnection.cpp:110-128)
                const double cs = cslot[sBase + tid];
                v[0] += cs;
                v[4] += cs;
                v[8] += cs;
            }
            bool bad = false;
#pragma unroll
            for (int e = 0; e < 9; ++e)
            {
                v[e] *= dt2;
                if (diag && (e == 0 || e == 4 || e == 8))
                    v[e] = RM[row] + v[e];
                bad = bad || !fin(v[e]);
            }
            const int full = SE[tid];
            if (upperOnly)
                SliceStore<M>::put(vals, F, full, v);
            else
            {
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    vals[BsrStore<M>::at(F, full, e)] = static_cast<M>(v[e]);
                if (!diag)
                {
                    const int mirror = SM[tid];
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y)
                            vals[BsrStore<M>::at(F, mirror, 3 * y + x)] = static_cast<M>(v[3 * x + y]);
                }
            }
            if (bad)
            {
                const int g = RG[row];
                gs[g >= 0 ? g : -1 - g].failed = 1;  // a non-finite block rejects the assembly
            }
            if (diag)
            {
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    D[9 * row + e] = v[e];
            }
        }
        __syncthreads();

        // phase 3 (thread tid reads D, written before the barrier above, and
        // RG[tid], its own: the next chunk's phase 1 may start meanwhile)
        if (tid < nR)
        {
            double A[9];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                A[e] = D[9 * tid + e];
            const int gr = RG[tid];
            double mi[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            if (gr >= 0)
            {
                bool pok = fin9(A);
                if (pok)
                {
                    const double d = det3(A);
                    pok = fin(d) && fabs(d) > static_cast<double>(1.0e-8f);
                }
                if (pok)
                {
                    inv3_sym(A, mi);  // the block is symmetric bit for bit
                    pok = fin9(mi);
                }
                if (!pok)
                    gs[gr].failed = 1;  // BuildInversePreconditioner fails the solve
            }
            minv_store(minv, static_cast<size_t>(nd), r0w + tid, mi);
        }
        c0 = c1;
        c1 = c2;
        c2 = c3;
        p0 = p1;
        l0 = l1;
        r0 = r1;
        j0 = j1;
        p1 = p2;
        l1 = l2;
    }
}

// Lanes per dynamic row in the right-hand side gather (a node has ~20 corners).
constexpr int kRowLanes = 4;

// b = dt (f + m g) - dt^2 K v per dynamic row from its corner entries (each
// dt f - dt^2 (K v) of one tet, k_element, added in tet order), then the CG
// start x = 0, r = b,
// z = d = M^-1 r with the (r.z, r.r) partials (ConjugateGradientSolver.cpp:
// 151-164).
template <class M>
__global__ __launch_bounds__(NT) void k_rhs(TileView tv, const int* __restrict__ nstart,
                                            const double* __restrict__ corner,
                                            const double* __restrict__ mrow, double gx, double gy, double gz,
                                            double dt, const M* __restrict__ minv, int nd,
                                            double* __restrict__ xs, double* __restrict__ rv, M* __restrict__ zv,
                                            M* __restrict__ d0, Partial* __restrict__ part,
                                            const double* __restrict__ crhs)
{
    const int t = tile_of(tv);
    if (t < 0)
        return;
    const RowTile tl = tv.tiles[t];
    double pa = 0.0, pb = 0.0;
    int flag = 0;
    constexpr int RL = kRowLanes, RPP = NT / kRowLanes;
    const int lane = threadIdx.x % RL;
    // every lane group runs the same number of passes (the shuffles below
    // need whole groups); rows past the tile are masked
    for (int row0 = tl.rowBegin; row0 < tl.rowEnd; row0 += RPP)
    {
        const int row = row0 + static_cast<int>(threadIdx.x) / RL;
        const bool live = row < tl.rowEnd;
        // the row's entries in list order (tet order, as the reference adds
        // them): the group's lanes load RL at a time, then every lane adds
        // them one by one (shuffles within the group)
        double e3[3] = {0, 0, 0};
        const int k0 = live ? nstart[row] : 0, k1 = live ? nstart[row + 1] : 0;
        for (int k = k0; k < k1; k += RL)
        {
            double u[3] = {0, 0, 0};
            if (k + lane < k1)
            {
                const double* e = corner + static_cast<size_t>(k + lane) * N_CORNER;
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    u[i] = e[i];
            }
            const int n = k1 - k < RL ? k1 - k : RL;
            for (int l = 0; l < n; ++l)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    e3[i] += __shfl(u[i], l, RL);
        }
        if (!live || lane != 0)
            continue;
        const double m = mrow[row];
        const double gv[3] = {gx, gy, gz};
        double b[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            b[i] = e3[i] + (gv[i] * m) * dt + (crhs ? crhs[3 * row + i] : 0.0);
        if (!(fin(b[0]) && fin(b[1]) && fin(b[2])))
            flag |= 1;
        minv_apply(minv, static_cast<size_t>(nd), row, b, z);
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag |= 2;
        M zs[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            zs[i] = static_cast<M>(z[i]);
            // x and r are read again after a whole SpMV: past the L2 (as in the CG)
            __builtin_nontemporal_store(0.0, &xs[3 * row + i]);
            __builtin_nontemporal_store(b[i], &rv[3 * row + i]);
            zv[3 * row + i] = zs[i];
            d0[3 * row + i] = zs[i];
        }
        pa += b[0] * static_cast<double>(zs[0]) + b[1] * static_cast<double>(zs[1]) + b[2] * static_cast<double>(zs[2]);
        pb += b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    }
    double unused = 0.0;
    block_sum3(pa, pb, unused);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
        part[t] = Partial{pa, pb, flag, 0};
}

__global__ __launch_bounds__(NT) void k_integrate(TileView tv, const int* __restrict__ rowNode,
                                                  const double* __restrict__ dvel, double* __restrict__ x,
                                                  double* __restrict__ v, double dt, GroupState* gs, int commit)
{
    const int t = tile_of(tv);
    if (t < 0)
        return;
    const RowTile tl = tv.tiles[t];
    GroupState& s = gs[tl.group];
    // a body still iterating (a substep graph integrates the bodies whose
    // solve has ended, the host the rest) or already integrated is left alone
    if (s.failed || (commit && s.nonfinite) || s.stopAt < 0 || s.integrated)
        return;
    bool ok = true;
    for (int row = tl.rowBegin + threadIdx.x; row < tl.rowEnd; row += NT)
    {
        const int n = rowNode[row];
        double nv[3], nx[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            const double xo = x[3 * n + i], vo = v[3 * n + i], dd = dvel[3 * row + i];
            nv[i] = vo + dd;
            nx[i] = xo + nv[i] * dt;
            ok = ok && fin(xo) && fin(vo) && fin(dd) && fin(nv[i]) && fin(nx[i]);
        }
        if (commit)
        {
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                v[3 * n + i] = nv[i];
                x[3 * n + i] = nx[i];
            }
        }
    }
    if (!commit && !ok)
        s.nonfinite = 1;
}

__global__ __launch_bounds__(NT) void k_explicit(TileView tv, const int* __restrict__ rowNode,
                                                 const int* __restrict__ nstart, const double* __restrict__ corner,
                                                 const double* __restrict__ mrow,
                                                 double gx, double gy, double gz, double dt, double* __restrict__ x,
                                                 double* __restrict__ v, const double* __restrict__ crhs)
{
    const int t = tile_of(tv);
    if (t < 0)
        return;
    const RowTile tl = tv.tiles[t];
    for (int row = tl.rowBegin + threadIdx.x; row < tl.rowEnd; row += NT)
    {
        double f[3] = {0, 0, 0};
        for (int k = nstart[row]; k < nstart[row + 1]; ++k)
        {
            const double* e = corner + static_cast<size_t>(k) * N_CORNER;
            const double e0 = e[0], e1 = e[1], e2 = e[2];
            if (fin(e0) && fin(e1) && fin(e2))
            {
                f[0] += e0;
                f[1] += e1;
                f[2] += e2;
            }
        }
        if (crhs)
        {
            f[0] += crhs[3 * row];
            f[1] += crhs[3 * row + 1];
            f[2] += crhs[3 * row + 2];
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
}

static inline unsigned grid_of(const TileView& tv) { return 8u * static_cast<unsigned>(tv.per); }

void launch_reset(const DeviceView& v, double tol, int maxIt)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(NT), 0, v.stream, v.gs, v.G, v.active, v.kbase, tol, maxIt);
}

void launch_element(const DeviceView& v, double dt, int implicit)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_element, dim3(xcd_grid((v.T + NT - 1) / NT)), dim3(NT), 0, v.stream, v.T, v.tets, v.live,
                           v.fem, v.prm, v.x, v.v, v.cpos, v.basis, v.corner, dt, implicit);
}

// Workgroups of the pair gather: as many as the current device holds at once
// (a multiple of 8, found once per device), at most one per chunk.
template <class M>
static unsigned pair_gather_grid(int nchunks)
{
    static int resident[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        dev = 0;
    int& n = resident[dev & 63];
    if (n == 0)
    {
        int per = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_assemble_pairs<M>, NT, kPairLds) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            per = cus = 0;
        n = per * cus >= 8 ? per * cus / 8 * 8 : 8;
    }
    const unsigned all = xcd_grid(nchunks);
    return all < static_cast<unsigned>(n) ? all : static_cast<unsigned>(n);
}

template <class M>
void launch_assemble(const DeviceView& v, double dt)
{
    const int n = static_cast<int>(v.nnzu);
    if (n <= 0 || v.nd <= 0)
        return;
    if (v.nchunks > 0)
    {
        hipLaunchKernelGGL((k_assemble_pairs<M>), dim3(pair_gather_grid<M>(v.nchunks)), dim3(NT), kPairLds, v.stream,
                           v.nchunks, v.chunks, v.nd, v.nlist, v.pairOff, v.slotRun,
                           v.upper ? v.slotEll : v.uFull, v.uMirror, v.rowGroup, v.basis, v.kdat, v.mrow, dt * dt,
                           static_cast<M*>(v.vals), v.upper ? v.nE : v.nnzf, static_cast<M*>(v.minv), v.gs, v.upper,
                           v.slotAdd);
        return;
    }
    const long long groups = (static_cast<long long>(v.nd) * kRowGatherLanes + NT - 1) / NT;
    // upper-only storage: slots map to their elements (of nE)
    hipLaunchKernelGGL((k_assemble<M>), dim3(xcd_grid(groups)), dim3(NT), 0, v.stream, n, v.nd, v.cstart, v.clist,
                       v.nstart, v.nlist, v.diagSlot, v.upper ? v.slotEll : v.uFull, v.uMirror, v.rowGroup, v.basis,
                       v.kdat, v.mrow, dt * dt, static_cast<M*>(v.vals), v.upper ? v.nE : v.nnzf,
                       static_cast<M*>(v.minv), v.gs, v.upper, v.slotAdd);
}

template <class M>
void launch_rhs(const DeviceView& v, double gx, double gy, double gz, double dt)
{
    if (v.vec.count > 0)
        hipLaunchKernelGGL((k_rhs<M>), dim3(grid_of(v.vec)), dim3(NT), 0, v.stream, v.vec, v.nstart, v.corner, v.mrow,
                           gx, gy, gz, dt, static_cast<const M*>(v.minv), v.nd, v.xs, v.r, static_cast<M*>(v.z),
                           static_cast<M*>(v.d[0]), v.pb[1], v.rowAdd);
}

// Marks the bodies whose solve has ended as integrated (after both
// k_integrate passes), so a later integration skips them.
__global__ void k_mark_integrated(GroupState* gs, int G)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        if (gs[g].stopAt >= 0)
            gs[g].integrated = 1;
}

void launch_mark_integrated(const DeviceView& v)
{
    if (v.G > 0)
        hipLaunchKernelGGL(k_mark_integrated, dim3((v.G + NT - 1) / NT), dim3(NT), 0, v.stream, v.gs, v.G);
}

void launch_integrate(const DeviceView& v, double dt, int commit)
{
    if (v.vec.count > 0)
        hipLaunchKernelGGL(k_integrate, dim3(grid_of(v.vec)), dim3(NT), 0, v.stream, v.vec, v.rowNode, v.xs, v.x, v.v,
                           dt, v.gs, commit);
}

void launch_explicit(const DeviceView& v, double gx, double gy, double gz, double dt)
{
    if (v.vec.count > 0)
        hipLaunchKernelGGL(k_explicit, dim3(grid_of(v.vec)), dim3(NT), 0, v.stream, v.vec, v.rowNode, v.nstart,
                           v.corner, v.mrow, gx, gy, gz, dt, v.x, v.v, v.rowAdd);
}

// Point connections of a substep (PointConnection.cpp:54-130): the
// barycentric point's position and velocity and the force
// k (target - point) + (-c) velocity, in the reference's order of terms. A
// connection whose point, velocity or target is not finite pushes no force
// (ok = 0; its blocks stay in the matrix).
__global__ __launch_bounds__(NT) void k_conn_eval(int n, const int4* __restrict__ nodes,
                                                  const double* __restrict__ data, const double* __restrict__ x,
                                                  const double* __restrict__ v, double* __restrict__ eval)
{
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= n)
        return;
    const int4 nd = nodes[c];
    const double* d = data + 9 * static_cast<size_t>(c);
    double p[3], u[3];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        p[i] = ((x[3 * nd.x + i] * d[0] + x[3 * nd.y + i] * d[1]) + x[3 * nd.z + i] * d[2]) + x[3 * nd.w + i] * d[3];
        u[i] = ((v[3 * nd.x + i] * d[0] + v[3 * nd.y + i] * d[1]) + v[3 * nd.z + i] * d[2]) + v[3 * nd.w + i] * d[3];
        ok = ok && fin(p[i]) && fin(u[i]) && fin(d[4 + i]);
    }
    double* e = eval + 7 * static_cast<size_t>(c);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        e[i] = ok ? (d[4 + i] - p[i]) * d[7] + u[i] * (-d[8]) : 0.0;
        e[3 + i] = ok ? u[i] : 0.0;
    }
    e[6] = ok ? 1.0 : 0.0;
}

// Each touched dynamic row's share of its connections, summed in connection
// order: w_a (dt F - dt^2 k u) for the implicit right-hand side (the force
// and the -dt^2 K v of the k w_a w_b I blocks over all four nodes, with u the
// point velocity), w_a F as a force for the explicit step.
__global__ __launch_bounds__(NT) void k_conn_rows(int n, const int* __restrict__ rows, const int* __restrict__ start,
                                                  const int2* __restrict__ entries, const double* __restrict__ data,
                                                  const double* __restrict__ eval, double dt, double dt2, int implicit,
                                                  double* __restrict__ crhs)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n)
        return;
    double acc[3] = {0, 0, 0};
    for (int k = start[i]; k < start[i + 1]; ++k)
    {
        const int2 en = entries[k];
        const double w = data[9 * static_cast<size_t>(en.x) + en.y];
        const double kc = data[9 * static_cast<size_t>(en.x) + 7];
        const double* e = eval + 7 * static_cast<size_t>(en.x);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            acc[j] += implicit ? w * (dt * e[j] - dt2 * kc * e[3 + j]) : w * e[j];
    }
    const int r = rows[i];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        crhs[3 * r + j] = acc[j];
}

__global__ __launch_bounds__(NT) void k_slot_scatter(int n, const int* __restrict__ idx, const double* __restrict__ val,
                                                     double* __restrict__ cslot)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n)
        cslot[idx[i]] = val ? val[i] : 0.0;
}

// Frame export: the node state as the reference's float Vec3 triples
// (Vec3.h:7-12, the value static_cast<float> gives, round to nearest even).
// Positions fill out[0 .. n3), velocities out[n3 .. 2 n3); each lane converts
// two consecutive doubles so a wave reads 1 KiB and writes 512 B contiguous.
__global__ __launch_bounds__(NT) void k_export_frame(long long n3, const double* __restrict__ x,
                                                     const double* __restrict__ v, float* __restrict__ out)
{
    const long long pairs = (n3 + 1) / 2;
    const long long stride = static_cast<long long>(gridDim.x) * NT;
    for (long long p = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; p < 2 * pairs; p += stride)
    {
        const bool vel = p >= pairs;
        const long long i = 2 * (vel ? p - pairs : p);
        const double* src = vel ? v : x;
        float* dst = out + (vel ? n3 : 0);
        if (src == nullptr)
            continue;
        if (i + 1 < n3)
        {
            dst[i] = static_cast<float>(src[i]);
            dst[i + 1] = static_cast<float>(src[i + 1]);
        }
        else
            dst[i] = static_cast<float>(src[i]);
    }
}

__global__ __launch_bounds__(NT) void k_zero_rows(int n, const int* __restrict__ rows, double* __restrict__ crhs)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n)
    {
        const int r = rows[i];
        crhs[3 * r] = crhs[3 * r + 1] = crhs[3 * r + 2] = 0.0;
    }
}

void launch_connections(const DeviceView& v, double dt, int implicit)
{
    if (v.nconn <= 0)
        return;
    hipLaunchKernelGGL(k_conn_eval, dim3((v.nconn + NT - 1) / NT), dim3(NT), 0, v.stream, v.nconn, v.connNodes,
                       v.connData, v.x, v.v, v.connEval);
    if (v.nconnRows > 0)
        hipLaunchKernelGGL(k_conn_rows, dim3((v.nconnRows + NT - 1) / NT), dim3(NT), 0, v.stream, v.nconnRows,
                           v.connRows, v.connRowStart, v.connEntries, v.connData, v.connEval, dt, dt * dt, implicit,
                           v.crhs);
}

// CollisionSphereComponent::UpdateSystem (CollisionSphereComponent.cpp:112-193)
// and AddPointConnectionFromContact (:48-66) for one candidate tet: for every
// sphere of positive radius holding the tet's centroid p (distance d < radius,
// a NaN distance included, as in the reference), a connection with weights
// 1/4 to the target p + n (radius - d), n = (p - centre)/d or (0, 0, 1) for
// d <= 1e-5, handled as PointConnection::UpdateSystem (PointConnection.cpp:54-130)
// handles it: nothing when p, the centroid velocity u or the target is not
// finite; else the force F = k (target - p) + (-c) u and the blocks
// k/16 I. Writes ck = sum k and, per corner before the 1/4 weight,
// cg = sum (dt F - dt^2 k u) (implicit) or sum F (explicit).
__global__ __launch_bounds__(NT) void k_sphere_contacts(int n, const int* __restrict__ cand,
                                                        const unsigned char* __restrict__ tact,
                                                        const int4* __restrict__ tets, const double* __restrict__ x,
                                                        const double* __restrict__ v, int nsph,
                                                        const double* __restrict__ sph, double dt, int implicit,
                                                        double* __restrict__ ck, double* __restrict__ cg)
{
    const int i = xcd_block() * NT + static_cast<int>(threadIdx.x);
    if (i >= n)
        return;
    const int tt = cand[i];
    if (!tact[tt])
    {
        // switched off: no contacts (CollisionSphereComponent.cpp:143-146)
        ck[i] = 0.0;
        cg[3 * static_cast<size_t>(i)] = cg[3 * static_cast<size_t>(i) + 1] = cg[3 * static_cast<size_t>(i) + 2] = 0.0;
        return;
    }
    const int4 nd = tets[tt];
    const double w = 0.25;
    double p[3], u[3];
    bool fin3 = true;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        p[j] = ((x[3 * nd.x + j] * w + x[3 * nd.y + j] * w) + x[3 * nd.z + j] * w) + x[3 * nd.w + j] * w;
        u[j] = ((v[3 * nd.x + j] * w + v[3 * nd.y + j] * w) + v[3 * nd.z + j] * w) + v[3 * nd.w + j] * w;
        fin3 = fin3 && fin(p[j]) && fin(u[j]);
    }
    double ksum = 0.0, g[3] = {0.0, 0.0, 0.0};
    const double dt2 = dt * dt;
    for (int s = 0; s < nsph; ++s)
    {
        const double* sp = sph + 6 * s;
        const double r = sp[3];
        if (!(r > 0.0))
            continue;
        const double d[3] = {p[0] - sp[0], p[1] - sp[1], p[2] - sp[2]};
        const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dist >= r)
            continue;
        const bool unit = dist > static_cast<double>(0.00001f);
        const double depth = r - dist;
        if (depth <= 0.0)
            continue;
        double t[3];
        bool ok = fin3;
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double nj = unit ? d[j] / dist : (j == 2 ? 1.0 : 0.0);
            t[j] = p[j] + nj * depth;
            ok = ok && fin(t[j]);
        }
        if (!ok)
            continue;
        const double k = sp[4], c = sp[5];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double F = (t[j] - p[j]) * k + u[j] * (-c);
            g[j] += implicit ? dt * F - dt2 * k * u[j] : F;
        }
        ksum += k;
    }
    ck[i] = ksum;
    cg[3 * static_cast<size_t>(i)] = g[0];
    cg[3 * static_cast<size_t>(i) + 1] = g[1];
    cg[3 * static_cast<size_t>(i) + 2] = g[2];
}

// Lanes per upper slot and per row of the two sums below: a slot gathers about
// 7 candidates on a tet mesh, a row about 24; each lane group walks its list
// with consecutive lanes on consecutive entries and combines the lanes'
// partial sums in a fixed order.
constexpr int kSphereSlotLanes = 4;
constexpr int kSphereRowLanes = 8;

// sslot[s] = base[s] + (1/16) sum of ck over the slot's list (each ordered
// corner pair of a tet adds k w_a w_b = k/16).
__global__ __launch_bounds__(NT) void k_sphere_slots(long long n, const int* __restrict__ start,
                                                     const int* __restrict__ list, const double* __restrict__ ck,
                                                     const double* __restrict__ base, double* __restrict__ sslot)
{
    const long long s = (static_cast<long long>(xcd_block()) * NT + threadIdx.x) / kSphereSlotLanes;
    const int lane = threadIdx.x % kSphereSlotLanes;
    if (s >= n)
        return;  // whole lane groups
    double acc = 0.0;
    const int k1 = start[s + 1];
    for (int k = start[s] + lane; k < k1; k += kSphereSlotLanes)
        acc += ck[list[k]];
#pragma unroll
    for (int o = kSphereSlotLanes / 2; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, kSphereSlotLanes);
    if (lane == 0)
        sslot[s] = (base ? base[s] : 0.0) + 0.0625 * acc;
}

// srhs[3 r + j] = base[3 r + j] + (1/4) sum of cg over the row's list.
__global__ __launch_bounds__(NT) void k_sphere_rows(int n, const int* __restrict__ start, const int* __restrict__ list,
                                                    const double* __restrict__ cg, const double* __restrict__ base,
                                                    double* __restrict__ srhs)
{
    const int r = (xcd_block() * NT + static_cast<int>(threadIdx.x)) / kSphereRowLanes;
    const int lane = threadIdx.x % kSphereRowLanes;
    if (r >= n)
        return;  // whole lane groups
    double acc[3] = {0.0, 0.0, 0.0};
    const int k1 = start[r + 1];
    for (int k = start[r] + lane; k < k1; k += kSphereRowLanes)
    {
        const double* c = cg + 3 * static_cast<size_t>(list[k]);
        acc[0] += c[0];
        acc[1] += c[1];
        acc[2] += c[2];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
#pragma unroll
        for (int o = kSphereRowLanes / 2; o > 0; o >>= 1)
            acc[j] += __shfl_xor(acc[j], o, kSphereRowLanes);
    }
    if (lane == 0)
    {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            srhs[3 * r + j] = (base ? base[3 * r + j] : 0.0) + 0.25 * acc[j];
    }
}

void launch_spheres(const DeviceView& v, double dt, int implicit)
{
    if (v.nsph <= 0)
        return;
    if (v.ncand > 0)
        hipLaunchKernelGGL(k_sphere_contacts, dim3(xcd_grid((v.ncand + NT - 1) / NT)), dim3(NT), 0, v.stream, v.ncand,
                           v.cand, v.tact, v.tets, v.x, v.v, v.nsph, v.sph, dt, implicit, v.ck, v.cg);
    if (v.nnzu > 0)
        hipLaunchKernelGGL(k_sphere_slots, dim3(xcd_grid((v.nnzu * kSphereSlotLanes + NT - 1) / NT)), dim3(NT), 0,
                           v.stream, v.nnzu, v.csStart, v.csList, v.ck, v.nconn > 0 ? v.cslot : nullptr, v.sslot);
    if (v.nd > 0)
        hipLaunchKernelGGL(k_sphere_rows,
                           dim3(xcd_grid((static_cast<long long>(v.nd) * kSphereRowLanes + NT - 1) / NT)), dim3(NT),
                           0, v.stream, v.nd, v.crStart, v.crList, v.cg, v.nconn > 0 ? v.crhs : nullptr, v.srhs);
}

// A tet switched off in place: its element outputs that the gather lists
// still reach (basis planes, its corner entries, its reverse-mode planes)
// become zeros, which stay since the element kernels skip tets not live.
__global__ __launch_bounds__(NT) void k_zero_tets(int n, const int* __restrict__ ids, int T,
                                                  const int* __restrict__ cpos, const int* __restrict__ apos,
                                                  double* __restrict__ basis, double* __restrict__ corner,
                                                  double* __restrict__ adj)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n)
        return;
    const int t = ids[i];
    const size_t TT = static_cast<size_t>(T);
    for (int k = 0; k < EB_COUNT; ++k)
        basis[static_cast<size_t>(t) * EB_COUNT + k] = 0.0;
    for (int a = 0; a < 4; ++a)
    {
        const int pos = cpos[a * TT + t];
        if (pos >= 0)
            for (int e = 0; e < N_CORNER; ++e)
                corner[static_cast<size_t>(pos) * N_CORNER + e] = 0.0;
    }
    if (adj)
        for (int a = 0; a < 4; ++a)
        {
            const int pos = apos[a * TT + t];
            if (pos >= 0)
                for (int e = 0; e < N_ADJ; ++e)
                    adj[static_cast<size_t>(pos) * N_ADJ + e] = 0.0;
        }
}

// live[ids[i]] = liveVal[i], tact[ids[i]] = actVal[i].
__global__ __launch_bounds__(NT) void k_tet_flags(int n, const int* __restrict__ ids,
                                                  const unsigned char* __restrict__ liveVal,
                                                  const unsigned char* __restrict__ actVal,
                                                  unsigned char* __restrict__ live, unsigned char* __restrict__ tact)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n)
    {
        live[ids[i]] = liveVal[i];
        tact[ids[i]] = actVal[i];
    }
}

void launch_zero_tets(const DeviceView& v, int n, const int* ids)
{
    if (n > 0)
        hipLaunchKernelGGL(k_zero_tets, dim3((n + NT - 1) / NT), dim3(NT), 0, v.stream, n, ids, v.T, v.cpos, v.apos,
                           v.basis, v.corner, v.adj);
}

void launch_tet_flags(hipStream_t stream, int n, const int* ids, const unsigned char* liveVal,
                      const unsigned char* actVal, unsigned char* live, unsigned char* tact)
{
    if (n > 0)
        hipLaunchKernelGGL(k_tet_flags, dim3((n + NT - 1) / NT), dim3(NT), 0, stream, n, ids, liveVal, actVal, live,
                           tact);
}

void launch_slot_scatter(hipStream_t stream, int n, const int* idx, const double* val, double* cslot)
{
    if (n > 0)
        hipLaunchKernelGGL(k_slot_scatter, dim3((n + NT - 1) / NT), dim3(NT), 0, stream, n, idx, val, cslot);
}

void launch_zero_rows(hipStream_t stream, int n, const int* rows, double* crhs)
{
    if (n > 0)
        hipLaunchKernelGGL(k_zero_rows, dim3((n + NT - 1) / NT), dim3(NT), 0, stream, n, rows, crhs);
}

void launch_export_frame(hipStream_t stream, long long n3, const double* x, const double* v, float* out)
{
    if (n3 <= 0)
        return;
    const long long work = 2 * ((n3 + 1) / 2);
    const long long blocks = (work + NT - 1) / NT;
    const long long grid = blocks < 8192 ? blocks : 8192;
    hipLaunchKernelGGL(k_export_frame, dim3(static_cast<unsigned>(grid)), dim3(NT), 0, stream, n3, x, v, out);
}

void launch_stiff_records(const DeviceView& v)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_stiff_records, dim3((v.T + NT - 1) / NT), dim3(NT), 0, v.stream, v.T, v.prm, v.kdat);
}

void launch_gdot(const DeviceView& v, double* prm)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_gdot, dim3((v.T + NT - 1) / NT), dim3(NT), 0, v.stream, v.T, prm);
}

void launch_element_records(const DeviceView& v, double* out)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_element_records, dim3((v.T + NT - 1) / NT), dim3(NT), 0, v.stream, v.T, v.tets, v.live,
                           v.fem, v.prm, v.kdat, v.x, v.v, out);
}

template void launch_assemble<double>(const DeviceView&, double);
template void launch_assemble<float>(const DeviceView&, double);
template void launch_rhs<double>(const DeviceView&, double, double, double, double);
template void launch_rhs<float>(const DeviceView&, double, double, double, double);

}  // namespace pk
