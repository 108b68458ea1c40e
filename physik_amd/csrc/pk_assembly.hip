// Per-substep assembly kernels for gfx950: element, slot gather, right-hand
// side with the CG start, and the integrators.
//
//   k_element   one lane per tet: F = Ds Dm^-1, Newton polar R, corotated
//               forces and the 10 blocks R K_ab R^T, K v per corner
//               (FEMModel.cpp:614-679, :377-424, :336-375, :115-163)
//   k_assemble  one lane per upper BSR slot: sum of the slot's element blocks
//               in tet order through the precomputed element->slot lists (no
//               atomics), A = M + dt^2 K written to both triangles, 3x3
//               block-Jacobi inverse on diagonal slots
//               (SolverData.cpp:308-465, :260-306)
//   k_rhs       b = dt (f + m g) - dt^2 K v per dynamic row, then the CG start
//               x = 0, r = b, z = d = M^-1 r with the (r.z, r.r) partials
//               (ConjugateGradientSolver.cpp:151-164)
//   k_integrate all-or-nothing v += dv, x += dt v per group (World.cpp:644-695)
//   k_explicit  explicit Euler (World.cpp:697-725)
#include "pk_device.h"
#include "pk_fem.h"
#include "pk_topology.h"

namespace pk
{

__global__ void k_reset(GroupState* gs, int G, int* active, double tol, int maxIt)
{
    __shared__ int cnt;
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
    {
        GroupState& s = gs[g];
        s.delta[0] = s.delta[1] = s.target = s.residual = 0.0;
        s.tol = tol;
        s.maxIt = maxIt;
        s.iters = 0;
        s.stopAt = s.hasRows ? -1 : 0;
        s.stopInUpdate = 0;
        s.converged = 0;
        s.failed = 0;
        s.nonfinite = 0;
        if (s.hasRows)
            atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *active = cnt;
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
    double* out = elem + static_cast<size_t>(t) * E_STRIDE;

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
        vals[BsrStore<M>::at(F, full, e)] = static_cast<M>(a[e]);
    if (mirror >= 0)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                vals[BsrStore<M>::at(F, mirror, 3 * i + j)] = static_cast<M>(a[3 * j + i]);
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

// z = M^-1 r with the block-Jacobi inverse stored as 9 planes of nd values.
__device__ inline void apply_minv(const double* __restrict__ minv, size_t nd, int row, const double* r, double* z)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        z[i] = minv[(3 * i) * nd + row] * r[0] + minv[(3 * i + 1) * nd + row] * r[1] +
               minv[(3 * i + 2) * nd + row] * r[2];
}

template <class M>
__global__ __launch_bounds__(NT) void k_rhs(TileView tv, const int* __restrict__ nstart,
                                            const int* __restrict__ nlist, const double* __restrict__ elem,
                                            const double* __restrict__ mrow, double gx, double gy, double gz,
                                            double dt, double dt2, const double* __restrict__ minv, int nd,
                                            double* __restrict__ xs, double* __restrict__ rv, M* __restrict__ zv,
                                            M* __restrict__ d0, Partial* __restrict__ part)
{
    const int t = tile_of(tv);
    if (t < 0)
        return;
    const RowTile tl = tv.tiles[t];
    double pa = 0.0, pb = 0.0;
    int flag = 0;
    for (int row = tl.rowBegin + threadIdx.x; row < tl.rowEnd; row += NT)
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
        apply_minv(minv, static_cast<size_t>(nd), row, b, z);
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag |= 2;
        M zs[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            zs[i] = static_cast<M>(z[i]);
            xs[3 * row + i] = 0.0;
            rv[3 * row + i] = b[i];
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
    if (s.failed || (commit && s.nonfinite))
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
                                                 const int* __restrict__ nstart, const int* __restrict__ nlist,
                                                 const double* __restrict__ elem, const double* __restrict__ mrow,
                                                 double gx, double gy, double gz, double dt, double* __restrict__ x,
                                                 double* __restrict__ v)
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
}

static inline unsigned grid_of(const TileView& tv) { return 8u * static_cast<unsigned>(tv.per); }

void launch_reset(const DeviceView& v, double tol, int maxIt)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(NT), 0, v.stream, v.gs, v.G, v.active, tol, maxIt);
}

void launch_element(const DeviceView& v)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_element, dim3((v.T + NT - 1) / NT), dim3(NT), 0, v.stream, v.T, v.tets, v.live, v.prm,
                           v.x, v.v, v.elem);
}

template <class M>
void launch_assemble(const DeviceView& v, double dt)
{
    const int n = static_cast<int>(v.nnzu);
    if (n > 0)
        hipLaunchKernelGGL((k_assemble<M>), dim3((n + NT - 1) / NT), dim3(NT), 0, v.stream, n, v.cstart, v.clist,
                           v.uFull, v.uMirror, v.slotRow, v.rowGroup, v.elem, v.mrow, dt * dt,
                           static_cast<M*>(v.vals), v.nnzf, v.minv, v.nd, v.gs);
}

template <class M>
void launch_rhs(const DeviceView& v, double gx, double gy, double gz, double dt)
{
    if (v.vec.count > 0)
        hipLaunchKernelGGL((k_rhs<M>), dim3(grid_of(v.vec)), dim3(NT), 0, v.stream, v.vec, v.nstart, v.nlist,
                           v.elem, v.mrow, gx, gy, gz, dt, dt * dt, v.minv, v.nd, v.xs, v.r, static_cast<M*>(v.z),
                           static_cast<M*>(v.d[0]), v.pb[1]);
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
                           v.nlist, v.elem, v.mrow, gx, gy, gz, dt, v.x, v.v);
}

template void launch_assemble<double>(const DeviceView&, double);
template void launch_assemble<float>(const DeviceView&, double);
template void launch_rhs<double>(const DeviceView&, double, double, double, double);
template void launch_rhs<float>(const DeviceView&, double, double, double, double);

}  // namespace pk
