// Reverse-mode kernels of the implicit-Euler substep for gfx950. The
// reference has no adjoint; the math below is the exact derivative of the
// forward map the forward kernels implement (FEMModel.cpp:377-424 forces,
// :336-375 blocks, SolverData.cpp:308-465 system, World.cpp:644-695 update).
//
// Substep (x, v) -> (x', v'): A(x) dv = b(x, v), v' = v + dv, x' = x + dt v'
// on dynamic nodes, with A = M + dt^2 K(x), b = dt f(x, v) - dt^2 K(x) v and
// per tet, for S(H) = lambda tr(H) I + mu (H + H^T), W = sum_b w_b g_b^T:
//   f_a = -V R S(R^T F - I) g_a - c v_a,   (K_e w)_a = V R S(R^T W) g_a.
// Given (xb', vb'):  u = vb' + dt xb' on dynamic rows, A lam = u, then
//   xb = xb' + dPhi/dx,  vb = (u | vb' off the dynamic rows) + dPhi/dv,
//   Phi_e = -dt V <L, R S(R^T F - I)> - dt^2 V <L, R S(R^T W')> - dt c lam.v,
//   L = sum_a lam_a g_a^T, W' = sum_a v'_a g_a^T (v' = the substep's result),
// and dPhi/dlogE_e is the elastic part of Phi_e (lambda, mu scale with E).
// R = polar(F) is differentiated as the exact polar factor: dR = R Omega,
// S Omega + Omega S = R^T dF - dF^T R with S = R^T F.
//
//   k_record_ok     per body: did the forward substep integrate
//   k_adj_rhs       u on dynamic rows and the CG start for A lam = u
//   k_adj_element   one lane per tet: dPhi/dlogE (accumulated in place) and
//                   the tet's 4 corner terms of dPhi/dx and dPhi/dv
//   k_adj_conn      one lane per point connection: its terms of dPhi/dx and
//                   dPhi/dv before the corner weight
//   k_adj_sphere    one lane per candidate tet: its sphere contacts' terms
//   k_adj_nodes     one lane per node: xb, vb from the corner terms of its
//                   tets in tet order, then of its connections in connection
//                   order and its candidate tets in candidate order (no
//                   atomics)
//
// Contacts. A point connection (PointConnection.cpp:54-130) on nodes n_a with
// weights w_a pushes w_a F, F = k G(p) - c u (p = sum w_a x_a, u = sum w_a
// v_a, G = t - p), and the constant blocks k w_a w_b I; with lam_c = sum_a
// w_a lam_a it adds dPhi/dx_a = dt k w_a (dG/dp)^T lam_c and dPhi/dv_a =
// -(dt c + dt^2 k) w_a lam_c. A connection's target is fixed (dG/dp = -I);
// a sphere contact (CollisionSphereComponent.cpp:112-193) targets
// t = p + n (r - |d|), d = p - o, n = d / |d| (or (0, 0, 1) within 1e-5 of
// the centre o), so G = d (r / |d| - 1) and dG/dp = (r / |d| - 1) I -
// r d d^T / |d|^3 (near the centre G = (0, 0, r - |d|), dG/dp = -e_z d^T / |d|).
// Which tets touch a sphere is piecewise constant in x: no derivative.
#include "pk_device.h"
#include "pk_fem.h"
#include "pk_topology.h"

namespace pk
{

__device__ inline void mm3(const double* a, const double* b, double* o)  // o = a b
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            o[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

__device__ inline void mtm3(const double* a, const double* b, double* o)  // o = a^T b
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            o[3 * i + j] = a[i] * b[j] + a[3 + i] * b[3 + j] + a[6 + i] * b[6 + j];
}

// lambda tr(H) I + mu (H + H^T)
__device__ inline void stress3(const double* H, double lam, double mu, double* S)
{
    const double tr = H[0] + H[4] + H[8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            S[3 * i + j] = mu * (H[3 * i + j] + H[3 * j + i]) + (i == j ? lam * tr : 0.0);
}

__device__ inline double frob3(const double* a, const double* b)
{
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 9; ++e)
        s += a[e] * b[e];
    return s;
}

// sum_a p_a g_a^T
__device__ inline void outer4(const double p[4][3], const double g[4][3], double* o)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            o[3 * i + j] = p[0][i] * g[0][j] + p[1][i] * g[1][j] + p[2][i] * g[2][j] + p[3][i] * g[3][j];
}

// dL/dF of R = polar(F) given G = dL/dR (kind as polar() reports it).
__device__ inline void polar_vjp(const double* F, const double* R, int kind, const double* G, double* Fb)
{
    if (kind == 2)
    {
#pragma unroll
        for (int e = 0; e < 9; ++e)
            Fb[e] = 0.0;
        return;
    }
    double Q[9], GQ[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
    {
        // R = Q diag(1, 1, -1) when the column was negated
        const double sgn = (kind == 1 && e % 3 == 2) ? -1.0 : 1.0;
        Q[e] = sgn * R[e];
        GQ[e] = sgn * G[e];
    }
    double QtF[9], QtG[9];
    mtm3(Q, F, QtF);
    mtm3(Q, GQ, QtG);
    double S[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            S[3 * i + j] = 0.5 * (QtF[3 * i + j] + QtF[3 * j + i]);
    // [y]x = skew(Q^T G);  S [z]x + [z]x S = [(tr S I - S) z]x
    const double y0 = 0.5 * (QtG[7] - QtG[5]), y1 = 0.5 * (QtG[2] - QtG[6]), y2 = 0.5 * (QtG[3] - QtG[1]);
    const double tr = S[0] + S[4] + S[8];
    double Mx[9];
#pragma unroll
    for (int e = 0; e < 9; ++e)
        Mx[e] = -S[e];
    Mx[0] += tr;
    Mx[4] += tr;
    Mx[8] += tr;
    const double d = det3(Mx);
    // Cramer's rule, columns of Mx replaced by y
    const double c0[9] = {y0, Mx[1], Mx[2], y1, Mx[4], Mx[5], y2, Mx[7], Mx[8]};
    const double c1[9] = {Mx[0], y0, Mx[2], Mx[3], y1, Mx[5], Mx[6], y2, Mx[8]};
    const double c2[9] = {Mx[0], Mx[1], y0, Mx[3], Mx[4], y1, Mx[6], Mx[7], y2};
    const double z0 = det3(c0) / d, z1 = det3(c1) / d, z2 = det3(c2) / d;
    const double Z[9] = {0.0, -z2, z1, z2, 0.0, -z0, -z1, z0, 0.0};
    double QZ[9];
    mm3(Q, Z, QZ);
#pragma unroll
    for (int e = 0; e < 9; ++e)
        Fb[e] = 2.0 * QZ[e];
}

__global__ void k_record_ok(const GroupState* __restrict__ gs, int G, int* __restrict__ ok)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        ok[g] = (gs[g].hasRows && !gs[g].failed && !gs[g].nonfinite) ? 1 : 0;
}

// u = vb + dt xb on dynamic rows (zero in bodies whose substep did not
// integrate), then the CG start x = 0, r = u, z = d = M^-1 r and the
// (r.z, r.r, flags) partials, as the forward right-hand side does.
template <class M>
__global__ __launch_bounds__(NT) void k_adj_rhs(TileView tv, const int* __restrict__ rowNode,
                                                const int* __restrict__ ok, const double* __restrict__ xbar,
                                                const double* __restrict__ vbar, double dt,
                                                const M* __restrict__ minv, int nd, double* __restrict__ xs,
                                                double* __restrict__ rv, M* __restrict__ zv, M* __restrict__ d0,
                                                Partial* __restrict__ part)
{
    const int t = tile_of(tv);
    if (t < 0)
        return;
    const RowTile tl = tv.tiles[t];
    const bool live = ok[tl.group] != 0;
    const size_t N = static_cast<size_t>(nd);
    double pa = 0.0, pb = 0.0;
    int flag = 0;
    for (int row = tl.rowBegin + threadIdx.x; row < tl.rowEnd; row += NT)
    {
        const int n = rowNode[row];
        double b[3], z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            b[i] = live ? vbar[3 * n + i] + xbar[3 * n + i] * dt : 0.0;
        if (!(fin(b[0]) && fin(b[1]) && fin(b[2])))
            flag |= 1;
        minv_apply(minv, N, row, b, z);
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

// Corner terms (dPhi/dx_a, then dPhi/dv_a) go to the tet's four entries of
// the node gather lists, at apos[a * T + t] (N_ADJ doubles each).
__global__ __launch_bounds__(NT) void k_adj_element(int T, const int4* __restrict__ tets,
                                                    const unsigned char* __restrict__ live,
                                                    const double* __restrict__ prm,
                                                    const int* __restrict__ nodeGroup, const int* __restrict__ ok,
                                                    const int* __restrict__ nodeRow, const double* __restrict__ x,
                                                    const double* __restrict__ vnext, const double* __restrict__ lam,
                                                    double dt, const int* __restrict__ apos,
                                                    double* __restrict__ elem, double* __restrict__ gradLogE)
{
    const int t = xcd_block() * NT + static_cast<int>(threadIdx.x);
    if (t >= T || !live[t])
        return;
    const int4 n4 = tets[t];
    const size_t TT = static_cast<size_t>(T);
    int pos[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        pos[a] = apos[a * TT + t];
    // corner a's entry, value e (0..2 dPhi/dx_a, 3..5 dPhi/dv_a); every
    // corner of a live tet has one
    auto out = [&](int a, int e) -> double& { return elem[static_cast<size_t>(pos[a]) * N_ADJ + e]; };
    if (!ok[nodeGroup[n4.x]])
        return;  // the body's substep was identity: its nodes are not gathered
    const int ids[4] = {n4.x, n4.y, n4.z, n4.w};
    double xs[4][3], ws[4][3], ls[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
        const int row = nodeRow[ids[a]];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xs[a][i] = x[3 * ids[a] + i];
            ws[a][i] = vnext[3 * ids[a] + i];
            ls[a][i] = row >= 0 ? lam[3 * row + i] : 0.0;
        }
    }
    double Dinv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        Dinv[k] = prm[(P_DMINV + k) * TT + t];
    const double vol = prm[P_VOL * TT + t];
    const double la = prm[P_LAMBDA * TT + t];
    const double mu = prm[P_MU * TT + t];
    const double damp = prm[P_DAMP * TT + t];

    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = (xs[1][i] - xs[0][i]) * Dinv[j] + (xs[2][i] - xs[0][i]) * Dinv[3 + j] +
                           (xs[3][i] - xs[0][i]) * Dinv[6 + j];
    if (!fin9(F))
    {
        // the forward skipped the tet: no terms
        for (int a = 0; a < 4; ++a)
            for (int e = 0; e < N_ADJ; ++e)
                out(a, e) = 0.0;
        return;
    }
    double R[9];
    int kind = 2;
    polar(F, R, &kind);
    double g[4][3];
    shape_gradients(Dinv, g);
    double L[9], W[9];
    outer4(ls, g, L);
    outer4(ws, g, W);
    double H1[9], RtW[9], A[9];
    mtm3(R, F, H1);
    H1[0] -= 1.0;
    H1[4] -= 1.0;
    H1[8] -= 1.0;
    mtm3(R, W, RtW);
    mtm3(R, L, A);
    double S1[9], S2[9], SA[9];
    stress3(H1, la, mu, S1);
    stress3(RtW, la, mu, S2);
    stress3(A, la, mu, SA);
    const double dt2 = dt * dt;
    gradLogE[t] += -dt * vol * frob3(A, S1) - dt2 * vol * frob3(A, S2);

    double t1[9], t2[9], t3[9], t4[9], GR[9];
    mm3(L, S1, t1);
    mm3(F, SA, t2);
    mm3(L, S2, t3);
    mm3(W, SA, t4);
#pragma unroll
    for (int e = 0; e < 9; ++e)
        GR[e] = -dt * vol * (t1[e] + t2[e]) - dt2 * vol * (t3[e] + t4[e]);
    double P[9], Fb[9];
    mm3(R, SA, P);
    polar_vjp(F, R, kind, GR, Fb);
#pragma unroll
    for (int e = 0; e < 9; ++e)
        Fb[e] += -dt * vol * P[e];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            out(a, i) = Fb[3 * i] * g[a][0] + Fb[3 * i + 1] * g[a][1] + Fb[3 * i + 2] * g[a][2];
            out(a, 3 + i) =
                -dt2 * vol * (P[3 * i] * g[a][0] + P[3 * i + 1] * g[a][1] + P[3 * i + 2] * g[a][2]) -
                dt * damp * ls[a][i];
        }
}

__global__ __launch_bounds__(NT) void k_adj_nodes(int N, const int* __restrict__ astart,
                                                  const double* __restrict__ elem,
                                                  const int* __restrict__ nodeRow, const int* __restrict__ nodeGroup,
                                                  const int* __restrict__ ok, const int* __restrict__ connStart,
                                                  const int2* __restrict__ connList,
                                                  const double* __restrict__ connData,
                                                  const double* __restrict__ connAdj,
                                                  const int* __restrict__ candStart,
                                                  const int* __restrict__ candList,
                                                  const double* __restrict__ candAdj, double dt,
                                                  double* __restrict__ xbar, double* __restrict__ vbar)
{
    const int n = xcd_block() * NT + static_cast<int>(threadIdx.x);
    if (n >= N || !ok[nodeGroup[n]])
        return;
    double xb[3], vb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        xb[i] = xbar[3 * n + i];
        vb[i] = vbar[3 * n + i];
    }
    if (nodeRow[n] >= 0)
    {
        // v' = v + dv and x' = x + dt v': the velocity adjoint is u
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vb[i] = vb[i] + xb[i] * dt;
    }
    // the node's corner entries: one contiguous run, in tet order
    const int k1 = astart[n + 1];
    for (int k = astart[n]; k < k1; ++k)
    {
        const double* e = elem + static_cast<size_t>(k) * N_ADJ;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            xb[i] += e[i];
            vb[i] += e[3 + i];
        }
    }
    if (connStart)
        for (int k = connStart[n]; k < connStart[n + 1]; ++k)
        {
            const int2 en = connList[k];  // (connection, corner)
            const double wa = connData[9 * static_cast<size_t>(en.x) + en.y];
            const double* e = connAdj + 6 * static_cast<size_t>(en.x);
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                xb[i] += wa * e[i];
                vb[i] += wa * e[3 + i];
            }
        }
    if (candStart)
        for (int k = candStart[n]; k < candStart[n + 1]; ++k)
        {
            const double* e = candAdj + 6 * static_cast<size_t>(candList[k]);
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                xb[i] += e[i];
                vb[i] += e[3 + i];
            }
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        xbar[3 * n + i] = xb[i];
        vbar[3 * n + i] = vb[i];
    }
}

// lam at node n: the adjoint solution on its dynamic row, 0 elsewhere.
__device__ inline double lam_at(const int* __restrict__ nodeRow, const double* __restrict__ lam, int n, int i)
{
    const int row = nodeRow[n];
    return row >= 0 ? lam[3 * row + i] : 0.0;
}

// Per point connection: gx = -dt k lam_c and gv = -(dt c + dt^2 k) lam_c
// (the node gather scales them by w_a); zeros when the forward pushed no
// force (point or velocity not finite).
__global__ __launch_bounds__(NT) void k_adj_conn(int n, const int4* __restrict__ nodes,
                                                 const double* __restrict__ data, const double* __restrict__ x,
                                                 const double* __restrict__ v, const int* __restrict__ nodeRow,
                                                 const double* __restrict__ lam, double dt, double* __restrict__ out)
{
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= n)
        return;
    const int4 nd = nodes[c];
    const int ids[4] = {nd.x, nd.y, nd.z, nd.w};
    const double* d = data + 9 * static_cast<size_t>(c);
    double lc[3] = {0.0, 0.0, 0.0};
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        // the point and velocity as k_conn_eval forms them
        const double p = ((x[3 * nd.x + i] * d[0] + x[3 * nd.y + i] * d[1]) + x[3 * nd.z + i] * d[2]) +
                         x[3 * nd.w + i] * d[3];
        const double u = ((v[3 * nd.x + i] * d[0] + v[3 * nd.y + i] * d[1]) + v[3 * nd.z + i] * d[2]) +
                         v[3 * nd.w + i] * d[3];
        ok = ok && fin(p) && fin(u) && fin(d[4 + i]);
#pragma unroll
        for (int a = 0; a < 4; ++a)
            lc[i] += d[a] * lam_at(nodeRow, lam, ids[a], i);
    }
    double* o = out + 6 * static_cast<size_t>(c);
    const double k = d[7], cd = d[8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        o[i] = ok ? -dt * k * lc[i] : 0.0;
        o[3 + i] = ok ? -(dt * cd + dt * dt * k) * lc[i] : 0.0;
    }
}

// Per candidate tet: the sum over the spheres it touches (the contacts
// k_sphere_contacts makes, tested the same way) of dt k (dG/dp)^T lam_t and
// -(dt c + dt^2 k) lam_t, lam_t = (1/4) sum_a lam_a, times the corner weight
// 1/4: what each of its four nodes receives.
__global__ __launch_bounds__(NT) void k_adj_sphere(int n, const int* __restrict__ cand,
                                                   const unsigned char* __restrict__ tact,
                                                   const int4* __restrict__ tets, const double* __restrict__ x,
                                                   const double* __restrict__ v, int nsph,
                                                   const double* __restrict__ sph, const int* __restrict__ nodeRow,
                                                   const double* __restrict__ lam, double dt,
                                                   double* __restrict__ out)
{
    const int i = xcd_block() * NT + static_cast<int>(threadIdx.x);
    if (i >= n)
        return;
    double* o = out + 6 * static_cast<size_t>(i);
    const int tt = cand[i];
    if (!tact[tt])
    {
        for (int j = 0; j < 6; ++j)
            o[j] = 0.0;
        return;
    }
    const int4 nd = tets[tt];
    const int ids[4] = {nd.x, nd.y, nd.z, nd.w};
    const double w = 0.25;
    double p[3], u[3], lt[3];
    bool fin3 = true;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        p[j] = ((x[3 * nd.x + j] * w + x[3 * nd.y + j] * w) + x[3 * nd.z + j] * w) + x[3 * nd.w + j] * w;
        u[j] = ((v[3 * nd.x + j] * w + v[3 * nd.y + j] * w) + v[3 * nd.z + j] * w) + v[3 * nd.w + j] * w;
        fin3 = fin3 && fin(p[j]) && fin(u[j]);
        lt[j] = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
            lt[j] += w * lam_at(nodeRow, lam, ids[a], j);
    }
    double gx[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
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
        bool ok = fin3;
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double nj = unit ? d[j] / dist : (j == 2 ? 1.0 : 0.0);
            ok = ok && fin(p[j] + nj * depth);
        }
        if (!ok)
            continue;
        const double k = sp[4], c = sp[5];
        double jl[3];
        if (unit)
        {
            const double dl = d[0] * lt[0] + d[1] * lt[1] + d[2] * lt[2];
            const double f = r / dist - 1.0, g = r * dl / (dist * dist * dist);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                jl[j] = f * lt[j] - g * d[j];
        }
        else
        {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                jl[j] = dist > 0.0 ? -lt[2] * d[j] / dist : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            gx[j] += dt * k * jl[j];
            gv[j] += -(dt * c + dt2 * k) * lt[j];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        o[j] = w * gx[j];
        o[3 + j] = w * gv[j];
    }
}

void launch_adj_contacts(const DeviceView& v, double dt)
{
    if (v.nconn > 0)
        hipLaunchKernelGGL(k_adj_conn, dim3((v.nconn + NT - 1) / NT), dim3(NT), 0, v.stream, v.nconn, v.connNodes,
                           v.connData, v.x, v.v, v.nodeRow, v.xs, dt, v.connAdj);
    if (v.nsph > 0 && v.ncand > 0)
        hipLaunchKernelGGL(k_adj_sphere, dim3(xcd_grid((v.ncand + NT - 1) / NT)), dim3(NT), 0, v.stream, v.ncand,
                           v.cand, v.tact, v.tets, v.x, v.v, v.nsph, v.sph, v.nodeRow, v.xs, dt, v.candAdj);
}

void launch_record_ok(const DeviceView& v, int* ok)
{
    if (v.G > 0)
        hipLaunchKernelGGL(k_record_ok, dim3((v.G + NT - 1) / NT), dim3(NT), 0, v.stream, v.gs, v.G, ok);
}

template <class M>
void launch_adj_rhs(const DeviceView& v, const int* ok, const double* xbar, const double* vbar, double dt)
{
    if (v.vec.count == 0)
        return;
    hipLaunchKernelGGL((k_adj_rhs<M>), dim3(8u * static_cast<unsigned>(v.vec.per)), dim3(NT), 0, v.stream, v.vec,
                       v.rowNode, ok, xbar, vbar, dt, static_cast<const M*>(v.minv), v.nd, v.xs, v.r, static_cast<M*>(v.z),
                       static_cast<M*>(v.d[0]), v.pb[1]);
}

// The assembly's per-body rejection flags, saved after a replayed substep's
// assembly and put back before the reverse step reuses its matrix.
__global__ void k_save_failed(const GroupState* __restrict__ gs, int G, int* __restrict__ out)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        out[g] = gs[g].failed;
}

__global__ void k_restore_failed(GroupState* __restrict__ gs, int G, const int* __restrict__ in)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        gs[g].failed = in[g];
}

void launch_save_failed(const DeviceView& v, int* out)
{
    if (v.G > 0)
        hipLaunchKernelGGL(k_save_failed, dim3((v.G + NT - 1) / NT), dim3(NT), 0, v.stream, v.gs, v.G, out);
}

void launch_restore_failed(const DeviceView& v, const int* in)
{
    if (v.G > 0)
        hipLaunchKernelGGL(k_restore_failed, dim3((v.G + NT - 1) / NT), dim3(NT), 0, v.stream, v.gs, v.G, in);
}

void launch_adj_element(const DeviceView& v, const int* ok, const double* vnext, double dt, double* gradLogE)
{
    if (v.T > 0)
        hipLaunchKernelGGL(k_adj_element, dim3(xcd_grid((v.T + NT - 1) / NT)), dim3(NT), 0, v.stream, v.T, v.tets,
                           v.live,
                           v.prm, v.nodeGroup, ok, v.nodeRow, v.x, vnext, v.xs, dt, v.apos, v.adj, gradLogE);
}

void launch_adj_nodes(const DeviceView& v, const int* ok, double dt, double* xbar, double* vbar)
{
    if (v.N > 0)
        hipLaunchKernelGGL(k_adj_nodes, dim3(xcd_grid((v.N + NT - 1) / NT)), dim3(NT), 0, v.stream, v.N,
                           v.astart, v.adj, v.nodeRow, v.nodeGroup, ok,
                           v.nconn > 0 ? v.connNodeStart : nullptr, v.connNodeList, v.connData, v.connAdj,
                           (v.nsph > 0 && v.ncand > 0) ? v.candNodeStart : nullptr, v.candNodeList, v.candAdj, dt,
                           xbar, vbar);
}

template void launch_adj_rhs<double>(const DeviceView&, const int*, const double*, const double*, double);
template void launch_adj_rhs<float>(const DeviceView&, const int*, const double*, const double*, double);

}  // namespace pk
