// Driver for the compiled reference (oracle/_ref/libphysik_ref.so).
//
// TEST INFRASTRUCTURE ONLY. It is linked together with the reference's own
// source files (compiled in place from the reference tree by oracle/Makefile)
// and exposes a small C ABI so that tests/golden/make_golden.py can record
// fixtures and bench.py can time the reference step on the host cores. The
// product library never loads it.
//
// Everything below drives the reference through its public C ABI
// (include/PhysiK/API/PhysiKAPI.h:55-398), except two read-outs of internal
// state that the reference keeps private: the implicit system assembled by
// World::PrecomputeSolve (World.cpp:625-630, SolverData.cpp:126-465) and the
// per-tet element contributions (FEMModel.cpp:614-679). Those are reached by
// building the same objects the reference builds.

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

// Read access to World's and SolverData's private members (layout is not
// affected by access specifiers in g++).
#define private public
#include "PhysiK/Core/Solvers/SolverData.h"
#include "PhysiK/Core/World/World.h"
#undef private

#include "PhysiK/API/PhysiKAPI.h"
#include "PhysiK/Components/TetMeshPhysicsComponent.h"
#include "PhysiK/Core/Physics/FEM/FEMModel.h"
#include "PhysiK/Core/Solvers/Linear/ConjugateGradientSolver.h"
#include "PhysiK/Core/Solvers/Linear/LinearSolver.h"
#include "PhysiK/Math/SparseBlockMatrix.h"

namespace
{
struct RefHandle
{
    PhysiK::WorldHandle world = nullptr;
    PhysiK::ComponentHandle comp;
    int nodeCount = 0;
    int tetCount = 0;
    std::unique_ptr<PhysiK::SolverData> sd;
    std::vector<PhysiK::ComponentHandle> spheres;
};

PhysiK::World* W(RefHandle* h) { return static_cast<PhysiK::World*>(h->world); }

PhysiK::TetMeshPhysicsComponent* C(RefHandle* h)
{
    return dynamic_cast<PhysiK::TetMeshPhysicsComponent*>(W(h)->GetComponent(h->comp));
}

void put_rowmajor(float* out, const PhysiK::Mat3& m)
{
    const float* c0 = &m.columns[0].x;
    const float* c1 = &m.columns[1].x;
    const float* c2 = &m.columns[2].x;
    for (int i = 0; i < 3; ++i)
    {
        out[3 * i + 0] = c0[i];
        out[3 * i + 1] = c1[i];
        out[3 * i + 2] = c2[i];
    }
}

PhysiK::Mat3 get_rowmajor(const float* in)
{
    return PhysiK::Mat3::FromColumns(PhysiK::Vec3{in[0], in[3], in[6]}, PhysiK::Vec3{in[1], in[4], in[7]},
                                     PhysiK::Vec3{in[2], in[5], in[8]});
}
}  // namespace

extern "C"
{
// One world with one tet-mesh physics component (PHYSIK_GenerateTetMesh +
// PHYSIK_CreateTetMeshPhysicsComponent, as the reference tests do at
// PointConnectionTests.cpp:134-155).
void* ref_create(const float* rest, int n, const int* tets, int t, float density, float young, float poisson,
                 float damping)
{
    auto* h = new RefHandle();
    h->world = PHYSIK_CreateWorld();
    PhysiK::GeneratedTetMeshHandle mesh =
        PHYSIK_GenerateTetMesh(reinterpret_cast<const PhysiK::Vec3*>(rest), n, tets, t);
    PhysikMaterialDesc mat{density, young, poisson, damping};
    h->comp = PHYSIK_CreateTetMeshPhysicsComponent(h->world, mesh, &mat);
    h->nodeCount = PHYSIK_GetGeneratedTetMeshVertexCount(mesh);
    h->tetCount = PHYSIK_GetGeneratedTetMeshTetCount(mesh);
    PHYSIK_DestroyGeneratedTetMesh(mesh);
    return h;
}

void ref_destroy(void* p)
{
    auto* h = static_cast<RefHandle*>(p);
    h->sd.reset();
    PHYSIK_DestroyWorld(h->world);
    delete h;
}

int ref_counts(void* p, int* nodes, int* tets)
{
    auto* h = static_cast<RefHandle*>(p);
    *nodes = h->nodeCount;
    *tets = h->tetCount;
    return PHYSIK_IsComponentHandleValid(h->world, h->comp);
}

void ref_set_params(void* p, float gx, float gy, float gz, int substeps, float tol, int max_iter, int implicit)
{
    auto* h = static_cast<RefHandle*>(p);
    PHYSIK_SetGravity(h->world, gx, gy, gz);
    PHYSIK_SetSubstepCount(h->world, substeps);
    PHYSIK_SetConjugateGradientTolerance(h->world, tol);
    PHYSIK_SetConjugateGradientMaxIterations(h->world, max_iter);
    PHYSIK_SetSolverMode(h->world, implicit ? 1 : 0);
}

void ref_set_fixed(void* p, const unsigned char* fixed)
{
    auto* h = static_cast<RefHandle*>(p);
    for (int i = 0; i < h->nodeCount; ++i)
        PHYSIK_SetNodeFixed(h->world, PHYSIK_GetTetMeshGlobalNodeIndex(h->world, h->comp, i), fixed[i] ? 1 : 0);
}

void ref_set_tet_active(void* p, const unsigned char* active)
{
    auto* h = static_cast<RefHandle*>(p);
    for (int i = 0; i < h->tetCount; ++i)
        PHYSIK_SetTetActive(h->world, h->comp, i, active[i] ? 1 : 0);
}

// Per-element Young's modulus (reachable only from C++: Tet::youngModulus +
// RebuildTetFemCache, TetMeshPhysicsComponent.h:35-64).
void ref_set_young(void* p, const float* young)
{
    auto* h = static_cast<RefHandle*>(p);
    PhysiK::TetMeshPhysicsComponent* c = C(h);
    for (int i = 0; i < h->tetCount; ++i)
        c->tets[i].youngModulus = young[i];
    c->RebuildTetFemCache();
}

// Position first (it zeroes the velocity, World.cpp:394-401), then velocity.
void ref_set_state(void* p, const float* x, const float* v)
{
    auto* h = static_cast<RefHandle*>(p);
    for (int i = 0; i < h->nodeCount; ++i)
    {
        const int g = PHYSIK_GetTetMeshGlobalNodeIndex(h->world, h->comp, i);
        if (x)
            PHYSIK_SetNodePosition(h->world, g, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        if (v)
            PHYSIK_SetNodeVelocity(h->world, g, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
}

void ref_get_state(void* p, float* x, float* v)
{
    auto* h = static_cast<RefHandle*>(p);
    for (int i = 0; i < h->nodeCount; ++i)
    {
        const int g = PHYSIK_GetTetMeshGlobalNodeIndex(h->world, h->comp, i);
        if (x)
            PHYSIK_GetNodePosition(h->world, g, x + 3 * i, x + 3 * i + 1, x + 3 * i + 2);
        if (v)
            PHYSIK_GetNodeVelocity(h->world, g, v + 3 * i, v + 3 * i + 1, v + 3 * i + 2);
    }
}

// PHYSIK_AddPointConnection with the component's local node ids mapped to
// world nodes (PHYSIK_GetTetMeshGlobalNodeIndex).
void ref_add_point_connection(void* p, const int* nodes, const float* w, const float* target, float stiffness,
                              float damping)
{
    auto* h = static_cast<RefHandle*>(p);
    int g[4];
    for (int a = 0; a < 4; ++a)
        g[a] = PHYSIK_GetTetMeshGlobalNodeIndex(h->world, h->comp, nodes[a]);
    PHYSIK_AddPointConnection(h->world, g[0], g[1], g[2], g[3], w[0], w[1], w[2], w[3], target[0], target[1],
                              target[2], stiffness, damping);
}

// PHYSIK_CreateCollisionSphereComponent + PHYSIK_SetCollisionSphereConnectionSettings;
// returns the sphere's index for the calls below.
int ref_add_sphere(void* p, float x, float y, float z, float radius, float stiffness, float damping)
{
    auto* h = static_cast<RefHandle*>(p);
    const PhysiK::ComponentHandle c = PHYSIK_CreateCollisionSphereComponent(h->world, x, y, z, radius);
    PHYSIK_SetCollisionSphereConnectionSettings(h->world, c, stiffness, damping);
    h->spheres.push_back(c);
    return static_cast<int>(h->spheres.size()) - 1;
}

// PHYSIK_GetCollisionSphereConnectionSettings
void ref_sphere_settings(void* p, int sphere, float* stiffness, float* damping)
{
    auto* h = static_cast<RefHandle*>(p);
    PHYSIK_GetCollisionSphereConnectionSettings(h->world, h->spheres[sphere], stiffness, damping);
}

// PHYSIK_SetCollisionComponentKinematicTarget (taken at the next PHYSIK_Step)
void ref_move_sphere(void* p, int sphere, float x, float y, float z)
{
    auto* h = static_cast<RefHandle*>(p);
    PHYSIK_SetCollisionComponentKinematicTarget(h->world, h->spheres[sphere], x, y, z);
}

// PHYSIK_GetCollisionSphereOverlaps: per overlap 8 ints (geometry type,
// primitive, nodes 0-3, node mask, node count) and 5 floats (centre, radius,
// minimum node distance). Returns the count written.
int ref_sphere_overlaps(void* p, int sphere, int max_overlaps, int* ints, float* floats)
{
    auto* h = static_cast<RefHandle*>(p);
    std::vector<PhysikCollisionSphereOverlap> o(static_cast<size_t>(std::max(1, max_overlaps)));
    const int n = PHYSIK_GetCollisionSphereOverlaps(h->world, h->spheres[sphere], o.data(), max_overlaps);
    for (int i = 0; i < n; ++i)
    {
        const PhysikCollisionSphereOverlap& e = o[i];
        const int iv[8] = {e.geometryType, e.primitiveIndex, e.node0, e.node1, e.node2, e.node3,
                           e.overlappedNodeMask, e.overlappedNodeCount};
        const float fv[5] = {e.sphereCenterX, e.sphereCenterY, e.sphereCenterZ, e.sphereRadius, e.minDistance};
        std::copy(iv, iv + 8, ints + 8 * i);
        std::copy(fv, fv + 5, floats + 5 * i);
    }
    return n;
}

int ref_sphere_overlap_count(void* p, int sphere)
{
    auto* h = static_cast<RefHandle*>(p);
    return PHYSIK_GetCollisionSphereOverlapCount(h->world, h->spheres[sphere]);
}

int ref_point_connection_count(void* p)
{
    return PHYSIK_GetPointConnectionCount(static_cast<RefHandle*>(p)->world);
}

// PHYSIK_Step; returns 1 when the reference throws (World.cpp:576).
int ref_step(void* p, float dt)
{
    auto* h = static_cast<RefHandle*>(p);
    try
    {
        PHYSIK_Step(h->world, dt);
    }
    catch (const std::exception&)
    {
        return 1;
    }
    return 0;
}

void ref_get_cg(void* p, int* iterations, float* residual, int* converged)
{
    auto* h = static_cast<RefHandle*>(p);
    *iterations = PHYSIK_GetLastConjugateGradientIterations(h->world);
    *residual = PHYSIK_GetLastConjugateGradientResidualNorm(h->world);
    *converged = PHYSIK_DidLastConjugateGradientSolveConverge(h->world);
}

// World::PrecomputeSolve on a fresh SolverData (World.cpp:625-630).
int ref_assemble(void* p, float dt, int* nd, int* nnzb_upper)
{
    auto* h = static_cast<RefHandle*>(p);
    h->sd.reset(new PhysiK::SolverData());
    try
    {
        W(h)->PrecomputeSolve(*h->sd, dt);
    }
    catch (const std::exception&)
    {
        return 1;
    }
    *nd = h->sd->dynamicBlockCount;
    *nnzb_upper = static_cast<int>(h->sd->matrix.colIndex.size());
    return 0;
}

void ref_get_system(void* p, int* dyn, int* row_start, int* col_index, float* vals, float* rhs, float* mass)
{
    auto* h = static_cast<RefHandle*>(p);
    const PhysiK::SolverData& sd = *h->sd;
    for (int i = 0; i < h->nodeCount; ++i)
    {
        const int g = PHYSIK_GetTetMeshGlobalNodeIndex(h->world, h->comp, i);
        if (dyn)
            dyn[i] = sd.GetDynamicBlockForNode(g);
        if (mass)
            mass[i] = sd.GetAssembledMassForNode(g);
    }
    for (size_t i = 0; i < sd.matrix.rowStart.size(); ++i)
        if (row_start)
            row_start[i] = sd.matrix.rowStart[i];
    for (size_t k = 0; k < sd.matrix.colIndex.size(); ++k)
    {
        if (col_index)
            col_index[k] = sd.matrix.colIndex[k];
        if (vals)
            put_rowmajor(vals + 9 * k, sd.matrix.values[k]);
    }
    for (size_t i = 0; i < sd.rhs.size(); ++i)
        if (rhs)
        {
            rhs[3 * i + 0] = sd.rhs[i].x;
            rhs[3 * i + 1] = sd.rhs[i].y;
            rhs[3 * i + 2] = sd.rhs[i].z;
        }
}

// Element -> slot map through the reference's own hash lookup
// (SparseBlockMatrix::FindBlockIndex, SparseBlockMatrix.cpp:270-282).
void ref_get_scatter(void* p, int* slots)
{
    auto* h = static_cast<RefHandle*>(p);
    const PhysiK::SolverData& sd = *h->sd;
    const PhysiK::TetMeshPhysicsComponent* c = C(h);
    for (int t = 0; t < h->tetCount; ++t)
    {
        const PhysiK::Tet& tet = c->tets[t];
        const int n[4] = {tet.node0, tet.node1, tet.node2, tet.node3};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
            {
                const int ra = sd.GetDynamicBlockForNode(c->GetGlobalNodeIndex(n[a]));
                const int rb = sd.GetDynamicBlockForNode(c->GetGlobalNodeIndex(n[b]));
                slots[16 * t + 4 * a + b] = (ra < 0 || rb < 0) ? -1 : sd.matrix.FindBlockIndex(ra, rb);
            }
    }
}

// Element contributions for the current state (FEMModel::ComputeForces on one
// tet at a time, using the component's synced positions and velocities).
void ref_get_elements(void* p, unsigned char* valid, float* f, float* K)
{
    auto* h = static_cast<RefHandle*>(p);
    PhysiK::TetMeshPhysicsComponent* c = C(h);
    c->SyncCurrentPositionsFromWorld(*W(h));
    for (int t = 0; t < h->tetCount; ++t)
    {
        std::vector<PhysiK::TetElementContribution> out;
        PhysiK::FEMModel::ComputeForces(c->GetFemModel(), {c->tets[t]}, {c->tetFemCache[t]},
                                        c->nodePositions, c->nodeVelocities, out);
        valid[t] = out.empty() ? 0 : 1;
        if (out.empty())
            continue;
        for (int a = 0; a < 4; ++a)
        {
            f[12 * t + 3 * a + 0] = out[0].forces[a].x;
            f[12 * t + 3 * a + 1] = out[0].forces[a].y;
            f[12 * t + 3 * a + 2] = out[0].forces[a].z;
            for (int b = 0; b < 4; ++b)
                put_rowmajor(K + 144 * t + 9 * (4 * a + b), out[0].stiffness[a][b]);
        }
    }
}

// Rest data as initialised by FEMModel::InitializeTetRestData (FEMModel.cpp:504-548).
void ref_get_rest(void* p, float* vol, float* dminv, float* grad)
{
    auto* h = static_cast<RefHandle*>(p);
    const PhysiK::TetMeshPhysicsComponent* c = C(h);
    for (int t = 0; t < h->tetCount; ++t)
    {
        const PhysiK::Tet& tet = c->tets[t];
        vol[t] = tet.restVolume;
        put_rowmajor(dminv + 9 * t, tet.restDmInverse);
        for (int a = 0; a < 4; ++a)
        {
            grad[12 * t + 3 * a + 0] = tet.shapeFunctionGradients[a].x;
            grad[12 * t + 3 * a + 1] = tet.shapeFunctionGradients[a].y;
            grad[12 * t + 3 * a + 2] = tet.shapeFunctionGradients[a].z;
        }
    }
}

// TetMeshPhysicsComponentDesc::femModel of the driver's component (the C ABI
// has no setter; FemModel, FEMModel.h:40-45).
void ref_set_fem_model(void* p, int model)
{
    auto* h = static_cast<RefHandle*>(p);
    C(h)->SetFemModel(static_cast<PhysiK::FemModel>(model));
}

// Symmetric-upper BSR SpMV (SparseBlockMatrix::Multiply, :209-268).
void ref_spmv(int n, const int* row_start, const int* col_index, const float* vals, const float* x, float* y)
{
    std::vector<std::pair<int, int>> coords;
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
            coords.push_back({r, col_index[k]});
    PhysiK::SparseBlockMatrix A;
    A.BuildPattern(n, coords);
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
            A.AddBlock(r, col_index[k], get_rowmajor(vals + 9 * k));
    std::vector<PhysiK::Vec3> xv(static_cast<size_t>(n)), yv;
    for (int i = 0; i < n; ++i)
        xv[i] = PhysiK::Vec3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    A.Multiply(xv, yv);
    for (int i = 0; i < n; ++i)
    {
        y[3 * i + 0] = yv[i].x;
        y[3 * i + 1] = yv[i].y;
        y[3 * i + 2] = yv[i].z;
    }
}

// The reference's LinearSolver plugin slot (GetCurrentLinearSolver().Solve,
// LinearSolver.cpp:93-136) on a symmetric-upper BSR. Returns the iterations,
// or -1 - iterations when the solver left the solution vector empty.
int ref_linear_solve(int n, const int* row_start, const int* col_index, const float* vals, const float* b,
                     int max_iterations, float tolerance, int use_jacobi, float* x, float* residual, int* converged)
{
    std::vector<std::pair<int, int>> coords;
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
            coords.push_back({r, col_index[k]});
    PhysiK::SparseBlockMatrix A;
    A.BuildPattern(n, coords);
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
            A.AddBlock(r, col_index[k], get_rowmajor(vals + 9 * k));
    std::vector<PhysiK::Vec3> bv(static_cast<size_t>(n)), xv;
    for (int i = 0; i < n; ++i)
        bv[i] = PhysiK::Vec3{b[3 * i], b[3 * i + 1], b[3 * i + 2]};
    PhysiK::LinearSolveSettings settings;
    settings.maxIterations = max_iterations;
    settings.tolerance = tolerance;
    settings.useJacobiPreconditioner = use_jacobi != 0;
    const PhysiK::LinearSolveResult res = PhysiK::GetCurrentLinearSolver().Solve(A, bv, xv, settings);
    for (size_t i = 0; i < xv.size() && i < static_cast<size_t>(n); ++i)
    {
        x[3 * i + 0] = xv[i].x;
        x[3 * i + 1] = xv[i].y;
        x[3 * i + 2] = xv[i].z;
    }
    *residual = res.residualNorm;
    *converged = res.converged ? 1 : 0;
    return static_cast<int>(xv.size()) == n ? res.iterations : -1 - res.iterations;
}
}
