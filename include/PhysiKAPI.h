/*
 * PhysiKAPI.h - drop-in PHYSIK_* C ABI for the soft-body implicit-Euler step
 * on MI355X.
 *
 * Every declaration in the first block has the signature, argument meaning
 * and error behaviour of the reference declaration cited beside it
 * (include/PhysiK/API/PhysiKAPI.h in the reference tree). A host that links
 * the reference for the physics subset can link libphysik_amd.so instead; the
 * implicit step then runs on the GPU through physik_amd.h.
 *
 * Semantics kept from the reference:
 *   - null or stale handles make a call a no-op that returns 0 / -1 / 0.0f /
 *     an invalid ComponentHandle (PhysiKAPI.cpp:310-1558);
 *   - setters clamp silently (World.cpp:157-201);
 *   - PHYSIK_Step throws std::runtime_error when a free active node has no
 *     mass (World.cpp:560-580); PHYSIK_AMD_TryStep reports the same condition
 *     as a status for callers that cannot catch C++ exceptions;
 *   - the default solver mode is explicit (World.h:133); mode 1 selects
 *     implicit Euler;
 *   - a failed solve leaves the state of that substep untouched.
 *
 * Not provided (out of scope: render and gameplay features): scripts,
 * surface extraction/visuals, visual meshes, tet mesh mappers.
 */
#ifndef PHYSIK_AMD_PHYSIKAPI_H
#define PHYSIK_AMD_PHYSIKAPI_H

#include <cstdint>

#define PHYSIK_API __attribute__((visibility("default")))

namespace PhysiK
{
/* Layout-compatible with include/PhysiK/Math/Vec3.h:7-12. */
struct Vec3
{
    float x = 0.0f;
    float y = 0.0f;
    float z = 0.0f;
};

/* include/PhysiK/API/Handles.h:7-19 */
using WorldHandle = void*;
using GeneratedTetMeshHandle = void*;

struct ComponentHandle
{
    std::uint32_t index = 0xFFFFFFFFu;
    std::uint32_t generation = 0u;

    bool IsValid() const { return index != 0xFFFFFFFFu && generation != 0u; }
};
}  // namespace PhysiK

/* include/PhysiK/API/PhysiKAPI.h:24-31 */
enum PhysikOverlapGeometryType
{
    PHYSIK_OverlapGeometry_Unknown = 0,
    PHYSIK_OverlapGeometry_Tetrahedron = 1,
    PHYSIK_OverlapGeometry_Triangle = 2,
    PHYSIK_OverlapGeometry_Sphere = 3,
    PHYSIK_OverlapGeometry_Node = 4
};

/* include/PhysiK/API/PhysiKAPI.h:33-53: one active tet of a tet-mesh physics
 * component with at least one node inside the sphere. */
struct PhysikCollisionSphereOverlap
{
    int geometryType;
    PhysiK::ComponentHandle component;
    int primitiveIndex;
    int node0;
    int node1;
    int node2;
    int node3;
    int overlappedNodeMask;
    int overlappedNodeCount;
    float sphereCenterX;
    float sphereCenterY;
    float sphereCenterZ;
    float sphereRadius;
    float minDistance;
};

/* include/PhysiK/API/PhysiKAPI.h:16-22 */
struct PhysikMaterialDesc
{
    float density;
    float youngModulus;
    float poissonRatio;
    float damping;
};

extern "C"
{
    /* ---- reference entry points (PhysiKAPI.h line of the declaration) ---- */
    PHYSIK_API PhysiK::WorldHandle PHYSIK_CreateWorld();                                   /* :57 */
    PHYSIK_API void PHYSIK_DestroyWorld(PhysiK::WorldHandle world);                        /* :58 */
    PHYSIK_API void PHYSIK_Step(PhysiK::WorldHandle world, float dt);                      /* :60 */
    PHYSIK_API void PHYSIK_SetSubstepCount(PhysiK::WorldHandle world, int substepCount);   /* :61 */
    PHYSIK_API void PHYSIK_SetSolverMode(PhysiK::WorldHandle world, int mode);             /* :62 */
    PHYSIK_API void PHYSIK_SetConjugateGradientTolerance(PhysiK::WorldHandle world, float tolerance);  /* :63 */
    PHYSIK_API float PHYSIK_GetConjugateGradientTolerance(PhysiK::WorldHandle world);      /* :66 */
    PHYSIK_API void PHYSIK_SetConjugateGradientMaxIterations(PhysiK::WorldHandle world, int maxIterations); /* :68 */
    PHYSIK_API int PHYSIK_GetConjugateGradientMaxIterations(PhysiK::WorldHandle world);    /* :71 */
    PHYSIK_API int PHYSIK_GetLastConjugateGradientIterations(PhysiK::WorldHandle world);   /* :73 */
    PHYSIK_API float PHYSIK_GetLastConjugateGradientResidualNorm(PhysiK::WorldHandle world); /* :75 */
    PHYSIK_API int PHYSIK_DidLastConjugateGradientSolveConverge(PhysiK::WorldHandle world); /* :77 */
    PHYSIK_API void PHYSIK_SetGravity(PhysiK::WorldHandle world, float x, float y, float z); /* :79 */
    PHYSIK_API int PHYSIK_AddNode(PhysiK::WorldHandle world, float x, float y, float z);   /* :84 */
    PHYSIK_API void PHYSIK_SetNodeFixed(PhysiK::WorldHandle world, int nodeIndex, int fixed); /* :90 */
    PHYSIK_API int PHYSIK_IsNodeFixed(PhysiK::WorldHandle world, int nodeIndex);           /* :95 */

    PHYSIK_API PhysiK::GeneratedTetMeshHandle PHYSIK_GenerateTetMesh(                      /* :99 */
        const PhysiK::Vec3* positions, int nodeCount, const int* tetLocalNodeIndices, int tetCount);
    PHYSIK_API int PHYSIK_IsGeneratedTetMeshHandleValid(PhysiK::GeneratedTetMeshHandle h); /* :105 */
    PHYSIK_API void PHYSIK_DestroyGeneratedTetMesh(PhysiK::GeneratedTetMeshHandle h);      /* :108 */
    PHYSIK_API int PHYSIK_GetGeneratedTetMeshVertexCount(PhysiK::GeneratedTetMeshHandle h); /* :111 */
    PHYSIK_API int PHYSIK_GetGeneratedTetMeshTetCount(PhysiK::GeneratedTetMeshHandle h);   /* :114 */
    PHYSIK_API int PHYSIK_GetGeneratedTetMeshTetIndexCount(PhysiK::GeneratedTetMeshHandle h); /* :117 */
    PHYSIK_API int PHYSIK_GetGeneratedTetMeshVertex(PhysiK::GeneratedTetMeshHandle h, int vertexIndex, /* :120 */
                                                    float* outX, float* outY, float* outZ);
    PHYSIK_API int PHYSIK_GetGeneratedTetMeshTetNodeIndex(PhysiK::GeneratedTetMeshHandle h, /* :127 */
                                                          int tetIndexArrayIndex, int* outNodeIndex);

    PHYSIK_API PhysiK::ComponentHandle PHYSIK_CreateTetMeshComponent(                      /* :132 */
        PhysiK::WorldHandle world, PhysiK::GeneratedTetMeshHandle generatedTetMeshHandle);
    PHYSIK_API PhysiK::ComponentHandle PHYSIK_CreateTetMeshPhysicsComponent(               /* :136 */
        PhysiK::WorldHandle world, PhysiK::GeneratedTetMeshHandle generatedTetMeshHandle,
        const PhysikMaterialDesc* material);
    PHYSIK_API int PHYSIK_GetTetMeshGlobalNodeBeginIndex(PhysiK::WorldHandle world,        /* :141 */
                                                         PhysiK::ComponentHandle tetMeshPhysicsHandle);
    PHYSIK_API int PHYSIK_GetTetMeshGlobalNodeCount(PhysiK::WorldHandle world,             /* :145 */
                                                    PhysiK::ComponentHandle tetMeshPhysicsHandle);
    PHYSIK_API int PHYSIK_GetTetMeshGlobalNodeIndex(PhysiK::WorldHandle world,             /* :149 */
                                                    PhysiK::ComponentHandle tetMeshPhysicsHandle,
                                                    int localNodeIndex);
    PHYSIK_API int PHYSIK_SetTetMeshLocalCurrentPosition(PhysiK::WorldHandle world,        /* :236 */
                                                         PhysiK::ComponentHandle tetMesh, int localNodeIndex,
                                                         float x, float y, float z);
    PHYSIK_API int PHYSIK_GetTetMeshLocalCurrentPosition(PhysiK::WorldHandle world,        /* :244 */
                                                         PhysiK::ComponentHandle tetMesh, int localNodeIndex,
                                                         float* outX, float* outY, float* outZ);
    PHYSIK_API void PHYSIK_SetTetMeshMaterial(PhysiK::WorldHandle world,                   /* :154 */
                                              PhysiK::ComponentHandle component,
                                              const PhysikMaterialDesc* material);
    PHYSIK_API PhysiK::ComponentHandle PHYSIK_CreateCollisionSphereComponent(              /* :159 */
        PhysiK::WorldHandle world, float x, float y, float z, float radius);
    PHYSIK_API void PHYSIK_SetCollisionSphereConnectionSettings(PhysiK::WorldHandle world, /* :288 */
                                                                PhysiK::ComponentHandle sphereComponent,
                                                                float stiffness, float damping);
    PHYSIK_API void PHYSIK_GetCollisionSphereConnectionSettings(PhysiK::WorldHandle world, /* :294 */
                                                                PhysiK::ComponentHandle sphereComponent,
                                                                float* outStiffness, float* outDamping);
    PHYSIK_API void PHYSIK_DestroyComponent(PhysiK::WorldHandle world,                     /* :300 */
                                            PhysiK::ComponentHandle component);
    PHYSIK_API int PHYSIK_IsComponentHandleValid(PhysiK::WorldHandle world,                /* :304 */
                                                 PhysiK::ComponentHandle component);
    PHYSIK_API int PHYSIK_GetTetMeshTetCount(PhysiK::WorldHandle world,                    /* :308 */
                                             PhysiK::ComponentHandle component);
    PHYSIK_API int PHYSIK_GetTetMeshNodeCount(PhysiK::WorldHandle world,                   /* :312 */
                                              PhysiK::ComponentHandle component);
    PHYSIK_API int PHYSIK_IsTetActive(PhysiK::WorldHandle world, PhysiK::ComponentHandle component, /* :316 */
                                      int tetIndex);
    PHYSIK_API void PHYSIK_SetTetActive(PhysiK::WorldHandle world, PhysiK::ComponentHandle component, /* :321 */
                                        int tetIndex, int active);
    PHYSIK_API void PHYSIK_DeactivateTet(PhysiK::WorldHandle world, PhysiK::ComponentHandle component, /* :327 */
                                         int tetIndex);
    PHYSIK_API int PHYSIK_GetActiveTetCount(PhysiK::WorldHandle world,                     /* :332 */
                                            PhysiK::ComponentHandle component);
    PHYSIK_API int PHYSIK_GetCollisionSphereOverlapCount(PhysiK::WorldHandle world,        /* :336 */
                                                         PhysiK::ComponentHandle sphereComponent);
    PHYSIK_API int PHYSIK_GetCollisionSphereOverlaps(PhysiK::WorldHandle world,            /* :340 */
                                                     PhysiK::ComponentHandle sphereComponent,
                                                     PhysikCollisionSphereOverlap* outOverlaps, int maxOverlaps);
    PHYSIK_API void PHYSIK_SetCollisionComponentKinematicTarget(PhysiK::WorldHandle world, /* :346 */
                                                                PhysiK::ComponentHandle component, float x,
                                                                float y, float z);
    PHYSIK_API int PHYSIK_GetPointConnectionCount(PhysiK::WorldHandle world);             /* :353 */
    PHYSIK_API void PHYSIK_AddPointConnection(PhysiK::WorldHandle world, int node0, int node1, /* :355 */
                                              int node2, int node3, float barycentricX, float barycentricY,
                                              float barycentricZ, float barycentricW, float targetX,
                                              float targetY, float targetZ, float stiffness, float damping);
    PHYSIK_API void PHYSIK_GetNodePosition(PhysiK::WorldHandle world, int nodeIndex,       /* :371 */
                                           float* outX, float* outY, float* outZ);
    PHYSIK_API void PHYSIK_GetNodeVelocity(PhysiK::WorldHandle world, int nodeIndex,       /* :378 */
                                           float* outX, float* outY, float* outZ);
    PHYSIK_API void PHYSIK_SetNodePosition(PhysiK::WorldHandle world, int nodeIndex,       /* :385 */
                                           float x, float y, float z);
    PHYSIK_API void PHYSIK_SetNodeVelocity(PhysiK::WorldHandle world, int nodeIndex,       /* :392 */
                                           float x, float y, float z);

    /* ---- additive entry points (no reference counterpart) ---- */

    /* Storage precision of the GPU system: 0 = FP64 (default), 1 = FP32
     * storage with FP64 dots and residual update, 2 = the same with only the
     * upper triangle of the matrix stored (identical results to 1), 3 = FP64
     * with only the upper triangle stored (identical results to 0). */
    PHYSIK_API int PHYSIK_AMD_SetPrecision(PhysiK::WorldHandle world, int precision);
    /* PHYSIK_Step that returns a status instead of throwing: 0 ok, 3 when a
     * free node has no mass (the reference throws), 1 bad handle, 2 GPU error. */
    PHYSIK_API int PHYSIK_AMD_TryStep(PhysiK::WorldHandle world, float dt);
    /* Outcome of the last substep's solve: 0 integrated, 1 solve failed,
     * 2 non-finite update rejected, 3 no dynamic node. */
    PHYSIK_API int PHYSIK_AMD_GetLastStepStatus(PhysiK::WorldHandle world);
    /* How many times the world has built its device system. PHYSIK_SetTetActive
     * switches a tet in place at the next step while the set of active nodes
     * stays the same; any other topology change builds the system again. */
    PHYSIK_API int PHYSIK_AMD_GetSystemBuildCount(PhysiK::WorldHandle world);
    /* Bulk node state in FP64 (count nodes, xyz triples). Copy returns the
     * number of nodes written; Set writes positions and/or velocities as given
     * (positions do not reset velocities here). */
    PHYSIK_API int PHYSIK_AMD_CopyNodeStateF64(PhysiK::WorldHandle world, double* outPositions,
                                               double* outVelocities, int count);
    PHYSIK_API int PHYSIK_AMD_SetNodeStateF64(PhysiK::WorldHandle world, const double* positions,
                                              const double* velocities, int count);
    /* PHYSIK_GenerateTetMesh with FP64 positions (same welding and
     * degenerate-tet removal, evaluated in FP64). */
    PHYSIK_API PhysiK::GeneratedTetMeshHandle PHYSIK_AMD_GenerateTetMeshF64(
        const double* positions, int nodeCount, const int* tetLocalNodeIndices, int tetCount);
    /* Per-tet Young's modulus of a physics component (count = its tet count). */
    PHYSIK_API int PHYSIK_AMD_SetTetMeshYoungModuli(PhysiK::WorldHandle world, PhysiK::ComponentHandle component,
                                                    const double* youngModuli, int count);
    /* FEM model of a physics component: the reference's
     * TetMeshPhysicsComponentDesc::femModel (TetMeshPhysicsComponent.h:13-17),
     * which its C ABI cannot set. 0 Linear, 1 Corotational (the default), 2
     * NeoHookean (no elastic contribution, as FEMModel::ComputeForces).
     * Returns 1, or 0 for a bad handle or model. */
    PHYSIK_API int PHYSIK_AMD_SetTetMeshFemModel(PhysiK::WorldHandle world, PhysiK::ComponentHandle component,
                                                 int femModel);
    /* Frame readback (the per-frame readers of the reference: visual
     * components in World::PostUpdate, World.cpp:539-549, and a host engine's
     * PHYSIK_GetNodePosition/Velocity loop). After each PHYSIK_Step the GPU
     * converts the node state to float triples and copies them into pinned
     * host memory without the host waiting; the first getter of the frame
     * waits for that copy, and the getters then read it. mask: 1 positions,
     * 2 velocities, 3 both (the default), 0 none (getters then download the
     * FP64 state). Returns 1, or 0 for a bad handle or mask. */
    PHYSIK_API int PHYSIK_AMD_SetFrameExport(PhysiK::WorldHandle world, int mask);
    /* Bulk PHYSIK_GetNodePosition / PHYSIK_GetNodeVelocity: nodes
     * [firstNode, firstNode + count) into out (clipped to the node count),
     * the same float values the per-node getters return. Returns the number
     * of nodes written (0 for a bad handle or range). */
    PHYSIK_API int PHYSIK_AMD_GetNodePositions(PhysiK::WorldHandle world, int firstNode, int count,
                                               PhysiK::Vec3* outPositions);
    PHYSIK_API int PHYSIK_AMD_GetNodeVelocities(PhysiK::WorldHandle world, int firstNode, int count,
                                                PhysiK::Vec3* outVelocities);
    /* Message of the last error raised by this library on this thread. */
    PHYSIK_API const char* PHYSIK_AMD_GetLastError();
}

#endif /* PHYSIK_AMD_PHYSIKAPI_H */
