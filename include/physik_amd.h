/*
 * physik_amd.h - the device-level C ABI of the MI355X implicit-Euler step.
 *
 * This layer has no single counterpart in the reference: it is the thin
 * C ABI between the PHYSIK_* host API (PhysiKAPI.h in this directory) and the
 * HIP kernels. It replaces, as one GPU-resident object, the per-frame work the
 * reference does on the host in
 *   World::Step substep loop ............ src/Core/World/World.cpp:32-73
 *   TetMeshPhysicsComponent::UpdateSystem  src/Components/TetMeshPhysicsComponent.cpp:293-354
 *   FEMModel corotational element ....... src/Core/Physics/FEM/FEMModel.cpp:614-679
 *   SolverData implicit assembly ........ src/Core/Solvers/SolverData.cpp:126-465
 *   block-Jacobi PCG .................... src/Core/Solvers/Linear/ConjugateGradientSolver.cpp:114-241
 *   IntegrateImplicitEuler .............. src/Core/World/World.cpp:644-695
 *
 * Conventions: opaque handles, int status (PK_OK = 0), no exceptions across
 * the ABI, caller-owned host buffers, system-owned device buffers. Vectors are
 * node-major xyz triples; 3x3 blocks are row-major. All entry points are
 * synchronous with respect to the host unless stated otherwise.
 *
 * A system may hold several independent bodies ("groups"): each group gets
 * its own CG solve with its own convergence test, all advanced by the same
 * kernel launches. One group reproduces the reference's single global solve
 * over every component of a World.
 */
#ifndef PHYSIK_AMD_H
#define PHYSIK_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

typedef struct pk_system pk_system;
typedef struct pk_pattern pk_pattern;

enum pk_status {
    PK_OK = 0,
    PK_ERR_ARG = 1,       /* invalid argument or size */
    PK_ERR_HIP = 2,       /* a HIP runtime call failed (message in pk_last_error) */
    PK_ERR_MASS = 3,      /* a free active node has no mass: the reference throws
                             std::runtime_error here (World.cpp:560-580) */
    PK_ERR_NO_DEVICE = 4, /* no GPU visible */
    PK_ERR_REBUILD = 5,   /* the change needs the system created again (nothing was changed) */
    PK_ERR_STATE = 6      /* nothing to read: no frame export since the state last changed */
};

/* What a frame export carries (pk_system_export_begin). */
enum pk_export {
    PK_EXPORT_POSITIONS = 1,
    PK_EXPORT_VELOCITIES = 2
};

enum pk_precision {
    PK_FP64 = 0,          /* everything in FP64 */
    PK_FP32_STORAGE = 1,  /* A, M^-1 and the SpMV vectors stored in FP32; dots,
                             residual update, solution and state in FP64 */
    PK_FP32_UPPER = 2,    /* as PK_FP32_STORAGE, with A stored as its upper
                             triangle only (each lower block read as the
                             transpose of its mirror); results identical to
                             PK_FP32_STORAGE, about 35 % fewer SpMV bytes */
    PK_FP64_UPPER = 3     /* as PK_FP64, with A stored as its upper triangle
                             only; results identical to PK_FP64 */
};

/* Per-group step outcome (pk_system_group_results). */
enum pk_group_status {
    PK_GROUP_INTEGRATED = 0,     /* solve usable, state advanced */
    PK_GROUP_SOLVE_FAILED = 1,   /* preconditioner or inputs rejected: state kept
                                    (SolverData.cpp:169-212, World.cpp:56-62) */
    PK_GROUP_NONFINITE = 2,      /* integration produced a non-finite value: state
                                    kept (World.cpp:667-677) */
    PK_GROUP_NO_DYNAMIC = 3      /* no dynamic node in the group */
};

/* Kernel ids for pk_system_kernel_time. */
enum pk_kernel {
    PK_K_ELEMENT = 0,     /* per-tet corotation, forces, R Ke R^T */
    PK_K_ASSEMBLE = 1,    /* atomic-free gather into BSR slots + block-Jacobi inverse */
    PK_K_RHS = 2,         /* right-hand side gather + CG start */
    PK_K_SPMV = 3,        /* q = A d fused with d.q */
    PK_K_UPDATE = 4,      /* x, r update, z = M^-1 r, r.z, r.r */
    PK_K_DIRECTION = 5,   /* d = z + beta d */
    PK_K_INTEGRATE = 6,   /* v += dv, x += dt v (check + commit) */
    PK_K_CONTACT = 7,     /* point connections and collision-sphere contacts */
    PK_K_COUNT = 8
};

typedef struct pk_mesh_desc {
    int node_count;
    const double* rest;              /* node_count x 3 rest positions */
    int tet_count;
    const int* tets;                 /* tet_count x 4 node ids */
    const double* young;             /* tet_count */
    const double* poisson;           /* tet_count */
    const double* damping;           /* tet_count */
    const double* density;           /* tet_count (the owning component's density) */
    const unsigned char* tet_active; /* tet_count, or NULL for all active */
    const unsigned char* node_fixed; /* node_count, or NULL for none fixed */
    const unsigned char* node_active;/* node_count, or NULL for all active */
    const int* node_group;           /* node_count non-decreasing group ids starting at 0,
                                        or NULL for a single group */
    int extra_pair_count;            /* node pairs added to the BSR pattern beside the
                                        tets' (for point connections between nodes no
                                        tet joins, SolverData.cpp:313-401) */
    const int* extra_pairs;          /* extra_pair_count x 2 node ids, or NULL */
} pk_mesh_desc;

/* Message of the last failed call on this thread. */
const char* pk_last_error(void);

/* Number of visible GPUs (0 when none; never fails). */
int pk_device_count(void);

/* Uploads a mesh and builds everything that only depends on topology and
 * material: rest data, lumped masses, dynamic-node numbering, the BSR pattern
 * and the element->slot gather lists. Positions start at rest, velocities at
 * zero. Runs on the current HIP device. */
int pk_system_create(const pk_mesh_desc* desc, int precision, pk_system** out);
void pk_system_destroy(pk_system* sys);

int pk_system_sizes(const pk_system* sys, int* nodes, int* tets, int* groups, int* dynamic_rows,
                    long long* nnzb_upper, long long* nnzb_full);

/* Launch shape of the CG kernels: lanes per BSR row in the SpMV (1, 4, 8, 16
 * or 32), rows per SpMV workgroup, rows per vector-kernel workgroup, and rows
 * each SpMV lane group loads before reducing (1, 2, or 4 up to 16 lanes).
 * Upper-only storage (PK_FP32_UPPER, PK_FP64_UPPER) runs one lane per row
 * over slices of 64 rows: lanes_per_row must be 1, spmv_rows a multiple of
 * 64, and unroll is not used (any accepted value runs the same). Results
 * of a body do not depend on the other bodies of the system, only on these
 * numbers; upper-only storage gives the results of full storage at one lane
 * per row. A default is chosen from the row lengths at creation; the
 * autotuner (physik_amd/autotune.py) searches them. */
int pk_system_tune(pk_system* sys, int lanes_per_row, int spmv_rows, int vector_rows, int unroll);
int pk_system_get_tune(const pk_system* sys, int* lanes_per_row, int* spmv_rows, int* vector_rows, int* unroll);

/* SpMV staging (no reference counterpart). mode 1 (the default) lets each
 * SpMV workgroup copy the rows of the direction vector its tile references
 * into LDS and read the direction blocks from there, up to a capacity: the
 * LDS share that still leaves a CU every workgroup of the selected SpMV
 * kernel its registers allow (so it depends on the storage, precision and
 * launch shape). When every tile's column window [first, last] fits, tiles
 * stage their windows; otherwise they stage up to four segments of rows (the
 * columns cut at their three widest gaps), and tiles that still need more
 * than the capacity gather from global memory. mode 0 gathers from global
 * memory. Results are identical in every mode. get reports the mode, the
 * rows a workgroup stages (0: none) and whether segments are in use;
 * capacity the rows the capacity holds (0 with staging off). */
int pk_system_set_spmv_staging(pk_system* sys, int mode);
int pk_system_get_spmv_staging(const pk_system* sys, int* mode, int* rows, int* segmented);
int pk_system_get_spmv_stage_capacity(const pk_system* sys, int* rows);

/* Sliced upper-only storage layout (no reference counterpart): the slices
 * of 64 rows, how many of them are stored compactly (a position's blocks
 * back to back in lane order, for slices whose ELL positions are less than
 * 97 % filled; the others keep one entry per lane and position), and the
 * elements held, padding included. All zero under full storage. */
int pk_system_get_slice_layout(const pk_system* sys, int* slices, int* compact_slices, long long* elements);

/* Node state, node_count x 3 each; either pointer may be NULL. */
int pk_system_set_state(pk_system* sys, const double* x, const double* v);
int pk_system_get_state(pk_system* sys, double* x, double* v);
/* The same with device pointers on the system's device (no host round trip):
 * the caller's writes of x and v must be complete before the call (e.g. its
 * stream synchronised); the copies are complete when the call returns. */
int pk_system_set_state_device(pk_system* sys, const double* x, const double* v);
int pk_system_get_state_device(pk_system* sys, double* x, double* v);

/* Frame export: the readback after a step for the reference's per-frame
 * readers (visual components in World::PostUpdate, World.cpp:539-549, and a
 * host engine's PHYSIK_GetNodePosition/Velocity loop, PhysiKAPI.cpp:1510-1545,
 * which return float). pk_system_export_begin enqueues on the system's stream
 * a device conversion of the node state to the reference's float Vec3 layout
 * (3 floats per node, each the value static_cast<float> gives) and its copy
 * into pinned host memory that the system owns, and returns without waiting;
 * `what` is PK_EXPORT_POSITIONS, PK_EXPORT_VELOCITIES or both. 12 bytes per
 * node and quantity cross PCIe instead of pk_system_get_state's 24.
 * pk_system_export_view waits for the copy and points at the float arrays
 * (node_count x 3; pass NULL for a quantity not wanted); the pointers stay
 * valid until the next export or the destruction of the system.
 * pk_system_export_read waits likewise and copies nodes [first, first+count).
 * Both return PK_ERR_STATE when the quantity asked for was not exported since
 * x or v last changed (a step, pk_system_set_state, record or backward). */
int pk_system_export_begin(pk_system* sys, int what);
int pk_system_export_view(pk_system* sys, const float** positions, const float** velocities);
int pk_system_export_read(pk_system* sys, int first, int count, float* positions, float* velocities);

int pk_system_set_gravity(pk_system* sys, double gx, double gy, double gz);
/* CG settings, applied to every group (ConjugateGradientSettings,
 * ConjugateGradientSolver.h:12-17). No clamping happens here. */
int pk_system_set_solver(pk_system* sys, double tolerance, int max_iterations, int use_jacobi);
/* 1 = implicit Euler (default), 0 = explicit Euler (World.cpp:697-725). */
int pk_system_set_mode(pk_system* sys, int implicit);
/* FEM model per tet (tet_count values, the reference's FemModel,
 * FEMModel.h:40-45): 0 linear (FEMModel::ComputeElasticForces,
 * FEMModel.cpp:559-612: K1 with R = I, no deformation-gradient check), 1
 * corotational (the default, TetMeshPhysicsComponent.h:13-17), 2 NeoHookean,
 * for which the reference's FEMModel::ComputeForces produces no element
 * contribution (the tet keeps its lumped mass). The reverse mode covers the
 * corotational model only: pk_system_record fails while any tet has another. */
int pk_system_set_fem_models(pk_system* sys, const unsigned char* models);
/* Per-tet Young's modulus (tet_count values); topology is unchanged. */
int pk_system_set_young(pk_system* sys, const double* young);

/* Advances frames x substeps substeps of length dt / substeps each, entirely
 * on the GPU (World::Step, World.cpp:32-73). dt is the frame length. Returns
 * PK_OK, or PK_ERR_MASS before any work when a free node has no mass. */
int pk_system_step(pk_system* sys, double dt, int substeps, int frames);

/* Outcome of the last substep, one entry per group (any pointer may be NULL). */
int pk_system_group_results(pk_system* sys, int* iterations, double* residual_norm, int* converged,
                            int* status);

/* Transient point springs (PHYSIK_AddPointConnection, PointConnection.cpp:
 * 54-130): `count` connections, each pulling the barycentric point
 * sum_a weights[4i+a] x[nodes[4i+a]] toward targets[3i..3i+2] with stiffness
 * k and damping c (force k (target - point) - c (point velocity), spread to
 * the nodes by weight; k w_a w_b I blocks into the implicit matrix). They act
 * in every substep of every following pk_system_step until replaced; count 0
 * clears them. As in the reference, a connection is ignored when a node is out
 * of range or inactive, k or c is negative or non-finite, or the target is
 * non-finite. Every pair of dynamic nodes a connection couples must lie in the
 * system's BSR pattern (pk_mesh_desc.extra_pairs adds pairs): PK_ERR_ARG
 * otherwise, with the connections left as they were. The reverse mode
 * (pk_system_record / backward) differentiates their forces (the targets,
 * stiffness and damping are constants of the run). */
int pk_system_set_connections(pk_system* sys, int count, const int* nodes, const double* weights,
                              const double* targets, const double* stiffness, const double* damping);

/* Switches tets on or off in place (TetMeshComponent::SetTetActive,
 * TetMeshComponent.cpp:153-169): count tet ids and their new active flags (the
 * last entry for a tet wins). A tet switched off stops producing forces,
 * stiffness, lumped mass and sphere contacts; its node pairs stay in the BSR
 * pattern as zero blocks, as the reference keeps its pattern
 * (PointConnectionTests.cpp:2780-2803). Masses are summed again as the
 * creation sums them, so the state after a switch equals that of a system
 * created with the new flags up to the order of additions in the CG. Node
 * activity is not changed here. PK_ERR_REBUILD, with nothing changed, when
 * the switch would change which nodes are dynamic (a free node left without
 * mass, or one gaining its first) or turn on a tet that was not live when the
 * system was created: create the system again with the new flags then. */
int pk_system_set_tets_active(pk_system* sys, int count, const int* tets, const unsigned char* active);

/* Collision spheres (CollisionSphereComponent::UpdateSystem,
 * CollisionSphereComponent.cpp:112-193): in every substep of every following
 * pk_system_step, each active tet with four active nodes whose centroid p lies
 * inside a sphere (distance d < radius) gets a spring with weights 1/4 from p
 * to p + n (radius - d), n = (p - centre) / d (or (0, 0, 1) when d <= 1e-5),
 * of the sphere's stiffness and damping, for that substep only; the contacts
 * are found on the device from the substep's starting state and act on top of
 * the point connections. `count` spheres: centers 3 x count, radii, stiffness
 * and damping finite and >= 0 (PK_ERR_ARG otherwise; the PHYSIK_* layer
 * clamps as the reference does, PhysiKAPI.cpp:685-700, 1180-1194). A sphere of
 * radius 0 touches nothing. count 0 removes them. The reverse mode
 * differentiates the contact forces (their targets move with the centroid;
 * which tets touch a sphere is piecewise constant and contributes no
 * derivative). */
int pk_system_set_spheres(pk_system* sys, int count, const double* centers, const double* radii,
                          const double* stiffness, const double* damping);

/* Reverse-mode memory (no reference counterpart). pk_system_backward replays
 * each checkpoint segment forward and then reverses it substep by substep;
 * when the segment's matrices (A and M^-1 of every substep) fit max_bytes
 * (default 8 GiB), the replay keeps them and the reverse steps solve with
 * them instead of assembling again. 0 turns this off. Gradients are identical
 * either way. */
int pk_system_set_matrix_store(pk_system* sys, long long max_bytes);

/* Launch mode (no reference counterpart); results are identical in every
 * mode. 0: every kernel launched from the host. 1: every solve replays its CG
 * four iterations at a time from HIP graphs captured on first use (direction,
 * SpMV and update launches, plus a batch's tile compaction). 2: each implicit
 * forward substep of pk_system_step is one captured graph -- reset,
 * element, contacts, assembly, right-hand side, enough CG chunks for the
 * previous solve's iteration count plus one, and the integration of every
 * body whose solve ended within them; a body that needs more iterations is
 * finished by launches from the host and integrated after (record and
 * backward replay CG chunks as in mode 1, and a profiled step launches its
 * kernels). A capture is redone when a setting, the launch shape, the step
 * or that chunk count changes. PK_ERR_ARG for another mode. */
int pk_system_set_graphs(pk_system* sys, int mode);

/* Kernel timing on the system's stream: enable, reset, read. The CG kernels
 * (SpMV, update) are timed from the wall clock read inside the kernel (first
 * workgroup start to last workgroup end, the span a kernel trace reports);
 * the others with HIP events around the launch. Only CG launches that did
 * work are counted (those after every body stopped are excluded). */
int pk_system_profile(pk_system* sys, int enable);
int pk_system_kernel_time(pk_system* sys, int kernel, long long* launches, double* total_ms);

/* Work the profiled SpMV launches did, summed over launches: full BSR blocks
 * multiplied, the upper-triangle blocks among them (the stored ones under
 * PK_FP32_UPPER), dynamic rows, and (body, launch) pairs. The rows of a body
 * whose CG loop has ended are skipped, so in a batched solve late launches
 * multiply only the bodies still iterating. Reset by pk_system_profile. */
int pk_system_spmv_work(pk_system* sys, long long* blocks, long long* upper_blocks, long long* rows,
                        long long* body_launches);

/* ---- inspection hooks used by the parity tests ---- */

/* Runs the element and assembly kernels of one substep of length dt (no
 * solve, no state change). */
int pk_system_assemble(pk_system* sys, double dt);
/* Assembled system: full BSR values (nnzb_full x 9, row-major blocks, in
 * pattern order), the CG residual vector (dynamic_rows x 3: the right-hand
 * side b right after pk_system_assemble, the final residual after a step)
 * and block-Jacobi inverses (dynamic_rows x 9). Any pointer may be NULL. */
int pk_system_get_matrix(pk_system* sys, double* full_vals, double* rhs, double* minv);
/* Per-tet element outputs at the current state (what the last assembly used
 * when the state has not changed since): the 10 blocks R K_ab R^T for a <= b
 * in order (0,0),(0,1),(0,2),(0,3),(1,1),(1,2),(1,3),(2,2),(2,3),(3,3)
 * (tet_count x 90), the element forces (tet_count x 12) and K v per corner
 * (tet_count x 12); zero for tets that contribute nothing. Any pointer may be
 * NULL. */
int pk_system_get_elements(pk_system* sys, double* blocks, double* forces, double* kv);
/* Per-tet rotation of the corotation at the current state (tet_count x 9,
 * row-major R) and how ExtractRotationPolar (FEMModel.cpp:115-163) found it
 * (tet_count): 0 the Newton limit, 1 the Newton limit with column 2 negated
 * (det < 0), 2 the Identity fallback (non-finite or |det| <= 1e-8), 3 the
 * linear model (R = I), -1 the tet contributes nothing (inactive, zero
 * volume, non-finite F or no elastic model; R reads 0).
 * Either pointer may be NULL. */
int pk_system_get_rotations(pk_system* sys, double* rotations, int* kind);
/* y = A x with the assembled full BSR (dynamic_rows x 3). */
int pk_system_spmv(pk_system* sys, const double* x, double* y);
/* Solution dv of the last solve (dynamic_rows x 3). */
int pk_system_get_solution(pk_system* sys, double* dv);

/* Reverse mode of the implicit step (no counterpart in the reference, which
 * has no adjoint; the role the reference's interface would give it is an
 * "adjoint_step" beside the forward step).
 *
 * pk_system_record runs `steps` implicit substeps of length dt from the
 * current state, exactly as pk_system_step would (the state advances), and
 * keeps device checkpoints (x, v) every `every` substeps. Any later change of
 * material, solver, gravity, mode or launch shape discards the recording.
 *
 * pk_system_backward propagates the adjoints of the recorded run's final
 * state -- xbar = dL/dx_S and vbar = dL/dv_S (node_count x 3 each; NULL means
 * zero) -- back to its initial state. Each checkpoint segment is replayed
 * forward on the device, then reversed substep by substep: the adjoint solve
 * A lam = vbar + dt xbar on the dynamic rows (same matrix and preconditioner,
 * block-Jacobi PCG to adj_tol in the same M^-1-norm rule, at most
 * adj_max_iter iterations), then the corotational terms including dR/dF.
 * Outputs (NULL to skip): dL/dlog E per tet (tet_count), dL/dx_0 and dL/dv_0
 * (node_count x 3), and the largest adjoint CG iteration count of each
 * substep (steps). Bodies of a batched system are reversed independently; a
 * body whose forward substep did not integrate passes its adjoints through.
 * The system is left at the recorded final state. Memory: checkpoints
 * (steps / every + 1) x node_count x 48 bytes, segment (every + 1) x
 * node_count x 48 bytes. */
int pk_system_record(pk_system* sys, double dt, int steps, int every);
int pk_system_backward(pk_system* sys, const double* xbar, const double* vbar, double adj_tol, int adj_max_iter,
                       double* grad_log_young, double* grad_x0, double* grad_v0, int* adj_iterations);
/* pk_system_backward with xbar, vbar and the three gradient arrays as device
 * pointers on the system's device (adj_iterations stays a host pointer);
 * results are identical. Same completion rules as pk_system_set_state_device. */
int pk_system_backward_device(pk_system* sys, const double* xbar, const double* vbar, double adj_tol,
                              int adj_max_iter, double* grad_log_young, double* grad_x0, double* grad_v0,
                              int* adj_iterations);
/* Identifies the recorded run pk_system_backward would reverse: a count of
 * this system's successful pk_system_record calls (so a later recording has
 * a larger id), or 0 when no valid recording exists (none made, or discarded
 * by a later change). A caller holding an id can detect that the system has
 * recorded another run since. */
int pk_system_tape_id(const pk_system* sys, long long* id);

/* ---- host-only topology (no GPU needed) ---- */

/* Builds the dynamic numbering and the BSR patterns exactly as pk_system_create
 * does, on the host. */
int pk_pattern_create(const pk_mesh_desc* desc, pk_pattern** out);
void pk_pattern_destroy(pk_pattern* p);
int pk_pattern_sizes(const pk_pattern* p, int* dynamic_rows, long long* nnzb_upper, long long* nnzb_full,
                     long long* contributions);
/* dyn: node_count (dynamic row or -1); upper_start/upper_col: symmetric-upper
 * pattern (the reference's SolverData matrix, SparseBlockMatrix.cpp:83-129);
 * full_start/full_col: both triangles; scatter: tet_count x 16 upper slot of
 * each local pair (a, b) or -1 (bit-identical to FindBlockIndex,
 * SparseBlockMatrix.cpp:270-282). Any pointer may be NULL. */
int pk_pattern_copy(const pk_pattern* p, int* dyn, int* upper_start, int* upper_col, int* full_start,
                    int* full_col, int* scatter);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* PHYSIK_AMD_H */
