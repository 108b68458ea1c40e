/*
 * physik_amd_linear.h - the block-Jacobi PCG solve of one block-sparse system
 * on the GPU: the reference's LinearSolver plugin slot.
 *
 * Reference interface replaced: PhysiK::LinearSolver::Solve(const
 * SparseBlockMatrix&, const std::vector<Vec3>& rhs, std::vector<Vec3>&
 * solution, const LinearSolveSettings&) -> LinearSolveResult
 * (include/PhysiK/Core/Solvers/Linear/LinearSolver.h:27-37), implemented by
 * CurrentLinearSolver::Solve (src/Core/Solvers/Linear/LinearSolver.cpp:93-131:
 * BuildInversePreconditioner :30-85, then SolvePreconditionedConjugateGradient,
 * ConjugateGradientSolver.cpp:114-241).
 *
 * Semantics kept: block-Jacobi preconditioner from the inverse of each
 * diagonal block (the solve is refused and the solution left empty when a
 * diagonal block is missing or non-finite, |det| <= 1e-8 or its inverse is
 * non-finite; Mat3::Inverse gives Identity for |det| <= 1e-6; Identity blocks
 * when Jacobi is off); maxIterations <= 0 means 3 n; x0 = 0; the stopping
 * rule r^T M^-1 r <= tol^2 r0^T M^-1 r0; iterations counts completed updates;
 * breakdown (d.q <= 0 or any non-finite scalar) ends the loop; residualNorm
 * is ||r||_2 of the final residual. Arithmetic is FP64 (the reference's is
 * FP32). Conventions as physik_amd.h: int status (0 = PK_OK), message in
 * pk_last_error(); a refused solve is PK_OK with solution_valid = 0.
 */
#ifndef PHYSIK_AMD_LINEAR_H
#define PHYSIK_AMD_LINEAR_H

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

/* LinearSolveResult (LinearSolver.h:20-25) plus whether a solution exists:
 * solution_valid = 0 where the reference clears the solution vector
 * (preconditioner refused, non-finite right-hand side, tolerance not finite
 * and > 0). */
typedef struct pk_linear_result {
    int iterations;
    int converged;
    int solution_valid;
    int pad;
    double residual_norm;
} pk_linear_result;

/* Largest system pk_bsr_pcg_noalloc solves: its work arrays live in a
 * file-scope device array sized for this many rows (44,053,760 bytes: 168
 * bytes per row plus the partial sums), which every GPU the library runs on
 * holds until the process ends. */
#define PK_LINEAR_NOALLOC_MAX_ROWS 262144

/* The plugin entry, host memory in the reference's own layout: a
 * SparseBlockMatrix of block_count block rows (row_start: block_count + 1
 * offsets, col_index: row_start[block_count] block columns, values: one
 * column-major Mat3 of 9 floats per stored block, as in Mat3.h:9-14; every
 * stored block (r, c) also acts as (c, r) transposed when r != c, as
 * SparseBlockMatrix::Multiply applies it, SparseBlockMatrix.cpp:209-268), rhs
 * and solution as block_count Vec3s (3 floats each). Solves on the current
 * GPU in FP64; solution receives static_cast<float> of the result when
 * result->solution_valid is set and is left untouched otherwise. Allocates
 * device memory for the matrix and the work arrays per call. */
int pk_linear_solve(int block_count, const int* row_start, const int* col_index, const float* values,
                    const float* rhs, float* solution, int max_iterations, float tolerance, int use_jacobi,
                    pk_linear_result* result);

/* The same solve on device memory of the current GPU, in the pk_* layout:
 * the full block matrix (both triangles: each off-diagonal block stored at
 * (r, c) and its transpose at (c, r)) as rows + 1 int32 row offsets, int32
 * block columns and row-major 3x3 FP64 blocks; b and x as rows x 3 doubles.
 * Work is issued on `stream` (a hipStream_t; NULL for the null stream) and
 * has finished when the call returns (the host reads the stop state every 8
 * iterations). The work arrays (about 168 bytes per row) are allocated with
 * hipMalloc and freed per call. x holds no solution when solution_valid ends
 * 0. */
int pk_bsr_pcg(int rows, const int* row_start, const int* col, const double* vals, const double* b, double* x,
               int max_iterations, double tolerance, int use_jacobi, void* stream, pk_linear_result* result);

/* pk_bsr_pcg without any allocation: no hipMalloc, hipFree or allocating
 * library call. No counterpart in the reference; it is for callers that
 * cannot allocate (code between HIP graph captures, serving loops, harnesses
 * that forbid device allocation). Identical kernels and results as
 * pk_bsr_pcg; the work arrays are one file-scope device array per GPU
 * (PK_LINEAR_NOALLOC_MAX_ROWS rows at most: larger systems are refused with
 * PK_ERR_ARG), located on first use on each device. Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (from several
 * host threads or streams) race and must be serialised by the caller. Each
 * call writes every work value it reads, so nothing carries over between
 * calls. */
int pk_bsr_pcg_noalloc(int rows, const int* row_start, const int* col, const double* vals, const double* b,
                       double* x, int max_iterations, double tolerance, int use_jacobi, void* stream,
                       pk_linear_result* result);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* PHYSIK_AMD_LINEAR_H */
