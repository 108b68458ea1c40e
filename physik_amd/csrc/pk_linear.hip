// The block-Jacobi PCG solve of one block-sparse system (include/physik_amd_linear.h):
// the reference's LinearSolver plugin slot, CurrentLinearSolver::Solve
// (LinearSolver.cpp:30-131) over SolvePreconditionedConjugateGradient
// (ConjugateGradientSolver.cpp:114-241), in FP64 on the GPU.
//
// Iteration k is three launches and no launch synchronises internally:
//   k_lin_spmv(k)    q = A d (4 lanes per block row, 256 rows per workgroup),
//                    a d.q partial per workgroup
//   k_lin_update(k)  every workgroup sums the d.q partials in the same order,
//                    so all take the same alpha and breakdown decision;
//                    x += alpha d, r -= alpha q, z = M^-1 r, (r.z, r.r) partials
//   k_lin_dir(k)     every workgroup sums those partials alike: delta, beta,
//                    d = z + beta d; workgroup 0 records the decisions
// State a launch reads was written by an earlier launch (delta is kept by
// iteration parity), except the stop index, which a launch of iteration k
// only ever moves to k or k + 1 and reads as "stop before k": a workgroup
// that sees the write of its own launch would have taken the same decision.
// The host issues 8 iterations, then reads the stop index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "physik_amd.h"
#include "physik_amd_linear.h"
#include "pk_device.h"

namespace pk
{
void set_error(const std::string& msg);

namespace lin
{

constexpr int kVecRows = 1024;  // rows per vector workgroup (4 per thread)
constexpr int kSpmvRows = 256;  // rows per SpMV workgroup: 4 passes of 64 rows, 4 lanes per row
constexpr int kChunk = 8;       // iterations issued between two reads of the stop index

enum Outcome
{
    kLoop = 0,           // the iteration loop ran (converged, capped or broke down)
    kConvergedAtStart = 1,
    kRefused = 2,        // the reference clears the solution
    kStoppedAtStart = 3  // x = 0 and an empty result (ConjugateGradientSolver.cpp:151-169)
};

struct State
{
    double delta[2];   // r.z entering iteration k, at index k & 1
    double target;     // tol^2 delta_0
    double deltaLast;  // the latest r.z computed (the reference's deltaNew)
    double rr;         // r.r of the current residual
    double residual;
    int stopBefore;    // the first iteration that does not run
    int iters;
    int outcome;
    int converged;
};

// Work arrays carved from one block of device memory.
struct Work
{
    double *minv, *r, *z, *d, *q, *pdq, *prz, *prr;
    int* pflag;
    State* st;
};

__host__ __device__ constexpr size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

__host__ __device__ constexpr size_t vec_groups(size_t n) { return (n + kVecRows - 1) / kVecRows; }
__host__ __device__ constexpr size_t spmv_groups(size_t n) { return (n + kSpmvRows - 1) / kSpmvRows; }

// Bytes of the work arrays for n rows.
__host__ __device__ constexpr size_t work_bytes(size_t n)
{
    return round256(9 * n * sizeof(double)) + 4 * round256(3 * n * sizeof(double)) +
           round256(spmv_groups(n) * sizeof(double)) + 2 * round256(vec_groups(n) * sizeof(double)) +
           round256(vec_groups(n) * sizeof(int)) + round256(sizeof(State));
}

inline Work carve(char* base, size_t n)
{
    Work w;
    char* p = base;
    auto take = [&p](size_t bytes) {
        char* q = p;
        p += round256(bytes);
        return q;
    };
    w.minv = reinterpret_cast<double*>(take(9 * n * sizeof(double)));
    w.r = reinterpret_cast<double*>(take(3 * n * sizeof(double)));
    w.z = reinterpret_cast<double*>(take(3 * n * sizeof(double)));
    w.d = reinterpret_cast<double*>(take(3 * n * sizeof(double)));
    w.q = reinterpret_cast<double*>(take(3 * n * sizeof(double)));
    w.pdq = reinterpret_cast<double*>(take(spmv_groups(n) * sizeof(double)));
    w.prz = reinterpret_cast<double*>(take(vec_groups(n) * sizeof(double)));
    w.prr = reinterpret_cast<double*>(take(vec_groups(n) * sizeof(double)));
    w.pflag = reinterpret_cast<int*>(take(vec_groups(n) * sizeof(int)));
    w.st = reinterpret_cast<State*>(take(sizeof(State)));
    return w;
}

// The stop index of the state, read once per workgroup (so every wave of the
// workgroup takes the same branch).
__device__ inline int stop_before(const State* st)
{
    __shared__ int s;
    if (threadIdx.x == 0)
        s = __atomic_load_n(&st->stopBefore, __ATOMIC_RELAXED);
    __syncthreads();
    const int v = s;
    __syncthreads();
    return v;
}

__device__ inline int or_flags(const int* p, int count)
{
    int f = 0;
    for (int i = threadIdx.x; i < count; i += NT)
        f |= p[i];
    return __syncthreads_or(f);
}

// z = M v with M row-major (Mat3 * Vec3 sums left to right, Mat3.h:40-47).
__device__ inline void mat_vec(const double* M, const double* v, double* z)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        z[i] = M[3 * i] * v[0] + M[3 * i + 1] * v[1] + M[3 * i + 2] * v[2];
}

__device__ inline double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Block-Jacobi inverse (BuildInversePreconditioner, LinearSolver.cpp:30-85),
// x = 0, r = b, z = d = M^-1 r and the (r.z, r.r) partials. Flags: 1 a
// non-finite b, 2 a non-finite M^-1 r, 4 the preconditioner refused.
__global__ __launch_bounds__(NT) void k_lin_init(int n, const int* __restrict__ rs, const int* __restrict__ col,
                                                 const double* __restrict__ vals, const double* __restrict__ b,
                                                 int jacobi, Work w, double* __restrict__ x)
{
    const int row0 = blockIdx.x * kVecRows;
    double rz = 0.0, rr = 0.0, unused = 0.0;
    int flag = 0;
    for (int i = threadIdx.x; i < kVecRows; i += NT)
    {
        const int row = row0 + i;
        if (row >= n)
            break;
        double M[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
        if (jacobi)
        {
            int e = -1;
            for (int k = rs[row]; k < rs[row + 1]; ++k)
                if (col[k] == row)
                {
                    e = k;
                    break;
                }
            if (e < 0)
                flag |= 4;
            else
            {
                double D[9];
#pragma unroll
                for (int j = 0; j < 9; ++j)
                    D[j] = vals[9 * static_cast<size_t>(e) + j];
                const double det = det3(D);
                if (!fin9(D) || !fin(det) || fabs(det) <= static_cast<double>(1.0e-8f))
                    flag |= 4;
                else
                {
                    inv3(D, M);
                    if (!fin9(M))
                        flag |= 4;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 9; ++j)
            w.minv[9 * static_cast<size_t>(row) + j] = M[j];
        double r[3], z[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            r[j] = b[3 * static_cast<size_t>(row) + j];
            x[3 * static_cast<size_t>(row) + j] = 0.0;
        }
        if (!(fin(r[0]) && fin(r[1]) && fin(r[2])))
            flag |= 1;
        mat_vec(M, r, z);
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag |= 2;
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            w.r[3 * static_cast<size_t>(row) + j] = r[j];
            w.z[3 * static_cast<size_t>(row) + j] = z[j];
            w.d[3 * static_cast<size_t>(row) + j] = z[j];
        }
        rz += dot3d(r, z);
        rr += dot3d(r, r);
    }
    block_sum3(rz, rr, unused);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
    {
        w.prz[blockIdx.x] = rz;
        w.prr[blockIdx.x] = rr;
        w.pflag[blockIdx.x] = flag;
    }
}

// The checks before the loop (ConjugateGradientSolver.cpp:136-176).
__global__ __launch_bounds__(NT) void k_lin_begin(int groups, Work w, double tol)
{
    const double rz = sum_doubles(w.prz, 0, groups);
    const double rr = sum_doubles(w.prr, 0, groups);
    const int flag = or_flags(w.pflag, groups);
    if (threadIdx.x != 0)
        return;
    State& s = *w.st;
    s.rr = rr;
    s.iters = 0;
    s.converged = 0;
    s.residual = 0.0;
    s.deltaLast = rz;
    s.delta[0] = rz;
    s.delta[1] = 0.0;
    s.target = 0.0;
    s.stopBefore = 0;
    if (flag & 5)
    {
        s.outcome = kRefused;
        return;
    }
    if ((flag & 2) || !(fin(rz) && rz >= 0.0))
    {
        s.outcome = kStoppedAtStart;
        return;
    }
    s.target = tol * tol * rz;
    if (!fin(s.target))
    {
        s.outcome = kStoppedAtStart;
        return;
    }
    if (rz <= s.target)
    {
        s.outcome = kConvergedAtStart;
        s.converged = 1;
        s.residual = norm_of(rr);
        return;
    }
    s.outcome = kLoop;
    s.stopBefore = INT_MAX;
}

// q = A d with 4 lanes per block row (lane-strided partial sums, then a
// butterfly: a fixed order), and the workgroup's d.q partial.
__global__ __launch_bounds__(NT) void k_lin_spmv(int n, const int* __restrict__ rs, const int* __restrict__ col,
                                                 const double* __restrict__ vals, Work w, int k)
{
    if (stop_before(w.st) <= k)
        return;
    const int lane = threadIdx.x & 3;
    const int slot = threadIdx.x >> 2;
    double dq = 0.0, u1 = 0.0, u2 = 0.0;
    for (int pass = 0; pass < kSpmvRows / (NT / 4); ++pass)
    {
        const int row = blockIdx.x * kSpmvRows + pass * (NT / 4) + slot;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (row < n)
            for (int e = rs[row] + lane; e < rs[row + 1]; e += 4)
            {
                const double* B = vals + 9 * static_cast<size_t>(e);
                const double* dc = w.d + 3 * static_cast<size_t>(col[e]);
                const double d0 = dc[0], d1 = dc[1], d2 = dc[2];
                a0 += B[0] * d0 + B[1] * d1 + B[2] * d2;
                a1 += B[3] * d0 + B[4] * d1 + B[5] * d2;
                a2 += B[6] * d0 + B[7] * d1 + B[8] * d2;
            }
        a0 += __shfl_xor(a0, 1, 64);
        a1 += __shfl_xor(a1, 1, 64);
        a2 += __shfl_xor(a2, 1, 64);
        a0 += __shfl_xor(a0, 2, 64);
        a1 += __shfl_xor(a1, 2, 64);
        a2 += __shfl_xor(a2, 2, 64);
        if (lane == 0 && row < n)
        {
            double* q = w.q + 3 * static_cast<size_t>(row);
            const double* d = w.d + 3 * static_cast<size_t>(row);
            q[0] = a0;
            q[1] = a1;
            q[2] = a2;
            dq += d[0] * a0 + d[1] * a1 + d[2] * a2;
        }
    }
    block_sum3(dq, u1, u2);
    if (threadIdx.x == 0)
        w.pdq[blockIdx.x] = dq;
}

// alpha, x += alpha d, r -= alpha q, z = M^-1 r and the (r.z, r.r) partials
// (ConjugateGradientSolver.cpp:186-209).
__global__ __launch_bounds__(NT) void k_lin_update(int n, int spmvGroups, Work w, int k, double* __restrict__ x)
{
    if (stop_before(w.st) <= k)
        return;
    const double dq = sum_doubles(w.pdq, 0, spmvGroups);
    const double alpha = w.st->delta[k & 1] / dq;
    if (!(fin(dq) && dq > 0.0) || !fin(alpha))
    {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            w.st->stopBefore = k;  // this iteration breaks down before any update
        return;
    }
    const int row0 = blockIdx.x * kVecRows;
    double rz = 0.0, rr = 0.0, unused = 0.0;
    int flag = 0;
    for (int i = threadIdx.x; i < kVecRows; i += NT)
    {
        const int row = row0 + i;
        if (row >= n)
            break;
        const size_t o = 3 * static_cast<size_t>(row);
        double r[3], z[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            x[o + j] = x[o + j] + w.d[o + j] * alpha;
            r[j] = w.r[o + j] + w.q[o + j] * -alpha;
            w.r[o + j] = r[j];
        }
        mat_vec(w.minv + 9 * static_cast<size_t>(row), r, z);
        if (!(fin(z[0]) && fin(z[1]) && fin(z[2])))
            flag = 1;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            w.z[o + j] = z[j];
        rz += dot3d(r, z);
        rr += dot3d(r, r);
    }
    block_sum3(rz, rr, unused);
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0)
    {
        w.prz[blockIdx.x] = rz;
        w.prr[blockIdx.x] = rr;
        w.pflag[blockIdx.x] = flag;
    }
}

// delta, beta, d = z + beta d, and the loop's bookkeeping
// (ConjugateGradientSolver.cpp:205-231).
__global__ __launch_bounds__(NT) void k_lin_dir(int n, int vecGroups, Work w, int k, int maxIt)
{
    if (stop_before(w.st) <= k)
        return;
    const double rz = sum_doubles(w.prz, 0, vecGroups);
    const double rr = sum_doubles(w.prr, 0, vecGroups);
    const int flag = or_flags(w.pflag, vecGroups);
    const double deltaOld = w.st->delta[k & 1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (flag)
    {
        // the preconditioned residual is not finite: the loop ends with r
        // (and x) updated and delta unchanged
        if (lead)
        {
            w.st->rr = rr;
            w.st->stopBefore = k + 1;
        }
        return;
    }
    const double beta = rz / deltaOld;
    if (!fin(rz) || !fin(deltaOld) || deltaOld <= 0.0 || !fin(beta))
    {
        if (lead)
        {
            w.st->rr = rr;
            w.st->deltaLast = rz;
            w.st->stopBefore = k + 1;
        }
        return;
    }
    const int row0 = blockIdx.x * kVecRows;
    for (int i = threadIdx.x; i < kVecRows; i += NT)
    {
        const int row = row0 + i;
        if (row >= n)
            break;
        const size_t o = 3 * static_cast<size_t>(row);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            w.d[o + j] = w.z[o + j] + w.d[o + j] * beta;
    }
    if (lead)
    {
        State& s = *w.st;
        s.rr = rr;
        s.deltaLast = rz;
        s.delta[(k + 1) & 1] = rz;
        s.iters = k + 1;
        if (rz <= s.target || k + 1 >= maxIt)
            s.stopBefore = k + 1;
    }
}

// The result after the loop (ConjugateGradientSolver.cpp:234-240).
__global__ void k_lin_finish(Work w)
{
    if (threadIdx.x != 0)
        return;
    State& s = *w.st;
    if (s.outcome != kLoop)
        return;
    s.converged = s.deltaLast <= s.target ? 1 : 0;
    s.residual = norm_of(s.rr);
}

}  // namespace lin

// The solve with its work arrays at `base` (lin::work_bytes(rows) bytes of
// device memory).
int linear_solve_on(char* base, int rows, const int* rs, const int* col, const double* vals, const double* b,
                    double* x, int maxIterations, double tol, int jacobi, hipStream_t stream, pk_linear_result* res)
{
    using namespace lin;
    const size_t n = static_cast<size_t>(rows);
    const Work w = carve(base, n);
    const int vg = static_cast<int>(vec_groups(n));
    const int sg = static_cast<int>(spmv_groups(n));
    const int maxIt = maxIterations > 0 ? maxIterations : 3 * rows;
    hipLaunchKernelGGL(k_lin_init, dim3(vg), dim3(NT), 0, stream, rows, rs, col, vals, b, jacobi, w, x);
    hipLaunchKernelGGL(k_lin_begin, dim3(1), dim3(NT), 0, stream, vg, w, tol);
    int stop = 0;
    if (hipMemcpyAsync(&stop, &w.st->stopBefore, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
    {
        set_error(std::string("linear solve: ") + hipGetErrorString(hipGetLastError()));
        return PK_ERR_HIP;
    }
    for (int k0 = 0; stop > k0 && k0 < maxIt; k0 += kChunk)
    {
        for (int k = k0; k < std::min(k0 + kChunk, maxIt); ++k)
        {
            hipLaunchKernelGGL(k_lin_spmv, dim3(sg), dim3(NT), 0, stream, rows, rs, col, vals, w, k);
            hipLaunchKernelGGL(k_lin_update, dim3(vg), dim3(NT), 0, stream, rows, sg, w, k, x);
            hipLaunchKernelGGL(k_lin_dir, dim3(vg), dim3(NT), 0, stream, rows, vg, w, k, maxIt);
        }
        if (hipMemcpyAsync(&stop, &w.st->stopBefore, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
        {
            set_error(std::string("linear solve: ") + hipGetErrorString(hipGetLastError()));
            return PK_ERR_HIP;
        }
    }
    hipLaunchKernelGGL(k_lin_finish, dim3(1), dim3(64), 0, stream, w);
    State st;
    if (hipMemcpyAsync(&st, w.st, sizeof(State), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess || hipGetLastError() != hipSuccess)
    {
        set_error("linear solve: a HIP call failed");
        return PK_ERR_HIP;
    }
    res->iterations = st.iters;
    res->converged = st.converged;
    res->residual_norm = st.residual;
    res->solution_valid = st.outcome == kRefused ? 0 : 1;
    return PK_OK;
}

// Argument checks shared by the entry points; returns true when the solve
// is decided without any GPU work (res filled).
bool linear_trivial(int rows, double tol, pk_linear_result* res, int* status)
{
    *res = pk_linear_result{};
    *status = PK_OK;
    if (rows < 0)
    {
        set_error("linear solve: rows must be >= 0");
        *status = PK_ERR_ARG;
        return true;
    }
    if (rows == 0)
    {
        res->converged = 1;  // an empty system (ConjugateGradientSolver.cpp:126-134)
        res->solution_valid = 1;
        return true;
    }
    if (!std::isfinite(tol) || tol <= 0.0)
        return true;  // refused: the reference clears the solution (:136-148)
    return false;
}

}  // namespace pk

// The work arrays of pk_bsr_pcg_noalloc: one copy per GPU, held until the
// process ends. Host code never passes `scratch` itself (its host-side name
// is not a device address); it uses the address hipGetSymbolAddress gives on
// each device.
constexpr size_t kNoallocBytes = pk::lin::work_bytes(PK_LINEAR_NOALLOC_MAX_ROWS);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace pk
{
namespace
{
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device (with it current) and kept.
char* scratch_on_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    {
        set_error("pk_bsr_pcg_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev])
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            set_error(std::string("pk_bsr_pcg_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = static_cast<char*>(p);
    }
    return g_scratch[dev];
}
}  // namespace
}  // namespace pk

using namespace pk;

extern "C" {

int pk_bsr_pcg_noalloc(int rows, const int* row_start, const int* col, const double* vals, const double* b,
                       double* x, int max_iterations, double tolerance, int use_jacobi, void* stream,
                       pk_linear_result* result)
{
    if (!result)
    {
        set_error("pk_bsr_pcg_noalloc: result is NULL");
        return PK_ERR_ARG;
    }
    if (rows > PK_LINEAR_NOALLOC_MAX_ROWS)
    {
        *result = pk_linear_result{};
        set_error("pk_bsr_pcg_noalloc: " + std::to_string(rows) + " rows exceed the " +
                  std::to_string(PK_LINEAR_NOALLOC_MAX_ROWS) + " its work arrays hold (use pk_bsr_pcg)");
        return PK_ERR_ARG;
    }
    int status;
    if (linear_trivial(rows, tolerance, result, &status))
        return status;
    if (!row_start || !col || !vals || !b || !x)
    {
        set_error("pk_bsr_pcg_noalloc: a device pointer is NULL");
        return PK_ERR_ARG;
    }
    char* base = scratch_on_current_device();
    if (!base)
        return PK_ERR_HIP;
    return linear_solve_on(base, rows, row_start, col, vals, b, x, max_iterations, tolerance, use_jacobi,
                           static_cast<hipStream_t>(stream), result);
}

int pk_bsr_pcg(int rows, const int* row_start, const int* col, const double* vals, const double* b, double* x,
               int max_iterations, double tolerance, int use_jacobi, void* stream, pk_linear_result* result)
{
    if (!result)
    {
        set_error("pk_bsr_pcg: result is NULL");
        return PK_ERR_ARG;
    }
    int status;
    if (linear_trivial(rows, tolerance, result, &status))
        return status;
    if (!row_start || !col || !vals || !b || !x)
    {
        set_error("pk_bsr_pcg: a device pointer is NULL");
        return PK_ERR_ARG;
    }
    char* base = nullptr;
    const size_t bytes = lin::work_bytes(static_cast<size_t>(rows));
    if (hipMalloc(reinterpret_cast<void**>(&base), bytes) != hipSuccess)
    {
        set_error("pk_bsr_pcg: allocating " + std::to_string(bytes) + " bytes of work arrays failed");
        return PK_ERR_HIP;
    }
    const int rc = linear_solve_on(base, rows, row_start, col, vals, b, x, max_iterations, tolerance, use_jacobi,
                                   static_cast<hipStream_t>(stream), result);
    (void)hipFree(base);
    return rc;
}

int pk_linear_solve(int block_count, const int* row_start, const int* col_index, const float* values,
                    const float* rhs, float* solution, int max_iterations, float tolerance, int use_jacobi,
                    pk_linear_result* result)
{
    if (!result)
    {
        set_error("pk_linear_solve: result is NULL");
        return PK_ERR_ARG;
    }
    int status;
    if (linear_trivial(block_count, tolerance, result, &status))
        return status;
    if (!row_start || !col_index || !rhs || !solution || row_start[0] != 0 || row_start[block_count] < 0 ||
        (row_start[block_count] > 0 && !values))
    {
        set_error("pk_linear_solve: missing arrays or bad row offsets");
        return PK_ERR_ARG;
    }
    const int n = block_count;
    const int stored = row_start[n];
    for (int r = 0; r < n; ++r)
        if (row_start[r + 1] < row_start[r])
        {
            set_error("pk_linear_solve: row offsets decrease");
            return PK_ERR_ARG;
        }
    for (int k = 0; k < stored; ++k)
        if (col_index[k] < 0 || col_index[k] >= n)
        {
            set_error("pk_linear_solve: a block column is out of range");
            return PK_ERR_ARG;
        }
    // Both triangles in the pk_* layout: a stored block B at (r, c) is
    // applied as (r, c, B) and, off the diagonal, (c, r, B^T), as
    // SparseBlockMatrix::Multiply applies it; each row sorted by column.
    std::vector<int> count(static_cast<size_t>(n) + 1, 0);
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
        {
            ++count[r + 1];
            if (col_index[k] != r)
                ++count[col_index[k] + 1];
        }
    for (int r = 0; r < n; ++r)
        count[r + 1] += count[r];
    const size_t full = static_cast<size_t>(count[n]);
    std::vector<int> fcol(full), fill(count.begin(), count.end() - 1);
    std::vector<int> src(full);  // stored block k, negated minus 1 when transposed
    for (int r = 0; r < n; ++r)
        for (int k = row_start[r]; k < row_start[r + 1]; ++k)
        {
            const int c = col_index[k];
            fcol[fill[r]] = c;
            src[fill[r]++] = k;
            if (c != r)
            {
                fcol[fill[c]] = r;
                src[fill[c]++] = -1 - k;
            }
        }
    std::vector<double> fvals(full * 9);
    std::vector<int> order;
    for (int r = 0; r < n; ++r)
    {
        const int a = count[r], e = count[r + 1];
        order.resize(static_cast<size_t>(e - a));
        for (int i = a; i < e; ++i)
            order[i - a] = i;
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return fcol[u] < fcol[v]; });
        std::vector<int> cs(order.size()), ss(order.size());
        for (size_t i = 0; i < order.size(); ++i)
        {
            cs[i] = fcol[order[i]];
            ss[i] = src[order[i]];
        }
        for (size_t i = 0; i < order.size(); ++i)
        {
            fcol[a + i] = cs[i];
            src[a + i] = ss[i];
        }
    }
    for (size_t i = 0; i < full; ++i)
    {
        const bool t = src[i] < 0;
        const float* m = values + 9 * static_cast<size_t>(t ? -1 - src[i] : src[i]);  // column-major Mat3
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 3; ++c)
                fvals[9 * i + 3 * row + c] = t ? m[3 * row + c] : m[3 * c + row];
    }
    std::vector<double> bd(static_cast<size_t>(n) * 3), xd(static_cast<size_t>(n) * 3);
    for (size_t i = 0; i < bd.size(); ++i)
        bd[i] = rhs[i];
    int *drs = nullptr, *dcol = nullptr;
    double *dvals = nullptr, *db = nullptr, *dx = nullptr;
    char* base = nullptr;
    int rc = PK_ERR_HIP;
    const size_t bytes = lin::work_bytes(static_cast<size_t>(n));
    if (hipMalloc(reinterpret_cast<void**>(&drs), count.size() * sizeof(int)) == hipSuccess &&
        hipMalloc(reinterpret_cast<void**>(&dcol), std::max<size_t>(full, 1) * sizeof(int)) == hipSuccess &&
        hipMalloc(reinterpret_cast<void**>(&dvals), std::max<size_t>(full, 1) * 9 * sizeof(double)) == hipSuccess &&
        hipMalloc(reinterpret_cast<void**>(&db), bd.size() * sizeof(double)) == hipSuccess &&
        hipMalloc(reinterpret_cast<void**>(&dx), xd.size() * sizeof(double)) == hipSuccess &&
        hipMalloc(reinterpret_cast<void**>(&base), bytes) == hipSuccess &&
        hipMemcpy(drs, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
        (full == 0 || hipMemcpy(dcol, fcol.data(), full * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
        (full == 0 ||
         hipMemcpy(dvals, fvals.data(), fvals.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) &&
        hipMemcpy(db, bd.data(), bd.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess)
    {
        rc = linear_solve_on(base, n, drs, dcol, dvals, db, dx, max_iterations, tolerance, use_jacobi, nullptr,
                             result);
        if (rc == PK_OK && result->solution_valid)
        {
            if (hipMemcpy(xd.data(), dx, xd.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                rc = PK_ERR_HIP;
            else
                for (size_t i = 0; i < xd.size(); ++i)
                    solution[i] = static_cast<float>(xd[i]);
        }
    }
    else
        set_error("pk_linear_solve: a device allocation or copy failed");
    for (void* p : {static_cast<void*>(drs), static_cast<void*>(dcol), static_cast<void*>(dvals),
                    static_cast<void*>(db), static_cast<void*>(dx), static_cast<void*>(base)})
        if (p)
            (void)hipFree(p);
    return rc;
}

}  // extern "C"
