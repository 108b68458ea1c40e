// Diagnostics library (libphysik_diag.so, not part of the product library):
// a plain streaming copy kernel of known size, used to calibrate the HBM byte
// counters (FETCH_SIZE, WRITE_SIZE) against the bytes a kernel must move.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "pk_diag.h"

namespace pk
{
namespace
{
thread_local std::string g_diag_error;
void set_error(const std::string& msg) { g_diag_error = msg; }
}  // namespace

__global__ __launch_bounds__(256) void k_copy(const double4* __restrict__ src, double4* __restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256ull)
        dst[i] = src[i];
}
}  // namespace pk

extern "C" const char* pk_diag_last_error(void) { return pk::g_diag_error.c_str(); }

extern "C" int pk_diag_copy(long long bytes, int reps, double* avg_ms)
{
    using namespace pk;
    if (bytes <= 0 || reps <= 0 || bytes % 32)
    {
        set_error("pk_diag_copy: bytes must be a positive multiple of 32");
        return 1;
    }
    const size_t n = static_cast<size_t>(bytes) / 32;
    double4 *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess)
    {
        if (a) (void)hipFree(a);
        set_error("pk_diag_copy: allocation failed");
        return 2;
    }
    (void)hipMemset(a, 0, bytes);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n);  // warm
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (avg_ms)
        *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
