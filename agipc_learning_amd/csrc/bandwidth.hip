// Practical memory-rate ceilings of the card, measured with plain streaming
// kernels: a read-only stream (what the iteration kernels mostly do: they
// read matrices and inverses and write vectors) and a copy.  Reported next to
// the solver kernels' rates so their fractions of the nominal HBM peak can be
// read against what a simple kernel reaches on the same card.
#include "bandwidth.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace masx
{
namespace
{
#define BW_HIP(x)                                                                                    \
    do                                                                                               \
    {                                                                                                \
        const hipError_t e_ = (x);                                                                   \
        if(e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while(0)

constexpr int kThreads = 256;
constexpr int kUnroll  = 4;

// Each thread sums kUnroll float4 per step of a grid-stride loop (all loads
// of a step in flight at once); the sum is stored only if it equals a value
// the caller's data never produces, so the loads stay.
__global__ void __launch_bounds__(kThreads) k_stream_read(const float4* __restrict__ a, long long n4, float* sink)
{
    float           s      = 0.f;
    const long long stride = (long long)gridDim.x * kThreads;
    long long       i      = (long long)blockIdx.x * kThreads + threadIdx.x;
    for(; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride)
    {
        float4 v[kUnroll];
#pragma unroll
        for(int u = 0; u < kUnroll; ++u) v[u] = a[i + u * stride];
#pragma unroll
        for(int u = 0; u < kUnroll; ++u) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for(; i < n4; i += stride)
    {
        const float4 v = a[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    if(s == -1.0f) sink[0] = s;  // the buffer holds zeros
}

__global__ void __launch_bounds__(kThreads) k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                                          long long n4)
{
    const long long stride = (long long)gridDim.x * kThreads;
    for(long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride) b[i] = a[i];
}
}  // namespace

double stream_rate(long long bytes, int kind, int reps)
{
    if(bytes < (1 << 20) || bytes % 16 != 0) throw std::invalid_argument("stream rate: bytes >= 1 MiB, multiple of 16");
    if(kind != 0 && kind != 1) throw std::invalid_argument("stream rate: kind 0 (read) or 1 (copy)");
    if(reps < 1) throw std::invalid_argument("stream rate: reps >= 1");
    int dev = 0, cus = 0;
    BW_HIP(hipGetDevice(&dev));
    BW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    float4 *a = nullptr, *b = nullptr;
    float*  sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    double      best = 0.0;
    try
    {
        BW_HIP(hipMalloc((void**)&a, (size_t)bytes));
        if(kind == 1) BW_HIP(hipMalloc((void**)&b, (size_t)bytes));
        BW_HIP(hipMalloc((void**)&sink, sizeof(float)));
        BW_HIP(hipMemset(a, 0, (size_t)bytes));
        BW_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        BW_HIP(hipEventCreate(&e0));
        BW_HIP(hipEventCreate(&e1));
        const long long n4   = bytes / 16;
        const unsigned  grid = (unsigned)std::min<long long>((long long)cus * 16, (n4 + kThreads - 1) / kThreads);
        auto launch = [&](bool timed) {
            if(kind == 0)
            {
                if(timed) hipExtLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kThreads), 0, st, e0, e1, 0, a, n4, sink);
                else hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kThreads), 0, st, a, n4, sink);
            }
            else
            {
                if(timed) hipExtLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kThreads), 0, st, e0, e1, 0, a, b, n4);
                else hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kThreads), 0, st, a, b, n4);
            }
            BW_HIP(hipGetLastError());
        };
        for(int i = 0; i < 3; ++i) launch(false);
        for(int i = 0; i < reps; ++i)
        {
            launch(true);
            BW_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            BW_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double moved = (double)bytes * (kind == 1 ? 2.0 : 1.0);
            if(ms > 0.f) best = std::max(best, moved / (ms * 1e-3) / 1e9);
        }
    }
    catch(...)
    {
        if(e0) (void)hipEventDestroy(e0);
        if(e1) (void)hipEventDestroy(e1);
        if(st) (void)hipStreamDestroy(st);
        if(a) (void)hipFree(a);
        if(b) (void)hipFree(b);
        if(sink) (void)hipFree(sink);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    if(b) (void)hipFree(b);
    (void)hipFree(sink);
    return best;
}
}  // namespace masx
