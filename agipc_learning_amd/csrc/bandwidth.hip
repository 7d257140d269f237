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
constexpr int kUnroll  = 4;  // float4 loads in flight per thread
constexpr int kRounds  = 4;  // rounds of kUnroll loads per workgroup (read stream)

typedef float f4v __attribute__((ext_vector_type(4)));

// Workgroup b reads the kThreads * kUnroll * kRounds float4 from
// b * kThreads * kUnroll * kRounds on, in kRounds rounds with kUnroll
// nontemporal 16-byte loads in flight per thread (each round a contiguous
// span).  Of the read streams tried on this card (grid-stride loops of 4 to
// 64 workgroups per CU, 4 to 16 loads in flight, 256 to 1024 threads,
// ordinary or nontemporal loads, contiguous spans per workgroup), this one
// reads fastest.  The sum is stored only if it equals a value the caller's
// data never produces, so the loads stay.
__global__ void __launch_bounds__(kThreads) k_stream_read(const float4* __restrict__ a, long long n4, float* sink)
{
    constexpr long long kPer = (long long)kThreads * kUnroll * kRounds;
    const long long     base = (long long)blockIdx.x * kPer + threadIdx.x;
    float               s    = 0.f;
#pragma unroll
    for(int r = 0; r < kRounds; ++r)
    {
        const long long i0 = base + (long long)r * kThreads * kUnroll;
        f4v v[kUnroll];
#pragma unroll
        for(int u = 0; u < kUnroll; ++u)
        {
            const long long i = i0 + (long long)u * kThreads;
            v[u] = i < n4 ? __builtin_nontemporal_load(reinterpret_cast<const f4v*>(a + i)) : f4v{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for(int u = 0; u < kUnroll; ++u) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
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
        constexpr long long kPer = (long long)kThreads * kUnroll * kRounds;
        const unsigned  grid_read = (unsigned)((n4 + kPer - 1) / kPer);
        auto launch = [&](bool timed) {
            if(kind == 0)
            {
                if(timed)
                    hipExtLaunchKernelGGL(k_stream_read, dim3(grid_read), dim3(kThreads), 0, st, e0, e1, 0, a, n4, sink);
                else hipLaunchKernelGGL(k_stream_read, dim3(grid_read), dim3(kThreads), 0, st, a, n4, sink);
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
