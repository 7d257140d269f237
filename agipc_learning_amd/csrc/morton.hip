// Morton-order vertex partition on the device (gl_main.cu:1176-1243): the
// bounding box, a 30-bit code per vertex (10 bits per axis, x | y << 1 |
// z << 2), vertices ordered by (code, index), banks padded with -1.
//
// The order is a stable least-significant-digit radix sort of the codes
// (four passes of 8 bits), so ties keep the index order without a key
// carrying the index, and every input distribution costs the same (a
// bucket-and-rank scheme would go quadratic on clustered points).  Per pass:
// per-tile digit counts, one exclusive scan over (digit, tile) in digit-major
// order, then a stable scatter in which each wavefront ranks equal digits by
// eight ballots.  Integer work, latency-bound at mesh sizes.
#include "morton.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace masx
{
namespace
{

constexpr int kThreads = 256;
constexpr int kItems   = 8;
constexpr int kTile    = kThreads * kItems;   // items per scatter tile
constexpr int kBins    = 256;                  // 8-bit digits
constexpr int kPasses  = 4;                    // 32 key bits (codes use 30)
constexpr int kScan    = 1024;                 // threads of the single-workgroup scan
constexpr int kBoxMax  = 1024;                 // partial bounding boxes

#define MT_HIP(x)                                                                                    \
    do                                                                                               \
    {                                                                                                \
        const hipError_t e_ = (x);                                                                   \
        if(e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while(0)

// per-block partial bounds: lo xyz, hi xyz
__global__ void __launch_bounds__(kThreads) k_mt_box_partial(int n, const double* __restrict__ pos, double* part)
{
    __shared__ double s[6][kThreads];
    const int t = threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(long long v = (long long)blockIdx.x * kThreads + t; v < n; v += (long long)gridDim.x * kThreads)
#pragma unroll
        for(int d = 0; d < 3; ++d)
        {
            const double p = pos[3 * v + d];
            lo[d] = fmin(lo[d], p);
            hi[d] = fmax(hi[d], p);
        }
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        s[d][t]     = lo[d];
        s[3 + d][t] = hi[d];
    }
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1)
    {
        if(t < off)
#pragma unroll
            for(int d = 0; d < 3; ++d)
            {
                s[d][t]     = fmin(s[d][t], s[d][t + off]);
                s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + off]);
            }
        __syncthreads();
    }
    if(t < 6) part[6 * blockIdx.x + t] = s[t][0];
}

// the partials -> the box (written after the partials, at part + 6 * g)
__global__ void __launch_bounds__(kThreads) k_mt_box_final(int g, double* part)
{
    __shared__ double s[6][kThreads];
    const int t = threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(int b = t; b < g; b += kThreads)
#pragma unroll
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = fmin(lo[d], part[6 * b + d]);
            hi[d] = fmax(hi[d], part[6 * b + 3 + d]);
        }
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        s[d][t]     = lo[d];
        s[3 + d][t] = hi[d];
    }
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1)
    {
        if(t < off)
#pragma unroll
            for(int d = 0; d < 3; ++d)
            {
                s[d][t]     = fmin(s[d][t], s[d][t + off]);
                s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + off]);
            }
        __syncthreads();
    }
    if(t < 6) part[6 * g + t] = s[t][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    uint32_t out = 0;
#pragma unroll
    for(int b = 0; b < 10; ++b) out |= ((v >> b) & 1u) << (3 * b);
    return out;
}

// code per vertex: t = clamp((p - lo) / (hi - lo), 0, 1) per axis (0 for a
// flat axis), q = (uint)(t * 1023 + 0.5) with the multiply and the add
// rounded separately (as the host forms it), 10 bits interleaved
__global__ void __launch_bounds__(kThreads) k_mt_codes(int n, const double* __restrict__ pos,
                                                       const double* __restrict__ box, uint32_t* key, int* idx)
{
    const int v = blockIdx.x * kThreads + threadIdx.x;
    if(v >= n) return;
    uint32_t code = 0;
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        const double lo = box[d], e = __dsub_rn(box[3 + d], lo);
        double t = 0.0;
        if(e > 0.0) t = fmin(fmax(__ddiv_rn(__dsub_rn(pos[3 * (long long)v + d], lo), e), 0.0), 1.0);
        const double q = __dadd_rn(__dmul_rn(t, 1023.0), 0.5);
        code |= spread10((uint32_t)q) << d;
    }
    key[v] = code;
    idx[v] = v;
}

// digit counts of one tile -> hist[digit * n_tiles + tile]
__global__ void __launch_bounds__(kThreads) k_mt_hist(int n, const uint32_t* __restrict__ key, int shift, int* hist,
                                                      int n_tiles)
{
    __shared__ int cnt[kBins];
    const int t = threadIdx.x, tile = blockIdx.x;
    cnt[t] = 0;
    __syncthreads();
#pragma unroll
    for(int j = 0; j < kItems; ++j)
    {
        const long long i = (long long)tile * kTile + j * kThreads + t;
        if(i < n) atomicAdd(&cnt[(key[i] >> shift) & (kBins - 1)], 1);
    }
    __syncthreads();
    hist[(long long)t * n_tiles + tile] = cnt[t];
}

// exclusive scan of a[0..m) in place, one workgroup: each thread sums a
// contiguous chunk, the chunk sums are scanned in LDS, each chunk is rewritten
__global__ void __launch_bounds__(kScan) k_mt_scan(int m, int* a)
{
    __shared__ int s[kScan];
    const int t = threadIdx.x;
    const int chunk = (m + kScan - 1) / kScan;
    const int b = min(m, t * chunk), e = min(m, b + chunk);
    int sum = 0;
    for(int i = b; i < e; ++i) sum += a[i];
    s[t] = sum;
    __syncthreads();
    for(int off = 1; off < kScan; off <<= 1)
    {
        const int v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = s[t] - sum;
    for(int i = b; i < e; ++i)
    {
        const int v = a[i];
        a[i] = run;
        run += v;
    }
}

// stable scatter of one tile: items in rounds of 256 consecutive indices;
// inside a round, a lane's rank among equal digits comes from eight ballots
// (lanes of its wavefront) plus the counts of the earlier wavefronts, and
// `running` carries each digit's count over the rounds
__global__ void __launch_bounds__(kThreads) k_mt_scatter(int n, const uint32_t* __restrict__ key_in,
                                                         const int* __restrict__ idx_in, uint32_t* key_out,
                                                         int* idx_out, int shift, const int* __restrict__ offs,
                                                         int n_tiles)
{
    __shared__ int running[kBins];
    __shared__ int wcnt[kThreads / 64][kBins];
    const int t = threadIdx.x, tile = blockIdx.x, lane = t & 63, wave = t >> 6;
    running[t] = 0;
    const int base = offs[(long long)t * n_tiles + tile];  // this tile's start for digit t
    __shared__ int sbase[kBins];
    sbase[t] = base;
    for(int j = 0; j < kItems; ++j)
    {
#pragma unroll
        for(int w = 0; w < kThreads / 64; ++w) wcnt[w][t] = 0;
        __syncthreads();
        const long long i = (long long)tile * kTile + j * kThreads + t;
        const bool valid = i < n;
        const uint32_t k = valid ? key_in[i] : 0u;
        const int d = (int)((k >> shift) & (kBins - 1));
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for(int b = 0; b < 8; ++b)
        {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const unsigned long long below = peers & ((1ull << lane) - 1ull);
        if(valid && below == 0ull) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if(valid)
        {
            int pre = running[d];
            for(int w = 0; w < wave; ++w) pre += wcnt[w][d];
            const int p = sbase[d] + pre + __popcll(below);
            key_out[p] = k;
            idx_out[p] = idx_in[i];
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for(int w = 0; w < kThreads / 64; ++w) add += wcnt[w][t];
        running[t] += add;
    }
}

__global__ void __launch_bounds__(kThreads) k_mt_maps(int n, int padded, const int* __restrict__ order, int* p2r,
                                                      int* r2p)
{
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if(s >= padded) return;
    const int v = s < n ? order[s] : -1;
    p2r[s] = v;
    if(s < n) r2p[v] = s;
}

template <typename T>
void grow(T*& p, size_t& cap, size_t need)
{
    if(need <= cap && p) return;
    if(p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    MT_HIP(hipMalloc((void**)&p, need * sizeof(T)));
    cap = need;
}

// Where one partition keeps its intermediate arrays.
struct MtScratch
{
    double*   box;      // 6 * (g + 1): per-block partial bounds, then the final 6
    uint32_t* key[2];   // n each, ping-pong
    int*      idx[2];   // n each, ping-pong
    int*      hist;     // kBins * n_tiles, scanned in place
};

struct MtShape
{
    long long padded;
    int       n_tiles, g;
};

MtShape mt_shape(int n, const double* pos, int bank, const int* part_to_real, long long part_cap,
                 const int* real_to_part)
{
    if(n < 1 || bank < 1 || !pos || !part_to_real || !real_to_part)
        throw std::invalid_argument("morton partition: n >= 1, bank >= 1 and device arrays required");
    MtShape sh;
    sh.padded = ((long long)n + bank - 1) / bank * bank;
    if(part_cap < sh.padded) throw std::invalid_argument("part_to_real buffer smaller than the padded vertex count");
    if(sh.padded > 0x7fffffffLL) throw std::invalid_argument("too many vertices");
    sh.n_tiles = (int)(((long long)n + kTile - 1) / kTile);
    sh.g       = (int)std::min<long long>(kBoxMax, ((long long)n + kThreads - 1) / kThreads);
    return sh;
}

// The launch sequence: every scratch word a kernel reads was written by an
// earlier kernel of the same sequence, so scratch needs no clearing.
void mt_launch(int n, const double* pos, int* part_to_real, int* real_to_part, hipStream_t st, const MtShape& sh,
               const MtScratch& s)
{
    const unsigned blocks = (unsigned)(((long long)n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_mt_box_partial, dim3(sh.g), dim3(kThreads), 0, st, n, pos, s.box);
    MT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mt_box_final, dim3(1), dim3(kThreads), 0, st, sh.g, s.box);
    MT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mt_codes, dim3(blocks), dim3(kThreads), 0, st, n, pos, s.box + 6 * sh.g, s.key[0], s.idx[0]);
    MT_HIP(hipGetLastError());
    int cur = 0;
    for(int p = 0; p < kPasses; ++p)
    {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_mt_hist, dim3(sh.n_tiles), dim3(kThreads), 0, st, n, s.key[cur], shift, s.hist,
                           sh.n_tiles);
        MT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mt_scan, dim3(1), dim3(kScan), 0, st, kBins * sh.n_tiles, s.hist);
        MT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mt_scatter, dim3(sh.n_tiles), dim3(kThreads), 0, st, n, s.key[cur], s.idx[cur],
                           s.key[1 - cur], s.idx[1 - cur], shift, s.hist, sh.n_tiles);
        MT_HIP(hipGetLastError());
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_mt_maps, dim3((unsigned)((sh.padded + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, n,
                       (int)sh.padded, s.idx[cur], part_to_real, real_to_part);
    MT_HIP(hipGetLastError());
}

// ---- the allocation-free variant's scratch ---------------------------------
// A static device array holds the scratch of morton_partition_noalloc: the
// box partials (6 * (kBoxMax + 1) doubles), two key and two index arrays of
// kMortonNoAllocMax words and the digit counts of kMortonNoAllocMax / kTile
// tiles, 8.7 MB in all, kept by every GPU the library runs on until the
// process ends.
constexpr size_t kAlign256  = 256;
constexpr size_t round256(size_t b) { return (b + kAlign256 - 1) / kAlign256 * kAlign256; }
constexpr size_t kNaBoxOff  = 0;
constexpr size_t kNaKeyOff0 = round256(sizeof(double) * 6 * (kBoxMax + 1));
constexpr size_t kNaArr     = round256(sizeof(uint32_t) * (size_t)kMortonNoAllocMax);
constexpr size_t kNaKeyOff1 = kNaKeyOff0 + kNaArr;
constexpr size_t kNaIdxOff0 = kNaKeyOff1 + kNaArr;
constexpr size_t kNaIdxOff1 = kNaIdxOff0 + kNaArr;
constexpr size_t kNaHistOff = kNaIdxOff1 + kNaArr;
constexpr size_t kNaBytes   = kNaHistOff + round256(sizeof(int) * (size_t)kBins * (kMortonNoAllocMax / kTile));
static_assert(kMortonNoAllocMax % kTile == 0, "the static scratch holds whole scatter tiles");

static __device__ __align__(256) uint32_t scratch[kNaBytes / sizeof(uint32_t)];

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mutex;
char*         g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call there.  The array's host-side name is only a handle for the
// runtime: kernels get this address, never the name.
char* noalloc_scratch()
{
    int dev = -1;
    MT_HIP(hipGetDevice(&dev));
    if(dev < 0 || dev >= kMaxDevices) throw std::runtime_error("morton partition: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if(!g_scratch_ptr[dev])
    {
        void*            ptr = nullptr;
        const hipError_t e   = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("morton partition: static scratch address: ") +
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char*>(ptr);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

DeviceMorton::~DeviceMorton()
{
    for(void* p : {(void*)box_, (void*)key_[0], (void*)key_[1], (void*)idx_[0], (void*)idx_[1], (void*)hist_,
                   (void*)part_})
        if(p) (void)hipFree(p);
}

void DeviceMorton::partition(int n, const double* pos, int bank, int* part_to_real, long long part_cap,
                             int* real_to_part, hipStream_t st)
{
    const MtShape sh = mt_shape(n, pos, bank, part_to_real, part_cap, real_to_part);
    grow(box_, box_cap_, 6 * (size_t)(sh.g + 1));
    for(int b = 0; b < 2; ++b)
    {
        grow(key_[b], key_cap_[b], (size_t)n);
        grow(idx_[b], idx_cap_[b], (size_t)n);
    }
    grow(hist_, hist_cap_, (size_t)kBins * sh.n_tiles);
    mt_launch(n, pos, part_to_real, real_to_part, st, sh, MtScratch{box_, {key_[0], key_[1]}, {idx_[0], idx_[1]}, hist_});
}

void morton_partition_noalloc(int n, const double* pos, int bank, int* part_to_real, long long part_cap,
                              int* real_to_part, hipStream_t st)
{
    const MtShape sh = mt_shape(n, pos, bank, part_to_real, part_cap, real_to_part);
    if(n > kMortonNoAllocMax)
        throw std::invalid_argument("morton partition without allocation: at most " +
                                    std::to_string(kMortonNoAllocMax) + " vertices (got " + std::to_string(n) + ")");
    char* base = noalloc_scratch();
    const MtScratch s{reinterpret_cast<double*>(base + kNaBoxOff),
                      {reinterpret_cast<uint32_t*>(base + kNaKeyOff0), reinterpret_cast<uint32_t*>(base + kNaKeyOff1)},
                      {reinterpret_cast<int*>(base + kNaIdxOff0), reinterpret_cast<int*>(base + kNaIdxOff1)},
                      reinterpret_cast<int*>(base + kNaHistOff)};
    mt_launch(n, pos, part_to_real, real_to_part, st, sh, s);
}

}  // namespace masx
