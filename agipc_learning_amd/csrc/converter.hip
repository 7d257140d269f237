// Triplets -> unique BCOO sorted by (row, col) with duplicates summed, on the
// GPU (the job of Converter::convert, linear_system/utils/converter.cu:137-284,
// and optionally Converter::ge2sym, :286-336, which keeps row <= col).
//
// Instead of a global radix sort of (row << 32 | col) keys, triplets are
// bucketed by block row (integer histogram, scan, scatter) and each row's
// bucket -- a few dozen entries for a mesh -- is ordered by (col, input
// index) with a rank sort in LDS, one wavefront per row.  The order inside a
// duplicate group is the input order, so every sum is formed in a fixed order
// and the output is bit-reproducible.
#include "converter.hpp"

#include <stdexcept>
#include <string>

namespace masx
{
namespace
{
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;               // elements per thread in the block scan
constexpr int kScanTile  = kScanBlock * kScanItems;
constexpr int kRowCap    = 1024;            // longest row sorted in LDS; longer rows sort in global scratch

enum ConvFlag : int { kBadIndex = 1 };

__global__ void __launch_bounds__(256) k_conv_count(int n, int nr, const int* rows, const int* cols, int drop_lower,
                                                    int* row_count, int* flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    const int r = rows[t], c = cols[t];
    if(r < 0 || r >= nr || c < 0 || c >= nr)
    {
        atomicOr(flag, (int)kBadIndex);
        return;
    }
    if(drop_lower && r > c) return;
    atomicAdd(&row_count[r], 1);
}

__global__ void __launch_bounds__(256) k_conv_scatter(int n, int nr, const int* rows, const int* cols, int drop_lower,
                                                      const int* row_start, int* cursor, uint32_t* bucket)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    const int r = rows[t], c = cols[t];
    if(r < 0 || r >= nr || c < 0 || c >= nr) return;
    if(drop_lower && r > c) return;
    bucket[row_start[r] + atomicAdd(&cursor[r], 1)] = (uint32_t)t;
}

// exclusive scan, level 1: each workgroup scans one tile and emits its total
__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(const int* in, int* out, int m, int* tile_sums)
{
    __shared__ int s[kScanBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], sum = 0;
#pragma unroll
    for(int k = 0; k < kScanItems; ++k)
    {
        v[k] = base + k < m ? in[base + k] : 0;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for(int off = 1; off < kScanBlock; off <<= 1)
    {
        const int add = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
    for(int k = 0; k < kScanItems; ++k)
    {
        if(base + k < m) out[base + k] = run;
        run += v[k];
    }
    if(threadIdx.x == kScanBlock - 1) tile_sums[blockIdx.x] = s[kScanBlock - 1];
}

// level 2: one workgroup scans the tile totals in place (exclusive), any count
__global__ void __launch_bounds__(kScanBlock) k_scan_totals(int* tile_sums, int tiles, int* total)
{
    __shared__ int s[kScanBlock];
    int carry = 0;
    for(int b = 0; b < tiles; b += kScanBlock)
    {
        const int i = b + threadIdx.x;
        const int v = i < tiles ? tile_sums[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for(int off = 1; off < kScanBlock; off <<= 1)
        {
            const int add = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if(i < tiles) tile_sums[i] = carry + s[threadIdx.x] - v;
        const int chunk_total = s[kScanBlock - 1];
        __syncthreads();
        carry += chunk_total;
    }
    if(threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_add(int* out, int m, const int* tile_sums, const int* total)
{
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    const int add = tile_sums[blockIdx.x];
#pragma unroll
    for(int k = 0; k < kScanItems; ++k)
        if(i + k * kScanBlock < m) out[i + k * kScanBlock] += add;
    if(blockIdx.x == 0 && threadIdx.x == 0) out[m] = *total;
}

// One wavefront per block row: order the row's triplets by (col, input
// index), mark the first entry of every column run, and record where each
// unique block's run starts (written into the row's own bucket range).
__global__ void __launch_bounds__(64) k_conv_row_sort(int nr, const int* row_start, uint32_t* bucket,
                                                      const int* cols, unsigned long long* gkeys,
                                                      unsigned long long* gsorted, int* seg, int* uniq)
{
    __shared__ unsigned long long s_keys[kRowCap], s_sorted[kRowCap];
    const int r = blockIdx.x;
    if(r >= nr) return;
    const int rs = row_start[r], L = row_start[r + 1] - rs;
    const int lane = threadIdx.x;
    if(L == 0)
    {
        if(lane == 0) uniq[r] = 0;
        return;
    }
    unsigned long long* keys   = L <= kRowCap ? s_keys : gkeys + rs;
    unsigned long long* sorted = L <= kRowCap ? s_sorted : gsorted + rs;
    for(int i = lane; i < L; i += 64)
    {
        const uint32_t t = bucket[rs + i];
        keys[i] = ((unsigned long long)(uint32_t)cols[t] << 32) | t;
    }
    __syncthreads();
    // keys are distinct (the input index is part of the key): rank = position
    for(int i = lane; i < L; i += 64)
    {
        const unsigned long long k = keys[i];
        int rank = 0;
        for(int j = 0; j < L; ++j) rank += keys[j] < k;
        sorted[rank] = k;
    }
    __syncthreads();
    int count = 0;  // unique blocks before this 64-entry slice
    for(int base = 0; base < L; base += 64)
    {
        const int i = base + lane;
        bool head = false;
        unsigned long long k = 0;
        if(i < L)
        {
            k = sorted[i];
            head = i == 0 || (k >> 32) != (sorted[i - 1] >> 32);
        }
        const unsigned long long heads = __ballot(head);
        const int before = __popcll(heads & ((1ull << lane) - 1ull));
        if(head) seg[rs + count + before] = rs + i;
        count += __popcll(heads);
        __syncthreads();  // every lane has read sorted[] of this slice before bucket[] is overwritten
        if(i < L) bucket[rs + i] = (uint32_t)(k & 0xffffffffu);
    }
    if(lane == 0) uniq[r] = count;
}

// One wavefront per block row: sum every run (input order) into its unique block.
__global__ void __launch_bounds__(64) k_conv_write(int nr, const int* row_start, const uint32_t* bucket,
                                                   const int* seg, const int* uniq, const int* out_start,
                                                   const int* cols, const double* vals, int* out_rows,
                                                   int* out_cols, double* out_vals)
{
    const int r = blockIdx.x;
    if(r >= nr) return;
    const int rs = row_start[r], re = row_start[r + 1], U = uniq[r], o0 = out_start[r];
    for(int u = threadIdx.x; u < U; u += 64)
    {
        const int s0 = seg[rs + u], s1 = u + 1 < U ? seg[rs + u + 1] : re;
        double acc[9];
#pragma unroll
        for(int c = 0; c < 9; ++c) acc[c] = 0.0;
        for(int k = s0; k < s1; ++k)
        {
            const double* B = vals + 9 * (long long)bucket[k];
#pragma unroll
            for(int c = 0; c < 9; ++c) acc[c] += B[c];
        }
        const long long o = o0 + u;
        out_rows[o] = r;
        out_cols[o] = cols[bucket[s0]];
#pragma unroll
        for(int c = 0; c < 9; ++c) out_vals[9 * o + c] = acc[c];
    }
}

unsigned grid_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

#define CONV_HIP(x)                                                                                      \
    do {                                                                                                 \
        const hipError_t e_ = (x);                                                                       \
        if(e_ != hipSuccess) throw std::runtime_error(std::string("converter: ") + hipGetErrorString(e_)); \
    } while(0)

template <typename T>
void grow(T*& p, size_t& cap, size_t need)
{
    if(need <= cap) return;
    if(p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    CONV_HIP(hipMalloc((void**)&p, need * sizeof(T)));
    cap = need;
}
}  // namespace

Converter::~Converter()
{
    for(void* p : {(void*)row_count_, (void*)row_start_, (void*)cursor_, (void*)bucket_, (void*)seg_, (void*)uniq_,
                   (void*)out_start_, (void*)tiles_, (void*)flag_, (void*)keys_, (void*)sorted_})
        if(p) (void)hipFree(p);
    if(h_pinned_) (void)hipHostFree(h_pinned_);
}

void Converter::scan(const int* in, int* out, int m, hipStream_t st)
{
    // out[0..m] = exclusive prefix sums of in[0..m-1]; out[m] = total
    const int tiles = std::max(1, (m + kScanTile - 1) / kScanTile);
    grow(tiles_, tiles_cap_, (size_t)tiles + 1);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, st, in, out, m, tiles_);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanBlock), 0, st, tiles_, tiles, tiles_ + tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanBlock), 0, st, out, m, tiles_, tiles_ + tiles);
    CONV_HIP(hipGetLastError());
}

int Converter::convert(int n, int n_block_rows, const int* rows, const int* cols, const double* vals, int drop_lower,
                       int* out_rows, int* out_cols, double* out_vals, hipStream_t st)
{
    if(n < 0 || n_block_rows <= 0) throw std::invalid_argument("bad triplet or block-row count");
    if(n == 0) return 0;
    if(!rows || !cols || !vals || !out_rows || !out_cols || !out_vals) throw std::invalid_argument("null pointer");
    const int nr = n_block_rows;
    grow(row_count_, row_count_cap_, (size_t)nr);
    grow(row_start_, row_start_cap_, (size_t)nr + 1);
    grow(cursor_, cursor_cap_, (size_t)nr);
    grow(uniq_, uniq_cap_, (size_t)nr);
    grow(out_start_, out_start_cap_, (size_t)nr + 1);
    grow(bucket_, bucket_cap_, (size_t)n);
    grow(seg_, seg_cap_, (size_t)n);
    grow(keys_, keys_cap_, (size_t)n);
    grow(sorted_, sorted_cap_, (size_t)n);
    grow(flag_, flag_cap_, 1);
    if(!h_pinned_) CONV_HIP(hipHostMalloc((void**)&h_pinned_, 2 * sizeof(int), hipHostMallocDefault));

    CONV_HIP(hipMemsetAsync(row_count_, 0, sizeof(int) * nr, st));
    CONV_HIP(hipMemsetAsync(cursor_, 0, sizeof(int) * nr, st));
    CONV_HIP(hipMemsetAsync(flag_, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_conv_count, dim3(grid_for(n, 256)), dim3(256), 0, st, n, nr, rows, cols, drop_lower,
                       row_count_, flag_);
    CONV_HIP(hipGetLastError());
    scan(row_count_, row_start_, nr, st);
    hipLaunchKernelGGL(k_conv_scatter, dim3(grid_for(n, 256)), dim3(256), 0, st, n, nr, rows, cols, drop_lower,
                       row_start_, cursor_, bucket_);
    hipLaunchKernelGGL(k_conv_row_sort, dim3(nr), dim3(64), 0, st, nr, row_start_, bucket_, cols, keys_, sorted_,
                       seg_, uniq_);
    CONV_HIP(hipGetLastError());
    scan(uniq_, out_start_, nr, st);
    hipLaunchKernelGGL(k_conv_write, dim3(nr), dim3(64), 0, st, nr, row_start_, bucket_, seg_, uniq_, out_start_, cols,
                       vals, out_rows, out_cols, out_vals);
    CONV_HIP(hipGetLastError());
    CONV_HIP(hipMemcpyAsync(h_pinned_, out_start_ + nr, sizeof(int), hipMemcpyDeviceToHost, st));
    CONV_HIP(hipMemcpyAsync(h_pinned_ + 1, flag_, sizeof(int), hipMemcpyDeviceToHost, st));
    CONV_HIP(hipStreamSynchronize(st));
    if(h_pinned_[1] & kBadIndex) throw std::invalid_argument("triplet block index outside [0, n_block_rows)");
    return h_pinned_[0];
}

}  // namespace masx
