// Structure analysis of one MAS system on the device (see devplan.hpp).
// Integer work: counting, stable sorts and scans over the matrix structure,
// bound by memory latency; every output is a pure function of the inputs (no
// order of atomics reaches an output: atomics only count or take a maximum).
#include "devplan.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device.hpp"
#include "spmv_layout.hpp"

namespace masx
{
namespace
{
constexpr int kThreads = 256;

inline int grid_for(long long n, int per = kThreads)
{
    const long long g = (n + per - 1) / per;
    return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

__device__ __forceinline__ int packed_index(int bank, int r, int c) { return bank * r - r * (r + 1) / 2 + c; }

// first entry of row r whose column is >= c (columns sorted within a row)
__device__ __forceinline__ int first_ge(const int* row_ptr, const int* col, int r, int c)
{
    int lo = row_ptr[r], hi = row_ptr[r + 1];
    while(lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if(col[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// the system whose range [off[s], off[s+1]) holds x (off increasing, n_sys >= 1)
__device__ __forceinline__ int system_of(const int* off, int n_sys, long long x)
{
    int lo = 0, hi = n_sys - 1;
    while(lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if(off[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

constexpr uint64_t kDpSentinel = ~(uint64_t)0;

// exclusive scan of v over a 256-thread workgroup (s: 256 ints of LDS);
// returns the exclusive prefix, *total the sum
__device__ int block_exclusive_scan(int v, int* s, int* total)
{
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for(int off = 1; off < kThreads; off <<= 1)
    {
        const int a = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    const int incl = s[t];
    *total = s[kThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void k_dp_fill(int* p, int value, long long n)
{
    for(long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_dp_csr_items(long long nb, const int* rows, const int* cols, DpSystems sys, const int* e2i,
                               int n_rows, int bits, uint64_t* keys, int* vals, int* row_cnt, int* diag_src, int* err)
{
    const uint64_t sentinel = (uint64_t)n_rows << bits;
    for(long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nb; k += (long long)gridDim.x * blockDim.x)
    {
        const int s = system_of(sys.src_off, sys.n_sys, k);
        const int e0 = sys.ext_off[s], n_ext = sys.ext_off[s + 1] - e0;
        const int r = rows[k], c = cols[k];
        if((unsigned)r >= (unsigned)n_ext || (unsigned)c >= (unsigned)n_ext)
        {
            atomicOr(err, 1);
            keys[2 * k] = keys[2 * k + 1] = sentinel;
            vals[2 * k] = vals[2 * k + 1] = -1;
            continue;
        }
        const int iR = e2i[e0 + r];
        if(r == c)
        {
            keys[2 * k]     = (uint64_t)iR << bits | (uint64_t)iR;
            vals[2 * k]     = (int)k;
            keys[2 * k + 1] = sentinel;
            vals[2 * k + 1] = -1;
            atomicAdd(&row_cnt[iR], 1);
            atomicMax(&diag_src[iR], (int)k);  // the last diagonal block of the row
        }
        else
        {
            const int iC    = e2i[e0 + c];
            keys[2 * k]     = (uint64_t)iR << bits | (uint64_t)iC;
            vals[2 * k]     = (int)k;
            keys[2 * k + 1] = (uint64_t)iC << bits | (uint64_t)iR;
            vals[2 * k + 1] = (int)k | (int)0x80000000u;
            atomicAdd(&row_cnt[iR], 1);
            atomicAdd(&row_cnt[iC], 1);
        }
    }
}

__global__ void k_dp_csr_fill(long long nnz, const uint64_t* keys, const int* vals, int bits, int* col, int* entry_src)
{
    const uint64_t mask = ((uint64_t)1 << bits) - 1;
    for(long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < nnz; x += (long long)gridDim.x * blockDim.x)
    {
        col[x]       = (int)(keys[x] & mask);
        entry_src[x] = vals[x];
    }
}

__global__ void k_dp_gal_items(int lvl, long long n, const int* rows, const int* cols, const int* EI, const int* EJ,
                               const int* ES, int e_prev, DpSystems sys, const int* gn_all, const int* lvl_begin,
                               int bits, uint64_t* keys, int* vals, int* n_valid)
{
    int valid = 0;
    for(long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    {
        const int s = lvl == 1 ? system_of(sys.src_off, sys.n_sys, k) : ES[e_prev + k];
        if(sys.levels[s] <= lvl)
        {
            keys[k] = kDpSentinel;
            vals[k] = -1;
            continue;
        }
        const int* gn = gn_all + sys.gn_off[s];
        int a, b, src;
        bool diag;
        if(lvl == 1)
        {
            const int r = rows[k], c = cols[k];
            a    = gn[r];
            b    = gn[c];
            diag = r == c;
            src  = (int)k;
        }
        else
        {
            const int pi = EI[e_prev + k], pj = EJ[e_prev + k];
            a    = gn[pi];
            b    = gn[pj];
            diag = pi == pj;
            src  = e_prev + (int)k;
        }
        int I, J, mode;
        if(a < b) { I = a; J = b; mode = 0; }
        else if(a > b) { I = b; J = a; mode = 1; }
        else { I = a; J = a; mode = diag ? 0 : 2; }
        const int lb = lvl_begin[s];
        keys[k] = (uint64_t)s << (2 * bits) | (uint64_t)(I - lb) << bits | (uint64_t)(J - lb);
        vals[k] = src << 2 | mode;
        ++valid;
    }
    if(valid) atomicAdd(n_valid, valid);
}

__global__ void k_dp_run_ends(long long n, const uint64_t* keys, int* flag)
{
    for(long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x <= n; x += (long long)gridDim.x * blockDim.x)
        flag[x] = x == n || keys[x] == kDpSentinel ? 0 : (x + 1 == n || keys[x + 1] != keys[x] ? 1 : 0);
}

__global__ void k_dp_gal_fill(long long n, const uint64_t* keys, const int* vals, const int* flag, const int* qpos,
                              int g_base, int e_base, const int* lvl_begin, int bits, int* gal_src, int* gal_ptr,
                              int* EI, int* EJ, int* ES)
{
    const uint64_t mask = ((uint64_t)1 << bits) - 1;
    for(long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x)
    {
        const uint64_t key = keys[x];
        if(key == kDpSentinel) continue;  // sorted last: x is past the level's items
        gal_src[g_base + x] = vals[x];
        if(flag[x])
        {
            const int s  = (int)(key >> (2 * bits));
            const int lb = lvl_begin[s];
            const int q  = e_base + qpos[x];
            EI[q]          = lb + (int)((key >> bits) & mask);
            EJ[q]          = lb + (int)(key & mask);
            ES[q]          = s;
            gal_ptr[q + 1] = g_base + (int)x + 1;
        }
    }
}

__global__ void k_dp_ent_count_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                                    int* row_cnt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if(r > n_slots) return;
    int c = 0;
    if(r < n_slots)
    {
        const int r0 = r / bank * bank;
        for(int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
            c += entry_src[e] >= 0 && col[e] >= r0 && col[e] < r0 + bank;
    }
    row_cnt[r] = c;  // row_cnt[n_slots] = 0 closes the scan
}

__global__ void k_dp_ent_fill_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                                   const int* row_off, int* pos, int* src)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if(r >= n_slots) return;
    const int r0 = r / bank * bank, lr = r - r0;
    int at = row_off[r];
    for(int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
    {
        const int lc = col[e] - r0;
        if(entry_src[e] < 0 || lc < 0 || lc >= bank) continue;  // transposed copy, or another bank
        pos[at] = (lr <= lc ? packed_index(bank, lr, lc) : packed_index(bank, lc, lr)) | (lr > lc ? 1 : 0) << 16;
        src[at] = entry_src[e];
        ++at;
    }
}

__device__ __forceinline__ int coarse_bank_of(const DpSystems& sys, int n_fine_banks, int bank, int s, int I)
{
    return n_fine_banks + (sys.coarse_base[s] + I - sys.n_slots[s]) / bank;
}

__global__ void k_dp_ent_flag_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* EJ,
                                     const int* ES, int* flag, int* bank_cnt)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e > n_gal) return;
    int f = 0;
    if(e < n_gal)
    {
        const int I = EI[e], J = EJ[e];
        f = I / bank == J / bank;
        if(f) atomicAdd(&bank_cnt[coarse_bank_of(sys, n_fine_banks, bank, ES[e], I)], 1);
    }
    flag[e] = f;
}

__global__ void k_dp_ent_first_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI,
                                      const int* ES, const int* flag, const int* fpos, int* bank_first)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= n_gal || !flag[e]) return;
    atomicMin(&bank_first[coarse_bank_of(sys, n_fine_banks, bank, ES[e], EI[e])], fpos[e]);
}

__global__ void k_dp_ent_fill_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI,
                                     const int* EJ, const int* ES, const int* flag, const int* fpos,
                                     const int* bank_first, const int* bank_ptr, int* pos, int* src)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= n_gal || !flag[e]) return;
    const int gb = coarse_bank_of(sys, n_fine_banks, bank, ES[e], EI[e]);
    const int at = bank_ptr[gb] + fpos[e] - bank_first[gb];
    pos[at] = packed_index(bank, EI[e] % bank, EJ[e] % bank);
    src[at] = e;
}

__global__ void k_dp_ent_bank_counts(int n_fine_banks, int bank, const int* row_off, int* bank_cnt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b < n_fine_banks) bank_cnt[b] = row_off[(b + 1) * bank] - row_off[b * bank];
}

// the greedy tile that starts at r0 (build_spmv_layout's cut, from any r0).
// A workgroup takes kThreads consecutive r0, whose tiles can only reach the
// next kThreads + kSpmvTileRows - 1 rows: their row pointers, and their
// column words when they fit kSlCols, are staged in LDS once, so each
// thread's per-row searches (two per row it adds) read LDS instead of HBM.
constexpr int kSlRows = kThreads + kSpmvTileRows;
constexpr int kSlCols = 8192;

__global__ void __launch_bounds__(kThreads) k_dp_sl_next(int n, const int* row_ptr, const int* col, const int* row_sys,
                                                         const int* row_off, int cap, int* nxt, int* kept)
{
    __shared__ int s_ptr[kSlRows + 1];
    __shared__ int s_col[kSlCols];
    const int R = blockIdx.x * kThreads, r0 = R + (int)threadIdx.x;
    const int rw = min(n, R + kSlRows);  // rows [R, rw) are staged
    for(int i = threadIdx.x; i <= rw - R; i += kThreads) s_ptr[i] = row_ptr[R + i];
    __syncthreads();
    const int e0 = s_ptr[0], ne = s_ptr[rw - R] - e0;
    const bool cached = ne <= kSlCols;  // uniform
    if(cached)
        for(int i = threadIdx.x; i < ne; i += kThreads) s_col[i] = col[e0 + i];
    __syncthreads();
    if(r0 >= n) return;
    // first entry of row r (R <= r < rw) whose column is >= c
    auto first_ge_s = [&](int r, int c) {
        int lo = s_ptr[r - R], hi = s_ptr[r + 1 - R];
        while(lo < hi)
        {
            const int mid = (lo + hi) >> 1;
            if((cached ? s_col[mid - e0] : col[mid]) < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    };
    const int r_end = row_off[row_sys[r0] + 1];  // the system's end
    int k = s_ptr[r0 + 1 - R] - s_ptr[r0 - R];
    int r1 = r0 + 1;
    while(r1 < r_end && r1 - r0 < kSpmvTileRows)
    {
        // r1 + 1 <= r0 + kSpmvTileRows <= R + kSlRows - 1 and r1 < n: inside the staged rows
        const int q = (s_ptr[r1 + 1 - R] - first_ge_s(r1, r1)) + (first_ge_s(r1, r0) - s_ptr[r1 - R]);
        if(k + q > cap) break;
        k += q;
        ++r1;
    }
    nxt[r0]  = r1;
    kept[r0] = k;
}

// The chain of tiles from row 0, walked by one thread through windows of the
// tile ends staged in LDS (each step r -> nxt[r] advances by 1..256 rows).
constexpr int kChainWindow = 32768;
__global__ void __launch_bounds__(1024) k_dp_sl_chain(int n, const int* nxt, int* ts, int* n_tiles)
{
    __shared__ unsigned char step[kChainWindow];
    __shared__ int s_r, s_t;
    if(threadIdx.x == 0)
    {
        s_r = 0;
        s_t = 0;
    }
    for(;;)
    {
        __syncthreads();
        const int w0 = s_r;
        if(w0 >= n) break;  // uniform: every thread read the same s_r
        const int wn = min(kChainWindow, n - w0);
        for(int i = threadIdx.x; i < wn; i += blockDim.x) step[i] = (unsigned char)(nxt[w0 + i] - (w0 + i) - 1);
        __syncthreads();
        if(threadIdx.x == 0)
        {
            int r = w0, t = s_t;
            while(r < w0 + wn)
            {
                ts[t++] = r;
                r += step[r - w0] + 1;
            }
            s_r = r;
            s_t = t;
        }
    }
    if(threadIdx.x == 0) *n_tiles = s_t;
}

__global__ void __launch_bounds__(kThreads) k_dp_sl_tile_counts(int n_tiles, int n, const int* ts, const int* kept,
                                                                const int* row_ptr, const int* col, int* kept_al,
                                                                int* trans, int* err)
{
    __shared__ int s[kThreads];
    const int t = blockIdx.x;
    if(t > n_tiles) return;  // uniform per workgroup
    if(t == n_tiles)
    {
        if(threadIdx.x == 0) kept_al[t] = trans[t] = 0;  // closes both scans
        return;
    }
    const int r0 = ts[t], r1 = t + 1 < n_tiles ? ts[t + 1] : n, r = r0 + (int)threadIdx.x;
    int q = 0;
    if(r < r1) q = first_ge(row_ptr, col, r, r1) - first_ge(row_ptr, col, r, r + 1);
    int total;
    (void)block_exclusive_scan(q, s, &total);
    if(threadIdx.x == 0)
    {
        trans[t]   = total;
        kept_al[t] = (kept[r0] + kSpmvAlign - 1) / kSpmvAlign * kSpmvAlign;
        if(total > kSpmvMaxCap) atomicOr(err, 4);  // tpos has 14 bits
    }
}

__global__ void __launch_bounds__(kThreads) k_dp_sl_fill(int n_tiles, int n, const int* ts, const int* kept,
                                                         const int* e_base, const int* q_base, const int* row_ptr,
                                                         const int* col, const int* entry_src, const int* row_sys,
                                                         int* L_row_ptr, int* L_tptr, int* L_col, int* L_src, int4* wg,
                                                         int2* wg_ent, int* err)
{
    __shared__ int s_scan[kThreads];
    __shared__ int s_tstart[kThreads + 1];
    const int t = blockIdx.x;
    const int r0 = ts[t], r1 = t + 1 < n_tiles ? ts[t + 1] : n, nr = r1 - r0;
    const int j = (int)threadIdx.x, r = r0 + j;
    int cnt = 0, kr = 0;
    if(j < nr)
    {
        const int d = first_ge(row_ptr, col, r, r), b = first_ge(row_ptr, col, r, r0);
        cnt = d - b;                                    // entries (r', r), r0 <= r' < r: by symmetry
        kr  = (row_ptr[r + 1] - d) + (b - row_ptr[r]);  // entries row r keeps
    }
    int n_trans, n_kept;
    const int tst  = block_exclusive_scan(cnt, s_scan, &n_trans);
    const int koff = block_exclusive_scan(kr, s_scan, &n_kept);
    s_tstart[j] = tst;
    if(j == 0) s_tstart[kThreads] = n_trans;
    __syncthreads();
    if(j == 0 && n_trans > kSpmvMaxCap) atomicOr(err, 4);
    if(j < nr && n_trans <= kSpmvMaxCap)
    {
        int o = e_base[t] + koff;
        L_row_ptr[r] = o;
        L_tptr[r]    = q_base[t] + tst;
        unsigned first = kSpmvRowStart;
        for(int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
        {
            const int c = col[e];
            unsigned word = (unsigned)c;
            if(c >= r0 && c < r1)
            {
                if(c < r) continue;  // the tile's lower copy: its upper twin covers it
                const int tpos = c == r ? 0
                                        : s_tstart[c - r0] +
                                              (first_ge(row_ptr, col, c, r) - first_ge(row_ptr, col, c, r0)) +
                                              (e - first_ge(row_ptr, col, r, c));
                word = kSpmvInTile | (unsigned)tpos << 16 | (unsigned)(r - r0) << 8 | (unsigned)(c - r0);
            }
            L_col[o] = (int)(word | first);
            L_src[o] = entry_src[e];
            first = 0;
            ++o;
        }
    }
    if(j == 0)
    {
        wg[t]     = make_int4(row_sys[r0], kWgPlain << 28, r0, r1);
        wg_ent[t] = make_int2(e_base[t], e_base[t] + kept[r0]);
        if(t == n_tiles - 1)
        {
            L_row_ptr[n] = e_base[t] + kept[r0];
            L_tptr[n]    = q_base[t] + n_trans;
        }
    }
}

__global__ void k_dp_sl_pad(long long total, int* L_col, int* L_src)
{
    for(long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x)
    {
        L_col[i] = 0;
        L_src[i] = kSpmvPadSrc;
    }
}
}  // namespace

size_t dp_sort_bytes(long long n, int end_bit)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (const int*)nullptr,
                                    (int*)nullptr, (size_t)n, 0u, (unsigned)end_bit);
    return bytes;
}

size_t dp_scan_bytes(long long n)
{
    size_t bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, bytes, (const int*)nullptr, (int*)nullptr, 0, (size_t)n,
                                  rocprim::plus<int>());
    return bytes;
}

hipError_t dp_sort(void* tmp, size_t tmp_bytes, const uint64_t* keys_in, uint64_t* keys_out, const int* vals_in,
                   int* vals_out, long long n, int end_bit, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u,
                                     (unsigned)end_bit, st);
}

hipError_t dp_scan(void* tmp, size_t tmp_bytes, const int* in, int* out, long long n, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), st);
}

hipError_t dp_fill(int* p, int value, long long n, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_fill, dim3(grid_for(n)), dim3(kThreads), 0, st, p, value, n);
    return hipGetLastError();
}

hipError_t dp_csr_items(long long nb, const int* rows, const int* cols, DpSystems sys, const int* e2i, int n_rows,
                        int bits, uint64_t* keys, int* vals, int* row_cnt, int* diag_src, int* err, hipStream_t st)
{
    if(nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_csr_items, dim3(grid_for(nb)), dim3(kThreads), 0, st, nb, rows, cols, sys, e2i, n_rows,
                       bits, keys, vals, row_cnt, diag_src, err);
    return hipGetLastError();
}

hipError_t dp_csr_fill(long long nnz, const uint64_t* keys, const int* vals, int bits, int* col, int* entry_src,
                       hipStream_t st)
{
    if(nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_csr_fill, dim3(grid_for(nnz)), dim3(kThreads), 0, st, nnz, keys, vals, bits, col, entry_src);
    return hipGetLastError();
}

hipError_t dp_gal_items(int lvl, long long n, const int* rows, const int* cols, const int* EI, const int* EJ,
                        const int* ES, int e_prev, DpSystems sys, const int* gn_all, const int* lvl_begin, int bits,
                        uint64_t* keys, int* vals, int* n_valid, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_gal_items, dim3(grid_for(n)), dim3(kThreads), 0, st, lvl, n, rows, cols, EI, EJ, ES,
                       e_prev, sys, gn_all, lvl_begin, bits, keys, vals, n_valid);
    return hipGetLastError();
}

hipError_t dp_run_ends(long long n, const uint64_t* keys, int* flag, hipStream_t st)
{
    hipLaunchKernelGGL(k_dp_run_ends, dim3(grid_for(n + 1)), dim3(kThreads), 0, st, n, keys, flag);
    return hipGetLastError();
}

hipError_t dp_gal_fill(long long n, const uint64_t* keys, const int* vals, const int* flag, const int* qpos,
                       int g_base, int e_base, const int* lvl_begin, int bits, int* gal_src, int* gal_ptr, int* EI,
                       int* EJ, int* ES, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_gal_fill, dim3(grid_for(n)), dim3(kThreads), 0, st, n, keys, vals, flag, qpos, g_base,
                       e_base, lvl_begin, bits, gal_src, gal_ptr, EI, EJ, ES);
    return hipGetLastError();
}

hipError_t dp_ent_count_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                             int* row_cnt, hipStream_t st)
{
    hipLaunchKernelGGL(k_dp_ent_count_rows, dim3((n_slots + kThreads) / kThreads), dim3(kThreads), 0, st, n_slots,
                       bank, row_ptr, col, entry_src, row_cnt);
    return hipGetLastError();
}

hipError_t dp_ent_fill_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                            const int* row_off, int* pos, int* src, hipStream_t st)
{
    if(n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_ent_fill_rows, dim3((n_slots + kThreads - 1) / kThreads), dim3(kThreads), 0, st, n_slots,
                       bank, row_ptr, col, entry_src, row_off, pos, src);
    return hipGetLastError();
}

hipError_t dp_ent_flag_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* EJ,
                              const int* ES, int* flag, int* bank_cnt, hipStream_t st)
{
    hipLaunchKernelGGL(k_dp_ent_flag_coarse, dim3((n_gal + kThreads) / kThreads), dim3(kThreads), 0, st, n_gal, bank,
                       n_fine_banks, sys, EI, EJ, ES, flag, bank_cnt);
    return hipGetLastError();
}

hipError_t dp_ent_first_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* ES,
                               const int* flag, const int* fpos, int* bank_first, hipStream_t st)
{
    if(n_gal <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_ent_first_coarse, dim3((n_gal + kThreads - 1) / kThreads), dim3(kThreads), 0, st, n_gal,
                       bank, n_fine_banks, sys, EI, ES, flag, fpos, bank_first);
    return hipGetLastError();
}

hipError_t dp_ent_fill_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* EJ,
                              const int* ES, const int* flag, const int* fpos, const int* bank_first,
                              const int* bank_ptr, int* pos, int* src, hipStream_t st)
{
    if(n_gal <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_ent_fill_coarse, dim3((n_gal + kThreads - 1) / kThreads), dim3(kThreads), 0, st, n_gal,
                       bank, n_fine_banks, sys, EI, EJ, ES, flag, fpos, bank_first, bank_ptr, pos, src);
    return hipGetLastError();
}

hipError_t dp_ent_bank_counts(int n_fine_banks, int bank, const int* row_off, int* bank_cnt, hipStream_t st)
{
    if(n_fine_banks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_ent_bank_counts, dim3((n_fine_banks + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                       n_fine_banks, bank, row_off, bank_cnt);
    return hipGetLastError();
}

hipError_t dp_sl_next(int n, const int* row_ptr, const int* col, const int* row_sys, const int* row_off, int cap,
                      int* nxt, int* kept, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_sl_next, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, n, row_ptr, col,
                       row_sys, row_off, cap, nxt, kept);
    return hipGetLastError();
}

hipError_t dp_sl_chain(int n, const int* nxt, int* ts, int* n_tiles, hipStream_t st)
{
    hipLaunchKernelGGL(k_dp_sl_chain, dim3(1), dim3(1024), 0, st, n, nxt, ts, n_tiles);
    return hipGetLastError();
}

hipError_t dp_sl_tile_counts(int n_tiles, int n, const int* ts, const int* kept, const int* row_ptr, const int* col,
                             int* kept_al, int* trans, int* err, hipStream_t st)
{
    hipLaunchKernelGGL(k_dp_sl_tile_counts, dim3(n_tiles + 1), dim3(kThreads), 0, st, n_tiles, n, ts, kept, row_ptr,
                       col, kept_al, trans, err);
    return hipGetLastError();
}

hipError_t dp_sl_fill(int n_tiles, int n, const int* ts, const int* kept, const int* e_base, const int* q_base,
                      const int* row_ptr, const int* col, const int* entry_src, const int* row_sys, int* L_row_ptr,
                      int* L_tptr, int* L_col, int* L_src, int4* wg, int2* wg_ent, int* err, hipStream_t st)
{
    if(n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_sl_fill, dim3(n_tiles), dim3(kThreads), 0, st, n_tiles, n, ts, kept, e_base, q_base,
                       row_ptr, col, entry_src, row_sys, L_row_ptr, L_tptr, L_col, L_src, wg, wg_ent, err);
    return hipGetLastError();
}

hipError_t dp_sl_pad(long long total, int* L_col, int* L_src, hipStream_t st)
{
    if(total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dp_sl_pad, dim3(grid_for(total)), dim3(kThreads), 0, st, total, L_col, L_src);
    return hipGetLastError();
}
}  // namespace masx
