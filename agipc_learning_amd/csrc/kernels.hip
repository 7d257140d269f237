// HIP kernels for the MAS-preconditioned PCG on gfx950 (CDNA4, wave64).
//
// Conventions
//  * vectors are float3 per internal block row (12 bytes, array of structs);
//  * SpMV matrix values are nine float planes (entry-major SoA) so that lanes
//    reading consecutive entries issue fully coalesced dword loads;
//  * bank inverses are stored diagonal-major: for block diagonal d (blocks
//    (I, I+d)), value c of every block is contiguous over I, so the lanes of
//    a bank read each diagonal with coalesced loads;
//  * reductions are deterministic and need no fences or atomics: every
//    workgroup writes its FP64 partial sum, and the consuming kernel (the next
//    launch) sums a system's partials in a fixed order in each workgroup;
//    the first workgroup of a system in a kernel writes the scalars that kernel
//    does not read;
//  * kernels of a finished system return at once, so a captured chunk of
//    iterations can run past convergence at negligible cost.
#include "kernels.hpp"
#include "mas_tiles.hpp"
#include "spmv_layout.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <type_traits>

namespace masx
{
namespace
{

__device__ __forceinline__ int packed_index(int bank, int r, int c) { return bank * r - r * (r + 1) / 2 + c; }

// Row of packed upper-triangle index pos (the largest r with
// packed_index(BANK, r, r) <= pos): the root of the row-start quadratic,
// corrected by at most a step either way for the float rounding.
template <int BANK>
__device__ __forceinline__ int packed_row(int pos)
{
    constexpr int K = 2 * BANK + 1;
    int r = (int)(((float)K - sqrtf((float)(K * K - 8 * pos))) * 0.5f);
    r = r < 0 ? 0 : (r > BANK - 1 ? BANK - 1 : r);
    if(r > 0 && packed_index(BANK, r, r) > pos) --r;
    if(r + 1 < BANK && packed_index(BANK, r + 1, r + 1) <= pos) ++r;
    return r;
}

// blocks stored before block diagonal d of a bank
__host__ __device__ __forceinline__ constexpr int diag_off(int bank, int d) { return d * bank - d * (d - 1) / 2; }

template <typename T> struct Store { static constexpr int kWords = 5; };  // value pairs (0,1)..(6,7),(8,-)
template <> struct Store<float> { static constexpr int kWords = 9; };

__device__ __forceinline__ float lo_half(uint32_t w, __half*) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float hi_half(uint32_t w, __half*) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }
__device__ __forceinline__ float lo_half(uint32_t w, __hip_bfloat16*) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_half(uint32_t w, __hip_bfloat16*) { return __uint_as_float(w & 0xffff0000u); }

__device__ __forceinline__ uint32_t half_bits(double v, __half*)
{
    return (uint32_t)__half_as_ushort(__float2half_rn((float)v));
}
__device__ __forceinline__ uint32_t half_bits(double v, __hip_bfloat16*)
{
    const __hip_bfloat16 h = __float2bfloat16((float)v);
    unsigned short u;
    __builtin_memcpy(&u, &h, 2);
    return (uint32_t)u;
}

// Loads of operands each iteration streams once (bank inverses, SpMV values
// and columns): nontemporal, so they pass through the caches without
// evicting the vectors that the gathers reuse.  On this card a read stream
// of nontemporal 16-byte loads runs ~15% faster than one of ordinary loads.
typedef float    nt_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t nt_u4 __attribute__((ext_vector_type(4)));
typedef int      nt_i4 __attribute__((ext_vector_type(4)));
typedef float    nt_f2 __attribute__((ext_vector_type(2)));
typedef int      nt_i2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float    ld_nt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ int      ld_nt(const int* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ uint32_t ld_nt(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float4 ld_nt(const float4* p)
{
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ld_nt(const uint4* p)
{
    const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(p));
    return uint4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ int4 ld_nt(const int4* p)
{
    const nt_i4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_i4*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld_nt(const float2* p)
{
    const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ int2 ld_nt(const int2* p)
{
    const nt_i2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_i2*>(p));
    return make_int2(v.x, v.y);
}

// Loads the nine values of block (I, I+d) of one bank (row-major within the block).
template <int BANK>
__device__ __forceinline__ void load_block(const float* M, int d, int I, float v[9])
{
    const int len = BANK - d;
    const float* p = M + 9 * diag_off(BANK, d) + I;
#pragma unroll
    for(int c = 0; c < 9; ++c) v[c] = p[c * len];
}
template <int BANK, typename H>
__device__ __forceinline__ void load_block_h(const uint32_t* M, int d, int I, float v[9])
{
    const int len = BANK - d;
    const uint32_t* p = M + 5 * diag_off(BANK, d) + I;
#pragma unroll
    for(int q = 0; q < 4; ++q)
    {
        const uint32_t w = p[q * len];
        v[2 * q]     = lo_half(w, (H*)nullptr);
        v[2 * q + 1] = hi_half(w, (H*)nullptr);
    }
    v[8] = lo_half(p[4 * len], (H*)nullptr);
}

// wave64 butterfly sum
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide FP64 sum in a fixed order; the result is returned to every thread.
__device__ double block_sum(double v, double* smem)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if(lane == 0) smem[wid] = v;
    __syncthreads();
    double t = 0.0;
    for(int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += smem[w];
    __syncthreads();
    return t;
}

// Block-wide FP64 sum in the same fixed order as block_sum, valid in thread
// 0 only; one barrier (for the last reduction of a kernel).
__device__ double block_sum_last(double v, double* smem)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if(lane == 0) smem[wid] = v;
    __syncthreads();
    double t = 0.0;
    if(threadIdx.x == 0)
        for(int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += smem[w];
    return t;
}

// [first, end) of system s's workgroups in a kernel's table (range = the
// per-system pairs, n_wg = the table's size).  With one system that is the
// whole table, known without a load, so the loads that depend on it can
// issue at once.
__device__ __forceinline__ int2 wg_range(const DevData& d, const int* range, int n_wg, int s)
{
    return d.single ? make_int2(0, n_wg) : make_int2(range[2 * s], range[2 * s + 1]);
}

// Sum of the partials of one system's workgroups of the producing kernel
// (r = their range).  Identical in every workgroup.
__device__ double sys_sum(const double* part, int2 r, double* smem)
{
    const int b = r.x, e = r.y;
    double v = 0.0;
    for(int i = b + (int)threadIdx.x; i < e; i += blockDim.x) v += part[i];
    return block_sum(v, smem);
}

// The loads of sys_sum without the reduction: this thread's share of the
// partials (eight loads in flight per pass).  Forming the total with
// block_sum later lets the caller issue its other loads in between.
__device__ __forceinline__ double partial_share(const double* part, int2 r)
{
    const int b = r.x, e = r.y;
    double v = 0.0;
#pragma unroll 1
    for(int i = b + (int)threadIdx.x; i < e; i += 8 * 256)
    {
        double q[8];
#pragma unroll
        for(int j = 0; j < 8; ++j) q[j] = i + 256 * j < e ? part[i + 256 * j] : 0.0;
        v += ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    }
    return v;
}

// a[k] for a small register array without dynamic indexing (no scratch)
template <int N>
__device__ __forceinline__ int pick(const int (&a)[N], int k)
{
    int v = a[0];
#pragma unroll
    for(int j = 1; j < N; ++j) v = k == j ? a[j] : v;
    return v;
}

// Workgroup order for row tiles: the hardware deals workgroups to the eight
// XCDs in turn, so this gives each XCD (and its L2) a contiguous run of tiles.
__device__ __forceinline__ int xcd_tile(int bid, int n)
{
    const int xcd = bid & 7, k = bid >> 3;
    const int per = n >> 3, rem = n & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

__device__ __forceinline__ bool finite_d(double v) { return isfinite(v); }

}  // namespace

// ===================================================================== build

// 12 x 12 body couplings of the SpMV's body tiles from the FP64 upper blocks:
// one thread per coupling and 3 x 3 sub-block (q, t), summing the sub-block's
// entries of the full CSR (duplicate blocks follow the first, transposed
// copies read their upper block transposed) in FP64 before rounding.
__global__ void __launch_bounds__(256) k_pack_abd(DevData d, const int* __restrict__ cp_body,
                                                  const int* __restrict__ cp_sub, long long n_sub, float* vals_abd)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(g >= n_sub) return;
    const long long k = g / 16;
    const int sub = (int)(g % 16), q = sub / 4, t = sub % 4;
    const int r = d.body_row[cp_body[k]] + q, target = d.body_row[d.abd_cp_col[k]] + t;
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int e_end = d.row_ptr[r + 1];
    for(int e = cp_sub[g]; e >= 0 && e < e_end && d.col[e] == target; ++e)
    {
        const int code = d.entry_src[e];
        const double* B = d.vals64 + 9 * (long long)(code & 0x7fffffff);
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j) acc[3 * i + j] += code >= 0 ? B[i + 3 * j] : B[j + 3 * i];
    }
    float* out = vals_abd + 144 * k + 12 * (3 * q) + 3 * t;
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) out[12 * i + j] = (float)acc[3 * i + j];
}

// Upper FP64 blocks (col-major) -> FP32 SoA planes of the SpMV layout, one
// thread per entry so the plane writes are contiguous; an entry marked
// transposed takes its upper block's transpose, a padding entry zeros.
__global__ void __launch_bounds__(256) k_pack_values(DevData d)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= d.nnz_s) return;
    const int code = d.sy_src[e];
    float* planes = const_cast<float*>(d.vals32);
    if(code == kSpmvPadSrc)
    {
#pragma unroll
        for(int q = 0; q < 9; ++q) planes[q * d.plane + e] = 0.f;
        return;
    }
    const double* B = d.vals64 + 9 * (long long)(code & 0x7fffffff);
    if(code >= 0)
    {
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j) planes[(3 * i + j) * d.plane + e] = (float)B[i + 3 * j];
    }
    else
    {
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j) planes[(3 * i + j) * d.plane + e] = (float)B[j + 3 * i];
    }
}

// Galerkin coarse entries of one level: E[e] = sum of its mapped sources
// (as is, transposed, or B + B^T).  G lanes per entry stride over its source
// list; their sums meet in a fixed shuffle tree.
template <int G>
__global__ void __launch_bounds__(256) k_gal_gather(const double* fine, double* E, const int* ptr, const int* src,
                                                    int e0, int e1, int src_is_fine)
{
    const int e   = e0 + (int)blockIdx.x * (256 / G) + (int)threadIdx.x / G;
    const int sub = threadIdx.x % G;
    const bool valid = e < e1;  // uniform over each group of G lanes
    double acc[9];
#pragma unroll
    for(int i = 0; i < 9; ++i) acc[i] = 0.0;
    if(valid)
    {
        const int k1 = ptr[e + 1];
#pragma unroll 2
        for(int k = ptr[e] + sub; k < k1; k += G)
        {
            const int code = src[k];
            const int idx = code >> 2, mode = code & 3;
            double B[9];
            if(src_is_fine)
            {
#pragma unroll
                for(int i = 0; i < 3; ++i)
#pragma unroll
                    for(int j = 0; j < 3; ++j) B[3 * i + j] = fine[9 * (long long)idx + i + 3 * j];
            }
            else
            {
#pragma unroll
                for(int i = 0; i < 9; ++i) B[i] = E[9 * (long long)idx + i];
            }
#pragma unroll
            for(int i = 0; i < 3; ++i)
#pragma unroll
                for(int j = 0; j < 3; ++j)
                {
                    const double bij = B[3 * i + j], bji = B[3 * j + i];
                    acc[3 * i + j] += mode == 0 ? bij : mode == 1 ? bji : bij + bji;
                }
        }
    }
#pragma unroll
    for(int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for(int i = 0; i < 9; ++i) acc[i] += __shfl_xor(acc[i], o, 64);
    if(valid && sub == 0)
#pragma unroll
        for(int i = 0; i < 9; ++i) E[9 * (long long)e + i] = acc[i];
}

// Where one element of a bank's inverse is stored, diagonal-major: (i, j),
// i >= j, is the upper element (j, i) of block (j / 3, i / 3); a diagonal
// block also takes its transpose position.  Half formats store
// 16-bit halves of 5 words per block (the tenth half is zero).
// Half-precision inverses of 32-slot banks are stored as MFMA tiles (the
// tile layout of mas_tiles.hpp) when the MFMA local kernel reads them
// (DevData::mas_waves == 1); otherwise, and for 16-slot banks and FP32, by
// block diagonals.
template <int BANK, typename T>
constexpr bool kTileCapable = BANK == 32 && !std::is_same<T, float>::value;

template <int BANK, typename T>
__device__ __forceinline__ int inverse_slots(int i, int j, bool tiles, int (&slot)[2])
{
    // Stored positions of element (i, j), i >= j, of a bank's inverse,
    // bank-relative, in units of the stored value (an FP32 float, or a
    // 16-bit half of the half-precision layouts): one, or two where the
    // element has a mirror in the same diagonal block or tile.
    if constexpr(kTileCapable<BANK, T>)
    {
        if(tiles)
        {
            // (i, j) is the upper element (j, i); a diagonal tile also takes its mirror (i, j)
            slot[0] = tile_half_index(j, i);
            if(i != j && i / kTileDim == j / kTileDim)
            {
                slot[1] = tile_half_index(i, j);
                return 2;
            }
            return 1;
        }
    }
    const int I = j / 3, J = i / 3, dd = J - I, len = BANK - dd;
    const int c0 = 3 * (j % 3) + i % 3, c1 = 3 * (i % 3) + j % 3;
    const int nc = (I == J && i != j) ? 2 : 1;
    for(int u = 0; u < nc; ++u)
    {
        const int c = u ? c1 : c0;
        if constexpr(std::is_same<T, float>::value)
            slot[u] = 9 * diag_off(BANK, dd) + c * len + I;
        else
            // value c in half c & 1 of word c / 2 (value 8's word keeps the
            // unused tenth value, zero, in its upper half)
            slot[u] = 2 * (5 * diag_off(BANK, dd) + (c >> 1) * len + I) + (c & 1);
    }
    return nc;
}

// Bytes of one bank's stored inverse (the MFMA tile layout, the half
// block-diagonal layout of five words per block, or FP32 nine floats per
// block); every position is either an element's (inverse_slots) or a zero.
template <int BANK, typename T>
__device__ __forceinline__ int bank_inverse_bytes(bool tiles)
{
    constexpr int NBLK = BANK * (BANK + 1) / 2;
    if constexpr(kTileCapable<BANK, T>)
        if(tiles) return kTileBankHalves * 2;
    return std::is_same<T, float>::value ? NBLK * 9 * 4 : NBLK * 5 * 4;
}

// The bank's inverse leaves the factor kernels as a dense lower triangle of
// floats in LDS (element (i, j), i >= j, at i (i + 1) / 2 + j; the half
// layouts' values already scaled); `map` (one entry per stored value, from
// k_inverse_map: the element each position holds, 0xffff for a zero) turns
// it into the stored layout, written in 16-byte stores.
template <int BANK, typename T>
__device__ __forceinline__ void write_bank_inverse(const float* L, const unsigned short* __restrict__ map, void* out,
                                                   int bank, bool tiles)
{
    const int n16 = bank_inverse_bytes<BANK, T>(tiles) / 16;
    auto val = [&](unsigned short e) { return e != 0xffffu ? L[e] : 0.f; };
    if constexpr(std::is_same<T, float>::value)
    {
        float4* dst = reinterpret_cast<float4*>(static_cast<char*>(out) + (size_t)bank * (n16 * 16));
        for(int q = (int)threadIdx.x; q < n16; q += (int)blockDim.x)
        {
            const ushort4 m = reinterpret_cast<const ushort4*>(map)[q];
            dst[q] = make_float4(val(m.x), val(m.y), val(m.z), val(m.w));
        }
    }
    else
    {
        uint4* dst = reinterpret_cast<uint4*>(static_cast<char*>(out) + (size_t)bank * (n16 * 16));
        for(int q = (int)threadIdx.x; q < n16; q += (int)blockDim.x)
        {
            const uint4 mw = reinterpret_cast<const uint4*>(map)[q];  // eight 16-bit entries
            const uint32_t m[4] = {mw.x, mw.y, mw.z, mw.w};
            uint32_t o[4];
#pragma unroll
            for(int h = 0; h < 4; ++h)
            {
                const unsigned short e0 = (unsigned short)(m[h] & 0xffffu), e1 = (unsigned short)(m[h] >> 16);
                const uint32_t b0 = e0 != 0xffffu ? half_bits((double)L[e0], (T*)nullptr) : 0u;
                const uint32_t b1 = e1 != 0xffffu ? half_bits((double)L[e1], (T*)nullptr) : 0u;
                o[h] = (b0 & 0xffffu) | (b1 << 16);
            }
            dst[q] = uint4{o[0], o[1], o[2], o[3]};
        }
    }
}

// map[position] = the dense lower-triangle element a bank's stored inverse
// keeps there (the map was filled with 0xffff: positions no element takes
// hold zero).  One thread per element.
template <int BANK, typename T>
__global__ void __launch_bounds__(256) k_inverse_map(unsigned short* map, int tiles)
{
    constexpr int DIM = 3 * BANK, NE = DIM * (DIM + 1) / 2;
    static_assert(NE < 0xffff, "element indices fit 16 bits");
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if(e >= NE) return;
    int i = 0;
    while((i + 1) * (i + 2) / 2 <= e) ++i;
    const int j = e - i * (i + 1) / 2;
    int slot[2];
    const int n = inverse_slots<BANK, T>(i, j, tiles != 0, slot);
    for(int u = 0; u < n; ++u) map[slot[u]] = (unsigned short)e;
}

// First row of assembly pass s of `npass` over the lower triangle of a
// dim x dim matrix (rows [row(s), row(s + 1)) hold about 1/npass of it).
__host__ __device__ constexpr int factor_pass_row(int dim, int npass, int s)
{
    const int target = (int)((long long)s * (dim * (dim + 1) / 2) / npass);
    int p = 0;
    while(p < dim && (p + 1) * (p + 2) / 2 <= target) ++p;
    return p;
}

__host__ __device__ constexpr int factor_pass_max(int dim, int npass)
{
    int m = 0;
    for(int s = 0; s < npass; ++s)
    {
        const int lo = factor_pass_row(dim, npass, s), hi = factor_pass_row(dim, npass, s + 1);
        const int n = hi * (hi + 1) / 2 - lo * (lo + 1) / 2;
        m = n > m ? n : m;
    }
    return m;
}

// Register slot of tile (r, c), r >= 2c, of the 8 x 16 grid's lower tiles.
__host__ __device__ constexpr int tile8x16(int rr, int r, int c) { return c * rr - c * (c - 1) + (r - 2 * c); }

// Dense local assembly and FP64 inversion of one bank per workgroup, written
// diagonal-major in the storage precision T.  The inversion is the symmetric
// sweep: pivot k turns a_ij into a_ij - a_ik a_kj / d (i, j != k), a_ik into
// a_ik / d and a_kk into -1 / d, so the matrix stays symmetric and, after
// every pivot, holds -A^-1.  A pivot (the Schur-complement diagonal) below
// max(1e-12, 1e-12 max|diag|) is raised to that floor, which inverts the
// same modified matrix as the pivot-floored Cholesky of
// pro/.../TraditionalMAS32Preconditioner.cu:1382-1439.  Thread (ty, tx) of
// an 8-row x 16-column grid of 128 threads keeps, in registers, elements
// (ty + 8 r, tx + 16 c) of the tiles that reach the lower triangle (r >= 2c;
// 42 for 32-slot banks), so each pivot's barrier, LDS exchange of the pivot
// column (double-buffered) and bookkeeping serve 42 FMAs per thread; the
// sweep is bound by instruction issue, not by the FMAs.  The bank matrix is
// assembled in two passes through half the LDS the whole lower triangle
// would take (zero diagonal -> 1), that buffer then holds the pivot columns
// and at the end the bank's stored inverse, copied out in 16-byte stores.  Half-precision storage is
// scaled by the largest stored entry of the bank's inverse in magnitude, so
// every stored value lies in [-1, 1] (also for an inverse that floored
// pivots left indefinite, whose off-diagonal entries can exceed its diagonal).
// One pass of a bank's dense assembly: the entries' 3 x 3 blocks added into
// rows [lo, hi) of the packed lower triangle H (LDS, offset by `base`).  A
// block's lower half only for a diagonal block; fine banks read the FP64
// input values (transposed for a lower block), coarse banks their Galerkin
// sums E (row-major).  Each thread takes entries tid, tid + THREADS, ... and
// loads kFactorEntriesPerRound of them (their words, then their blocks)
// before adding them in entry order: one dependent memory round trip per
// round instead of one per entry, and the same sums in the same order.
constexpr int kFactorEntriesPerRound = 2;  // 64-system batch build: 4 equal, 6 and 8 slower

template <int BANK, int THREADS>
__device__ __forceinline__ void factor_assemble_pass(const DevData& d, const double* E, bool fine, int e_begin,
                                                     int e_end, const int* ent_pos, const int* ent_src, int lo,
                                                     int hi, int base, double* H)
{
    constexpr int R = kFactorEntriesPerRound;
    const double* vsrc = fine ? d.vals64 : E;
    for(int k = e_begin + (int)threadIdx.x; k < e_end; k += THREADS * R)
    {
        int pw[R], sr[R];
#pragma unroll
        for(int u = 0; u < R; ++u)
        {
            const bool has = k + THREADS * u < e_end;
            pw[u] = has ? ent_pos[k + THREADS * u] : -1;
            sr[u] = has ? ent_src[k + THREADS * u] : 0;
        }
        double v9[R][9];
#pragma unroll
        for(int u = 0; u < R; ++u)
        {
            const double* blk = vsrc + 9 * (long long)sr[u];
#pragma unroll
            for(int q = 0; q < 9; ++q) v9[u][q] = pw[u] >= 0 ? blk[q] : 0.0;
        }
#pragma unroll
        for(int u = 0; u < R; ++u)
        {
            if(pw[u] < 0) break;
            const int  pos      = pw[u] & 0xffff;
            const bool rowmajor = !fine || (pw[u] >> 16) != 0;
            const int  I        = packed_row<BANK>(pos);
            const int  J        = pos - packed_index(BANK, I, I) + I;
#pragma unroll
            for(int aa = 0; aa < 3; ++aa)
#pragma unroll
                for(int bb = 0; bb < 3; ++bb)
                {
                    const int p = 3 * I + aa, q = 3 * J + bb;
                    if(I == J && p < q) continue;  // diagonal block: its lower half carries it
                    const int row = p > q ? p : q, col = p > q ? q : p;
                    if(row < lo || row >= hi) continue;
                    const double v = rowmajor ? v9[u][3 * aa + bb] : v9[u][aa + 3 * bb];
                    atomicAdd(&H[row * (row + 1) / 2 + col - base], v);
                }
        }
    }
}

template <int BANK, typename T>
__global__ void __launch_bounds__(128) k_mas_factor_sym(DevData d, const double* E, const int* ent_ptr,
                                                        const int* ent_pos, const int* ent_src, void* out,
                                                        float* scale_out, int* status, int bank_begin)
{
    constexpr int DIM   = 3 * BANK;
    constexpr int RR    = DIM / 8;    // tile rows (8 rows each)
    constexpr int RC    = DIM / 16;   // tile columns (16 columns each)
    constexpr int NT    = tile8x16(RR, 2 * RC, RC);  // slots of every tile (r, c), r >= 2c
    constexpr int NPASS = DIM > 48 ? 2 : 1;
    constexpr int NH    = factor_pass_max(DIM, NPASS);
    constexpr int NS    = NH > 2 * DIM + 4 ? NH : 2 * DIM + 4;
    constexpr bool kHalf = !std::is_same<T, float>::value;
    const bool tiles = kTileCapable<BANK, T> && d.mas_waves == 1;  // the MFMA local kernel's layout
    static_assert(DIM % 16 == 0, "the 8 x 16 grid must tile the bank matrix");
    __shared__ __align__(16) double S[NS];
    __shared__ double red[2];
    double* H  = S;
    double* vk = S;              // during the sweep: two pivot columns ...
    double* pv = S + 2 * DIM;    // ... the pivots' reciprocals and floored flags
    const int bank = bank_begin + (int)blockIdx.x, tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double a[NT];
#pragma unroll
    for(int t = 0; t < NT; ++t) a[t] = 0.0;
    // ---- assembly, NPASS row ranges of the lower triangle through H
    const bool fine = bank < d.n_fine_banks;
    const int  e_begin = ent_ptr[bank], e_end = ent_ptr[bank + 1];
    double dmax = 0.0;
#pragma unroll
    for(int pass = 0; pass < NPASS; ++pass)
    {
        const int lo = factor_pass_row(DIM, NPASS, pass), hi = factor_pass_row(DIM, NPASS, pass + 1);
        const int base = lo * (lo + 1) / 2, n = hi * (hi + 1) / 2 - base;
        for(int q = tid; q < n; q += 128) H[q] = 0.0;
        __syncthreads();
        factor_assemble_pass<BANK, 128>(d, E, fine, e_begin, e_end, ent_pos, ent_src, lo, hi, base, H);
        __syncthreads();
        // zero diagonal -> 1 (padding slots behave as identity)
        for(int q = lo + tid; q < hi; q += 128)
        {
            double& h = H[q * (q + 1) / 2 + q - base];
            if(h == 0.0) h = 1.0;
            dmax = fmax(dmax, fabs(h));
        }
        __syncthreads();
#pragma unroll
        for(int c = 0; c < RC; ++c)
#pragma unroll
            for(int r = 2 * c; r < RR; ++r)
            {
                const int i = ty + 8 * r, j = tx + 16 * c;
                if(i >= lo && i < hi && i >= j) a[tile8x16(RR, r, c)] = H[i * (i + 1) / 2 + j - base];
            }
        __syncthreads();  // H becomes the pivot-column buffer
    }
    // pivot floor max(1e-12, 1e-12 max |diag|)
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    if((tid & 63) == 0) red[tid >> 6] = dmax;
    __syncthreads();
    const double floor_value = fmax(1.0e-12, 1.0e-12 * fmax(red[0], red[1]));
    int bad = 0;
    // ---- sweep: pivot k = 8 rk + k8 = kt + 16 cc (tile row rk, tile column
    // cc = rk / 2, kt = k8 + 8 (rk & 1))
#pragma unroll
    for(int rk = 0; rk < RR; ++rk)
    {
        const int cc = rk / 2, h = rk & 1;
        {
#pragma unroll 1
            for(int k8 = 0; k8 < 8; ++k8)
            {
                const int kt = k8 + 8 * h, k = kt + 16 * cc, buf = k & 1;
                double* col = vk + buf * DIM;
                // column k from row k down (lower elements (i, k), i >= k) ...
                if(tx == kt)
                {
#pragma unroll
                    for(int r = rk; r < RR; ++r)
                        if(r > rk || ty >= k8) col[ty + 8 * r] = a[tile8x16(RR, r, cc)];
                }
                // ... and left of the pivot (lower elements (k, j), j < k); the pivot
                if(ty == k8)
                {
#pragma unroll
                    for(int c = 0; c <= cc; ++c)
                        if(c < cc || tx < kt) col[tx + 16 * c] = a[tile8x16(RR, rk, c)];
                    if(tx == kt)
                    {
                        const double p = a[tile8x16(RR, rk, cc)];
                        const bool floored = !(p >= floor_value) || isinf(p);  // also catches NaN
                        pv[buf]     = 1.0 / (floored ? floor_value : p);
                        pv[2 + buf] = floored ? 1.0 : 0.0;
                    }
                }
                __syncthreads();
                const double ip = pv[buf];
                bad += pv[2 + buf] != 0.0 ? 1 : 0;
                double g[RC];
#pragma unroll
                for(int c = 0; c < RC; ++c) g[c] = col[tx + 16 * c] * ip;
#pragma unroll
                for(int r = 0; r < RR; ++r)
                {
                    const double f = col[ty + 8 * r];
#pragma unroll
                    for(int c = 0; c < RC; ++c)
                        if(2 * c <= r) a[tile8x16(RR, r, c)] -= f * g[c];
                }
                // column k below the pivot and row k left of it take v / d, the pivot -1 / d
                if(tx == kt)
                {
#pragma unroll
                    for(int r = rk; r < RR; ++r)
                        if(r > rk || ty > k8) a[tile8x16(RR, r, cc)] = col[ty + 8 * r] * ip;
                }
                if(ty == k8)
                {
#pragma unroll
                    for(int c = 0; c <= cc; ++c)
                        if(c < cc || tx < kt) a[tile8x16(RR, rk, c)] = g[c];
                    if(tx == kt) a[tile8x16(RR, rk, cc)] = -ip;
                }
            }
        }
    }
    // ---- the inverse (-a) from the registers
    double scale = 1.0;
    if constexpr(kHalf)
    {
        double m = 0.0;
#pragma unroll
        for(int c = 0; c < RC; ++c)
#pragma unroll
            for(int r = 2 * c; r < RR; ++r)
                if(ty + 8 * r >= tx + 16 * c) m = fmax(m, fabs(a[tile8x16(RR, r, c)]));
#pragma unroll
        for(int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        __syncthreads();  // red's floor reads are done
        if((tid & 63) == 0) red[tid >> 6] = m;
        __syncthreads();
        m = fmax(red[0], red[1]);
        scale = m > 0.0 ? m : 1.0;
    }
    if(tid == 0) scale_out[bank] = (float)scale;
    const double inv_scale = 1.0 / scale;
    static_assert(sizeof(S) >= (size_t)(3 * BANK) * (3 * BANK + 1) / 2 * 4, "the dense inverse fits the LDS buffer");
    float* Lf = reinterpret_cast<float*>(S);
    __syncthreads();  // the sweep's pivot columns are read
#pragma unroll
    for(int c = 0; c < RC; ++c)
#pragma unroll
        for(int r = 2 * c; r < RR; ++r)
        {
            const int i = ty + 8 * r, j = tx + 16 * c;
            const double v = -a[tile8x16(RR, r, c)];
            if(i >= j) Lf[i * (i + 1) / 2 + j] = (float)(kHalf ? v * inv_scale : v);
        }
    __syncthreads();
    write_bank_inverse<BANK, T>(Lf, d.inv_map, out, bank, tiles);
    if(tid == 0 && bad && status) atomicAdd(status, bad);
}

// The same sweep by one wave per bank: lane (ty, tx) of an 8 x 8 grid keeps
// elements (ty + 8 r, tx + 8 c) of the 8 x 8 tiles r >= c (78 for 32-slot
// banks) in registers.  Each pivot's exchange of the pivot column stays
// within the wave (no barrier between waves on the pivot chain), and two
// banks share a SIMD.
__host__ __device__ constexpr int tile8(int r, int c) { return r * (r + 1) / 2 + c; }


// value of v in lane `lane` (uniform over the wave)
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int BANK, typename T>
__global__ void __launch_bounds__(64) k_mas_factor_wave(DevData d, const double* E, const int* ent_ptr,
                                                        const int* ent_pos, const int* ent_src, void* out,
                                                        float* scale_out, int* status, int bank_begin)
{
    constexpr int DIM   = 3 * BANK;
    constexpr int NB    = DIM / 8;              // tile rows = tile columns
    constexpr int NT    = NB * (NB + 1) / 2;    // lower tiles
    constexpr int NPASS = DIM > 48 ? 2 : 1;
    constexpr int NH    = factor_pass_max(DIM, NPASS);
    constexpr int NS    = NH > 2 * DIM ? NH : 2 * DIM;
    constexpr bool kHalf = !std::is_same<T, float>::value;
    const bool tiles = kTileCapable<BANK, T> && d.mas_waves == 1;  // the MFMA local kernel's layout
    static_assert(DIM % 8 == 0, "the 8 x 8 grid must tile the bank matrix");
    __shared__ __align__(16) double S[NS];
    double* H  = S;
    double* vk = S;              // during the sweep: two pivot columns
    const int bank = bank_begin + (int)blockIdx.x, tid = threadIdx.x, tx = tid & 7, ty = tid >> 3;
    double a[NT];
#pragma unroll
    for(int t = 0; t < NT; ++t) a[t] = 0.0;
    // ---- assembly, NPASS row ranges of the lower triangle through H
    const bool fine = bank < d.n_fine_banks;
    const int  e_begin = ent_ptr[bank], e_end = ent_ptr[bank + 1];
    double dmax = 0.0;
#pragma unroll
    for(int pass = 0; pass < NPASS; ++pass)
    {
        const int lo = factor_pass_row(DIM, NPASS, pass), hi = factor_pass_row(DIM, NPASS, pass + 1);
        const int base = lo * (lo + 1) / 2, n = hi * (hi + 1) / 2 - base;
        for(int q = tid; q < n; q += 64) H[q] = 0.0;
        __syncthreads();
        factor_assemble_pass<BANK, 64>(d, E, fine, e_begin, e_end, ent_pos, ent_src, lo, hi, base, H);
        __syncthreads();
        // zero diagonal -> 1 (padding slots behave as identity)
        for(int q = lo + tid; q < hi; q += 64)
        {
            double& h = H[q * (q + 1) / 2 + q - base];
            if(h == 0.0) h = 1.0;
            dmax = fmax(dmax, fabs(h));
        }
        __syncthreads();
#pragma unroll
        for(int r = 0; r < NB; ++r)
#pragma unroll
            for(int c = 0; c <= r; ++c)
            {
                const int i = ty + 8 * r, j = tx + 8 * c;
                if(i >= lo && i < hi && i >= j) a[tile8(r, c)] = H[i * (i + 1) / 2 + j - base];
            }
        __syncthreads();  // H becomes the pivot-column buffer
    }
    // pivot floor max(1e-12, 1e-12 max |diag|)
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    const double floor_value = fmax(1.0e-12, 1.0e-12 * dmax);
    int bad = 0;
    // ---- sweep: pivot k = 8 rk + k8 (tile row and column rk).  Each pivot's
    // value is read from its lane (readlane) as soon as the previous pivot's
    // update has formed it, and every lane takes its reciprocal: the division
    // overlaps the exchange of the pivot column instead of preceding it.
    double pk = readlane_f64(a[tile8(0, 0)], 0);
#pragma unroll
    for(int rk = 0; rk < NB; ++rk)
    {
#pragma unroll 1
        for(int k8 = 0; k8 < 8; ++k8)
        {
            const int k = k8 + 8 * rk, buf = k & 1;
            double* col = vk + buf * DIM;
            const bool floored = !(pk >= floor_value) || isinf(pk);  // also catches NaN
            const double ip = 1.0 / (floored ? floor_value : pk);
            bad += floored ? 1 : 0;
            // column k from row k down (lower elements (i, k), i >= k) ...
            if(tx == k8)
            {
#pragma unroll
                for(int r = rk; r < NB; ++r)
                    if(r > rk || ty >= k8) col[ty + 8 * r] = a[tile8(r, rk)];
            }
            // ... and left of the pivot (lower elements (k, j), j < k)
            if(ty == k8)
            {
#pragma unroll
                for(int c = 0; c <= rk; ++c)
                    if(c < rk || tx < k8) col[tx + 8 * c] = a[tile8(rk, c)];
            }
            __syncthreads();
            double g[NB];
#pragma unroll
            for(int c = 0; c < NB; ++c) g[c] = col[tx + 8 * c] * ip;
#pragma unroll
            for(int r = 0; r < NB; ++r)
            {
                const double f = col[ty + 8 * r];
#pragma unroll
                for(int c = 0; c <= r; ++c) a[tile8(r, c)] -= f * g[c];
            }
            // column k below the pivot and row k left of it take v / d, the pivot -1 / d
            if(tx == k8)
            {
#pragma unroll
                for(int r = rk; r < NB; ++r)
                    if(r > rk || ty > k8) a[tile8(r, rk)] = col[ty + 8 * r] * ip;
            }
            if(ty == k8)
            {
#pragma unroll
                for(int c = 0; c <= rk; ++c)
                    if(c < rk || tx < k8) a[tile8(rk, c)] = g[c];
                if(tx == k8) a[tile8(rk, rk)] = -ip;
            }
            // the next pivot (k + 1, k + 1), untouched by this pivot's row and column
            if(k8 < 7)
                pk = readlane_f64(a[tile8(rk, rk)], 9 * (k8 + 1));
            else if(rk + 1 < NB)
                pk = readlane_f64(a[tile8(rk + 1 < NB ? rk + 1 : rk, rk + 1 < NB ? rk + 1 : rk)], 0);
        }
    }
    // ---- the inverse (-a) from the registers
    double scale = 1.0;
    if constexpr(kHalf)
    {
        double m = 0.0;
#pragma unroll
        for(int r = 0; r < NB; ++r)
#pragma unroll
            for(int c = 0; c <= r; ++c)
                if(ty + 8 * r >= tx + 8 * c) m = fmax(m, fabs(a[tile8(r, c)]));
#pragma unroll
        for(int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        scale = m > 0.0 ? m : 1.0;
    }
    if(tid == 0) scale_out[bank] = (float)scale;
    const double inv_scale = 1.0 / scale;
    static_assert(sizeof(S) >= (size_t)(3 * BANK) * (3 * BANK + 1) / 2 * 4, "the dense inverse fits the LDS buffer");
    float* Lf = reinterpret_cast<float*>(S);
    __syncthreads();  // the sweep's pivot columns are read
#pragma unroll
    for(int r = 0; r < NB; ++r)
#pragma unroll
        for(int c = 0; c <= r; ++c)
        {
            const int i = ty + 8 * r, j = tx + 8 * c;
            const double v = -a[tile8(r, c)];
            if(i >= j) Lf[i * (i + 1) / 2 + j] = (float)(kHalf ? v * inv_scale : v);
        }
    __syncthreads();
    write_bank_inverse<BANK, T>(Lf, d.inv_map, out, bank, tiles);
    if(tid == 0 && bad && status) atomicAdd(status, bad);
}

// ABD 12x12 block-Jacobi: gather, Gauss-Jordan with partial pivoting in FP64
// (cuda_tools/cuda_eigen.h:129-165), FP32 inverse.
__global__ void __launch_bounds__(64) k_abd_factor(DevData d, const int* body_src, float* inv)
{
    __shared__ double A[12][12], X[12][12];
    __shared__ int piv;
    const int body = blockIdx.x, tid = threadIdx.x;
    for(int t = tid; t < 144; t += blockDim.x)
    {
        const int i = t / 12, j = t % 12;
        const int q = i / 3, a = i % 3, u = j / 3, b = j % 3;
        const int code = body_src[16 * body + 4 * q + u];
        double v = 0.0;
        if(code != -1)
        {
            const int src = code & 0x7fffffff;
            const bool tr = code < 0;
            v = tr ? d.vals64[9 * (long long)src + b + 3 * a] : d.vals64[9 * (long long)src + a + 3 * b];
        }
        A[i][j] = v;
        X[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    for(int c = 0; c < 12; ++c)
    {
        if(tid == 0)
        {
            int p = c;
            double m = fabs(A[c][c]);
            for(int r = c + 1; r < 12; ++r)
                if(fabs(A[r][c]) > m) { m = fabs(A[r][c]); p = r; }
            piv = p;
        }
        __syncthreads();
        if(piv != c && tid < 12)
        {
            double t = A[c][tid]; A[c][tid] = A[piv][tid]; A[piv][tid] = t;
            t = X[c][tid]; X[c][tid] = X[piv][tid]; X[piv][tid] = t;
        }
        __syncthreads();
        const double dg = A[c][c];
        __syncthreads();
        if(tid < 12)
        {
            A[c][tid] /= dg;
            X[c][tid] /= dg;
        }
        __syncthreads();
        double fr[12];
        if(tid < 12)
            for(int r = 0; r < 12; ++r) fr[r] = A[r][c];
        __syncthreads();
        if(tid < 12)
            for(int r = 0; r < 12; ++r)
            {
                if(r == c) continue;
                A[r][tid] -= fr[r] * A[c][tid];
                X[r][tid] -= fr[r] * X[c][tid];
            }
        __syncthreads();
    }
    for(int t = tid; t < 144; t += blockDim.x) inv[144 * (long long)body + t] = (float)X[t / 12][t % 12];
}

// Diagonal 3x3 block inverses (FP64 cofactors, FP32 store) for rows that no
// local preconditioner covers.
__global__ void __launch_bounds__(256) k_diag_factor(DevData d, const int* diag_src, const int* needs, float* inv)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if(row >= d.n_rows) return;
    float* o = inv + 9 * (long long)row;
    const int src = diag_src[row];
    if(!needs[row] || src < 0)
    {
        for(int e = 0; e < 9; ++e) o[e] = (e % 4 == 0) ? 1.f : 0.f;
        return;
    }
    double m[9];
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j) m[3 * i + j] = d.vals64[9 * (long long)src + i + 3 * j];
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double id = 1.0 / (m[0] * c00 + m[1] * c01 + m[2] * c02);
    o[0] = (float)(c00 * id);
    o[1] = (float)((m[2] * m[7] - m[1] * m[8]) * id);
    o[2] = (float)((m[1] * m[5] - m[2] * m[4]) * id);
    o[3] = (float)(c01 * id);
    o[4] = (float)((m[0] * m[8] - m[2] * m[6]) * id);
    o[5] = (float)((m[2] * m[3] - m[0] * m[5]) * id);
    o[6] = (float)(c02 * id);
    o[7] = (float)((m[1] * m[6] - m[0] * m[7]) * id);
    o[8] = (float)((m[0] * m[4] - m[1] * m[3]) * id);
}

// ================================================================ solve I/O

__global__ void __launch_bounds__(256) k_load_rhs(DevData d, const int* ext_to_int, const double* b, int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const int R = ext_to_int[i];
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
        const double v = b[3 * (long long)i + c];
        d.b64[3 * (long long)R + c] = v;
        d.rt[3 * (long long)R + c]  = (float)v;
    }
}

__global__ void __launch_bounds__(256) k_store_x(DevData d, const int* ext_to_int, double* x, int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const int R = ext_to_int[i];
#pragma unroll
    for(int c = 0; c < 3; ++c) x[3 * (long long)i + c] = (double)d.x[3 * (long long)R + c];
}

// Zeroes the iteration vectors and loads b (FP64 and FP32 copies) in the
// internal row order; padding rows get b = 0.
__global__ void __launch_bounds__(256) k_solve_prepare(DevData d, const int* int_to_ext, const double* b)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if(row >= d.n_rows) return;
    const int e = int_to_ext[row];
    const long long o = 3 * (long long)row;
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
        const double v = e >= 0 ? b[3 * (long long)e + c] : 0.0;
        d.b64[o + c] = v;
        d.rt[o + c]  = (float)v;
        d.x[o + c] = 0.f; d.r[o + c] = 0.f; d.z[o + c] = 0.f;
        d.p0[o + c] = 0.f; d.p1[o + c] = 0.f; d.Ap[o + c] = 0.f;
    }
}

// Number of systems not yet done, written to host-visible memory with a
// sequence number (one workgroup, last node of every captured chunk):
// host[0] = count, then host[1] = ++(*seq).
// (unfinished count, sequence number) to host memory, each store made
// visible to the host before the next
__device__ __forceinline__ void export_to_host(int unfinished, int* host, int* seq)
{
    const int q = *seq + 1;
    *seq = q;
    *(volatile int*)host = unfinished;
    __threadfence_system();
    *(volatile int*)(host + 1) = q;
    __threadfence_system();
}

__global__ void __launch_bounds__(256) k_export_status(DevData d, int* host, int* seq)
{
    __shared__ int s[256];
    int n = 0;
    for(int i = threadIdx.x; i < d.n_sys; i += blockDim.x) n += d.state[i].status != kDone;
    s[threadIdx.x] = n;
    __syncthreads();
    for(int off = 128; off > 0; off >>= 1)
    {
        if((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if(threadIdx.x == 0) export_to_host(s[0], host, seq);
}

__global__ void k_init_state(DevData d, int rule, double tol, const int* max_iter, int max_override, int check_every,
                             int frozen)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= d.n_sys) return;
    SysState& st = d.state[s];
    st.rz_prev = st.rz_cur = st.rz0 = st.pAp = st.alpha = st.beta = st.rr = st.bb = st.true_rr = 0.0;
    st.tol = tol;
    st.k = 0;
    st.status = kInit;
    st.failure = kOk;
    st.rule = rule;
    st.max_iter = max_override > 0 ? max_override : max_iter[s];
    st.k_check = 0;
    st.check_every = check_every;
    st.beta_zero = 1;
    st.need_check = 0;
    st.restarts = 0;
    st.n_true_checks = 0;
    st.frozen = frozen;
    st.rejected = 0;
}

// Completes the init / restart phase: bb (init), status -> running.
// With `host` (one system only), the block also does k_export_status's
// work after the transition, saving the chunk a launch.
__global__ void __launch_bounds__(256) k_phase_finish(DevData d, int* host, int* seq)
{
    __shared__ double smem[8];
    const int s = blockIdx.x;
    SysState& st = d.state[s];
    if(st.status == kInit || st.status == kRestarting)  // uniform over the block
    {
        const double v = sys_sum(d.part_upd, wg_range(d, d.wg_upd_range, d.n_wg_upd, s), smem);
        if(threadIdx.x == 0)
        {
            if(!isfinite(v))
            {
                // a non-finite b (init) or b - A x (restart) ends the solve
                if(st.status == kInit) st.bb = v;
                st.rr = v;
                st.status = kDone;
                st.failure = kNonFinite;
            }
            else if(st.status == kInit)
            {
                st.bb = v;
                st.rr = v;
                st.status = v > 0.0 ? kRunning : kDone;
            }
            else
            {
                st.rr = v;
                st.status = kRunning;
            }
        }
    }
    if(host && threadIdx.x == 0) export_to_host(st.status != kDone ? 1 : 0, host, seq);
}

// ========================================================= iteration kernels

// MAS local solve: Z = Minv * R for every bank of every level.  A workgroup
// holds 64/BANK banks; each bank is spread over BANK lanes (lane I = block
// row I) and the four waves take block diagonals d = wave (mod 4).  For the
// stored block (I, I+d) lane I adds B r_{I+d} to its own row and sends
// B^T r_I to lane I+d by a shuffle; the NW partial results are summed in a
// fixed order.  NW (4 or 8 waves per workgroup, i.e. the waves that share a
// bank's block diagonals) is chosen by the autotuner.  Residuals of level >= 2
// banks are formed here from the level-1 residuals of their descendants.
// Workgroups run the coarse banks (the end of the bank list) first, so their
// longer gathers overlap the fine banks.
//
// STAGED (FP32): the workgroup's inverses (contiguous in memory) are copied to
// LDS with 16-byte loads, fully coalesced, and each lane then reads its
// blocks' values from LDS; otherwise every lane loads its own values (4-byte
// loads, a diagonal's values of two banks per wave instruction).
template <int BANK, typename T, int NW, bool STAGED = false>
// four waves per SIMD: the FP32 variant fits 128 VGPRs without spilling
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4, 8))) k_mas_local(DevData d, int force)
{
    static_assert(!STAGED || std::is_same<T, float>::value, "LDS staging is the FP32 layout's");
    constexpr int THREADS = 64 * NW;
    constexpr int BPW   = 64 / BANK;
    constexpr int NBLK  = BANK * (BANK + 1) / 2;
    constexpr int WORDS = Store<T>::kWords;
    constexpr int ND    = BANK / NW;          // block diagonals per wave
    constexpr int TPN   = THREADS / BANK;     // threads per node in the descendant sums
    constexpr int CHUNK = 1024;               // descendant entries staged per pass
    static_assert(BANK % NW == 0 && CHUNK % THREADS == 0, "diagonals and staged entries split evenly");
    __shared__ float sR[64 * 3 + 3];  // + a zero row for the lanes past a diagonal's end
    constexpr int kM4    = NBLK * 9 / 4;                                   // float4 per FP32 bank
    constexpr int kStage = STAGED ? (BPW * kM4 + THREADS - 1) / THREADS : 1;  // float4 per thread
    static_assert(!STAGED || (NBLK * 9 % 4 == 0 && BPW * NBLK * 9 >= CHUNK * 3),
                  "a bank is whole float4s; the staged inverses' area holds the descendant chunk");
    __shared__ __align__(16) float sM[STAGED ? BPW * NBLK * 9 : 4];
    __shared__ float sD0[STAGED ? 1 : CHUNK * 3];
    float* sD = STAGED ? sM : sD0;  // the descendant sums run before the inverses reach LDS
    // the waves' partial results (NW - 1 rows of 64 x 3) reuse the staging
    // area once the products are done (staged 32-slot banks: 38.8 KB of LDS
    // per workgroup, so four workgroups share a CU)
    float* sZ = STAGED ? sM : sD0;
    static_assert((NW - 1) * 64 * 3 <= (STAGED ? BPW * NBLK * 9 : CHUNK * 3), "partials fit the staging area");
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lb = lane / BANK, I = lane % BANK;
    const int bank0 = (int)(gridDim.x - 1 - blockIdx.x) * BPW;
    const int b  = bank0 + lb;
    const int bl = b < d.n_banks ? b : d.n_banks - 1;  // in-bounds address for idle lanes
    // per-bank metadata: uniform over the workgroup (scalar loads)
    int s_live[BPW], s_kind[BPW], s_vec[BPW];
#pragma unroll
    for(int k = 0; k < BPW; ++k)
    {
        const int bb = bank0 + k;
        s_live[k] = s_kind[k] = s_vec[k] = 0;
        if(bb < d.n_banks)
        {
            s_kind[k] = d.bank_kind[bb];
            s_vec[k]  = d.bank_vec[bb];
            s_live[k] = force || d.state[d.single ? 0 : d.bank_sys[bb]].status == kRunning;
        }
    }
    {
        bool any = false;
#pragma unroll
        for(int k = 0; k < BPW; ++k) any |= s_live[k] != 0;
        if(!any) return;  // systems already done: leave before streaming the inverse
    }
    // this wave's inverse blocks: their addresses depend on the bank only, so
    // every load is issued before the residuals are known
    uint32_t raw[STAGED ? 1 : ND][WORDS];
    float4   stage[kStage];
    if constexpr(STAGED)
    {
        // the workgroup's live banks, 16 bytes per lane and load
#pragma unroll
        for(int u = 0; u < kStage; ++u)
        {
            const int i = tid + THREADS * u, k = i / kM4;
            stage[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if(i < BPW * kM4 && pick(s_live, k))
                stage[u] = ld_nt(reinterpret_cast<const float4*>(d.minv) + (size_t)(bank0 + k) * kM4 + i % kM4);
        }
    }
    else
    {
#pragma unroll
        for(int kk = 0; kk < ND; ++kk)
        {
            const int dd = wave + NW * kk, len = BANK - dd;
            const int Ic = I < len ? I : 0;
            const uint32_t* M = reinterpret_cast<const uint32_t*>(d.minv) + (size_t)bl * NBLK * WORDS +
                                WORDS * diag_off(BANK, dd) + Ic;
#pragma unroll
            for(int c = 0; c < WORDS; ++c) raw[kk][c] = M[c * len];
        }
    }
    // residuals of level-0 and level-1 banks
    for(int i = tid; i < 64 * 3; i += blockDim.x)
    {
        const int k = i / (BANK * 3);
        float v = 0.f;
        if(pick(s_live, k) && pick(s_kind, k) < 2)
            v = (pick(s_kind, k) == 0 ? d.r : d.Rc)[3 * (long long)pick(s_vec, k) + i % (BANK * 3)];
        sR[i] = v;
    }
    if(tid < 3) sR[64 * 3 + tid] = 0.f;
    // level >= 2: R(node) = sum of the partials the update kernel formed for
    // it, one per 256-row group holding some of its level-1 descendants.  The
    // lists of a bank's consecutive nodes are contiguous in deep_list; they
    // pass through LDS in chunks with every load of a chunk in flight, then
    // each node's TPN threads sum its entries in list order.
#pragma unroll 1
    for(int k = 0; k < BPW; ++k)
    {
        if(!(pick(s_live, k) && pick(s_kind, k) == 2)) continue;  // uniform over the workgroup
        const int c0 = pick(s_vec, k);
        const int lo = d.deep_ptr[c0], hi = d.deep_ptr[c0 + BANK];
        const int node = tid / TPN, sub = tid % TPN;
        const int nb = d.deep_ptr[c0 + node], ne = d.deep_ptr[c0 + node + 1];
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 1
        for(int base = lo; base < hi; base += CHUNK)
        {
            const int n = min(CHUNK, hi - base);
            int   src[CHUNK / THREADS];
            float v[CHUNK / THREADS][3];
#pragma unroll
            for(int j = 0; j < CHUNK / THREADS; ++j)
            {
                const int e = tid + THREADS * j;
                src[j] = e < n ? d.deep_list[base + e] : -1;
            }
#pragma unroll
            for(int j = 0; j < CHUNK / THREADS; ++j)
            {
                if(src[j] >= 0)
                {
                    const long long c1 = 3 * (long long)src[j];
                    v[j][0] = d.part_deep[c1]; v[j][1] = d.part_deep[c1 + 1]; v[j][2] = d.part_deep[c1 + 2];
                }
            }
            __syncthreads();  // the previous chunk has been consumed
#pragma unroll
            for(int j = 0; j < CHUNK / THREADS; ++j)
            {
                const int e = tid + THREADS * j;
                if(e < n)
                {
                    sD[3 * e] = v[j][0]; sD[3 * e + 1] = v[j][1]; sD[3 * e + 2] = v[j][2];
                }
            }
            __syncthreads();
            const int q0 = max(nb, base), q1 = min(ne, base + n);
            for(int q = q0 + sub; q < q1; q += TPN)
            {
                a0 += sD[3 * (q - base)];
                a1 += sD[3 * (q - base) + 1];
                a2 += sD[3 * (q - base) + 2];
            }
        }
#pragma unroll
        for(int o = TPN / 2; o > 0; o >>= 1)
        {
            a0 += __shfl_xor(a0, o, 64);
            a1 += __shfl_xor(a1, o, 64);
            a2 += __shfl_xor(a2, o, 64);
        }
        if(sub == 0)
        {
            const int i = k * BANK + node;
            sR[3 * i] = a0; sR[3 * i + 1] = a1; sR[3 * i + 2] = a2;
        }
    }
    if constexpr(STAGED)
    {
        __syncthreads();  // the descendant chunk in the same area is consumed
#pragma unroll
        for(int u = 0; u < kStage; ++u)
        {
            const int i = tid + THREADS * u;
            if(i < BPW * kM4) reinterpret_cast<float4*>(sM)[i] = stage[u];
        }
    }
    __syncthreads();
    const bool live = b < d.n_banks && pick(s_live, lb);
    const float r0 = sR[3 * lane], r1 = sR[3 * lane + 1], r2 = sR[3 * lane + 2];
    float z0 = 0.f, z1 = 0.f, z2 = 0.f;
#pragma unroll
    for(int kk = 0; kk < ND; ++kk)
    {
        const int dd  = wave + NW * kk;
        // a lane past the diagonal's end (or of a finished system) holds a
        // finite in-bounds block and multiplies the zero row, adding exact
        // zeros; its transposed product would go to a lane past the bank's
        // end, which the shuffle never reads.  Without a select on the
        // values, each conversion folds into the FMAs that use it.
        const bool act = live && I < BANK - dd;
        float v[9];
        if constexpr(STAGED)
        {
            const int len = BANK - dd, Ic = I < len ? I : 0;
            const float* Ms = sM + lb * NBLK * 9 + 9 * diag_off(BANK, dd) + Ic;
#pragma unroll
            for(int c = 0; c < 9; ++c) v[c] = Ms[c * len];
        }
        else if constexpr(WORDS == 9)
        {
#pragma unroll
            for(int c = 0; c < 9; ++c) v[c] = __uint_as_float(raw[kk][c]);
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 4; ++q)
            {
                v[2 * q]     = lo_half(raw[kk][q], (T*)nullptr);
                v[2 * q + 1] = hi_half(raw[kk][q], (T*)nullptr);
            }
            v[8] = lo_half(raw[kk][4], (T*)nullptr);
        }
        const int j = act ? lane + dd : 64;
        const float q0 = sR[3 * j], q1 = sR[3 * j + 1], q2 = sR[3 * j + 2];
        z0 += v[0] * q0 + v[1] * q1 + v[2] * q2;
        z1 += v[3] * q0 + v[4] * q1 + v[5] * q2;
        z2 += v[6] * q0 + v[7] * q1 + v[8] * q2;
        // transposed contribution to row I + dd
        float t0 = v[0] * r0 + v[3] * r1 + v[6] * r2;
        float t1 = v[1] * r0 + v[4] * r1 + v[7] * r2;
        float t2 = v[2] * r0 + v[5] * r1 + v[8] * r2;
        t0 = __shfl_up(t0, dd, BANK);
        t1 = __shfl_up(t1, dd, BANK);
        t2 = __shfl_up(t2, dd, BANK);
        if(dd > 0 && I >= dd)
        {
            z0 += t0; z1 += t1; z2 += t2;
        }
    }
    if constexpr(STAGED) __syncthreads();  // every wave has read its blocks from the staged inverses
    if(wave > 0)
    {
        float* o = sZ + (wave - 1) * 64 * 3 + 3 * lane;
        o[0] = z0; o[1] = z1; o[2] = z2;
    }
    __syncthreads();
    if(wave == 0 && live)
    {
#pragma unroll
        for(int w = 0; w < NW - 1; ++w)
        {
            const float* o = sZ + w * 64 * 3 + 3 * lane;
            z0 += o[0]; z1 += o[1]; z2 += o[2];
        }
        if(WORDS != 9)
        {
            const float sc = d.minv_scale[b];
            z0 *= sc; z1 *= sc; z2 *= sc;
        }
        float* dst = (pick(s_kind, lb) == 0 ? d.z : d.Zc) + 3 * (long long)(pick(s_vec, lb) + I);
        dst[0] = z0; dst[1] = z1; dst[2] = z2;
    }
}

// ---- MAS local solve on MFMA (32-slot banks, FP16 / BF16 inverse) ---------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaHalf;
template <> struct MfmaHalf<__half>
{
    typedef _Float16 elem;
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ mfma_f4 mma(v8 a, v8 b, mfma_f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaHalf<__hip_bfloat16>
{
    typedef __bf16 elem;
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ mfma_f4 mma(v8 a, v8 b, mfma_f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// LDS copy of a wave's off-diagonal stored tiles (the 15 of tiles 0..17 that
// are not (0,0), (2,2), (4,4)), for the transposed operands: tile t in slot
// t - [t > 0] - [t > 10] - [t > 16] at slot * kTrTileStride, its 16-byte
// lane chunk c at c * 16 + (c / 8) * 64, so that the four 64-byte chunk runs
// one transposed read takes from a tile fall on alternate halves of the LDS
// banks.
constexpr int kTrTileStride = 704;
constexpr int kTrWaveBytes  = 15 * kTrTileStride;  // 10,560
__host__ __device__ constexpr bool tr_tile_kept(int t) { return t != 0 && t != 10 && t != 16; }
__host__ __device__ constexpr int tr_tile_offset(int t)
{
    return (t - (t > 0) - (t > 10) - (t > 16)) * kTrTileStride;
}
__host__ __device__ constexpr int tr_chunk_offset(int c) { return c * 16 + (c >> 3) * 64; }

typedef short mfma_s4 __attribute__((ext_vector_type(4)));

// The transposed half-operand: lane m (< 32) of the half gets row m % 16,
// columns 8 (m / 16) .. + 7 of T^T, i.e. T(8 (m / 16) + e, m % 16), e = 0..7.
// ds_read_b64_tr_b16 hands lane i of each 16-lane group element i % 4 of the
// 8-byte rows its group's lanes i / 4 + 4 q (q = 0..3) address; lane s of a
// group therefore addresses T(8 (m / 16) + 4 h + s / 4, 4 (s % 4) .. + 3),
// one call for e = 4 h .. 4 h + 3 (h = 0, 1).
template <typename V>
__device__ __forceinline__ V read_transposed(const char* tile, int m)
{
    const int a = m & 3, q = (m & 15) >> 2, ch = m >> 4;
    const char* base = tile + (a & 1) * 8;
    const int   c0   = (a >> 1) * 16 + 8 * ch + q;
    typedef __attribute__((address_space(3))) mfma_s4 lds_s4;
    const mfma_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4*)(base + tr_chunk_offset(c0)));
    const mfma_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4*)(base + tr_chunk_offset(c0 + 4)));
    typedef short s8 __attribute__((ext_vector_type(8)));
    const s8 both = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(V, both);
}

// Z = Minv R for every 32-slot bank with the FP16 / BF16 inverse on MFMA.
// One wave per bank, four banks per workgroup.  z_I (rows 16 I .. 16 I + 15
// of the bank's 96) = sum over column pairs P of [M(I, 2P) M(I, 2P + 1)]
// times residual entries 32 P .. 32 P + 31: one v_mfma_f32_16x16x32 per
// (I, P) and residual part.  M(I, J) is the stored tile T(I, J) for I <= J,
// loaded straight into operand registers, and T(J, I)^T below the diagonal,
// read transposed from the wave's LDS copy of its tiles.  The residual
// enters as two half-precision parts (r' = hi + lo to about 22 bits, r'
// = r 2^-e with 2^e above the bank's largest |r|) in column 0 of the B
// operand; FP32 accumulation; z = s 2^e D(:, 0) with s the inverse's scale.
// Residuals of level >= 2 banks come from the update kernel's partials as
// in k_mas_local.
template <typename T>
__global__ void __launch_bounds__(256) k_mas_local_mfma(DevData d, int force)
{
    using Op   = MfmaHalf<T>;
    using V8   = typename Op::v8;
    using E    = typename Op::elem;
    constexpr int BANK  = 32;
    constexpr int NB    = 4;          // banks per workgroup, one per wave
    constexpr int TPN   = 256 / BANK;
    constexpr int CHUNK = 1024;
    static_assert(CHUNK * 3 * 4 <= NB * kTrWaveBytes, "the deep-residual staging fits the tile buffer");
    __shared__ float sR[NB * 96];
    __shared__ __align__(16) char sT[NB * kTrWaveBytes];   // deep-residual staging, then the tiles
    __shared__ __align__(16) E sB[NB][2][96];               // residual parts hi, lo
    __shared__ int sMeta[3][NB];                            // live, kind, vector start of each bank
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int b    = (int)(gridDim.x - 1 - blockIdx.x) * NB + wave;
    int live = 0, kind = 0, vec = 0;
    if(b < d.n_banks)
    {
        kind = d.bank_kind[b];
        vec  = d.bank_vec[b];
        live = force || d.state[d.single ? 0 : d.bank_sys[b]].status == kRunning;
    }
    // this wave's inverse (a bank of a finished system is not read): nine
    // tile pairs (lanes 0-31 the first tile, 32-63 the second) and the
    // diagonal tiles (1,1), (3,3), (5,5) (lanes 32-63; lanes 0-31 load the
    // same chunks), issued before the residuals are known
    uint4 pr[kTilePairs], sg[3];
    if(live)
    {
        const uint4* M = reinterpret_cast<const uint4*>(d.minv) + (size_t)b * (kTileBankHalves / 8);
#pragma unroll
        for(int q = 0; q < kTilePairs; ++q) pr[q] = ld_nt(M + q * 64 + lane);
#pragma unroll
        for(int q = 0; q < 3; ++q) sg[q] = ld_nt(M + (2 * kTilePairs + q) * 32 + (lane & 31));
    }
    else
    {
#pragma unroll
        for(int q = 0; q < kTilePairs; ++q) pr[q] = uint4{0u, 0u, 0u, 0u};
#pragma unroll
        for(int q = 0; q < 3; ++q) sg[q] = uint4{0u, 0u, 0u, 0u};
    }
    if(lane == 0)
    {
        sMeta[0][wave] = live;
        sMeta[1][wave] = kind;
        sMeta[2][wave] = vec;
    }
    // residuals of level-0 and level-1 banks
    {
        float va = 0.f, vb = 0.f;
        if(live && kind < 2)
        {
            const float* src = (kind == 0 ? d.r : d.Rc) + 3 * (long long)vec;
            va = src[lane];
            if(lane < 32) vb = src[64 + lane];
        }
        sR[wave * 96 + lane] = va;
        if(lane < 32) sR[wave * 96 + 64 + lane] = vb;
    }
    __syncthreads();
    if(!(sMeta[0][0] | sMeta[0][1] | sMeta[0][2] | sMeta[0][3])) return;  // systems already done
    // level >= 2: sums of the update kernel's partials, in list order
    float* sD = reinterpret_cast<float*>(sT);
#pragma unroll 1
    for(int k = 0; k < NB; ++k)
    {
        if(!(sMeta[0][k] && sMeta[1][k] == 2)) continue;  // uniform over the workgroup
        const int c0 = sMeta[2][k];
        const int lo = d.deep_ptr[c0], hi = d.deep_ptr[c0 + BANK];
        const int node = tid / TPN, sub = tid % TPN;
        const int nb = d.deep_ptr[c0 + node], ne = d.deep_ptr[c0 + node + 1];
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 1
        for(int base = lo; base < hi; base += CHUNK)
        {
            const int n = min(CHUNK, hi - base);
            int   src[CHUNK / 256];
            float v[CHUNK / 256][3];
#pragma unroll
            for(int j = 0; j < CHUNK / 256; ++j)
            {
                const int e = tid + 256 * j;
                src[j] = e < n ? d.deep_list[base + e] : -1;
            }
#pragma unroll
            for(int j = 0; j < CHUNK / 256; ++j)
            {
                if(src[j] >= 0)
                {
                    const long long c1 = 3 * (long long)src[j];
                    v[j][0] = d.part_deep[c1]; v[j][1] = d.part_deep[c1 + 1]; v[j][2] = d.part_deep[c1 + 2];
                }
            }
            __syncthreads();
#pragma unroll
            for(int j = 0; j < CHUNK / 256; ++j)
            {
                const int e = tid + 256 * j;
                if(e < n)
                {
                    sD[3 * e] = v[j][0]; sD[3 * e + 1] = v[j][1]; sD[3 * e + 2] = v[j][2];
                }
            }
            __syncthreads();
            const int q0 = max(nb, base), q1 = min(ne, base + n);
            for(int q = q0 + sub; q < q1; q += TPN)
            {
                a0 += sD[3 * (q - base)];
                a1 += sD[3 * (q - base) + 1];
                a2 += sD[3 * (q - base) + 2];
            }
        }
#pragma unroll
        for(int o = TPN / 2; o > 0; o >>= 1)
        {
            a0 += __shfl_xor(a0, o, 64);
            a1 += __shfl_xor(a1, o, 64);
            a2 += __shfl_xor(a2, o, 64);
        }
        if(sub == 0)
        {
            const int i = k * 96 + 3 * node;
            sR[i] = a0; sR[i + 1] = a1; sR[i + 2] = a2;
        }
        __syncthreads();  // the staging area is reused by the next bank
    }
    __syncthreads();  // residuals complete; the staging area is free
    // residual parts: r' = r 2^-e, hi = half(r'), lo = half(r' - hi)
    float ra = sR[wave * 96 + lane], rb = lane < 32 ? sR[wave * 96 + 64 + lane] : 0.f;
    float mx = fmaxf(fabsf(ra), fabsf(rb));
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    int ex = 0;
    if(mx > 0.f && isfinite(mx)) (void)frexpf(mx, &ex);
    const float down = ldexpf(1.f, -ex);
    ra *= down;
    rb *= down;
    {
        const E ha = (E)ra, hb = (E)rb;
        sB[wave][0][lane] = ha;
        sB[wave][1][lane] = (E)(ra - (float)ha);
        if(lane < 32)
        {
            sB[wave][0][64 + lane] = hb;
            sB[wave][1][64 + lane] = (E)(rb - (float)hb);
        }
    }
    // the wave's tile pairs into LDS for the transposed operands
    char* myT = sT + wave * kTrWaveBytes;
#pragma unroll
    for(int q = 0; q < kTilePairs; ++q)
    {
        const int t = 2 * q + (lane >> 5);
        if(tr_tile_kept(t)) *reinterpret_cast<uint4*>(myT + tr_tile_offset(t) + tr_chunk_offset(lane & 31)) = pr[q];
    }
    __syncthreads();
    // B operands: lane l holds B(8 (l / 16) + j, l % 16) of column pair P;
    // only column 0 is nonzero
    const bool col0 = (lane & 15) == 0;
    V8 bh[3], blo[3];
#pragma unroll
    for(int P = 0; P < 3; ++P)
    {
        const int k0 = 32 * P + 8 * (lane >> 4);
        const V8 h = *reinterpret_cast<const V8*>(&sB[wave][0][k0]);
        const V8 l = *reinterpret_cast<const V8*>(&sB[wave][1][k0]);
        const V8 z = {};
        bh[P]  = col0 ? h : z;
        blo[P] = col0 ? l : z;
    }
    mfma_f4 acc[6];
#pragma unroll
    for(int I = 0; I < 6; ++I)
    {
        acc[I] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for(int P = 0; P < 3; ++P)
        {
            V8 a;
            if(I <= 2 * P)
                a = __builtin_bit_cast(V8, pr[tile_pair_slot(I, P)]);
            else
            {
                // lanes 0-31: T(2P, I)^T; lanes 32-63: T(2P + 1, I)^T, or the
                // stored diagonal tile when 2P + 1 == I
                const bool upper_t = 2 * P + 1 < I && lane >= 32;
                const int  t = upper_t ? tile_pos(2 * P + 1, I) : tile_pos(2 * P, I);
                a = read_transposed<V8>(myT + tr_tile_offset(t), lane & 31);
                if(2 * P + 1 == I && lane >= 32) a = __builtin_bit_cast(V8, sg[I / 2]);
            }
            acc[I] = Op::mma(a, bh[P], acc[I]);
            acc[I] = Op::mma(a, blo[P], acc[I]);
        }
    }
    // column 0 of each D: lanes 0, 16, 32, 48 hold rows 4 (l / 16) .. + 3
    if(live && col0)
    {
        const float sc = d.minv_scale[b] * ldexpf(1.f, ex);
        float* dst = (kind == 0 ? d.z : d.Zc) + 3 * (long long)vec + 4 * (lane >> 4);
#pragma unroll
        for(int I = 0; I < 6; ++I)
        {
            float4 o;
            o.x = acc[I][0] * sc; o.y = acc[I][1] * sc; o.z = acc[I][2] * sc; o.w = acc[I][3] * sc;
            *reinterpret_cast<float4*>(dst + 16 * I) = o;
        }
    }
}

// Convergence decision after the previous update (or init / restart), then
// finish the preconditioner: prolong the coarse corrections into the level-0
// result (MAS rows), 12x12 block-Jacobi (ABD bodies), identity or 3x3
// block-Jacobi (plain rows); partial r.z.  The row data (and the coarse
// corrections they need) are loaded while the r.r partials are reduced.
__global__ void __launch_bounds__(256) k_precond_finish(DevData d, int force)
{
    // rows by XCD as in the SpMV (each XCD a contiguous run of workgroups)
    const int wgi = xcd_tile(blockIdx.x, gridDim.x);
    __shared__ double smem[8];
    const int4 w = d.wg_pre[wgi];
    const int  s = d.single ? 0 : w.x;
    SysState&  st = d.state[s];
    if(!force && st.status != kRunning) return;
    const double share = force ? 0.0 : partial_share(d.part_upd, wg_range(d, d.wg_upd_range, d.n_wg_upd, s));
    const int type = w.y >> 28;
    // row loads, independent of the decision
    const int row = w.z + threadIdx.x;
    const bool row_ok = type != kWgAbd && row < w.w;
    const long long o = 3 * (long long)(row_ok ? row : w.z);
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, z0 = 0.f, z1 = 0.f, z2 = 0.f;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    float dg[9];
    int mode = 0;
    // ABD bodies: 12 threads per body, thread `dof` holds row dof of its 12x12 inverse
    const int  abd_body = w.z + (int)threadIdx.x / 12, abd_dof = (int)threadIdx.x % 12;
    const bool abd_ok = type == kWgAbd && (int)threadIdx.x < 12 * kBodiesPerWg && abd_body < w.w;
    float abd_p[12], abd_r[12];
    long long abd_base = 0;
    if(abd_ok)
    {
        const float* P = d.abd_inv + 144 * (long long)abd_body + 12 * abd_dof;
        abd_base = 3 * (long long)d.body_row[abd_body];
#pragma unroll
        for(int j = 0; j < 12; ++j)
        {
            abd_p[j] = P[j];
            abd_r[j] = d.r[abd_base + j];
        }
    }
    if(row_ok)
    {
        r0 = d.r[o]; r1 = d.r[o + 1]; r2 = d.r[o + 2];
        if(type == kWgFemMas)
        {
            z0 = d.z[o]; z1 = d.z[o + 1]; z2 = d.z[o + 2];
            // the row's level-1 node and that node's ancestors (a table of the
            // coarse nodes, shared by the rows of a component)
            const int g1 = d.row_l1[row];
            int4 a = make_int4(-1, -1, -1, -1), e = make_int4(-1, -1, -1, -1);
            if(g1 >= 0)
            {
                a = reinterpret_cast<const int4*>(d.anc)[2 * (long long)g1];
                e = reinterpret_cast<const int4*>(d.anc)[2 * (long long)g1 + 1];
            }
            const int ids[8] = {a.x, a.y, a.z, a.w, e.x, e.y, e.z, e.w};
            float g[8][3];
#pragma unroll
            for(int k = 0; k < 8; ++k)
            {
                const long long gi = 3 * (long long)(ids[k] >= 0 ? ids[k] : 0);
                g[k][0] = d.Zc[gi]; g[k][1] = d.Zc[gi + 1]; g[k][2] = d.Zc[gi + 2];
            }
#pragma unroll
            for(int k = 0; k < 8; ++k)
                if(ids[k] >= 0)
                {
                    c0 += g[k][0]; c1 += g[k][1]; c2 += g[k][2];
                }
        }
        else
        {
            mode = d.row_mode[row];
            if(mode)
            {
                const float* D = d.diag_inv + 9 * (long long)row;
#pragma unroll
                for(int q = 0; q < 9; ++q) dg[q] = D[q];
            }
        }
    }
    if(!force)
    {
        const double rr = block_sum(share, smem);
        int ns = kRunning, fail = kOk, chk = 0;
        if(!st.frozen)
        {
            if(!finite_d(rr)) { ns = kDone; fail = kNonFinite; }
            else if(st.rule == kRuleRefRtz)
            {
                // pcg_solver.cu:277: the gate tests the r.z of the previous iteration
                if(st.k >= 1 && fabs(st.rz_cur) <= st.tol * st.rz0) ns = kDone;
                else if(st.k >= 1 && st.k + 1 >= st.max_iter) { ns = kDone; fail = kIterationCap; }
            }
            else
            {
                if(rr <= st.tol * st.tol * st.bb) { ns = kCandidate; chk = 1; }
                else if(st.k >= st.max_iter) { ns = kDone; fail = kIterationCap; }
                else if(st.check_every > 0 && st.k - st.k_check >= st.check_every) chk = 1;
            }
        }
        if((int)wgi == wg_range(d, d.wg_pre_range, d.n_wg_pre, s).x && threadIdx.x == 0)
        {
            st.rr = rr;
            if(ns != kRunning) st.status = ns;
            if(fail != kOk) st.failure = fail;
            if(chk) st.need_check = 1;
        }
        if(ns != kRunning) return;
    }
    double part = 0.0;
    if(type == kWgAbd)
    {
        if(abd_ok)
        {
            float acc = 0.f;
#pragma unroll
            for(int j = 0; j < 12; ++j) acc += abd_p[j] * abd_r[j];
            d.z[abd_base + abd_dof] = acc;
            part = (double)abd_r[abd_dof] * (double)acc;
        }
    }
    else if(row_ok)
    {
        if(type == kWgFemMas)
        {
            z0 += c0; z1 += c1; z2 += c2;
        }
        else if(mode)
        {
            z0 = dg[0] * r0 + dg[1] * r1 + dg[2] * r2;
            z1 = dg[3] * r0 + dg[4] * r1 + dg[5] * r2;
            z2 = dg[6] * r0 + dg[7] * r1 + dg[8] * r2;
        }
        else
        {
            z0 = r0; z1 = r1; z2 = r2;
        }
        d.z[o] = z0; d.z[o + 1] = z1; d.z[o + 2] = z2;
        part = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
    }
    if(force) return;
    part = block_sum_last(part, smem);
    if(threadIdx.x == 0) d.part_pre[wgi] = part;
}

// FP64 true residual r_true = b - A x of rows [r_begin, r_end) (x the FP32
// iterate), from the FP64 input blocks over the full block CSR; writes
// r_true and returns this thread's share of |r_true|^2.  With `drift`, also
// adds this thread's share of |r_true - r|^2 (r the recursive residual of the
// same iterate) to *drift.  G lanes share a row, G the largest power of two
// <= 64 with G * rows <= the workgroup (so the few long rows of a body tile
// are not one thread's serial loop): each lane sums a strided share of the
// row's entries and the group adds them by a butterfly, a fixed order for a
// given range, so the result is deterministic.
__device__ double true_residual_rows(const DevData& d, int r_begin, int r_end, double* drift = nullptr)
{
    const int nr = r_end - r_begin;
    int G = 1;
    while(G < 64 && 2 * G * nr <= (int)blockDim.x) G *= 2;
    const int sub = (int)threadIdx.x % G, per = (int)blockDim.x / G;
    double part = 0.0;
    for(int base = 0; base < nr; base += per)  // uniform trip count: the butterflies see whole groups
    {
        const int  row   = r_begin + base + (int)threadIdx.x / G;
        const bool valid = row < r_end;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        const int e_end = valid ? d.row_ptr[row + 1] : 0;
        for(int e = valid ? d.row_ptr[row] + sub : 0; e < e_end; e += G)
        {
            const int code = d.entry_src[e];
            const double* B = d.vals64 + 9 * (long long)(code & 0x7fffffff);
            const long long c = 3 * (long long)d.col[e];
            const double x0 = d.x[c], x1 = d.x[c + 1], x2 = d.x[c + 2];
            if(code >= 0)
            {
                a0 += B[0] * x0 + B[3] * x1 + B[6] * x2;
                a1 += B[1] * x0 + B[4] * x1 + B[7] * x2;
                a2 += B[2] * x0 + B[5] * x1 + B[8] * x2;
            }
            else
            {
                a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
                a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
                a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
            }
        }
        for(int off = G >> 1; off > 0; off >>= 1)
        {
            a0 += __shfl_xor(a0, off, 64);
            a1 += __shfl_xor(a1, off, 64);
            a2 += __shfl_xor(a2, off, 64);
        }
        if(valid && sub == 0)
        {
            const long long o = 3 * (long long)row;
            const double t0 = d.b64[o] - a0, t1 = d.b64[o + 1] - a1, t2 = d.b64[o + 2] - a2;
            if(drift)
            {
                const double e0 = t0 - (double)d.r[o], e1 = t1 - (double)d.r[o + 1], e2 = t2 - (double)d.r[o + 2];
                *drift += e0 * e0 + e1 * e1 + e2 * e2;
            }
            d.rt[o] = (float)t0; d.rt[o + 1] = (float)t1; d.rt[o + 2] = (float)t2;
            part += t0 * t0 + t1 * t1 + t2 * t2;
        }
    }
    return part;
}

// The final report can reuse the last check when it was taken at the final x
// (a TRUE-rule system confirmed done by a true-residual check).
__device__ __forceinline__ bool report_is_current(const SysState& st)
{
    return st.rule == kRuleTrueRel && st.status == kDone && st.n_true_checks > 0 && st.k_check == st.k &&
           st.failure == kOk;
}

template <int U>
struct SpmvChunk
{
    int   cix[U];    // column words of the thread's run of U consecutive entries
    float v[U][9];
    float g[U][6];   // z and old p of a column outside the tile
};

// a load of streamed data: nontemporal (NT) or ordinary
template <bool NT, typename V>
__device__ __forceinline__ V ld_stream(const V* p)
{
    if constexpr(NT)
        return ld_nt(p);
    else
        return *p;
}

// U consecutive floats (ints) of the streamed layout from a run start aligned
// to min(U, 4) entries
template <int U, bool NT, typename T>
__device__ __forceinline__ void load_run(const T* __restrict__ p, T (&o)[U])
{
    if constexpr(U == 1)
        o[0] = ld_stream<NT>(p);
    else if constexpr(U == 2)
    {
        using V2 = std::conditional_t<std::is_same_v<T, int>, int2, float2>;
        const V2 t = ld_stream<NT>(reinterpret_cast<const V2*>(p));
        o[0] = t.x; o[1] = t.y;
    }
    else
    {
        using V4 = std::conditional_t<std::is_same_v<T, int>, int4, float4>;
#pragma unroll
        for(int h = 0; h < U / 4; ++h)
        {
            const V4 t = ld_stream<NT>(reinterpret_cast<const V4*>(p) + h);
            o[4 * h] = t.x; o[4 * h + 1] = t.y; o[4 * h + 2] = t.z; o[4 * h + 3] = t.w;
        }
    }
}

// This thread's run: entries [base + U t, base + U t + U) of the chunk (those
// at or past n are read -- the layout's spare entries keep that in bounds --
// and ignored)
template <int U, bool NT>
__device__ __forceinline__ void spmv_load_chunk(const DevData& d, const float* __restrict__ p_old, int base, int n,
                                                SpmvChunk<U>& c)
{
    const int k0 = U * (int)threadIdx.x;
    if(k0 >= n)
    {
#pragma unroll
        for(int j = 0; j < U; ++j) c.cix[j] = 0;
        return;
    }
    const int e = base + k0;
    load_run<U, NT>(d.sy_col + e, c.cix);
#pragma unroll
    for(int q = 0; q < 9; ++q)
    {
        float t[U];
        load_run<U, NT>(d.vals32 + q * d.plane + e, t);
#pragma unroll
        for(int j = 0; j < U; ++j) c.v[j][q] = t[j];
    }
#pragma unroll
    for(int j = 0; j < U; ++j)
    {
        // columns inside the tile come from its x in LDS, not from memory
        const unsigned w = (unsigned)c.cix[j];
        if(!(w & kSpmvInTile) && k0 + j < n)
        {
            const long long x = 3 * (long long)(w & kSpmvColMask);
            c.g[j][0] = d.z[x]; c.g[j][1] = d.z[x + 1]; c.g[j][2] = d.z[x + 2];
            c.g[j][3] = p_old[x]; c.g[j][4] = p_old[x + 1]; c.g[j][5] = p_old[x + 2];
        }
    }
}

// Products of the thread's run, summed per row in registers (a row-start
// flag ends the running sum); each sum goes to LDS at the last entry of its
// row within the run.  An entry inside the tile takes x_c from the tile's x
// (sX) and, off the diagonal, also gives A^T x_r for its column (slot tpos).
// Then each owning row adds the run ends of its entries in order, then
// (tb < te) its transposed products.
template <int U>
__device__ __forceinline__ void spmv_reduce_chunk(const SpmvChunk<U>& c, float beta, int base, int n, bool has_row,
                                                  int rb, int re, int tb, int te, const float* sX, float* sP,
                                                  float* sQ, float acc[3])
{
    const int k0 = U * (int)threadIdx.x;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
    for(int j = 0; j < U; ++j)
    {
        const int k = k0 + j;
        if(k < n)
        {
            const float* v = c.v[j];
            const unsigned w = (unsigned)c.cix[j];
            if(j > 0 && (w & kSpmvRowStart))
            {
                sP[3 * (k - 1)] = s0; sP[3 * (k - 1) + 1] = s1; sP[3 * (k - 1) + 2] = s2;
                s0 = s1 = s2 = 0.f;
            }
            float x0, x1, x2;
            if(!(w & kSpmvInTile))
            {
                x0 = c.g[j][0] + beta * c.g[j][3];
                x1 = c.g[j][1] + beta * c.g[j][4];
                x2 = c.g[j][2] + beta * c.g[j][5];
            }
            else
            {
                const int cl = w & 0xff, rl = (w >> 8) & 0xff, tpos = (w >> 16) & 0x3fff;
                x0 = sX[3 * cl]; x1 = sX[3 * cl + 1]; x2 = sX[3 * cl + 2];
                if(rl != cl)
                {
                    const float y0 = sX[3 * rl], y1 = sX[3 * rl + 1], y2 = sX[3 * rl + 2];
                    sQ[3 * tpos]     = v[0] * y0 + v[3] * y1 + v[6] * y2;
                    sQ[3 * tpos + 1] = v[1] * y0 + v[4] * y1 + v[7] * y2;
                    sQ[3 * tpos + 2] = v[2] * y0 + v[5] * y1 + v[8] * y2;
                }
            }
            s0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            s1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
            s2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
        }
    }
    if(k0 < n)
    {
        const int last = min(n, k0 + U) - 1;
        sP[3 * last] = s0; sP[3 * last + 1] = s1; sP[3 * last + 2] = s2;
    }
    __syncthreads();
    if(has_row)
    {
        int q        = max(rb, base) - base;
        const int qb = min(re, base + n) - base;
        while(q < qb)
        {
            const int end = min(qb, (q / U + 1) * U);
            acc[0] += sP[3 * (end - 1)];
            acc[1] += sP[3 * (end - 1) + 1];
            acc[2] += sP[3 * (end - 1) + 2];
            q = end;
        }
        for(int t = tb; t < te; ++t)
        {
            acc[0] += sQ[3 * t];
            acc[1] += sQ[3 * t + 1];
            acc[2] += sQ[3 * t + 2];
        }
    }
    __syncthreads();
}

// Ap = A p with p = z + beta p_old formed on the fly (and stored), over one
// tile of the SpMV layout: at most 256 rows, one thread per row; a tile of
// several rows fits one chunk of 256 * U entries, a longer single row takes
// several chunks (it has no transposed products).  Also the p.Ap partial.
// NT: the matrix is read with nontemporal loads (the choice for matrices far
// larger than the caches).  BODIES: the layout has tiles of affine bodies
// (their 12 x 12 couplings' registers would otherwise set every tile's
// register count and occupancy).
template <int U, bool NT, bool BODIES>
__global__ void __launch_bounds__(256) k_spmv_fused(DevData d, const float* __restrict__ p_old,
                                                    float* __restrict__ p_new)
{
    constexpr int CH = 256 * U;
    __shared__ float  sP[CH * 3];
    __shared__ float  sQ[CH * 3];
    __shared__ float  sX[kSpmvTileRows * 3];
    __shared__ double smem[8];
    const int  wg = xcd_tile(blockIdx.x, gridDim.x);
    const int4 w  = d.wg_spmv[wg];
    const int  s  = d.single ? 0 : w.x;
    SysState&  st = d.state[s];
    if(st.status != kRunning)
    {
        // a convergence candidate from this iteration's preconditioner step:
        // the tile's rows of the FP64 true residual, decided by k_update
        if(st.status == kCandidate && st.need_check == 1)
        {
            double part = true_residual_rows(d, w.z, w.w);
            part = block_sum_last(part, smem);
            if(threadIdx.x == 0) d.part_tr[wg] = part;
        }
        return;
    }
    // state read before any store of this kernel (scalar loads)
    const int    st_rule = st.rule, st_frozen = st.frozen, st_beta_zero = st.beta_zero, st_k = st.k,
                 st_restarts = st.restarts;
    const double st_rz_prev = st.rz_prev;
    const double share = partial_share(d.part_pre, wg_range(d, d.wg_pre_range, d.n_wg_pre, s));
    const int r0 = w.z, r1 = w.w;
    const int2 er = d.wg_spmv_ent[wg];
    const int e0 = er.x, e1 = er.y;
    const int row = r0 + (int)threadIdx.x;
    const bool has_row = row < r1;
    int rb = 0, re = 0, tb = 0, te = 0;
    float own[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if(has_row)
    {
        rb = d.sy_row_ptr[row];
        re = d.sy_row_ptr[row + 1];
        const int t0 = d.sy_tptr[r0];
        tb = d.sy_tptr[row] - t0;
        te = d.sy_tptr[row + 1] - t0;
        const long long o = 3 * (long long)row;
        own[0] = d.z[o]; own[1] = d.z[o + 1]; own[2] = d.z[o + 2];
        own[3] = p_old[o]; own[4] = p_old[o + 1]; own[5] = p_old[o + 2];
    }
    SpmvChunk<U> c;
    const int n0 = min(CH, e1 - e0);
    if(n0 > 0) spmv_load_chunk<U, NT>(d, p_old, e0, n0, c);
    const double rz = block_sum(share, smem);
    const bool bad = !finite_d(rz) || (st_rule == kRuleTrueRel && rz <= 0.0);
    if(wg == wg_range(d, d.wg_spmv_range, d.n_wg_spmv, s).x && threadIdx.x == 0)
    {
        if(bad && !st_frozen)
        {
            st.status = kDone;
            st.failure = kBadRz;
        }
        st.rz_cur = rz;
        st.beta = st_beta_zero ? 0.0 : rz / st_rz_prev;
        if(st_k == 0 && st_restarts == 0) st.rz0 = rz;
    }
    if(bad && !st_frozen) return;
    const float beta = st_beta_zero ? 0.f : (float)(rz / st_rz_prev);
    // this row's p, for the store below and for the tile's own columns
    const float y0 = own[0] + beta * own[3];
    const float y1 = own[1] + beta * own[4];
    const float y2 = own[2] + beta * own[5];
    if(has_row)
    {
        sX[3 * threadIdx.x] = y0; sX[3 * threadIdx.x + 1] = y1; sX[3 * threadIdx.x + 2] = y2;
    }
    __syncthreads();
    float acc[3] = {0.f, 0.f, 0.f};
    if(n0 > 0) spmv_reduce_chunk<U>(c, beta, e0, n0, has_row, rb, re, tb, te, sX, sP, sQ, acc);
    // a single row longer than one chunk
#pragma unroll 1
    for(int base = e0 + CH; base < e1; base += CH)
    {
        const int n = min(CH, e1 - base);
        spmv_load_chunk<U, NT>(d, p_old, base, n, c);
        spmv_reduce_chunk<U>(c, beta, base, n, has_row, rb, re, 0, 0, sX, sP, sQ, acc);
    }
    if(BODIES && (w.y >> 28) == kWgAbd)  // uniform: a tile of affine bodies
    {
        // their 12 x 12 couplings: 48 threads per body, thread 48 b + 12 g + q
        // takes scalar row q of the tile's body b over every fourth of its
        // couplings (g, g + 4, ...), one coupling block row (three float4
        // loads) each, x from the tile's own p or gathered as z + beta p_old
        // (float4 loads: a body's 12 values start 16-byte aligned); the four
        // partial sums of a row are added in group order, after the row's
        // 3 x 3 entries
        constexpr int G = 4, BATCH = 2;
        const int b0 = w.y & 0x0fffffff, nbod = (r1 - r0) / 4;
        const int bl = (int)threadIdx.x / (12 * G), g = ((int)threadIdx.x / 12) % G, q = (int)threadIdx.x % 12;
        float a = 0.f;
        if(bl < nbod)
        {
            const int b = b0 + bl;
            const int k1 = d.abd_cp_ptr[b + 1];
            // BATCH couplings at a time: their column rows, then every value and
            // x load, then the products (the loads of a batch overlap)
#pragma unroll 1
            for(int kb = d.abd_cp_ptr[b] + g; kb < k1; kb += G * BATCH)
            {
                int    jr[BATCH];
                float4 v[BATCH][3], zz[BATCH][3], pp[BATCH][3];
#pragma unroll
                for(int u = 0; u < BATCH; ++u)
                {
                    const int k = kb + G * u;
                    jr[u] = k < k1 ? d.abd_cp_row[k] : -1;
                }
#pragma unroll
                for(int u = 0; u < BATCH; ++u)
                {
                    const int k = kb + G * u;
                    if(jr[u] < 0) continue;
                    const float4* V = reinterpret_cast<const float4*>(d.vals_abd + 144 * (long long)k + 12 * q);
                    v[u][0] = ld_stream<NT>(V); v[u][1] = ld_stream<NT>(V + 1); v[u][2] = ld_stream<NT>(V + 2);
                    if(jr[u] < r0 || jr[u] >= r1)
                    {
                        const float4* Zg = reinterpret_cast<const float4*>(d.z + 3 * (long long)jr[u]);
                        const float4* Pg = reinterpret_cast<const float4*>(p_old + 3 * (long long)jr[u]);
                        zz[u][0] = Zg[0]; zz[u][1] = Zg[1]; zz[u][2] = Zg[2];
                        pp[u][0] = Pg[0]; pp[u][1] = Pg[1]; pp[u][2] = Pg[2];
                    }
                }
#pragma unroll
                for(int u = 0; u < BATCH; ++u)
                {
                    if(jr[u] < 0) continue;
                    float x[12];
                    if(jr[u] >= r0 && jr[u] < r1)
                    {
#pragma unroll
                        for(int c = 0; c < 12; ++c) x[c] = sX[3 * (jr[u] - r0) + c];
                    }
                    else
                    {
#pragma unroll
                        for(int c = 0; c < 3; ++c)
                        {
                            x[4 * c]     = zz[u][c].x + beta * pp[u][c].x;
                            x[4 * c + 1] = zz[u][c].y + beta * pp[u][c].y;
                            x[4 * c + 2] = zz[u][c].z + beta * pp[u][c].z;
                            x[4 * c + 3] = zz[u][c].w + beta * pp[u][c].w;
                        }
                    }
                    a += v[u][0].x * x[0] + v[u][0].y * x[1] + v[u][0].z * x[2] + v[u][0].w * x[3] +
                         v[u][1].x * x[4] + v[u][1].y * x[5] + v[u][1].z * x[6] + v[u][1].w * x[7] +
                         v[u][2].x * x[8] + v[u][2].y * x[9] + v[u][2].z * x[10] + v[u][2].w * x[11];
                }
            }
        }
        sP[threadIdx.x] = a;  // the chunk sums above are consumed
        __syncthreads();
        if(has_row)
        {
            const int lr = row - r0, base = 12 * G * (lr / 4) + 3 * (lr % 4);
#pragma unroll
            for(int gg = 0; gg < G; ++gg)
            {
                acc[0] += sP[base + 12 * gg];
                acc[1] += sP[base + 12 * gg + 1];
                acc[2] += sP[base + 12 * gg + 2];
            }
        }
    }
    double part = 0.0;
    if(has_row)
    {
        const long long o = 3 * (long long)row;
        p_new[o] = y0; p_new[o + 1] = y1; p_new[o + 2] = y2;
        d.Ap[o] = acc[0]; d.Ap[o + 1] = acc[1]; d.Ap[o + 2] = acc[2];
        part = (double)y0 * acc[0] + (double)y1 * acc[1] + (double)y2 * acc[2];
    }
    part = block_sum_last(part, smem);
    if(threadIdx.x == 0) d.part_spmv[wg] = part;
}

// mode 0: alpha from the p.Ap partials, x += alpha p, r -= alpha Ap;
// mode 1 (init / restart): r := rt.  Then the level-1 restriction of the MAS
// rows and the partial r.r (b.b in FP64 at init).  The row vectors and the
// component lists are loaded while the p.Ap partials are reduced.
__global__ void __launch_bounds__(256) k_update(DevData d, const float* __restrict__ p, int mode)
{
    // rows by XCD as in the SpMV (each XCD a contiguous run of workgroups)
    const int wgi = xcd_tile(blockIdx.x, gridDim.x);
    __shared__ double smem[8];
    const int4 w = d.wg_upd[wgi];
    const int  s = d.single ? 0 : w.x;
    SysState&  st = d.state[s];
    bool init = false;
    if(mode == 0)
    {
        if(st.status != kRunning)
        {
            // confirm a candidate whose true residual the SpMV just formed:
            // done if it meets the tolerance, else left to the chunk's check
            // (need_check 2: no second in-iteration attempt)
            if(st.status == kCandidate && st.need_check == 1 &&
               (int)wgi == wg_range(d, d.wg_upd_range, d.n_wg_upd, s).x)
            {
                const double tr = sys_sum(d.part_tr, wg_range(d, d.wg_spmv_range, d.n_wg_spmv, s), smem);
                if(threadIdx.x == 0)
                {
                    if(tr <= st.tol * st.tol * st.bb)
                    {
                        st.true_rr = tr;
                        st.n_true_checks += 1;
                        st.k_check = st.k;
                        st.need_check = 0;
                        st.status = kDone;
                    }
                    else
                        st.need_check = 2;
                }
            }
            return;
        }
    }
    else
    {
        if(st.status != kRestarting && st.status != kInit) return;
        init = st.status == kInit;
    }
    // read before this kernel's first store, so it stays a scalar load
    const int4 grp = d.wg_upd_grp[wgi];  // uniform over the workgroup
    const double share = mode == 0 ? partial_share(d.part_spmv, wg_range(d, d.wg_spmv_range, d.n_wg_spmv, s)) : 0.0;
    const int type = w.y >> 28;
    const int aux  = w.y & 0x0fffffff;
    const int row  = w.z + threadIdx.x;
    const bool row_ok = row < w.w;
    const long long o = 3 * (long long)(row_ok ? row : w.z);
    float x[3] = {0.f, 0.f, 0.f}, pv[3] = {0.f, 0.f, 0.f}, rv[3] = {0.f, 0.f, 0.f}, ap[3] = {0.f, 0.f, 0.f};
    double bv[3] = {0.0, 0.0, 0.0};
    if(row_ok)
    {
        if(mode == 0)
        {
#pragma unroll
            for(int c = 0; c < 3; ++c)
            {
                x[c] = d.x[o + c]; pv[c] = p[o + c]; rv[c] = d.r[o + c]; ap[c] = d.Ap[o + c];
            }
        }
        else
        {
#pragma unroll
            for(int c = 0; c < 3; ++c) rv[c] = d.rt[o + c];
            if(init)
#pragma unroll
                for(int c = 0; c < 3; ++c) bv[c] = d.b64[o + c];
        }
    }
    // level-1 restriction set-up: the bank's lanes (a power of two, <= 64)
    const int bank_shift = __ffs(d.bank) - 1;
    const bool restrict_rows = type == kWgFemMas && w.z + (int)((threadIdx.x >> bank_shift) << bank_shift) < w.w;
    const int gb = aux + (int)(threadIdx.x >> bank_shift);
    int cb = 0, ce = 0;
    if(restrict_rows)
    {
        cb = d.comp_ptr[gb];
        ce = d.comp_ptr[gb + 1];
    }
    // lane j of a bank holds the bank's component cb + j (at most one per lane)
    const int blane = threadIdx.x & (d.bank - 1);
    const bool has_comp = restrict_rows && cb + blane < ce;
    uint32_t my_mask = 0;
    int      my_parent = 0;
    int4     my_ds = make_int4(-1, -1, -1, -1);
    if(has_comp)
    {
        my_mask   = d.comp_mask[cb + blane];
        my_parent = d.comp_parent[cb + blane];
        my_ds     = reinterpret_cast<const int4*>(d.comp_dslot)[cb + blane];
    }
    float alpha = 0.f;
    if(mode == 0)
    {
        const double pAp = block_sum(share, smem);
        const bool bad = !finite_d(pAp) || (st.rule == kRuleTrueRel && pAp <= 0.0);
        if((int)wgi == wg_range(d, d.wg_upd_range, d.n_wg_upd, s).x && threadIdx.x == 0)
        {
            if(bad && !st.frozen)
            {
                st.status = kDone;
                st.failure = kBadCurvature;
            }
            else
            {
                st.pAp = pAp;
                st.alpha = st.rz_cur / pAp;
                st.k = st.k + 1;
                st.rz_prev = st.rz_cur;
                st.beta_zero = 0;
            }
        }
        if(bad && !st.frozen) return;
        alpha = (float)(st.rz_cur / pAp);
    }
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    double part = 0.0;
    if(row_ok)
    {
        if(mode == 0)
        {
            d.x[o] = x[0] + alpha * pv[0];
            d.x[o + 1] = x[1] + alpha * pv[1];
            d.x[o + 2] = x[2] + alpha * pv[2];
            r0 = rv[0] - alpha * ap[0];
            r1 = rv[1] - alpha * ap[1];
            r2 = rv[2] - alpha * ap[2];
        }
        else
        {
            r0 = rv[0]; r1 = rv[1]; r2 = rv[2];
        }
        d.r[o] = r0; d.r[o + 1] = r1; d.r[o + 2] = r2;
        if(init)
            part = bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2];
        else
            part = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    }
    // level-1 restriction: sum the residual rows of each component of the bank
    __shared__ float sC[256 * 3];
    __shared__ int4  sDs[256];

    if(restrict_rows)
    {
        const int lane  = blane;
        const int gbase = (int)(threadIdx.x & 63) & ~(d.bank - 1);  // the bank's first lane in the wave
        for(int c = cb; c < ce; ++c)
        {
            const uint32_t mask   = (uint32_t)__shfl((int)my_mask, gbase + (c - cb), 64);
            const int      parent = __shfl(my_parent, gbase + (c - cb), 64);
            const bool in = (mask >> lane) & 1u;
            float sx = in ? r0 : 0.f, sy = in ? r1 : 0.f, sz = in ? r2 : 0.f;
            for(int off = d.bank >> 1; off > 0; off >>= 1)
            {
                sx += __shfl_xor(sx, off, 64);
                sy += __shfl_xor(sy, off, 64);
                sz += __shfl_xor(sz, off, 64);
            }
            if(lane == 0)
            {
                const long long pc = 3 * (long long)parent;
                d.Rc[pc] = sx; d.Rc[pc + 1] = sy; d.Rc[pc + 2] = sz;
                const int j = c - grp.z;
                sC[3 * j] = sx; sC[3 * j + 1] = sy; sC[3 * j + 2] = sz;
            }
        }
        if(has_comp && grp.y > 0) sDs[cb + blane - grp.z] = my_ds;
    }
    // the r.r partial of each wave goes to LDS ahead of the barrier the deeper
    // partials need; thread 0 adds the four after it
    part = wave_sum(part);
    if((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = part;
    __syncthreads();
    // one partial per deeper ancestor of the group's components (levels >= 2),
    // summed over the components in order
    if(grp.y > 0)
    {
        for(int t = threadIdx.x; t < grp.y; t += blockDim.x)
        {
            float ax = 0.f, ay = 0.f, az = 0.f;
            for(int c = grp.z; c < grp.w; ++c)
            {
                const int4 ds = sDs[c - grp.z];
                if(ds.x == t || ds.y == t || ds.z == t || ds.w == t)
                {
                    const int j = c - grp.z;
                    ax += sC[3 * j]; ay += sC[3 * j + 1]; az += sC[3 * j + 2];
                }
            }
            const long long o = 3 * (long long)(grp.x + t);
            d.part_deep[o] = ax; d.part_deep[o + 1] = ay; d.part_deep[o + 2] = az;
        }
    }
    if(threadIdx.x == 0)
    {
        double t = 0.0;
        for(int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += smem[w];
        d.part_upd[wgi] = t;
    }
}

// r_true = b - A x in FP64 from the FP64 input blocks (rows of the SpMV
// tiling) for systems that asked for a check (or all, when reporting).
__global__ void __launch_bounds__(256) k_true_residual(DevData d, int report_only)
{
    __shared__ double smem[8];
    const int4 w = d.wg_spmv[blockIdx.x];
    const int  s = w.x;
    const SysState& st = d.state[s];
    if(!report_only && !st.need_check) return;
    if(report_only && report_is_current(st)) return;
    double drift = 0.0;
    double part = true_residual_rows(d, w.z, w.w, &drift);
    part = block_sum(part, smem);
    drift = block_sum_last(drift, smem);
    if(threadIdx.x == 0)
    {
        d.part_tr[blockIdx.x] = part;
        d.part_drift[blockIdx.x] = drift;
    }
}

// One workgroup per system: true residual decision (done / restart on drift).
__global__ void __launch_bounds__(256) k_check_finish(DevData d, int report_only)
{
    __shared__ double smem[8];
    const int s = blockIdx.x;
    SysState& st = d.state[s];
    if(!report_only && !st.need_check) return;
    if(report_only && report_is_current(st)) return;
    const int2 tiles = wg_range(d, d.wg_spmv_range, d.n_wg_spmv, s);
    const double tr = sys_sum(d.part_tr, tiles, smem);
    // |r_true - r|^2: the true and the recursive residual of the same iterate
    const double drift2 = (!report_only && st.status == kRunning) ? sys_sum(d.part_drift, tiles, smem) : 0.0;
    if(threadIdx.x != 0) return;
    st.true_rr = tr;
    if(report_only) return;
    st.n_true_checks += 1;
    st.need_check = 0;
    st.k_check = st.k;
    const double tol2 = st.tol * st.tol * st.bb;
    bool restart = false;
    if(!isfinite(tr))
    {
        st.status = kDone;
        st.failure = kNonFinite;
        return;
    }
    if(st.status == kCandidate)
    {
        if(tr <= tol2) st.status = kDone;
        else
        {
            restart = true;
            st.rejected += 1;
        }
    }
    else if(st.status == kRunning)
    {
        // restart when the recursion has drifted from b - A x by more than 10 tol ||b||
        restart = drift2 > 100.0 * tol2;
    }
    if(restart)
    {
        if(st.k >= st.max_iter)
        {
            st.status = kDone;
            st.failure = kIterationCap;
        }
        else
        {
            st.status = kRestarting;
            st.beta_zero = 1;
            st.restarts += 1;
        }
    }
}

// ================================================================ FP64 API

// y_int = A x_int (FP64 full block CSR over the FP64 input blocks).
__global__ void __launch_bounds__(256) k_spmv64(DevData d, const double* x, double* y)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if(row >= d.n_rows) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for(int e = d.row_ptr[row]; e < d.row_ptr[row + 1]; ++e)
    {
        const int code = d.entry_src[e];
        const double* B = d.vals64 + 9 * (long long)(code & 0x7fffffff);
        const long long c = 3 * (long long)d.col[e];
        const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
        if(code >= 0)
        {
            a0 += B[0] * x0 + B[3] * x1 + B[6] * x2;
            a1 += B[1] * x0 + B[4] * x1 + B[7] * x2;
            a2 += B[2] * x0 + B[5] * x1 + B[8] * x2;
        }
        else
        {
            a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
            a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
            a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
        }
    }
    y[3 * (long long)row] = a0;
    y[3 * (long long)row + 1] = a1;
    y[3 * (long long)row + 2] = a2;
}

__global__ void __launch_bounds__(256) k_gather_ext64(const int* ext_to_int, const double* src_ext, double* dst_int,
                                                      int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const long long R = ext_to_int[i];
    for(int c = 0; c < 3; ++c) dst_int[3 * R + c] = src_ext[3 * (long long)i + c];
}

__global__ void __launch_bounds__(256) k_axpby_int_to_ext(const int* ext_to_int, const double* y_int, double* y_ext,
                                                          double a, double b, int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const long long R = ext_to_int[i];
    for(int c = 0; c < 3; ++c)
    {
        const long long o = 3 * (long long)i + c;
        y_ext[o] = a * y_int[3 * R + c] + (b != 0.0 ? b * y_ext[o] : 0.0);
    }
}

__global__ void __launch_bounds__(256) k_z32_to_ext(const int* ext_to_int, const float* z, double* dst_ext, int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const long long R = ext_to_int[i];
    for(int c = 0; c < 3; ++c) dst_ext[3 * (long long)i + c] = (double)z[3 * R + c];
}

// values[k] = src[indices[k]] for 9-double blocks (subsystem selection of
// LocalPreconditioner::calculate_subsystem_bcoo_indices, i_preconditioner.cu:54-89)
__global__ void __launch_bounds__(256) k_gather_blocks(const double* src, const uint32_t* idx, double* dst, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= 9 * n) return;
    const long long k = t / 9;
    dst[t] = src[9 * (long long)idx[k] + t % 9];
}

// Up to kCopySegments device arrays of doubles copied into one array in a
// single launch (a batch's per-system values): block b copies chunk
// b - first[s] of segment s, kCopyChunk doubles, read with nontemporal loads
// (each value is read once).
__global__ void __launch_bounds__(256) k_copy_segments(CopySegments c, double* __restrict__ out)
{
    const int b = (int)blockIdx.x;
    int lo = 0, hi = c.count - 1;  // the last segment whose first block is <= b
    while(lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if(c.first[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const long long base = (long long)(b - c.first[lo]) * kCopyChunk;
    const long long m    = min(c.n[lo] - base, (long long)kCopyChunk);
    const double* src = c.src[lo] + base;
    double*       dst = out + c.dst[lo] + base;
    constexpr int U = kCopyChunk / 256;
    double v[U];
#pragma unroll
    for(int u = 0; u < U; ++u)
    {
        const int i = (int)threadIdx.x + 256 * u;
        v[u] = i < m ? __builtin_nontemporal_load(src + i) : 0.0;
    }
#pragma unroll
    for(int u = 0; u < U; ++u)
    {
        const int i = (int)threadIdx.x + 256 * u;
        if(i < m) dst[i] = v[u];
    }
}

// ============================================================== launchers

#define MASX_CHECK_LAUNCH() \
    do { const hipError_t e_ = hipGetLastError(); if(e_ != hipSuccess) return e_; } while(0)

static inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_pack_values(const DevData& d, hipStream_t st)
{
    if(d.nnz_s <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_values, dim3(blocks_for(d.nnz_s, 256)), dim3(256), 0, st, d);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_pack_abd(const DevData& d, const int* cp_body, const int* cp_sub, long long n_couplings,
                           float* vals_abd, hipStream_t st)
{
    if(n_couplings <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_abd, dim3(blocks_for(16 * n_couplings, 256)), dim3(256), 0, st, d, cp_body, cp_sub,
                       16 * n_couplings, vals_abd);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_gal_gather(const double* fine, double* E, const int* ptr, const int* src, int e0, int e1,
                             int src_is_fine, int lanes, hipStream_t st)
{
    if(e1 <= e0) return hipSuccess;
    if(lanes == 64)
        hipLaunchKernelGGL(k_gal_gather<64>, dim3(blocks_for(e1 - e0, 4)), dim3(256), 0, st, fine, E, ptr, src, e0,
                           e1, src_is_fine);
    else if(lanes == 16)
        hipLaunchKernelGGL(k_gal_gather<16>, dim3(blocks_for(e1 - e0, 16)), dim3(256), 0, st, fine, E, ptr, src, e0,
                           e1, src_is_fine);
    else
        return hipErrorInvalidValue;
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

size_t mas_bank_bytes(int bank, int precision)
{
    if(bank == 32 && precision != kPrecFp32) return (size_t)kTileBankHalves * 2;  // either layout fits
    const size_t nblk = (size_t)bank * (bank + 1) / 2;
    return nblk * (precision == kPrecFp32 ? 9 : 5) * 4;
}

hipError_t launch_mas_factor(const DevData& d, const double* E, const int* ent_ptr, const int* ent_pos,
                             const int* ent_src, void* out, float* scale, int precision, int* status, int bank_begin,
                             int bank_end, hipStream_t st)
{
    if(bank_begin < 0 || bank_end > d.n_banks || bank_begin > bank_end) return hipErrorInvalidValue;
    if(bank_end == bank_begin) return hipSuccess;
    const dim3 g(bank_end - bank_begin);
    const bool wave = d.factor_waves == 1;  // one wave per bank, else two
#define MASX_FACTOR(BK, T)                                                                                          \
    do                                                                                                             \
    {                                                                                                              \
        if(wave)                                                                                                   \
            hipLaunchKernelGGL((k_mas_factor_wave<BK, T>), g, dim3(64), 0, st, d, E, ent_ptr, ent_pos, ent_src,    \
                               out, scale, status, bank_begin);                                                    \
        else                                                                                                       \
            hipLaunchKernelGGL((k_mas_factor_sym<BK, T>), g, dim3(128), 0, st, d, E, ent_ptr, ent_pos, ent_src,    \
                               out, scale, status, bank_begin);                                                    \
    } while(0)
    if(d.bank == 32)
    {
        if(precision == kPrecFp16) MASX_FACTOR(32, __half);
        else if(precision == kPrecBf16) MASX_FACTOR(32, __hip_bfloat16);
        else MASX_FACTOR(32, float);
    }
    else if(d.bank == 16)
    {
        if(precision == kPrecFp16) MASX_FACTOR(16, __half);
        else if(precision == kPrecBf16) MASX_FACTOR(16, __hip_bfloat16);
        else MASX_FACTOR(16, float);
    }
    else
        return hipErrorInvalidValue;
#undef MASX_FACTOR
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

int inverse_map_entries(int bank, int precision, bool tiles)
{
    const int nblk = bank * (bank + 1) / 2;
    if(precision == kPrecFp32) return nblk * 9;
    if(tiles && bank == 32) return kTileBankHalves;
    return nblk * 10;
}

hipError_t launch_inverse_map(int bank, int precision, bool tiles, unsigned short* map, hipStream_t st)
{
    const int n = inverse_map_entries(bank, precision, tiles);
    hipError_t e = hipMemsetAsync(map, 0xff, sizeof(unsigned short) * (size_t)n, st);
    if(e != hipSuccess) return e;
    const int ne = 3 * bank * (3 * bank + 1) / 2;
    const dim3 g((ne + 255) / 256);
#define MASX_MAP(BK, T) hipLaunchKernelGGL((k_inverse_map<BK, T>), g, dim3(256), 0, st, map, tiles ? 1 : 0)
    if(bank == 32)
    {
        if(precision == kPrecFp16) MASX_MAP(32, __half);
        else if(precision == kPrecBf16) MASX_MAP(32, __hip_bfloat16);
        else MASX_MAP(32, float);
    }
    else if(bank == 16)
    {
        if(precision == kPrecFp16) MASX_MAP(16, __half);
        else if(precision == kPrecBf16) MASX_MAP(16, __hip_bfloat16);
        else MASX_MAP(16, float);
    }
    else
        return hipErrorInvalidValue;
#undef MASX_MAP
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_abd_factor(const DevData& d, const int* body_src, float* inv, int n_bodies, hipStream_t st)
{
    if(n_bodies <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_abd_factor, dim3(n_bodies), dim3(64), 0, st, d, body_src, inv);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_diag_factor(const DevData& d, const int* diag_src, const int* needs, float* inv, hipStream_t st)
{
    if(d.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diag_factor, dim3(blocks_for(d.n_rows, 256)), dim3(256), 0, st, d, diag_src, needs, inv);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_load_rhs(const DevData& d, const int* e2i, const double* b, int n_ext, hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load_rhs, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, d, e2i, b, n_ext);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_store_x(const DevData& d, const int* e2i, double* x, int n_ext, hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_store_x, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, d, e2i, x, n_ext);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_init_state(const DevData& d, int rule, double tol, const int* max_iter, int max_override,
                             int check_every, int frozen, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_state, dim3(blocks_for(d.n_sys, 64)), dim3(64), 0, st, d, rule, tol, max_iter,
                       max_override, check_every, frozen);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_solve_prepare(const DevData& d, const int* int_to_ext, const double* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_solve_prepare, dim3(blocks_for(d.n_rows, 256)), dim3(256), 0, st, d, int_to_ext, b);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_export_status(const DevData& d, int* host, int* seq, hipStream_t st)
{
    hipLaunchKernelGGL(k_export_status, dim3(1), dim3(256), 0, st, d, host, seq);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_phase_finish(const DevData& d, hipStream_t st, int* host, int* seq)
{
    if(host && d.n_sys != 1) return hipErrorInvalidValue;  // the export counts one system
    hipLaunchKernelGGL(k_phase_finish, dim3(d.n_sys), dim3(256), 0, st, d, host, seq);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

// plain launch, or one whose events are stamped by the dispatch itself
#define MASX_DISPATCH(K, G, B, ...)                                                        \
    do                                                                                      \
    {                                                                                       \
        if(ev) hipExtLaunchKernelGGL(K, G, B, 0, st, ev->start, ev->stop, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(K, G, B, 0, st, __VA_ARGS__);                               \
    } while(0)

hipError_t launch_mas_local(const DevData& d, int precision, int force, hipStream_t st, const LaunchEvents* ev)
{
    if(d.n_banks <= 0) return hipSuccess;
    const bool mfma = d.bank == 32 && precision != kPrecFp32 && d.mas_waves == 1;
    const bool staged = precision == kPrecFp32 && (d.mas_waves == 5 || d.mas_waves == 9);
    if(!mfma && !staged && d.mas_waves != 4 && d.mas_waves != 8) return hipErrorInvalidValue;
    const int bpw = 64 / d.bank;
    const dim3 g(blocks_for(d.n_banks, bpw)), b(256);
    const dim3 b8(512);
#define MASX_LOCAL(BK, T)                                                          \
    do                                                                             \
    {                                                                              \
        if(d.mas_waves == 8) MASX_DISPATCH((k_mas_local<BK, T, 8>), g, b8, d, force); \
        else MASX_DISPATCH((k_mas_local<BK, T, 4>), g, b, d, force);                \
    } while(0)
#define MASX_LOCAL_F32(BK)                                                                      \
    do                                                                                          \
    {                                                                                           \
        if(d.mas_waves == 9) MASX_DISPATCH((k_mas_local<BK, float, 8, true>), g, b8, d, force);    \
        else if(d.mas_waves == 5) MASX_DISPATCH((k_mas_local<BK, float, 4, true>), g, b, d, force); \
        else MASX_LOCAL(BK, float);                                                             \
    } while(0)
    if(d.bank == 32)
    {
        const dim3 gm(blocks_for(d.n_banks, 4));
        if(mfma && precision == kPrecFp16) MASX_DISPATCH((k_mas_local_mfma<__half>), gm, b, d, force);
        else if(mfma) MASX_DISPATCH((k_mas_local_mfma<__hip_bfloat16>), gm, b, d, force);
        else if(precision == kPrecFp16) MASX_LOCAL(32, __half);
        else if(precision == kPrecBf16) MASX_LOCAL(32, __hip_bfloat16);
        else MASX_LOCAL_F32(32);
    }
    else if(d.bank == 16)
    {
        if(precision == kPrecFp16) MASX_LOCAL(16, __half);
        else if(precision == kPrecBf16) MASX_LOCAL(16, __hip_bfloat16);
        else MASX_LOCAL_F32(16);
    }
    else
        return hipErrorInvalidValue;
#undef MASX_LOCAL
#undef MASX_LOCAL_F32
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_precond_finish(const DevData& d, int n_wg, int force, hipStream_t st, const LaunchEvents* ev)
{
    if(n_wg <= 0) return hipSuccess;
    MASX_DISPATCH(k_precond_finish, dim3(n_wg), dim3(256), d, force);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

bool spmv_tiling_supported(int entries_per_thread)
{
    return entries_per_thread == 2 || entries_per_thread == 4 || entries_per_thread == 8;
}

hipError_t launch_spmv_fused(const DevData& d, int n_wg, int entries_per_thread, const float* p_old, float* p_new,
                             hipStream_t st, const LaunchEvents* ev)
{
    if(n_wg <= 0) return hipSuccess;
    void (*kern)(DevData, const float*, float*) = nullptr;
    const bool nt = d.spmv_nt != 0, bodies = d.spmv_bodies != 0;
#define MASX_SPMV_PICK(UU)                                                                 \
    kern = bodies ? (nt ? k_spmv_fused<UU, true, true> : k_spmv_fused<UU, false, true>) \
                  : (nt ? k_spmv_fused<UU, true, false> : k_spmv_fused<UU, false, false>)
    switch(entries_per_thread)
    {
        case 2: MASX_SPMV_PICK(2); break;
        case 4: MASX_SPMV_PICK(4); break;
        case 8: MASX_SPMV_PICK(8); break;
        default: return hipErrorInvalidValue;
    }
#undef MASX_SPMV_PICK
    MASX_DISPATCH(kern, dim3(n_wg), dim3(256), d, p_old, p_new);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_update(const DevData& d, int n_wg, const float* p, int mode, hipStream_t st, const LaunchEvents* ev)
{
    if(n_wg <= 0) return hipSuccess;
    MASX_DISPATCH(k_update, dim3(n_wg), dim3(256), d, p, mode);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}
#undef MASX_DISPATCH

hipError_t launch_true_residual(const DevData& d, int n_wg, int report_only, hipStream_t st)
{
    if(n_wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_true_residual, dim3(n_wg), dim3(256), 0, st, d, report_only);
    MASX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_check_finish, dim3(d.n_sys), dim3(256), 0, st, d, report_only);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_spmv64(const DevData& d, const double* x, double* y, hipStream_t st)
{
    if(d.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spmv64, dim3(blocks_for(d.n_rows, 256)), dim3(256), 0, st, d, x, y);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_gather_ext64(const int* e2i, const double* src, double* dst, int n_ext, hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_ext64, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, e2i, src, dst, n_ext);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_axpby_int_to_ext(const int* e2i, const double* y_int, double* y_ext, double a, double b, int n_ext,
                                   hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_axpby_int_to_ext, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, e2i, y_int, y_ext, a, b,
                       n_ext);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_z32_to_ext(const int* e2i, const float* z, double* dst, int n_ext, hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_z32_to_ext, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, e2i, z, dst, n_ext);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_gather_blocks(const double* src, const uint32_t* idx, double* dst, long long n, hipStream_t st)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_blocks, dim3(blocks_for(9 * n, 256)), dim3(256), 0, st, src, idx, dst, n);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_copy_segments(const CopySegments& c, double* out, hipStream_t st)
{
    if(c.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_segments, dim3((unsigned)c.first[c.count]), dim3(256), 0, st, c, out);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

}  // namespace masx
