// HIP kernels for the MAS-preconditioned PCG on gfx950 (CDNA4, wave64).
//
// Conventions
//  * vectors are float3 per internal block row (12 bytes, array of structs);
//  * matrix values for the SpMV are nine float planes (entry-major SoA) so that
//    lanes reading consecutive entries issue fully coalesced dword loads;
//  * every reduction is deterministic: per-workgroup partial sums in FP64 are
//    combined in a fixed order by the last workgroup of each system (detected
//    with a per-system counter), never with floating point atomics;
//  * kernels of a finished system return at once, so a captured chunk of
//    iterations can run past convergence at negligible cost.
#include "kernels.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace masx
{
namespace
{

__device__ __forceinline__ int packed_index(int bank, int r, int c) { return bank * r - r * (r + 1) / 2 + c; }

template <typename T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <> __device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }

template <typename T> __device__ __forceinline__ T from_d(double v);
template <> __device__ __forceinline__ float from_d<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __half from_d<__half>(double v) { return __float2half((float)v); }
template <> __device__ __forceinline__ __hip_bfloat16 from_d<__hip_bfloat16>(double v)
{
    return __float2bfloat16((float)v);
}

// wave64 butterfly sum
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide FP64 sum in a fixed order; result valid in thread 0.
__device__ double block_sum(double v, double* smem)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if(lane == 0) smem[wid] = v;
    __syncthreads();
    double t = 0.0;
    if(threadIdx.x == 0)
        for(int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += smem[w];
    __syncthreads();
    return t;
}

// Publishes this workgroup's partial and reports whether it is the last
// workgroup of system `s` for counter `cnt`.  In the last workgroup every
// thread has executed an acquire fence, so other workgroups' global writes
// are visible.
__device__ bool publish_partial(const DevData& d, int s, int cnt, double v, int nwg, bool* flag)
{
    if(threadIdx.x == 0)
    {
        d.partial[blockIdx.x] = v;
        __threadfence();
        const unsigned prev = atomicAdd(&d.state[s].counter[cnt], 1u);
        *flag = prev == (unsigned)(nwg - 1);
    }
    __syncthreads();
    const bool last = *flag;
    if(last) __threadfence();
    return last;
}

// Sum of the partials of workgroups [b, e) in a fixed order (whole block
// participates; result valid in all threads).
__device__ double sum_partials(const DevData& d, int b, int e, double* smem)
{
    double v = 0.0;
    for(int i = b + (int)threadIdx.x; i < e; i += blockDim.x) v += d.partial[i];
    v = block_sum(v, smem);
    if(threadIdx.x == 0) smem[8] = v;
    __syncthreads();
    v = smem[8];
    __syncthreads();
    return v;
}

__device__ __forceinline__ bool finite_d(double v) { return isfinite(v); }

}  // namespace

// ===================================================================== build

// Upper FP64 blocks (col-major) -> FP32 SoA planes of the full block CSR.
__global__ void __launch_bounds__(256) k_pack_values(DevData d, const int* dest1, const int* dest2, long long nb)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(g >= nb) return;
    const double* B = d.vals64 + 9 * g;
    float v[9];
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) v[3 * i + j] = (float)B[i + 3 * j];
    float* planes = const_cast<float*>(d.vals32);
    const int e1 = dest1[g], e2 = dest2[g];
    if(e1 >= 0)
#pragma unroll
        for(int c = 0; c < 9; ++c) planes[c * d.plane + e1] = v[c];
    if(e2 >= 0)
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j) planes[(3 * i + j) * d.plane + e2] = v[3 * j + i];
}

// Galerkin coarse entries of one level: E[e] = sum of mapped sources.
__global__ void __launch_bounds__(256) k_gal_gather(const double* fine, double* E, const int* ptr, const int* src,
                                                    int e0, int e1, int src_is_fine)
{
    const int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= e1) return;
    double acc[9];
#pragma unroll
    for(int i = 0; i < 9; ++i) acc[i] = 0.0;
    for(int k = ptr[e]; k < ptr[e + 1]; ++k)
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
#pragma unroll
    for(int i = 0; i < 9; ++i) E[9 * (long long)e + i] = acc[i];
}

// Dense local assembly, FP64 Cholesky with a pivot floor
// (pro/.../TraditionalMAS32Preconditioner.cu:1382-1439), triangular inverse
// and A^-1 = L^-T L^-1, stored as packed upper 3x3 blocks (block row-major)
// in the storage precision T.  One workgroup per bank; the symmetric matrix
// lives in LDS as a packed lower triangle.
template <int BANK, typename T>
__global__ void __launch_bounds__(256) k_mas_factor(DevData d, const double* E, const int* ent_ptr, const int* ent_pos,
                                                    const int* ent_src, T* out, int* status)
{
    constexpr int DIM  = 3 * BANK;
    constexpr int NL   = DIM * (DIM + 1) / 2;
    constexpr int NBLK = BANK * (BANK + 1) / 2;
    __shared__ double L[NL];
    __shared__ double colbuf[DIM];
    __shared__ double red[2];
    const int bank = blockIdx.x;
    const int tid  = threadIdx.x;
    auto at = [](int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; };
    for(int i = tid; i < NL; i += blockDim.x) L[i] = 0.0;
    __syncthreads();
    const bool fine = bank < d.n_fine_banks;
    for(int k = ent_ptr[bank] + tid; k < ent_ptr[bank + 1]; k += blockDim.x)
    {
        const int pos = ent_pos[k] & 0xffff, tr = ent_pos[k] >> 16;
        const int src = ent_src[k];
        int I = 0;
        while(I + 1 < BANK && packed_index(BANK, I + 1, I + 1) <= pos) ++I;
        const int J = pos - packed_index(BANK, I, I) + I;
        for(int a = 0; a < 3; ++a)
            for(int b = 0; b < 3; ++b)
            {
                double v;
                if(fine)
                    v = tr ? d.vals64[9 * (long long)src + b + 3 * a] : d.vals64[9 * (long long)src + a + 3 * b];
                else
                    v = E[9 * (long long)src + 3 * a + b];
                const int p = 3 * I + a, q = 3 * J + b;
                if(I == J && p < q) continue;  // diagonal block: its lower half carries it
                atomicAdd(&L[at(p, q)], v);
            }
    }
    __syncthreads();
    // zero diagonal -> 1 (padding slots behave as identity), pivot floor
    if(tid < DIM && L[at(tid, tid)] == 0.0) L[at(tid, tid)] = 1.0;
    __syncthreads();
    if(tid == 0)
    {
        double s = 0.0;
        for(int i = 0; i < DIM; ++i) s = fmax(s, fabs(L[at(i, i)]));
        red[0] = fmax(1.0e-12, 1.0e-12 * s);
    }
    __syncthreads();
    const double floor_value = red[0];
    int bad = 0;
    for(int k = 0; k < DIM; ++k)
    {
        if(tid == 0)
        {
            double v = L[at(k, k)];
            if(!isfinite(v)) { v = floor_value; ++bad; }
            if(v < floor_value) { v = floor_value; ++bad; }
            L[at(k, k)] = sqrt(v);
        }
        __syncthreads();
        const double piv = L[at(k, k)];
        for(int i = k + 1 + tid; i < DIM; i += blockDim.x)
        {
            double v = L[at(i, k)] / piv;
            if(!isfinite(v)) v = 0.0;
            L[at(i, k)] = v;
            colbuf[i] = v;
        }
        __syncthreads();
        // trailing update of the lower triangle: 8 row groups x 32 column lanes
        for(int i = k + 1 + (tid >> 5); i < DIM; i += (int)(blockDim.x >> 5))
        {
            const double ci = colbuf[i];
            for(int j = k + 1 + (tid & 31); j <= i; j += 32) L[i * (i + 1) / 2 + j] -= ci * colbuf[j];
        }
        __syncthreads();
    }
    // in-place inverse of the lower factor, columns from last to first
    for(int j = DIM - 1; j >= 0; --j)
    {
        if(tid == 0) L[at(j, j)] = 1.0 / L[at(j, j)];
        for(int i = j + 1 + tid; i < DIM; i += blockDim.x) colbuf[i] = L[at(i, j)];
        __syncthreads();
        const double ajj = -L[at(j, j)];
        for(int i = j + 1 + tid; i < DIM; i += blockDim.x)
        {
            double s = 0.0;
            for(int k = j + 1; k <= i; ++k) s += L[at(i, k)] * colbuf[k];
            L[at(i, j)] = ajj * s;
        }
        __syncthreads();
    }
    // A^-1(p,q) = sum_{k >= max(p,q)} X(k,p) X(k,q), X = L^-1 (lower)
    T* dst = out + (size_t)bank * NBLK * 9;
    for(int t = tid; t < NBLK * 9; t += blockDim.x)
    {
        const int blk = t / 9, e = t % 9;
        int I = 0;
        while(I + 1 < BANK && packed_index(BANK, I + 1, I + 1) <= blk) ++I;
        const int J = blk - packed_index(BANK, I, I) + I;
        const int p = 3 * I + e / 3, q = 3 * J + e % 3;
        const int k0 = p > q ? p : q;
        double s = 0.0;
        for(int k = k0; k < DIM; ++k) s += L[at(k, p)] * L[at(k, q)];
        dst[t] = from_d<T>(s);
    }
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

__global__ void k_init_state(DevData d, int rule, double tol, const int* max_iter, int check_every)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= d.n_sys) return;
    SysState& st = d.state[s];
    st.rz = st.rz0 = st.pAp = st.alpha = st.beta = st.rr = st.bb = st.true_rr = 0.0;
    st.tol = tol;
    st.k = 0;
    st.status = kInit;
    st.failure = kOk;
    st.rule = rule;
    st.max_iter = max_iter[s];
    st.k_check = 0;
    st.check_every = check_every;
    st.beta_zero = 1;
    st.need_check = 0;
    st.restarts = 0;
    st.n_true_checks = 0;
    for(int c = 0; c < 6; ++c) st.counter[c] = 0u;
}

// ========================================================= iteration kernels

// Coarse restriction of levels >= 2 for system s, by one workgroup:
// R(parent) = sum of the R of the component's lanes in the bank below.
__device__ void restrict_coarse_levels(const DevData& d, int s)
{
    const int lvl_b = d.sys_level_ptr[s], lvl_e = d.sys_level_ptr[s + 1];
    const int group = threadIdx.x >> 5, lane = threadIdx.x & 31, ngroups = blockDim.x >> 5;
    for(int lv = lvl_b; lv + 1 < lvl_e; ++lv)  // banks of level lv feed level lv+1
    {
        const int b0 = d.sys_level_banks[2 * lv], b1 = d.sys_level_banks[2 * lv + 1];
        for(int b = b0 + group; b < b1; b += ngroups)
        {
            const int base = d.bank_vec[b] + lane;
            const float rx = d.Rc[3 * base + 0], ry = d.Rc[3 * base + 1], rz = d.Rc[3 * base + 2];
            for(int c = d.comp_ptr[b]; c < d.comp_ptr[b + 1]; ++c)
            {
                const bool in = (d.comp_mask[c] >> lane) & 1u;
                float sx = in ? rx : 0.f, sy = in ? ry : 0.f, sz = in ? rz : 0.f;
#pragma unroll
                for(int o = 16; o > 0; o >>= 1)
                {
                    sx += __shfl_xor(sx, o, 32);
                    sy += __shfl_xor(sy, o, 32);
                    sz += __shfl_xor(sz, o, 32);
                }
                if(lane == 0)
                {
                    const int pc = d.comp_parent[c];
                    d.Rc[3 * pc + 0] = sx;
                    d.Rc[3 * pc + 1] = sy;
                    d.Rc[3 * pc + 2] = sz;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// MAS local solve: Z = Minv * R for every bank of every level.  Each
// workgroup stages its banks' packed inverses and residuals in LDS with
// 16-byte loads, then thread t of a bank forms scalar row t of the product,
// using the stored upper blocks directly and transposed for the lower part.
template <int BANK, typename T>
__global__ void __launch_bounds__(192) k_mas_local(DevData d, int force)
{
    constexpr int DIM  = 3 * BANK;
    constexpr int NBLK = BANK * (BANK + 1) / 2;
    constexpr int BPW  = 192 / DIM;                  // banks per workgroup
    constexpr int VALS = NBLK * 9;                   // values per bank
    constexpr int VEC16 = VALS * (int)sizeof(T) / 16;
    static_assert(VALS * sizeof(T) % 16 == 0, "bank storage must be a multiple of 16 bytes");
    __shared__ uint4 smat[BPW * VEC16];
    __shared__ float sR[BPW * DIM];
    const int bank0 = blockIdx.x * BPW;
    const int nb = min(BPW, d.n_banks - bank0);
    // a workgroup's banks may belong to different systems
    bool any_live = force != 0;
    for(int i = 0; i < nb && !any_live; ++i) any_live = d.state[d.bank_sys[bank0 + i]].status == kRunning;
    if(!any_live) return;
    const uint4* g = reinterpret_cast<const uint4*>(reinterpret_cast<const T*>(d.minv) + (size_t)bank0 * VALS);
    for(int i = threadIdx.x; i < nb * VEC16; i += blockDim.x) smat[i] = g[i];
    for(int i = threadIdx.x; i < nb * DIM; i += blockDim.x)
    {
        const int b = bank0 + i / DIM;
        const float* src = b < d.n_fine_banks ? d.r : d.Rc;
        sR[i] = src[3 * (long long)d.bank_vec[b] + i % DIM];
    }
    __syncthreads();
    const int lb = threadIdx.x / DIM;
    if(lb >= nb) return;
    const int b = bank0 + lb;
    if(!force && d.state[d.bank_sys[b]].status != kRunning) return;
    const int t = threadIdx.x % DIM;
    const int I = t / 3, a = t % 3;
    const T*     M = reinterpret_cast<const T*>(smat) + (size_t)lb * VALS;
    const float* R = sR + lb * DIM;
    float acc = 0.f;
    int   pd  = packed_index(BANK, I, I);  // block (I, J) for J >= I is pd + (J - I)
    int   pt  = I;                          // block (J, I) for J < I, advanced per J
#pragma unroll 4
    for(int J = 0; J < BANK; ++J)
    {
        const bool up  = J >= I;
        const int  blk = up ? pd + (J - I) : pt;
        const int  off = up ? blk * 9 + 3 * a : blk * 9 + a;
        const int  st  = up ? 1 : 3;
        acc += to_f(M[off]) * R[3 * J] + to_f(M[off + st]) * R[3 * J + 1] + to_f(M[off + 2 * st]) * R[3 * J + 2];
        if(!up) pt += BANK - (J + 1);
    }
    float* dst = b < d.n_fine_banks ? d.z : d.Zc;
    dst[3 * (long long)d.bank_vec[b] + t] = acc;
}

// Finish the preconditioner: prolong coarse corrections into the level-0
// result (MAS rows), 12x12 block-Jacobi (ABD bodies), identity or 3x3
// block-Jacobi (plain rows); then r.z and beta.
__global__ void __launch_bounds__(256) k_precond_finish(DevData d, int force)
{
    __shared__ double smem[16];
    __shared__ bool   flag;
    const int4 w = d.wg_pre[blockIdx.x];
    const int  s = w.x;
    SysState&  st = d.state[s];
    if(!force && st.status != kRunning) return;
    const int type = w.y >= 0 ? (w.y >> 28) : 0;
    const int aux  = w.y & 0x0fffffff;
    double part = 0.0;
    if(type == kWgAbd)
    {
        const int t = threadIdx.x;
        const int body = w.z + t / 12, dof = t % 12;
        if(t < 12 * kBodiesPerWg && body < w.w)
        {
            const float* P  = d.abd_inv + 144 * (long long)body + 12 * dof;
            const long long base = 3 * (long long)d.body_row[body];
            float acc = 0.f;
#pragma unroll
            for(int j = 0; j < 12; ++j) acc += P[j] * d.r[base + j];
            d.z[base + dof] = acc;
            part = (double)d.r[base + dof] * (double)acc;
        }
    }
    else
    {
        const int row = w.z + threadIdx.x;
        if(row < w.w)
        {
            const long long o = 3 * (long long)row;
            const float r0 = d.r[o], r1 = d.r[o + 1], r2 = d.r[o + 2];
            float z0, z1, z2;
            if(type == kWgFemMas)
            {
                z0 = d.z[o]; z1 = d.z[o + 1]; z2 = d.z[o + 2];
                const int gb = aux + (int)(threadIdx.x >> 5);
                const int lane = threadIdx.x & 31;
                int c = -1;
                for(int k = d.comp_ptr[gb]; k < d.comp_ptr[gb + 1]; ++k)
                    if((d.comp_mask[k] >> lane) & 1u) { c = d.comp_parent[k]; break; }
                while(c >= 0)
                {
                    z0 += d.Zc[3 * c];
                    z1 += d.Zc[3 * c + 1];
                    z2 += d.Zc[3 * c + 2];
                    c = d.coarse_parent[c];
                }
            }
            else if(d.row_mode[row])
            {
                const float* D = d.diag_inv + 9 * (long long)row;
                z0 = D[0] * r0 + D[1] * r1 + D[2] * r2;
                z1 = D[3] * r0 + D[4] * r1 + D[5] * r2;
                z2 = D[6] * r0 + D[7] * r1 + D[8] * r2;
            }
            else
            {
                z0 = r0; z1 = r1; z2 = r2;
            }
            d.z[o] = z0; d.z[o + 1] = z1; d.z[o + 2] = z2;
            part = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
        }
    }
    if(force) return;
    part = block_sum(part, smem);
    const int* rg = d.wg_pre_range + 2 * s;
    if(!publish_partial(d, s, kCntPre, part, rg[1] - rg[0], &flag)) return;
    const double rz = sum_partials(d, rg[0], rg[1], smem);
    if(threadIdx.x == 0)
    {
        st.counter[kCntPre] = 0u;
        if(st.k == 0 && st.restarts == 0) st.rz0 = rz;
        if(!finite_d(rz) || (st.rule == kRuleTrueRel && rz <= 0.0))
        {
            st.status = kDone;
            st.failure = kBadRz;
        }
        st.beta = st.beta_zero ? 0.0 : rz / st.rz;
        st.beta_zero = 0;
        st.rz = rz;
    }
}

// Ap = A p_new with p_new = z + beta * p_old formed on the fly (the p update
// of the previous iteration is fused here), p.Ap, alpha.
template <int G>
__global__ void __launch_bounds__(256) k_spmv_fused(DevData d, const float* __restrict__ p_old,
                                                    float* __restrict__ p_new)
{
    __shared__ double smem[16];
    __shared__ bool   flag;
    const int4 w = d.wg_spmv[blockIdx.x];
    const int  s = w.x;
    SysState&  st = d.state[s];
    if(st.status != kRunning) return;
    const float beta = (float)st.beta;
    const int   sub = threadIdx.x % G;
    double      part = 0.0;
    for(int row = w.z + (int)threadIdx.x / G; row < w.w; row += 256 / G)
    {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        const int e_end = d.row_ptr[row + 1];
        for(int e = d.row_ptr[row] + sub; e < e_end; e += G)
        {
            const long long c = 3 * (long long)d.col[e];
            const float x0 = d.z[c] + beta * p_old[c];
            const float x1 = d.z[c + 1] + beta * p_old[c + 1];
            const float x2 = d.z[c + 2] + beta * p_old[c + 2];
            const float* v = d.vals32 + e;
            a0 += v[0] * x0 + v[d.plane] * x1 + v[2 * d.plane] * x2;
            a1 += v[3 * d.plane] * x0 + v[4 * d.plane] * x1 + v[5 * d.plane] * x2;
            a2 += v[6 * d.plane] * x0 + v[7 * d.plane] * x1 + v[8 * d.plane] * x2;
        }
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
        {
            a0 += __shfl_xor(a0, o, G);
            a1 += __shfl_xor(a1, o, G);
            a2 += __shfl_xor(a2, o, G);
        }
        if(sub == 0)
        {
            const long long o = 3 * (long long)row;
            const float q0 = d.z[o] + beta * p_old[o];
            const float q1 = d.z[o + 1] + beta * p_old[o + 1];
            const float q2 = d.z[o + 2] + beta * p_old[o + 2];
            p_new[o] = q0; p_new[o + 1] = q1; p_new[o + 2] = q2;
            d.Ap[o] = a0; d.Ap[o + 1] = a1; d.Ap[o + 2] = a2;
            part += (double)q0 * a0 + (double)q1 * a1 + (double)q2 * a2;
        }
    }
    part = block_sum(part, smem);
    const int* rg = d.wg_spmv_range + 2 * s;
    if(!publish_partial(d, s, kCntSpmv, part, rg[1] - rg[0], &flag)) return;
    const double pAp = sum_partials(d, rg[0], rg[1], smem);
    if(threadIdx.x == 0)
    {
        st.counter[kCntSpmv] = 0u;
        st.pAp = pAp;
        st.k += 1;
        if(!finite_d(pAp) || (st.rule == kRuleTrueRel && pAp <= 0.0))
        {
            st.status = kDone;
            st.failure = kBadCurvature;
            st.alpha = 0.0;
        }
        else
            st.alpha = st.rz / pAp;
    }
}

// x += alpha p, r -= alpha Ap (mode 0), or r := rt (restart / init, mode 1);
// level-1 restriction of the MAS rows; r.r; convergence decision; coarse
// restriction of levels >= 2 by the last workgroup of each system.
__global__ void __launch_bounds__(256) k_update(DevData d, const float* __restrict__ p, int mode)
{
    __shared__ double smem[16];
    __shared__ bool   flag;
    const int4 w = d.wg_upd[blockIdx.x];
    const int  s = w.x;
    SysState&  st = d.state[s];
    const int want = mode == 0 ? kRunning : -1;
    if(mode == 0 ? st.status != want : !(st.status == kRestarting || st.status == kInit)) return;
    const int   type  = w.y >= 0 ? (w.y >> 28) : 0;
    const int   aux   = w.y & 0x0fffffff;
    const float alpha = (float)st.alpha;
    const int   row   = w.z + threadIdx.x;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    double part = 0.0;
    if(row < w.w)
    {
        const long long o = 3 * (long long)row;
        if(mode == 0)
        {
            d.x[o]     += alpha * p[o];
            d.x[o + 1] += alpha * p[o + 1];
            d.x[o + 2] += alpha * p[o + 2];
            r0 = d.r[o] - alpha * d.Ap[o];
            r1 = d.r[o + 1] - alpha * d.Ap[o + 1];
            r2 = d.r[o + 2] - alpha * d.Ap[o + 2];
        }
        else
        {
            r0 = d.rt[o]; r1 = d.rt[o + 1]; r2 = d.rt[o + 2];
        }
        d.r[o] = r0; d.r[o + 1] = r1; d.r[o + 2] = r2;
        if(st.status == kInit)
            part = d.b64[o] * d.b64[o] + d.b64[o + 1] * d.b64[o + 1] + d.b64[o + 2] * d.b64[o + 2];
        else
            part = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    }
    if(type == kWgFemMas)
    {
        const int gb = aux + (int)(threadIdx.x >> 5);
        const int lane = threadIdx.x & 31;
        if(w.z + (int)(threadIdx.x & ~31) < w.w)
            for(int c = d.comp_ptr[gb]; c < d.comp_ptr[gb + 1]; ++c)
            {
                const bool in = (d.comp_mask[c] >> lane) & 1u;
                float sx = in ? r0 : 0.f, sy = in ? r1 : 0.f, sz = in ? r2 : 0.f;
#pragma unroll
                for(int o = 16; o > 0; o >>= 1)
                {
                    sx += __shfl_xor(sx, o, 32);
                    sy += __shfl_xor(sy, o, 32);
                    sz += __shfl_xor(sz, o, 32);
                }
                if(lane == 0)
                {
                    const int pc = d.comp_parent[c];
                    d.Rc[3 * pc] = sx; d.Rc[3 * pc + 1] = sy; d.Rc[3 * pc + 2] = sz;
                }
            }
    }
    part = block_sum(part, smem);
    const int* rg = d.wg_upd_range + 2 * s;
    if(!publish_partial(d, s, kCntUpd, part, rg[1] - rg[0], &flag)) return;
    const double rr = sum_partials(d, rg[0], rg[1], smem);
    __shared__ int next_status;
    if(threadIdx.x == 0)
    {
        st.counter[kCntUpd] = 0u;
        int ns = kRunning;
        if(st.status == kInit)
        {
            st.bb = rr;
            st.rr = rr;
            if(rr == 0.0) ns = kDone;  // zero right-hand side: x = 0
        }
        else if(st.status == kRestarting)
        {
            st.rr = rr;
        }
        else
        {
            st.rr = rr;
            if(!finite_d(rr))
            {
                ns = kDone;
                st.failure = kNonFinite;
            }
            else if(st.rule == kRuleRefRtz)
            {
                // pcg_solver.cu:277: the gate tests the r.z of the previous iteration
                if(fabs(st.rz) <= st.tol * st.rz0) ns = kDone;
                else if(st.k + 1 >= st.max_iter) { ns = kDone; st.failure = kIterationCap; }
            }
            else
            {
                if(rr <= st.tol * st.tol * st.bb) { ns = kCandidate; st.need_check = 1; }
                else if(st.k >= st.max_iter) { ns = kDone; st.failure = kIterationCap; }
                else if(st.check_every > 0 && st.k - st.k_check >= st.check_every) st.need_check = 1;
            }
        }
        st.status   = ns;
        next_status = ns;
    }
    __syncthreads();
    if(next_status == kRunning) restrict_coarse_levels(d, s);
}

// r_true = b - A x in FP64 from the FP64 input blocks (rows of the SpMV
// tiling) for systems that asked for a check; decides convergence / restart.
__global__ void __launch_bounds__(256) k_true_residual(DevData d, int report_only)
{
    __shared__ double smem[16];
    __shared__ bool   flag;
    const int4 w = d.wg_spmv[blockIdx.x];
    const int  s = w.x;
    SysState&  st = d.state[s];
    if(!report_only && !st.need_check) return;
    double part = 0.0;
    for(int row = w.z + (int)threadIdx.x; row < w.w; row += blockDim.x)
    {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for(int e = d.row_ptr[row]; e < d.row_ptr[row + 1]; ++e)
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
        const long long o = 3 * (long long)row;
        const double t0 = d.b64[o] - a0, t1 = d.b64[o + 1] - a1, t2 = d.b64[o + 2] - a2;
        d.rt[o] = (float)t0; d.rt[o + 1] = (float)t1; d.rt[o + 2] = (float)t2;
        part += t0 * t0 + t1 * t1 + t2 * t2;
    }
    part = block_sum(part, smem);
    const int* rg = d.wg_spmv_range + 2 * s;
    if(!publish_partial(d, s, kCntTrue, part, rg[1] - rg[0], &flag)) return;
    const double tr = sum_partials(d, rg[0], rg[1], smem);
    if(threadIdx.x == 0)
    {
        st.counter[kCntTrue] = 0u;
        st.true_rr = tr;
        if(report_only) return;
        st.n_true_checks += 1;
        st.need_check = 0;
        st.k_check = st.k;
        const double tol2 = st.tol * st.tol * st.bb;
        if(st.status == kCandidate)
        {
            if(tr <= tol2) st.status = kDone;
            else
            {
                st.status = kRestarting;
                st.beta_zero = 1;
                st.restarts += 1;
            }
        }
        else if(st.status == kRunning)
        {
            const double drift = fabs(sqrt(tr) - sqrt(st.rr));
            if(drift > 10.0 * st.tol * sqrt(st.bb))
            {
                st.status = kRestarting;
                st.beta_zero = 1;
                st.restarts += 1;
            }
        }
        if(st.status == kRestarting && st.k >= st.max_iter)
        {
            st.status = kDone;
            st.failure = kIterationCap;
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

__global__ void __launch_bounds__(256) k_ext_to_r32(const int* ext_to_int, const double* src_ext, float* r, int n_ext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n_ext) return;
    const long long R = ext_to_int[i];
    for(int c = 0; c < 3; ++c) r[3 * R + c] = (float)src_ext[3 * (long long)i + c];
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

// ============================================================== launchers

#define MASX_CHECK_LAUNCH() \
    do { const hipError_t e_ = hipGetLastError(); if(e_ != hipSuccess) return e_; } while(0)

static inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_pack_values(const DevData& d, const int* dest1, const int* dest2, long long nb, hipStream_t st)
{
    if(nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_values, dim3(blocks_for(nb, 256)), dim3(256), 0, st, d, dest1, dest2, nb);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_gal_gather(const double* fine, double* E, const int* ptr, const int* src, int e0, int e1,
                             int src_is_fine, hipStream_t st)
{
    if(e1 <= e0) return hipSuccess;
    hipLaunchKernelGGL(k_gal_gather, dim3(blocks_for(e1 - e0, 256)), dim3(256), 0, st, fine, E, ptr, src, e0, e1,
                       src_is_fine);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_mas_factor(const DevData& d, const double* E, const int* ent_ptr, const int* ent_pos,
                             const int* ent_src, void* out, int precision, int* status, hipStream_t st)
{
    if(d.n_banks <= 0) return hipSuccess;
    const dim3 g(d.n_banks), b(256);
#define MASX_FACTOR(BK, T) hipLaunchKernelGGL((k_mas_factor<BK, T>), g, b, 0, st, d, E, ent_ptr, ent_pos, ent_src, (T*)out, status)
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

hipError_t launch_init_state(const DevData& d, int rule, double tol, const int* max_iter, int check_every,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_init_state, dim3(blocks_for(d.n_sys, 64)), dim3(64), 0, st, d, rule, tol, max_iter,
                       check_every);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_mas_local(const DevData& d, int precision, int force, hipStream_t st)
{
    if(d.n_banks <= 0) return hipSuccess;
    const int bpw = 192 / (3 * d.bank);
    const dim3 g(blocks_for(d.n_banks, bpw)), b(192);
#define MASX_LOCAL(BK, T) hipLaunchKernelGGL((k_mas_local<BK, T>), g, b, 0, st, d, force)
    if(d.bank == 32)
    {
        if(precision == kPrecFp16) MASX_LOCAL(32, __half);
        else if(precision == kPrecBf16) MASX_LOCAL(32, __hip_bfloat16);
        else MASX_LOCAL(32, float);
    }
    else if(d.bank == 16)
    {
        if(precision == kPrecFp16) MASX_LOCAL(16, __half);
        else if(precision == kPrecBf16) MASX_LOCAL(16, __hip_bfloat16);
        else MASX_LOCAL(16, float);
    }
    else
        return hipErrorInvalidValue;
#undef MASX_LOCAL
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_precond_finish(const DevData& d, int n_wg, int force, hipStream_t st)
{
    if(n_wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_precond_finish, dim3(n_wg), dim3(256), 0, st, d, force);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_spmv_fused(const DevData& d, int n_wg, int lanes_per_row, const float* p_old, float* p_new,
                             hipStream_t st)
{
    if(n_wg <= 0) return hipSuccess;
    switch(lanes_per_row)
    {
        case 4: hipLaunchKernelGGL((k_spmv_fused<4>), dim3(n_wg), dim3(256), 0, st, d, p_old, p_new); break;
        case 8: hipLaunchKernelGGL((k_spmv_fused<8>), dim3(n_wg), dim3(256), 0, st, d, p_old, p_new); break;
        case 16: hipLaunchKernelGGL((k_spmv_fused<16>), dim3(n_wg), dim3(256), 0, st, d, p_old, p_new); break;
        case 32: hipLaunchKernelGGL((k_spmv_fused<32>), dim3(n_wg), dim3(256), 0, st, d, p_old, p_new); break;
        default: return hipErrorInvalidValue;
    }
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_update(const DevData& d, int n_wg, const float* p, int mode, hipStream_t st)
{
    if(n_wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_update, dim3(n_wg), dim3(256), 0, st, d, p, mode);
    MASX_CHECK_LAUNCH();
    return hipSuccess;
}

hipError_t launch_true_residual(const DevData& d, int n_wg, int report_only, hipStream_t st)
{
    if(n_wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_true_residual, dim3(n_wg), dim3(256), 0, st, d, report_only);
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

hipError_t launch_ext_to_r32(const int* e2i, const double* src, float* r, int n_ext, hipStream_t st)
{
    if(n_ext <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ext_to_r32, dim3(blocks_for(n_ext, 256)), dim3(256), 0, st, e2i, src, r, n_ext);
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

}  // namespace masx
