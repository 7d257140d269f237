// MAS hierarchy on the device: the rebuild the reference runs every Newton
// step with contacts (ReorderRealtime, TraditionalMAS32Preconditioner.cu:
// 2189-2230, with the component kernels :463-1057 and the collision
// connections :1550-1960).
//
// Level by level, with the level sizes kept on the device (no host round trip
// until the end): mesh edges and the vertex pairs of every contact set lane
// bits in per-slot adjacency masks (integer atomics); one wavefront per two
// banks closes the masks transitively with lane shuffles, so every slot ends
// with its component's mask; the lowest lane leads a component; a scan over
// the banks numbers the next level by (bank, leader lane).  The arrays equal
// the host build (build_hierarchy in plan.cpp) bit for bit.
#include "hierarchy.hpp"
#include "pair_code.hpp"
#include "plan.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace masx
{
namespace
{
constexpr int kMaxLevels = 6;

enum HierFlag : int { kBadPartition = 1, kBadNeighbor = 2, kBadList = 4, kAsymmetric = 8 };

// Device level table, 32 ints: count of level l at [l], first slot at
// [8 + l], error flags at [16], listed components at [17].
struct Lv
{
    int* v;
    __device__ int  count(int l) const { return v[l]; }
    __device__ int  begin(int l) const { return v[8 + l]; }
    __device__ bool failed() const { return v[16] != 0; }
};

__device__ __forceinline__ int pad_dev(int n, int b) { return (n + b - 1) / b * b; }

__device__ __forceinline__ uint32_t group_bits(unsigned long long ballot, int gbase, int bank)
{
    const uint32_t g = (uint32_t)(ballot >> gbase);
    return bank == 32 ? g : (g & ((1u << bank) - 1u));
}

// Exclusive scan of in[0..n) into out[0..n) by one 256-thread workgroup
// (8 consecutive items per thread); returns the total to every thread.
__device__ int block_scan_256x8(const int* in, int n, int* out)
{
    __shared__ int s[256];
    int carry = 0;
    for(int t0 = 0; t0 < n; t0 += 2048)
    {
        const int base = t0 + 8 * threadIdx.x;
        int v[8], sum = 0;
#pragma unroll
        for(int k = 0; k < 8; ++k)
        {
            v[k] = base + k < n ? in[base + k] : 0;
            sum += v[k];
        }
        s[threadIdx.x] = sum;
        __syncthreads();
        for(int o = 1; o < 256; o <<= 1)
        {
            const int add = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        int run = carry + s[threadIdx.x] - sum;
#pragma unroll
        for(int k = 0; k < 8; ++k)
        {
            if(base + k < n) out[base + k] = run;
            run += v[k];
        }
        carry += s[255];
        __syncthreads();
    }
    return carry;
}

// Level table (level 0 = n_slots slots at 0), going_next = -1, bank counts = 0.
__global__ void __launch_bounds__(256) k_h_init(Lv lv, int n_slots, long long max_slots, long long max_banks,
                                                int* going_next, int* ncomp_all)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if(t < 32) lv.v[t] = t == 0 ? n_slots : 0;
    if(t < max_slots) going_next[t] = -1;
    if(t < max_banks) ncomp_all[t] = 0;
}

// Input checks, the level-0 adjacency (own bit for slots holding a vertex) and
// the starting node of every vertex.
__global__ void __launch_bounds__(256) k_h_validate(Lv lv, int n_fem, int n_slots, int bank, const uint32_t* start,
                                                    const uint32_t* num, const uint32_t* list, long long list_len,
                                                    const int* p2r, const int* r2p, int* id, uint32_t* adj)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int flag = 0;
    if(t < n_slots)
    {
        const int r = p2r[t];
        if(r < -1 || r >= n_fem || (r >= 0 && r2p[r] != t)) flag |= kBadPartition;
        adj[t] = r >= 0 ? (1u << (t % bank)) : 0u;
    }
    if(t < n_fem)
    {
        const int s = r2p[t];
        if(s < 0 || s >= n_slots || p2r[s] != t) flag |= kBadPartition;
        id[t] = s >= 0 && s < n_slots ? s : 0;
        const unsigned long long b = start[t], c = num[t];
        if(b + c > (unsigned long long)list_len)
            flag |= kBadList;
        else
            for(unsigned long long i = 0; i < c; ++i)
            {
                const uint32_t u = list[b + i];
                if(u >= (uint32_t)n_fem)
                {
                    flag |= kBadNeighbor;
                    continue;
                }
                // t must appear in u's list (u's own range is checked by u's thread)
                const unsigned long long bu = start[u], cu = num[u];
                if(bu + cu > (unsigned long long)list_len) continue;
                bool back = false;
                for(unsigned long long k = 0; k < cu && !back; ++k) back = list[bu + k] == (uint32_t)t;
                if(!back) flag |= kAsymmetric;
            }
    }
    if(flag) atomicOr(&lv.v[16], flag);
}

// Levels with at most kLdsNodes nodes pool link masks per workgroup in LDS
// (fast shared-memory atomics) and flush each touched node word once.
constexpr int kLdsNodes = 4096;

__device__ __forceinline__ void lds_or_begin(uint32_t* acc, int count)
{
    for(int i = threadIdx.x; i < count; i += blockDim.x) acc[i] = 0u;
    __syncthreads();
}

__device__ __forceinline__ void lds_or_flush(const uint32_t* acc, int count, uint32_t* dst)
{
    __syncthreads();
    for(int i = threadIdx.x; i < count; i += blockDim.x)
    {
        const uint32_t m = acc[i];
        if(m) atomicOr(&dst[i], m);
    }
}

// ORs `m` into adj[a] once per distinct `a` of the wavefront: lanes sharing a
// target pool their bits and one lane issues the atomic (coarse levels map
// many fine vertices onto few nodes, and per-edge atomics on one word would
// serialize).  Lanes with m == 0 take no part.  Wave-uniform loop.
__device__ __forceinline__ void or_into_dedup(int a, uint32_t m, uint32_t* adj)
{
    const int lane = threadIdx.x & 63;
    bool act = m != 0u;
    while(true)
    {
        const unsigned long long bal = __ballot(act);
        if(bal == 0ull) break;
        const int leader = __ffsll((long long)bal) - 1;
        const int target = __shfl(a, leader, 64);
        const bool match = act && a == target;
        uint32_t v = match ? m : 0u;
        for(int o = 32; o >= 1; o >>= 1) v |= (uint32_t)__shfl_xor((int)v, o, 64);
        if(lane == leader) atomicOr(&adj[target], v);
        act = act && !match;
    }
}

// Mesh links at the current level, one thread per level-0 slot: the vertex
// ORs the lanes of its neighbours' nodes that share its node's bank (a
// symmetric list makes every edge count from both ends, as an undirected
// link), then the masks are pooled per node.  Slot offsets of every level are
// bank multiples, so global slots share a bank exactly when their
// level-local indices do.
__global__ void __launch_bounds__(256) k_h_mesh_links(Lv lv, int level, int n_slots, const int* p2r,
                                                      const uint32_t* start, const uint32_t* num, const uint32_t* list,
                                                      const int* id, int bank, uint32_t* adj)
{
    if(lv.failed()) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = -1;
    uint32_t m = 0u;
    const int v = s < n_slots ? p2r[s] : -1;
    if(v >= 0)
    {
        a = id[v];
        const uint32_t b0 = start[v], n = num[v];
        for(uint32_t i = 0; i < n; ++i)
        {
            const int b = id[list[b0 + i]];
            if(b != a && b / bank == a / bank) m |= 1u << (b % bank);
        }
    }
    const int count = lv.count(level), begin = lv.begin(level);
    if(level == 0)
    {
        if(m) atomicOr(&adj[a], m);  // one vertex per slot: no two lanes share a target
    }
    else if(count <= kLdsNodes)
    {
        __shared__ uint32_t acc[kLdsNodes];
        lds_or_begin(acc, count);
        if(m) atomicOr(&acc[a - begin], m);
        lds_or_flush(acc, count, adj + begin);
    }
    else
        or_into_dedup(a, m, adj);
}

// Contact links: every pair of distinct in-range vertices of a contact joins
// their nodes when those share a bank (both directions).
__global__ void __launch_bounds__(256) k_h_pair_links(Lv lv, int level, int n_pairs, const int* pairs, int node_offset, int n_fem,
                                                      const int* id, int bank, uint32_t* adj)
{
    if(lv.failed()) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int cp[4] = {-1, -1, -1, -1};
    uint32_t msk[4] = {0u, 0u, 0u, 0u};
    if(t < n_pairs)
    {
        int v[4];
        const int cnt = decode_pair_code(pairs + 4 * (long long)t, v);
        for(int i = 0; i < 4; ++i)
        {
            const int u = i < cnt ? v[i] - node_offset : -1;
            cp[i] = u >= 0 && u < n_fem ? id[u] : -1;
        }
        for(int i = 0; i < 4; ++i)
            for(int j = i + 1; j < 4; ++j)
                if(cp[i] >= 0 && cp[j] >= 0 && cp[i] != cp[j] && cp[i] / bank == cp[j] / bank)
                {
                    msk[i] |= 1u << (cp[j] % bank);
                    msk[j] |= 1u << (cp[i] % bank);
                }
    }
    const int count = lv.count(level), begin = lv.begin(level);
    if(count <= kLdsNodes)
    {
        __shared__ uint32_t acc[kLdsNodes];
        lds_or_begin(acc, count);
        for(int i = 0; i < 4; ++i)
            if(msk[i]) atomicOr(&acc[cp[i] - begin], msk[i]);
        lds_or_flush(acc, count, adj + begin);
    }
    else
        for(int i = 0; i < 4; ++i)
            if(msk[i]) atomicOr(&adj[cp[i]], msk[i]);
}

// One wavefront per 64/bank banks: transitive closure of the adjacency masks
// (each pass doubles the path length covered; 2^6 > 32 lanes), then the
// per-bank component count.
__global__ void __launch_bounds__(64) k_h_components(Lv lv, int level, int bank, const uint32_t* adj,
                                                     uint32_t* compm, int* ncomp)
{
    if(lv.failed()) return;
    const int count = lv.count(level), begin = lv.begin(level);
    const int lane = threadIdx.x, per = 64 / bank;
    const int bk = blockIdx.x * per + lane / bank, l = lane % bank;
    const int c = bk * bank + l;
    const bool in = c < pad_dev(count, bank);
    uint32_t m = in ? adj[begin + c] : 0u;
    const int gbase = lane & ~(bank - 1);
    for(int pass = 0; pass < 6; ++pass)
    {
        uint32_t r = m;
        for(int j = 0; j < bank; ++j)
        {
            const uint32_t mj = (uint32_t)__shfl((int)m, gbase + j, 64);
            if((m >> j) & 1u) r |= mj;
        }
        const bool grew = r != m;
        m = r;
        if(__ballot(grew) == 0ull) break;  // closed: a further pass changes nothing
    }
    if(in) compm[begin + c] = m;
    const bool leader = m != 0u && (int)__builtin_ctz(m) == l;
    const uint32_t mine = group_bits(__ballot(leader), gbase, bank);
    if(in && l == 0) ncomp[bk] = __popc(mine);
}

// Numbers the next level: exclusive scan of the per-bank component counts,
// next level's count and first slot.
__global__ void __launch_bounds__(256) k_h_scan_level(Lv lv, int level, int bank, int n_slots, const int* ncomp,
                                                      int* off)
{
    if(lv.failed()) return;
    const int count = lv.count(level), begin = lv.begin(level);
    const int total = block_scan_256x8(ncomp, pad_dev(count, bank) / bank, off);
    if(threadIdx.x == 0)
    {
        lv.v[level + 1]     = total;
        lv.v[8 + level + 1] = level == 0 ? pad_dev(n_slots, bank) : begin + pad_dev(count, bank);
    }
}

// Parent (global slot of the next-level node) of every slot; per-bank counts
// of the listed components (none for the top level); level-0 prefix.
__global__ void __launch_bounds__(64) k_h_assign(Lv lv, int level, int bank, int top, const uint32_t* compm,
                                                 const int* off, const int* ncomp, int* parent, int* ncomp_all,
                                                 int* prefix)
{
    if(lv.failed()) return;
    const int count = lv.count(level), begin = lv.begin(level), next_begin = lv.begin(level + 1);
    const int lane = threadIdx.x, per = 64 / bank;
    const int bk = blockIdx.x * per + lane / bank, l = lane % bank;
    const int c = bk * bank + l;
    const bool in = c < pad_dev(count, bank);
    const uint32_t m = in ? compm[begin + c] : 0u;
    const int gbase = lane & ~(bank - 1);
    const bool leader = m != 0u && (int)__builtin_ctz(m) == l;
    const uint32_t mine = group_bits(__ballot(leader), gbase, bank);
    if(!in) return;
    int p = -1;
    if(m != 0u)
    {
        const int lead = (int)__builtin_ctz(m);
        p = next_begin + off[bk] + __popc(mine & ((1u << lead) - 1u));
    }
    parent[begin + c] = p;
    if(l == 0)
    {
        ncomp_all[begin / bank + bk] = top ? 0 : ncomp[bk];
        if(level == 0) prefix[bk] = ncomp[bk];
    }
}

// Moves every vertex to its next-level node, records going_next (level 0 by
// real vertex, coarse levels by slot), the level-0 fine masks, and clears the
// next level's adjacency.
__global__ void __launch_bounds__(256) k_h_advance(Lv lv, int level, int bank, int n_fem, const int* r2p,
                                                   const uint32_t* compm, const int* parent, int* id,
                                                   uint32_t* fine_mask, int* going_next, uint32_t* adj)
{
    if(lv.failed()) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t < n_fem)
    {
        if(level == 0)
        {
            const int s = r2p[t];
            fine_mask[t]  = compm[s];
            going_next[t] = parent[s];
        }
        id[t] = parent[id[t]];
    }
    if(level > 0 && t < lv.count(level))
    {
        const int s   = lv.begin(level) + t;
        going_next[s] = parent[s];
    }
    const int next_count = lv.count(level + 1);
    if(t < pad_dev(next_count, bank)) adj[lv.begin(level + 1) + t] = t < next_count ? (1u << (t % bank)) : 0u;
}

__global__ void __launch_bounds__(256) k_h_coarse_table(Lv lv, int n_fem, int levels, const int* going_next,
                                                        int* table)
{
    if(lv.failed()) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= n_fem) return;
    int cur = v;
    for(int l = 0; l < kMaxLevels; ++l)
    {
        if(l < levels - 1) cur = going_next[cur];
        table[kMaxLevels * (long long)v + l] = l < levels - 1 ? cur : 0;
    }
}

__global__ void __launch_bounds__(256) k_h_comp_scan(Lv lv, int levels, int bank, const int* ncomp_all,
                                                     int* comp_ptr)
{
    if(lv.failed()) return;
    const int n_banks = lv.begin(levels) / bank;
    const int total   = block_scan_256x8(ncomp_all, n_banks, comp_ptr);
    if(threadIdx.x == 0)
    {
        comp_ptr[n_banks] = total;
        lv.v[17]          = total;
    }
}

// Component lists of every bank of levels 0..L-2: (lane mask, parent slot),
// leaders in lane order.
__global__ void __launch_bounds__(64) k_h_comp_lists(Lv lv, int levels, int bank, const uint32_t* compm,
                                                     const int* parent, const int* comp_ptr, const int* ncomp_all,
                                                     uint32_t* cmask, int* cparent)
{
    if(lv.failed()) return;
    const int n_banks = lv.begin(levels) / bank;
    const int lane = threadIdx.x, per = 64 / bank;
    const int bk = blockIdx.x * per + lane / bank, l = lane % bank;
    const bool in = bk < n_banks;
    const int s = bk * bank + l;
    const uint32_t m = in ? compm[s] : 0u;
    const int gbase = lane & ~(bank - 1);
    const bool leader = in && ncomp_all[bk] > 0 && m != 0u && (int)__builtin_ctz(m) == l;
    const uint32_t mine = group_bits(__ballot(leader), gbase, bank);
    if(leader)
    {
        const int k = comp_ptr[bk] + __popc(mine & ((1u << l) - 1u));
        cmask[k]    = m;
        cparent[k]  = parent[s];
    }
}

void hier_check(hipError_t e)
{
    if(e != hipSuccess) throw std::runtime_error(std::string("device hierarchy: ") + hipGetErrorString(e));
}

template <typename T>
void grow(T*& p, size_t& cap, size_t need)
{
    if(need <= cap) return;
    if(p) (void)hipFree(p);
    p   = nullptr;
    cap = 0;
    hier_check(hipMalloc((void**)&p, need * sizeof(T)));
    cap = need;
}

unsigned blocks(long long n, int b) { return (unsigned)std::max<long long>(1, (n + b - 1) / b); }
}  // namespace

DeviceHierarchy::~DeviceHierarchy()
{
    for(void* p : {(void*)lv_, (void*)id_, (void*)adj_, (void*)compm_, (void*)parent_, (void*)ncomp_, (void*)off_,
                   (void*)ncomp_all_, (void*)comp_ptr_, (void*)cmask_, (void*)cparent_, (void*)fine_mask_,
                   (void*)prefix_, (void*)going_next_, (void*)table_})
        if(p) (void)hipFree(p);
    if(h_lv_) (void)hipHostFree(h_lv_);
}

void DeviceHierarchy::build(int n_fem, int n_slots, int bank, int levels, const uint32_t* start, const uint32_t* num,
                            const uint32_t* list, long long list_len, const int* p2r, const int* r2p,
                            const int* pairs, int n_pairs, int node_offset, hipStream_t st)
{
    if(bank != 32 && bank != 16) throw std::invalid_argument("bank size must be 32 or 16");
    if(n_fem <= 0 || n_slots < n_fem || n_slots % bank) throw std::invalid_argument("bad vertex or partition size");
    if(n_pairs < 0 || (n_pairs > 0 && !pairs)) throw std::invalid_argument("bad contact pairs");
    if(!start || !num || !p2r || !r2p || list_len < 0 || (list_len > 0 && !list))
        throw std::invalid_argument("null topology array");
    const int L = levels > 0 ? std::min(levels, kMaxLevels) : reference_level_count(n_fem, n_slots, bank);
    // every level holds at most n_slots nodes (a level never has more nodes
    // than the one below), so L levels fit in (L + 1) * n_slots slots
    const size_t max_slots = (size_t)(L + 1) * n_slots + (size_t)bank * (L + 2);
    const size_t max_banks = max_slots / bank + 1;
    grow(lv_, lv_cap_, 32);
    grow(id_, id_cap_, (size_t)n_fem);
    grow(adj_, adj_cap_, max_slots);
    grow(compm_, compm_cap_, max_slots);
    grow(parent_, parent_cap_, max_slots);
    grow(ncomp_, ncomp_cap_, (size_t)n_slots / bank + 1);
    grow(off_, off_cap_, (size_t)n_slots / bank + 1);
    grow(ncomp_all_, ncomp_all_cap_, max_banks);
    grow(comp_ptr_, comp_ptr_cap_, max_banks + 1);
    grow(cmask_, cmask_cap_, max_slots);
    grow(cparent_, cparent_cap_, max_slots);
    grow(going_next_, going_next_cap_, max_slots);
    grow(fine_mask_, fine_mask_cap_, (size_t)n_fem);
    grow(prefix_, prefix_cap_, (size_t)n_slots / bank);
    grow(table_, table_cap_, (size_t)kMaxLevels * n_fem);
    if(!h_lv_) hier_check(hipHostMalloc((void**)&h_lv_, 32 * sizeof(int), hipHostMallocDefault));

    const Lv lv{lv_};
    const int per = 64 / bank;
    hipLaunchKernelGGL(k_h_init, dim3(blocks((long long)max_slots, 256)), dim3(256), 0, st, lv, n_slots,
                       (long long)max_slots, (long long)max_banks, going_next_, ncomp_all_);
    hipLaunchKernelGGL(k_h_validate, dim3(blocks(std::max(n_fem, n_slots), 256)), dim3(256), 0, st, lv, n_fem,
                       n_slots, bank, start, num, list, list_len, p2r, r2p, id_, adj_);
    // grids are sized by the bound n_slots; kernels read the level sizes
    for(int level = 0; level < L; ++level)
    {
        hipLaunchKernelGGL(k_h_mesh_links, dim3(blocks(n_slots, 256)), dim3(256), 0, st, lv, level, n_slots, p2r,
                           start, num, list, id_, bank, adj_);
        if(n_pairs > 0)
            hipLaunchKernelGGL(k_h_pair_links, dim3(blocks(n_pairs, 256)), dim3(256), 0, st, lv, level, n_pairs, pairs,
                               node_offset, n_fem, id_, bank, adj_);
        hipLaunchKernelGGL(k_h_components, dim3(blocks(n_slots / bank, per)), dim3(64), 0, st, lv, level, bank,
                           adj_, compm_, ncomp_);
        hipLaunchKernelGGL(k_h_scan_level, dim3(1), dim3(256), 0, st, lv, level, bank, n_slots, ncomp_, off_);
        hipLaunchKernelGGL(k_h_assign, dim3(blocks(n_slots / bank, per)), dim3(64), 0, st, lv, level, bank,
                           level == L - 1 ? 1 : 0, compm_, off_, ncomp_, parent_, ncomp_all_, prefix_);
        hipLaunchKernelGGL(k_h_advance, dim3(blocks(std::max(n_fem, n_slots), 256)), dim3(256), 0, st, lv, level,
                           bank, n_fem, r2p, compm_, parent_, id_, fine_mask_, going_next_, adj_);
    }
    hipLaunchKernelGGL(k_h_coarse_table, dim3(blocks(n_fem, 256)), dim3(256), 0, st, lv, n_fem, L, going_next_,
                       table_);
    hipLaunchKernelGGL(k_h_comp_scan, dim3(1), dim3(256), 0, st, lv, L, bank, ncomp_all_, comp_ptr_);
    hipLaunchKernelGGL(k_h_comp_lists, dim3(blocks(max_banks, per)), dim3(64), 0, st, lv, L, bank, compm_, parent_,
                       comp_ptr_, ncomp_all_, cmask_, cparent_);
    const hipError_t launch = hipGetLastError();
    hier_check(hipMemcpyAsync(h_lv_, lv_, 32 * sizeof(int), hipMemcpyDeviceToHost, st));
    hier_check(hipStreamSynchronize(st));
    hier_check(launch);
    levels_ = 0;
    if(h_lv_[16] & kBadPartition) throw std::invalid_argument("part_to_real and real_to_part disagree");
    if(h_lv_[16] & kBadList) throw std::invalid_argument("neighbor range past the end of the list");
    if(h_lv_[16] & kBadNeighbor) throw std::invalid_argument("neighbor id out of range");
    if(h_lv_[16] & kAsymmetric) throw std::invalid_argument("neighbor lists are not symmetric");
    levels_  = L;
    n_fem_   = n_fem;
    n_slots_ = n_slots;
    bank_    = bank;
    level_count_.assign(L + 1, 0);
    level_begin_.assign(L + 1, 0);
    for(int l = 0; l <= L; ++l)
    {
        level_count_[l] = h_lv_[l];
        level_begin_[l] = h_lv_[8 + l];
    }
    total_   = level_begin_[L];
    n_banks_ = total_ / bank;
    n_comp_  = h_lv_[17];
}

void DeviceHierarchy::download(int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix, int* going_next,
                               int* coarse_table, int* comp_ptr, uint32_t* comp_mask, int* comp_parent,
                               hipStream_t st) const
{
    if(levels_ <= 0) throw std::runtime_error("device hierarchy not built");
    if(level_count) std::copy(level_count_.begin(), level_count_.end(), level_count);
    if(level_begin) std::copy(level_begin_.begin(), level_begin_.end(), level_begin);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if(dst && bytes) hier_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    };
    d2h(fine_mask, fine_mask_, sizeof(uint32_t) * n_fem_);
    d2h(prefix, prefix_, sizeof(int) * (n_slots_ / bank_));
    d2h(going_next, going_next_, sizeof(int) * std::max(total_, n_fem_));
    d2h(coarse_table, table_, sizeof(int) * kMaxLevels * (size_t)n_fem_);
    d2h(comp_ptr, comp_ptr_, sizeof(int) * (n_banks_ + 1));
    d2h(comp_mask, cmask_, sizeof(uint32_t) * n_comp_);
    d2h(comp_parent, cparent_, sizeof(int) * n_comp_);
    hier_check(hipStreamSynchronize(st));
}

}  // namespace masx
