// Host-side structure analysis (see plan.hpp).
#include "plan.hpp"
#include "pair_code.hpp"
#include "phase_timer.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

namespace masx
{
namespace
{
inline int pad_to(int v, int b) { return (v + b - 1) / b * b; }
inline int packed_index(int bank, int r, int c) { return bank * r - r * (r + 1) / 2 + c; }
inline uint32_t lane_bit(int l) { return 1u << l; }
}  // namespace

// Stable parallel counting sort: item k (k < n) goes to up to two buckets
// (emit(k, sink) calls sink(bucket, slot) for each); buckets keep items in k
// order.  Returns the bucket pointers (nbuckets + 1); place(k, slot, pos)
// receives every item's final position.
template <class Emit, class Place>
std::vector<long long> bucket_scatter(long long n, int nbuckets, Emit emit, Place place)
{
    const int T = std::max(1, std::min(omp_get_max_threads(), (int)((n + 65535) / 65536)));
    // per-thread counters, zeroed (first touched) by the thread that uses them
    std::unique_ptr<long long[]> cnt(new long long[(size_t)T * (nbuckets + 1)]);
    const long long chunk = (n + T - 1) / T;
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for(int t = 0; t < T; ++t)
    {
        long long* c = cnt.get() + (size_t)t * (nbuckets + 1);
        std::fill(c, c + nbuckets + 1, 0LL);
        for(long long k = t * chunk; k < std::min(n, (t + 1) * chunk); ++k) emit(k, [&](int b, int) { ++c[b]; });
    }
    // bucket totals (parallel over buckets), their scan, then each thread's
    // start inside every bucket
    std::vector<long long> ptr(nbuckets + 1, 0);
#pragma omp parallel for schedule(static)
    for(int b = 0; b < nbuckets; ++b)
    {
        long long acc = 0;
        for(int t = 0; t < T; ++t) acc += cnt[(size_t)t * (nbuckets + 1) + b];
        ptr[b + 1] = acc;
    }
    for(int b = 0; b < nbuckets; ++b) ptr[b + 1] += ptr[b];
#pragma omp parallel for schedule(static)
    for(int b = 0; b < nbuckets; ++b)
    {
        long long acc = ptr[b];
        for(int t = 0; t < T; ++t)
        {
            const long long v = cnt[(size_t)t * (nbuckets + 1) + b];
            cnt[(size_t)t * (nbuckets + 1) + b] = acc;
            acc += v;
        }
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for(int t = 0; t < T; ++t)
    {
        long long* c = cnt.get() + (size_t)t * (nbuckets + 1);
        for(long long k = t * chunk; k < std::min(n, (t + 1) * chunk); ++k)
            emit(k, [&](int b, int slot) { place(k, slot, c[b]++); });
    }
    return ptr;
}

int reference_level_count(int vert_num, int part_map_size, int bank)
{
    // TraditionalMAS32Preconditioner.cu:2142-2161 with the max of :2898
    const int max_nodes = std::max(vert_num, part_map_size);
    int       levels    = 1;
    int       sz        = pad_to(max_nodes, bank);
    while(sz > bank)
    {
        sz = pad_to(sz / bank, bank);
        ++levels;
    }
    return std::min(levels + 1, 6);
}

int decode_pair(const int* p, int v[4]) { return decode_pair_code(p, v); }

void morton_partition(int n, const double* pos, int bank, std::vector<int>& p2r, std::vector<int>& r2p)
{
    // gl_main.cu:1176-1243: 10 bits per axis over the bounding box, x | y<<1 | z<<2,
    // ties broken by index, banks padded with -1 at the end
    auto spread = [](uint32_t v) {
        uint64_t x = v & 0x3ffu, out = 0;
        for(int b = 0; b < 10; ++b) out |= ((x >> b) & 1u) << (3 * b);
        return (uint32_t)out;
    };
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for(int d = 0; d < 3 && n > 0; ++d) lo[d] = hi[d] = pos[d];
    for(int v = 1; v < n; ++v)
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = std::min(lo[d], pos[3 * v + d]);
            hi[d] = std::max(hi[d], pos[3 * v + d]);
        }
    std::vector<std::pair<uint32_t, int>> key(n);
    for(int v = 0; v < n; ++v)
    {
        uint32_t code = 0;
        for(int d = 0; d < 3; ++d)
        {
            // separate statements: no fused multiply-add, so the device
            // partition (morton.hip) forms the same codes bit for bit
            const double e = hi[d] - lo[d];
            const double t = e > 0.0 ? std::clamp((pos[3 * v + d] - lo[d]) / e, 0.0, 1.0) : 0.0;
            const double s = t * 1023.0;
            const double q = s + 0.5;
            code |= spread((uint32_t)q) << d;
        }
        key[v] = {code, v};
    }
    std::sort(key.begin(), key.end());
    p2r.assign(pad_to(n, bank), -1);
    r2p.assign(n, 0);
    for(int s = 0; s < n; ++s)
    {
        p2r[s]               = key[s].second;
        r2p[key[s].second] = s;
    }
}

// Components inside one bank by label propagation over lane bit sets.
// adj[l] holds lane l's own bit plus its neighbours' bits.
static void bank_components(const uint32_t* adj, uint32_t valid, uint32_t* comp)
{
    uint32_t done = 0;
    for(int l = 0; l < 32; ++l)
    {
        if(!(valid & lane_bit(l)) || (done & lane_bit(l))) continue;
        uint32_t set = lane_bit(l), frontier = lane_bit(l);
        while(frontier)
        {
            uint32_t next = 0;
            for(uint32_t f = frontier; f; f &= f - 1) next |= adj[__builtin_ctz(f)];
            frontier = next & ~set;
            set |= next;
        }
        for(uint32_t f = set; f; f &= f - 1) comp[__builtin_ctz(f)] = set;
        done |= set;
    }
}

void build_hierarchy(Hierarchy& h, int n_fem, int n_slots, int bank, int levels, const uint32_t* nstart,
                     const uint32_t* nnum, const uint32_t* nlist, const int* p2r, const int* r2p,
                     const int* pairs, int n_pairs, int node_offset)
{
    if(bank < 2 || bank > 32) throw std::invalid_argument("bank size must be in [2, 32]");
    if(n_slots % bank || n_slots < n_fem) throw std::invalid_argument("partition size must be a bank multiple >= vertex count");
    h.bank    = bank;
    h.n_fem   = n_fem;
    h.n_slots = n_slots;
    h.levels  = levels > 0 ? std::min(levels, 6) : reference_level_count(n_fem, n_slots, bank);
    const int L = h.levels;
    h.level_count.assign(L + 1, 0);
    h.level_begin.assign(L + 1, 0);
    h.level_count[0] = n_slots;
    h.fine_mask.assign(n_fem, 0);
    h.prefix.assign(n_slots / bank, 0);
    h.coarse_table.assign((size_t)n_fem * 6, 0);

    // undirected links between fine vertices: mesh edges + every pair of
    // distinct in-range vertices of each contact
    std::vector<std::pair<int, int>> links;
    links.reserve(nstart ? (size_t)n_fem * 8 : 0);
    for(int v = 0; v < n_fem && nstart; ++v)
        for(uint32_t i = 0; i < nnum[v]; ++i)
        {
            const int u = (int)nlist[nstart[v] + i];
            if(u < 0 || u >= n_fem) throw std::invalid_argument("neighbor id out of range");
            if(u > v) links.push_back({v, u});
        }
    for(int k = 0; k < n_pairs; ++k)
    {
        int       vv[4];
        const int cnt = decode_pair(pairs + 4 * k, vv);
        for(int i = 0; i < cnt; ++i) vv[i] -= node_offset;
        for(int i = 0; i < cnt; ++i)
            for(int j = i + 1; j < cnt; ++j)
                if(vv[i] >= 0 && vv[j] >= 0 && vv[i] < n_fem && vv[j] < n_fem && vv[i] != vv[j])
                    links.push_back({vv[i], vv[j]});
    }

    std::vector<int> id(r2p, r2p + n_fem);  // fine vertex -> node at the current level
    int count = n_slots, begin = 0;
    std::vector<uint32_t> adj, comp;
    std::vector<int>      newid;
    std::vector<std::vector<std::pair<uint32_t, int>>> bank_comps;  // per slot bank, all levels
    const int max_slots = n_slots * 2 + bank * (L + 2);
    bank_comps.reserve(max_slots / bank + 1);
    h.going_next.assign(max_slots, -1);
    for(int level = 0; level < L; ++level)
    {
        const int nbanks = (count + bank - 1) / bank;
        adj.assign((size_t)nbanks * bank, 0u);
        comp.assign((size_t)nbanks * bank, 0u);
        std::vector<uint32_t> valid(nbanks, 0u);
        for(int c = 0; c < count; ++c)
        {
            const bool ok = level > 0 || p2r[c] >= 0;
            if(!ok) continue;
            adj[c] |= lane_bit(c % bank);
            valid[c / bank] |= lane_bit(c % bank);
        }
        for(const auto& e : links)
        {
            const int a = id[e.first], b = id[e.second];
            if(a != b && a / bank == b / bank)
            {
                adj[a] |= lane_bit(b % bank);
                adj[b] |= lane_bit(a % bank);
            }
        }
        newid.assign((size_t)nbanks * bank, -1);
        int next = 0;
        const int next_begin = level == 0 ? pad_to(n_slots, bank) : begin + pad_to(count, bank);
        for(int b = 0; b < nbanks; ++b)
        {
            bank_components(&adj[(size_t)b * bank], valid[b], &comp[(size_t)b * bank]);
            int ncomp = 0;
            std::vector<std::pair<uint32_t, int>> comps;
            for(int l = 0; l < bank; ++l)
            {
                if(!(valid[b] & lane_bit(l))) continue;
                const uint32_t m = comp[(size_t)b * bank + l];
                if(__builtin_ctz(m) == l)  // lowest lane leads the component
                {
                    comps.push_back({m, next_begin + next});
                    for(uint32_t f = m; f; f &= f - 1) newid[(size_t)b * bank + __builtin_ctz(f)] = next;
                    ++next;
                    ++ncomp;
                }
            }
            if(level == 0) h.prefix[b] = ncomp;
            if(level + 1 < L)
                bank_comps.push_back(std::move(comps));
            else
                bank_comps.push_back({});
        }
        if(level == 0)
            for(int v = 0; v < n_fem; ++v) h.fine_mask[v] = comp[r2p[v]];
        h.level_count[level + 1 <= L ? level + 1 : L] = level + 1 <= L ? next : 0;
        if(level + 1 <= L) h.level_begin[level + 1] = next_begin;
        if(level == 0)
        {
            for(int v = 0; v < n_fem; ++v) h.going_next[v] = newid[r2p[v]] + next_begin;
        }
        else
        {
            if(begin + count > (int)h.going_next.size()) h.going_next.resize(begin + count + bank, -1);
            for(int c = 0; c < count; ++c) h.going_next[begin + c] = newid[c] + next_begin;
        }
        for(int v = 0; v < n_fem; ++v)
        {
            id[v] = newid[id[v]];
            if(level < L - 1) h.coarse_table[6 * (size_t)v + level] = 0;  // filled below
        }
        // drop links that became internal to one node
        links.erase(std::remove_if(links.begin(), links.end(),
                                   [&](const std::pair<int, int>& e) { return id[e.first] == id[e.second]; }),
                    links.end());
        begin = next_begin;
        count = next;
    }
    h.total = h.level_begin[L];
    h.going_next.resize(std::max(h.total, n_fem), -1);
    for(int v = 0; v < n_fem; ++v)
    {
        int cur = v;
        for(int l = 0; l < L - 1; ++l)
        {
            cur                             = h.going_next[cur];
            h.coarse_table[6 * (size_t)v + l] = cur;
        }
    }
    // component lists for the banks of levels 0..L-2 (slots [0, level_begin[L-1]))
    const int total_banks = h.total / bank;
    h.comp_ptr.assign(total_banks + 1, 0);
    h.comp_mask.clear();
    h.comp_parent.clear();
    // bank_comps is ordered by level then bank; map to slot banks
    int idx = 0;
    std::vector<int> bank_of_entry;
    for(int level = 0; level < L; ++level)
    {
        const int b0 = (level == 0 ? 0 : h.level_begin[level]) / bank;
        const int nb = (level == 0 ? n_slots : pad_to(h.level_count[level], bank)) / bank;
        for(int b = 0; b < nb; ++b, ++idx)
        {
            h.comp_ptr[b0 + b + 1] = (int)bank_comps[idx].size();
        }
    }
    for(int b = 0; b < total_banks; ++b) h.comp_ptr[b + 1] += h.comp_ptr[b];
    h.comp_mask.resize(h.comp_ptr[total_banks]);
    h.comp_parent.resize(h.comp_ptr[total_banks]);
    idx = 0;
    for(int level = 0; level < L; ++level)
    {
        const int b0 = (level == 0 ? 0 : h.level_begin[level]) / bank;
        const int nb = (level == 0 ? n_slots : pad_to(h.level_count[level], bank)) / bank;
        for(int b = 0; b < nb; ++b, ++idx)
            for(size_t k = 0; k < bank_comps[idx].size(); ++k)
            {
                h.comp_mask[h.comp_ptr[b0 + b] + k]   = bank_comps[idx][k].first;
                h.comp_parent[h.comp_ptr[b0 + b] + k] = bank_comps[idx][k].second;
            }
    }
}

// ----------------------------------------------------------------- plan ----


void build_plan(Plan& P, std::vector<SystemDesc>& sys, bool structure_on_device)
{
    PhaseTimer tm("plan");
    const int S = (int)sys.size();
    if(S == 0) throw std::invalid_argument("no systems");
    for(int s = 0; structure_on_device && s < S; ++s)
        if(!sys[s].use_mas || sys[s].n_abd != 0 || !sys[s].hier_ready)
            throw std::logic_error("the device structure build takes MAS systems without bodies");
    const int bank = sys[0].bank;
    for(auto& s : sys)
        if(s.use_mas && s.bank != bank) throw std::invalid_argument("all systems of a batch share one bank size");
    P     = Plan{};
    P.bank = bank;
    P.n_sys = S;
    P.structure_on_device = structure_on_device;
    P.row_off.assign(S + 1, 0);
    P.ext_off.assign(S + 1, 0);
    P.fem_row0.assign(S, 0);
    P.abd_row0.assign(S, 0);
    P.src_off.assign(S + 1, 0);
    P.coarse_base.assign(S + 1, 0);
    // ---- rows
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        if(d.fem_offset != 4 * d.n_abd || d.block_rows != d.fem_offset + d.n_fem)
            throw std::invalid_argument("block rows must be [4 per affine body][FEM vertices]");
        if(d.use_mas)
        {
            if((int)d.part_to_real.size() % bank || (int)d.real_to_part.size() != d.n_fem)
                throw std::invalid_argument("partition arrays have the wrong size");
        }
        const int abd_pad = pad_to(4 * d.n_abd, bank);
        const int fem_rows = d.use_mas ? (int)d.part_to_real.size() : pad_to(d.n_fem, bank);
        P.abd_row0[s]     = P.row_off[s];
        P.fem_row0[s]     = P.row_off[s] + abd_pad;
        P.row_off[s + 1]  = P.row_off[s] + abd_pad + fem_rows;
        P.ext_off[s + 1]  = P.ext_off[s] + d.block_rows;
        P.src_off[s + 1]  = P.src_off[s] + d.nb;
    }
    P.n_rows   = P.row_off[S];
    P.n_ext    = P.ext_off[S];
    P.nb_total = P.src_off[S];
    if(P.nb_total >= (1ll << 31) - 1) throw std::invalid_argument("too many blocks for 32-bit indices");
    P.ext_to_int.assign(P.n_ext, -1);
    P.row_sys.assign(P.n_rows, 0);
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        for(int r = 0; r < 4 * d.n_abd; ++r) P.ext_to_int[P.ext_off[s] + r] = P.abd_row0[s] + r;
        for(int v = 0; v < d.n_fem; ++v)
            P.ext_to_int[P.ext_off[s] + d.fem_offset + v] = P.fem_row0[s] + (d.use_mas ? d.real_to_part[v] : v);
        for(int r = P.row_off[s]; r < P.row_off[s + 1]; ++r) P.row_sys[r] = s;
    }
    tm.mark("rows");
    // ---- full block CSR (both triangles)
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        if(d.rows_on_device)
        {
            // checked on the device by the block CSR build (out-of-range indices are flagged there)
            if(!structure_on_device) throw std::logic_error("device-held structure needs the device analysis");
            continue;
        }
        if((long long)d.rows.size() != d.nb || (long long)d.cols.size() != d.nb)
            throw std::invalid_argument("row/col arrays do not match the block count");
        int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
        for(long long k = 0; k < d.nb; ++k)
            if(d.rows[k] < 0 || d.rows[k] >= d.block_rows || d.cols[k] < 0 || d.cols[k] >= d.block_rows) bad |= 1;
        if(bad) throw std::invalid_argument("block index out of range");
    }
    if(!structure_on_device)
    {
        // items: the upper blocks of all systems; block k of system s lands in
        // row R as (C, g) and, off the diagonal, in row C as (R, g | T)
        std::vector<long long> item_base(S + 1, 0);
        for(int s = 0; s < S; ++s) item_base[s + 1] = item_base[s] + sys[s].nb;
        // internal (row, col) of every upper block, once (items are the
        // global upper block indices: item k is block k of the concatenation)
        std::unique_ptr<int[]> iR(new int[item_base[S] + 1]), iC(new int[item_base[S] + 1]);
        for(int s = 0; s < S; ++s)
        {
            const auto& d = sys[s];
            const int* e2i = P.ext_to_int.data() + P.ext_off[s];
            const long long b = item_base[s];
#pragma omp parallel for schedule(static)
            for(long long kk = 0; kk < d.nb; ++kk)
            {
                iR[b + kk] = e2i[d.rows[kk]];
                iC[b + kk] = d.rows[kk] == d.cols[kk] ? -1 : e2i[d.cols[kk]];  // -1: diagonal
            }
        }
        auto emit = [&](long long k, auto&& sink) {
            sink(iR[k], 0);
            if(iC[k] >= 0) sink(iC[k], 1);
        };
        P.col.assign(2 * item_base[S], 0);
        P.entry_src.assign(2 * item_base[S], 0);
        auto ptr = bucket_scatter(item_base[S], P.n_rows, emit, [&](long long k, int slot, long long pos) {
            const int g = (int)k;  // == src_off[s] + local index
            P.col[pos] = slot == 0 ? (iC[k] >= 0 ? iC[k] : iR[k]) : iR[k];
            P.entry_src[pos] = slot == 0 ? g : (g | (int)0x80000000u);
        });
        P.row_ptr.assign(P.n_rows + 1, 0);
        for(int r = 0; r <= P.n_rows; ++r) P.row_ptr[r] = (int)ptr[r];
        P.nnz = ptr[P.n_rows];
        P.col.resize(P.nnz);
        P.entry_src.resize(P.nnz);
        // sort each row by column (stable) for gather locality
#pragma omp parallel for schedule(dynamic, 512)
        for(int r = 0; r < P.n_rows; ++r)
        {
            thread_local std::vector<std::pair<int, int>> tmp;
            const int b = P.row_ptr[r], e = P.row_ptr[r + 1];
            tmp.resize(e - b);
            for(int i = b; i < e; ++i) tmp[i - b] = {P.col[i], P.entry_src[i]};
            if(tmp.size() <= 64)  // rows are short: stable insertion sort, no allocation
            {
                for(size_t i = 1; i < tmp.size(); ++i)
                {
                    const auto v = tmp[i];
                    size_t j = i;
                    for(; j > 0 && tmp[j - 1].first > v.first; --j) tmp[j] = tmp[j - 1];
                    tmp[j] = v;
                }
            }
            else
                std::stable_sort(tmp.begin(), tmp.end(), [](auto& x, auto& y) { return x.first < y.first; });
            for(int i = b; i < e; ++i)
            {
                P.col[i]       = tmp[i - b].first;
                P.entry_src[i] = tmp[i - b].second;
            }
        }
    }
    tm.mark("block CSR");
    // ---- global 3x3 block-Jacobi and ABD bodies
    P.diag_src.assign(structure_on_device ? 0 : P.n_rows, -1);
    P.row_needs_diag.assign(P.n_rows, 0);
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        for(long long k = 0; !structure_on_device && k < d.nb; ++k)
            if(d.rows[k] == d.cols[k]) P.diag_src[P.ext_to_int[P.ext_off[s] + d.rows[k]]] = (int)(P.src_off[s] + k);
        for(int r = P.row_off[s]; r < P.row_off[s + 1]; ++r)
        {
            const bool abd = r >= P.abd_row0[s] && r < P.abd_row0[s] + 4 * d.n_abd;
            const bool fem = r >= P.fem_row0[s];
            P.row_needs_diag[r] = d.global_diag && !abd && !(fem && d.use_mas) ? 1 : 0;
        }
        if(d.n_abd > 0)
        {
            std::unordered_map<long long, int> where;
            for(long long k = 0; k < d.nb; ++k)
                if(d.rows[k] < 4 * d.n_abd && d.rows[k] / 4 == d.cols[k] / 4)
                    where[(long long)d.rows[k] * d.block_rows + d.cols[k]] = (int)(P.src_off[s] + k);
            for(int b = 0; b < d.n_abd; ++b)
            {
                P.body_row.push_back(P.abd_row0[s] + 4 * b);
                P.body_sys.push_back(s);
                for(int q = 0; q < 4; ++q)
                    for(int t = 0; t < 4; ++t)
                    {
                        int src = -1;
                        auto it = where.find((long long)(4 * b + q) * d.block_rows + 4 * b + t);
                        if(it != where.end())
                            src = it->second;
                        else
                        {
                            it = where.find((long long)(4 * b + t) * d.block_rows + 4 * b + q);
                            if(it != where.end()) src = it->second | (int)0x80000000u;
                        }
                        P.body_src.push_back(src);
                    }
            }
        }
    }
    P.n_bodies = (int)P.body_row.size();

    tm.mark("bodies");
    // ---- MAS: hierarchy per system, banks, components, Galerkin maps
    int n_fine_banks = 0, n_coarse = 0;
    std::vector<int> fine_bank_base(S, 0);
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        fine_bank_base[s] = n_fine_banks;
        P.coarse_base[s]  = n_coarse;
        if(!d.use_mas) continue;
        if(!d.hier_ready)
            build_hierarchy(d.hier, d.n_fem, (int)d.part_to_real.size(), bank, d.levels, d.nbr_start.data(),
                            d.nbr_num.data(), d.nbr_list.data(), d.part_to_real.data(), d.real_to_part.data(),
                            d.pairs.data(), (int)d.pairs.size() / 4, d.node_offset);
        n_fine_banks += d.hier.n_slots / bank;
        n_coarse += d.hier.total - d.hier.n_slots;
    }
    P.coarse_base[S] = n_coarse;
    P.n_fine_banks   = n_fine_banks;
    P.n_coarse       = n_coarse;
    P.n_banks        = n_fine_banks + n_coarse / bank;
    P.bank_vec.assign(P.n_banks, 0);
    P.bank_sys.assign(P.n_banks, 0);
    P.coarse_parent.assign(n_coarse, -1);
    P.comp_ptr.assign(P.n_banks + 1, 0);
    auto gbank_of_slot = [&](int s, int slot) {
        const auto& h = sys[s].hier;
        return slot < h.n_slots ? fine_bank_base[s] + slot / bank
                                : n_fine_banks + (P.coarse_base[s] + slot - h.n_slots) / bank;
    };
    auto gcoarse = [&](int s, int slot) { return P.coarse_base[s] + slot - sys[s].hier.n_slots; };
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        if(!d.use_mas) continue;
        const auto& h = d.hier;
        for(int b = 0; b < h.total / bank; ++b)
        {
            const int gb = gbank_of_slot(s, b * bank);
            P.bank_sys[gb] = s;
            P.bank_vec[gb] = b * bank < h.n_slots ? P.fem_row0[s] + b * bank : gcoarse(s, b * bank);
            P.comp_ptr[gb + 1] = h.comp_ptr[b + 1] - h.comp_ptr[b];
        }
        for(int slot = h.n_slots; slot < h.total; ++slot)
        {
            // parent exists below the top level
            const int lvl_top_begin = h.level_begin[h.levels - 1];
            if(slot < lvl_top_begin && h.going_next[slot] >= 0) P.coarse_parent[gcoarse(s, slot)] = gcoarse(s, h.going_next[slot]);
        }
    }
    for(int b = 0; b < P.n_banks; ++b) P.comp_ptr[b + 1] += P.comp_ptr[b];
    // level of each bank, level-1 descendants of deeper nodes, ancestors of
    // level-1 nodes, level-1 parent of each fine row
    P.bank_kind.assign(P.n_banks, 0);
    P.anc.assign(8 * (size_t)std::max(n_coarse, 1), -1);
    P.row_l1.assign(P.n_rows, -1);
    // systems are independent here (disjoint banks, coarse nodes and rows)
#pragma omp parallel for schedule(dynamic, 1)
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        if(!d.use_mas) continue;
        const auto& h = d.hier;
        for(int l = 1; l < h.levels; ++l)
        {
            const int b0 = gbank_of_slot(s, h.level_begin[l]);
            const int nb = pad_to(h.level_count[l], bank) / bank;
            for(int b = b0; b < b0 + nb; ++b) P.bank_kind[b] = l == 1 ? 1 : 2;
        }
        for(int c = 0; h.levels > 1 && c < h.level_count[1]; ++c)
        {
            const int slot1 = h.level_begin[1] + c;
            const int g1    = gcoarse(s, slot1);
            int       cur   = slot1, k = 0;
            const int top_begin = h.level_begin[h.levels - 1];
            while(cur >= 0 && k < 8)
            {
                P.anc[8 * (size_t)g1 + k++] = gcoarse(s, cur);
                if(cur >= top_begin) break;
                cur = h.going_next[cur];
            }
        }
        for(int slot = 0; slot < h.n_slots; ++slot)
        {
            const int v = d.part_to_real[slot];
            if(v < 0 || h.levels < 2) continue;
            const int g1 = gcoarse(s, h.going_next[v]);
            P.row_l1[P.fem_row0[s] + slot] = g1;
        }
    }
    // level-1 descendants of every deeper node, in level-1 node order (a
    // counting pass over the ancestor lists, then the fill)
    std::vector<std::pair<int, int>> l1_ranges;  // per system: its level-1 nodes [first, end)
    for(int s = 0; s < S; ++s)
    {
        const auto& h = sys[s].hier;
        if(sys[s].use_mas && h.levels > 1)
            l1_ranges.push_back({gcoarse(s, h.level_begin[1]), gcoarse(s, h.level_begin[1]) + h.level_count[1]});
    }
    P.desc_ptr.assign(n_coarse + 1, 0);
    for(const auto& rg : l1_ranges)
        for(int g1 = rg.first; g1 < rg.second; ++g1)
            for(int k = 1; k < 8 && P.anc[8 * (size_t)g1 + k] >= 0; ++k) ++P.desc_ptr[P.anc[8 * (size_t)g1 + k] + 1];
    for(int c = 0; c < n_coarse; ++c) P.desc_ptr[c + 1] += P.desc_ptr[c];
    P.desc_list.resize(P.desc_ptr[n_coarse]);
    {
        std::vector<int> at(P.desc_ptr.begin(), P.desc_ptr.end() - 1);
        for(const auto& rg : l1_ranges)
            for(int g1 = rg.first; g1 < rg.second; ++g1)
                for(int k = 1; k < 8 && P.anc[8 * (size_t)g1 + k] >= 0; ++k)
                    P.desc_list[at[P.anc[8 * (size_t)g1 + k]]++] = g1;
    }
    // self-check: every level-1 node lists itself first and appears exactly
    // once among the descendants of each deeper level
    for(int s = 0; s < S; ++s)
    {
        const auto& d = sys[s];
        if(!d.use_mas || d.hier.levels < 2) continue;
        const auto& h = d.hier;
        for(int c = 0; c < h.level_count[1]; ++c)
        {
            const int g1 = gcoarse(s, h.level_begin[1] + c);
            if(P.anc[8 * (size_t)g1] != g1) throw std::logic_error("ancestor list of a level-1 node is wrong");
        }
        for(int l = 2; l < h.levels; ++l)
        {
            long long n = 0;
            for(int c = 0; c < h.level_count[l]; ++c)
            {
                const int g = gcoarse(s, h.level_begin[l] + c);
                n += P.desc_ptr[g + 1] - P.desc_ptr[g];
            }
            if(n != h.level_count[1]) throw std::logic_error("descendant lists do not cover level 1");
        }
    }
    P.comp_mask.resize(P.comp_ptr[P.n_banks]);
    P.comp_parent.resize(P.comp_ptr[P.n_banks]);
    P.sys_level_ptr.assign(S + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for(int s = 0; s < S; ++s)
    {
        const auto& d = sys[s];
        if(!d.use_mas) continue;
        const auto& h = d.hier;
        for(int b = 0; b < h.total / bank; ++b)
        {
            const int gb = gbank_of_slot(s, b * bank);
            for(int k = h.comp_ptr[b]; k < h.comp_ptr[b + 1]; ++k)
            {
                P.comp_mask[P.comp_ptr[gb] + k - h.comp_ptr[b]]   = h.comp_mask[k];
                P.comp_parent[P.comp_ptr[gb] + k - h.comp_ptr[b]] = gcoarse(s, h.comp_parent[k]);
            }
        }
    }
    for(int s = 0; s < S; ++s)
    {
        auto& d = sys[s];
        if(d.use_mas)
        {
            const auto& h = d.hier;
            // coarse levels 1..L-1 as global coarse bank ranges
            for(int l = 1; l < h.levels; ++l)
            {
                const int b0 = gbank_of_slot(s, h.level_begin[l]);
                const int nb = pad_to(h.level_count[l], bank) / bank;
                P.sys_level_banks.push_back(b0);
                P.sys_level_banks.push_back(b0 + nb);
            }
        }
        P.sys_level_ptr[s + 1] = (int)P.sys_level_banks.size() / 2;
    }

    tm.mark("banks, ancestors, components");
    // Residuals of levels >= 2, pre-reduced per update group: the update
    // kernel treats 256 rows (256/bank fine banks) of one system per workgroup
    // and sums the level-1 component residuals of those banks.  Each group
    // also emits one partial per distinct deeper ancestor of its components
    // (slots), so a level >= 2 node gathers its groups' partials instead of all
    // its level-1 descendants.
    {
        const int per_group = 256 / bank;
        P.bank_ugroup.assign(std::max(n_fine_banks, 1), -1);
        P.comp_dslot.assign(4 * (size_t)std::max(P.comp_ptr[n_fine_banks], 1), -1);
        P.ugroup.clear();
        std::vector<std::vector<int>> deep(n_coarse);
        int slots = 0;
        std::vector<long long> covered(n_coarse, 0);
        for(int s = 0; s < S; ++s)
        {
            const auto& d = sys[s];
            if(!d.use_mas) continue;
            const int fb0 = fine_bank_base[s], nfb = d.hier.n_slots / bank;
            for(int q = 0; q < nfb; q += per_group)
            {
                const int g = (int)P.ugroup.size() / 4;
                const int b0 = fb0 + q, b1 = fb0 + std::min(nfb, q + per_group);
                std::vector<int> nodes;  // distinct deep ancestors, first appearance order
                for(int b = b0; b < b1; ++b)
                {
                    P.bank_ugroup[b] = g;
                    for(int c = P.comp_ptr[b]; c < P.comp_ptr[b + 1]; ++c)
                    {
                        const int g1 = P.comp_parent[c];
                        for(int k = 1; k < 8; ++k)
                        {
                            const int a = P.anc[8 * (size_t)g1 + k];
                            if(a < 0) break;
                            if(k > 4) throw std::invalid_argument("more than 6 MAS levels");
                            int slot = -1;
                            for(size_t t = 0; t < nodes.size(); ++t)
                                if(nodes[t] == a) slot = (int)t;
                            if(slot < 0)
                            {
                                slot = (int)nodes.size();
                                nodes.push_back(a);
                            }
                            P.comp_dslot[4 * (size_t)c + k - 1] = slot;
                            ++covered[a];
                        }
                    }
                }
                if(P.comp_ptr[b1] - P.comp_ptr[b0] > 256) throw std::logic_error("update group with > 256 components");
                P.ugroup.insert(P.ugroup.end(), {slots, (int)nodes.size(), P.comp_ptr[b0], P.comp_ptr[b1]});
                for(size_t t = 0; t < nodes.size(); ++t) deep[nodes[t]].push_back(slots + (int)t);
                slots += (int)nodes.size();
            }
        }
        P.n_deep_slots = slots;
        P.deep_ptr.assign(n_coarse + 1, 0);
        for(int c = 0; c < n_coarse; ++c) P.deep_ptr[c + 1] = P.deep_ptr[c] + (int)deep[c].size();
        P.deep_list.clear();
        for(int c = 0; c < n_coarse; ++c) P.deep_list.insert(P.deep_list.end(), deep[c].begin(), deep[c].end());
        // self-check: every level-1 descendant of a deep node is counted once
        for(int c = 0; c < n_coarse; ++c)
            if(covered[c] != P.desc_ptr[c + 1] - P.desc_ptr[c])
                throw std::logic_error("update-group partials do not cover the level-1 descendants");
    }

    tm.mark("update groups");
    if(structure_on_device) return;
    // Galerkin coarse operators A_l = P_l^T A P_l (all entries, not only in-bank),
    // built level by level from the previous level; group g holds level g+1 of
    // every system.
    int max_levels = 0;
    for(auto& d : sys)
        if(d.use_mas) max_levels = std::max(max_levels, d.hier.levels);
    struct Ent { long long key; int srcmode; int pos; };  // srcmode = source << 2 | mode
    // level-(l) entries per system: (I,J) slots, global E index
    std::vector<std::vector<std::pair<int, int>>> prev_ij(S);   // (I,J) of the previous level's entries
    std::vector<std::vector<int>>                 prev_idx(S);  // their global E index
    P.gal_level_ptr.assign(1, 0);
    P.gal_ptr.assign(1, 0);
    P.gal_src.clear();
    std::vector<std::pair<int, int>> E_ij;     // all E entries (slots, system-local)
    std::vector<int>                 E_sys;
    for(int lvl = 1; lvl < std::max(max_levels, 1); ++lvl)
    {
        for(int s = 0; s < S; ++s)
        {
            auto& d = sys[s];
            if(!d.use_mas || lvl >= d.hier.levels) continue;
            const auto& h = d.hier;
            // entries (I, J, source, mode) in canonical (min, max) order: as
            // is, transposed, or B + B^T when both ends merge into one node;
            // bucketed by I (stable), then each bucket sorted by J (stable),
            // which is the (I, J) order with sources in input order
            const int lb = h.level_begin[lvl], lc = h.level_count[lvl];
            const long long n_src = lvl == 1 ? d.nb : (long long)prev_ij[s].size();
            auto canon = [&](long long k, int& I, int& J, int& src, int& mode) -> bool {
                int a, b;
                bool diag;
                if(lvl == 1)
                {
                    const int r = d.rows[k] - d.fem_offset, c = d.cols[k] - d.fem_offset;
                    if(r < 0 || c < 0) return false;
                    a = h.going_next[r];
                    b = h.going_next[c];
                    diag = r == c;
                    src = (int)(P.src_off[s] + k);
                }
                else
                {
                    a = h.going_next[prev_ij[s][k].first];
                    b = h.going_next[prev_ij[s][k].second];
                    diag = prev_ij[s][k].first == prev_ij[s][k].second;
                    src = prev_idx[s][k];
                }
                if(a < b) { I = a; J = b; mode = 0; }
                else if(a > b) { I = b; J = a; mode = 1; }
                else { I = a; J = a; mode = diag ? 0 : 2; }
                return true;
            };
            std::unique_ptr<Ent[]> ents(new Ent[n_src > 0 ? n_src : 1]);  // every slot used is written
            const auto bp = bucket_scatter(
                n_src, lc,
                [&](long long k, auto&& sink) {
                    int I, J, src, mode;
                    if(canon(k, I, J, src, mode)) sink(I - lb, 0);
                },
                [&](long long k, int, long long pos) {
                    int I, J, src, mode;
                    canon(k, I, J, src, mode);
                    ents[pos] = {(long long)J, (src << 2) | mode, 0};
                });
            const long long n_ents = bp[lc];
            std::vector<int> bptr(lc + 1);
            for(int i = 0; i <= lc; ++i) bptr[i] = (int)bp[i];
            std::vector<int> n_uniq(lc + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
            for(int i = 0; i < lc; ++i)
            {
                Ent* b = ents.get() + bptr[i];
                const int n = bptr[i + 1] - bptr[i];
                if(n <= 32)
                    for(int x = 1; x < n; ++x)
                    {
                        const Ent v = b[x];
                        int y = x;
                        for(; y > 0 && b[y - 1].key > v.key; --y) b[y] = b[y - 1];
                        b[y] = v;
                    }
                else
                {
                    // (key, input position): the stable order, without the
                    // temporary buffer a stable sort allocates per bucket
                    for(int x = 0; x < n; ++x) b[x].pos = x;
                    std::sort(b, b + n, [](const Ent& u, const Ent& v) {
                        return u.key != v.key ? u.key < v.key : u.pos < v.pos;
                    });
                }
                int u = 0;
                for(int x = 0; x < n; ++x)
                    if(x == 0 || b[x].key != b[x - 1].key) ++u;
                n_uniq[i + 1] = u;
            }
            for(int i = 0; i < lc; ++i) n_uniq[i + 1] += n_uniq[i];
            const int e_base = (int)E_ij.size(), g_base = (int)P.gal_src.size();
            const int n_new = n_uniq[lc];
            E_ij.resize(e_base + n_new);
            E_sys.resize(e_base + n_new, s);
            prev_ij[s].assign(n_new, {0, 0});
            prev_idx[s].assign(n_new, 0);
            P.gal_src.resize(g_base + n_ents);
            const size_t gp_base = P.gal_ptr.size();
            P.gal_ptr.resize(gp_base + n_new);
#pragma omp parallel for schedule(dynamic, 64)
            for(int i = 0; i < lc; ++i)
            {
                const Ent* b = ents.get() + bptr[i];
                const int n = bptr[i + 1] - bptr[i];
                int q = n_uniq[i];
                for(int x = 0; x < n; ++x)
                {
                    P.gal_src[g_base + bptr[i] + x] = b[x].srcmode;
                    if(x + 1 == n || b[x + 1].key != b[x].key)
                    {
                        const int I = lb + i, J = (int)b[x].key;
                        E_ij[e_base + q] = {I, J};
                        prev_ij[s][q] = {I, J};
                        prev_idx[s][q] = e_base + q;
                        P.gal_ptr[gp_base + q] = g_base + bptr[i] + x + 1;
                        ++q;
                    }
                }
            }
        }
        P.gal_level_ptr.push_back((int)E_ij.size());
        P.gal_src_is_fine.push_back(lvl == 1 ? 1 : 0);
    }
    P.n_gal = (long long)E_ij.size();
    if((long long)P.gal_src.size() >= (1ll << 29)) throw std::invalid_argument("Galerkin map too large");
    for(int e : P.gal_level_ptr) P.gal_level_src.push_back(P.gal_ptr[e]);

    tm.mark("Galerkin maps");
    // per-bank entries for the dense local assembly: a fine bank's upper
    // blocks (from its rows of the full CSR, as-is entries with both ends in
    // the bank), then the coarse entries with both ends in one coarse bank.
    // Every entry has its own packed position, so the order within a bank
    // does not change the assembled matrix.
    P.bank_ent_ptr.assign(P.n_banks + 1, 0);
    std::vector<int> fine_row0_of_bank(n_fine_banks, 0);
    for(int s = 0; s < S; ++s)
    {
        if(!sys[s].use_mas) continue;
        for(int b = 0; b < sys[s].hier.n_slots / bank; ++b) fine_row0_of_bank[fine_bank_base[s] + b] = P.fem_row0[s] + b * bank;
    }
    auto fine_bank_entries = [&](int gb, auto&& out) {
        const int r0 = fine_row0_of_bank[gb];
        for(int lr = 0; lr < bank; ++lr)
            for(int e = P.row_ptr[r0 + lr]; e < P.row_ptr[r0 + lr + 1]; ++e)
            {
                const int lc_ = P.col[e] - r0;
                if(P.entry_src[e] < 0 || lc_ < 0 || lc_ >= bank) continue;  // transposed copy, or another bank
                const int pos = lr <= lc_ ? packed_index(bank, lr, lc_) : packed_index(bank, lc_, lr);
                out(pos | ((lr > lc_ ? 1 : 0) << 16), P.entry_src[e]);
            }
    };
#pragma omp parallel for schedule(dynamic, 64)
    for(int gb = 0; gb < n_fine_banks; ++gb)
    {
        int c = 0;
        fine_bank_entries(gb, [&](int, int) { ++c; });
        P.bank_ent_ptr[gb + 1] = c;
    }
    for(size_t e = 0; e < E_ij.size(); ++e)
    {
        const int I = E_ij[e].first, J = E_ij[e].second;
        if(I / bank == J / bank) ++P.bank_ent_ptr[gbank_of_slot(E_sys[e], I) + 1];
    }
    for(int b = 0; b < P.n_banks; ++b) P.bank_ent_ptr[b + 1] += P.bank_ent_ptr[b];
    P.bank_ent_pos.resize(P.bank_ent_ptr[P.n_banks]);
    P.bank_ent_src.resize(P.bank_ent_ptr[P.n_banks]);
#pragma omp parallel for schedule(dynamic, 64)
    for(int gb = 0; gb < n_fine_banks; ++gb)
    {
        int at = P.bank_ent_ptr[gb];
        fine_bank_entries(gb, [&](int code, int src) {
            P.bank_ent_pos[at] = code;
            P.bank_ent_src[at] = src;
            ++at;
        });
    }
    {
        std::vector<int> cur(P.bank_ent_ptr.begin() + n_fine_banks, P.bank_ent_ptr.end() - 1);
        for(size_t e = 0; e < E_ij.size(); ++e)
        {
            const int I = E_ij[e].first, J = E_ij[e].second;
            if(I / bank != J / bank) continue;
            const int at = cur[gbank_of_slot(E_sys[e], I) - n_fine_banks]++;
            P.bank_ent_pos[at] = packed_index(bank, I % bank, J % bank);
            P.bank_ent_src[at] = (int)e;
        }
    }
    tm.mark("bank entries");
}


void validate_plan(const Plan& P)
{
    auto fail = [](const char* what) { throw std::logic_error(std::string("plan check failed: ") + what); };
    const long long nb = P.nb_total;
    const int nblk = P.bank * (P.bank + 1) / 2;
    const bool host = !P.structure_on_device;  // else the arrays the device builds exist only there
    if(host && ((int)P.row_ptr.size() != P.n_rows + 1 || P.row_ptr[0] != 0 || P.row_ptr.back() != P.nnz))
        fail("row_ptr");
    for(int r = 0; host && r < P.n_rows; ++r)
        if(P.row_ptr[r + 1] < P.row_ptr[r]) fail("row_ptr order");
    int bad_e = 0;
    const long long nnz_host = host ? P.nnz : 0;
#pragma omp parallel for reduction(| : bad_e) schedule(static)
    for(long long e = 0; e < nnz_host; ++e)
    {
        if(P.col[e] < 0 || P.col[e] >= P.n_rows) bad_e |= 1;
        if((P.entry_src[e] & 0x7fffffff) >= nb) bad_e |= 2;
    }
    if(bad_e & 1) fail("col");
    if(bad_e & 2) fail("entry_src");
    for(int v : P.ext_to_int)
        if(v < 0 || v >= P.n_rows) fail("ext_to_int");
    if((int)P.bank_vec.size() != P.n_banks || (int)P.bank_kind.size() != P.n_banks) fail("bank arrays");
    for(int b = 0; b < P.n_banks; ++b)
    {
        const bool fine = b < P.n_fine_banks;
        if(fine != (P.bank_kind[b] == 0)) fail("bank_kind");
        if(P.bank_vec[b] < 0 || P.bank_vec[b] + P.bank > (fine ? P.n_rows : P.n_coarse)) fail("bank_vec");
        if(P.bank_sys[b] < 0 || P.bank_sys[b] >= P.n_sys) fail("bank_sys");
        if(P.comp_ptr[b + 1] < P.comp_ptr[b]) fail("comp_ptr");
        for(int k = P.comp_ptr[b]; k < P.comp_ptr[b + 1]; ++k)
            if(P.comp_parent[k] < 0 || P.comp_parent[k] >= P.n_coarse) fail("comp_parent");
        if(!host) continue;
        if(P.bank_ent_ptr[b + 1] < P.bank_ent_ptr[b]) fail("bank_ent_ptr");
        for(int k = P.bank_ent_ptr[b]; k < P.bank_ent_ptr[b + 1]; ++k)
        {
            if((P.bank_ent_pos[k] & 0xffff) >= nblk) fail("bank_ent_pos");
            if(P.bank_ent_src[k] < 0 || P.bank_ent_src[k] >= (fine ? nb : P.n_gal)) fail("bank_ent_src");
        }
    }
    if((int)P.deep_ptr.size() != P.n_coarse + 1 || P.deep_ptr.back() != (int)P.deep_list.size()) fail("deep_ptr");
    for(int c = 0; c < P.n_coarse; ++c)
        if(P.deep_ptr[c + 1] < P.deep_ptr[c]) fail("deep_ptr order");
    for(int v : P.deep_list)
        if(v < 0 || v >= P.n_deep_slots) fail("deep_list");
    for(size_t g = 0; g < P.ugroup.size() / 4; ++g)
    {
        const int* u = &P.ugroup[4 * g];
        if(u[0] < 0 || u[0] + u[1] > P.n_deep_slots || u[2] < 0 || u[3] < u[2] || u[3] - u[2] > 256 ||
           u[3] > P.comp_ptr[P.n_fine_banks])
            fail("ugroup");
        for(int c = u[2]; c < u[3]; ++c)
            for(int k = 0; k < 4; ++k)
                if(P.comp_dslot[4 * (size_t)c + k] < -1 || P.comp_dslot[4 * (size_t)c + k] >= u[1]) fail("comp_dslot");
    }
    for(int b = 0; b < P.n_fine_banks; ++b)
        if(P.bank_ugroup[b] < -1 || P.bank_ugroup[b] >= (int)P.ugroup.size() / 4) fail("bank_ugroup");
    if((int)P.desc_ptr.size() != P.n_coarse + 1) fail("desc_ptr");
    for(int c = 0; c < P.n_coarse; ++c)
        if(P.desc_ptr[c + 1] < P.desc_ptr[c]) fail("desc_ptr order");
    for(int v : P.desc_list)
        if(v < 0 || v >= P.n_coarse) fail("desc_list");
    for(int v : P.anc)
        if(v < -1 || v >= P.n_coarse) fail("anc");
    if(P.row_l1.size() != (size_t)P.n_rows) fail("row_l1");
    for(int v : P.row_l1)
        if(v < -1 || v >= P.n_coarse) fail("row_l1 range");
    for(size_t g = 0; host && g + 1 < P.gal_level_ptr.size(); ++g)
        for(int e = P.gal_level_ptr[g]; e < P.gal_level_ptr[g + 1]; ++e)
            for(int k = P.gal_ptr[e]; k < P.gal_ptr[e + 1]; ++k)
            {
                const int idx = P.gal_src[k] >> 2;
                if(idx < 0 || idx >= (P.gal_src_is_fine[g] ? nb : P.gal_level_ptr[g])) fail("gal_src");
            }
    for(int b = 0; b < P.n_bodies; ++b)
    {
        if(P.body_row[b] < 0 || P.body_row[b] + 4 > P.n_rows) fail("body_row");
        for(int k = 0; k < 16; ++k)
            if(P.body_src[16 * b + k] != -1 && (P.body_src[16 * b + k] & 0x7fffffff) >= nb) fail("body_src");
    }
    for(int v : P.diag_src)
        if(v < -1 || v >= nb) fail("diag_src");
}

// ------------------------------------------------------------ SpMV layout --

void build_spmv_layout(SpmvLayout& L, int n_sys, const std::vector<int>& row_off, const std::vector<int>& row_ptr,
                       const std::vector<int>& col, const std::vector<int>& entry_src, int cap,
                       const SpmvBodies* bodies)
{
    if(cap < 1 || cap > kSpmvMaxCap) throw std::invalid_argument("SpMV tile capacity out of range");
    const int n = (int)row_ptr.size() - 1;
    if((unsigned)n > kSpmvColMask) throw std::length_error("too many rows for the SpMV layout");
    // first entry of row r with column >= c (rows are sorted by column)
    auto first_ge = [&](int r, int c) {
        return (int)(std::lower_bound(col.begin() + row_ptr[r], col.begin() + row_ptr[r + 1], c) - col.begin());
    };
    // entries row r keeps in a tile starting at r0: columns < r0 or >= r
    auto kept = [&](int r, int r0) { return (row_ptr[r + 1] - first_ge(r, r)) + (first_ge(r, r0) - row_ptr[r]); };
    // body rows [a0, a1) of system s (empty without bodies); a body row keeps
    // its entries outside the bodies' columns (the 12 x 12 couplings take those)
    auto body_rows = [&](int s, int& a0, int& a1) {
        a0 = a1 = row_off[s];
        if(bodies && bodies->n_abd[s] > 0)
        {
            a0 = bodies->abd_row0[s];
            a1 = a0 + 4 * bodies->n_abd[s];
        }
    };
    auto kept_body = [&](int r, int a0, int a1) {
        return (row_ptr[r + 1] - row_ptr[r]) - (first_ge(r, a1) - first_ge(r, a0));
    };
    L = SpmvLayout{};
    L.cap = cap;
    L.sys_range.resize(2 * (size_t)n_sys);
    std::vector<long long> t_kept;
    std::vector<int>       t_body;  // per tile: first row of its system's bodies, -1 for a row tile
    std::vector<int>       t_body_end;
    for(int s = 0; s < n_sys; ++s)
    {
        L.sys_range[2 * s] = (int)(L.tiles.size() / 4);
        int r0 = row_off[s];
        int a0, a1;
        body_rows(s, a0, a1);
        if(a1 > a0)
        {
            if(a0 != row_off[s] || a1 > row_off[s + 1]) throw std::logic_error("SpMV layout: bodies must lead the rows");
            const int nab = bodies->n_abd[s];
            auto body_kept = [&](int b) {
                long long k = 0;
                for(int q = 0; q < 4; ++q) k += kept_body(a0 + 4 * b + q, a0, a1);
                return k;
            };
            for(int b0 = 0; b0 < nab;)
            {
                long long k = body_kept(b0);
                int b1 = b0 + 1;
                while(b1 < nab && b1 - b0 < kBodiesPerTile)
                {
                    const long long q = body_kept(b1);
                    if(k + q > cap) break;
                    k += q;
                    ++b1;
                }
                L.tiles.insert(L.tiles.end(), {s, kSpmvBodyTile << 28 | (bodies->body_base[s] + b0), a0 + 4 * b0,
                                               a0 + 4 * b1});
                t_kept.push_back(k);
                t_body.push_back(a0);
                t_body_end.push_back(a1);
                b0 = b1;
            }
            r0 = a1;
        }
        while(r0 < row_off[s + 1])
        {
            long long k = row_ptr[r0 + 1] - row_ptr[r0];
            int r1 = r0 + 1;
            while(r1 < row_off[s + 1] && r1 - r0 < kSpmvTileRows)
            {
                const int q = kept(r1, r0);
                if(k + q > cap) break;
                k += q;
                ++r1;
            }
            L.tiles.insert(L.tiles.end(), {s, 0, r0, r1});
            t_kept.push_back(k);
            t_body.push_back(-1);
            t_body_end.push_back(-1);
            r0 = r1;
        }
        L.sys_range[2 * s + 1] = (int)(L.tiles.size() / 4);
    }
    const int T = (int)t_kept.size();
    // transposed products per tile: pairs r < c inside it (a body tile's
    // entries all lie outside it)
    std::vector<long long> t_trans(T);
#pragma omp parallel for schedule(dynamic, 64)
    for(int t = 0; t < T; ++t)
    {
        const int r0 = L.tiles[4 * t + 2], r1 = L.tiles[4 * t + 3];
        long long q = 0;
        for(int r = r0; t_body[t] < 0 && r < r1; ++r) q += first_ge(r, r1) - first_ge(r, r + 1);
        t_trans[t] = q;
    }
    // tile starts aligned to kSpmvAlign entries; the gaps are padding
    std::vector<long long> e_base(T + 1, 0), e_end(T, 0), q_base(T + 1, 0);
    long long stored = 0;
    for(int t = 0; t < T; ++t)
    {
        e_end[t]      = e_base[t] + t_kept[t];
        e_base[t + 1] = (e_end[t] + kSpmvAlign - 1) / kSpmvAlign * kSpmvAlign;
        q_base[t + 1] = q_base[t] + t_trans[t];
        stored += t_kept[t];
    }
    const long long total = e_base[T] + kSpmvSlack;
    if(total > 0x7fffffffLL || q_base[T] > 0x7fffffffLL) throw std::length_error("SpMV layout too large");
    L.total     = total;
    L.n_entries = stored;
    L.n_in_tile = q_base[T];
    L.row_ptr.assign(n + 1, 0);
    L.tptr.assign(n + 1, 0);
    L.col.assign(total, 0);
    L.src.assign(total, kSpmvPadSrc);
    L.tile_ent.resize(2 * (size_t)T);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for(int t = 0; t < T; ++t)
    {
        const int r0 = L.tiles[4 * t + 2], r1 = L.tiles[4 * t + 3], nr = r1 - r0;
        int tstart[kSpmvTileRows + 1];
        std::fill(tstart, tstart + nr + 1, 0);
        for(int r = r0; t_body[t] < 0 && r < r1; ++r)
            for(int e = first_ge(r, r + 1), e_end = first_ge(r, r1); e < e_end; ++e) ++tstart[col[e] - r0 + 1];
        for(int j = 0; j < nr; ++j) tstart[j + 1] += tstart[j];
        if(t_trans[t] != tstart[nr]) bad |= 1;
        if(nr > (t_body[t] >= 0 ? 4 : 1) && t_kept[t] > cap) bad |= 2;  // a body may take several chunks
        if(tstart[nr] > kSpmvMaxCap) bad |= 4;
        int tcur[kSpmvTileRows];
        std::copy(tstart, tstart + nr, tcur);
        long long o = e_base[t];
        for(int r = r0; r < r1; ++r)
        {
            L.row_ptr[r] = (int)o;
            L.tptr[r]    = (int)(q_base[t] + tstart[r - r0]);
            unsigned first = kSpmvRowStart;
            for(int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
            {
                const int c = col[e];
                unsigned word = (unsigned)c;
                if(t_body[t] >= 0)
                {
                    if(c >= t_body[t] && c < t_body_end[t]) continue;  // in a 12 x 12 coupling
                }
                else if(c >= r0 && c < r1)
                {
                    if(c < r) continue;  // the tile's lower copy: its upper twin covers it
                    const int tpos = c == r ? 0 : tcur[c - r0]++;
                    word = kSpmvInTile | (unsigned)tpos << 16 | (unsigned)(r - r0) << 8 | (unsigned)(c - r0);
                }
                L.col[o] = (int)(word | first);
                L.src[o] = entry_src[e];
                first = 0;
                ++o;
            }
        }
        if(o != e_end[t]) bad |= 8;
        L.tile_ent[2 * t]     = (int)e_base[t];
        L.tile_ent[2 * t + 1] = (int)e_end[t];
    }
    if(bad) throw std::logic_error("SpMV layout: tile codes inconsistent (check mask " + std::to_string(bad) + ")");
    L.row_ptr[n] = (int)(T ? e_end[T - 1] : 0);
    L.tptr[n]    = (int)q_base[T];
    // 12 x 12 couplings of every body, in increasing column body
    L.cp_ptr.assign(1, 0);
    for(int s = 0; bodies && s < n_sys; ++s)
    {
        int a0, a1;
        body_rows(s, a0, a1);
        for(int b = 0; b < (a1 - a0) / 4; ++b)
        {
            std::vector<int> js;
            for(int q = 0; q < 4; ++q)
            {
                const int r = a0 + 4 * b + q;
                for(int e = first_ge(r, a0), e1 = first_ge(r, a1); e < e1; ++e) js.push_back((col[e] - a0) / 4);
            }
            std::sort(js.begin(), js.end());
            js.erase(std::unique(js.begin(), js.end()), js.end());
            for(const int j : js)
            {
                L.cp_col.push_back(bodies->body_base[s] + j);
                L.cp_body.push_back(bodies->body_base[s] + b);
                for(int q = 0; q < 4; ++q)
                    for(int t = 0; t < 4; ++t)
                    {
                        const int r = a0 + 4 * b + q, target = a0 + 4 * j + t;
                        const int e = first_ge(r, target);
                        L.cp_sub.push_back(e < row_ptr[r + 1] && col[e] == target ? e : -1);
                    }
            }
            L.cp_ptr.push_back((int)L.cp_col.size());
        }
    }
}

void spmv_layout_apply(const SpmvLayout& L, const double* upper, const double* x, double* y, const int* full_row_ptr,
                       const int* full_col, const int* full_src, const int* body_row)
{
    const int T = (int)L.tiles.size() / 4;
    const int U = std::max(1, L.cap / 256);
    // block of entry e as row-major 3x3 (an entry flagged transposed takes its
    // upper block's transpose; upper blocks are col-major)
    auto block = [&](long long e, double B[9]) {
        const int code = L.src[e];
        const double* Ub = upper + 9 * (long long)(code & 0x7fffffff);
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 3; ++j) B[3 * i + j] = code >= 0 ? Ub[i + 3 * j] : Ub[j + 3 * i];
    };
#pragma omp parallel for schedule(dynamic, 16)
    for(int t = 0; t < T; ++t)
    {
        const int r0 = L.tiles[4 * t + 2], r1 = L.tiles[4 * t + 3];
        const int e0 = L.tile_ent[2 * t], e1 = L.tile_ent[2 * t + 1];
        std::vector<double> acc(3 * (size_t)(r1 - r0), 0.0), Q(3 * (size_t)(L.tptr[r1] - L.tptr[r0]));
        const int CH = 256 * U;
        for(int base = e0; base < e1; base += CH)
        {
            const int n_ch = std::min(CH, e1 - base);
            // run ends: a thread's entries [U*th, U*th+U) summed per row run
            std::vector<double> P(3 * (size_t)n_ch, 0.0);
            for(int th = 0; th * U < n_ch; ++th)
            {
                double run[3] = {0.0, 0.0, 0.0};
                const int k_end = std::min(n_ch, th * U + U);
                for(int k = th * U; k < k_end; ++k)
                {
                    const long long e = base + k;
                    const unsigned w  = (unsigned)L.col[e];
                    if((w & kSpmvRowStart) && k > th * U)
                    {
                        for(int i = 0; i < 3; ++i) P[3 * (k - 1) + i] = run[i];
                        run[0] = run[1] = run[2] = 0.0;
                    }
                    double B[9];
                    block(e, B);
                    const double* xc;
                    if(w & kSpmvInTile)
                    {
                        const int cl = w & 0xff, rl = (w >> 8) & 0xff, tpos = (w >> 16) & 0x3fff;
                        xc = x + 3 * (long long)(r0 + cl);
                        if(cl != rl)
                        {
                            const double* xr = x + 3 * (long long)(r0 + rl);
                            for(int i = 0; i < 3; ++i)
                                Q[3 * tpos + i] = B[i] * xr[0] + B[3 + i] * xr[1] + B[6 + i] * xr[2];
                        }
                    }
                    else
                        xc = x + 3 * (long long)(w & kSpmvColMask);
                    for(int i = 0; i < 3; ++i) run[i] += B[3 * i] * xc[0] + B[3 * i + 1] * xc[1] + B[3 * i + 2] * xc[2];
                }
                for(int i = 0; i < 3; ++i) P[3 * (k_end - 1) + i] = run[i];
            }
            for(int r = r0; r < r1; ++r)
            {
                int q        = std::max(L.row_ptr[r], base) - base;
                const int qb = std::min(L.row_ptr[r + 1], base + n_ch) - base;
                while(q < qb)
                {
                    const int end = std::min(qb, (q / U + 1) * U);
                    for(int i = 0; i < 3; ++i) acc[3 * (r - r0) + i] += P[3 * (end - 1) + i];
                    q = end;
                }
            }
        }
        for(int r = r0; r < r1; ++r)
        {
            for(int q = L.tptr[r] - L.tptr[r0]; q < L.tptr[r + 1] - L.tptr[r0]; ++q)
                for(int i = 0; i < 3; ++i) acc[3 * (r - r0) + i] += Q[3 * q + i];
            for(int i = 0; i < 3; ++i) y[3 * (long long)r + i] = acc[3 * (r - r0) + i];
        }
        // a body tile: then its bodies' 12 x 12 couplings, sub-blocks summed
        // over their duplicate entries in the full CSR
        if((L.tiles[4 * t + 1] >> 28) != kSpmvBodyTile) continue;
        if(!full_row_ptr || !full_col || !full_src || !body_row)
            throw std::invalid_argument("body couplings need the full CSR and the body rows");
        const int b0 = L.tiles[4 * t + 1] & 0x0fffffff, b1 = b0 + (r1 - r0) / 4;
        for(int b = b0; b < b1; ++b)
            for(int k = L.cp_ptr[b]; k < L.cp_ptr[b + 1]; ++k)
                for(int q = 0; q < 4; ++q)
                    for(int tt = 0; tt < 4; ++tt)
                    {
                        const int r = body_row[b] + q, target = body_row[L.cp_col[k]] + tt;
                        const double* xc = x + 3 * (long long)target;
                        for(int e = L.cp_sub[16 * k + 4 * q + tt]; e >= 0 && e < full_row_ptr[r + 1] &&
                                                                   full_col[e] == target;
                            ++e)
                        {
                            double B[9];
                            const int code = full_src[e];
                            const double* Ub = upper + 9 * (long long)(code & 0x7fffffff);
                            for(int i = 0; i < 3; ++i)
                                for(int j = 0; j < 3; ++j) B[3 * i + j] = code >= 0 ? Ub[i + 3 * j] : Ub[j + 3 * i];
                            for(int i = 0; i < 3; ++i)
                                y[3 * (long long)r + i] += B[3 * i] * xc[0] + B[3 * i + 1] * xc[1] + B[3 * i + 2] * xc[2];
                        }
                    }
    }
}

}  // namespace masx
