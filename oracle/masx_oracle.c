/*
 * masx oracle implementation -- TEST INFRASTRUCTURE ONLY (see masx_oracle.h).
 * Plain C11, FP64.  Straight-line restatements of the reference kernels: each
 * GPU kernel of the reference becomes a serial loop over its thread index in
 * the same order, so the integer results (hierarchy) are bit-identical and
 * the floating point results differ only by summation order / FP64 vs FP32.
 */
#include "masx_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* helpers                                                                  */
/* ------------------------------------------------------------------------ */

/* Eigen column-major 3x3: element (i,j) at [i + 3*j]. */
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the parallel loops (bank inverses, local solves, row SpMV,
 * vector updates).  Every parallel loop writes disjoint outputs and every
 * reduction sums fixed-size chunks in a fixed order, so results do not
 * depend on the thread count. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline double blk(const double* m, int i, int j) { return m[i + 3 * j]; }

static inline int pad_to(int v, int bank) { return (v + bank - 1) / bank * bank; }

static inline int packed_index(int bank, int r, int c) { return bank * r - r * (r + 1) / 2 + c; }

static int popc32(uint32_t v) { return __builtin_popcount(v); }
static int ffs32(uint32_t v) { return __builtin_ffs((int)v); } /* 1-based, 0 if none */
static uint32_t lanemask_lt(int lane) { return lane >= 32 ? 0xffffffffu : ((1u << lane) - 1u); }
static uint32_t lanebit(int lane) { return 1u << lane; }

/* ------------------------------------------------------------------------ */
/* SpMV  (linear_system/utils/spmv.cu:28-53, :207-233)                      */
/* ------------------------------------------------------------------------ */

void orc_sym_bcoo_spmv(int nb, const double* vals, const int* rows, const int* cols,
                       double a, const double* x, double b, double* y, int dof)
{
    /* scale_y_kernel / fill_y_zero_kernel (spmv.cu:14-26) */
    if(b != 0.0)
        for(int i = 0; i < dof; ++i) y[i] *= b;
    else
        for(int i = 0; i < dof; ++i) y[i] = 0.0;
    for(int k = 0; k < nb; ++k)
    {
        const double* B = vals + 9 * (size_t)k;
        const int     i = rows[k], j = cols[k];
        for(int r = 0; r < 3; ++r)
        {
            double s = 0.0;
            for(int c = 0; c < 3; ++c) s += blk(B, r, c) * x[3 * j + c];
            y[3 * i + r] += a * s;
        }
        if(i != j)
            for(int r = 0; r < 3; ++r)
            {
                double s = 0.0;
                for(int c = 0; c < 3; ++c) s += blk(B, c, r) * x[3 * i + c];
                y[3 * j + r] += a * s;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Morton partition (gl_main.cu:1176-1243)                                  */
/* ------------------------------------------------------------------------ */

static uint32_t expand_bits_10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static uint32_t quantize10(double value, double lo, double hi)
{
    const double extent = hi - lo;
    double       t      = extent > 0.0 ? (value - lo) / extent : 0.0;
    if(t < 0.0) t = 0.0;
    if(t > 1.0) t = 1.0;
    return (uint32_t)(t * 1023.0 + 0.5);
}

typedef struct { uint32_t code; int index; } morton_key;

static int morton_cmp(const void* a, const void* b)
{
    const morton_key* x = (const morton_key*)a;
    const morton_key* y = (const morton_key*)b;
    if(x->code != y->code) return x->code < y->code ? -1 : 1;
    return (x->index > y->index) - (x->index < y->index);
}

int orc_morton_partition(int n, const double* pos, int bank, int* part_to_real, int* real_to_part)
{
    if(n <= 0) return 0;
    double lo[3], hi[3];
    for(int d = 0; d < 3; ++d) lo[d] = hi[d] = pos[d];
    for(int v = 1; v < n; ++v)
        for(int d = 0; d < 3; ++d)
        {
            const double p = pos[3 * v + d];
            if(p < lo[d]) lo[d] = p;
            if(p > hi[d]) hi[d] = p;
        }
    morton_key* keys = (morton_key*)malloc(sizeof(morton_key) * (size_t)n);
    for(int v = 0; v < n; ++v)
    {
        const uint32_t x = expand_bits_10(quantize10(pos[3 * v + 0], lo[0], hi[0]));
        const uint32_t y = expand_bits_10(quantize10(pos[3 * v + 1], lo[1], hi[1]));
        const uint32_t z = expand_bits_10(quantize10(pos[3 * v + 2], lo[2], hi[2]));
        keys[v].code  = x | (y << 1) | (z << 2);
        keys[v].index = v;
    }
    /* stable_sort with ties broken by index == total order on (code, index) */
    qsort(keys, (size_t)n, sizeof(morton_key), morton_cmp);
    const int padded = pad_to(n, bank);
    for(int s = 0; s < padded; ++s) part_to_real[s] = -1;
    for(int s = 0; s < n; ++s)
    {
        part_to_real[s]             = keys[s].index;
        real_to_part[keys[s].index] = s;
    }
    free(keys);
    return padded;
}

/* ------------------------------------------------------------------------ */
/* Neighbor lists (load_mesh.cpp:934-1040)                                  */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t a, b; } edge_pair;

static int edge_cmp(const void* x, const void* y)
{
    const edge_pair* p = (const edge_pair*)x;
    const edge_pair* q = (const edge_pair*)y;
    if(p->a != q->a) return p->a < q->a ? -1 : 1;
    return (p->b > q->b) - (p->b < q->b);
}

int orc_mesh_neighbors(int n, int ntet, const int* tets, int ntri, const int* tris,
                       unsigned* start, unsigned* num, unsigned* list, int list_cap)
{
    /* The reference keeps each undirected edge in the orientation it was first
     * inserted and iterates an ordered set of (first, second). */
    if(n <= 0 || ntet < 0 || ntri < 0) return 0;
    const size_t cap   = (size_t)ntet * 6 + (size_t)ntri * 3;
    edge_pair*   edges = (edge_pair*)malloc(sizeof(edge_pair) * (cap ? cap : 1));
    size_t       ne    = 0;
    static const int tet_edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for(int t = 0; t < ntet; ++t)
        for(int e = 0; e < 6; ++e)
        {
            edges[ne].a = (uint64_t)tets[4 * t + tet_edges[e][0]];
            edges[ne].b = (uint64_t)tets[4 * t + tet_edges[e][1]];
            ++ne;
        }
    static const int tri_edges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for(int t = 0; t < ntri; ++t)
        for(int e = 0; e < 3; ++e)
        {
            edges[ne].a = (uint64_t)tris[3 * t + tri_edges[e][0]];
            edges[ne].b = (uint64_t)tris[3 * t + tri_edges[e][1]];
            ++ne;
        }
    /* first-inserted orientation: canonicalise to find duplicates, remember
     * the insertion rank to keep the first orientation */
    typedef struct { uint64_t lo, hi; size_t rank; } canon;
    canon* cz = (canon*)malloc(sizeof(canon) * (ne ? ne : 1));
    for(size_t i = 0; i < ne; ++i)
    {
        cz[i].lo   = edges[i].a < edges[i].b ? edges[i].a : edges[i].b;
        cz[i].hi   = edges[i].a < edges[i].b ? edges[i].b : edges[i].a;
        cz[i].rank = i;
    }
    /* sort by (lo, hi, rank) with a simple merge sort for determinism */
    for(size_t width = 1; width < ne; width *= 2)
    {
        canon* tmp = (canon*)malloc(sizeof(canon) * ne);
        for(size_t lo = 0; lo < ne; lo += 2 * width)
        {
            size_t mid = lo + width < ne ? lo + width : ne;
            size_t hi  = lo + 2 * width < ne ? lo + 2 * width : ne;
            size_t i = lo, j = mid, k = lo;
            while(i < mid && j < hi)
            {
                const canon* p = cz + i;
                const canon* q = cz + j;
                int less = p->lo != q->lo ? p->lo < q->lo
                         : p->hi != q->hi ? p->hi < q->hi : p->rank < q->rank;
                tmp[k++] = less ? cz[i++] : cz[j++];
            }
            while(i < mid) tmp[k++] = cz[i++];
            while(j < hi) tmp[k++] = cz[j++];
        }
        memcpy(cz, tmp, sizeof(canon) * ne);
        free(tmp);
    }
    size_t     nu     = 0;
    edge_pair* unique = (edge_pair*)malloc(sizeof(edge_pair) * (ne ? ne : 1));
    for(size_t i = 0; i < ne; ++i)
    {
        if(i > 0 && cz[i].lo == cz[i - 1].lo && cz[i].hi == cz[i - 1].hi) continue;
        if(cz[i].lo == cz[i].hi) continue; /* degenerate */
        unique[nu++] = edges[cz[i].rank];  /* first inserted orientation */
    }
    free(cz);
    free(edges);
    edges = unique;
    qsort(edges, nu, sizeof(edge_pair), edge_cmp);
    if((long long)nu * 2 > (long long)list_cap)
    {
        free(edges);
        return -1;
    }
    for(int v = 0; v < n; ++v) num[v] = 0;
    for(size_t i = 0; i < nu; ++i)
    {
        num[edges[i].a]++;
        num[edges[i].b]++;
    }
    unsigned off = 0;
    for(int v = 0; v < n; ++v)
    {
        start[v] = off;
        off += num[v];
    }
    unsigned* fill = (unsigned*)calloc((size_t)n, sizeof(unsigned));
    for(size_t i = 0; i < nu; ++i)
    {
        const uint64_t f = edges[i].a, s = edges[i].b;
        list[start[f] + fill[f]++] = (unsigned)s;
        list[start[s] + fill[s]++] = (unsigned)f;
    }
    free(fill);
    free(edges);
    return (int)(nu * 2);
}

/* ------------------------------------------------------------------------ */
/* Level count and capacity                                                 */
/* ------------------------------------------------------------------------ */

int orc_num_levels(int vert_num, int part_map_size, int bank)
{
    const int max_nodes = part_map_size > vert_num ? part_map_size : vert_num;
    int       nlevel    = 1;
    int       level_sz  = pad_to(max_nodes, bank);
    while(level_sz > bank)
    {
        level_sz /= bank;
        nlevel++;
        level_sz = pad_to(level_sz, bank);
    }
    nlevel = nlevel + 1;
    return nlevel > 6 ? 6 : nlevel;
}

long long orc_hierarchy_capacity(long long vertices, long long mapped, long long levels, long long bank)
{
    const long long base = vertices > mapped ? vertices : mapped;
    if(!base || !levels) return 0;
    const long long maximum = (long long)(SIZE_MAX / sizeof(unsigned int) > (size_t)INT64_MAX
                                              ? INT64_MAX
                                              : SIZE_MAX / sizeof(unsigned int));
    if(!bank || base > maximum || bank - 1 > maximum - base) return -1;
    const long long padded = (base + bank - 1) / bank * bank;
    if(padded > maximum / levels) return -1;
    return padded * levels;
}

/* ------------------------------------------------------------------------ */
/* MAS object                                                               */
/* ------------------------------------------------------------------------ */

struct orc_mas
{
    int bank, n_fem, n_slots, levels;
    int nblk; /* packed blocks per bank */
    /* topology (copies) */
    unsigned *nstart, *nnum_init, *nlist_init;
    int       nlist_len;
    int *     part_to_real, *real_to_part;
    /* hierarchy */
    unsigned* nnum;        /* working (compacted) neighbor counts */
    unsigned* nlist;       /* working neighbor list */
    unsigned* fine_mask;   /* n_fem */
    int*      prefix;      /* components per level-0 bank */
    int*      level_count; /* levels+1 */
    int*      level_begin; /* levels+1 */
    int*      going_next;  /* capacity */
    int*      coarse_space;/* n_fem*levels */
    int*      coarse_table;/* n_fem*6 */
    int       total_clusters;
    int       capacity;
    /* matrices */
    double* local;   /* banks * nblk * 9 */
    float*  inverse; /* banks * nblk * 9 */
};

orc_mas* orc_mas_create(int n_fem, int n_slots, int bank, int levels_override,
                        const unsigned* start, const unsigned* num, const unsigned* list,
                        const int* part_to_real, const int* real_to_part)
{
    if(bank < 2 || bank > 32 || n_fem < 1 || n_slots < n_fem || n_slots % bank) return NULL;
    orc_mas* m   = (orc_mas*)calloc(1, sizeof(orc_mas));
    m->bank      = bank;
    m->n_fem     = n_fem;
    m->n_slots   = n_slots;
    m->levels    = levels_override > 0 ? levels_override : orc_num_levels(n_fem, n_slots, bank);
    if(m->levels > 6) m->levels = 6;
    m->nblk      = bank * (bank + 1) / 2;
    m->nlist_len = 0;
    for(int v = 0; v < n_fem; ++v)
        if((int)(start[v] + num[v]) > m->nlist_len) m->nlist_len = (int)(start[v] + num[v]);
    m->nstart     = (unsigned*)malloc(sizeof(unsigned) * n_fem);
    m->nnum_init  = (unsigned*)malloc(sizeof(unsigned) * n_fem);
    m->nnum       = (unsigned*)malloc(sizeof(unsigned) * n_fem);
    m->nlist_init = (unsigned*)malloc(sizeof(unsigned) * (m->nlist_len + 1));
    m->nlist      = (unsigned*)malloc(sizeof(unsigned) * (m->nlist_len + 1));
    memcpy(m->nstart, start, sizeof(unsigned) * n_fem);
    memcpy(m->nnum_init, num, sizeof(unsigned) * n_fem);
    memcpy(m->nlist_init, list, sizeof(unsigned) * m->nlist_len);
    m->part_to_real = (int*)malloc(sizeof(int) * n_slots);
    m->real_to_part = (int*)malloc(sizeof(int) * n_fem);
    memcpy(m->part_to_real, part_to_real, sizeof(int) * n_slots);
    memcpy(m->real_to_part, real_to_part, sizeof(int) * n_fem);
    m->fine_mask    = (unsigned*)calloc(n_fem, sizeof(unsigned));
    m->prefix       = (int*)calloc(n_slots / bank, sizeof(int));
    m->level_count  = (int*)calloc(m->levels + 2, sizeof(int));
    m->level_begin  = (int*)calloc(m->levels + 2, sizeof(int));
    m->coarse_space = (int*)calloc((size_t)n_fem * m->levels, sizeof(int));
    m->coarse_table = (int*)calloc((size_t)n_fem * 6, sizeof(int));
    /* every level holds at most as many nodes as the level below, plus a bank
     * of padding per level */
    m->capacity   = n_slots * (m->levels + 1) + bank * (m->levels + 2);
    m->going_next = (int*)malloc(sizeof(int) * m->capacity);
    for(int i = 0; i < m->capacity; ++i) m->going_next[i] = -1;
    return m;
}

void orc_mas_destroy(orc_mas* m)
{
    if(!m) return;
    free(m->nstart); free(m->nnum_init); free(m->nnum); free(m->nlist_init); free(m->nlist);
    free(m->part_to_real); free(m->real_to_part); free(m->fine_mask); free(m->prefix);
    free(m->level_count); free(m->level_begin); free(m->going_next); free(m->coarse_space);
    free(m->coarse_table); free(m->local); free(m->inverse);
    free(m);
}

int orc_mas_levels(const orc_mas* m) { return m->levels; }
int orc_mas_total_clusters(const orc_mas* m) { return m->total_clusters; }

/* decode a sign-encoded collision pair into its vertex list
 * (TraditionalMAS32Preconditioner.cu:1550-1960) */
static int decode_pair(const int* p, int v[4])
{
    int x = p[0], y = p[1], z = p[2], w = p[3];
    if(x >= 0)
    {
        if(w < 0) w = -w - 1;
        v[0] = x; v[1] = y; v[2] = z; v[3] = w;
        return 4;
    }
    x = -x - 1;
    if(z < 0)
    {
        if(y < 0)
        {
            y = -y - 1; z = -z - 1; w = -w - 1;
            v[0] = x; v[1] = y; v[2] = z; v[3] = w;
            return 4;
        }
        v[0] = x; v[1] = y;
        return 2;
    }
    if(w < 0)
    {
        if(y < 0)
        {
            y = -y - 1; w = -w - 1;
            v[0] = x; v[1] = y; v[2] = z; v[3] = w;
            return 4;
        }
        v[0] = x; v[1] = y; v[2] = z;
        return 3;
    }
    v[0] = x; v[1] = y; v[2] = z; v[3] = w;
    return 4;
}

/* _buildCollisionConnection_new.  level < 0: fine level, masks indexed by real
 * vertex, ids are slots; level >= 1: masks indexed by level-`level` id. */
static void collision_connection(orc_mas* m, unsigned* conn, int level, int cp_num,
                                 const int* pairs, int node_offset)
{
    const int bank = m->bank;
    for(int k = 0; k < cp_num; ++k)
    {
        int       vid[4];
        const int cnt = decode_pair(pairs + 4 * k, vid);
        int       cp[4];
        for(int i = 0; i < cnt; ++i)
        {
            vid[i] -= node_offset;
            if(vid[i] >= 0 && vid[i] < m->n_fem)
                cp[i] = level < 0 ? m->real_to_part[vid[i]]
                                  : m->coarse_space[(size_t)(level - 1) * m->n_fem + vid[i]];
            else
                cp[i] = -1;
        }
        unsigned msk[4] = {0, 0, 0, 0};
        for(int i = 0; i < cnt; ++i)
            for(int j = i + 1; j < cnt; ++j)
            {
                const unsigned my = (unsigned)cp[i], ot = (unsigned)cp[j];
                if(my == ot) continue;
                if(my / (unsigned)bank == ot / (unsigned)bank)
                {
                    msk[i] |= lanebit((int)(ot % (unsigned)bank));
                    msk[j] |= lanebit((int)(my % (unsigned)bank));
                }
            }
        for(int i = 0; i < cnt; ++i)
        {
            if(level < 0)
            {
                if(vid[i] >= 0 && vid[i] < m->n_fem) conn[vid[i]] |= msk[i];
            }
            else if(cp[i] >= 0)
                conn[cp[i]] |= msk[i];
        }
    }
}

/* connected-component closure of one bank by bit-BFS over the original
 * adjacency (the loops of _preparePrefixSumL0_new / _nextLevelCluster) */
static uint32_t bank_closure(const uint32_t* adj, int lane, int bank_lanes_valid_unused)
{
    (void)bank_lanes_valid_unused;
    uint32_t mask    = adj[lane];
    uint32_t visited = lanebit(lane);
    while(mask != 0xffffffffu)
    {
        const uint32_t todo = visited ^ mask;
        if(!todo) break;
        const int next = ffs32(todo) - 1;
        visited |= lanebit(next);
        mask |= adj[next];
    }
    return mask;
}

int orc_mas_reorder(orc_mas* m, int cp_num, const int* pairs, int node_offset)
{
    const int bank   = m->bank;
    const int n      = m->n_fem;
    const int nslots = m->n_slots;
    const int banks0 = nslots / bank;
    memcpy(m->nnum, m->nnum_init, sizeof(unsigned) * n);
    memcpy(m->nlist, m->nlist_init, sizeof(unsigned) * m->nlist_len);
    for(int i = 0; i <= m->levels; ++i) m->level_count[i] = m->level_begin[i] = 0;
    for(int i = 0; i < m->capacity; ++i) m->going_next[i] = -1;

    /* _buildCML0_new (:463-503) */
    for(int s = 0; s < nslots; ++s)
    {
        const int idx = m->part_to_real[s];
        if(idx < 0) continue;
        const int      b = s / bank, lane = s % bank;
        uint32_t       msk = lanebit(lane);
        unsigned       nk  = 0;
        const unsigned st  = m->nstart[idx];
        for(unsigned i = 0; i < m->nnum[idx]; ++i)
        {
            const unsigned u  = m->nlist[st + i];
            const int      su = m->real_to_part[u];
            if(su / bank == b)
                msk |= lanebit(su % bank);
            else
                m->nlist[st + nk++] = u;
        }
        m->nnum[idx]      = nk;
        m->fine_mask[idx] = msk;
    }
    if(cp_num && pairs) collision_connection(m, m->fine_mask, -1, cp_num, pairs, node_offset);

    /* _preparePrefixSumL0_new (:560-620) */
    uint32_t adj[32];
    for(int b = 0; b < banks0; ++b)
    {
        for(int l = 0; l < bank; ++l)
        {
            const int idx = m->part_to_real[b * bank + l];
            adj[l]        = idx >= 0 ? m->fine_mask[idx] : 0u;
        }
        int count = 0;
        for(int l = 0; l < bank; ++l)
        {
            const int idx = m->part_to_real[b * bank + l];
            if(idx < 0) continue;
            const uint32_t closure = bank_closure(adj, l, 0);
            m->fine_mask[idx]      = closure;
            if(popc32(closure & lanemask_lt(l)) == 0) count++;
        }
        m->prefix[b] = count;
    }

    /* _buildLevel1_new (:681-743) with the exclusive scan of BuildLevel1 */
    int* prefix_sum = (int*)malloc(sizeof(int) * (banks0 + 1));
    prefix_sum[0]   = 0;
    for(int b = 0; b < banks0; ++b) prefix_sum[b + 1] = prefix_sum[b] + m->prefix[b];
    m->level_count[1] = prefix_sum[banks0];
    m->level_begin[1] = pad_to(nslots, bank);
    for(int b = 0; b < banks0; ++b)
    {
        uint32_t elected = 0;
        for(int l = 0; l < bank; ++l)
        {
            const int idx = m->part_to_real[b * bank + l];
            if(idx >= 0 && popc32(m->fine_mask[idx] & lanemask_lt(l)) == 0) elected |= lanebit(l);
        }
        for(int l = 0; l < bank; ++l)
        {
            const int idx = m->part_to_real[b * bank + l];
            if(idx < 0) continue;
            const int leader = ffs32(m->fine_mask[idx]) - 1;
            const int id     = prefix_sum[b] + popc32(elected & lanemask_lt(leader));
            m->coarse_space[idx] = id;
            m->going_next[idx]   = id + m->level_begin[1];
        }
    }
    free(prefix_sum);

    unsigned* next_mask = (unsigned*)malloc(sizeof(unsigned) * (nslots + bank));
    for(int level = 1; level < m->levels; ++level)
    {
        const int count = m->level_count[level];
        const int begin = m->level_begin[level];
        memset(next_mask, 0, sizeof(unsigned) * (nslots + bank));
        /* _buildConnectMaskLx_new (:821-902): mask of the coarse node owned by
         * each fine vertex's level-`level` id, over the not-yet-merged list */
        const int* cs_prev = m->coarse_space + (size_t)(level - 1) * n;
        for(int s = 0; s < nslots; ++s)
        {
            const int idx = m->part_to_real[s];
            if(idx < 0) continue;
            const int      c   = cs_prev[idx];
            uint32_t       msk = 0;
            unsigned       nk  = 0;
            const unsigned st  = m->nstart[idx];
            for(unsigned i = 0; i < m->nnum[idx]; ++i)
            {
                const unsigned u  = m->nlist[st + i];
                const int      cu = cs_prev[u];
                if(c / bank == cu / bank)
                    msk |= lanebit(cu % bank);
                else
                    m->nlist[st + nk++] = u;
            }
            m->nnum[idx] = nk;
            next_mask[c] |= msk;
        }
        if(cp_num && pairs) collision_connection(m, next_mask, level, cp_num, pairs, node_offset);

        /* _nextLevelCluster (:905-956) */
        const int nb_level = (count + bank - 1) / bank;
        int*      nprefix  = (int*)calloc(nb_level + 1, sizeof(int));
        for(int b = 0; b < nb_level; ++b)
        {
            const int lanes = count - b * bank < bank ? count - b * bank : bank;
            for(int l = 0; l < bank; ++l) adj[l] = 0;
            for(int l = 0; l < lanes; ++l) adj[l] = lanebit(l) | next_mask[b * bank + l];
            int comps = 0;
            for(int l = 0; l < lanes; ++l)
            {
                const uint32_t closure = bank_closure(adj, l, 0);
                next_mask[b * bank + l] = closure;
                if(popc32(closure & lanemask_lt(l)) == 0) comps++;
            }
            nprefix[b] = comps;
        }
        /* _prefixSumLx (:958-1009) */
        int* nsum = (int*)calloc(nb_level + 1, sizeof(int));
        for(int b = 0; b < nb_level; ++b) nsum[b + 1] = nsum[b] + nprefix[b];
        m->level_count[level + 1] = nsum[nb_level];
        m->level_begin[level + 1] = begin + pad_to(count, bank);
        int* newid = (int*)malloc(sizeof(int) * (count + 1));
        for(int b = 0; b < nb_level; ++b)
        {
            const int lanes   = count - b * bank < bank ? count - b * bank : bank;
            uint32_t  elected = 0;
            for(int l = 0; l < lanes; ++l)
                if(popc32(next_mask[b * bank + l] & lanemask_lt(l)) == 0) elected |= lanebit(l);
            for(int l = 0; l < lanes; ++l)
            {
                const int leader = ffs32(next_mask[b * bank + l]) - 1;
                const int id     = nsum[b] + popc32(elected & lanemask_lt(leader));
                newid[b * bank + l] = id;
                m->going_next[begin + b * bank + l] = id + m->level_begin[level + 1];
            }
        }
        /* _computeNextLevel (:1011-1022) */
        int* cs_next = m->coarse_space + (size_t)level * n;
        for(int v = 0; v < n; ++v) cs_next[v] = newid[cs_prev[v]];
        free(newid);
        free(nsum);
        free(nprefix);
    }
    free(next_mask);
    m->total_clusters = m->level_begin[m->levels];

    /* _aggregationKernel (:1024-1057) */
    for(int v = 0; v < n; ++v)
    {
        int cur = v;
        for(int l = 0; l < 6; ++l) m->coarse_table[6 * v + l] = 0;
        for(int l = 0; l < m->levels - 1; ++l)
        {
            cur                        = m->going_next[cur];
            m->coarse_table[6 * v + l] = cur;
        }
    }
    return m->total_clusters;
}

void orc_mas_get_hierarchy(const orc_mas* m, int* level_count, int* level_begin, unsigned* fine_mask,
                           int* prefix, int* going_next, int* coarse_table)
{
    if(level_count) memcpy(level_count, m->level_count, sizeof(int) * (m->levels + 1));
    if(level_begin) memcpy(level_begin, m->level_begin, sizeof(int) * (m->levels + 1));
    if(fine_mask) memcpy(fine_mask, m->fine_mask, sizeof(unsigned) * m->n_fem);
    if(prefix) memcpy(prefix, m->prefix, sizeof(int) * (m->n_slots / m->bank));
    if(going_next) memcpy(going_next, m->going_next, sizeof(int) * m->total_clusters);
    if(coarse_table) memcpy(coarse_table, m->coarse_table, sizeof(int) * 6 * (size_t)m->n_fem);
}

/* ------------------------------------------------------------------------ */
/* Galerkin scatter / aggregate (PrepareHessian_bcoo, :2234-2525)           */
/* ------------------------------------------------------------------------ */

static void add_block(double* dst, const double* H, int transpose, double scale)
{
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j)
            dst[i + 3 * j] += scale * (transpose ? H[j + 3 * i] : H[i + 3 * j]);
}

int orc_mas_assemble(orc_mas* m, const double* vals, const int* rows, const int* cols,
                     const unsigned* indices, int offset, int count)
{
    const int    bank  = m->bank;
    const int    banks = m->total_clusters / bank;
    const size_t bsz   = (size_t)m->nblk * 9;
    free(m->local);
    m->local = (double*)calloc((size_t)banks * bsz, sizeof(double));

    /* prepare_hessian_bcoo_kernel */
    for(int I = 0; I < count; ++I)
    {
        const int     index = indices ? (int)indices[I] : I;
        const int     rr    = rows[index] - offset;
        const int     cr    = cols[index] - offset;
        const double* H     = vals + 9 * (size_t)index;
        int           vc    = m->real_to_part[cr];
        int           vr    = m->real_to_part[rr];
        if(vc / bank == vr / bank)
        {
            double* dst;
            int     tr;
            if(vc >= vr)
            {
                dst = m->local + (size_t)(vc / bank) * bsz + 9 * (size_t)packed_index(bank, vr % bank, vc % bank);
                tr  = 0;
            }
            else
            {
                dst = m->local + (size_t)(vc / bank) * bsz + 9 * (size_t)packed_index(bank, vc % bank, vr % bank);
                tr  = 1;
            }
            for(int i = 0; i < 3; ++i)
                for(int j = 0; j < 3; ++j) dst[i + 3 * j] = tr ? H[j + 3 * i] : H[i + 3 * j];
        }
        else
        {
            for(int level = 1; level < m->levels; ++level)
            {
                if(level == 1)
                {
                    vc = m->going_next[cr];
                    vr = m->going_next[rr];
                }
                else
                {
                    vc = m->going_next[vc];
                    vr = m->going_next[vr];
                }
                if(vc / bank != vr / bank) continue;
                double* base = m->local + (size_t)(vc / bank) * bsz;
                if(vc >= vr)
                {
                    double* dst = base + 9 * (size_t)packed_index(bank, vr % bank, vc % bank);
                    add_block(dst, H, 0, 1.0);
                    if(vc == vr) add_block(dst, H, 1, 1.0);
                }
                else
                    add_block(base + 9 * (size_t)packed_index(bank, vc % bank, vr % bank), H, 1, 1.0);
            }
        }
    }

    /* prepare_hessian_bcoo_sum_kernel: thread (bank h, row lr, col lc) */
    const int banks0 = m->n_slots / bank;
    double    mat3[9], rowsum[9];
    for(int h = 0; h < banks0; ++h)
    {
        for(int lr = 0; lr < bank; ++lr)
        {
            const int rdx = m->part_to_real[h * bank + lr];
            if(m->prefix[h] == 1)
            {
                /* warp segmented sum over the row, delivered by the row's lane
                 * 0 (which must itself be mapped) to the diagonal block of
                 * every ancestor */
                if(rdx < 0 || m->part_to_real[h * bank + 0] < 0) continue;
                for(int e = 0; e < 9; ++e) rowsum[e] = 0.0;
                for(int lc = 0; lc < bank; ++lc)
                {
                    if(m->part_to_real[h * bank + lc] < 0) continue;
                    const double* src = m->local + (size_t)h * bsz
                                        + 9 * (size_t)(lc >= lr ? packed_index(bank, lr, lc)
                                                                : packed_index(bank, lc, lr));
                    for(int i = 0; i < 3; ++i)
                        for(int j = 0; j < 3; ++j)
                            rowsum[i + 3 * j] += lc >= lr ? src[i + 3 * j] : src[j + 3 * i];
                }
                int next = m->going_next[rdx];
                for(int level = 1; level < m->levels; ++level)
                {
                    double* dst = m->local + (size_t)(next / bank) * bsz
                                  + 9 * (size_t)packed_index(bank, next % bank, next % bank);
                    add_block(dst, rowsum, 0, 1.0);
                    next = m->going_next[next];
                }
            }
            else
            {
                for(int lc = 0; lc < bank; ++lc)
                {
                    int cdx = m->part_to_real[h * bank + lc];
                    int r2  = rdx;
                    if(r2 < 0 || cdx < 0) continue;
                    const double* src = m->local + (size_t)h * bsz
                                        + 9 * (size_t)(lc >= lr ? packed_index(bank, lr, lc)
                                                                : packed_index(bank, lc, lr));
                    for(int i = 0; i < 3; ++i)
                        for(int j = 0; j < 3; ++j)
                            mat3[i + 3 * j] = lc >= lr ? src[i + 3 * j] : src[j + 3 * i];
                    for(int level = 1; level < m->levels; ++level)
                    {
                        r2  = m->going_next[r2];
                        cdx = m->going_next[cdx];
                        if(r2 / bank == cdx / bank && cdx >= r2)
                            add_block(m->local + (size_t)(cdx / bank) * bsz
                                          + 9 * (size_t)packed_index(bank, r2 % bank, cdx % bank),
                                      mat3, 0, 1.0);
                    }
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Local inverse                                                            */
/* ------------------------------------------------------------------------ */

/* expand_symmetric_mas32 (:1062-1097): full row-major dim x dim */
static void expand_packed(const double* packed, int bank, double* full)
{
    const int dim = 3 * bank;
    for(int row = 0; row < dim; ++row)
    {
        for(int col = 0; col < dim; ++col)
        {
            const int br = row / 3, bc = col / 3;
            if(bc >= br)
                full[row * dim + col] = packed[9 * packed_index(bank, br, bc) + (row % 3) + 3 * (col % 3)];
            else
                full[row * dim + col] = packed[9 * packed_index(bank, bc, br) + (col % 3) + 3 * (row % 3)];
        }
        if(full[row * dim + row] == 0.0) full[row * dim + row] = 1.0;
    }
}

/* __inverse6_P96x96 (:1099-1146): in-place Gauss-Jordan, no pivoting */
static void gauss_jordan_nopivot(double* m, int dim, double* column)
{
    for(int j = 0; j < dim; ++j)
    {
        const double pivot = m[j * dim + j];
        for(int i = 0; i < dim; ++i) column[i] = m[i * dim + j];
        for(int i = 0; i < dim; ++i) m[i * dim + j] = i == j ? 1.0 : 0.0;
        for(int i = 0; i < dim; ++i) m[j * dim + i] /= pivot;
        for(int i = 0; i < dim; ++i)
            for(int k = 0; k < dim; ++k)
                if(k != j) m[k * dim + i] += -column[k] * m[j * dim + i];
    }
}

/* pro factor_mas32_cholesky (:1382-1439) + invert_mas32_cholesky (:1475-1510):
 * lower Cholesky with pivot floor, inverse column by column. */
static void cholesky_inverse(double* m, int dim, double* sol)
{
    double scale = 0.0;
    for(int i = 0; i < dim; ++i)
        if(fabs(m[i * dim + i]) > scale) scale = fabs(m[i * dim + i]);
    const double floor_value = fmax(1.0e-12, scale * 1.0e-12);
    for(int k = 0; k < dim; ++k)
    {
        double value = m[k * dim + k];
        if(!isfinite(value)) value = floor_value;
        if(value < floor_value) value = floor_value;
        const double pivot = sqrt(value);
        m[k * dim + k]     = pivot;
        for(int i = k + 1; i < dim; ++i)
        {
            double v = m[i * dim + k] / pivot;
            if(!isfinite(v)) v = 0.0;
            m[i * dim + k] = v;
        }
        for(int i = k + 1; i < dim; ++i)
            for(int j = k + 1; j <= i; ++j) m[i * dim + j] -= m[i * dim + k] * m[j * dim + k];
    }
    /* inverse into the upper triangle storage of a fresh matrix */
    double* inv = sol + dim;
    for(int col = 0; col < dim; ++col)
    {
        for(int i = 0; i < dim; ++i)
        {
            double s = i == col ? 1.0 : 0.0;
            for(int j = 0; j < i; ++j) s -= m[i * dim + j] * sol[j];
            sol[i] = s / m[i * dim + i];
        }
        for(int i = dim - 1; i >= 0; --i)
        {
            double s = sol[i];
            for(int j = i + 1; j < dim; ++j) s -= m[j * dim + i] * sol[j];
            sol[i] = s / m[i * dim + i];
        }
        for(int row = 0; row < dim; ++row) inv[row * dim + col] = sol[row];
    }
    memcpy(m, inv, sizeof(double) * dim * dim);
}

int orc_mas_invert(orc_mas* m, int method)
{
    const int    bank  = m->bank, dim = 3 * bank;
    const int    banks = m->total_clusters / bank;
    const size_t bsz   = (size_t)m->nblk * 9;
    free(m->inverse);
    m->inverse = (float*)calloc((size_t)banks * bsz, sizeof(float));
#pragma omp parallel
    {
    double* full = (double*)malloc(sizeof(double) * dim * dim);
    double* work = (double*)malloc(sizeof(double) * (dim + dim * dim));
#pragma omp for schedule(dynamic, 4)
    for(int b = 0; b < banks; ++b)
    {
        expand_packed(m->local + (size_t)b * bsz, bank, full);
        if(method == 1)
            cholesky_inverse(full, dim, work);
        else
            gauss_jordan_nopivot(full, dim, work);
        /* store block_col >= block_row of the computed inverse as FP32 */
        float* dst = m->inverse + (size_t)b * bsz;
        for(int row = 0; row < dim; ++row)
            for(int col = 0; col < dim; ++col)
            {
                const int br = row / 3, bc = col / 3;
                if(bc >= br)
                    dst[9 * packed_index(bank, br, bc) + (row % 3) + 3 * (col % 3)] = (float)full[row * dim + col];
            }
    }
    free(full);
    free(work);
    }
    return 0;
}

void orc_mas_get_local(const orc_mas* m, double* out)
{
    memcpy(out, m->local, sizeof(double) * (size_t)(m->total_clusters / m->bank) * m->nblk * 9);
}

void orc_mas_get_inverse(const orc_mas* m, float* out)
{
    memcpy(out, m->inverse, sizeof(float) * (size_t)(m->total_clusters / m->bank) * m->nblk * 9);
}

void orc_mas_set_inverse(orc_mas* m, const float* in)
{
    const size_t n = (size_t)(m->total_clusters / m->bank) * m->nblk * 9;
    free(m->inverse);
    m->inverse = (float*)malloc(sizeof(float) * n);
    memcpy(m->inverse, in, sizeof(float) * n);
}

/* ------------------------------------------------------------------------ */
/* MAS apply (:1149-1297, :1375-1445, :2684-2725)                           */
/* ------------------------------------------------------------------------ */

void orc_mas_apply(const orc_mas* m, const double* r, double* z)
{
    const int    bank  = m->bank;
    const int    total = m->total_clusters;
    const size_t bsz   = (size_t)m->nblk * 9;
    double*      R     = (double*)calloc((size_t)total * 3, sizeof(double));
    double*      Z     = (double*)calloc((size_t)total * 3, sizeof(double));
    /* restriction: level 0 is the slot copy, each coarse slot sums its members */
#pragma omp parallel for schedule(static)
    for(int s = 0; s < m->n_slots; ++s)
    {
        const int v = m->part_to_real[s];
        if(v < 0) continue;
        for(int d = 0; d < 3; ++d) R[3 * s + d] = r[3 * v + d];
    }
    for(int v = 0; v < m->n_fem; ++v)
        for(int l = 0; l < m->levels - 1; ++l)
        {
            const int c = m->coarse_table[6 * v + l];
            for(int d = 0; d < 3; ++d) R[3 * c + d] += r[3 * v + d];
        }
    /* _schwarzLocalXSym6: per bank, Z = Minv_sym * R */
#pragma omp parallel for schedule(static)
    for(int b = 0; b < total / bank; ++b)
    {
        const float* inv = m->inverse + (size_t)b * bsz;
        for(int lr = 0; lr < bank; ++lr)
            for(int lc = 0; lc < bank; ++lc)
            {
                const int    vr = b * bank + lr, vc = b * bank + lc;
                const float* B  = inv + 9 * (lc >= lr ? packed_index(bank, lr, lc) : packed_index(bank, lc, lr));
                for(int i = 0; i < 3; ++i)
                {
                    double s = 0.0;
                    for(int j = 0; j < 3; ++j)
                        s += (double)(lc >= lr ? B[i + 3 * j] : B[j + 3 * i]) * R[3 * vc + j];
                    Z[3 * vr + i] += s;
                }
            }
    }
    /* __collectFinalZ_new */
#pragma omp parallel for schedule(static)
    for(int v = 0; v < m->n_fem; ++v)
    {
        const int s = m->real_to_part[v];
        for(int d = 0; d < 3; ++d)
        {
            double acc = Z[3 * s + d];
            for(int l = 1; l < m->levels; ++l) acc += Z[3 * m->coarse_table[6 * v + l - 1] + d];
            z[3 * v + d] = acc;
        }
    }
    free(R);
    free(Z);
}

/* ------------------------------------------------------------------------ */
/* Block-Jacobi pieces                                                      */
/* ------------------------------------------------------------------------ */

int orc_inverse_gj_pivot(int n, const double* a_in, double* inv)
{
    double* a = (double*)malloc(sizeof(double) * n * n);
    memcpy(a, a_in, sizeof(double) * n * n);
    for(int i = 0; i < n * n; ++i) inv[i] = 0.0;
    for(int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for(int col = 0; col < n; ++col)
    {
        int    pivot = col;
        double maxv  = fabs(a[col * n + col]);
        for(int row = col + 1; row < n; ++row)
        {
            const double v = fabs(a[row * n + col]);
            if(v > maxv)
            {
                maxv  = v;
                pivot = row;
            }
        }
        if(pivot != col)
            for(int k = 0; k < n; ++k)
            {
                double t = a[col * n + k]; a[col * n + k] = a[pivot * n + k]; a[pivot * n + k] = t;
                t = inv[col * n + k]; inv[col * n + k] = inv[pivot * n + k]; inv[pivot * n + k] = t;
            }
        const double diag = a[col * n + col];
        for(int k = 0; k < n; ++k)
        {
            a[col * n + k] /= diag;
            inv[col * n + k] /= diag;
        }
        for(int row = 0; row < n; ++row)
        {
            if(row == col) continue;
            const double f = a[row * n + col];
            for(int k = 0; k < n; ++k)
            {
                a[row * n + k] -= f * a[col * n + k];
                inv[row * n + k] -= f * inv[col * n + k];
            }
        }
    }
    free(a);
    return 0;
}

void orc_abd_block_inverse(int nbody, int nb, const double* vals, const int* rows, const int* cols,
                           double* inv)
{
    double* P = (double*)calloc((size_t)nbody * 144, sizeof(double));
    for(int k = 0; k < nb; ++k)
    {
        const int row = rows[k], col = cols[k];
        if(row / 4 != col / 4 || row / 4 >= nbody) continue;
        double*       Pb = P + (size_t)(row / 4) * 144;
        const double* H  = vals + 9 * (size_t)k;
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 3; ++j)
            {
                Pb[((row % 4) * 3 + i) * 12 + (col % 4) * 3 + j] = H[i + 3 * j];
                if(row != col) Pb[((col % 4) * 3 + j) * 12 + (row % 4) * 3 + i] = H[i + 3 * j];
            }
    }
    for(int b = 0; b < nbody; ++b) orc_inverse_gj_pivot(12, P + (size_t)b * 144, inv + (size_t)b * 144);
    free(P);
}

void orc_diag3_inverse(int block_rows, int nb, const double* vals, const int* rows, const int* cols,
                       double* inv)
{
    for(int i = 0; i < block_rows; ++i)
        for(int e = 0; e < 9; ++e) inv[9 * i + e] = (e % 4 == 0) ? 1.0 : 0.0;
    for(int k = 0; k < nb; ++k)
    {
        if(rows[k] != cols[k]) continue;
        double a[9];
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 3; ++j) a[3 * i + j] = vals[9 * (size_t)k + i + 3 * j];
        orc_inverse_gj_pivot(3, a, inv + 9 * (size_t)rows[k]);
    }
}

/* ------------------------------------------------------------------------ */
/* PCG                                                                      */
/* ------------------------------------------------------------------------ */

/* Sum of a[i] b[i]: partial sums over fixed chunks of 4,096 entries (in
 * parallel), then the partials in chunk order. */
static double dot(const double* a, const double* b, int n)
{
    enum { CH = 4096 };
    const int nc = (n + CH - 1) / CH;
    double    stackbuf[64];
    double*   part = nc <= 64 ? stackbuf : (double*)malloc(sizeof(double) * (size_t)nc);
#pragma omp parallel for schedule(static) if(nc > 1)
    for(int c = 0; c < nc; ++c)
    {
        double    s  = 0.0;
        const int e1 = (c + 1) * CH < n ? (c + 1) * CH : n;
        for(int i = c * CH; i < e1; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for(int c = 0; c < nc; ++c) s += part[c];
    if(part != stackbuf) free(part);
    return s;
}

/* The symmetric upper-BCOO product of orc_sym_bcoo_spmv (a = 1, b = 0) as a
 * row-parallel loop: row i's contributions (B x_j of its upper blocks, B^T
 * x_k of the blocks (k, i) above it) are kept in increasing block index, the
 * order in which the serial loop adds them, so y is bitwise the same. */
typedef struct
{
    int  n_rows;
    int* ptr;  /* n_rows + 1 */
    int* ent;  /* block index << 1 | transposed */
} row_spmv;

static void row_spmv_build(row_spmv* R, int n_rows, int nb, const int* rows, const int* cols)
{
    R->n_rows = n_rows;
    R->ptr    = (int*)calloc((size_t)n_rows + 1, sizeof(int));
    for(int k = 0; k < nb; ++k)
    {
        R->ptr[rows[k] + 1]++;
        if(rows[k] != cols[k]) R->ptr[cols[k] + 1]++;
    }
    for(int i = 0; i < n_rows; ++i) R->ptr[i + 1] += R->ptr[i];
    R->ent   = (int*)malloc(sizeof(int) * ((size_t)R->ptr[n_rows] + 1));
    int* pos = (int*)malloc(sizeof(int) * ((size_t)n_rows + 1));
    memcpy(pos, R->ptr, sizeof(int) * ((size_t)n_rows + 1));
    for(int k = 0; k < nb; ++k)
    {
        R->ent[pos[rows[k]]++] = k << 1;
        if(rows[k] != cols[k]) R->ent[pos[cols[k]]++] = (k << 1) | 1;
    }
    free(pos);
}

static void row_spmv_free(row_spmv* R)
{
    free(R->ptr);
    free(R->ent);
}

static void row_spmv_apply(const row_spmv* R, const double* vals, const int* rows, const int* cols,
                           const double* x, double* y)
{
#pragma omp parallel for schedule(static, 256)
    for(int i = 0; i < R->n_rows; ++i)
    {
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        for(int e = R->ptr[i]; e < R->ptr[i + 1]; ++e)
        {
            const int     k = R->ent[e] >> 1;
            const double* B = vals + 9 * (size_t)k;
            double        s[3];
            if(R->ent[e] & 1)
            {
                const int c = rows[k];
                for(int r = 0; r < 3; ++r)
                    s[r] = blk(B, 0, r) * x[3 * c] + blk(B, 1, r) * x[3 * c + 1] + blk(B, 2, r) * x[3 * c + 2];
            }
            else
            {
                const int c = cols[k];
                for(int r = 0; r < 3; ++r)
                    s[r] = blk(B, r, 0) * x[3 * c] + blk(B, r, 1) * x[3 * c + 1] + blk(B, r, 2) * x[3 * c + 2];
            }
            y0 += s[0];
            y1 += s[1];
            y2 += s[2];
        }
        y[3 * i] = y0;
        y[3 * i + 1] = y1;
        y[3 * i + 2] = y2;
    }
}

typedef struct
{
    int            block_rows, global_diag, n_abd, fem_offset;
    const double*  diag_inv;
    const double*  abd_inv;
    const orc_mas* mas;
} precond_ctx;

/* GlobalLinearSystem::apply_preconditioner (global_linear_system.cu:293-306) */
static void apply_precond(const precond_ctx* p, const double* r, double* z)
{
    const int dof = 3 * p->block_rows;
    if(p->global_diag)
    {
#pragma omp parallel for schedule(static)
        for(int i = 0; i < p->block_rows; ++i)
            for(int a = 0; a < 3; ++a)
            {
                double s = 0.0;
                for(int b = 0; b < 3; ++b) s += p->diag_inv[9 * i + 3 * a + b] * r[3 * i + b];
                z[3 * i + a] = s;
            }
    }
    else
        memcpy(z, r, sizeof(double) * dof);
    for(int body = 0; body < p->n_abd; ++body)
    {
        const double* P = p->abd_inv + 144 * (size_t)body;
        for(int a = 0; a < 12; ++a)
        {
            double s = 0.0;
            for(int b = 0; b < 12; ++b) s += P[12 * a + b] * r[12 * body + b];
            z[12 * body + a] = s;
        }
    }
    if(p->mas) orc_mas_apply(p->mas, r + 3 * p->fem_offset, z + 3 * p->fem_offset);
}

int orc_pcg(int block_rows, int nb, const double* vals, const int* rows, const int* cols,
            const double* b, double* x, int global_diag, int n_abd, const double* abd_inv,
            const orc_mas* mas, int fem_offset, double tol, int rule, int max_iter,
            orc_pcg_stats* st)
{
    const int dof = 3 * block_rows;
    precond_ctx pc;
    pc.block_rows  = block_rows;
    pc.global_diag = global_diag;
    pc.n_abd       = n_abd;
    pc.fem_offset  = fem_offset;
    pc.abd_inv     = abd_inv;
    pc.mas         = mas;
    double* diag   = NULL;
    if(global_diag)
    {
        diag = (double*)malloc(sizeof(double) * 9 * (size_t)block_rows);
        orc_diag3_inverse(block_rows, nb, vals, rows, cols, diag);
    }
    pc.diag_inv = diag;
    double* r  = (double*)malloc(sizeof(double) * dof);
    double* z  = (double*)malloc(sizeof(double) * dof);
    double* p  = (double*)malloc(sizeof(double) * dof);
    double* Ap = (double*)malloc(sizeof(double) * dof);
    row_spmv A;
    row_spmv_build(&A, block_rows, nb, rows, cols);
    if(max_iter <= 0) max_iter = (int)(0.3 * (double)dof);
    memset(x, 0, sizeof(double) * dof);
    memcpy(r, b, sizeof(double) * dof);
    orc_pcg_stats s;
    memset(&s, 0, sizeof(s));
    int k = 0;
    if(rule == ORC_RULE_REF_RTZ)
    {
        /* pcg_solver.cu:192-326 */
        apply_precond(&pc, r, z);
        double rz = dot(r, z, dof), rz0 = rz;
        memcpy(p, z, sizeof(double) * dof);
        s.rz0 = rz0;
        for(k = 1; k < max_iter; ++k)
        {
            row_spmv_apply(&A, vals, rows, cols, p, Ap);
            const double pAp   = dot(p, Ap, dof);
            const double alpha = rz / pAp;
#pragma omp parallel for schedule(static)
            for(int i = 0; i < dof; ++i)
            {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            if(fabs(rz) <= tol * rz0) break;
            apply_precond(&pc, r, z);
            const double rz_new = dot(r, z, dof);
            const double beta   = rz_new / rz;
#pragma omp parallel for schedule(static)
            for(int i = 0; i < dof; ++i) p[i] = z[i] + beta * p[i];
            rz = rz_new;
        }
        s.rz_last = rz;
    }
    else
    {
        /* agipc_coarse_replay.cu:120-160 */
        apply_precond(&pc, r, z);
        memcpy(p, z, sizeof(double) * dof);
        const double initial2  = dot(b, b, dof);
        double       residual2 = initial2, rz = dot(r, z, dof);
        const double target    = tol * tol * initial2;
        s.rz0 = rz;
        while(k < max_iter && residual2 > target)
        {
            if(!isfinite(rz) || rz <= 0) { s.failure = 1; break; }
            row_spmv_apply(&A, vals, rows, cols, p, Ap);
            const double curv = dot(p, Ap, dof);
            if(!isfinite(curv) || curv <= 0) { s.failure = 2; break; }
            const double alpha = rz / curv;
#pragma omp parallel for schedule(static)
            for(int i = 0; i < dof; ++i)
            {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            ++k;
            residual2 = dot(r, r, dof);
            if(!isfinite(residual2)) { s.failure = 3; break; }
            if(residual2 <= target) break;
            apply_precond(&pc, r, z);
            const double next_rz = dot(r, z, dof);
            if(!isfinite(next_rz) || next_rz <= 0 || fabs(rz) < 1e-30) { s.failure = 1; break; }
            const double beta = next_rz / rz;
#pragma omp parallel for schedule(static)
            for(int i = 0; i < dof; ++i) p[i] = z[i] + beta * p[i];
            rz = next_rz;
        }
        if(!s.failure && residual2 > target) s.failure = 4;
        s.rz_last = rz;
    }
    /* residual report */
    const double bb = dot(b, b, dof);
    s.rec_rel_res   = bb > 0 ? sqrt(dot(r, r, dof) / bb) : 0.0;
    row_spmv_apply(&A, vals, rows, cols, x, Ap);
    for(int i = 0; i < dof; ++i) Ap[i] = b[i] - Ap[i];
    const double t2 = dot(Ap, Ap, dof);
    s.true_rel_res = bb > 0 ? sqrt(t2 / bb) : sqrt(t2);
    s.iterations   = k;
    if(st) *st = s;
    row_spmv_free(&A);
    free(r); free(z); free(p); free(Ap); free(diag);
    return k;
}
