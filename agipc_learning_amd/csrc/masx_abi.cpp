// C ABI of the masx library (include/masx.h).
#include "masx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "converter.hpp"
#include "fem.hpp"
#include "hierarchy.hpp"
#include "morton.hpp"
#include "bandwidth.hpp"
#include "solver.hpp"

namespace
{
thread_local std::string g_error;

template <typename F>
int guard(F&& f)
{
    try
    {
        f();
        return 0;
    }
    catch(const std::exception& e)
    {
        g_error = e.what();
        return 1;
    }
    catch(...)
    {
        g_error = "unknown error";
        return 1;
    }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

void require(bool ok, const char* msg)
{
    if(!ok) throw std::invalid_argument(msg);
}
}  // namespace

struct masx_solver
{
    masx::Solver impl;
    // Morton partition on the device for masx_set_topology_morton
    std::unique_ptr<masx::DeviceMorton> morton;
    masx::DeviceArray<double>           mt_pos;
    masx::DeviceArray<int>              mt_p2r, mt_r2p;
};

namespace
{
// meshes this large sort on the device when one is present (the same arrays
// as the host sort, bit for bit; ~20x faster at 69k vertices)
constexpr int kDeviceMortonMin = 4096;

bool device_present()
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

void hip_ok(hipError_t e)
{
    if(e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e));
}
}  // namespace

struct masx_mas
{
    int bank = 32;
    int vert_num = 0, node_offset = 0, neighbor_total = 0, part_map_size = 0, levels = 0;
    std::vector<uint32_t> nbr_list, nbr_start, nbr_num;
    std::vector<int>      part_to_real, real_to_part;
    bool uploaded = false, matrix_ready = false;
    // structure of the last set_preconditioner_bcoo call
    std::vector<int>      rows, cols, pairs;
    std::unique_ptr<masx_solver> solver;
    uint32_t* d_sel = nullptr;  // selected block indices (device)
    size_t    d_sel_n = 0;
    ~masx_mas()
    {
        if(d_sel) (void)hipFree(d_sel);
    }
};

extern "C" {

const char* masx_last_error(void) { return g_error.c_str(); }

int masx_device_info(char* name, int name_len, int* cu_count, long long* hbm_bytes)
{
    return guard([&] {
        int dev = 0;
        if(hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device");
        hipDeviceProp_t p;
        if(hipGetDeviceProperties(&p, dev) != hipSuccess) throw std::runtime_error("hipGetDeviceProperties failed");
        if(name && name_len > 0)
        {
            std::strncpy(name, p.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if(cu_count) *cu_count = p.multiProcessorCount;
        if(hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
    });
}

void masx_default_solve_params(masx_solve_params* p)
{
    masx::SolveParams d;
    p->rule        = d.rule;
    p->tol         = d.tol;
    p->max_iter    = d.max_iter;
    p->check_every = d.check_every;
    p->chunk       = d.chunk;
    p->use_graph   = d.use_graph;
}

int masx_create(masx_solver** out)
{
    return guard([&] {
        require(out != nullptr, "null output pointer");
        *out = new masx_solver();
    });
}

int masx_destroy(masx_solver* s)
{
    return guard([&] { delete s; });
}

int masx_add_system(masx_solver* s, const masx_system_desc* desc, const int* rows, const int* cols, int* sys)
{
    return guard([&] {
        require(s && desc, "null argument");
        require(desc->nb >= 0 && desc->block_rows > 0, "empty system");
        require(desc->nb == 0 || (rows && cols), "null rows/cols");
        masx::SystemDesc d;
        d.block_rows  = desc->block_rows;
        d.nb          = desc->nb;
        d.n_abd       = desc->n_abd;
        d.fem_offset  = 4 * desc->n_abd;
        d.n_fem       = desc->n_fem;
        d.use_mas     = desc->use_mas != 0;
        d.global_diag = desc->global_diag != 0;
        d.bank        = desc->bank > 0 ? desc->bank : 32;
        d.levels      = desc->levels;
        require(d.bank == 32 || d.bank == 16, "bank size must be 32 or 16");
        d.rows.assign(rows, rows + desc->nb);
        d.cols.assign(cols, cols + desc->nb);
        const int idx = s->impl.add_system(d);
        if(sys) *sys = idx;
    });
}

int masx_set_structure(masx_solver* s, int sys, long long nb, const int* rows, const int* cols)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        require(nb >= 0 && (nb == 0 || (rows && cols)), "null rows/cols");
        auto& d = s->impl.system(sys);
        d.nb = nb;
        d.rows.assign(rows, rows + nb);
        d.cols.assign(cols, cols + nb);
        d.rows_on_device = false;
        s->impl.structure_changed(sys);
    });
}

int masx_set_structure_device(masx_solver* s, int sys, long long nb, const int* d_rows, const int* d_cols,
                              void* stream)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        require(nb >= 0 && (nb == 0 || (d_rows && d_cols)), "null rows/cols");
        s->impl.set_structure_device(sys, nb, d_rows, d_cols, as_stream(stream));
    });
}

int masx_set_topology(masx_solver* s, int sys, const uint32_t* nbr_start, const uint32_t* nbr_num,
                      const uint32_t* nbr_list, int list_len, const int* part_to_real, int part_map_size,
                      const int* real_to_part)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        auto& d = s->impl.system(sys);
        require(part_to_real && real_to_part && part_map_size >= d.n_fem, "bad partition");
        require(part_map_size % d.bank == 0, "partition size must be a multiple of the bank size");
        if(d.n_fem > 0) require(nbr_start && nbr_num, "null neighbor arrays");
        d.nbr_start.assign(nbr_start, nbr_start + d.n_fem);
        d.nbr_num.assign(nbr_num, nbr_num + d.n_fem);
        d.nbr_list.assign(nbr_list, nbr_list + list_len);
        for(int v = 0; v < d.n_fem; ++v)
            require((long long)nbr_start[v] + nbr_num[v] <= list_len, "neighbor range outside the list");
        d.part_to_real.assign(part_to_real, part_to_real + part_map_size);
        d.real_to_part.assign(real_to_part, real_to_part + d.n_fem);
        for(int v = 0; v < d.n_fem; ++v)
        {
            const int slot = real_to_part[v];
            require(slot >= 0 && slot < part_map_size && part_to_real[slot] == v, "inconsistent partition");
        }
        ++d.topo_version;
        s->impl.structure_changed(sys);
    });
}

int masx_set_topology_morton(masx_solver* s, int sys, const uint32_t* nbr_start, const uint32_t* nbr_num,
                             const uint32_t* nbr_list, int list_len, const double* positions)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        auto& d = s->impl.system(sys);
        require(positions != nullptr || d.n_fem == 0, "positions required");
        std::vector<int> p2r, r2p;
        if(d.n_fem >= kDeviceMortonMin && device_present())
        {
            const size_t n = (size_t)d.n_fem, padded = (n + d.bank - 1) / d.bank * d.bank;
            if(!s->morton) s->morton = std::make_unique<masx::DeviceMorton>();
            s->mt_pos.alloc(3 * n);
            s->mt_p2r.alloc(padded);
            s->mt_r2p.alloc(n);
            hip_ok(hipMemcpy(s->mt_pos.ptr, positions, 3 * n * sizeof(double), hipMemcpyHostToDevice));
            s->morton->partition((int)n, s->mt_pos.ptr, d.bank, s->mt_p2r.ptr, (long long)padded, s->mt_r2p.ptr,
                                 nullptr);
            p2r.resize(padded);
            r2p.resize(n);
            hip_ok(hipMemcpy(p2r.data(), s->mt_p2r.ptr, padded * sizeof(int), hipMemcpyDeviceToHost));
            hip_ok(hipMemcpy(r2p.data(), s->mt_r2p.ptr, n * sizeof(int), hipMemcpyDeviceToHost));
        }
        else
            masx::morton_partition(d.n_fem, positions, d.bank, p2r, r2p);
        if(masx_set_topology(s, sys, nbr_start, nbr_num, nbr_list, list_len, p2r.data(), (int)p2r.size(),
                             r2p.data()))
            throw std::runtime_error(g_error);
    });
}

int masx_morton_partition(int n, const double* positions, int bank, int* part_to_real, int part_cap,
                          int* real_to_part)
{
    return guard([&] {
        require(n >= 1 && positions && part_to_real && real_to_part, "bad argument");
        require(bank == 32 || bank == 16, "bank size must be 32 or 16");
        std::vector<int> p2r, r2p;
        masx::morton_partition(n, positions, bank, p2r, r2p);
        require(part_cap >= (int)p2r.size(), "part_to_real buffer too small");
        std::memcpy(part_to_real, p2r.data(), sizeof(int) * p2r.size());
        std::memcpy(real_to_part, r2p.data(), sizeof(int) * r2p.size());
    });
}

int masx_set_contacts(masx_solver* s, int sys, const int* pairs, int n_pairs, int node_offset)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        require(n_pairs == 0 || pairs, "null pairs");
        auto& d = s->impl.system(sys);
        d.pairs.assign(pairs, pairs + 4 * (size_t)n_pairs);
        d.node_offset = node_offset;
        s->impl.structure_changed(sys);
    });
}

int masx_analyze(masx_solver* s)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.analyze();
    });
}

int masx_setup(masx_solver* s)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.setup();
    });
}

int masx_set_values(masx_solver* s, int sys, const double* d_values, void* stream)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        s->impl.set_values(sys, d_values, as_stream(stream));
    });
}

int masx_set_values_many(masx_solver* s, int n, const int* systems, const double* const* d_values, void* stream)
{
    return guard([&] {
        require(s && n >= 0 && (n == 0 || (systems && d_values)), "null argument");
        for(int q = 0; q < n; ++q) require(d_values[q] != nullptr, "null values");
        s->impl.set_values_many(n, systems, d_values, as_stream(stream));
    });
}

int masx_set_values_host(masx_solver* s, int sys, const double* h_values, void* stream)
{
    return guard([&] {
        require(s && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        s->impl.set_values_host(sys, h_values, as_stream(stream));
    });
}

int masx_build(masx_solver* s, int precision, void* stream)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.build(precision, as_stream(stream));
    });
}

int masx_solve(masx_solver* s, const double* d_b, double* d_x, const masx_solve_params* params,
               masx_solve_stats* stats, void* stream)
{
    return guard([&] {
        require(s && d_b && d_x, "null argument");
        masx::SolveParams p;
        if(params)
        {
            p.rule        = params->rule;
            p.tol         = params->tol;
            p.max_iter    = params->max_iter;
            p.check_every = params->check_every;
            p.chunk       = params->chunk;
            p.use_graph   = params->use_graph;
        }
        require(p.rule == MASX_RULE_REF_RTZ || p.rule == MASX_RULE_TRUE_REL, "unknown stopping rule");
        require(p.tol >= 0.0, "negative tolerance");
        std::vector<masx::SolveStats> st;
        s->impl.solve(d_b, d_x, p, stats ? &st : nullptr, as_stream(stream));
        if(stats)
            for(size_t i = 0; i < st.size(); ++i)
            {
                stats[i].iterations        = st[i].iterations;
                stats[i].status            = st[i].status == masx::kDone ? MASX_DONE : MASX_RUNNING;
                stats[i].failure           = st[i].failure;
                stats[i].restarts          = st[i].restarts;
                stats[i].true_checks       = st[i].true_checks;
                stats[i].rz0               = st[i].rz0;
                stats[i].rz                = st[i].rz;
                stats[i].rel_res_recursive = st[i].rel_res_recursive;
                stats[i].rel_res_true      = st[i].rel_res_true;
                stats[i].rejected          = st[i].rejected;
            }
    });
}

int masx_spmv(masx_solver* s, double a, const double* d_x, double b, double* d_y, void* stream)
{
    return guard([&] {
        require(s && d_x && d_y, "null argument");
        s->impl.spmv(a, d_x, b, d_y, as_stream(stream));
    });
}

int masx_apply_preconditioner(masx_solver* s, const double* d_r, double* d_z, void* stream)
{
    return guard([&] {
        require(s && d_r && d_z, "null argument");
        s->impl.apply_preconditioner(d_r, d_z, as_stream(stream));
    });
}

int masx_set_spmv_tiling(masx_solver* s, int entries_per_thread)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.set_spmv_tiling(entries_per_thread);
    });
}

int masx_autotune(masx_solver* s, int reps, int* entries_per_thread, double* ms)
{
    return guard([&] {
        require(s != nullptr && reps > 0, "bad argument");
        const auto t = s->impl.autotune_spmv(reps);
        if(entries_per_thread) *entries_per_thread = t.entries;
        if(ms) *ms = t.ms;
    });
}

int masx_set_factor_waves(masx_solver* s, int waves)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.set_factor_waves(waves);
    });
}

int masx_set_spmv_loads(masx_solver* s, int mode)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.set_spmv_loads(mode);
    });
}

int masx_spmv_loads_used(masx_solver* s, int* nontemporal)
{
    return guard([&] {
        require(s != nullptr && nontemporal != nullptr, "null argument");
        *nontemporal = s->impl.spmv_nontemporal() ? 1 : 0;
    });
}

int masx_set_mas_waves(masx_solver* s, int waves)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.set_mas_waves(waves);
    });
}

int masx_autotune_mas(masx_solver* s, int reps, int* waves, double* ms)
{
    return guard([&] {
        require(s != nullptr && reps > 0, "bad argument");
        const auto t = s->impl.autotune_mas(reps);
        if(waves) *waves = t.entries;
        if(ms) *ms = t.ms;
    });
}

int masx_time_kernel(masx_solver* s, int which, int reps, double* ms)
{
    return guard([&] {
        require(s && ms && reps > 0 && reps <= 65536, "bad argument (reps in 1..65536)");
        *ms = s->impl.time_kernel(which, reps);
    });
}

int masx_set_device_structure(masx_solver* s, int on)
{
    return guard([&] {
        require(s != nullptr, "null solver");
        s->impl.set_device_structure(on);
    });
}

int masx_device_structure_used(masx_solver* s, int* used)
{
    return guard([&] {
        require(s && used, "null argument");
        *used = s->impl.device_structure_used() ? 1 : 0;
    });
}

int masx_plan_array(masx_solver* s, int which, int* out, long long capacity, long long* n)
{
    return guard([&] {
        require(s && n && s->impl.is_setup(), "solver not set up");
        std::vector<int> v;
        s->impl.download_plan_array(which, v);
        *n = (long long)v.size();
        if(out)
        {
            require((long long)v.size() <= capacity, "output buffer too small");
            std::memcpy(out, v.data(), v.size() * sizeof(int));
        }
    });
}

int masx_plan_info_get(masx_solver* s, masx_plan_info* info)
{
    return guard([&] {
        require(s && info && s->impl.is_analyzed(), "solver not analyzed");
        const auto& P       = s->impl.plan();
        info->n_rows        = P.n_rows;
        info->n_ext         = P.n_ext;
        info->n_sys         = P.n_sys;
        info->bank          = P.bank;
        info->nnz           = P.nnz;
        info->nb_total      = P.nb_total;
        info->n_banks       = P.n_banks;
        info->n_fine_banks  = P.n_fine_banks;
        info->n_coarse      = P.n_coarse;
        info->n_bodies      = P.n_bodies;
        info->n_gal         = P.n_gal;
        info->inverse_bytes = s->impl.is_setup() ? (long long)s->impl.inverse_bytes() : 0;
    });
}

int masx_get_hierarchy(masx_solver* s, int sys, int* levels, int* total, int* level_count, int* level_begin,
                       uint32_t* fine_mask, int* prefix, int* going_next, int* coarse_table)
{
    return guard([&] {
        require(s && s->impl.is_analyzed() && sys >= 0 && sys < s->impl.n_systems(), "bad system index");
        require(s->impl.system(sys).use_mas, "system has no MAS hierarchy");
        const auto& h = s->impl.hierarchy(sys);
        if(levels) *levels = h.levels;
        if(total) *total = h.total;
        if(level_count)
        {
            level_count[0] = 0;
            for(int l = 1; l <= h.levels; ++l) level_count[l] = h.level_count[l];
        }
        if(level_begin) std::memcpy(level_begin, h.level_begin.data(), sizeof(int) * (h.levels + 1));
        if(fine_mask) std::memcpy(fine_mask, h.fine_mask.data(), sizeof(uint32_t) * h.n_fem);
        if(prefix) std::memcpy(prefix, h.prefix.data(), sizeof(int) * h.prefix.size());
        if(going_next) std::memcpy(going_next, h.going_next.data(), sizeof(int) * h.total);
        if(coarse_table) std::memcpy(coarse_table, h.coarse_table.data(), sizeof(int) * 6 * (size_t)h.n_fem);
    });
}

int masx_get_inverse(masx_solver* s, float* out, long long capacity)
{
    return guard([&] {
        require(s && out, "null argument");
        std::vector<float> v;
        s->impl.download_inverse(v);
        require((long long)v.size() <= capacity, "output buffer too small");
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    });
}

int masx_bank_of_slot(masx_solver* s, int sys, int slot, int* bank)
{
    return guard([&] {
        require(s && bank && s->impl.is_setup(), "bad argument");
        const auto& P = s->impl.plan();
        const auto& d = s->impl.system(sys);
        require(d.use_mas, "system has no MAS hierarchy");
        const auto& h = d.hier;
        require(slot >= 0 && slot < h.total, "slot out of range");
        if(slot < h.n_slots)
        {
            int fine_base = 0;
            for(int t = 0; t < sys; ++t)
                if(s->impl.system(t).use_mas) fine_base += s->impl.system(t).hier.n_slots / P.bank;
            *bank = fine_base + slot / P.bank;
        }
        else
            *bank = P.n_fine_banks + (P.coarse_base[sys] + slot - h.n_slots) / P.bank;
    });
}

int masx_factor_status(masx_solver* s, long long* count)
{
    return guard([&] {
        require(s && count, "null argument");
        *count = s->impl.factor_status();
    });
}

int masx_hierarchy_host(int n_fem, int n_slots, int bank, int levels, const uint32_t* nbr_start,
                        const uint32_t* nbr_num, const uint32_t* nbr_list, const int* part_to_real,
                        const int* real_to_part, const int* pairs, int n_pairs, int node_offset, int* out_levels,
                        int* out_total, int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix,
                        int* going_next, long long going_next_cap, int* coarse_table)
{
    return guard([&] {
        require(n_fem >= 1 && n_slots >= n_fem && part_to_real && real_to_part, "bad argument");
        masx::Hierarchy h;
        masx::build_hierarchy(h, n_fem, n_slots, bank, levels, nbr_start, nbr_num, nbr_list, part_to_real,
                              real_to_part, pairs, n_pairs, node_offset);
        if(out_levels) *out_levels = h.levels;
        if(out_total) *out_total = h.total;
        if(!going_next) return;
        require(going_next_cap >= (long long)std::max(h.total, n_fem), "going_next buffer too small");
        if(level_count)
        {
            level_count[0] = 0;
            for(int l = 1; l <= h.levels; ++l) level_count[l] = h.level_count[l];
        }
        if(level_begin) std::memcpy(level_begin, h.level_begin.data(), sizeof(int) * (h.levels + 1));
        if(fine_mask) std::memcpy(fine_mask, h.fine_mask.data(), sizeof(uint32_t) * n_fem);
        if(prefix) std::memcpy(prefix, h.prefix.data(), sizeof(int) * h.prefix.size());
        std::memcpy(going_next, h.going_next.data(), sizeof(int) * std::max(h.total, n_fem));
        if(coarse_table) std::memcpy(coarse_table, h.coarse_table.data(), sizeof(int) * 6 * (size_t)n_fem);
    });
}

struct masx_morton
{
    masx::DeviceMorton impl;
};

int masx_morton_create(masx_morton** out)
{
    return guard([&] {
        require(out != nullptr, "bad argument");
        *out = new masx_morton();
    });
}

int masx_morton_destroy(masx_morton* m)
{
    delete m;
    return 0;
}

int masx_morton_partition_device(masx_morton* m, int n, const double* d_positions, int bank, int* d_part_to_real,
                                 long long part_cap, int* d_real_to_part, void* stream)
{
    return guard([&] {
        require(m != nullptr, "bad argument");
        m->impl.partition(n, d_positions, bank, d_part_to_real, part_cap, d_real_to_part, (hipStream_t)stream);
    });
}

int masx_stream_rate(long long bytes, int kind, int reps, double* gb_per_s)
{
    return guard([&] {
        require(gb_per_s != nullptr, "bad argument");
        *gb_per_s = masx::stream_rate(bytes, kind, reps);
    });
}

int masx_morton_partition_device_noalloc(int n, const double* d_positions, int bank, int* d_part_to_real,
                                         long long part_cap, int* d_real_to_part, void* stream)
{
    return guard([&] {
        masx::morton_partition_noalloc(n, d_positions, bank, d_part_to_real, part_cap, d_real_to_part,
                                       (hipStream_t)stream);
    });
}

int masx_spmv_layout_check_bodies(int n_rows, int n_bodies, long long nb, const int* rows, const int* cols,
                                  int entries_per_thread, unsigned seed, double* max_rel_err, long long* n_entries,
                                  long long* n_in_tile)
{
    return guard([&] {
        require(n_rows >= 1 && nb >= 1 && rows && cols && max_rel_err, "bad argument");
        require(n_bodies >= 0 && 4LL * n_bodies <= n_rows, "bodies must fit the rows");
        require(entries_per_thread >= 1 && 256LL * entries_per_thread <= masx::kSpmvMaxCap, "bad entries per thread");
        // full CSR: upper (r, c) -> row r as is, row c transposed
        std::vector<std::array<int, 3>> ent;
        ent.reserve(2 * (size_t)nb);
        for(long long k = 0; k < nb; ++k)
        {
            const int r = rows[k], c = cols[k];
            require(r >= 0 && c < n_rows && r <= c, "upper structure expected (0 <= rows <= cols < n_rows)");
            ent.push_back({r, c, (int)k});
            if(r != c) ent.push_back({c, r, (int)(k | 0x80000000LL)});
        }
        std::sort(ent.begin(), ent.end(),
                  [](const auto& a, const auto& b) { return a[0] != b[0] ? a[0] < b[0] : a[1] < b[1]; });
        for(size_t i = 1; i < ent.size(); ++i)
            require(ent[i][0] != ent[i - 1][0] || ent[i][1] != ent[i - 1][1], "duplicate block");
        std::vector<int> row_ptr(n_rows + 1, 0), col(ent.size()), src(ent.size());
        for(size_t i = 0; i < ent.size(); ++i)
        {
            ++row_ptr[ent[i][0] + 1];
            col[i] = ent[i][1];
            src[i] = ent[i][2];
        }
        for(int r = 0; r < n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
        masx::SpmvLayout L;
        masx::SpmvBodies bodies;  // bodies occupy rows [0, 4 n_bodies)
        bodies.abd_row0  = {0};
        bodies.n_abd     = {n_bodies};
        bodies.body_base = {0};
        std::vector<int> body_row(n_bodies);
        for(int b = 0; b < n_bodies; ++b) body_row[b] = 4 * b;
        masx::build_spmv_layout(L, 1, {0, n_rows}, row_ptr, col, src, 256 * entries_per_thread, &bodies);
        // seeded values (64-bit LCG, uniform in [-1, 1))
        uint64_t z = 0x9e3779b97f4a7c15ULL ^ seed;
        auto uni = [&z]() {
            z = z * 6364136223846793005ULL + 1442695040888963407ULL;
            return (double)(z >> 11) * (1.0 / 4503599627370496.0) - 1.0;
        };
        std::vector<double> upper(9 * (size_t)nb), x(3 * (size_t)n_rows), y(3 * (size_t)n_rows),
            yf(3 * (size_t)n_rows, 0.0);
        for(auto& v : upper) v = uni();
        for(auto& v : x) v = uni();
        masx::spmv_layout_apply(L, upper.data(), x.data(), y.data(), row_ptr.data(), col.data(), src.data(),
                                body_row.data());
        double scale = 0.0, err = 0.0;
        for(int r = 0; r < n_rows; ++r)
            for(int e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
            {
                const double* U = upper.data() + 9 * (long long)(src[e] & 0x7fffffff);
                const double* xc = x.data() + 3 * (long long)col[e];
                for(int i = 0; i < 3; ++i)
                    for(int j = 0; j < 3; ++j) yf[3 * r + i] += (src[e] >= 0 ? U[i + 3 * j] : U[j + 3 * i]) * xc[j];
            }
        for(size_t i = 0; i < y.size(); ++i)
        {
            scale = std::max(scale, std::fabs(yf[i]));
            err   = std::max(err, std::fabs(y[i] - yf[i]));
        }
        *max_rel_err = scale > 0.0 ? err / scale : err;
        if(n_entries) *n_entries = L.n_entries;
        if(n_in_tile) *n_in_tile = L.n_in_tile;
    });
}

int masx_spmv_layout_check(int n_rows, long long nb, const int* rows, const int* cols, int entries_per_thread,
                           unsigned seed, double* max_rel_err, long long* n_entries, long long* n_in_tile)
{
    return masx_spmv_layout_check_bodies(n_rows, 0, nb, rows, cols, entries_per_thread, seed, max_rel_err, n_entries,
                                         n_in_tile);
}

// ------------------------------------------------------------------ MAS ----

int masx_mas_create(masx_mas** out, int bank)
{
    return guard([&] {
        require(out != nullptr, "null output pointer");
        require(bank == 32 || bank == 16, "bank size must be 32 or 16");
        auto* m = new masx_mas();
        m->bank = bank;
        *out    = m;
    });
}

int masx_mas_destroy(masx_mas* m)
{
    return guard([&] { delete m; });
}

int masx_mas_init_neighbor(masx_mas* m, int vert_num, int collision_node_offset, int total_neighbor_num,
                           int part_map_size)
{
    return guard([&] {
        require(m && vert_num >= 1 && total_neighbor_num >= 0, "bad argument");
        require(part_map_size >= vert_num && part_map_size % m->bank == 0, "bad partition size");
        m->vert_num       = vert_num;
        m->node_offset    = collision_node_offset;
        m->neighbor_total = total_neighbor_num;
        m->part_map_size  = part_map_size;
    });
}

int masx_mas_upload(masx_mas* m, const uint32_t* neighbor_list, const uint32_t* neighbor_start,
                    const uint32_t* neighbor_num, const int* part_to_real, const int* real_to_part)
{
    return guard([&] {
        require(m && m->vert_num > 0, "init_neighbor must come first");
        require(neighbor_start && neighbor_num && part_to_real && real_to_part, "null argument");
        require(m->neighbor_total == 0 || neighbor_list, "null neighbor list");
        m->nbr_list.assign(neighbor_list, neighbor_list + m->neighbor_total);
        m->nbr_start.assign(neighbor_start, neighbor_start + m->vert_num);
        m->nbr_num.assign(neighbor_num, neighbor_num + m->vert_num);
        m->part_to_real.assign(part_to_real, part_to_real + m->part_map_size);
        m->real_to_part.assign(real_to_part, real_to_part + m->vert_num);
        m->uploaded = true;
    });
}

int masx_mas_init_matrix(masx_mas* m, int levels)
{
    return guard([&] {
        require(m && m->uploaded, "upload must come first");
        m->levels       = levels;
        m->matrix_ready = true;
    });
}

int masx_mas_set_preconditioner_bcoo(masx_mas* m, const double* d_values, const int* d_rows, const int* d_cols,
                                     const uint32_t* d_indices, int offset, int triplet_num, const int* d_pairs,
                                     int cp_num, int precision, void* stream)
{
    return guard([&] {
        require(m && m->matrix_ready, "init_matrix must come first");
        require(d_values && d_rows && d_cols && triplet_num >= 0, "bad matrix arguments");
        hipStream_t st = as_stream(stream);
        // structure back to the host (the reference also synchronizes here:
        // i_preconditioner.cu:83-86 and TraditionalMAS32Preconditioner.cu:2211)
        std::vector<uint32_t> idx(triplet_num);
        if(d_indices)
        {
            if(triplet_num)
                if(hipMemcpyAsync(idx.data(), d_indices, sizeof(uint32_t) * triplet_num, hipMemcpyDeviceToHost, st)
                   != hipSuccess)
                    throw std::runtime_error("index copy failed");
        }
        else
            for(int i = 0; i < triplet_num; ++i) idx[i] = (uint32_t)i;
        if(hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("stream sync failed");
        uint32_t max_idx = 0;
        for(uint32_t v : idx) max_idx = std::max(max_idx, v);
        const size_t n_all = triplet_num ? (size_t)max_idx + 1 : 0;
        std::vector<int> all_r(n_all), all_c(n_all);
        if(n_all)
        {
            if(hipMemcpy(all_r.data(), d_rows, sizeof(int) * n_all, hipMemcpyDeviceToHost) != hipSuccess ||
               hipMemcpy(all_c.data(), d_cols, sizeof(int) * n_all, hipMemcpyDeviceToHost) != hipSuccess)
                throw std::runtime_error("structure copy failed");
        }
        std::vector<int> rows(triplet_num), cols(triplet_num), pairs(4 * (size_t)cp_num);
        for(int i = 0; i < triplet_num; ++i)
        {
            rows[i] = all_r[idx[i]] - offset;
            cols[i] = all_c[idx[i]] - offset;
            require(rows[i] >= 0 && cols[i] >= 0 && rows[i] < m->vert_num && cols[i] < m->vert_num,
                    "selected block outside the FEM range");
        }
        if(cp_num)
        {
            require(d_pairs != nullptr, "null contact pairs");
            if(hipMemcpy(pairs.data(), d_pairs, sizeof(int) * 4 * cp_num, hipMemcpyDeviceToHost) != hipSuccess)
                throw std::runtime_error("pair copy failed");
        }
        const bool same = m->solver && rows == m->rows && cols == m->cols && pairs == m->pairs;
        if(!same && m->solver)
        {
            // a new structure on the same mesh (a new contact set): the solver
            // keeps its topology and device buffers (ReorderRealtime runs again)
            auto& d = m->solver->impl.system(0);
            d.nb    = triplet_num;
            d.rows  = rows;
            d.cols  = cols;
            d.pairs = pairs;
            m->solver->impl.structure_changed(0);
            m->solver->impl.setup();
            m->rows  = std::move(rows);
            m->cols  = std::move(cols);
            m->pairs = std::move(pairs);
        }
        else if(!same)
        {
            m->solver = std::make_unique<masx_solver>();
            masx::SystemDesc d;
            d.block_rows   = m->vert_num;
            d.nb           = triplet_num;
            d.n_abd        = 0;
            d.fem_offset   = 0;
            d.n_fem        = m->vert_num;
            d.use_mas      = true;
            d.bank         = m->bank;
            d.levels       = m->levels;
            d.rows         = rows;
            d.cols         = cols;
            d.nbr_start    = m->nbr_start;
            d.nbr_num      = m->nbr_num;
            d.nbr_list     = m->nbr_list;
            d.part_to_real = m->part_to_real;
            d.real_to_part = m->real_to_part;
            d.pairs        = pairs;
            d.node_offset  = m->node_offset;
            m->solver->impl.add_system(d);
            m->solver->impl.setup();
            m->rows  = std::move(rows);
            m->cols  = std::move(cols);
            m->pairs = std::move(pairs);
            if(m->d_sel) (void)hipFree(m->d_sel);
            m->d_sel   = nullptr;
            m->d_sel_n = 0;
        }
        if(triplet_num)
        {
            if(m->d_sel_n < (size_t)triplet_num)
            {
                if(m->d_sel) (void)hipFree(m->d_sel);
                if(hipMalloc((void**)&m->d_sel, sizeof(uint32_t) * triplet_num) != hipSuccess)
                    throw std::runtime_error("allocation failed");
                m->d_sel_n = triplet_num;
            }
            // blocking copy: the gather below runs on the library's stream
            if(hipMemcpy(m->d_sel, idx.data(), sizeof(uint32_t) * triplet_num, hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("index upload failed");
            m->solver->impl.gather_values(0, d_values, m->d_sel, st);
        }
        m->solver->impl.build(precision, st);
        if(hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("build failed");
    });
}

int masx_mas_preconditioning(masx_mas* m, const double* d_R, double* d_Z, void* stream)
{
    return guard([&] {
        require(m && m->solver, "set_preconditioner_bcoo must come first");
        m->solver->impl.apply_preconditioner(d_R, d_Z, as_stream(stream));
    });
}

masx_solver* masx_mas_solver(masx_mas* m) { return m ? m->solver.get() : nullptr; }

struct masx_converter
{
    masx::Converter impl;
};

int masx_converter_create(masx_converter** out)
{
    return guard([&] {
        require(out != nullptr, "null output");
        *out = new masx_converter();
    });
}

int masx_converter_destroy(masx_converter* c)
{
    return guard([&] { delete c; });
}

int masx_convert_triplets(masx_converter* c, int n, int n_block_rows, const int* d_rows, const int* d_cols,
                          const double* d_vals, int drop_lower, int* d_out_rows, int* d_out_cols, double* d_out_vals,
                          int* n_unique, void* stream)
{
    return guard([&] {
        require(c != nullptr && n_unique != nullptr, "null argument");
        *n_unique = c->impl.convert(n, n_block_rows, d_rows, d_cols, d_vals, drop_lower, d_out_rows, d_out_cols,
                                    d_out_vals, as_stream(stream));
    });
}

struct masx_fem
{
    masx::FemHessian impl;
};

int masx_fem_create(masx_fem** out, int n_tets, const int* tets, long long nb, const int* rows, const int* cols)
{
    return guard([&] {
        require(out != nullptr, "null output");
        auto f = std::make_unique<masx_fem>();
        f->impl.setup(n_tets, tets, nb, rows, cols);
        *out = f.release();
    });
}

int masx_fem_destroy(masx_fem* f)
{
    return guard([&] { delete f; });
}

int masx_fem_assemble(masx_fem* f, const double* d_dminv, const double* d_vol, const double* d_F, const double* d_mu,
                      const double* d_lambda, const double* d_base, double* d_vals, void* stream)
{
    return guard([&] {
        require(f != nullptr, "null fem object");
        f->impl.assemble(d_dminv, d_vol, d_F, d_mu, d_lambda, d_base, d_vals, as_stream(stream));
    });
}

int masx_fem_param_grad(masx_fem* f, const double* d_dminv, const double* d_vol, const double* d_F,
                        const double* d_mu, const double* d_lambda, const double* d_x, const double* d_y,
                        double* d_grad_mu, double* d_grad_lambda, void* stream)
{
    return guard([&] {
        require(f != nullptr, "null fem object");
        f->impl.param_grad(d_dminv, d_vol, d_F, d_mu, d_lambda, d_x, d_y, d_grad_mu, d_grad_lambda,
                           as_stream(stream));
    });
}

struct masx_hier
{
    masx::DeviceHierarchy impl;
};

int masx_hier_create(masx_hier** out)
{
    return guard([&] {
        require(out != nullptr, "null output");
        *out = new masx_hier();
    });
}

int masx_hier_destroy(masx_hier* h)
{
    return guard([&] { delete h; });
}

int masx_hier_build(masx_hier* h, int n_fem, int n_slots, int bank, int levels, const uint32_t* d_nbr_start,
                    const uint32_t* d_nbr_num, const uint32_t* d_nbr_list, long long list_len,
                    const int* d_part_to_real, const int* d_real_to_part, const int* d_pairs, int n_pairs,
                    int node_offset, int* out_levels, int* out_total, int* out_components, void* stream)
{
    return guard([&] {
        require(h != nullptr, "null hierarchy");
        h->impl.build(n_fem, n_slots, bank, levels, d_nbr_start, d_nbr_num, d_nbr_list, list_len, d_part_to_real,
                      d_real_to_part, d_pairs, n_pairs, node_offset, as_stream(stream));
        if(out_levels) *out_levels = h->impl.levels();
        if(out_total) *out_total = h->impl.total();
        if(out_components) *out_components = h->impl.n_components();
    });
}

int masx_hier_download(masx_hier* h, int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix,
                       int* going_next, int* coarse_table, int* comp_ptr, uint32_t* comp_mask, int* comp_parent,
                       void* stream)
{
    return guard([&] {
        require(h != nullptr, "null hierarchy");
        h->impl.download(level_count, level_begin, fine_mask, prefix, going_next, coarse_table, comp_ptr, comp_mask,
                         comp_parent, as_stream(stream));
        if(level_count) level_count[0] = 0;
    });
}

int masx_hier_device(masx_hier* h, masx_hier_arrays* out)
{
    return guard([&] {
        require(h != nullptr && out != nullptr, "null argument");
        require(h->impl.levels() > 0, "hierarchy not built");
        out->going_next   = h->impl.d_going_next();
        out->coarse_table = h->impl.d_coarse_table();
        out->fine_mask    = h->impl.d_fine_mask();
        out->prefix       = h->impl.d_prefix();
        out->comp_ptr     = h->impl.d_comp_ptr();
        out->comp_mask    = h->impl.d_comp_mask();
        out->comp_parent  = h->impl.d_comp_parent();
    });
}

}  // extern "C"
