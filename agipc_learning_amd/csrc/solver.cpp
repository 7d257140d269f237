// Host orchestrator of the MAS-PCG solve (see solver.hpp).
#include "solver.hpp"
#include "mas_tiles.hpp"
#include "phase_timer.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "kernels.hpp"

namespace masx
{

#define MASX_HIP(expr)                                                                                  \
    do                                                                                                  \
    {                                                                                                   \
        const hipError_t e_ = (expr);                                                                   \
        if(e_ != hipSuccess)                                                                            \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while(0)



template <typename T>
void DeviceArray<T>::alloc(size_t count)
{
    if(count == 0) count = 1;  // keep a valid pointer for empty arrays
    if(ptr && count <= cap)    // a new structure of no larger size reuses the buffer
    {
        n = count;
        return;
    }
    release();
    MASX_HIP(hipMalloc((void**)&ptr, count * sizeof(T)));
    n = cap = count;
}

template <typename T>
void DeviceArray<T>::upload(const std::vector<T>& h, hipStream_t st)
{
    alloc(h.size());
    if(!h.empty()) MASX_HIP(hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

template struct DeviceArray<int>;
template struct DeviceArray<float>;
template struct DeviceArray<double>;
template struct DeviceArray<char>;
template struct DeviceArray<uint32_t>;
template struct DeviceArray<int4>;
template struct DeviceArray<int2>;
template struct DeviceArray<SysState>;
template struct DeviceArray<uint64_t>;

Solver::Solver() = default;

void Solver::ensure_stream()
{
    if(work_) return;
    MASX_HIP(hipStreamCreateWithFlags(&work_, hipStreamNonBlocking));
    MASX_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    MASX_HIP(hipEventCreate(&ev_a_));
    MASX_HIP(hipEventCreate(&ev_b_));
    MASX_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    MASX_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    for(auto& e : chunk_ev_) MASX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void Solver::analyze()
{
    dev_struct_ = false;
    build_plan(plan_, systems_);
    PhaseTimer tm("analyze");
    validate_plan(plan_);
    tm.mark("validate");
    analyzed_ = true;
}

void Solver::structure_changed(int s)
{
    systems_.at(s).hier_ready = false;
    if(s == hier_dev_sys_) hier_dev_sys_ = -1;
    analyzed_   = false;
    setup_done_ = false;
    built_      = false;
}

// The hierarchy of every MAS system from the device builder (the same arrays
// as the host build); a system the builder rejects (neighbour lists that are
// not symmetric, which the host build reads as undirected) keeps the host path.
void Solver::build_hierarchies_on_device()
{
    hipStream_t st = work_;
    for(auto& d : systems_)
    {
        if(!d.use_mas || d.n_fem <= 0 || d.hier_ready) continue;
        if(!hier_dev_) hier_dev_ = std::make_unique<DeviceHierarchy>();
        const int sys = (int)(&d - systems_.data());
        // the mesh topology and partition stay on the device per system (a new
        // contact set leaves them unchanged)
        while((int)topo_.size() <= sys) topo_.push_back(std::make_unique<TopoCache>());
        TopoCache& tc = *topo_[sys];
        if(tc.version != d.topo_version)
        {
            tc.start.upload(d.nbr_start, st);
            tc.num.upload(d.nbr_num, st);
            tc.list.upload(d.nbr_list, st);
            tc.p2r.upload(d.part_to_real, st);
            tc.r2p.upload(d.real_to_part, st);
            tc.version = d.topo_version;
        }
        d_h_pairs_.upload(d.pairs, st);
        const int n_slots = (int)d.part_to_real.size(), n_pairs = (int)d.pairs.size() / 4;
        try
        {
            hier_dev_->build(d.n_fem, n_slots, d.bank, d.levels, tc.start.ptr, tc.num.ptr, tc.list.ptr,
                             (long long)d.nbr_list.size(), tc.p2r.ptr, tc.r2p.ptr, n_pairs > 0 ? d_h_pairs_.ptr : nullptr,
                             n_pairs, d.node_offset, st);
        }
        catch(const std::invalid_argument&)
        {
            hier_dev_sys_ = -1;
            continue;
        }
        hier_dev_sys_ = sys;
        Hierarchy& h = d.hier;
        const DeviceHierarchy& g = *hier_dev_;
        h.bank    = d.bank;
        h.levels  = g.levels();
        h.n_fem   = d.n_fem;
        h.n_slots = n_slots;
        h.total   = g.total();
        h.level_count.assign(h.levels + 1, 0);
        h.level_begin.assign(h.levels + 1, 0);
        h.going_next.assign(std::max(h.total, d.n_fem), -1);
        h.comp_ptr.assign(h.total / d.bank + 1, 0);
        h.comp_mask.assign(g.n_components(), 0u);
        h.comp_parent.assign(g.n_components(), 0);
        // the plan reads the levels, going_next and the components; with one
        // system the per-vertex tables stay on the device until asked for
        // (the builder keeps them until the next structure)
        h.tables_pending = systems_.size() == 1;
        h.fine_mask.assign(h.tables_pending ? 0 : d.n_fem, 0u);
        h.prefix.assign(h.tables_pending ? 0 : n_slots / d.bank, 0);
        h.coarse_table.assign(h.tables_pending ? 0 : 6 * (size_t)d.n_fem, 0);
        g.download(h.level_count.data(), h.level_begin.data(), h.tables_pending ? nullptr : h.fine_mask.data(),
                   h.tables_pending ? nullptr : h.prefix.data(), h.going_next.data(),
                   h.tables_pending ? nullptr : h.coarse_table.data(), h.comp_ptr.data(), h.comp_mask.data(),
                   h.comp_parent.data(), st);
        d.hier_ready = true;
    }
}

const Hierarchy& Solver::hierarchy(int s)
{
    SystemDesc& d = systems_.at(s);
    Hierarchy&  h = d.hier;
    if(h.tables_pending)
    {
        if(hier_dev_sys_ != s || !hier_dev_) throw std::logic_error("hierarchy tables are no longer on the device");
        h.fine_mask.assign(d.n_fem, 0u);
        h.prefix.assign(h.n_slots / d.bank, 0);
        h.coarse_table.assign(6 * (size_t)d.n_fem, 0);
        ensure_stream();
        hier_dev_->download(nullptr, nullptr, h.fine_mask.data(), h.prefix.data(), nullptr, h.coarse_table.data(),
                            nullptr, nullptr, nullptr, work_);
        h.tables_pending = false;
    }
    return h;
}

Solver::~Solver()
{
    if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if(h_state_) (void)hipHostFree(h_state_);
    if(h_export_) (void)hipHostFree(h_export_);
    if(h_dp_) (void)hipHostFree(h_dp_);
    if(ev_a_) (void)hipEventDestroy(ev_a_);
    if(ev_b_) (void)hipEventDestroy(ev_b_);
    if(ev_fork_) (void)hipEventDestroy(ev_fork_);
    if(ev_join_) (void)hipEventDestroy(ev_join_);
    for(auto e : timing_ev_) (void)hipEventDestroy(e);
    for(auto e : chunk_ev_)
        if(e) (void)hipEventDestroy(e);
    if(side_) (void)hipStreamDestroy(side_);
    if(work_) (void)hipStreamDestroy(work_);
}

hipStream_t Solver::enter(hipStream_t user)
{
    // with no stream given, order after the legacy default stream (the
    // library's stream is non-blocking and would not otherwise wait for it)
    MASX_HIP(hipEventRecord(ev_a_, user));
    MASX_HIP(hipStreamWaitEvent(work_, ev_a_, 0));
    return work_;
}

void Solver::leave(hipStream_t user)
{
    if(user)
    {
        MASX_HIP(hipEventRecord(ev_b_, work_));
        MASX_HIP(hipStreamWaitEvent(user, ev_b_, 0));
    }
    else
        MASX_HIP(hipStreamSynchronize(work_));
}

int Solver::add_system(const SystemDesc& d)
{
    systems_.push_back(d);
    setup_done_ = false;
    analyzed_   = false;
    built_      = false;
    return (int)systems_.size() - 1;
}

void Solver::set_spmv_tiling(int entries_per_thread)
{
    if(!spmv_tiling_supported(entries_per_thread)) throw std::invalid_argument("SpMV entries per thread must be 2, 4 or 8");
    entries_ = entries_per_thread;
    if(setup_done_)
    {
        build_work_tables();
        make_devdata();
        if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
        // the value planes follow the layout: pack them again from the FP64 blocks
        if(built_)
        {
            MASX_HIP(launch_pack_values(dd_, work_));
            MASX_HIP(launch_pack_abd(dd_, d_cp_body_.ptr, d_cp_sub_.ptr, (long long)spmv_layout_.cp_col.size(),
                                     d_vals_abd_.ptr, work_));
            MASX_HIP(hipStreamSynchronize(work_));
        }
    }
}

// From this many banks on the MAS local kernel defaults to the MFMA kernel
// (half-precision 32-slot banks) or the LDS-staged FP32 kernel; below, the
// register-fed VALU kernel's shorter latency wins (measured on 2,270 and
// 91,066 banks).
constexpr int kMfmaAutoBanks = 16384;

int Solver::effective_mas_waves() const
{
    const bool mfma_ok = plan_.bank == 32 && precision_ != kPrecFp32;
    const bool fp32    = precision_ == kPrecFp32;
    if(mas_waves_ == 1) return mfma_ok ? 1 : 4;
    if(mas_waves_ == 4 || mas_waves_ == 8) return mas_waves_;
    if(mas_waves_ == 5 || mas_waves_ == 9) return fp32 ? mas_waves_ : mas_waves_ - 1;  // LDS staging is FP32's
    // automatic: on large batches the MFMA kernel (half) or the LDS-staged
    // FP32 kernel stream the inverse fastest; on a few thousand banks the
    // register-fed 4-wave kernel's shorter latency wins
    if(plan_.n_banks >= kMfmaAutoBanks) return mfma_ok ? 1 : (fp32 ? 9 : 4);
    return 4;
}

void Solver::set_mas_waves(int waves)
{
    if(waves != 0 && waves != 1 && waves != 4 && waves != 5 && waves != 8 && waves != 9)
        throw std::invalid_argument(
            "MAS local variant must be 0 (automatic), 1 (MFMA), 4 or 8 (waves per workgroup), 5 or 9 (4 or 8 waves, "
            "FP32 inverse staged through LDS)");
    const bool tiles_before = effective_mas_waves() == 1;
    mas_waves_    = waves;
    dd_.mas_waves = effective_mas_waves();
    if(built_ && (dd_.mas_waves == 1) != tiles_before) built_ = false;  // the inverse is in the other layout
    if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
}

Solver::Tiling Solver::autotune_mas(int reps)
{
    if(!built_) throw std::runtime_error("build() must precede autotune");
    if(plan_.n_banks == 0) return Tiling{effective_mas_waves(), 0.0};
    // the automatic choice first; another variant replaces the best only when
    // it is more than 2 % faster, so near-ties keep one choice from run to run
    const int first = effective_mas_waves();
    std::vector<int> cand = {first};
    for(const int w : {1, 4, 8, 5, 9})
    {
        if(w == first || (w == 1 && !(plan_.bank == 32 && precision_ != kPrecFp32)) ||
           ((w == 5 || w == 9) && precision_ != kPrecFp32))
            continue;
        cand.push_back(w);
    }
    Tiling best{first, 1e30};
    for(const int w : cand)
    {
        set_mas_waves(w);
        if(!built_) build(precision_, nullptr);
        const double ms = time_kernel(0, reps);
        if(ms < 0.98 * best.ms) best = Tiling{w, ms};
    }
    set_mas_waves(best.entries);
    if(!built_) build(precision_, nullptr);
    return best;
}

Solver::Tiling Solver::autotune_spmv(int reps)
{
    if(!built_) throw std::runtime_error("build() must precede autotune");
    // the current tiling first; another replaces the best only when more than
    // 2 % faster (near-ties keep one choice from run to run)
    const int first = entries_;
    Tiling best{first, 1e30};
    for(const int u : {first, 2, 4, 8})
    {
        if(u == first && best.ms < 1e30) continue;
        set_spmv_tiling(u);
        const double ms = time_kernel(2, reps);
        if(ms < 0.98 * best.ms) best = Tiling{u, ms};
    }
    set_spmv_tiling(best.entries);
    // then the other load policy on the chosen tiling, by the same rule
    const bool nt   = effective_spmv_nt();
    const int  mode = spmv_loads_;
    set_spmv_loads(nt ? 2 : 1);
    const double ms = time_kernel(2, reps);
    if(ms < 0.98 * best.ms)
        best.ms = ms;
    else
        set_spmv_loads(mode);
    return best;
}

// The streamed SpMV matrix (value planes, column words, body couplings)
// from which nontemporal loads pay: on matrices that the 256 MB Infinity
// Cache could hold between iterations, ordinary loads keep them there.
constexpr long long kSpmvCachedBytes = 256ll << 20;

bool Solver::effective_spmv_nt() const
{
    if(spmv_loads_ == 1) return true;
    if(spmv_loads_ == 2) return false;
    const long long bytes = 40ll * spmv_layout_.total + 576ll * (long long)spmv_layout_.cp_col.size();
    return bytes > kSpmvCachedBytes;
}

void Solver::set_spmv_loads(int mode)
{
    if(mode < 0 || mode > 2)
        throw std::invalid_argument("SpMV loads must be 0 (automatic), 1 (nontemporal) or 2 (ordinary)");
    spmv_loads_ = mode;
    dd_.spmv_nt = effective_spmv_nt() ? 1 : 0;
    if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
}

void Solver::setup()
{
    PhaseTimer tm("setup");
    ensure_stream();
    hipStream_t st = work_;
    if(!analyzed_)
    {
        build_hierarchies_on_device();
        tm.mark("device hierarchy");
        if(device_structure_eligible())
        {
            build_plan(plan_, systems_, true);
            validate_plan(plan_);
            dev_struct_ = true;
            analyzed_   = true;
            tm.mark("host plan (MAS tables)");
        }
        else
        {
            download_device_structures();  // the host analysis reads host rows/cols
            analyze();
            tm.mark("host plan");
        }
    }
    const Plan& P = plan_;
    d_e2i_.upload(P.ext_to_int, st);
    if(dev_struct_)
    {
        build_structure_on_device(st);
        tm.mark("device structure");
    }
    else
    {
        d_row_ptr_.upload(P.row_ptr, st);
        d_col_.upload(P.col, st);
        d_entry_src_.upload(P.entry_src, st);
    }
    d_row_sys_.upload(P.row_sys, st);
    d_vals64_.alloc(9 * (size_t)P.nb_total);
    d_E_.alloc(9 * (size_t)std::max<long long>(P.n_gal, 1));
    const size_t nv = 3 * (size_t)P.n_rows;
    d_b64_.alloc(nv);
    d_tmp64a_.alloc(nv);
    d_tmp64b_.alloc(nv);
    for(auto* a : {&d_x_, &d_r_, &d_z_, &d_p0_, &d_p1_, &d_Ap_, &d_rt_}) a->alloc(nv);
    d_Rc_.alloc(3 * (size_t)std::max(P.n_coarse, 1));
    d_Zc_.alloc(3 * (size_t)std::max(P.n_coarse, 1));
    MASX_HIP(hipMemsetAsync(d_Rc_.ptr, 0, d_Rc_.n * sizeof(float), st));
    MASX_HIP(hipMemsetAsync(d_Zc_.ptr, 0, d_Zc_.n * sizeof(float), st));
    d_bank_vec_.upload(P.bank_vec, st);
    d_bank_sys_.upload(P.bank_sys, st);
    d_bank_kind_.upload(P.bank_kind, st);
    d_comp_ptr_.upload(P.comp_ptr, st);
    d_comp_mask_.upload(P.comp_mask, st);
    d_comp_parent_.upload(P.comp_parent, st);
    d_deep_ptr_.upload(P.deep_ptr, st);
    d_deep_list_.upload(P.deep_list, st);
    d_bank_ugroup_.upload(P.bank_ugroup, st);
    d_ugroup_.upload(P.ugroup, st);
    d_comp_dslot_.upload(P.comp_dslot, st);
    d_part_deep_.alloc(3 * (size_t)std::max(P.n_deep_slots, 1));
    d_row_l1_.upload(P.row_l1, st);
    d_anc_.upload(P.anc, st);
    d_minv_scale_.alloc(std::max(P.n_banks, 1));
    // the bank count may have grown with a new structure (alloc keeps a buffer that fits)
    d_minv_.alloc(std::max<size_t>(inverse_bytes(), 16));
    if(!dev_struct_)
    {
        d_gal_ptr_.upload(P.gal_ptr, st);
        d_gal_src_.upload(P.gal_src, st);
        d_ent_ptr_.upload(P.bank_ent_ptr, st);
        d_ent_pos_.upload(P.bank_ent_pos, st);
        d_ent_src_.upload(P.bank_ent_src, st);
        d_diag_src_.upload(P.diag_src, st);
        n_bank_ent_ = P.bank_ent_ptr.empty() ? 0 : P.bank_ent_ptr.back();
    }
    d_body_row_.upload(P.body_row, st);
    d_body_src_.upload(P.body_src, st);
    d_needs_diag_.upload(P.row_needs_diag, st);
    d_row_mode_.upload(P.row_needs_diag, st);
    d_abd_inv_.alloc(144 * (size_t)std::max(P.n_bodies, 1));
    d_diag_inv_.alloc(9 * (size_t)P.n_rows);
    d_state_.alloc(P.n_sys);
    {
        std::vector<int> mi(P.n_sys), i2e(P.n_rows, -1);
        for(int q = 0; q < P.n_sys; ++q) mi[q] = std::max(1, (int)(0.3 * 3.0 * systems_[q].block_rows));
        for(int e = 0; e < P.n_ext; ++e) i2e[P.ext_to_int[e]] = e;
        d_max_iter_.upload(mi, st);
        d_i2e_.upload(i2e, st);
    }
    if(!h_export_)
    {
        MASX_HIP(hipHostMalloc((void**)&h_export_, 64, hipHostMallocCoherent | hipHostMallocMapped));
        MASX_HIP(hipHostGetDevicePointer((void**)&d_export_, h_export_, 0));
    }
    h_export_[0] = h_export_[1] = 0;
    exports_ = 0;
    d_export_seq_.alloc(1);
    MASX_HIP(hipMemsetAsync(d_export_seq_.ptr, 0, sizeof(int), st));
    d_factor_status_.alloc(1);
    MASX_HIP(hipMemsetAsync(d_state_.ptr, 0, d_state_.n * sizeof(SysState), st));
    if(h_state_cap_ < P.n_sys)  // pinned allocations are slow: kept across structures that fit
    {
        if(h_state_) (void)hipHostFree(h_state_);
        h_state_ = nullptr;
        MASX_HIP(hipHostMalloc((void**)&h_state_, sizeof(SysState) * P.n_sys * (kDepth + 1), hipHostMallocDefault));
        h_state_cap_ = P.n_sys;
    }
    tm.mark("uploads");
    build_work_tables();
    tm.mark("work tables + SpMV layout");
    make_devdata();
    MASX_HIP(hipStreamSynchronize(st));
    tm.mark("device sync");
    if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
    setup_done_ = true;
    built_      = false;
}

void Solver::build_work_tables()
{
    const Plan& P = plan_;
    const int   S = P.n_sys;
    std::vector<int4> pre, spmv, upd;
    std::vector<int>  pre_r(2 * S), spmv_r(2 * S), upd_r(2 * S);
    int body_base = 0;
    std::vector<int> fine_bank_base(S, 0);
    {
        int acc = 0;
        for(int s = 0; s < S; ++s)
        {
            fine_bank_base[s] = acc;
            if(systems_[s].use_mas) acc += (P.row_off[s + 1] - P.fem_row0[s]) / P.bank;
        }
    }
    auto rows_chunks = [&](std::vector<int4>& out, int s, int b, int e, int type, int step, int bank_base) {
        for(int r = b; r < e; r += step)
        {
            const int aux = type == kWgFemMas ? bank_base + (r - b) / P.bank : 0;
            out.push_back(make_int4(s, (type << 28) | aux, r, std::min(e, r + step)));
        }
    };
    for(int s = 0; s < S; ++s)
    {
        const auto& d = systems_[s];
        const int abd_rows_end = P.abd_row0[s] + 4 * d.n_abd;
        const int fem_type = d.use_mas ? kWgFemMas : kWgPlain;
        // update kernel
        upd_r[2 * s] = (int)upd.size();
        rows_chunks(upd, s, P.abd_row0[s], P.fem_row0[s], kWgPlain, kBlock, 0);
        rows_chunks(upd, s, P.fem_row0[s], P.row_off[s + 1], fem_type, kBlock, fine_bank_base[s]);
        upd_r[2 * s + 1] = (int)upd.size();
        // preconditioner finish
        pre_r[2 * s] = (int)pre.size();
        for(int b = 0; b < d.n_abd; b += kBodiesPerWg)
            pre.push_back(make_int4(s, kWgAbd << 28, body_base + b, body_base + std::min(d.n_abd, b + kBodiesPerWg)));
        rows_chunks(pre, s, abd_rows_end, P.fem_row0[s], kWgPlain, kBlock, 0);
        rows_chunks(pre, s, P.fem_row0[s], P.row_off[s + 1], fem_type, kBlock, fine_bank_base[s]);
        pre_r[2 * s + 1] = (int)pre.size();
        body_base += d.n_abd;
        if(upd_r[2 * s] == upd_r[2 * s + 1] || pre_r[2 * s] == pre_r[2 * s + 1]) throw std::runtime_error("empty system");
    }
    // SpMV: the layout's tiles (at most 256 rows; several rows only within
    // 256 * entries_ entries, a longer row gets a tile of its own)
    SpmvLayout& L = spmv_layout_;
    hipStream_t st = work_;
    if(dev_struct_)
        build_spmv_layout_on_device(256 * entries_, st);  // fills the layout and the SpMV work tables
    else
    {
        // rows of affine bodies get body tiles with 12 x 12 couplings
        static_assert(kSpmvBodyTile == kWgAbd, "body tiles carry the ABD work type");
        SpmvBodies bodies;
        for(int s = 0, base = 0; s < S; ++s)
        {
            bodies.abd_row0.push_back(P.abd_row0[s]);
            bodies.n_abd.push_back(systems_[s].n_abd);
            bodies.body_base.push_back(base);
            base += systems_[s].n_abd;
        }
        build_spmv_layout(L, S, P.row_off, P.row_ptr, P.col, P.entry_src, 256 * entries_, &bodies);
        for(size_t t = 0; t < L.tiles.size() / 4; ++t)
            spmv.push_back(make_int4(L.tiles[4 * t], L.tiles[4 * t + 1], L.tiles[4 * t + 2], L.tiles[4 * t + 3]));
    }
    d_cp_ptr_.upload(L.cp_ptr, st);
    d_cp_col_.upload(L.cp_col, st);
    {
        std::vector<int> cp_row(L.cp_col.size());
        for(size_t k = 0; k < cp_row.size(); ++k) cp_row[k] = P.body_row[L.cp_col[k]];
        d_cp_row_.upload(cp_row, st);
    }
    d_cp_body_.upload(L.cp_body, st);
    d_cp_sub_.upload(L.cp_sub, st);
    d_vals_abd_.alloc(144 * std::max<size_t>(L.cp_col.size(), 1));
    spmv_r = L.sys_range;
    for(int s = 0; s < S; ++s)
        if(spmv_r[2 * s] == spmv_r[2 * s + 1]) throw std::runtime_error("empty system");
    // the update workgroups of MAS rows must be exactly the plan's update groups
    std::vector<int4> upd_grp(upd.size(), make_int4(0, 0, 0, 0));
    for(size_t k = 0; k < upd.size(); ++k)
    {
        const int4& w = upd[k];
        if((w.y >> 28) != kWgFemMas) continue;
        const int aux = w.y & 0x0fffffff, nbw = (w.w - w.z) / P.bank;
        const int g = aux < (int)P.bank_ugroup.size() ? P.bank_ugroup[aux] : -1;
        if(g < 0 || P.ugroup[4 * g + 2] != P.comp_ptr[aux] || P.ugroup[4 * g + 3] != P.comp_ptr[aux + nbw])
            throw std::logic_error("update workgroups do not match the plan's update groups");
        upd_grp[k] = make_int4(P.ugroup[4 * g], P.ugroup[4 * g + 1], P.ugroup[4 * g + 2], P.ugroup[4 * g + 3]);
    }
    if(!dev_struct_)
    {
        std::vector<int2> spmv_ent(spmv.size());
        for(size_t k = 0; k < spmv.size(); ++k) spmv_ent[k] = make_int2(L.tile_ent[2 * k], L.tile_ent[2 * k + 1]);
        d_sy_row_ptr_.upload(L.row_ptr, st);
        d_sy_col_.upload(L.col, st);
        d_sy_src_.upload(L.src, st);
        d_sy_tptr_.upload(L.tptr, st);
        d_wg_spmv_.upload(spmv, st);
        d_wg_spmv_ent_.upload(spmv_ent, st);
        n_wg_spmv_ = (int)spmv.size();
    }
    d_vals32_.alloc(9 * (size_t)std::max<long long>(64, (L.total + 63) / 64 * 64));
    d_wg_pre_.upload(pre, st);
    d_wg_upd_.upload(upd, st);
    d_wg_upd_grp_.upload(upd_grp, st);
    d_wg_pre_range_.upload(pre_r, st);
    d_wg_spmv_range_.upload(spmv_r, st);
    d_wg_upd_range_.upload(upd_r, st);
    n_wg_pre_  = (int)pre.size();
    n_wg_upd_  = (int)upd.size();
    d_part_pre_.alloc(n_wg_pre_);
    d_part_spmv_.alloc(n_wg_spmv_);
    d_part_upd_.alloc(n_wg_upd_);
    d_part_tr_.alloc(n_wg_spmv_);
    d_part_drift_.alloc(n_wg_spmv_);
    MASX_HIP(hipStreamSynchronize(st));
}

void Solver::make_devdata()
{
    const Plan& P = plan_;
    DevData& d = dd_;
    d.n_rows = P.n_rows;
    d.n_sys = P.n_sys;
    d.single = P.n_sys == 1 ? 1 : 0;
    d.n_wg_pre = n_wg_pre_;
    d.n_wg_spmv = n_wg_spmv_;
    d.n_wg_upd = n_wg_upd_;
    d.bank = P.bank;
    d.n_banks = P.n_banks;
    d.mas_waves = effective_mas_waves();
    d.n_fine_banks = P.n_fine_banks;
    d.nnz = P.nnz;
    d.nnz_s = spmv_layout_.total;
    d.spmv_nt = effective_spmv_nt() ? 1 : 0;
    d.spmv_bodies = P.n_bodies > 0 ? 1 : 0;
    d.factor_waves = effective_factor_waves();
    d.inv_map = d_inv_map_.ptr;
    d.plane = std::max<long long>(64, (d.nnz_s + 63) / 64 * 64);
    d.row_ptr = d_row_ptr_.ptr;
    d.col = d_col_.ptr;
    d.entry_src = d_entry_src_.ptr;
    d.sy_row_ptr = d_sy_row_ptr_.ptr;
    d.sy_col = d_sy_col_.ptr;
    d.sy_src = d_sy_src_.ptr;
    d.sy_tptr = d_sy_tptr_.ptr;
    d.abd_cp_ptr = d_cp_ptr_.ptr;
    d.abd_cp_col = d_cp_col_.ptr;
    d.abd_cp_row = d_cp_row_.ptr;
    d.vals_abd = d_vals_abd_.ptr;
    d.row_sys = d_row_sys_.ptr;
    d.vals32 = d_vals32_.ptr;
    d.vals64 = d_vals64_.ptr;
    d.x = d_x_.ptr;
    d.r = d_r_.ptr;
    d.z = d_z_.ptr;
    d.p0 = d_p0_.ptr;
    d.p1 = d_p1_.ptr;
    d.Ap = d_Ap_.ptr;
    d.rt = d_rt_.ptr;
    d.b64 = d_b64_.ptr;
    d.Rc = d_Rc_.ptr;
    d.Zc = d_Zc_.ptr;
    d.minv = d_minv_.ptr;
    d.minv_scale = d_minv_scale_.ptr;
    d.bank_vec = d_bank_vec_.ptr;
    d.bank_sys = d_bank_sys_.ptr;
    d.bank_kind = d_bank_kind_.ptr;
    d.comp_ptr = d_comp_ptr_.ptr;
    d.comp_mask = d_comp_mask_.ptr;
    d.comp_parent = d_comp_parent_.ptr;
    d.deep_ptr = d_deep_ptr_.ptr;
    d.deep_list = d_deep_list_.ptr;
    d.bank_ugroup = d_bank_ugroup_.ptr;
    d.ugroup = reinterpret_cast<const int4*>(d_ugroup_.ptr);
    d.comp_dslot = d_comp_dslot_.ptr;
    d.part_deep = d_part_deep_.ptr;
    d.row_l1 = d_row_l1_.ptr;
    d.anc = d_anc_.ptr;
    d.abd_inv = d_abd_inv_.ptr;
    d.body_row = d_body_row_.ptr;
    d.diag_inv = d_diag_inv_.ptr;
    d.row_mode = d_row_mode_.ptr;
    d.wg_pre = d_wg_pre_.ptr;
    d.wg_spmv = d_wg_spmv_.ptr;
    d.wg_upd = d_wg_upd_.ptr;
    d.wg_upd_grp = d_wg_upd_grp_.ptr;
    d.wg_spmv_ent = d_wg_spmv_ent_.ptr;
    d.wg_pre_range = d_wg_pre_range_.ptr;
    d.wg_spmv_range = d_wg_spmv_range_.ptr;
    d.wg_upd_range = d_wg_upd_range_.ptr;
    d.part_pre = d_part_pre_.ptr;
    d.part_spmv = d_part_spmv_.ptr;
    d.part_upd = d_part_upd_.ptr;
    d.part_tr = d_part_tr_.ptr;
    d.part_drift = d_part_drift_.ptr;
    d.state = d_state_.ptr;
}

void Solver::set_values(int s, const double* d_values, hipStream_t st)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede set_values()");
    const long long nb = systems_.at(s).nb;
    hipStream_t w = enter(st);
    if(nb) MASX_HIP(hipMemcpyAsync(d_vals64_.ptr + 9 * plan_.src_off[s], d_values, 9 * nb * sizeof(double),
                                   hipMemcpyDeviceToDevice, w));
    leave(st);
    built_ = false;
}

void Solver::set_structure_device(int s, long long nb, const int* d_rows, const int* d_cols, hipStream_t user)
{
    SystemDesc& d = systems_.at(s);
    if(nb < 0 || (nb > 0 && (!d_rows || !d_cols))) throw std::invalid_argument("null rows/cols");
    ensure_stream();
    while(d_sys_rows_.size() < systems_.size())
    {
        d_sys_rows_.push_back(std::make_unique<DeviceArray<int>>());
        d_sys_cols_.push_back(std::make_unique<DeviceArray<int>>());
    }
    d_sys_rows_[s]->alloc(std::max<long long>(nb, 1));
    d_sys_cols_[s]->alloc(std::max<long long>(nb, 1));
    hipStream_t w = enter(user);
    if(nb)
    {
        MASX_HIP(hipMemcpyAsync(d_sys_rows_[s]->ptr, d_rows, sizeof(int) * nb, hipMemcpyDeviceToDevice, w));
        MASX_HIP(hipMemcpyAsync(d_sys_cols_[s]->ptr, d_cols, sizeof(int) * nb, hipMemcpyDeviceToDevice, w));
    }
    leave(user);
    d.nb = nb;
    d.rows.clear();
    d.cols.clear();
    d.rows_on_device = true;
    structure_changed(s);
}

void Solver::download_device_structures()
{
    for(size_t s = 0; s < systems_.size(); ++s)
    {
        SystemDesc& d = systems_[s];
        if(!d.rows_on_device) continue;
        d.rows.resize(d.nb);
        d.cols.resize(d.nb);
        MASX_HIP(hipStreamSynchronize(work_));
        if(d.nb)
        {
            MASX_HIP(hipMemcpy(d.rows.data(), d_sys_rows_[s]->ptr, sizeof(int) * d.nb, hipMemcpyDeviceToHost));
            MASX_HIP(hipMemcpy(d.cols.data(), d_sys_cols_[s]->ptr, sizeof(int) * d.nb, hipMemcpyDeviceToHost));
        }
        d.rows_on_device = false;
    }
}

void Solver::set_values_many(int n, const int* systems, const double* const* d_values, hipStream_t st)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede set_values()");
    for(int q = 0; q < n; ++q) (void)systems_.at(systems[q]);  // range checks before any copy
    hipStream_t w = enter(st);
    // one launch per kCopySegmentsMax systems instead of one copy per system
    CopySegments c{};
    auto flush = [&]() {
        MASX_HIP(launch_copy_segments(c, d_vals64_.ptr, w));
        c = CopySegments{};
    };
    for(int q = 0; q < n; ++q)
    {
        const int s = systems[q];
        const long long nb = systems_[s].nb;
        if(!nb) continue;
        if(!d_values[q]) throw std::invalid_argument("null values");
        const long long blocks = (9 * nb + kCopyChunk - 1) / kCopyChunk;
        if(c.count == kCopySegmentsMax || (long long)c.first[c.count] + blocks > (long long)INT_MAX) flush();
        c.src[c.count] = d_values[q];
        c.dst[c.count] = 9 * plan_.src_off[s];
        c.n[c.count]   = 9 * nb;
        c.first[c.count + 1] = c.first[c.count] + (int)blocks;
        ++c.count;
    }
    flush();
    leave(st);
    built_ = false;
}

void Solver::set_values_host(int s, const double* h_values, hipStream_t st)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede set_values()");
    const long long nb = systems_.at(s).nb;
    hipStream_t w = enter(st);
    if(nb) MASX_HIP(hipMemcpyAsync(d_vals64_.ptr + 9 * plan_.src_off[s], h_values, 9 * nb * sizeof(double),
                                   hipMemcpyHostToDevice, w));
    MASX_HIP(hipStreamSynchronize(w));  // the host buffer may go away after return
    leave(st);
    built_ = false;
}

void Solver::gather_values(int s, const double* d_values, const uint32_t* d_indices, hipStream_t st)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede set_values()");
    hipStream_t w = enter(st);
    MASX_HIP(launch_gather_blocks(d_values, d_indices, d_vals64_.ptr + 9 * plan_.src_off[s], systems_.at(s).nb, w));
    leave(st);
    built_ = false;
}

size_t Solver::inverse_bytes() const
{
    return (size_t)plan_.n_banks * mas_bank_bytes(plan_.bank, precision_);
}

// From this many banks the MAS factor runs one wave per bank: more banks
// per SIMD and no barrier between waves on the pivot chain (measured on
// 91,066 banks: 4.3 against 5.5 ms); on a few thousand banks the build waits
// on one bank's latency, which two waves per bank shorten (2,270 banks: 0.16
// against 0.19 ms).
constexpr int kFactorWaveBanks = 16384;

int Solver::effective_factor_waves() const
{
    if(factor_waves_ == 1 || factor_waves_ == 2) return factor_waves_;
    return plan_.n_banks >= kFactorWaveBanks ? 1 : 2;
}

void Solver::set_factor_waves(int waves)
{
    if(waves < 0 || waves > 2) throw std::invalid_argument("MAS factor waves per bank must be 0 (automatic), 1 or 2");
    factor_waves_    = waves;
    dd_.factor_waves = effective_factor_waves();
}

// From this many banks build() overlaps its memory-bound kernels with the
// MAS factor (a factor of milliseconds); measured: 91,066 banks gain, 2,270
// banks lose the ~50 us the extra launch and the stream waits cost.
constexpr int kOverlapBanks = 16384;

void Solver::build(int precision, hipStream_t user)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede build()");
    if(precision < kPrecFp32 || precision > kPrecBf16) throw std::invalid_argument("unknown precision");
    hipStream_t st = enter(user);
    const bool prec_changed = precision != precision_;
    precision_ = precision;
    dd_.mas_waves    = effective_mas_waves();  // the factor writes the layout the local kernel reads
    dd_.factor_waves = effective_factor_waves();
    const char* old_minv = d_minv_.ptr;
    d_minv_.alloc(std::max<size_t>(inverse_bytes(), 16));  // reuses the buffer when it fits
    if(inverse_bytes() > d_minv_.cap) throw std::logic_error("inverse buffer smaller than the banks need");
    if(prec_changed || d_minv_.ptr != old_minv)
    {
        // the captured chunk holds the old pointer (and the old precision's kernel)
        make_devdata();
        if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
    const Plan& P = plan_;
    {
        // the factor writes the inverse through this layout's map
        const bool tiles = P.bank == 32 && precision_ != kPrecFp32 && dd_.mas_waves == 1;
        const int  key   = (P.bank * 4 + precision_) * 2 + (tiles ? 1 : 0);
        if(key != inv_map_key_ && P.n_banks > 0)
        {
            d_inv_map_.alloc((size_t)inverse_map_entries(P.bank, precision_, tiles));
            MASX_HIP(launch_inverse_map(P.bank, precision_, tiles, d_inv_map_.ptr, st));
            inv_map_key_ = key;
        }
        dd_.inv_map = d_inv_map_.ptr;
    }
    // On large systems the MAS factor of the fine banks (compute-bound,
    // milliseconds) runs on the work stream while the side stream forms the
    // coarse levels' Galerkin sums and factors the coarse banks, then packs
    // the SpMV value planes and factors the block-Jacobi blocks (which only
    // the solve needs): the side stream's kernels fill the CUs the factor's
    // register-limited waves leave free.  On small systems the cross-stream
    // waits would cost more than they hide: one stream, in that order.
    const bool overlap = P.n_banks >= kOverlapBanks;
    hipStream_t aux = overlap ? side_ : st;
    MASX_HIP(hipMemsetAsync(d_factor_status_.ptr, 0, sizeof(int), st));
    if(overlap)
    {
        MASX_HIP(hipEventRecord(ev_fork_, st));
        MASX_HIP(hipStreamWaitEvent(side_, ev_fork_, 0));
    }
    for(size_t g = 0; g + 1 < P.gal_level_ptr.size(); ++g)
    {
        // a whole wave per entry where the lists are long (the top levels)
        const int e0 = P.gal_level_ptr[g], e1 = P.gal_level_ptr[g + 1];
        const long long srcs = (long long)P.gal_level_src[g + 1] - P.gal_level_src[g];
        const int lanes = srcs > 32ll * (e1 - e0) ? 64 : 16;
        MASX_HIP(launch_gal_gather(d_vals64_.ptr, d_E_.ptr, d_gal_ptr_.ptr, d_gal_src_.ptr, e0, e1,
                                   P.gal_src_is_fine[g], lanes, aux));
    }
    const int split = overlap ? P.n_fine_banks : P.n_banks;
    if(overlap)
        MASX_HIP(launch_mas_factor(dd_, d_E_.ptr, d_ent_ptr_.ptr, d_ent_pos_.ptr, d_ent_src_.ptr, d_minv_.ptr,
                                   d_minv_scale_.ptr, precision_, d_factor_status_.ptr, split, P.n_banks, side_));
    MASX_HIP(launch_pack_values(dd_, aux));
    MASX_HIP(launch_pack_abd(dd_, d_cp_body_.ptr, d_cp_sub_.ptr, (long long)spmv_layout_.cp_col.size(),
                             d_vals_abd_.ptr, aux));
    MASX_HIP(launch_abd_factor(dd_, d_body_src_.ptr, d_abd_inv_.ptr, P.n_bodies, aux));
    MASX_HIP(launch_diag_factor(dd_, d_diag_src_.ptr, d_needs_diag_.ptr, d_diag_inv_.ptr, aux));
    if(overlap) MASX_HIP(hipEventRecord(ev_join_, side_));
    MASX_HIP(launch_mas_factor(dd_, d_E_.ptr, d_ent_ptr_.ptr, d_ent_pos_.ptr, d_ent_src_.ptr, d_minv_.ptr,
                               d_minv_scale_.ptr, precision_, d_factor_status_.ptr, 0, split, st));
    if(overlap) MASX_HIP(hipStreamWaitEvent(st, ev_join_, 0));
    leave(user);
    built_ = true;
}

long long Solver::factor_status() const
{
    int v = 0;
    MASX_HIP(hipStreamSynchronize(work_));
    MASX_HIP(hipMemcpyAsync(&v, d_factor_status_.ptr, sizeof(int), hipMemcpyDeviceToHost, work_));
    MASX_HIP(hipStreamSynchronize(work_));
    return v;
}

void Solver::launch_iteration(hipStream_t st, int i)
{
    const float* p_old = (i & 1) ? d_p1_.ptr : d_p0_.ptr;
    float*       p_new = (i & 1) ? d_p0_.ptr : d_p1_.ptr;
    MASX_HIP(launch_mas_local(dd_, precision_, 0, st));
    MASX_HIP(launch_precond_finish(dd_, n_wg_pre_, 0, st));
    MASX_HIP(launch_spmv_fused(dd_, n_wg_spmv_, entries_, p_old, p_new, st));
    MASX_HIP(launch_update(dd_, n_wg_upd_, p_new, 0, st));
}

void Solver::run_chunk(hipStream_t st, int chunk, bool with_check)
{
    for(int i = 0; i < chunk; ++i) launch_iteration(st, i);
    // one system: the check's last kernel also exports the status
    const bool fold_export = with_check && plan_.n_sys == 1;
    if(with_check)
    {
        MASX_HIP(launch_true_residual(dd_, n_wg_spmv_, 0, st));
        MASX_HIP(launch_update(dd_, n_wg_upd_, d_p0_.ptr, 1, st));
        MASX_HIP(launch_phase_finish(dd_, st, fold_export ? d_export_ : nullptr,
                                     fold_export ? d_export_seq_.ptr : nullptr));
    }
    if(!fold_export) MASX_HIP(launch_export_status(dd_, d_export_, d_export_seq_.ptr, st));
}

void Solver::solve(const double* d_b, double* d_x, const SolveParams& prm, std::vector<SolveStats>* stats,
                   hipStream_t user)
{
    if(!built_) throw std::runtime_error("build() must precede solve()");
    const Plan& P = plan_;
    hipStream_t st = enter(user);
    const int chunk = std::max(2, prm.chunk + (prm.chunk & 1));
    int max_total = 0;
    for(int s = 0; s < P.n_sys; ++s)
        max_total = std::max(max_total, prm.max_iter > 0 ? prm.max_iter
                                                         : std::max(1, (int)(0.3 * 3.0 * systems_[s].block_rows)));
    MASX_HIP(launch_init_state(dd_, prm.rule, prm.tol, d_max_iter_.ptr, prm.max_iter,
                               prm.rule == kRuleTrueRel ? prm.check_every : 0, 0, st));
    MASX_HIP(launch_solve_prepare(dd_, d_i2e_.ptr, d_b, st));
    MASX_HIP(launch_update(dd_, n_wg_upd_, d_p0_.ptr, 1, st));
    MASX_HIP(launch_phase_finish(dd_, st));
    const bool with_check = prm.rule == kRuleTrueRel;
    if(prm.use_graph && (!graph_exec_ || graph_chunk_ != chunk || graph_check_ != with_check))
    {
        if(graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
        hipGraph_t g = nullptr;
        MASX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        run_chunk(st, chunk, with_check);
        MASX_HIP(hipStreamEndCapture(st, &g));
        MASX_HIP(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        graph_chunk_ = chunk;
        graph_check_ = with_check;
    }
    const int max_chunks = max_total / chunk + 8 + (with_check ? max_total / 10 : 0);
    // Chunks are launched ahead of the status checks: kDepth in flight, and
    // up to 80% of the previous solve's iterations before the first check.
    // A chunk launched after convergence costs only its early-exit launches.
    const size_t state_bytes = sizeof(SysState) * P.n_sys;
    // Every chunk ends by writing (unfinished count, sequence number) to host
    // memory; the host polls that instead of a per-chunk event or copy.
    const long long seq0 = exports_;
    int launched = 0, checked = 0;
    auto launch_one = [&]() {
        if(prm.use_graph)
            MASX_HIP(hipGraphLaunch(graph_exec_, st));
        else
            run_chunk(st, chunk, with_check);
        ++launched;
        ++exports_;
    };
    auto wait_export = [&](long long target) {
        for(unsigned spins = 1;; ++spins)
        {
            if((long long)*(volatile int*)(h_export_ + 1) >= target) return;
            if((spins & 255) == 0)
            {
                const hipError_t q = hipStreamQuery(st);
                if(q == hipSuccess)
                {
                    if((long long)*(volatile int*)(h_export_ + 1) >= target) return;
                    throw std::runtime_error("solve: chunk status was not exported");
                }
                if(q != hipErrorNotReady) MASX_HIP(q);
                std::this_thread::yield();
            }
        }
    };
    const int ahead = std::min(max_chunks, std::max(kDepth, (int)(0.8 * last_iters_) / chunk));
    while(launched < ahead) launch_one();
    checked = std::max(0, launched - kDepth);
    for(;;)
    {
        while(launched < max_chunks && launched - checked < kDepth) launch_one();
        if(checked >= launched) break;
        // a later chunk may already have overwritten the count, which only moves it on
        wait_export(seq0 + checked + 1);
        if(*(volatile int*)h_export_ == 0) break;
        ++checked;
    }
    MASX_HIP(launch_store_x(dd_, d_e2i_.ptr, d_x, P.n_ext, st));
    MASX_HIP(launch_true_residual(dd_, n_wg_spmv_, 1, st));
    SysState* fin = h_state_ + (size_t)kDepth * P.n_sys;
    MASX_HIP(hipMemcpyAsync(fin, d_state_.ptr, state_bytes, hipMemcpyDeviceToHost, st));
    MASX_HIP(hipStreamSynchronize(st));
    last_iters_ = 0;
    for(int s = 0; s < P.n_sys; ++s) last_iters_ = std::max(last_iters_, fin[s].k);
    if(stats)
    {
        stats->resize(P.n_sys);
        for(int s = 0; s < P.n_sys; ++s)
        {
            const SysState& h = fin[s];
            SolveStats& o = (*stats)[s];
            o.iterations = h.k;
            if(h.rule == kRuleRefRtz && h.failure == kIterationCap) o.iterations = h.max_iter;
            o.status = h.status;
            o.failure = h.failure;
            // the host ran out of chunks (repeated restarts) before the system
            // stopped: report it as capped, not as running without a failure
            if(h.status != kDone && h.failure == kOk) o.failure = kIterationCap;
            o.rejected = h.rejected;
            o.restarts = h.restarts;
            o.true_checks = h.n_true_checks;
            o.rz0 = h.rz0;
            o.rz = h.rz_cur;
            o.rel_res_recursive = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
            o.rel_res_true = h.bb > 0 ? std::sqrt(h.true_rr / h.bb) : std::sqrt(h.true_rr);
        }
    }
    leave(user);
}

void Solver::spmv(double a, const double* d_x, double b, double* d_y, hipStream_t user)
{
    if(!setup_done_) throw std::runtime_error("setup() must precede spmv()");
    hipStream_t st = enter(user);
    MASX_HIP(hipMemsetAsync(d_tmp64a_.ptr, 0, 3 * (size_t)plan_.n_rows * sizeof(double), st));
    MASX_HIP(launch_gather_ext64(d_e2i_.ptr, d_x, d_tmp64a_.ptr, plan_.n_ext, st));
    MASX_HIP(launch_spmv64(dd_, d_tmp64a_.ptr, d_tmp64b_.ptr, st));
    MASX_HIP(launch_axpby_int_to_ext(d_e2i_.ptr, d_tmp64b_.ptr, d_y, a, b, plan_.n_ext, st));
    leave(user);
}

void Solver::apply_preconditioner(const double* d_r, double* d_z, hipStream_t user)
{
    if(!built_) throw std::runtime_error("build() must precede apply_preconditioner()");
    hipStream_t st = enter(user);
    const size_t nvb = 3 * (size_t)plan_.n_rows * sizeof(float);
    MASX_HIP(hipMemsetAsync(d_rt_.ptr, 0, nvb, st));
    MASX_HIP(hipMemsetAsync(d_b64_.ptr, 0, 3 * (size_t)plan_.n_rows * sizeof(double), st));
    MASX_HIP(launch_init_state(dd_, kRuleTrueRel, 0.0, d_max_iter_.ptr, 1, 0, 1, st));
    MASX_HIP(launch_load_rhs(dd_, d_e2i_.ptr, d_r, plan_.n_ext, st));
    MASX_HIP(launch_update(dd_, n_wg_upd_, d_p0_.ptr, 1, st));
    MASX_HIP(launch_phase_finish(dd_, st));
    MASX_HIP(launch_mas_local(dd_, precision_, 1, st));
    MASX_HIP(launch_precond_finish(dd_, n_wg_pre_, 1, st));
    MASX_HIP(launch_z32_to_ext(d_e2i_.ptr, d_z_.ptr, d_z, plan_.n_ext, st));
    leave(user);
}

void Solver::download_inverse(std::vector<float>& out)
{
    // to packed upper blocks (block row-major), unscaled
    const int    B    = plan_.bank;
    const size_t nblk = (size_t)B * (B + 1) / 2;
    out.assign((size_t)plan_.n_banks * nblk * 9, 0.f);
    MASX_HIP(hipStreamSynchronize(work_));
    std::vector<float> scale(plan_.n_banks, 1.f);
    if(plan_.n_banks)
        MASX_HIP(hipMemcpy(scale.data(), d_minv_scale_.ptr, sizeof(float) * plan_.n_banks, hipMemcpyDeviceToHost));
    const size_t bank_bytes = mas_bank_bytes(B, precision_);
    // words per bank as the local kernels index them: FP32 9 per block, the
    // half block-diagonal layout 5 per block, the MFMA tiles kTileBankHalves / 2
    const size_t stride_words = precision_ == kPrecFp32 ? nblk * 9
                                : (B == 32 && dd_.mas_waves == 1) ? (size_t)kTileBankHalves / 2
                                                                   : nblk * 5;
    std::vector<uint32_t> raw((size_t)plan_.n_banks * bank_bytes / 4);
    if(!raw.empty()) MASX_HIP(hipMemcpy(raw.data(), d_minv_.ptr, raw.size() * 4, hipMemcpyDeviceToHost));
    auto half_to_float = [&](uint32_t v16) {
        float f;
        if(precision_ == kPrecBf16)
        {
            const uint32_t u = v16 << 16;
            std::memcpy(&f, &u, 4);
            return f;
        }
        const uint32_t sgn = (v16 >> 15) & 1u, ex = (v16 >> 10) & 0x1fu, man = v16 & 0x3ffu;
        float m;
        if(ex == 0) m = std::ldexp((float)man, -24);
        else if(ex == 31) m = man ? NAN : INFINITY;
        else m = std::ldexp((float)(man | 0x400u), (int)ex - 25);
        return sgn ? -m : m;
    };
    for(int b = 0; b < plan_.n_banks; ++b)
    {
        const uint32_t* bw = raw.data() + (size_t)b * stride_words;
        float* dst = out.data() + (size_t)b * nblk * 9;
        for(int dd = 0; dd < B; ++dd)
        {
            const int len = B - dd, off = dd * B - dd * (dd - 1) / 2;
            for(int I = 0; I < len; ++I)
            {
                const int blk = B * I - I * (I + 1) / 2 + I + dd;  // packed index of (I, I+dd)
                for(int c = 0; c < 9; ++c)
                {
                    float v;
                    if(precision_ == kPrecFp32)
                    {
                        std::memcpy(&v, &bw[9 * off + c * len + I], 4);
                    }
                    else if(B == 32 && dd_.mas_waves == 1)
                    {
                        // MFMA tiles: element (3 I + c / 3, 3 (I + dd) + c % 3) of the upper triangle
                        const int r = 3 * I + c / 3, col = 3 * (I + dd) + c % 3;
                        const int hidx = r <= col ? tile_half_index(r, col) : tile_half_index(col, r);
                        const uint32_t w = bw[hidx / 2];
                        v = half_to_float(hidx % 2 ? (w >> 16) : (w & 0xffffu)) * scale[b];
                    }
                    else
                    {
                        const uint32_t w = bw[5 * off + (c / 2) * len + I];
                        v = half_to_float(c % 2 ? (w >> 16) : (w & 0xffffu)) * scale[b];
                    }
                    dst[9 * (size_t)blk + c] = v;
                }
            }
        }
    }
}

void Solver::download_coarse_entries(std::vector<double>& out)
{
    out.resize(9 * (size_t)plan_.n_gal);
    MASX_HIP(hipStreamSynchronize(work_));
    if(plan_.n_gal) MASX_HIP(hipMemcpy(out.data(), d_E_.ptr, out.size() * sizeof(double), hipMemcpyDeviceToHost));
}

double Solver::time_kernel(int which, int reps)
{
    if(!built_) throw std::runtime_error("build() must precede time_kernel()");
    hipStream_t st = enter(nullptr);
    MASX_HIP(hipStreamSynchronize(st));
    // put every system in the running state with finite scalars
    std::vector<SysState> h(plan_.n_sys);
    MASX_HIP(hipMemcpy(h.data(), d_state_.ptr, sizeof(SysState) * plan_.n_sys, hipMemcpyDeviceToHost));
    for(auto& s : h)
    {
        s.status = kRunning;
        s.rz0 = 1.0;
        s.beta = 0.5;
        s.alpha = 0.0;
        s.max_iter = 1 << 30;
        s.k = 1;
        s.need_check = 0;
        s.check_every = 0;
        s.tol = 0.0;
        s.beta_zero = 0;
        s.rz_prev = s.rz_cur = 1.0;
        s.frozen = 1;
    }
    MASX_HIP(hipMemcpy(d_state_.ptr, h.data(), sizeof(SysState) * plan_.n_sys, hipMemcpyHostToDevice));
    auto launch = [&](const LaunchEvents* ev) {
        switch(which)
        {
            case 0: MASX_HIP(launch_mas_local(dd_, precision_, 0, st, ev)); break;
            case 1: MASX_HIP(launch_precond_finish(dd_, n_wg_pre_, 0, st, ev)); break;
            case 2:
                MASX_HIP(launch_spmv_fused(dd_, n_wg_spmv_, entries_, d_p0_.ptr, d_p1_.ptr, st, ev));
                break;
            case 3: MASX_HIP(launch_update(dd_, n_wg_upd_, d_p1_.ptr, 0, st, ev)); break;
            default: throw std::invalid_argument("unknown kernel id");
        }
    };
    for(int i = 0; i < std::min(reps, 20); ++i) launch(nullptr);  // untimed warm-up
    MASX_HIP(hipStreamSynchronize(st));
    // each launch carries events stamped by its own dispatch, so the average
    // is of the kernel's execution time (what a kernel trace reports), not of
    // the launch rate; the frozen state keeps every system running
    while(timing_ev_.size() < 2 * (size_t)reps)
    {
        hipEvent_t e = nullptr;
        MASX_HIP(hipEventCreate(&e));
        timing_ev_.push_back(e);
    }
    for(int i = 0; i < reps; ++i)
    {
        const LaunchEvents ev{timing_ev_[2 * i], timing_ev_[2 * i + 1]};
        launch(&ev);
    }
    MASX_HIP(hipStreamSynchronize(st));
    double total = 0.0;
    for(int i = 0; i < reps; ++i)
    {
        float ms = 0.f;
        MASX_HIP(hipEventElapsedTime(&ms, timing_ev_[2 * i], timing_ev_[2 * i + 1]));
        total += ms;
    }
    return total / reps;
}

}  // namespace masx
