// The solver's structure analysis on the device (devplan.hpp): for one MAS
// system without affine bodies, the arrays that build_plan and
// build_spmv_layout make on the host for the same structure.
#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>

#include "devplan.hpp"
#include "solver.hpp"
#include "spmv_layout.hpp"

namespace masx
{
namespace
{
void check(hipError_t e, const char* what)
{
    if(e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int bits_for(long long n)  // smallest b >= 1 with n <= 2^b
{
    int b = 1;
    while((1ll << b) < n) ++b;
    return b;
}

}  // namespace

// device ints into the pinned read-back slots 0..n-1, one synchronization
const int* Solver::read_ints(std::initializer_list<const int*> ptrs, hipStream_t st)
{
    if(!h_dp_) check(hipHostMalloc((void**)&h_dp_, 16 * sizeof(int), hipHostMallocDefault), "pinned read-back buffer");
    if(ptrs.size() > 16) throw std::logic_error("too many read-backs");
    int k = 0;
    for(const int* p : ptrs) check(hipMemcpyAsync(h_dp_ + k++, p, sizeof(int), hipMemcpyDeviceToHost, st), "read back");
    check(hipStreamSynchronize(st), "stream sync");
    return h_dp_;
}

bool Solver::device_structure_eligible() const
{
    if(!dev_struct_mode_ || systems_.empty()) return false;
    for(const SystemDesc& d : systems_)
        if(!d.use_mas || d.n_abd != 0 || d.n_fem <= 0 || !d.hier_ready) return false;
    return true;
}

void Solver::build_structure_on_device(hipStream_t st)
{
    Plan&           P      = plan_;
    const int       S      = (int)systems_.size();
    const int       n_rows = P.n_rows, bank = P.bank;
    const long long nb     = P.nb_total;

    // ---- the systems' ranges (one small array: DpSystems points into it)
    int max_levels = 1;
    for(const SystemDesc& d : systems_) max_levels = std::max(max_levels, d.hier.levels);
    std::vector<int> meta;  // src_off, ext_off, row_off (S+1 each), levels, gn_off, coarse_base, n_slots (S each),
                            // then the first slot of every level of every system (S per level)
    std::vector<int> gn_off(S + 1, 0);
    for(int s = 0; s < S; ++s)
        gn_off[s + 1] = gn_off[s] + (int)systems_[s].hier.going_next.size();
    for(int s = 0; s <= S; ++s) meta.push_back((int)P.src_off[s]);
    for(int s = 0; s <= S; ++s) meta.push_back(P.ext_off[s]);
    for(int s = 0; s <= S; ++s) meta.push_back(P.row_off[s]);
    for(int s = 0; s < S; ++s) meta.push_back(systems_[s].hier.levels);
    for(int s = 0; s < S; ++s) meta.push_back(gn_off[s]);
    for(int s = 0; s < S; ++s) meta.push_back(P.coarse_base[s]);
    for(int s = 0; s < S; ++s) meta.push_back(systems_[s].hier.n_slots);
    const size_t lvl0 = meta.size();
    for(int l = 0; l < max_levels; ++l)
        for(int s = 0; s < S; ++s)
            meta.push_back(l < systems_[s].hier.levels ? systems_[s].hier.level_begin[l] : 0);
    d_dp_sys_.upload(meta, st);
    const int* m = d_dp_sys_.ptr;
    DpSystems sys{S, m, m + (S + 1), m + 2 * (S + 1), m + 3 * (S + 1), m + 3 * (S + 1) + S,
                  m + 3 * (S + 1) + 2 * S, m + 3 * (S + 1) + 3 * S};
    // going_next of every system: the device builder's array when it holds the
    // only system, else the host copies (downloaded with the hierarchies)
    const int* gn_all = nullptr;
    if(S == 1 && hier_dev_sys_ == 0 && hier_dev_)
        gn_all = hier_dev_->d_going_next();
    else
    {
        d_dp_gn_.alloc(std::max(gn_off[S], 1));
        for(int s = 0; s < S; ++s)
        {
            const auto& g = systems_[s].hier.going_next;
            if(!g.empty())
                check(hipMemcpyAsync(d_dp_gn_.ptr + gn_off[s], g.data(), g.size() * sizeof(int), hipMemcpyHostToDevice,
                                     st),
                      "going_next upload");
        }
        gn_all = d_dp_gn_.ptr;
    }

    // ---- full block CSR: stable radix sort of (row, column) keys in block order
    d_dp_rows_.alloc(std::max<long long>(nb, 1));
    d_dp_cols_.alloc(std::max<long long>(nb, 1));
    for(int s = 0; s < S; ++s)  // each system's structure at its block offset
    {
        const SystemDesc& d = systems_[s];
        if(d.nb <= 0) continue;
        const bool dev = d.rows_on_device;  // a device-held structure: a device copy, no upload
        check(hipMemcpyAsync(d_dp_rows_.ptr + P.src_off[s], dev ? (const void*)d_sys_rows_[s]->ptr : d.rows.data(),
                             sizeof(int) * d.nb, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
              "structure upload");
        check(hipMemcpyAsync(d_dp_cols_.ptr + P.src_off[s], dev ? (const void*)d_sys_cols_[s]->ptr : d.cols.data(),
                             sizeof(int) * d.nb, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
              "structure upload");
    }
    const int       bits    = bits_for((long long)n_rows + 1);  // the sentinel row n_rows fits
    const long long n_items = 2 * nb;
    d_dp_keys_.alloc(n_items);
    d_dp_keys2_.alloc(n_items);
    d_dp_vals_.alloc(n_items);
    d_dp_vals2_.alloc(n_items);
    d_dp_cnt_.alloc((size_t)n_rows + 1);
    d_diag_src_.alloc(n_rows);
    d_dp_err_.alloc(2);
    check(dp_fill(d_dp_cnt_.ptr, 0, (long long)n_rows + 1, st), "fill");
    check(dp_fill(d_diag_src_.ptr, -1, n_rows, st), "fill");
    check(dp_fill(d_dp_err_.ptr, 0, 2, st), "fill");
    check(dp_csr_items(nb, d_dp_rows_.ptr, d_dp_cols_.ptr, sys, d_e2i_.ptr, n_rows, bits, d_dp_keys_.ptr,
                       d_dp_vals_.ptr, d_dp_cnt_.ptr, d_diag_src_.ptr, d_dp_err_.ptr, st),
          "block CSR items");
    size_t tmp = std::max(dp_sort_bytes(n_items, 2 * bits), dp_scan_bytes((long long)n_rows + 1));
    d_dp_tmp_.alloc(tmp);
    check(dp_sort(d_dp_tmp_.ptr, tmp, d_dp_keys_.ptr, d_dp_keys2_.ptr, d_dp_vals_.ptr, d_dp_vals2_.ptr, n_items,
                  2 * bits, st),
          "block CSR sort");
    d_row_ptr_.alloc((size_t)n_rows + 1);
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_cnt_.ptr, d_row_ptr_.ptr, (long long)n_rows + 1, st), "row scan");
    const int* rb = read_ints({d_dp_err_.ptr, d_row_ptr_.ptr + n_rows}, st);
    if(rb[0]) throw std::invalid_argument("block index out of range");
    P.nnz = rb[1];
    d_col_.alloc(P.nnz);
    d_entry_src_.alloc(P.nnz);
    check(dp_csr_fill(P.nnz, d_dp_keys2_.ptr, d_dp_vals2_.ptr, bits, d_col_.ptr, d_entry_src_.ptr, st), "CSR fill");

    // ---- Galerkin maps, level by level over every system that has the level:
    // items sorted by (system, I, J) stably, runs of equal keys are the
    // coarse entries (the host's order: level, system, I, J, source)
    const long long cap = std::max<long long>(1, nb * std::max(1, max_levels - 1));
    d_gal_src_.alloc(cap);
    d_gal_ptr_.alloc(cap + 1);
    d_gal_EI_.alloc(cap);
    d_gal_EJ_.alloc(cap);
    d_gal_ES_.alloc(cap);
    check(dp_fill(d_gal_ptr_.ptr, 0, 1, st), "fill");
    P.gal_level_ptr.assign(1, 0);
    P.gal_level_src.assign(1, 0);
    P.gal_src_is_fine.clear();
    const int sbits = bits_for((long long)S + 1);  // a system field of all ones is the sentinel's only
    long long g_base = 0, e_base = 0, e_prev = 0, n_prev = nb;
    for(int lvl = 1; lvl < max_levels; ++lvl)
    {
        int lc = 1;
        for(const SystemDesc& d : systems_)
            if(lvl < d.hier.levels) lc = std::max(lc, d.hier.level_count[lvl]);
        const long long n  = lvl == 1 ? nb : n_prev;
        const int       kb = bits_for(lc);
        if(sbits + 2 * kb > 63) throw std::length_error("coarse level too large for the device Galerkin keys");
        d_dp_flag_.alloc(n + 1);
        d_dp_scan_.alloc(n + 1);
        tmp = std::max(dp_sort_bytes(n, sbits + 2 * kb), dp_scan_bytes(n + 1));
        d_dp_tmp_.alloc(tmp);
        check(dp_fill(d_dp_err_.ptr + 1, 0, 1, st), "fill");
        check(dp_gal_items(lvl, n, d_dp_rows_.ptr, d_dp_cols_.ptr, d_gal_EI_.ptr, d_gal_EJ_.ptr, d_gal_ES_.ptr,
                           (int)e_prev, sys, gn_all, m + lvl0 + (size_t)S * lvl, kb, d_dp_keys_.ptr, d_dp_vals_.ptr,
                           d_dp_err_.ptr + 1, st),
              "Galerkin items");
        check(dp_sort(d_dp_tmp_.ptr, tmp, d_dp_keys_.ptr, d_dp_keys2_.ptr, d_dp_vals_.ptr, d_dp_vals2_.ptr, n,
                      sbits + 2 * kb, st),
              "Galerkin sort");
        check(dp_run_ends(n, d_dp_keys2_.ptr, d_dp_flag_.ptr, st), "run ends");
        check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_flag_.ptr, d_dp_scan_.ptr, n + 1, st), "run scan");
        rb = read_ints({d_dp_scan_.ptr + n, d_dp_err_.ptr + 1}, st);
        const int n_new = rb[0], n_valid = rb[1];
        check(dp_gal_fill(n, d_dp_keys2_.ptr, d_dp_vals2_.ptr, d_dp_flag_.ptr, d_dp_scan_.ptr, (int)g_base,
                          (int)e_base, m + lvl0 + (size_t)S * lvl, kb, d_gal_src_.ptr, d_gal_ptr_.ptr, d_gal_EI_.ptr,
                          d_gal_EJ_.ptr, d_gal_ES_.ptr, st),
              "Galerkin fill");
        e_prev = e_base;
        n_prev = n_new;
        g_base += n_valid;
        e_base += n_new;
        P.gal_level_ptr.push_back((int)e_base);
        P.gal_level_src.push_back((int)g_base);
        P.gal_src_is_fine.push_back(lvl == 1 ? 1 : 0);
    }
    if(g_base >= (1ll << 29)) throw std::invalid_argument("Galerkin map too large");
    P.n_gal = e_base;
    // the last level's fill still reads the scan buffer that may be replaced below
    check(hipStreamSynchronize(st), "stream sync");

    // ---- bank entry lists: fine banks from their rows (every row is a fine
    // slot here), then the coarse entries with both ends in one bank, ranked
    // in entry order within their bank
    const int n_fine = P.n_fine_banks, n_banks = P.n_banks, n_gal = (int)P.n_gal;
    if((long long)n_fine * bank != n_rows) throw std::logic_error("fine banks do not cover the rows");
    d_dp_cnt_.alloc((size_t)n_rows + 1);
    d_dp_scan_.alloc((size_t)n_rows + 1);
    d_dp_flag_.alloc((size_t)n_gal + 1);
    d_dp_scan2_.alloc((size_t)n_gal + 1);
    d_dp_bcnt_.alloc((size_t)n_banks + 1);
    d_dp_bfirst_.alloc((size_t)n_banks + 1);
    d_ent_ptr_.alloc((size_t)n_banks + 1);
    tmp = std::max({dp_scan_bytes((long long)n_rows + 1), dp_scan_bytes((long long)n_gal + 1),
                    dp_scan_bytes((long long)n_banks + 1)});
    d_dp_tmp_.alloc(tmp);
    check(dp_fill(d_dp_bcnt_.ptr, 0, (long long)n_banks + 1, st), "fill");
    check(dp_fill(d_dp_bfirst_.ptr, INT_MAX, (long long)n_banks + 1, st), "fill");
    check(dp_ent_count_rows(n_rows, bank, d_row_ptr_.ptr, d_col_.ptr, d_entry_src_.ptr, d_dp_cnt_.ptr, st),
          "entry counts");
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_cnt_.ptr, d_dp_scan_.ptr, (long long)n_rows + 1, st), "entry scan");
    check(dp_ent_bank_counts(n_fine, bank, d_dp_scan_.ptr, d_dp_bcnt_.ptr, st), "bank counts");
    check(dp_ent_flag_coarse(n_gal, bank, n_fine, sys, d_gal_EI_.ptr, d_gal_EJ_.ptr, d_gal_ES_.ptr, d_dp_flag_.ptr,
                             d_dp_bcnt_.ptr, st),
          "coarse entry flags");
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_flag_.ptr, d_dp_scan2_.ptr, (long long)n_gal + 1, st), "coarse scan");
    check(dp_ent_first_coarse(n_gal, bank, n_fine, sys, d_gal_EI_.ptr, d_gal_ES_.ptr, d_dp_flag_.ptr, d_dp_scan2_.ptr,
                              d_dp_bfirst_.ptr, st),
          "coarse bank starts");
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_bcnt_.ptr, d_ent_ptr_.ptr, (long long)n_banks + 1, st), "bank scan");
    n_bank_ent_ = read_ints({d_ent_ptr_.ptr + n_banks}, st)[0];
    d_ent_pos_.alloc(n_bank_ent_);
    d_ent_src_.alloc(n_bank_ent_);
    check(dp_ent_fill_rows(n_rows, bank, d_row_ptr_.ptr, d_col_.ptr, d_entry_src_.ptr, d_dp_scan_.ptr, d_ent_pos_.ptr,
                           d_ent_src_.ptr, st),
          "fine entries");
    check(dp_ent_fill_coarse(n_gal, bank, n_fine, sys, d_gal_EI_.ptr, d_gal_EJ_.ptr, d_gal_ES_.ptr, d_dp_flag_.ptr,
                             d_dp_scan2_.ptr, d_dp_bfirst_.ptr, d_ent_ptr_.ptr, d_ent_pos_.ptr, d_ent_src_.ptr, st),
          "coarse entries");
    d_gal_src_.n = (size_t)std::max<long long>(g_base, 1);
    d_gal_ptr_.n = (size_t)P.n_gal + 1;
}

void Solver::build_spmv_layout_on_device(int cap, hipStream_t st)
{
    if(cap < 1 || cap > kSpmvMaxCap) throw std::invalid_argument("SpMV tile capacity out of range");
    const Plan& P = plan_;
    const int   n = P.n_rows, S = P.n_sys;
    if((unsigned)n > kSpmvColMask) throw std::length_error("too many rows for the SpMV layout");
    const int* row_off = d_dp_sys_.ptr + 2 * (S + 1);  // the systems' internal row ranges
    d_dp_nxt_.alloc(n);
    d_dp_kept_.alloc(n);
    d_dp_ts_.alloc(n);
    d_dp_err_.alloc(2);
    check(dp_fill(d_dp_err_.ptr, 0, 2, st), "fill");
    check(dp_sl_next(n, d_row_ptr_.ptr, d_col_.ptr, d_row_sys_.ptr, row_off, cap, d_dp_nxt_.ptr, d_dp_kept_.ptr, st),
          "tile ends");
    check(dp_sl_chain(n, d_dp_nxt_.ptr, d_dp_ts_.ptr, d_dp_err_.ptr + 1, st), "tile chain");
    const int T = read_ints({d_dp_err_.ptr + 1}, st)[0];
    d_dp_tk_.alloc((size_t)T + 1);
    d_dp_tq_.alloc((size_t)T + 1);
    d_dp_eb_.alloc((size_t)T + 1);
    d_dp_qb_.alloc((size_t)T + 1);
    const size_t tmp = dp_scan_bytes((long long)T + 1);
    d_dp_tmp_.alloc(tmp);
    check(dp_sl_tile_counts(T, n, d_dp_ts_.ptr, d_dp_kept_.ptr, d_row_ptr_.ptr, d_col_.ptr, d_dp_tk_.ptr, d_dp_tq_.ptr,
                            d_dp_err_.ptr, st),
          "tile counts");
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_tk_.ptr, d_dp_eb_.ptr, (long long)T + 1, st), "entry scan");
    check(dp_scan(d_dp_tmp_.ptr, tmp, d_dp_tq_.ptr, d_dp_qb_.ptr, (long long)T + 1, st), "product scan");
    const int* rb = read_ints({d_dp_eb_.ptr + T, d_dp_qb_.ptr + T, d_dp_err_.ptr}, st);
    if(rb[2]) throw std::logic_error("SpMV layout: a tile's transposed products exceed the column word's field");
    const long long aligned = rb[0], n_trans = rb[1];
    const long long total   = aligned + kSpmvSlack;
    if(total > INT_MAX || aligned < 0 || n_trans < 0) throw std::length_error("SpMV layout too large");
    d_sy_col_.alloc(total);
    d_sy_src_.alloc(total);
    d_sy_row_ptr_.alloc((size_t)n + 1);
    d_sy_tptr_.alloc((size_t)n + 1);
    d_wg_spmv_.alloc(T);
    d_wg_spmv_ent_.alloc(T);
    check(dp_sl_pad(total, d_sy_col_.ptr, d_sy_src_.ptr, st), "layout padding");
    check(dp_sl_fill(T, n, d_dp_ts_.ptr, d_dp_kept_.ptr, d_dp_eb_.ptr, d_dp_qb_.ptr, d_row_ptr_.ptr, d_col_.ptr,
                     d_entry_src_.ptr, d_row_sys_.ptr, d_sy_row_ptr_.ptr, d_sy_tptr_.ptr, d_sy_col_.ptr, d_sy_src_.ptr,
                     d_wg_spmv_.ptr, d_wg_spmv_ent_.ptr, d_dp_err_.ptr, st),
          "layout fill");
    SpmvLayout& L = spmv_layout_;
    L           = SpmvLayout{};
    L.cap       = cap;
    L.total     = total;
    L.n_in_tile = n_trans;
    // each system's tiles: the tile starts are in row order, systems in order
    L.sys_range.assign(2 * (size_t)S, 0);
    if(S == 1)
        L.sys_range = {0, T};
    else
    {
        std::vector<int> ts(T);
        if(T) check(hipMemcpyAsync(ts.data(), d_dp_ts_.ptr, sizeof(int) * T, hipMemcpyDeviceToHost, st), "tile starts");
        check(hipStreamSynchronize(st), "stream sync");
        for(int s = 0, t = 0; s < S; ++s)
        {
            L.sys_range[2 * s] = t;
            while(t < T && ts[t] < P.row_off[s + 1]) ++t;
            L.sys_range[2 * s + 1] = t;
        }
    }
    n_wg_spmv_ = T;
}

void Solver::download_plan_array(int which, std::vector<int>& out)
{
    if(!setup_done_) throw std::runtime_error("setup() must come first");
    const Plan& P = plan_;
    const int   T = n_wg_spmv_;
    const int*  src = nullptr;
    size_t      n   = 0;
    switch(which)
    {
    case 0: src = d_row_ptr_.ptr; n = (size_t)P.n_rows + 1; break;
    case 1: src = d_col_.ptr; n = (size_t)P.nnz; break;
    case 2: src = d_entry_src_.ptr; n = (size_t)P.nnz; break;
    case 3: src = d_diag_src_.ptr; n = (size_t)P.n_rows; break;
    case 4: src = d_gal_ptr_.ptr; n = (size_t)P.n_gal + 1; break;
    case 5: src = d_gal_src_.ptr; n = (size_t)P.gal_level_src.back(); break;
    case 6: src = d_ent_ptr_.ptr; n = (size_t)P.n_banks + 1; break;
    case 7: src = d_ent_pos_.ptr; n = (size_t)n_bank_ent_; break;
    case 8: src = d_ent_src_.ptr; n = (size_t)n_bank_ent_; break;
    case 9: src = d_sy_row_ptr_.ptr; n = (size_t)P.n_rows + 1; break;
    case 10: src = d_sy_col_.ptr; n = (size_t)spmv_layout_.total; break;
    case 11: src = d_sy_src_.ptr; n = (size_t)spmv_layout_.total; break;
    case 12: src = d_sy_tptr_.ptr; n = (size_t)P.n_rows + 1; break;
    case 13: src = reinterpret_cast<const int*>(d_wg_spmv_.ptr); n = 4 * (size_t)T; break;
    case 14: src = reinterpret_cast<const int*>(d_wg_spmv_ent_.ptr); n = 2 * (size_t)T; break;
    default: throw std::invalid_argument("unknown plan array");
    }
    out.assign(n, 0);
    if(n) check(hipMemcpy(out.data(), src, n * sizeof(int), hipMemcpyDeviceToHost), "plan array copy");
}
}  // namespace masx
