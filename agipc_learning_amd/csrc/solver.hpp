// Host orchestrator: a batch of independent linear systems solved by
// MAS-preconditioned CG on one GPU.  Owns all device memory of the batch.
#pragma once
#include <hip/hip_runtime.h>

#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "hierarchy.hpp"
#include "plan.hpp"

namespace masx
{

template <typename T>
struct DeviceArray
{
    T*     ptr = nullptr;
    size_t n   = 0;
    size_t cap = 0;   // allocated elements (a smaller alloc() reuses the buffer)
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() { release(); }
    void release()
    {
        if(ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n   = 0;
        cap = 0;
    }
    void alloc(size_t count);
    void upload(const std::vector<T>& h, hipStream_t st);
};

struct SolveParams
{
    int    rule        = kRuleTrueRel;
    double tol         = 1e-4;
    int    max_iter    = 0;     // <= 0: floor(0.3 * dof) per system (pcg_solver.h:15)
    int    check_every = 50;    // true-residual recompute period (TRUE rule)
    int    chunk       = 8;     // iterations per captured graph (even)
    int    use_graph   = 1;
};

struct SolveStats
{
    int    iterations;
    int    status;
    int    failure;
    int    restarts;
    int    true_checks;
    double rz0, rz;
    double rel_res_recursive;
    double rel_res_true;
    int    rejected;  // convergence candidates whose FP64 true residual missed the tolerance
};

class Solver
{
  public:
    Solver();
    ~Solver();

    int  add_system(const SystemDesc& d);
    SystemDesc& system(int s) { return systems_.at(s); }
    // system s's MAS hierarchy with every table on the host
    const Hierarchy& hierarchy(int s);
    int  n_systems() const { return (int)systems_.size(); }
    void analyze();                    // host-only structure analysis (no device needed)
    // analyze (if needed; the MAS hierarchies are then built on the device)
    // + allocation and upload
    void setup();
    // topology or contacts of system s changed: the next setup() re-analyzes
    void structure_changed(int s);
    // system s's upper structure from device arrays (copied on the library's
    // stream, ordered after `user`); the device structure analysis reads the
    // copy, the host analysis downloads it if it is ever needed
    void set_structure_device(int s, long long nb, const int* d_rows, const int* d_cols, hipStream_t user);
    bool is_setup() const { return setup_done_; }
    bool is_analyzed() const { return analyzed_; }
    void set_values(int s, const double* d_values, hipStream_t st);      // device FP64 col-major
    // set_values for several systems, one stream hand-off for all of them
    void set_values_many(int n, const int* systems, const double* const* d_values, hipStream_t st);
    void set_values_host(int s, const double* h_values, hipStream_t st);
    void gather_values(int s, const double* d_values, const uint32_t* d_indices, hipStream_t st);
    void build(int precision, hipStream_t st);
    void solve(const double* d_b, double* d_x, const SolveParams& prm, std::vector<SolveStats>* stats,
               hipStream_t st);
    void spmv(double a, const double* d_x, double b, double* d_y, hipStream_t st);
    void apply_preconditioner(const double* d_r, double* d_z, hipStream_t st);

    // introspection for tests and the benchmark
    const Plan& plan() const { return plan_; }
    int  precision() const { return precision_; }
    void download_inverse(std::vector<float>& out);  // FP32 view of the packed bank inverses
    void download_coarse_entries(std::vector<double>& out);
    size_t inverse_bytes() const;
    void set_spmv_tiling(int entries_per_thread);
    struct Tiling
    {
        int    entries;  // SpMV entries per thread per chunk
        double ms;
    };
    // times every supported SpMV tiling on the built system and keeps the
    // fastest, then the other load policy on it
    Tiling autotune_spmv(int reps);
    // SpMV matrix loads: 1 nontemporal, 2 ordinary, 0 automatic (nontemporal
    // once the streamed matrix outgrows kSpmvCachedBytes)
    void set_spmv_loads(int mode);
    bool spmv_nontemporal() const { return effective_spmv_nt(); }
    // MAS factor: 1 or 2 waves per bank, 0 automatic (1 from kFactorWaveBanks
    // banks on); takes effect at the next build()
    void set_factor_waves(int waves);
    int  factor_waves() const { return effective_factor_waves(); }
    // MAS local kernel: 4 or 8 waves per workgroup (VALU; each workgroup holds
    // 64 / bank banks, its waves split their block diagonals), 5 or 9 (the
    // same with 4 or 8 waves, the FP32 inverse staged through LDS with
    // 16-byte loads; half precision falls back to 4 or 8), 1 (the MFMA
    // kernel, one wave per bank: FP16 / BF16 inverses of 32-slot banks; other
    // cases fall back to 4), or 0 (automatic: from kMfmaAutoBanks banks on
    // MFMA for half-precision 32-slot banks and 9 for FP32, else 4).  A change that moves the
    // half-precision inverse between the two layouts needs a new build().
    void set_mas_waves(int waves);
    int  mas_waves() const { return effective_mas_waves(); }
    // times every MAS local variant the built precision has and keeps the
    // fastest (rebuilding when the variant's inverse layout differs)
    Tiling autotune_mas(int reps);
    // average device time of one kernel of the iteration over `reps` launches
    // (0 MAS local, 1 preconditioner finish, 2 SpMV, 3 update), milliseconds
    double time_kernel(int which, int reps);
    long long factor_status() const;
    // Structure analysis of a new structure: 1 (default) builds the block
    // CSR, Galerkin maps, bank entry lists and SpMV layout on the device when
    // every system of the solver is a MAS system without bodies; 0 always on
    // the host.  Takes effect at the next analysis.
    void set_device_structure(int on) { dev_struct_mode_ = on ? 1 : 0; }
    bool device_structure_used() const { return dev_struct_; }
    // host copy of a device plan array (tests): 0 row_ptr, 1 col, 2 entry_src,
    // 3 diag_src, 4 gal_ptr, 5 gal_src, 6 bank_ent_ptr, 7 bank_ent_pos,
    // 8 bank_ent_src, 9 SpMV row_ptr, 10 SpMV col, 11 SpMV src, 12 SpMV tptr,
    // 13 SpMV work table (4 ints per tile), 14 SpMV tile entry ranges (2 per tile)
    void download_plan_array(int which, std::vector<int>& out);

  private:
    void build_work_tables();
    void make_devdata();
    hipStream_t enter(hipStream_t user);  // work stream, ordered after `user`
    void leave(hipStream_t user);         // orders `user` after the work stream, or waits
    void run_chunk(hipStream_t st, int chunk, bool with_check);
    void launch_iteration(hipStream_t st, int i);

    std::vector<SystemDesc> systems_;
    Plan plan_;
    bool setup_done_ = false;
    bool analyzed_   = false;
    void ensure_stream();
    int  precision_  = kPrecFp32;
    int  entries_    = 4;
    int  mas_waves_  = 0;
    int  effective_mas_waves() const;
    int  spmv_loads_ = 0;
    bool effective_spmv_nt() const;
    int  factor_waves_ = 0;
    int  effective_factor_waves() const;
    bool built_      = false;

    // device
    DeviceArray<int>       d_row_ptr_, d_col_, d_entry_src_, d_row_sys_, d_e2i_, d_i2e_;
    SpmvLayout             spmv_layout_;
    DeviceArray<int>       d_sy_row_ptr_, d_sy_col_, d_sy_src_, d_sy_tptr_;
    DeviceArray<int>       d_cp_ptr_, d_cp_col_, d_cp_row_, d_cp_body_, d_cp_sub_;  // 12 x 12 body couplings
    DeviceArray<float>     d_vals_abd_;
    DeviceArray<unsigned short> d_inv_map_;  // the current inverse layout's position -> element map
    std::vector<std::unique_ptr<DeviceArray<int>>> d_sys_rows_, d_sys_cols_;  // device-held structures (set_structure_device)
    void download_device_structures();  // for the host analysis
    int inv_map_key_ = -1;
    DeviceArray<float>     d_vals32_;
    DeviceArray<double>    d_vals64_, d_E_, d_b64_, d_tmp64a_, d_tmp64b_;
    DeviceArray<float>     d_x_, d_r_, d_z_, d_p0_, d_p1_, d_Ap_, d_rt_, d_Rc_, d_Zc_, d_part_deep_;
    DeviceArray<char>      d_minv_;
    DeviceArray<float>     d_minv_scale_;
    DeviceArray<int>       d_bank_vec_, d_bank_sys_, d_bank_kind_, d_comp_ptr_, d_comp_parent_;
    DeviceArray<uint32_t>  d_comp_mask_;
    DeviceArray<int>       d_deep_ptr_, d_deep_list_, d_row_l1_, d_anc_, d_bank_ugroup_, d_ugroup_, d_comp_dslot_;
    DeviceArray<int>       d_gal_ptr_, d_gal_src_, d_ent_ptr_, d_ent_pos_, d_ent_src_;
    DeviceArray<float>     d_abd_inv_, d_diag_inv_;
    DeviceArray<int>       d_body_row_, d_body_src_, d_diag_src_, d_needs_diag_, d_row_mode_;
    DeviceArray<int4>      d_wg_pre_, d_wg_spmv_, d_wg_upd_, d_wg_upd_grp_;
    DeviceArray<int2>      d_wg_spmv_ent_;
    DeviceArray<int>       d_wg_pre_range_, d_wg_spmv_range_, d_wg_upd_range_;
    DeviceArray<double>    d_part_pre_, d_part_spmv_, d_part_upd_, d_part_tr_, d_part_drift_;
    DeviceArray<SysState>  d_state_;
    DeviceArray<int>       d_max_iter_, d_factor_status_;
    // device MAS hierarchy build of setup(): inputs and the builder's scratch
    void build_hierarchies_on_device();
    int  hier_dev_sys_    = -1;     // system whose hierarchy hier_dev_ holds
    struct TopoCache  // one system's mesh topology and partition on the device
    {
        DeviceArray<uint32_t> start, num, list;
        DeviceArray<int>      p2r, r2p;
        long long             version = -1;  // the system's topo_version they hold
    };
    std::vector<std::unique_ptr<TopoCache>> topo_;
    int  dev_struct_mode_ = 1;
    bool dev_struct_      = false;  // the current plan's structure arrays were built on the device
    bool device_structure_eligible() const;
    void build_structure_on_device(hipStream_t st);
    void build_spmv_layout_on_device(int cap, hipStream_t st);
    int  n_bank_ent_ = 0;
    int* h_dp_ = nullptr;  // pinned read-back slots of the device structure build
    const int* read_ints(std::initializer_list<const int*> ptrs, hipStream_t st);
    DeviceArray<int>       d_dp_rows_, d_dp_cols_, d_dp_vals_, d_dp_vals2_, d_dp_cnt_, d_dp_err_, d_dp_flag_,
                           d_dp_scan_, d_dp_scan2_, d_dp_bcnt_, d_dp_bfirst_, d_gal_EI_, d_gal_EJ_, d_gal_ES_, d_dp_nxt_,
                           d_dp_kept_, d_dp_ts_, d_dp_tk_, d_dp_tq_, d_dp_eb_, d_dp_qb_, d_dp_sys_, d_dp_gn_;
    DeviceArray<uint64_t>  d_dp_keys_, d_dp_keys2_;
    DeviceArray<char>      d_dp_tmp_;
    DeviceArray<int>       d_h_pairs_;
    std::unique_ptr<DeviceHierarchy> hier_dev_;
    int n_wg_pre_ = 0, n_wg_spmv_ = 0, n_wg_upd_ = 0;
    DevData dd_{};

    hipStream_t work_ = nullptr;
    hipEvent_t  ev_a_ = nullptr, ev_b_ = nullptr;
    // build(): the Galerkin sums, the coarse banks' factor, the SpMV value
    // packing and the block-Jacobi factors run on side_ (default priority)
    // beside the fine banks' MAS factor
    hipStream_t side_ = nullptr;
    hipEvent_t  ev_fork_ = nullptr, ev_join_ = nullptr;
    std::vector<hipEvent_t> timing_ev_;    // time_kernel: one start/stop pair per launch
    static constexpr int kDepth = 2;       // graph chunks in flight during a solve
    hipEvent_t  chunk_ev_[kDepth] = {};
    SysState*   h_state_ = nullptr;        // pinned: kDepth + 1 slots of n_sys states
    int         h_state_cap_ = 0;          // systems it holds
    int*        h_export_ = nullptr;       // coherent pinned: unfinished-system count after each chunk
    int*        d_export_ = nullptr;       // its device address
    long long   exports_ = 0;              // status exports launched so far (host mirror of the sequence)
    DeviceArray<int> d_export_seq_;        // device sequence counter of the exports
    int         last_iters_ = 0;           // iterations of the previous solve (launch-ahead hint)
    hipGraphExec_t graph_exec_ = nullptr;
    int graph_chunk_ = 0;
    bool graph_check_ = false;
};

}  // namespace masx
