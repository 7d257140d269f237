/*
 * masx -- MI355X-native MAS-preconditioned conjugate gradient (C ABI).
 *
 * Drop-in boundary for StiffGIPC's Newton-step linear solve.  Plain C types
 * only; device pointers are HIP device pointers; every call returns 0 on
 * success and a nonzero code on failure, with the message available from
 * masx_last_error() (the reference aborts instead: cuda_tools/cuda_tools.h:38-143).
 * Streams are explicit (hipStream_t passed as void*; NULL = the library's
 * own stream, synchronous for the caller).
 *
 * Matrix contract (linear_system/linear_system/global_matrix.h:18-203): 3x3
 * block triplets, 9 doubles per block in Eigen column-major order, symmetric
 * semantics y += B x_j at (i,j) and, for i != j, B^T x_i at (j,i)
 * (linear_system/utils/spmv.cu:28-53); block rows ordered [4 per affine body]
 * [FEM vertices] (GIPC.cu:10271-10272).
 */
#ifndef MASX_H
#define MASX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct masx_solver masx_solver;
typedef struct masx_mas    masx_mas;

enum masx_precision { MASX_FP32 = 0, MASX_FP16 = 1, MASX_BF16 = 2 };
enum masx_rule      { MASX_RULE_REF_RTZ = 0, MASX_RULE_TRUE_REL = 1 };
enum masx_status    { MASX_RUNNING = 0, MASX_DONE = 1 };
/* Failure of a stopped solve (cf. the device PCG statuses of
 * pro/candidate-paper-t57-mas-fused-init/StiffGIPC/agipc/agipc_device_pcg.cuh:12-13):
 * BAD_RZ: r.z not finite, or not positive under the TRUE rule (InvalidRz);
 * BAD_CURVATURE: p.Ap not finite, or not positive under the TRUE rule
 * (NonfiniteCurvature / NonpositiveCurvature); NONFINITE: ||b||, ||r|| or the
 * true residual not finite (NonfiniteResidual); ITERATION_CAP: max_iter
 * reached (IterationCap), also reported when the host stops launching work
 * before a system has stopped.  Vectors are left as the last complete
 * iteration wrote them. */
enum masx_failure   { MASX_OK = 0, MASX_BAD_RZ = 1, MASX_BAD_CURVATURE = 2, MASX_NONFINITE = 3,
                      MASX_ITERATION_CAP = 4 };

typedef struct masx_system_desc
{
    int       block_rows;   /* 4 * n_abd + n_fem */
    long long nb;           /* number of 3x3 block triplets */
    int       n_abd;        /* affine bodies: block rows [0, 4*n_abd) get the 12x12 block-Jacobi */
    int       n_fem;        /* FEM vertices: block rows [4*n_abd, block_rows) */
    int       use_mas;      /* 1: MAS preconditioner on the FEM range */
    int       global_diag;  /* 1: 3x3 block-Jacobi on rows no local preconditioner covers */
    int       bank;         /* MAS subdomain size: 32 (MAS32) or 16 (CEMAS16) */
    int       levels;       /* MAS levels, <= 0 for the reference rule */
} masx_system_desc;

typedef struct masx_solve_params
{
    int    rule;         /* MASX_RULE_REF_RTZ: |r.z| <= tol * r0.z0 with the reference's one-iteration
                            lag (linear_system/solver/pcg_solver.cu:277);
                            MASX_RULE_TRUE_REL: ||b - A x|| <= tol ||b||, FP64 true residual */
    double tol;          /* 1e-4 in the reference (gipc/gipc.cu:77-79) */
    int    max_iter;     /* <= 0: floor(0.3 * dof) (linear_system/solver/pcg_solver.h:15) */
    int    check_every;  /* TRUE rule: true-residual recompute period (restart on 10x tol drift) */
    int    chunk;        /* iterations per captured hipGraph launch (even, >= 2) */
    int    use_graph;    /* 1: replay a captured graph per chunk */
} masx_solve_params;

typedef struct masx_solve_stats
{
    int    iterations;   /* SpMV count, the value PCGSolver::solve returns */
    int    status;       /* masx_status */
    int    failure;      /* masx_failure */
    int    restarts;
    int    true_checks;
    double rz0;
    double rz;
    double rel_res_recursive;
    double rel_res_true; /* ||b - A x|| / ||b|| in FP64 */
    int    rejected;     /* convergence candidates (recursive residual at the tolerance) whose FP64
                            true residual missed it; each one restarts the recursion from b - A x */
} masx_solve_stats;

const char* masx_last_error(void);
int         masx_device_info(char* name, int name_len, int* cu_count, long long* hbm_bytes);
void        masx_default_solve_params(masx_solve_params* p);

/* ---------------------------------------------------------------------------
 * Global solver (GlobalLinearSystem + PCGSolver).  A solver holds a batch of
 * independent systems; a single Newton system is a batch of one.
 * -------------------------------------------------------------------------*/

/* GlobalLinearSystem ctor (linear_system/linear_system/global_linear_system.h:14-119) */
int masx_create(masx_solver** out);
int masx_destroy(masx_solver* s);

/* Structure of one system: host rows/cols of the upper block triplets after
 * Converter::convert (linear_system/utils/converter.cu:137-156).  Returns the
 * system index in *sys. */
int masx_add_system(masx_solver* s, const masx_system_desc* desc, const int* rows, const int* cols, int* sys);

/* Replace system `sys`'s upper block structure (nb blocks, host rows/cols as
 * in masx_add_system), e.g. when a Newton step brings a new contact set: the
 * structure is analysed again at the next masx_setup, on the same solver, so
 * its host and device buffers are reused.  No counterpart in the reference,
 * whose BCOO is rebuilt each step by Converter::convert
 * (converter.cu:137-284) before MASPreconditioner's set-up. */
int masx_set_structure(masx_solver* s, int sys, long long nb, const int* rows, const int* cols);
/* masx_set_structure from device arrays (d_rows, d_cols: nb ints each, as
 * the reference's BCOO keeps its structure on the GPU in
 * m_block_row_indices / m_block_col_indices, global_matrix.h:26-27,
 * read through block_row_indices() / block_col_indices(), :93-107): copied to the
 * library on `stream` (the arrays may be reused once the stream passes this
 * call; NULL: the call returns after the copy).  The device structure
 * analysis (masx_set_device_structure) reads the copy without a host round
 * trip; an analysis on the host downloads it.  Index range errors are
 * reported by the next masx_setup. */
int masx_set_structure_device(masx_solver* s, int sys, long long nb, const int* d_rows, const int* d_cols,
                              void* stream);
/* MAS topology: neighbor CSR (load_mesh.cpp:934-1040) and the 32-slot vertex
 * partition (gl_main.cu:1201-1243), host arrays, as uploaded by
 * gl_main.cu:1502-1536 before initPreconditioner_Matrix. */
int masx_set_topology(masx_solver* s, int sys, const uint32_t* nbr_start, const uint32_t* nbr_num,
                      const uint32_t* nbr_list, int list_len, const int* part_to_real, int part_map_size,
                      const int* real_to_part);
/* Same, with the partition computed from vertex positions (host array) by
 * Morton order; meshes of 4,096 vertices or more are sorted on the device
 * when one is present (masx_morton_partition_device: identical arrays). */
int masx_set_topology_morton(masx_solver* s, int sys, const uint32_t* nbr_start, const uint32_t* nbr_num,
                             const uint32_t* nbr_list, int list_len, const double* positions);
/* The Morton partition masx_set_topology_morton uses (gl_main.cu:1176-1243),
 * host arrays: part_to_real (part_cap >= n rounded up to the bank, -1 in
 * padding slots), real_to_part (n). */
int masx_morton_partition(int n, const double* positions, int bank, int* part_to_real, int part_cap,
                          int* real_to_part);
/* Contact pairs (int4 per pair, sign-encoded types, global vertex ids;
 * node_offset is subtracted), the _collisonPairs of GIPC.cuh:49-51 consumed by
 * TraditionalMAS32Preconditioner.cu:1550-1960.  Host array. */
int masx_set_contacts(masx_solver* s, int sys, const int* pairs, int n_pairs, int node_offset);

/* Host-only structure analysis: hierarchy (ReorderRealtime,
 * TraditionalMAS32Preconditioner.cu:2189-2230), internal slot order, block
 * layout, Galerkin maps, with a bounds check of every kernel index.  Needs no
 * device. */
int masx_analyze(masx_solver* s);
/* masx_analyze (if not done) + device allocation and upload. */
int masx_setup(masx_solver* s);

/* Block values of one system (device, FP64 col-major, same order as rows/cols). */
int masx_set_values(masx_solver* s, int sys, const double* d_values, void* stream);
/* Host-side variant of masx_set_values. */
int masx_set_values_host(masx_solver* s, int sys, const double* h_values, void* stream);
/* masx_set_values for n systems at once (systems[q] takes d_values[q]; both
 * arrays on the host, the values on the device): one hand-off to the
 * library's stream instead of one per system, and one copy launch per 64
 * systems instead of one copy per system, for batches.  No counterpart
 * in the reference (one system per solve there). */
int masx_set_values_many(masx_solver* s, int n, const int* systems, const double* const* d_values, void* stream);

/* Preconditioner assembly: GlobalLinearSystem::build_linear_system's
 * preconditioner step (global_linear_system.cu:37-146), i.e.
 * ABDPreconditioner::assemble (abd_preconditioner.cu:33),
 * TraditionalMAS32_Preconditioner::assemble (traditional_mas32_preconditioner.cu:20-31),
 * DiagPreconditioner::assemble (diag_preconditioner.cu:79-86). */
int masx_build(masx_solver* s, int precision, void* stream);

/* IterativeSolver::solve(x, b) (linear_system/linear_system/i_linear_system_solver.h:43-44)
 * -> PCGSolver::solve (pcg_solver.cu:160-173).  b, x: device FP64, systems
 * concatenated in order.  stats: one per system (may be NULL). */
int masx_solve(masx_solver* s, const double* d_b, double* d_x, const masx_solve_params* params,
               masx_solve_stats* stats, void* stream);

/* IterativeSolver::spmv(a, x, b, y) (i_linear_system_solver.h:46), FP64. */
int masx_spmv(masx_solver* s, double a, const double* d_x, double b, double* d_y, void* stream);

/* IterativeSolver::apply_preconditioner(z, r) (i_linear_system_solver.h:52-53)
 * -> GlobalLinearSystem::apply_preconditioner (global_linear_system.cu:293-306). */
int masx_apply_preconditioner(masx_solver* s, const double* d_r, double* d_z, void* stream);

/* SpMV tiling: a workgroup owns whole rows (at most 256) holding at most
 * 256 * entries_per_thread block entries (2, 4 or 8); a longer row gets a
 * workgroup of its own. */
int masx_set_spmv_tiling(masx_solver* s, int entries_per_thread);
/* Times every supported SpMV tiling on the built system (`reps` launches
 * each) and keeps the fastest, then times the other load policy
 * (masx_set_spmv_loads) on it and keeps it if faster; reports the tiling and
 * the time per launch (ms).  Leaves the iteration state frozen: call
 * masx_solve afterwards as usual. */
int masx_autotune(masx_solver* s, int reps, int* entries_per_thread, double* ms);
/* How the SpMV reads the matrix (value planes, column words, body
 * couplings), which it streams once per iteration: mode 1 nontemporal loads
 * (they pass the caches without evicting the vectors), 2 ordinary loads (a
 * matrix the 256 MB Infinity Cache holds is then re-read from it in the next
 * iteration), 0 (the default) nontemporal once the matrix exceeds 256 MiB.
 * Results are identical under every mode.  No counterpart in the reference
 * (its SpMV uses ordinary loads). */
int masx_set_spmv_loads(masx_solver* s, int mode);
/* *nontemporal: whether the SpMV currently uses nontemporal loads. */
int masx_spmv_loads_used(masx_solver* s, int* nontemporal);
/* MAS local solve variant.  waves = 4 or 8: the VALU kernel, that many
 * waves per workgroup sharing the block diagonals of its banks; 5 or 9: the
 * same with 4 or 8 waves, the workgroup's FP32 inverses staged through LDS
 * with 16-byte loads (half precision runs 4 or 8); 1: the MFMA kernel, one
 * wave per bank (FP16 / BF16 inverses of 32-slot banks; other cases run 4);
 * 0 (the default): from 16,384 banks on MFMA for half-precision 32-slot
 * banks and 9 for FP32, else 4.  The two kernels store a half-precision inverse in
 * different layouts: a change between them needs a new masx_build before the
 * next solve (masx_solve reports the missing build). */
int masx_set_mas_waves(masx_solver* s, int waves);
/* MAS factor (the dense assembly and FP64 inverse of every bank at
 * masx_build): waves = 1, one wave per bank (more banks per SIMD; the faster
 * on many banks), 2, two waves per bank (a bank's inverse sooner; the faster
 * on a few thousand), 0 (the default): 1 from 16,384 banks, else 2.  Takes
 * effect at the next masx_build; the inverses agree to FP64 rounding.  No
 * counterpart in the reference (__inverse6_P96x96,
 * TraditionalMAS32Preconditioner.cu:1099-1147, has one form). */
int masx_set_factor_waves(masx_solver* s, int waves);
/* Times every MAS local variant of the built precision (`reps` launches
 * each; for half-precision 32-slot banks 1, 4 and 8, else 4 and 8), keeps the
 * fastest (rebuilding the preconditioner for a layout change) and reports it
 * and its time per launch (ms).  Leaves the iteration state frozen like
 * masx_autotune. */
int masx_autotune_mas(masx_solver* s, int reps, int* waves, double* ms);
/* Average execution time of one kernel (0 MAS local, 1 preconditioner finish,
 * 2 fused SpMV, 3 update) over `reps` back-to-back launches, milliseconds:
 * each launch is timed by events stamped by its own dispatch, so the figure
 * excludes the gaps between launches. */
int masx_time_kernel(masx_solver* s, int which, int reps, double* ms);

/* Where a new structure is analyzed (ReorderRealtime + PrepareHessian_bcoo,
 * TraditionalMAS32Preconditioner.cu:2189-2230, 2234-2525, which rebuild it on
 * the GPU every Newton step): on = 1 (the default) builds the block CSR, the
 * diagonal sources, the Galerkin maps, the bank entry lists and the SpMV tile
 * layout on the device when every system of the solver is a MAS system
 * without affine bodies (one system or a batch; the hierarchies are always
 * built there); 0 builds them on the host.
 * Takes effect at the next masx_setup after a structure change.  The two give
 * the same arrays bit for bit.  *used: whether the current set-up built them
 * on the device. */
int masx_set_device_structure(masx_solver* s, int on);
int masx_device_structure_used(masx_solver* s, int* used);
/* Copy of an internal structure array (for tests; no reference counterpart):
 * which = 0 row_ptr, 1 col, 2 entry_src, 3 diag_src, 4 gal_ptr, 5 gal_src,
 * 6 bank_ent_ptr, 7 bank_ent_pos, 8 bank_ent_src, 9..12 the SpMV layout's
 * row_ptr, col, src, tptr, 13 its work table (4 ints per tile), 14 its tile
 * entry ranges (2 ints per tile).  *n receives the length; out may be null
 * to query it. */
int masx_plan_array(masx_solver* s, int which, int* out, long long capacity, long long* n);

/* Practical streaming rate of the current GPU (no counterpart in the
 * reference; a measurement aid): the fastest of `reps` launches of a plain
 * kernel over `bytes` of device memory (>= 1 MiB, a multiple of 16), kind 0 =
 * read-only stream (bytes read; nontemporal 16-byte loads, a contiguous span
 * per workgroup), kind 1 = copy (bytes read + written; a grid-stride loop of
 * ordinary 16-byte loads and stores), in GB/s.  Allocates and frees its
 * buffers. */
int masx_stream_rate(long long bytes, int kind, int reps, double* gb_per_s);

/* Introspection (tests, benchmark). */
typedef struct masx_plan_info
{
    int       n_rows, n_ext, n_sys, bank;
    long long nnz, nb_total;
    int       n_banks, n_fine_banks, n_coarse, n_bodies;
    long long n_gal, inverse_bytes;
} masx_plan_info;
int masx_plan_info_get(masx_solver* s, masx_plan_info* info);
/* Hierarchy of one system, same arrays as the oracle: level_count and
 * level_begin (levels+1), fine_mask (n_fem), prefix (n_slots/bank), going_next
 * (total; [0,n_fem) by real vertex, [n_slots,total) by slot), coarse_table
 * (n_fem*6).  Any pointer may be NULL; *levels and *total are always set. */
int masx_get_hierarchy(masx_solver* s, int sys, int* levels, int* total, int* level_count, int* level_begin,
                       uint32_t* fine_mask, int* prefix, int* going_next, int* coarse_table);
/* Packed bank inverses (FP32 view), bank order = fine banks of all systems,
 * then coarse banks; within a bank, packed upper 3x3 blocks, row-major. */
int masx_get_inverse(masx_solver* s, float* out, long long capacity);
/* Global bank index of a system's slot (fine slots and coarse slots). */
int masx_bank_of_slot(masx_solver* s, int sys, int slot, int* bank);
/* Number of Cholesky pivots floored or non-finite in the last build. */
int masx_factor_status(masx_solver* s, long long* count);

/* Host-only hierarchy build (no device needed): the same computation
 * masx_setup runs for each MAS system, exposed for CPU-side verification.
 * Call once with going_next == NULL to obtain *levels and *total, then again
 * with buffers of level_count/level_begin (levels+1), fine_mask (n_fem),
 * prefix (n_slots/bank), going_next (going_next_cap >= max(total, n_fem)),
 * coarse_table (n_fem*6). */
int masx_hierarchy_host(int n_fem, int n_slots, int bank, int levels, const uint32_t* nbr_start,
                        const uint32_t* nbr_num, const uint32_t* nbr_list, const int* part_to_real,
                        const int* real_to_part, const int* pairs, int n_pairs, int node_offset, int* out_levels,
                        int* out_total, int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix,
                        int* going_next, long long going_next_cap, int* coarse_table);

/* ---- Morton vertex partition on the device ------------------------------
 * gl_main.cu:1176-1243 (the mesh reordering the MAS banks are cut from) on
 * the GPU: the same result as masx_morton_partition bit for bit, from device
 * positions (3 doubles per vertex) into device arrays part_to_real
 * (pad(n, bank) ints, -1 in the padding; part_cap >= that) and real_to_part
 * (n ints).  Asynchronous on `stream`; scratch is kept in the object and
 * grows as needed. */
typedef struct masx_morton masx_morton;
int masx_morton_create(masx_morton** out);
int masx_morton_destroy(masx_morton* m);
int masx_morton_partition_device(masx_morton* m, int n, const double* d_positions, int bank, int* d_part_to_real,
                                 long long part_cap, int* d_real_to_part, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): masx_morton_partition_device with the same arguments and
 * results, minus the object, that allocates nothing.  Its scratch is a static
 * device array of the library, 8.7 MB on every GPU it runs on, held until the
 * process ends; n <= MASX_MORTON_NOALLOC_MAX, larger n is an error.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (other
 * streams, other host threads) race. */
#define MASX_MORTON_NOALLOC_MAX 524288
int masx_morton_partition_device_noalloc(int n, const double* d_positions, int bank, int* d_part_to_real,
                                         long long part_cap, int* d_real_to_part, void* stream);

/* Check of the FP32 SpMV's tile layout on the host (no device): from the
 * upper block structure (rows[k] <= cols[k], unique) of one system of n_rows
 * block rows, build the full block CSR, then the layout for tiles of
 * 256 * entries_per_thread entries (pairs inside a tile stored once), and
 * apply both to seeded random upper blocks and x in FP64.  Outputs: the
 * largest |y_layout - y_full| over max |y_full|, the layout's entry count and
 * its pairs stored once.  Not part of the solve; the layout itself has no
 * counterpart in the reference (MASPreconditioner.cu and the BCOO SpMV of
 * pcg_solver.cu store and multiply the full matrix). */
int masx_spmv_layout_check(int n_rows, long long nb, const int* rows, const int* cols, int entries_per_thread,
                           unsigned seed, double* max_rel_err, long long* n_entries, long long* n_in_tile);
/* The same check with n_bodies affine bodies in rows [0, 4 n_bodies): their
 * rows get body tiles whose couplings to bodies are 12 x 12 blocks (the
 * layout's ABD path; rows 4i..4i+3 are body i, GIPC.cu:10271-10272). */
int masx_spmv_layout_check_bodies(int n_rows, int n_bodies, long long nb, const int* rows, const int* cols,
                                  int entries_per_thread, unsigned seed, double* max_rel_err, long long* n_entries,
                                  long long* n_in_tile);

/* ---- Stable Neo-Hookean element Hessians and their parameter gradient ---
 * The FEM part of the Newton matrix (femEnergy.cu: first Piola stress
 * :441-459 with rates u = 4 mu/3, r = lambda + 5 mu/6 (gl_main.cu:586-590),
 * PSD projection by the analytic eigensystem :1641-1717, element matrix
 * vol * PFPX^T H PFPX :1719-1855, here without the dt^2 factor: the system is
 * M/dt^2 + K) assembled on the device, and the backward kernel of an adjoint
 * solve: for H(theta) x = b and H y = dL/dx, dL/dtheta_e = -y_e^T (dK_e /
 * dtheta_e) x_e, with the derivative of the PSD projection in closed form.
 * create: host tets (4 block-row ids each) and the host upper block structure
 * (rows[k] <= cols[k], the order of the values); every tet edge and vertex
 * needs its block.  Element inputs (device, per tet): dminv (9, row-major
 * Dm^-1), vol, F (9, row-major), mu, lambda.  assemble: vals (nb x 9,
 * col-major blocks) = base (may be NULL) + sum of the element matrices, in a
 * fixed order (deterministic).  param_grad: x, y are FP64 block vectors (3
 * per block row); writes one gradient per tet. */
typedef struct masx_fem masx_fem;
int masx_fem_create(masx_fem** out, int n_tets, const int* tets, long long nb, const int* rows, const int* cols);
int masx_fem_destroy(masx_fem* f);
int masx_fem_assemble(masx_fem* f, const double* d_dminv, const double* d_vol, const double* d_F, const double* d_mu,
                      const double* d_lambda, const double* d_base, double* d_vals, void* stream);
int masx_fem_param_grad(masx_fem* f, const double* d_dminv, const double* d_vol, const double* d_F,
                        const double* d_mu, const double* d_lambda, const double* d_x, const double* d_y,
                        double* d_grad_mu, double* d_grad_lambda, void* stream);

/* ---- MAS hierarchy on the device ---------------------------------------
 * TraditionalMAS32Preconditioner::ReorderRealtime
 * (TraditionalMAS32Preconditioner.cu:2189-2230: BuildConnectMaskL0 :1963,
 * PreparePrefixSumL0 :1991, BuildLevel1 :2014, BuildConnectMaskLx :2055,
 * NextLevelCluster :2091, ComputeNextLevel :2101, PrefixSumLx :2112,
 * AggregationKernel :2130, BuildCollisionConnection :2163; kernels :463-1057
 * and :1550-1960) run on the device.  Inputs are device arrays: nbr_start and
 * nbr_num (n_fem), nbr_list (list_len), part_to_real (n_slots), real_to_part
 * (n_fem), pairs (4 * n_pairs ints, sign-encoded; node_offset subtracted).
 * The result equals masx_hierarchy_host bit for bit and stays on the device
 * (masx_hier_device); masx_hier_download copies it out.  The build
 * synchronizes `stream` once, at the end, to return the level sizes.  An
 * inconsistent partition or an out-of-range neighbour id is an error (no
 * kernel then writes the outputs).  Device scratch grows as needed and is
 * kept in the object. */
typedef struct masx_hier masx_hier;
int masx_hier_create(masx_hier** out);
int masx_hier_destroy(masx_hier* h);
int masx_hier_build(masx_hier* h, int n_fem, int n_slots, int bank, int levels, const uint32_t* d_nbr_start,
                    const uint32_t* d_nbr_num, const uint32_t* d_nbr_list, long long list_len,
                    const int* d_part_to_real, const int* d_real_to_part, const int* d_pairs, int n_pairs,
                    int node_offset, int* out_levels, int* out_total, int* out_components, void* stream);
/* Host copies; any pointer may be NULL.  level_count (levels+1, [0] = 0 as in
 * masx_hierarchy_host), level_begin (levels+1), fine_mask (n_fem), prefix
 * (n_slots/bank), going_next (max(total, n_fem)), coarse_table (n_fem*6), and
 * the per-bank component lists of levels 0..levels-2: comp_ptr (total/bank + 1), comp_mask and comp_parent
 * (*out_components each; masks in leader-lane order, parent = next-level slot). */
int masx_hier_download(masx_hier* h, int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix,
                       int* going_next, int* coarse_table, int* comp_ptr, uint32_t* comp_mask, int* comp_parent,
                       void* stream);
/* Device views of the result (valid until the next build or destroy). */
typedef struct masx_hier_arrays
{
    const int*      going_next;
    const int*      coarse_table;
    const uint32_t* fine_mask;
    const int*      prefix;
    const int*      comp_ptr;
    const uint32_t* comp_mask;
    const int*      comp_parent;
} masx_hier_arrays;
int masx_hier_device(masx_hier* h, masx_hier_arrays* out);

/* ---- triplet -> BCOO conversion ---------------------------------------
 * Converter::convert (linear_system/utils/converter.cu:137-284): n triplets
 * (device int32 rows, int32 cols, col-major Matrix3d values) become unique
 * blocks sorted by (row, col) with duplicates summed in input order;
 * drop_lower != 0 also removes blocks with row > col (Converter::ge2sym,
 * :286-336).  Outputs are device arrays with capacity n; *n_unique receives
 * the count (the call synchronizes `stream`, as the reference does to read
 * h_unique_key_number).  Indices outside [0, n_block_rows) are an error.
 * Scratch is kept in the converter object and grows as needed. */
typedef struct masx_converter masx_converter;
int masx_converter_create(masx_converter** out);
int masx_converter_destroy(masx_converter* c);
int masx_convert_triplets(masx_converter* c, int n, int n_block_rows, const int* d_rows, const int* d_cols,
                          const double* d_vals, int drop_lower, int* d_out_rows, int* d_out_cols,
                          double* d_out_vals, int* n_unique, void* stream);

/* ---------------------------------------------------------------------------
 * MAS preconditioner object (TraditionalMAS32Preconditioner /
 * MASPreconditioner API, TraditionalMAS32Preconditioner.cuh:78-116), for a
 * caller that keeps its own PCG and BCOO matrix.
 * -------------------------------------------------------------------------*/

int masx_mas_create(masx_mas** out, int bank);
int masx_mas_destroy(masx_mas* m);  /* FreeMAS */
/* initPreconditioner_Neighbor (TraditionalMAS32Preconditioner.cu:2889-2926) */
int masx_mas_init_neighbor(masx_mas* m, int vert_num, int collision_node_offset, int total_neighbor_num,
                           int part_map_size);
/* the host uploads of gl_main.cu:1502-1536 */
int masx_mas_upload(masx_mas* m, const uint32_t* neighbor_list, const uint32_t* neighbor_start,
                    const uint32_t* neighbor_num, const int* part_to_real, const int* real_to_part);
/* initPreconditioner_Matrix (:2928-2977); levels <= 0 keeps the reference rule */
int masx_mas_init_matrix(masx_mas* m, int levels);
/* setPreconditioner_bcoo (:2643-2681): device values/rows/cols/indices of the
 * global BCOO, FEM block offset, selected count, device contact pairs. */
int masx_mas_set_preconditioner_bcoo(masx_mas* m, const double* d_values, const int* d_rows, const int* d_cols,
                                     const uint32_t* d_indices, int offset, int triplet_num, const int* d_pairs,
                                     int cp_num, int precision, void* stream);
/* preconditioning(R, Z) (:2684-2725): device double3 arrays of the FEM range. */
int masx_mas_preconditioning(masx_mas* m, const double* d_R, double* d_Z, void* stream);
/* the solver the MAS object drives (introspection) */
masx_solver* masx_mas_solver(masx_mas* m);

#ifdef __cplusplus
}
#endif
#endif
