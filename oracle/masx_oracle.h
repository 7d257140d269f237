/*
 * masx oracle -- FP64 CPU restatement of StiffGIPC's Newton-step linear solve
 * (MAS-preconditioned PCG).  TEST INFRASTRUCTURE ONLY: the product library
 * never links, loads or calls anything in this directory.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only as the checker / the CPU timing reference.
 *
 * Every function cites the reference source it restates (paths relative to
 * the reference repository root).  Storage conventions follow the reference:
 *   - 3x3 blocks are Eigen::Matrix3d, i.e. 9 doubles in column-major order;
 *   - MAS local matrices are packed upper-triangular block storage with block
 *     index  bank*r - r*(r+1)/2 + c  for block row r <= block col c;
 *   - int4 contact pairs use the sign encoding of
 *     StiffGIPC/TraditionalMAS32Preconditioner.cu:1550-1960.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_sym_bcoo_spmv      20-row closed-form KAT (linear_system/utils/spmv.cu:301-459)
 *                          + dense brute force
 *   orc_morton_partition   bit-interleave identities + sortedness + permutation
 *   orc_mesh_neighbors     edge count closed form for Kuhn grids
 *   orc_num_levels /
 *   orc_hierarchy_capacity KATs of tests/runtime_options_tests.cpp:32-35 and C0/C1
 *   orc_mas_reorder        35-node KAT (TraditionalMAS32Preconditioner.cu:257-427)
 *                          + brute-force union-find per bank per level
 *   orc_mas_assemble       dense Galerkin P^T A P restricted to banks
 *   orc_mas_invert         ||A B - I|| checks and Sherman-Morrison closed forms
 *   orc_mas_apply          explicit dense sum_l P_l A_l^-1 P_l^T
 *   orc_pcg                exact diagonal solves, dense CG, and the 257-block
 *                          chain iteration counts held by the reference
 *                          (reports/agipc_reproduction/AGIPC_COARSE_MAS32_REPLAY.md:41-54)
 */
#ifndef MASX_ORACLE_H
#define MASX_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Threads of the oracle's parallel loops (OpenMP); results do not depend on
 * the count.  Test infrastructure and the CPU baseline's thread setting. */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* y = a*(U + U^T - D) x + b*y over block triplets (any triangle), where each
 * block (i,j) adds B x_j to y_i and, for i != j, B^T x_i to y_j.
 * linear_system/utils/spmv.cu:28-53 (legacy) and :55-118 (SRBK). */
void orc_sym_bcoo_spmv(int nb, const double* vals, const int* rows, const int* cols,
                       double a, const double* x, double b, double* y, int dof);

/* Morton 32-slot partition.  gl_main.cu:1176-1243.  Writes part_to_real
 * (length = returned padded size, -1 = padding) and real_to_part (length n). */
int orc_morton_partition(int n, const double* pos, int bank,
                         int* part_to_real, int* real_to_part);

/* Unique undirected tet/triangle edges -> per-vertex neighbor CSR in the
 * ordered-set order of load_mesh.cpp:934-1040.  Returns the list length, or
 * -1 if list_cap is too small. */
int orc_mesh_neighbors(int n, int ntet, const int* tets, int ntri, const int* tris,
                       unsigned* start, unsigned* num, unsigned* list, int list_cap);

/* TraditionalMAS32Preconditioner.cu:2142-2161 (and :2896-2899 for the max). */
int orc_num_levels(int vert_num, int part_map_size, int bank);
/* gipc/hierarchy_capacity.h:10-22 (returns -1 on overflow). */
long long orc_hierarchy_capacity(long long vertices, long long mapped,
                                 long long levels, long long bank);

typedef struct orc_mas orc_mas;

/* initPreconditioner_Neighbor + initPreconditioner_Matrix
 * (TraditionalMAS32Preconditioner.cu:2889-2977).  levels_override <= 0 uses
 * the reference level rule. Arrays are copied. */
orc_mas* orc_mas_create(int n_fem, int n_slots, int bank, int levels_override,
                        const unsigned* start, const unsigned* num, const unsigned* list,
                        const int* part_to_real, const int* real_to_part);
void orc_mas_destroy(orc_mas* m);

/* ReorderRealtime (TraditionalMAS32Preconditioner.cu:2189-2230 with the GROUP
 * kernels :463-1057 and :1550-1960).  Returns total cluster slots. */
int orc_mas_reorder(orc_mas* m, int cp_num, const int* pairs, int node_offset);

int orc_mas_levels(const orc_mas* m);
int orc_mas_total_clusters(const orc_mas* m);
/* level_count/level_begin: levels+1 entries; fine_mask: n_fem; prefix: n_slots/bank;
 * going_next: total_clusters (entries [0,n_fem) are indexed by real vertex id,
 * entries >= n_slots by cluster slot); coarse_table: n_fem*6. Any may be NULL. */
void orc_mas_get_hierarchy(const orc_mas* m, int* level_count, int* level_begin,
                           unsigned* fine_mask, int* prefix, int* going_next,
                           int* coarse_table);

/* PrepareHessian_bcoo scatter + aggregate (TraditionalMAS32Preconditioner.cu
 * :2234-2464).  indices may be NULL (= 0..count-1). */
int orc_mas_assemble(orc_mas* m, const double* vals, const int* rows, const int* cols,
                     const unsigned* indices, int offset, int count);

/* expand_symmetric_mas32 + __inverse6_P96x96 (method 0, no-pivot
 * Gauss-Jordan, :1062-1146) or the Cholesky-with-floor route of
 * pro/candidate-paper-t57-mas-fused-init/StiffGIPC/TraditionalMAS32Preconditioner.cu
 * :1382-1500 (method 1).  The inverse is rounded to FP32.  Returns 0. */
int orc_mas_invert(orc_mas* m, int method);

/* banks*nblk*9 doubles (col-major blocks) / floats. */
void orc_mas_get_local(const orc_mas* m, double* out);
void orc_mas_get_inverse(const orc_mas* m, float* out);
void orc_mas_set_inverse(orc_mas* m, const float* in);

/* preconditioning(R, Z) (:2684-2725): restrict, banked local solve with the
 * packed inverse (lower blocks = transposed upper), prolong.  r, z: n_fem*3. */
void orc_mas_apply(const orc_mas* m, const double* r, double* z);

/* Gauss-Jordan with partial pivoting, row-major n x n
 * (cuda_tools/cuda_eigen.h:129-165).  Returns 0. */
int orc_inverse_gj_pivot(int n, const double* a, double* inv);

/* ABD 12x12 block-Jacobi: gathers each body's diagonal 12x12 from the blocks
 * with row/4 == col/4 and inverts it (row-major 144 doubles per body).
 * abd_system/abd_system_function/setup_abd_system_gradient_and_hessian.cu:407-441. */
void orc_abd_block_inverse(int nbody, int nb, const double* vals, const int* rows,
                           const int* cols, double* inv);

/* Diagonal 3x3 block-Jacobi inverse (identity where no diagonal block),
 * row-major 9 doubles per block row.  linear_system/preconditioner/diag_preconditioner.cu:10-91. */
void orc_diag3_inverse(int block_rows, int nb, const double* vals, const int* rows,
                       const int* cols, double* inv);

typedef struct
{
    int    iterations;
    int    failure;        /* 0 ok, 1 nonpositive rz, 2 bad curvature, 3 nonfinite, 4 cap */
    double rz0;
    double rz_last;
    double rec_rel_res;    /* ||r_recursive|| / ||b|| */
    double true_rel_res;   /* ||b - A x|| / ||b|| */
} orc_pcg_stats;

enum { ORC_RULE_REF_RTZ = 0, ORC_RULE_TRUE_REL = 1 };

/* PCG with the reference's composition of preconditioners
 * (global_linear_system.cu:293-306): z = r (or 3x3 block-Jacobi if
 * global_diag), then the ABD range [0, 4*n_abd) block rows gets the 12x12
 * inverse, then the FEM range [fem_offset, fem_offset+n_fem) gets the MAS.
 * rule REF_RTZ: linear_system/solver/pcg_solver.cu:192-326 (break on
 * |rz_prev| <= tol*rz0, returns the SpMV count k).  rule TRUE_REL:
 * agipc/agipc_coarse_replay.cu:120-160 (break on ||r||^2 <= tol^2 ||b||^2,
 * returns the update count).  max_iter <= 0 uses floor(0.3*dof). */
int orc_pcg(int block_rows, int nb, const double* vals, const int* rows, const int* cols,
            const double* b, double* x, int global_diag, int n_abd, const double* abd_inv,
            const orc_mas* mas, int fem_offset, double tol, int rule, int max_iter,
            orc_pcg_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
