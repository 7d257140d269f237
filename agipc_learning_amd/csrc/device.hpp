// Device-side data layout shared by the kernels and the host orchestrator.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace masx
{

enum Status : int
{
    kRunning    = 0,
    kDone       = 1,
    kCandidate  = 2,  // recursive residual met the tolerance; awaiting a true-residual check
    kRestarting = 3,  // r := b - A x is pending
    kInit       = 4,  // first restriction pending
};

enum Failure : int
{
    kOk = 0, kBadRz = 1, kBadCurvature = 2, kNonFinite = 3, kIterationCap = 4
};

enum Rule : int
{
    kRuleRefRtz  = 0,  // pcg_solver.cu:277 -- |rz_prev| <= tol * rz0, one-iteration lag
    kRuleTrueRel = 1,  // ||r|| <= tol ||b||, FP64 true residual confirms, restart on drift
};

enum Precision : int
{
    kPrecFp32 = 0,
    kPrecFp16 = 1,
    kPrecBf16 = 2,
};

// One per system; 128-byte aligned so systems never share a cache line.
// Within one kernel launch no field is both read and written: each kernel's
// first workgroup of a system writes only fields that kernel does not read.
struct alignas(128) SysState
{
    double rz_prev;   // r.z of the previous iteration (beta denominator): written by update
    double rz_cur;    // r.z of this iteration: written by the SpMV kernel
    double rz0, pAp, alpha, beta, rr, bb, true_rr, tol;
    int    k, status, failure, rule;
    int    max_iter, k_check, check_every, beta_zero;
    int    need_check, restarts, n_true_checks, frozen;
    int    rejected;  // candidates whose true residual missed the tolerance (each ends in a restart)
};

enum WgType : int { kWgPlain = 0, kWgFemMas = 1, kWgAbd = 2 };

constexpr int kBlock      = 256;  // threads per workgroup of the row kernels
constexpr int kBodiesPerWg = 21;  // 21 bodies x 12 DOF rows = 252 threads

struct DevData
{
    // sizes
    int n_rows, n_sys, bank, n_banks, n_fine_banks;
    int mas_waves;                      // waves per workgroup of the FP32 MAS local kernel (4 or 8)
    int spmv_nt;                        // 1: the SpMV reads the matrix with nontemporal loads
    int spmv_bodies;                    // 1: the SpMV layout has tiles of affine bodies
    int factor_waves;                   // waves per bank of the MAS factor (1 or 2)
    const unsigned short* inv_map;      // stored position of a bank inverse -> element (launch_inverse_map)
    int single;                         // 1: one system (system index 0, whole work tables)
    int n_wg_pre, n_wg_spmv, n_wg_upd;  // workgroups of the iteration kernels' tables
    long long nnz;          // entries of the full block CSR
    long long nnz_s, plane; // entries of the FP32 SpMV layout; padded entry stride of its value planes
    // structure: full block CSR (FP64 paths)
    const int* row_ptr;
    const int* col;
    const int* entry_src;   // full entry -> upper block | transposed<<31
    const int* row_sys;
    // FP32 SpMV layout (tile-local symmetric pairs stored once; plan.hpp)
    const int* sy_row_ptr;  // n_rows+1
    const int* sy_col;      // column word: >= 0 outside the tile, else tile-local code
    const int* sy_src;      // entry -> upper block | transposed<<31
    const int* sy_tptr;     // n_rows+1: cumulative transposed products
    // values
    const float*  vals32;   // 9 planes of `plane` floats over the SpMV layout's entries, block row-major (r*3+c)
    const double* vals64;   // upper blocks, col-major (r + 3c), global upper order
    // vectors (float3 per internal row)
    float*  x;
    float*  r;
    float*  z;
    float*  p0;
    float*  p1;
    float*  Ap;
    float*  rt;             // true residual (restart source)
    double* b64;
    float*  Rc;             // coarse restriction, float3 per coarse node
    float*  Zc;
    // MAS
    const void*     minv;   // n_banks banks in diagonal-major layout of the storage precision
    const float*    minv_scale;  // per bank (half-precision storage is scaled to [-1, 1])
    const int*      bank_vec;
    const int*      bank_sys;
    const int*      bank_kind;   // 0 level 0, 1 level 1, 2 level >= 2
    const int*      comp_ptr;
    const uint32_t* comp_mask;
    const int*      comp_parent;
    const int*      deep_ptr;    // level >= 2 coarse node -> its update-group partial slots
    const int*      deep_list;
    const int*      bank_ugroup; // fine bank -> update group
    const int4*     ugroup;      // {slot base, slot count, component begin, component end}
    const int*      comp_dslot;  // 4 per fine-bank component: group-local slot per deep level, -1 none
    float*          part_deep;   // float3 per slot, written by the update kernel
    const int*      row_l1;      // per internal row: its level-1 node, -1 none
    const int*      anc;         // 8 per coarse node (used for level-1 nodes): itself and its ancestors, -1 padded
    // ABD and diagonal block-Jacobi
    const float* abd_inv;   // 144 per body, row-major
    const int*   body_row;
    // 12 x 12 body couplings of the SpMV's body tiles: per body its couplings
    // [abd_cp_ptr[i], abd_cp_ptr[i+1]), column body abd_cp_col, values 144
    // floats per coupling (row-major over the 12 scalar rows and columns)
    const int*   abd_cp_ptr;
    const int*   abd_cp_col;
    const int*   abd_cp_row;   // per coupling: first internal row of its column body
    const float* vals_abd;
    const float* diag_inv;  // 9 per internal row, row-major (identity where unused)
    const int*   row_mode;  // per internal row: 0 identity, 1 diag inverse
    // work distribution: int4 {sys, aux, begin, end}
    const int4* wg_pre;
    const int4* wg_spmv;
    const int4* wg_upd;
    const int2* wg_spmv_ent;    // per SpMV tile: its entry range [e0, e1)
    const int4* wg_upd_grp;     // per update workgroup: its update group (zeros for non-MAS rows)
    const int*  wg_pre_range;   // per system [begin,end) of its workgroups in wg_pre (2 ints)
    const int*  wg_spmv_range;
    const int*  wg_upd_range;
    double*     part_pre;       // per-workgroup partial sums, consumed by the next kernel
    double*     part_spmv;
    double*     part_upd;
    double*     part_tr;
    double*     part_drift; // per SpMV workgroup: |r_true - r|^2 of its rows at a check
    SysState*   state;
};

}  // namespace masx
