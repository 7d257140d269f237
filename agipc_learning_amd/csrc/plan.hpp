// Host-side structure analysis for the MAS-PCG solve.
//
// Everything here depends only on the sparsity structure, the mesh topology,
// the contact pairs and the vertex partition -- not on matrix values.  It runs
// once per structure ("new graph"); value refreshes with an unchanged
// structure reuse it (the reference's refresh_fixed_graph_bcoo path,
// TraditionalMAS32Preconditioner.cu:2877-2887).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "spmv_layout.hpp"

namespace masx
{

// --------------------------------------------------------------- hierarchy --
// Multilevel additive-Schwarz hierarchy over fixed-size banks of slots
// (TraditionalMAS32Preconditioner.cu:2142-2230 defines the algorithm: level 0
// banks are consecutive slots of the vertex partition; the connected
// components inside each bank -- through mesh edges and contact pairs --
// become the nodes of the next level, numbered by (bank, lowest lane)).
struct Hierarchy
{
    int bank   = 32;
    int levels = 0;
    int n_fem  = 0;
    int n_slots = 0;
    std::vector<int>      level_count;   // levels+1 (index 0 unused = n_slots)
    std::vector<int>      level_begin;   // levels+1, slot offset of each level
    int                   total = 0;     // level_begin[levels]
    std::vector<uint32_t> fine_mask;     // per real vertex: component lanes in its level-0 bank
    std::vector<int>      prefix;        // components per level-0 bank
    std::vector<int>      going_next;    // [0,n_fem): real vertex -> level-1 slot;
                                         // [n_slots,total): coarse slot -> next-level slot
    std::vector<int>      coarse_table;  // n_fem*6: ancestor slots at levels 1..levels-1
    // per bank (all levels but the top): components as (lane mask, parent slot)
    std::vector<int>      comp_ptr;      // total/bank + 1
    std::vector<uint32_t> comp_mask;
    std::vector<int>      comp_parent;   // slot at the next level
    bool                  tables_pending = false;  // fine_mask, prefix, coarse_table still only on the device
};

int  reference_level_count(int vert_num, int part_map_size, int bank);
// pairs: int4 sign-encoded collision pairs (global vertex ids), node_offset is
// subtracted (TraditionalMAS32Preconditioner.cu:1550-1960).
void build_hierarchy(Hierarchy& h, int n_fem, int n_slots, int bank, int levels,
                     const uint32_t* nbr_start, const uint32_t* nbr_num, const uint32_t* nbr_list,
                     const int* part_to_real, const int* real_to_part, const int* pairs, int n_pairs,
                     int node_offset);
int  decode_pair(const int* p, int v[4]);
void morton_partition(int n, const double* pos, int bank, std::vector<int>& part_to_real,
                      std::vector<int>& real_to_part);

// ---------------------------------------------------------------- systems --
struct SystemDesc
{
    int  block_rows = 0;   // external block rows (= 4*n_abd + n_fem for FEM at the end)
    long long nb    = 0;   // upper 3x3 blocks
    int  n_abd      = 0;   // affine bodies occupying block rows [0, 4*n_abd)
    int  fem_offset = 0;   // first FEM block row
    int  n_fem      = 0;   // FEM vertices [fem_offset, fem_offset + n_fem)
    bool use_mas    = false;
    bool global_diag = false;
    int  bank       = 32;
    int  levels     = 0;   // <= 0: reference rule
    // host copies of structure (empty while rows_on_device: the solver then
    // holds them on the device, for the device structure analysis)
    std::vector<int>      rows, cols;
    bool                  rows_on_device = false;
    std::vector<uint32_t> nbr_start, nbr_num, nbr_list;
    std::vector<int>      part_to_real, real_to_part;
    std::vector<int>      pairs;        // int4 per pair
    int                   node_offset = 0;
    long long             topo_version = 0;  // bumped by every new topology (neighbours, partition)
    Hierarchy             hier;
    bool                  hier_ready = false;  // hier already built (on the device) for this structure
};

// Global packed layout over a batch of systems (everything the kernels index).
struct Plan
{
    int bank = 32;
    int n_sys = 0;
    // rows: internal block rows, per system [ABD rows padded to bank][FEM slots]
    std::vector<int> row_off;        // n_sys+1 internal row offsets
    std::vector<int> ext_off;        // n_sys+1 external row offsets (concatenated b/x)
    std::vector<int> ext_to_int;     // per external row (global) -> internal row (global)
    int n_rows = 0;                  // internal rows total
    int n_ext  = 0;
    std::vector<int> fem_row0;       // per system: first internal FEM row (global)
    std::vector<int> abd_row0;       // per system: first internal ABD row (global)
    std::vector<int> row_sys;        // per internal row: system
    // full (both-triangle) block CSR over internal rows
    std::vector<int> row_ptr;        // n_rows+1
    std::vector<int> col;            // nnz
    std::vector<long long> src_off;  // n_sys+1: offset of each system's upper blocks in the
                                     // concatenated value array
    std::vector<int> entry_src;      // per full entry: global upper block | transposed flag (bit 31)
    long long nnz = 0, nb_total = 0;
    // MAS banks: fine banks first (all systems), then coarse banks
    int n_fine_banks = 0, n_banks = 0, n_coarse = 0;   // n_coarse = coarse slots total
    std::vector<int> bank_vec;       // per bank: float3 offset into fine vector or coarse array
    std::vector<int> bank_sys;
    std::vector<int> coarse_base;    // per system: global coarse index of its level-1 begin
    std::vector<int> comp_ptr;       // per bank (global) -> components
    std::vector<uint32_t> comp_mask;
    std::vector<int> comp_parent;    // global coarse index
    std::vector<int> coarse_parent;  // per global coarse node: parent global coarse index or -1
    std::vector<int> bank_kind;      // 0 level 0 (R from r), 1 level 1 (R in Rc), 2 level >= 2 (R from level-1 sums)
    std::vector<int> desc_ptr;       // per global coarse node (level >= 2): range of its level-1 descendants
    std::vector<int> desc_list;
    std::vector<int> anc;            // per global coarse node, 8 ints: itself and its ancestors, -1 padded
    std::vector<int> row_l1;         // per internal row: its level-1 node (global coarse index), -1 none;
                                     // the node's ancestors are its anc list
    // level >= 2 residuals pre-reduced per update group (256 rows of one system)
    std::vector<int> bank_ugroup;    // per fine bank: its update group (-1 if none)
    std::vector<int> ugroup;         // per group, 4 ints: slot base, slot count, component begin, end
    std::vector<int> comp_dslot;     // per fine-bank component, 4 ints: group-local slot of its ancestor
                                     // at levels 2..5, -1 if none
    std::vector<int> deep_ptr;       // per global coarse node: range of its partial slots in deep_list
    std::vector<int> deep_list;
    int n_deep_slots = 0;
    // restriction schedule for levels >= 2 (per system, coarse banks grouped by level)
    std::vector<int> sys_level_banks;   // flattened: per system, per level l>=1, [bank begin, end)
    std::vector<int> sys_level_ptr;     // n_sys+1 offsets into sys_level_banks (pairs)
    // Galerkin: coarse operator entries E (all systems, all levels >= 1), 3x3 row-major FP64
    std::vector<int> gal_level_ptr;     // per "global level step" boundaries in E (launch groups)
    std::vector<int> gal_ptr;           // per E entry: range in gal_src
    std::vector<int> gal_src;           // source index << 2 | mode (0 as is, 1 transpose, 2 sym)
    std::vector<int> gal_src_is_fine;   // per launch group: 1 = sources are fine upper blocks
    std::vector<int> gal_level_src;     // per launch group boundary: gal_ptr there (source offsets)
    long long n_gal = 0;
    // per bank entry lists for the dense local assembly
    std::vector<int> bank_ent_ptr;      // n_banks+1
    std::vector<int> bank_ent_pos;      // packed block position | transposed << 16
    std::vector<int> bank_ent_src;      // fine upper block index (fine banks) or E index (coarse)
    // ABD
    int n_bodies = 0;
    std::vector<int> body_row;          // internal row of each body's first block row
    std::vector<int> body_src;          // 16 per body: upper block index of (q,t) q<=t in the 4x4
                                        // grid, with transposed flag (bit 31), -1 if absent
    std::vector<int> body_sys;
    // diagonal 3x3 inverse (global block-Jacobi): per internal row, diagonal source or -1
    std::vector<int> diag_src;
    std::vector<int> row_needs_diag;    // 1 if no local preconditioner covers the row
    // The block CSR (row_ptr, col, entry_src), diag_src, the Galerkin maps
    // (gal_ptr, gal_src) and the bank entry lists are built on the device
    // (devplan.hpp) and exist only there; their sizes (nnz, n_gal,
    // gal_level_ptr, gal_level_src) are filled in by that build.
    bool structure_on_device = false;
};

// structure_on_device: MAS systems without affine bodies whose hierarchies
// are ready; the host analysis then skips what the device builds
void build_plan(Plan& plan, std::vector<SystemDesc>& systems, bool structure_on_device = false);
// Bounds of every index array the kernels dereference; throws on violation.
void validate_plan(const Plan& plan);

// ------------------------------------------------------------ SpMV layout --
// The FP32 SpMV's matrix layout.  Rows are cut into tiles of consecutive rows
// of one system (one thread per row, at most kSpmvTileRows); a tile of several
// rows holds at most `cap` entries.  A symmetric pair whose two rows share a
// tile is stored once, as its upper block: it yields A x_c for row r and
// A^T x_r for row c, both from the tile's own x.  Every other off-diagonal
// block keeps both copies, one in each row, so each row is still summed by
// the tile that owns it.  Row r of a tile starting at r0 keeps its entries
// with column < r0 or >= r, in column order.
//
// Entry column words and tile limits: spmv_layout.hpp.

// Rows of affine bodies get tiles of their own (up to kBodiesPerTile bodies):
// their couplings to bodies of the same system are dense 12 x 12 blocks
// (cp_*), the rest of their entries (couplings to FEM vertices) 3 x 3 entries
// of the layout as usual.  Per system: first internal row of its bodies,
// their count and the global index of its first body.
struct SpmvBodies
{
    std::vector<int> abd_row0, n_abd, body_base;
};
constexpr int kBodiesPerTile = 5;   // 5 bodies x 12 scalar rows x 4 coupling groups = 240 threads
constexpr int kSpmvBodyTile  = 2;   // the work-table type of a body tile (the kernels' kWgAbd)

struct SpmvLayout
{
    int cap = 0;
    std::vector<int> tiles;       // 4 per tile: system, type << 28 | aux (body tiles: kWgAbd, first body),
                                  // first row, end row
    std::vector<int> tile_ent;    // 2 per tile: entry range [e0, e1)
    std::vector<int> sys_range;   // 2 per system: tile range
    std::vector<int> row_ptr;     // n_rows+1
    std::vector<int> col;         // column word per entry (padding and slack included)
    std::vector<int> src;         // per entry: upper block | transposed flag (bit 31), as entry_src;
                                  // kSpmvPadSrc for padding and slack
    std::vector<int> tptr;        // n_rows+1: cumulative transposed products
    long long total = 0;          // entry slots (col.size(): padding and slack included)
    long long n_entries = 0;      // stored matrix entries (no padding)
    long long n_in_tile = 0;      // off-diagonal pairs stored once
    // 12 x 12 body couplings: body i (global) has couplings [cp_ptr[i],
    // cp_ptr[i+1]) in increasing column body cp_col (global), each with, per
    // 3 x 3 sub-block (q, t) (q row, t column within the bodies), the first
    // full-CSR entry of row 4i+q with column 4j+t (-1 if none; duplicate
    // blocks follow it in the row)
    std::vector<int> cp_ptr, cp_col, cp_body, cp_sub;
};

// From the full block CSR (rows sorted by column) and its entry sources.
// Throws if a tile's codes would not fit their fields.
void build_spmv_layout(SpmvLayout& L, int n_sys, const std::vector<int>& row_off, const std::vector<int>& row_ptr,
                       const std::vector<int>& col, const std::vector<int>& entry_src, int cap,
                       const SpmvBodies* bodies = nullptr);
// y = A x with the layout (FP64, host), from FP64 upper blocks (col-major),
// read the way the kernel reads it: each thread's run of cap/256 consecutive
// entries summed per row (restarting at row-start flags), the runs' ends
// summed by the owning row, then its transposed products.  For tests of the
// layout.
// Body tiles add their 12 x 12 couplings, read through the full CSR
// (row pointers, columns, entry sources) and the first row of every body.
void spmv_layout_apply(const SpmvLayout& L, const double* upper, const double* x, double* y,
                       const int* full_row_ptr = nullptr, const int* full_col = nullptr,
                       const int* full_src = nullptr, const int* body_row = nullptr);

}  // namespace masx
