// Structure analysis of MAS systems on the device: the full block CSR, the
// diagonal sources, the Galerkin maps of every coarse level, the per-bank
// entry lists of the local assembly and the FP32 SpMV's tile layout -- the
// arrays the host analysis (plan.cpp: build_plan, build_spmv_layout) makes
// for the same structure, bit for bit, without a host round trip of the
// matrix structure (TraditionalMAS32Preconditioner.cu:2189-2230 and
// :2234-2525 rebuild the same information on the GPU every Newton step).
//
// Every launcher below is asynchronous on `st`; sizes the host needs are read
// back by the caller.  Index widths: rows, columns and block indices are
// int32 (the host analysis rejects larger systems).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace masx
{
// Scratch bytes of one stable radix sort of n (uint64 key, int value) pairs
// over key bits [0, end_bit), and of one exclusive int scan of n items.
size_t dp_sort_bytes(long long n, int end_bit);
size_t dp_scan_bytes(long long n);
hipError_t dp_sort(void* tmp, size_t tmp_bytes, const uint64_t* keys_in, uint64_t* keys_out, const int* vals_in,
                   int* vals_out, long long n, int end_bit, hipStream_t st);
hipError_t dp_scan(void* tmp, size_t tmp_bytes, const int* in, int* out, long long n, hipStream_t st);

hipError_t dp_fill(int* p, int value, long long n, hipStream_t st);

// Systems of a batch (concatenated; one system is a batch of one): per
// system s, its upper blocks [src_off[s], src_off[s+1]) of the concatenated
// rows/cols (system-local external indices), its external rows [ext_off[s],
// ext_off[s+1]), its internal rows [row_off[s], row_off[s+1]), its MAS
// levels, the offset of its going_next in the concatenation gn_all, its
// first coarse node (global) and its fine slots.
struct DpSystems
{
    int        n_sys;
    const int* src_off;
    const int* ext_off;
    const int* row_off;
    const int* levels;
    const int* gn_off;
    const int* coarse_base;
    const int* n_slots;
};

// Block CSR items: upper block k of system s (external rows[k], cols[k] in
// [0, n_ext_s)) becomes key (R << bits | C) with value k in row R =
// e2i[ext_off[s] + rows[k]] and, off the diagonal, key (C << bits | R) with
// value k | bit 31 in row C; a diagonal block's second slot is the sentinel
// row n_rows.  Counts entries per row (row_cnt, zeroed), records the last
// diagonal block of each row (diag_src, filled with -1) and flags an index
// out of range in *err.
hipError_t dp_csr_items(long long nb, const int* rows, const int* cols, DpSystems sys, const int* e2i, int n_rows,
                        int bits, uint64_t* keys, int* vals, int* row_cnt, int* diag_src, int* err, hipStream_t st);
// col / entry_src of the first nnz sorted items
hipError_t dp_csr_fill(long long nnz, const uint64_t* keys, const int* vals, int bits, int* col, int* entry_src,
                       hipStream_t st);

// Galerkin items of coarse level `lvl` (>= 1) of every system that has it:
// level 1 from the n upper blocks (rows/cols, FEM-local vertices), deeper
// levels from the previous level's n entries EI/EJ/ES [e_prev, e_prev + n)
// (system-local slots, system); going_next (gn_all at the system's offset)
// maps a vertex (level 1) or a coarse slot to its next-level slot.  Key
// (s << 2 bits | (I - lb_s) << bits | (J - lb_s)), lb_s = lvl_begin[s], value
// (source << 2 | mode) with mode 0 as is, 1 transposed, 2 a block whose two
// ends merged (B + B^T); a system without the level gives the sentinel key
// (all ones), sorted last.  *n_valid counts the other items.
hipError_t dp_gal_items(int lvl, long long n, const int* rows, const int* cols, const int* EI, const int* EJ,
                        const int* ES, int e_prev, DpSystems sys, const int* gn_all, const int* lvl_begin, int bits,
                        uint64_t* keys, int* vals, int* n_valid, hipStream_t st);
// flag[x] = 1 where sorted item x ends a run of equal keys other than the
// sentinel (flag[n] = 0)
hipError_t dp_run_ends(long long n, const uint64_t* keys, int* flag, hipStream_t st);
// the level's sources and entries: gal_src[g_base + x] = vals[x] for the
// valid items; run end x with run index q = qpos[x] gives entry e_base + q =
// (lb_s + I, lb_s + J) of system s and gal_ptr[e_base + q + 1] = g_base + x + 1
hipError_t dp_gal_fill(long long n, const uint64_t* keys, const int* vals, const int* flag, const int* qpos,
                       int g_base, int e_base, const int* lvl_begin, int bits, int* gal_src, int* gal_ptr, int* EI,
                       int* EJ, int* ES, hipStream_t st);

// Bank entry lists.  Fine banks: the as-is entries of each row of the full
// CSR whose column lies in the row's bank (row_cnt per row, then written at
// row_off[r]); coarse banks: the coarse entries with both ends in one bank,
// in entry order (flag per entry; bank_cnt per bank, atomically counted).
hipError_t dp_ent_count_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                             int* row_cnt, hipStream_t st);
hipError_t dp_ent_fill_rows(int n_slots, int bank, const int* row_ptr, const int* col, const int* entry_src,
                            const int* row_off, int* pos, int* src, hipStream_t st);
// a coarse entry of system s lies in bank n_fine_banks + (coarse_base[s] + I -
// n_slots[s]) / bank; the entries of one bank are consecutive in entry order
// (one system and level, increasing I), so its first flagged entry's scan
// value (bank_first, atomic min) ranks the others
hipError_t dp_ent_flag_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* EJ,
                              const int* ES, int* flag, int* bank_cnt, hipStream_t st);
hipError_t dp_ent_first_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* ES,
                               const int* flag, const int* fpos, int* bank_first, hipStream_t st);
hipError_t dp_ent_fill_coarse(int n_gal, int bank, int n_fine_banks, DpSystems sys, const int* EI, const int* EJ,
                              const int* ES, const int* flag, const int* fpos, const int* bank_first,
                              const int* bank_ptr, int* pos, int* src, hipStream_t st);
// per-bank counts: fine bank b = sum of its rows' counts
hipError_t dp_ent_bank_counts(int n_fine_banks, int bank, const int* row_off, int* bank_cnt, hipStream_t st);

// SpMV tile layout (rows [0, n); a tile never crosses a system's end).
// nxt[r0] = end of the tile the greedy cut starts at r0 (at most
// kSpmvTileRows rows, several rows only within `cap` kept entries) and
// kept[r0] its kept entries; the chain walk from row 0 lists the tile starts
// (ts, *n_tiles).
hipError_t dp_sl_next(int n, const int* row_ptr, const int* col, const int* row_sys, const int* row_off, int cap,
                      int* nxt, int* kept, hipStream_t st);
hipError_t dp_sl_chain(int n, const int* nxt, int* ts, int* n_tiles, hipStream_t st);
// per tile: aligned kept count (for the entry scan) and transposed products;
// err bit 2 when a tile's transposed products exceed the column word's field
hipError_t dp_sl_tile_counts(int n_tiles, int n, const int* ts, const int* kept, const int* row_ptr, const int* col,
                             int* kept_al, int* trans, int* err, hipStream_t st);
// entries, row pointers, transposed-product pointers, SpMV work table and
// tile entry ranges (a tile over the field's limit writes no entries)
hipError_t dp_sl_fill(int n_tiles, int n, const int* ts, const int* kept, const int* e_base, const int* q_base,
                      const int* row_ptr, const int* col, const int* entry_src, const int* row_sys, int* L_row_ptr,
                      int* L_tptr, int* L_col, int* L_src, int4* wg, int2* wg_ent, int* err, hipStream_t st);
hipError_t dp_sl_pad(long long total, int* L_col, int* L_src, hipStream_t st);
}  // namespace masx
