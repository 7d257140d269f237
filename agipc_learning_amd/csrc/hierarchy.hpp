// MAS hierarchy built on the device (ReorderRealtime,
// TraditionalMAS32Preconditioner.cu:2189-2230): the arrays of the host build
// (plan.hpp, build_hierarchy), bit for bit, from device copies of the mesh
// topology, the vertex partition and the contact pairs.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace masx
{
class DeviceHierarchy
{
  public:
    DeviceHierarchy() = default;
    ~DeviceHierarchy();
    DeviceHierarchy(const DeviceHierarchy&) = delete;
    DeviceHierarchy& operator=(const DeviceHierarchy&) = delete;

    // All pointers are device arrays: nbr_start/nbr_num (n_fem), nbr_list
    // (list_len), part_to_real (n_slots), real_to_part (n_fem), pairs
    // (4 * n_pairs, may be null when n_pairs == 0).  levels <= 0 selects the
    // reference level rule.  The call synchronizes `st` once, at the end, to
    // read the level sizes; it throws on an inconsistent partition or a
    // neighbour index outside [0, n_fem) (the kernels skip all work then).
    void build(int n_fem, int n_slots, int bank, int levels, const uint32_t* nbr_start, const uint32_t* nbr_num,
               const uint32_t* nbr_list, long long list_len, const int* part_to_real, const int* real_to_part,
               const int* pairs, int n_pairs, int node_offset, hipStream_t st);

    int levels() const { return levels_; }
    int total() const { return total_; }
    int n_components() const { return n_comp_; }
    int n_banks() const { return n_banks_; }

    // Host copies (any pointer may be null): level_count/level_begin
    // (levels+1; level_count[0] = n_slots), fine_mask (n_fem), prefix
    // (n_slots/bank), going_next (max(total, n_fem)), coarse_table (6*n_fem),
    // comp_ptr (total/bank + 1), comp_mask/comp_parent (n_components).
    void download(int* level_count, int* level_begin, uint32_t* fine_mask, int* prefix, int* going_next,
                  int* coarse_table, int* comp_ptr, uint32_t* comp_mask, int* comp_parent, hipStream_t st) const;

    // Device arrays, valid until the next build or destruction.
    const int*      d_going_next() const { return going_next_; }
    const int*      d_coarse_table() const { return table_; }
    const uint32_t* d_fine_mask() const { return fine_mask_; }
    const int*      d_prefix() const { return prefix_; }
    const int*      d_comp_ptr() const { return comp_ptr_; }
    const uint32_t* d_comp_mask() const { return cmask_; }
    const int*      d_comp_parent() const { return cparent_; }

  private:
    int levels_ = 0, total_ = 0, n_comp_ = 0, n_banks_ = 0, n_fem_ = 0, n_slots_ = 0, bank_ = 32;
    std::vector<int> level_count_, level_begin_;

    int*      lv_ = nullptr;          // device level sizes and the error flag
    int*      id_ = nullptr;          // per fine vertex: its node (global slot) at the current level
    uint32_t* adj_ = nullptr;         // per slot: own lane bit | linked lanes of its bank
    uint32_t* compm_ = nullptr;       // per slot: lanes of its component
    int*      parent_ = nullptr;      // per slot: slot of its component at the next level
    int*      ncomp_ = nullptr;       // per bank of the current level: component count
    int*      off_ = nullptr;         // exclusive scan of ncomp_
    int*      ncomp_all_ = nullptr;   // per slot bank of all levels: listed components
    int*      comp_ptr_ = nullptr;
    uint32_t* cmask_ = nullptr;
    int*      cparent_ = nullptr;
    uint32_t* fine_mask_ = nullptr;
    int*      prefix_ = nullptr;
    int*      going_next_ = nullptr;
    int*      table_ = nullptr;
    int*      h_lv_ = nullptr;        // pinned copy of lv_
    size_t lv_cap_ = 0, id_cap_ = 0, adj_cap_ = 0, compm_cap_ = 0, parent_cap_ = 0, ncomp_cap_ = 0, off_cap_ = 0,
           ncomp_all_cap_ = 0, comp_ptr_cap_ = 0, cmask_cap_ = 0, cparent_cap_ = 0, fine_mask_cap_ = 0,
           prefix_cap_ = 0, going_next_cap_ = 0, table_cap_ = 0;
};
}  // namespace masx
