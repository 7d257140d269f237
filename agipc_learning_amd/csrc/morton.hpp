// Morton-order vertex partition on the device (the reference's mesh
// reordering, gl_main.cu:1176-1243): the same result as the host
// morton_partition (plan.cpp) bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace masx
{
// Largest vertex count morton_partition_noalloc takes (its static scratch).
constexpr int kMortonNoAllocMax = 1 << 19;

class DeviceMorton
{
  public:
    DeviceMorton() = default;
    ~DeviceMorton();
    DeviceMorton(const DeviceMorton&) = delete;
    DeviceMorton& operator=(const DeviceMorton&) = delete;

    // n positions (device, 3 doubles per vertex) -> part_to_real (pad(n, bank)
    // ints, -1 in the padding; part_cap >= that) and real_to_part (n ints),
    // both device arrays.  Asynchronous on `st`; scratch grows as needed.
    void partition(int n, const double* pos, int bank, int* part_to_real, long long part_cap, int* real_to_part,
                   hipStream_t st);

  private:
    double*   box_  = nullptr;   // per-block partial bounds, then the final 6
    uint32_t* key_[2] = {nullptr, nullptr};
    int*      idx_[2] = {nullptr, nullptr};
    int*      hist_ = nullptr;   // per (digit, tile) counts, scanned in place
    int*      part_ = nullptr;   // scan partials
    size_t box_cap_ = 0, key_cap_[2] = {0, 0}, idx_cap_[2] = {0, 0}, hist_cap_ = 0, part_cap_ = 0;
};

// DeviceMorton::partition without allocating: the scratch is a static
// device array of the library (n <= kMortonNoAllocMax), one per GPU, shared
// by every call on that GPU, so calls that overlap on one GPU race.  No
// hipMalloc, hipFree or memset; asynchronous on `st`.
void morton_partition_noalloc(int n, const double* pos, int bank, int* part_to_real, long long part_cap,
                              int* real_to_part, hipStream_t st);
}  // namespace masx
