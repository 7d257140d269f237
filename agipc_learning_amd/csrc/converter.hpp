// GPU triplet -> BCOO converter (sorted by (row, col), duplicates summed).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace masx
{
class Converter
{
  public:
    Converter() = default;
    ~Converter();
    Converter(const Converter&) = delete;
    Converter& operator=(const Converter&) = delete;

    // n triplets (device arrays: int32 rows, int32 cols, 9 doubles per block)
    // -> unique blocks in (row, col) order written to the outputs (capacity n
    // each); drop_lower != 0 discards blocks with row > col.  Returns the
    // unique count (the call synchronizes the stream, as the count is needed
    // on the host).  Throws on an index outside [0, n_block_rows).
    int convert(int n, int n_block_rows, const int* rows, const int* cols, const double* vals, int drop_lower,
                int* out_rows, int* out_cols, double* out_vals, hipStream_t st);

  private:
    void scan(const int* in, int* out, int m, hipStream_t st);

    int*                row_count_ = nullptr;
    int*                row_start_ = nullptr;
    int*                cursor_    = nullptr;
    uint32_t*           bucket_    = nullptr;
    int*                seg_       = nullptr;
    int*                uniq_      = nullptr;
    int*                out_start_ = nullptr;
    int*                tiles_     = nullptr;
    int*                flag_      = nullptr;
    unsigned long long* keys_      = nullptr;
    unsigned long long* sorted_    = nullptr;
    size_t row_count_cap_ = 0, row_start_cap_ = 0, cursor_cap_ = 0, bucket_cap_ = 0, seg_cap_ = 0, uniq_cap_ = 0,
           out_start_cap_ = 0, tiles_cap_ = 0, flag_cap_ = 0, keys_cap_ = 0, sorted_cap_ = 0;
    int* h_pinned_ = nullptr;
};
}  // namespace masx
