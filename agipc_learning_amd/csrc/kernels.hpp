// Launchers of the HIP kernels (kernels.hip).  All return the launch error.
#pragma once
#include "device.hpp"

namespace masx
{
// bytes of one bank inverse in the diagonal-major layout of `precision`
size_t mas_bank_bytes(int bank, int precision);
bool   spmv_tiling_supported(int entries_per_thread);

hipError_t launch_pack_values(const DevData& d, hipStream_t st);
hipError_t launch_pack_abd(const DevData& d, const int* cp_body, const int* cp_sub, long long n_couplings,
                           float* vals_abd, hipStream_t st);
// lanes per coarse entry: 16 or 64
hipError_t launch_gal_gather(const double* fine, double* E, const int* ptr, const int* src, int e0, int e1,
                             int src_is_fine, int lanes, hipStream_t st);
// the map from stored position of a bank's inverse to its element (one
// entry per stored value: FP32 floats or 16-bit halves), for the layout of
// this bank size, precision and (half, 32-slot banks) MFMA tiles or not;
// the factor kernels write through it (DevData::inv_map)
int inverse_map_entries(int bank, int precision, bool tiles);
hipError_t launch_inverse_map(int bank, int precision, bool tiles, unsigned short* map, hipStream_t st);
// banks [bank_begin, bank_end) of the bank list
hipError_t launch_mas_factor(const DevData& d, const double* E, const int* ent_ptr, const int* ent_pos,
                             const int* ent_src, void* out, float* scale, int precision, int* status, int bank_begin,
                             int bank_end, hipStream_t st);
hipError_t launch_abd_factor(const DevData& d, const int* body_src, float* inv, int n_bodies, hipStream_t st);
hipError_t launch_diag_factor(const DevData& d, const int* diag_src, const int* needs, float* inv, hipStream_t st);
hipError_t launch_load_rhs(const DevData& d, const int* e2i, const double* b, int n_ext, hipStream_t st);
hipError_t launch_store_x(const DevData& d, const int* e2i, double* x, int n_ext, hipStream_t st);
hipError_t launch_init_state(const DevData& d, int rule, double tol, const int* max_iter, int max_override,
                             int check_every, int frozen, hipStream_t st);
hipError_t launch_solve_prepare(const DevData& d, const int* int_to_ext, const double* b, hipStream_t st);
// host[0] = systems not done, then host[1] = ++*seq (device counter)
hipError_t launch_export_status(const DevData& d, int* host, int* seq, hipStream_t st);
// host/seq (one system only): also export the chunk status, as launch_export_status
hipError_t launch_phase_finish(const DevData& d, hipStream_t st, int* host = nullptr, int* seq = nullptr);
// Events stamped by one kernel's own dispatch (start and end of its
// execution, via hipExtLaunchKernelGGL): their elapsed time is the kernel's
// duration alone, without the gaps between launches.  Null: a plain launch.
struct LaunchEvents
{
    hipEvent_t start, stop;
};
hipError_t launch_mas_local(const DevData& d, int precision, int force, hipStream_t st,
                            const LaunchEvents* ev = nullptr);
hipError_t launch_precond_finish(const DevData& d, int n_wg, int force, hipStream_t st,
                                 const LaunchEvents* ev = nullptr);
// row-aligned tiles of at most 256 rows and 256 * entries_per_thread entries (2, 4 or 8)
hipError_t launch_spmv_fused(const DevData& d, int n_wg, int entries_per_thread, const float* p_old, float* p_new,
                             hipStream_t st, const LaunchEvents* ev = nullptr);
hipError_t launch_update(const DevData& d, int n_wg, const float* p, int mode, hipStream_t st,
                         const LaunchEvents* ev = nullptr);
// true residual of the checked (or, report_only, all) systems and its decision
hipError_t launch_true_residual(const DevData& d, int n_wg, int report_only, hipStream_t st);
hipError_t launch_spmv64(const DevData& d, const double* x, double* y, hipStream_t st);
hipError_t launch_gather_ext64(const int* e2i, const double* src, double* dst, int n_ext, hipStream_t st);
hipError_t launch_axpby_int_to_ext(const int* e2i, const double* y_int, double* y_ext, double a, double b, int n_ext,
                                   hipStream_t st);
hipError_t launch_z32_to_ext(const int* e2i, const float* z, double* dst, int n_ext, hipStream_t st);
hipError_t launch_gather_blocks(const double* src, const uint32_t* idx, double* dst, long long n, hipStream_t st);

// Device arrays of doubles copied into one array by one launch: segment s,
// n[s] doubles from src[s], goes to out + dst[s]; first[s] is its first block
// (chunks of kCopyChunk doubles), first[count] the grid size.
constexpr int kCopySegmentsMax = 64;
constexpr int kCopyChunk       = 2048;
struct CopySegments
{
    const double* src[kCopySegmentsMax];
    long long     dst[kCopySegmentsMax];
    long long     n[kCopySegmentsMax];
    int           first[kCopySegmentsMax + 1];
    int           count;
};
hipError_t launch_copy_segments(const CopySegments& c, double* out, hipStream_t st);
}  // namespace masx
