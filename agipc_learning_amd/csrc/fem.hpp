// Stable Neo-Hookean tet Hessians on the device: assembly into the values of
// an upper 3x3 block structure, and the parameter gradient of an adjoint
// solve.  The element matrix follows femEnergy.cu (first Piola stress
// :441-459, PSD projection through the analytic eigensystem :1641-1717,
// element matrix vol * PFPX^T H PFPX :1719-1855) without the dt^2 factor.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace masx
{
class FemHessian
{
  public:
    FemHessian() = default;
    ~FemHessian();
    FemHessian(const FemHessian&) = delete;
    FemHessian& operator=(const FemHessian&) = delete;

    // Host topology: tets (4 block-row ids each) and the upper block
    // structure (rows[k] <= cols[k]) the assembled values are laid out in.
    // Every tet edge and vertex must have its block.  Uploads the
    // contribution lists (per target block, in tet order).
    void setup(int n_tets, const int* tets, long long nb, const int* rows, const int* cols);

    // vals (nb x 9, col-major blocks) = base + sum of the element matrices.
    // Element inputs per tet: dminv (9, row-major Dm^-1), vol, F (9,
    // row-major), mu, lambda.  All device arrays.
    void assemble(const double* dminv, const double* vol, const double* F, const double* mu, const double* lambda,
                  const double* base, double* vals, hipStream_t st);

    // g_mu[e] = -y_e^T (dK_e/dmu) x_e, g_lambda likewise; x, y are block
    // vectors (3 doubles per block row).  Device arrays.
    void param_grad(const double* dminv, const double* vol, const double* F, const double* mu, const double* lambda,
                    const double* x, const double* y, double* g_mu, double* g_lambda, hipStream_t st) const;

    int       n_tets() const { return n_tets_; }
    long long nb() const { return nb_; }
    int       n_rows() const { return n_rows_; }

  private:
    int       n_tets_ = 0, n_rows_ = 0;
    long long nb_ = 0;
    int*      d_tets_ = nullptr;     // 4 per tet
    long long* d_ptr_ = nullptr;     // nb + 1
    int*      d_contrib_ = nullptr;  // (tet * 10 + block) << 1 | transposed
    double*   d_elem_ = nullptr;     // 10 blocks x 9 per tet
    size_t    elem_cap_ = 0;
};
}  // namespace masx
