// Stable Neo-Hookean tet Hessians on the device (see fem.hpp).
//
// Each element matrix is formed in the rotated frame of F = U S V^T, where the
// Hessian of the reference's energy (femEnergy.cu:441-459, rates u = 4 mu / 3,
// r = lambda + 5 mu / 6) is block diagonal: three twist/flip pairs on
// (X_ij, X_ji) with eigenvalues u + s S_k and u - s S_k (s = r (J - 1) - u,
// k the index outside {i, j}), and a 3x3 block on the diagonal of X with
// A_ii = u + r (S_j S_k)^2, A_ij = S_k (r (2 J - 1) - u).  Clamping those
// eigenvalues at zero is the reference's PSD projection
// (__project_StabbleNHK_H_3D1_makePD, femEnergy.cu:1641-1717), and the same
// structure gives the derivative of the projected matrix in closed form for
// the parameter gradient (divided differences of max(., 0) on the 3x3 block).
//
// Assembly is deterministic: element matrices go to a scratch array, then one
// thread per target block sums its contributions in tet order.
#include "fem.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>

namespace masx
{
namespace
{
constexpr int kBlocksPerTet = 10;  // upper 3x3 blocks (a <= b) of a 12x12 element matrix

__host__ __device__ constexpr int tet_block(int a, int b) { return 4 * a - a * (a - 1) / 2 + (b - a); }

// One Jacobi rotation annihilating A[p][q]; Q accumulates the eigenvectors (columns).
__device__ __forceinline__ void jrot(double A[3][3], double Q[3][3], int p, int q)
{
    const double apq = A[p][q];
    if(apq == 0.0) return;
    const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        const double akp = A[k][p], akq = A[k][q];
        A[k][p] = c * akp - s * akq;
        A[k][q] = s * akp + c * akq;
    }
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        const double apk = A[p][k], aqk = A[q][k];
        A[p][k] = c * apk - s * aqk;
        A[q][k] = s * apk + c * aqk;
    }
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        const double qkp = Q[k][p], qkq = Q[k][q];
        Q[k][p] = c * qkp - s * qkq;
        Q[k][q] = s * qkp + c * qkq;
    }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi: A = Q diag(w) Q^T.
__device__ void eig3(double A[3][3], double w[3], double Q[3][3])
{
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) Q[i][j] = i == j ? 1.0 : 0.0;
    double scale = 0.0;
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) scale += A[i][j] * A[i][j];
    for(int sweep = 0; sweep < 16; ++sweep)
    {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        if(!(off > 1e-32 * scale)) break;
        jrot(A, Q, 0, 1);
        jrot(A, Q, 0, 2);
        jrot(A, Q, 1, 2);
    }
#pragma unroll
    for(int i = 0; i < 3; ++i) w[i] = A[i][i];
}

__device__ __forceinline__ void sort_pair(double w[3], double V[3][3], int i, int j)
{
    if(w[j] > w[i])
    {
        const double t = w[i];
        w[i] = w[j];
        w[j] = t;
#pragma unroll
        for(int k = 0; k < 3; ++k)
        {
            const double v = V[k][i];
            V[k][i] = V[k][j];
            V[k][j] = v;
        }
    }
}

// F = U diag(S) V^T with U, V rotations (S[2] carries the sign of det F).
__device__ void svd3(const double F[3][3], double U[3][3], double S[3], double V[3][3])
{
    double C[3][3], e[3];
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    eig3(C, e, V);
    sort_pair(e, V, 0, 1);
    sort_pair(e, V, 0, 2);
    sort_pair(e, V, 1, 2);
    const double detv = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1])
                        - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0])
                        + V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if(detv < 0.0)
#pragma unroll
        for(int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
    double a0[3], a1[3];
#pragma unroll
    for(int i = 0; i < 3; ++i)
    {
        a0[i] = F[i][0] * V[0][0] + F[i][1] * V[1][0] + F[i][2] * V[2][0];
        a1[i] = F[i][0] * V[0][1] + F[i][1] * V[1][1] + F[i][2] * V[2][1];
    }
    const double n0 = sqrt(a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2]);
    double u0[3];
    if(n0 > 1e-300)
    {
        for(int i = 0; i < 3; ++i) u0[i] = a0[i] / n0;
    }
    else
    {
        u0[0] = 1.0;
        u0[1] = u0[2] = 0.0;
    }
    const double p = u0[0] * a1[0] + u0[1] * a1[1] + u0[2] * a1[2];
    for(int i = 0; i < 3; ++i) a1[i] -= p * u0[i];
    double n1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    if(!(n1 > 1e-300))
    {
        // rank <= 1: any unit vector orthogonal to u0
        const double e0[3] = {fabs(u0[0]) < 0.9 ? 1.0 : 0.0, fabs(u0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        const double q = u0[0] * e0[0] + u0[1] * e0[1] + u0[2] * e0[2];
        for(int i = 0; i < 3; ++i) a1[i] = e0[i] - q * u0[i];
        n1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    }
    double u1[3], u2[3];
    for(int i = 0; i < 3; ++i) u1[i] = a1[i] / n1;
    u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
    u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
    u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
#pragma unroll
    for(int i = 0; i < 3; ++i)
    {
        U[i][0] = u0[i];
        U[i][1] = u1[i];
        U[i][2] = u2[i];
    }
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        double acc = 0.0;
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j) acc += U[i][k] * F[i][j] * V[j][k];
        S[k] = acc;
    }
}

// Per-tet quantities shared by the element matrix and the gradient.
struct TetFrame
{
    double U[3][3], S[3], J;
    double Gh[4][3];  // shape gradients rotated by V: G V
    double lt[3], lf[3];  // twist / flip eigenvalues of pair k (pair k = the two indices other than k)
    double a[3], W[3][3];  // eigen-decomposition of the 3x3 diagonal block
};

__device__ void tet_frame(const double* __restrict__ Fg, const double* __restrict__ dm, double mu, double lambda,
                          TetFrame& T)
{
    double F[3][3], V[3][3];
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j) F[i][j] = Fg[3 * i + j];
    svd3(F, T.U, T.S, V);
    const double u = 4.0 * mu / 3.0, r = lambda + 5.0 * mu / 6.0;
    const double J = T.S[0] * T.S[1] * T.S[2];
    T.J = J;
    const double s = r * (J - 1.0) - u;
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        T.lt[k] = u + s * T.S[k];
        T.lf[k] = u - s * T.S[k];
    }
    double A[3][3];
    const double e2 = r * (2.0 * J - 1.0) - u;
    A[0][0] = u + r * (T.S[1] * T.S[2]) * (T.S[1] * T.S[2]);
    A[1][1] = u + r * (T.S[0] * T.S[2]) * (T.S[0] * T.S[2]);
    A[2][2] = u + r * (T.S[0] * T.S[1]) * (T.S[0] * T.S[1]);
    A[0][1] = A[1][0] = e2 * T.S[2];
    A[0][2] = A[2][0] = e2 * T.S[1];
    A[1][2] = A[2][1] = e2 * T.S[0];
    eig3(A, T.a, T.W);
    double G[4][3];
#pragma unroll
    for(int j = 0; j < 3; ++j)
    {
        G[0][j] = -(dm[j] + dm[3 + j] + dm[6 + j]);
        G[1][j] = dm[j];
        G[2][j] = dm[3 + j];
        G[3][j] = dm[6 + j];
    }
#pragma unroll
    for(int a = 0; a < 4; ++a)
#pragma unroll
        for(int j = 0; j < 3; ++j) T.Gh[a][j] = G[a][0] * V[0][j] + G[a][1] * V[1][j] + G[a][2] * V[2][j];
}

__global__ void __launch_bounds__(256) k_fem_element(int n_tets, const double* __restrict__ dminv,
                                                     const double* __restrict__ vol, const double* __restrict__ Fg,
                                                     const double* __restrict__ mu, const double* __restrict__ lambda,
                                                     double* __restrict__ elem)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n_tets) return;
    TetFrame T;
    tet_frame(Fg + 9 * (long long)t, dminv + 9 * (long long)t, mu[t], lambda[t], T);
    // projected Hessian in the rotated frame: pair k on (X_ij, X_ji) has
    // diagonal c[k] and coupling d[k]; the diagonal block is PA
    double c[3], d[3], PA[3][3];
#pragma unroll
    for(int k = 0; k < 3; ++k)
    {
        const double pt = fmax(T.lt[k], 0.0), pf = fmax(T.lf[k], 0.0);
        c[k] = 0.5 * (pt + pf);
        d[k] = 0.5 * (pf - pt);
    }
#pragma unroll
    for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
            double acc = 0.0;
#pragma unroll
            for(int m = 0; m < 3; ++m) acc += T.W[i][m] * fmax(T.a[m], 0.0) * T.W[j][m];
            PA[i][j] = acc;
        }
    const double v = vol[t];
    double* out = elem + (long long)t * kBlocksPerTet * 9;
#pragma unroll
    for(int a = 0; a < 4; ++a)
#pragma unroll
        for(int b = a; b < 4; ++b)
        {
            // rotated block Kh[i][k] = sum_{j,l} Gh[a][j] H_X[(i,j),(k,l)] Gh[b][l]
            double Kh[3][3];
#pragma unroll
            for(int i = 0; i < 3; ++i)
#pragma unroll
                for(int k = 0; k < 3; ++k)
                {
                    double acc = T.Gh[a][i] * T.Gh[b][k] * PA[i][k];
                    if(i == k)
                    {
#pragma unroll
                        for(int j = 0; j < 3; ++j)
                            if(j != i) acc += T.Gh[a][j] * T.Gh[b][j] * c[3 - i - j];
                    }
                    else
                        acc += T.Gh[a][k] * T.Gh[b][i] * d[3 - i - k];
                    Kh[i][k] = acc;
                }
            // K_ab = vol U Kh U^T, stored column-major
            double KU[3][3];
#pragma unroll
            for(int i = 0; i < 3; ++i)
#pragma unroll
                for(int k = 0; k < 3; ++k) KU[i][k] = Kh[i][0] * T.U[k][0] + Kh[i][1] * T.U[k][1] + Kh[i][2] * T.U[k][2];
            double* blk = out + 9 * tet_block(a, b);
#pragma unroll
            for(int i = 0; i < 3; ++i)
#pragma unroll
                for(int k = 0; k < 3; ++k)
                    blk[i + 3 * k] = v * (T.U[i][0] * KU[0][k] + T.U[i][1] * KU[1][k] + T.U[i][2] * KU[2][k]);
        }
}

__global__ void __launch_bounds__(256) k_fem_gather(long long nb, const long long* __restrict__ ptr,
                                                    const int* __restrict__ contrib, const double* __restrict__ elem,
                                                    const double* __restrict__ base, double* __restrict__ vals)
{
    const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if(j >= nb) return;
    double acc[9];
#pragma unroll
    for(int q = 0; q < 9; ++q) acc[q] = base ? base[9 * j + q] : 0.0;
    const long long e0 = ptr[j], e1 = ptr[j + 1];
    for(long long e = e0; e < e1; ++e)
    {
        const int code = contrib[e];
        const double* b = elem + (long long)(code >> 1) * 9;
        if(code & 1)
        {
#pragma unroll
            for(int i = 0; i < 3; ++i)
#pragma unroll
                for(int m = 0; m < 3; ++m) acc[i + 3 * m] += b[m + 3 * i];
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 9; ++q) acc[q] += b[q];
        }
    }
#pragma unroll
    for(int q = 0; q < 9; ++q) vals[9 * j + q] = acc[q];
}

// g_theta = -vol <Xy, dP[dH/dtheta] Xx> in the rotated frame, Xx = U^T dx G V
// for the element displacement x (and y likewise).
__global__ void __launch_bounds__(256) k_fem_param_grad(int n_tets, const int* __restrict__ tets,
                                                        const double* __restrict__ dminv, const double* __restrict__ vol,
                                                        const double* __restrict__ Fg, const double* __restrict__ mu,
                                                        const double* __restrict__ lambda, const double* __restrict__ x,
                                                        const double* __restrict__ y, double* __restrict__ g_mu,
                                                        double* __restrict__ g_lambda)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n_tets) return;
    TetFrame T;
    tet_frame(Fg + 9 * (long long)t, dminv + 9 * (long long)t, mu[t], lambda[t], T);
    double Xx[3][3] = {}, Xy[3][3] = {};
#pragma unroll
    for(int a = 0; a < 4; ++a)
    {
        const long long v = tets[4 * (long long)t + a];
        double xa[3], ya[3], xh[3], yh[3];
#pragma unroll
        for(int i = 0; i < 3; ++i)
        {
            xa[i] = x[3 * v + i];
            ya[i] = y[3 * v + i];
        }
#pragma unroll
        for(int i = 0; i < 3; ++i)
        {
            xh[i] = T.U[0][i] * xa[0] + T.U[1][i] * xa[1] + T.U[2][i] * xa[2];
            yh[i] = T.U[0][i] * ya[0] + T.U[1][i] * ya[1] + T.U[2][i] * ya[2];
        }
#pragma unroll
        for(int i = 0; i < 3; ++i)
#pragma unroll
            for(int j = 0; j < 3; ++j)
            {
                Xx[i][j] += xh[i] * T.Gh[a][j];
                Xy[i][j] += yh[i] * T.Gh[a][j];
            }
    }
    // diagonal block in its eigenbasis
    double al[3], be[3];
#pragma unroll
    for(int m = 0; m < 3; ++m)
    {
        al[m] = T.W[0][m] * Xx[0][0] + T.W[1][m] * Xx[1][1] + T.W[2][m] * Xx[2][2];
        be[m] = T.W[0][m] * Xy[0][0] + T.W[1][m] * Xy[1][1] + T.W[2][m] * Xy[2][2];
    }
    double Gam[3][3];
#pragma unroll
    for(int m = 0; m < 3; ++m)
#pragma unroll
        for(int n = 0; n < 3; ++n)
        {
            const double am = T.a[m], an = T.a[n];
            const double tol = 1e-12 * fmax(1.0, fmax(fabs(am), fabs(an)));
            Gam[m][n] = fabs(am - an) > tol ? (fmax(am, 0.0) - fmax(an, 0.0)) / (am - an) : (am > 0.0 ? 1.0 : 0.0);
        }
    const double J = T.J, v = vol[t];
    double out[2];
#pragma unroll
    for(int th = 0; th < 2; ++th)
    {
        const double du = th == 0 ? 4.0 / 3.0 : 0.0, dr = th == 0 ? 5.0 / 6.0 : 1.0;
        const double ds = dr * (J - 1.0) - du;
        double val = 0.0;
#pragma unroll
        for(int k = 0; k < 3; ++k)
        {
            const int i = k == 0 ? 1 : 0, j = k == 2 ? 1 : 2;
            const double tx = Xx[i][j] - Xx[j][i], fx = Xx[i][j] + Xx[j][i];
            const double ty = Xy[i][j] - Xy[j][i], fy = Xy[i][j] + Xy[j][i];
            const double dlt = T.lt[k] > 0.0 ? du + ds * T.S[k] : 0.0;
            const double dlf = T.lf[k] > 0.0 ? du - ds * T.S[k] : 0.0;
            val += 0.5 * (dlt * tx * ty + dlf * fx * fy);
        }
        double dA[3][3];
        const double e2 = dr * (2.0 * J - 1.0) - du;
        dA[0][0] = du + dr * (T.S[1] * T.S[2]) * (T.S[1] * T.S[2]);
        dA[1][1] = du + dr * (T.S[0] * T.S[2]) * (T.S[0] * T.S[2]);
        dA[2][2] = du + dr * (T.S[0] * T.S[1]) * (T.S[0] * T.S[1]);
        dA[0][1] = dA[1][0] = e2 * T.S[2];
        dA[0][2] = dA[2][0] = e2 * T.S[1];
        dA[1][2] = dA[2][1] = e2 * T.S[0];
#pragma unroll
        for(int m = 0; m < 3; ++m)
#pragma unroll
            for(int n = 0; n < 3; ++n)
            {
                double M = 0.0;
#pragma unroll
                for(int p = 0; p < 3; ++p)
#pragma unroll
                    for(int q = 0; q < 3; ++q) M += T.W[p][m] * dA[p][q] * T.W[q][n];
                val += be[m] * Gam[m][n] * M * al[n];
            }
        out[th] = -v * val;
    }
    g_mu[t]     = out[0];
    g_lambda[t] = out[1];
}

void fem_check(hipError_t e)
{
    if(e != hipSuccess) throw std::runtime_error(std::string("fem: ") + hipGetErrorString(e));
}

unsigned blocks_of(long long n, int b) { return (unsigned)std::max<long long>(1, (n + b - 1) / b); }
}  // namespace

FemHessian::~FemHessian()
{
    for(void* p : {(void*)d_tets_, (void*)d_ptr_, (void*)d_contrib_, (void*)d_elem_})
        if(p) (void)hipFree(p);
}

void FemHessian::setup(int n_tets, const int* tets, long long nb, const int* rows, const int* cols)
{
    if(n_tets <= 0 || !tets) throw std::invalid_argument("no tets");
    if(n_tets > (1 << 30) / kBlocksPerTet) throw std::invalid_argument("too many tets");
    if(nb <= 0 || !rows || !cols) throw std::invalid_argument("empty block structure");
    int n_rows = 0;
    for(long long k = 0; k < nb; ++k)
    {
        if(rows[k] < 0 || cols[k] < 0) throw std::invalid_argument("negative block index");
        if(rows[k] > cols[k]) throw std::invalid_argument("block structure must be upper (row <= col)");
        n_rows = std::max(n_rows, cols[k] + 1);
    }
    // (row, col) -> block: blocks sorted by (row, col) with a row index
    std::vector<long long> order(nb);
    std::iota(order.begin(), order.end(), 0LL);
    std::sort(order.begin(), order.end(), [&](long long a, long long b) {
        return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
    });
    std::vector<long long> row_ptr(n_rows + 1, 0);
    for(long long k = 0; k < nb; ++k)
    {
        if(k > 0 && rows[order[k]] == rows[order[k - 1]] && cols[order[k]] == cols[order[k - 1]])
            throw std::invalid_argument("duplicate block in the structure");
        ++row_ptr[rows[order[k]] + 1];
    }
    for(int r = 0; r < n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
    auto find = [&](int p, int q) -> long long {
        if(p >= n_rows) return -1;
        long long lo = row_ptr[p], hi = row_ptr[p + 1];
        while(lo < hi)
        {
            const long long mid = (lo + hi) / 2;
            if(cols[order[mid]] < q)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo < row_ptr[p + 1] && cols[order[lo]] == q ? order[lo] : -1;
    };
    // contributions (tet, a, b) with tets[a] <= tets[b], in tet order
    std::vector<long long> target;
    std::vector<int>       code;
    target.reserve((size_t)n_tets * kBlocksPerTet);
    code.reserve((size_t)n_tets * kBlocksPerTet);
    for(int t = 0; t < n_tets; ++t)
    {
        const int* tv = tets + 4 * (size_t)t;
        for(int a = 0; a < 4; ++a)
            if(tv[a] < 0 || tv[a] >= n_rows) throw std::invalid_argument("tet vertex outside the block structure");
        for(int a = 0; a < 4; ++a)
            for(int b = 0; b < 4; ++b)
            {
                // orientation p <= q; a repeated vertex (p == q, a != b) puts
                // both orderings on the diagonal block
                const int p = tv[a], q = tv[b];
                if(p > q) continue;
                const long long j = find(p, q);
                if(j < 0) throw std::invalid_argument("a tet edge or vertex has no block in the structure");
                target.push_back(j);
                code.push_back(((t * kBlocksPerTet + tet_block(std::min(a, b), std::max(a, b))) << 1) | (a > b ? 1 : 0));
            }
    }
    std::vector<long long> ptr(nb + 1, 0);
    for(long long j : target) ++ptr[j + 1];
    for(long long j = 0; j < nb; ++j) ptr[j + 1] += ptr[j];
    std::vector<int>       sorted(code.size());
    std::vector<long long> cur(ptr.begin(), ptr.end() - 1);
    for(size_t e = 0; e < code.size(); ++e) sorted[cur[target[e]]++] = code[e];

    for(void* p : {(void*)d_tets_, (void*)d_ptr_, (void*)d_contrib_})
        if(p) (void)hipFree(p);
    d_tets_ = nullptr;
    d_ptr_ = nullptr;
    d_contrib_ = nullptr;
    fem_check(hipMalloc((void**)&d_tets_, sizeof(int) * 4 * (size_t)n_tets));
    fem_check(hipMalloc((void**)&d_ptr_, sizeof(long long) * (nb + 1)));
    fem_check(hipMalloc((void**)&d_contrib_, sizeof(int) * std::max<size_t>(1, sorted.size())));
    fem_check(hipMemcpy(d_tets_, tets, sizeof(int) * 4 * (size_t)n_tets, hipMemcpyHostToDevice));
    fem_check(hipMemcpy(d_ptr_, ptr.data(), sizeof(long long) * (nb + 1), hipMemcpyHostToDevice));
    if(!sorted.empty())
        fem_check(hipMemcpy(d_contrib_, sorted.data(), sizeof(int) * sorted.size(), hipMemcpyHostToDevice));
    n_tets_ = n_tets;
    nb_     = nb;
    n_rows_ = n_rows;
}

void FemHessian::assemble(const double* dminv, const double* vol, const double* F, const double* mu,
                          const double* lambda, const double* base, double* vals, hipStream_t st)
{
    if(n_tets_ <= 0) throw std::runtime_error("fem: setup must come first");
    if(!dminv || !vol || !F || !mu || !lambda || !vals) throw std::invalid_argument("fem: null argument");
    const size_t need = (size_t)n_tets_ * kBlocksPerTet * 9;
    if(need > elem_cap_)
    {
        if(d_elem_) (void)hipFree(d_elem_);
        d_elem_   = nullptr;
        elem_cap_ = 0;
        fem_check(hipMalloc((void**)&d_elem_, need * sizeof(double)));
        elem_cap_ = need;
    }
    hipLaunchKernelGGL(k_fem_element, dim3(blocks_of(n_tets_, 256)), dim3(256), 0, st, n_tets_, dminv, vol, F, mu,
                       lambda, d_elem_);
    hipLaunchKernelGGL(k_fem_gather, dim3(blocks_of(nb_, 256)), dim3(256), 0, st, nb_, d_ptr_, d_contrib_, d_elem_,
                       base, vals);
    fem_check(hipGetLastError());
}

void FemHessian::param_grad(const double* dminv, const double* vol, const double* F, const double* mu,
                            const double* lambda, const double* x, const double* y, double* g_mu, double* g_lambda,
                            hipStream_t st) const
{
    if(n_tets_ <= 0) throw std::runtime_error("fem: setup must come first");
    if(!dminv || !vol || !F || !mu || !lambda || !x || !y || !g_mu || !g_lambda)
        throw std::invalid_argument("fem: null argument");
    hipLaunchKernelGGL(k_fem_param_grad, dim3(blocks_of(n_tets_, 256)), dim3(256), 0, st, n_tets_, d_tets_, dminv,
                       vol, F, mu, lambda, x, y, g_mu, g_lambda);
    fem_check(hipGetLastError());
}

}  // namespace masx
