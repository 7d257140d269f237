// Synthetic StiffGIPC-style Newton systems (see masx_gen.h).
#include "masx_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
thread_local std::string g_error;

// ---------------------------------------------------------------- random ----
// Counter-based: value(seed, stream, index) does not depend on call order or
// thread count.
inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t h = mix64(seed ^ mix64(stream * 0x632be59bd9b4e019ull + mix64(i)));
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
}
inline double normal(uint64_t seed, uint64_t stream, uint64_t i)
{
    const double u1 = uniform01(seed, stream, 2 * i);
    const double u2 = uniform01(seed, stream, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
enum Stream : uint64_t
{
    kStreamF = 1, kStreamE = 2, kStreamContact = 3, kStreamBody = 4, kStreamRhs = 5,
    kStreamBodyPair = 6, kStreamCoupling = 7
};

// ------------------------------------------------------------- 9x9 SNH -----
// vec(F) is column-major: index i + 3*j for F(i,j).
inline int eps3(int a, int b, int c)
{
    if(a == b || b == c || a == c) return 0;
    return ((a + 1) % 3 == b) ? 1 : -1;  // (0,1,2),(1,2,0),(2,0,1) -> +1
}

// Symmetric 9x9 Jacobi eigen-decomposition: A = Q diag(w) Q^T (Q columns).
void jacobi_eigen9(double A[9][9], double w[9], double Q[9][9])
{
    for(int i = 0; i < 9; ++i)
        for(int j = 0; j < 9; ++j) Q[i][j] = i == j ? 1.0 : 0.0;
    double scale = 0.0;
    for(int i = 0; i < 9; ++i)
        for(int j = 0; j < 9; ++j) scale += A[i][j] * A[i][j];
    const double tol = 1e-30 * scale;
    for(int sweep = 0; sweep < 60; ++sweep)
    {
        double off = 0.0;
        for(int p = 0; p < 9; ++p)
            for(int q = p + 1; q < 9; ++q) off += A[p][q] * A[p][q];
        if(off <= tol) break;
        for(int p = 0; p < 9; ++p)
            for(int q = p + 1; q < 9; ++q)
            {
                if(A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for(int k = 0; k < 9; ++k)
                {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for(int k = 0; k < 9; ++k)
                {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for(int k = 0; k < 9; ++k)
                {
                    const double qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p] = c * qkp - s * qkq;
                    Q[k][q] = s * qkp + c * qkq;
                }
            }
    }
    for(int i = 0; i < 9; ++i) w[i] = A[i][i];
}

bool cholesky9_ok(const double H[9][9])
{
    double L[9][9] = {};
    for(int j = 0; j < 9; ++j)
    {
        double d = H[j][j];
        for(int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if(!(d > 0.0)) return false;
        L[j][j] = std::sqrt(d);
        for(int i = j + 1; i < 9; ++i)
        {
            double s = H[i][j];
            for(int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = s / L[j][j];
        }
    }
    return true;
}

// Stable Neo-Hookean in the reference's rates: u = 4 mu / 3,
// r = lambda + 5 mu / 6 (gl_main.cu:586-590), P = u F + r (J - 1 - u/r) cof F
// (femEnergy.cu:441-459).  Hessian = u I + r g g^T + r (J - 1 - u/r) d2J,
// clamped to PSD.
void snh_hessian_psd(const double F[3][3], double mu, double lambda, double H[9][9])
{
    const double u = 4.0 * mu / 3.0;
    const double r = lambda + 5.0 * mu / 6.0;
    double cof[3][3];
    cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    cof[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
    cof[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
    cof[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
    cof[2][0] = F[0][1] * F[1][2] - F[1][1] * F[0][2];
    cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const double J = F[0][0] * cof[0][0] + F[0][1] * cof[0][1] + F[0][2] * cof[0][2];
    const double s = r * (J - 1.0 - u / r);
    double g[9];
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j) g[i + 3 * j] = cof[i][j];
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j)
            for(int k = 0; k < 3; ++k)
                for(int l = 0; l < 3; ++l)
                {
                    // d2J/dF_ij dF_kl = sum_mn eps_ikm eps_jln F_mn
                    double d2 = 0.0;
                    for(int m = 0; m < 3; ++m)
                    {
                        const int e1 = eps3(i, k, m);
                        if(!e1) continue;
                        for(int n = 0; n < 3; ++n)
                        {
                            const int e2 = eps3(j, l, n);
                            if(e2) d2 += e1 * e2 * F[m][n];
                        }
                    }
                    const int a = i + 3 * j, b = k + 3 * l;
                    H[a][b] = (a == b ? u : 0.0) + r * g[a] * g[b] + s * d2;
                }
    if(cholesky9_ok(H)) return;
    double A[9][9], w[9], Q[9][9];
    std::memcpy(A, H, sizeof(A));
    jacobi_eigen9(A, w, Q);
    for(int i = 0; i < 9; ++i)
        if(w[i] < 0.0) w[i] = 0.0;
    for(int a = 0; a < 9; ++a)
        for(int b = 0; b < 9; ++b)
        {
            double acc = 0.0;
            for(int k = 0; k < 9; ++k) acc += Q[a][k] * w[k] * Q[b][k];
            H[a][b] = acc;
        }
    for(int a = 0; a < 9; ++a)
        for(int b = a + 1; b < 9; ++b) H[a][b] = H[b][a] = 0.5 * (H[a][b] + H[b][a]);
}

// K (12x12 row-major) = vol * D^T H D, D = d vec(F) / d x with
// dF_ij / dx_{a,i} = G[a][j], G[0][j] = -sum_k Dm^-1[k][j], G[a][j] = Dm^-1[a-1][j].
void element_stiffness(const double dminv[9], double vol, const double H[9][9], double K[144])
{
    double G[4][3];
    for(int j = 0; j < 3; ++j)
    {
        G[0][j] = -(dminv[0 * 3 + j] + dminv[1 * 3 + j] + dminv[2 * 3 + j]);
        for(int a = 1; a < 4; ++a) G[a][j] = dminv[(a - 1) * 3 + j];
    }
    double T[9][12];  // H D
    for(int f = 0; f < 9; ++f)
        for(int a = 0; a < 4; ++a)
            for(int i = 0; i < 3; ++i)
            {
                double acc = 0.0;
                for(int j = 0; j < 3; ++j) acc += H[f][i + 3 * j] * G[a][j];
                T[f][3 * a + i] = acc;
            }
    for(int a = 0; a < 4; ++a)
        for(int i = 0; i < 3; ++i)
            for(int c = 0; c < 12; ++c)
            {
                double acc = 0.0;
                for(int j = 0; j < 3; ++j) acc += G[a][j] * T[i + 3 * j][c];
                K[(3 * a + i) * 12 + c] = vol * acc;
            }
    for(int p = 0; p < 12; ++p)
        for(int q = p + 1; q < 12; ++q) K[p * 12 + q] = K[q * 12 + p] = 0.5 * (K[p * 12 + q] + K[q * 12 + p]);
}

void inverse3(const double m[9], double inv[9])
{
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const double id  = 1.0 / det;
    inv[0] = c00 * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = c01 * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = c02 * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

// ---------------------------------------------------------- block matrix ---
struct BlockMatrix
{
    std::vector<long long> row_ptr;
    std::vector<int>       col;
    std::vector<double>    val;  // 9 per block, col-major

    long long find(int r, int c) const
    {
        const int* b = col.data() + row_ptr[r];
        const int* e = col.data() + row_ptr[r + 1];
        const int* p = std::lower_bound(b, e, c);
        if(p == e || *p != c) throw std::runtime_error("block not in structure");
        return row_ptr[r] + (p - b);
    }
    // adds the row-major 3x3 `blkrm` (rows of vertex a, cols of vertex b) to
    // the upper storage
    void add(int a, int b, const double* blkrm, int stride)
    {
        if(a <= b)
        {
            double* d = val.data() + 9 * find(a, b);
            for(int i = 0; i < 3; ++i)
                for(int j = 0; j < 3; ++j) d[i + 3 * j] += blkrm[i * stride + j];
        }
        else
        {
            double* d = val.data() + 9 * find(b, a);
            for(int i = 0; i < 3; ++i)
                for(int j = 0; j < 3; ++j) d[j + 3 * i] += blkrm[i * stride + j];
        }
    }
};

struct ContactRec { int v[4]; };
struct BodyPair { int a, b; };
struct Coupling { int body, vertex; };
}  // namespace

struct masx_gen_system
{
    masx_gen_params     p{};
    masx_gen_sizes      sz{};
    std::vector<double> pos, dminv, vol, defgrad, mu, lambda, mass, rhs;
    std::vector<int>    tets, tris, pairs;
    BlockMatrix         A;
    std::vector<int>    rows;  // expanded row per block
};

extern "C" void masx_gen_default_params(masx_gen_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 8;
    p->spacing             = 0.0125;
    p->young               = 1e5;
    p->young_sigma         = 0.0;
    p->poisson             = 0.4;
    p->density             = 1000.0;
    p->dt                  = 1.0 / 60.0;
    p->f_noise             = 0.1;
    p->contact_stiffness   = 1e3;
    p->contact_sigma       = 0.5;
    p->body_shift          = 1e4;
    p->seed                = 0;
}

extern "C" const char* masx_gen_last_error(void) { return g_error.c_str(); }

extern "C" int masx_gen_element_hessian(const double* dm_inv, double volume, const double* Fin, double mu,
                                        double lambda, double* H12)
{
    double F[3][3], H[9][9];
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j) F[i][j] = Fin[3 * i + j];
    snh_hessian_psd(F, mu, lambda, H);
    element_stiffness(dm_inv, volume, H, H12);
    return 0;
}

static void build(masx_gen_system& S)
{
    const masx_gen_params& p = S.p;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const bool fem = nx > 0 && ny > 0 && nz > 0;
    const int  nvx = nx + 1, nvy = ny + 1;
    const int  nv = fem ? nvx * nvy * (nz + 1) : 0;
    const int  ncube = fem ? nx * ny * nz : 0;
    const int  ntet = 6 * ncube;
    const int  nbody = p.n_bodies;
    const int  foff = 4 * nbody;
    const int  block_rows = foff + nv;
    auto vid = [&](int i, int j, int k) { return i + nvx * (j + nvy * k); };

    // geometry and tets (Kuhn split of each cube along its main diagonal)
    S.pos.resize(3 * (size_t)nv);
    for(int k = 0; k <= nz && fem; ++k)
        for(int j = 0; j <= ny; ++j)
            for(int i = 0; i <= nx; ++i)
            {
                const int v = vid(i, j, k);
                S.pos[3 * v + 0] = i * p.spacing;
                S.pos[3 * v + 1] = j * p.spacing;
                S.pos[3 * v + 2] = k * p.spacing;
            }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    S.tets.resize(4 * (size_t)ntet);
    for(int c = 0; c < ncube; ++c)
    {
        const int ci = c % nx, cj = (c / nx) % ny, ck = c / (nx * ny);
        for(int t = 0; t < 6; ++t)
        {
            int o[3] = {0, 0, 0};
            int* tv = &S.tets[4 * (size_t)(6 * c + t)];
            tv[0] = vid(ci, cj, ck);
            o[perms[t][0]] = 1;
            tv[1] = vid(ci + o[0], cj + o[1], ck + o[2]);
            o[perms[t][1]] = 1;
            tv[2] = vid(ci + o[0], cj + o[1], ck + o[2]);
            tv[3] = vid(ci + 1, cj + 1, ck + 1);
        }
    }
    // boundary triangles: tet faces lying in one boundary plane of the box
    auto on_plane = [&](int v, int axis, int value) {
        const int i = v % nvx, j = (v / nvx) % nvy, k = v / (nvx * nvy);
        const int c = axis == 0 ? i : axis == 1 ? j : k;
        return c == value;
    };
    static const int faces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    const int extent[3] = {nx, ny, nz};
    for(int t = 0; t < ntet; ++t)
    {
        const int* tv = &S.tets[4 * (size_t)t];
        for(int f = 0; f < 4; ++f)
        {
            const int a = tv[faces[f][0]], b = tv[faces[f][1]], c = tv[faces[f][2]];
            bool boundary = false;
            for(int axis = 0; axis < 3 && !boundary; ++axis)
                for(int value : {0, extent[axis]})
                    if(on_plane(a, axis, value) && on_plane(b, axis, value) && on_plane(c, axis, value))
                        boundary = true;
            if(boundary)
            {
                S.tris.push_back(a);
                S.tris.push_back(b);
                S.tris.push_back(c);
            }
        }
    }
    const int ntri = (int)S.tris.size() / 3;
    std::vector<int> surf;
    {
        std::vector<char> is_s(nv, 0);
        for(int v : S.tris) is_s[v] = 1;
        for(int v = 0; v < nv; ++v)
            if(is_s[v]) surf.push_back(v);
    }

    // per-tet material and deformation
    const double lame_mu0 = p.young / (2.0 * (1.0 + p.poisson));
    const double lame_la0 = p.young * p.poisson / ((1.0 + p.poisson) * (1.0 - 2.0 * p.poisson));
    S.dminv.resize(9 * (size_t)ntet);
    S.vol.resize(ntet);
    S.defgrad.resize(9 * (size_t)ntet);
    S.mu.resize(ntet);
    S.lambda.resize(ntet);
#pragma omp parallel for schedule(static)
    for(int t = 0; t < ntet; ++t)
    {
        const int* tv = &S.tets[4 * (size_t)t];
        double Dm[9];
        for(int i = 0; i < 3; ++i)
            for(int a = 0; a < 3; ++a) Dm[3 * i + a] = S.pos[3 * tv[a + 1] + i] - S.pos[3 * tv[0] + i];
        inverse3(Dm, &S.dminv[9 * (size_t)t]);
        const double det = Dm[0] * (Dm[4] * Dm[8] - Dm[5] * Dm[7]) - Dm[1] * (Dm[3] * Dm[8] - Dm[5] * Dm[6])
                           + Dm[2] * (Dm[3] * Dm[7] - Dm[4] * Dm[6]);
        S.vol[t] = std::fabs(det) / 6.0;
        for(int e = 0; e < 9; ++e)
            S.defgrad[9 * (size_t)t + e] = (e % 4 == 0 ? 1.0 : 0.0) + p.f_noise * normal(p.seed, kStreamF, 9ull * t + e);
        const double scale = p.young_sigma > 0 ? std::exp(p.young_sigma * normal(p.seed, kStreamE, t)) : 1.0;
        S.mu[t]     = lame_mu0 * scale;
        S.lambda[t] = lame_la0 * scale;
    }
    S.mass.assign(nv, 0.0);
    for(int t = 0; t < ntet; ++t)
        for(int a = 0; a < 4; ++a) S.mass[S.tets[4 * (size_t)t + a]] += p.density * S.vol[t] / 4.0;

    // contacts
    for(int k = 0; k < p.n_contacts && ntri > 0 && !surf.empty(); ++k)
    {
        ContactRec c;
        uint64_t   draw = 0;
        for(;;)
        {
            const int pv = surf[std::min((int)(uniform01(p.seed, kStreamContact, 64ull * k + draw++) * surf.size()),
                                         (int)surf.size() - 1)];
            const int tr = std::min((int)(uniform01(p.seed, kStreamContact, 64ull * k + draw++) * ntri), ntri - 1);
            const int* tv = &S.tris[3 * (size_t)tr];
            if(pv == tv[0] || pv == tv[1] || pv == tv[2])
            {
                if(draw > 60) throw std::runtime_error("contact sampling failed");
                continue;
            }
            c.v[0] = pv;
            c.v[1] = tv[0];
            c.v[2] = tv[1];
            c.v[3] = tv[2];
            break;
        }
        for(int i = 0; i < 4; ++i) S.pairs.push_back(c.v[i]);
    }
    const int npair = (int)S.pairs.size() / 4;

    std::vector<BodyPair> bpairs;
    for(int k = 0; k < p.n_body_pairs && nbody > 1; ++k)
    {
        int a = std::min((int)(uniform01(p.seed, kStreamBodyPair, 2ull * k) * nbody), nbody - 1);
        int b = std::min((int)(uniform01(p.seed, kStreamBodyPair, 2ull * k + 1) * (nbody - 1)), nbody - 2);
        if(b >= a) ++b;
        bpairs.push_back({std::min(a, b), std::max(a, b)});
    }
    std::vector<Coupling> couplings;
    if(p.n_coupled_bodies > 0 && nbody > 0 && !surf.empty())
    {
        std::vector<int> order(nbody);
        for(int i = 0; i < nbody; ++i) order[i] = i;
        // partial Fisher-Yates with counter-based draws
        const int m = std::min(p.n_coupled_bodies, nbody);
        for(int i = 0; i < m; ++i)
        {
            const int j = i + std::min((int)(uniform01(p.seed, kStreamCoupling, 2ull * i) * (nbody - i)), nbody - i - 1);
            std::swap(order[i], order[j]);
            const int v = surf[std::min((int)(uniform01(p.seed, kStreamCoupling, 2ull * i + 1) * surf.size()),
                                        (int)surf.size() - 1)];
            couplings.push_back({order[i], v});
        }
    }

    // structure
    std::vector<std::vector<int>> rc(block_rows);
    for(int b = 0; b < nbody; ++b)
        for(int q = 0; q < 4; ++q)
            for(int s = q; s < 4; ++s) rc[4 * b + q].push_back(4 * b + s);
    for(const auto& bp : bpairs)
        for(int q = 0; q < 4; ++q)
            for(int s = 0; s < 4; ++s) rc[4 * bp.a + q].push_back(4 * bp.b + s);
    for(const auto& cp : couplings)
        for(int q = 0; q < 4; ++q) rc[4 * cp.body + q].push_back(foff + cp.vertex);
    for(int v = 0; v < nv; ++v) rc[foff + v].push_back(foff + v);
    for(int t = 0; t < ntet; ++t)
    {
        const int* tv = &S.tets[4 * (size_t)t];
        for(int a = 0; a < 4; ++a)
            for(int b = a + 1; b < 4; ++b)
            {
                const int u = std::min(tv[a], tv[b]), w = std::max(tv[a], tv[b]);
                rc[foff + u].push_back(foff + w);
            }
    }
    for(int k = 0; k < npair; ++k)
    {
        const int* cv = &S.pairs[4 * (size_t)k];
        for(int a = 0; a < 4; ++a)
            for(int b = a + 1; b < 4; ++b)
                rc[foff + std::min(cv[a], cv[b])].push_back(foff + std::max(cv[a], cv[b]));
    }
    S.A.row_ptr.assign(block_rows + 1, 0);
    for(int r = 0; r < block_rows; ++r)
    {
        auto& v = rc[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        S.A.row_ptr[r + 1] = S.A.row_ptr[r] + (long long)v.size();
    }
    const long long nb = S.A.row_ptr[block_rows];
    S.A.col.resize(nb);
    for(int r = 0; r < block_rows; ++r)
    {
        std::copy(rc[r].begin(), rc[r].end(), S.A.col.begin() + S.A.row_ptr[r]);
        std::vector<int>().swap(rc[r]);
    }
    S.A.val.assign(9 * (size_t)nb, 0.0);

    // values: mass
    const double idt2 = 1.0 / (p.dt * p.dt);
    for(int v = 0; v < nv; ++v)
    {
        double* d = S.A.val.data() + 9 * S.A.find(foff + v, foff + v);
        d[0] += S.mass[v] * idt2;
        d[4] += S.mass[v] * idt2;
        d[8] += S.mass[v] * idt2;
    }
    // values: elements, 8-colored by cube parity so that cubes processed
    // concurrently share no vertex (deterministic summation order)
    for(int color = 0; color < 8 && fem; ++color)
    {
        const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
        const int cx = (nx - px + 1) / 2, cy = (ny - py + 1) / 2, cz = (nz - pz + 1) / 2;
        const long long ncol = (long long)cx * cy * cz;
#pragma omp parallel for schedule(dynamic, 64)
        for(long long q = 0; q < ncol; ++q)
        {
            const int ci = px + 2 * (int)(q % cx), cj = py + 2 * (int)((q / cx) % cy), ck = pz + 2 * (int)(q / ((long long)cx * cy));
            const int c = ci + nx * (cj + ny * ck);
            for(int t = 6 * c; t < 6 * c + 6; ++t)
            {
                double F[3][3], H[9][9], K[144];
                for(int i = 0; i < 3; ++i)
                    for(int j = 0; j < 3; ++j) F[i][j] = S.defgrad[9 * (size_t)t + 3 * i + j];
                snh_hessian_psd(F, S.mu[t], S.lambda[t], H);
                element_stiffness(&S.dminv[9 * (size_t)t], S.vol[t], H, K);
                const int* tv = &S.tets[4 * (size_t)t];
                for(int a = 0; a < 4; ++a)
                    for(int b = a; b < 4; ++b) S.A.add(foff + tv[a], foff + tv[b], &K[(3 * a) * 12 + 3 * b], 12);
            }
        }
    }
    // values: contacts (s G^T G, G 3x12)
    for(int k = 0; k < npair; ++k)
    {
        double G[3][12], K[144];
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 12; ++j) G[i][j] = normal(p.seed, kStreamContact + 100, 64ull * k + 12 * i + j);
        const double s = p.contact_stiffness * std::exp(p.contact_sigma * normal(p.seed, kStreamContact + 200, k));
        for(int a = 0; a < 12; ++a)
            for(int b = 0; b < 12; ++b)
            {
                double acc = 0.0;
                for(int i = 0; i < 3; ++i) acc += G[i][a] * G[i][b];
                K[a * 12 + b] = s * acc;
            }
        const int* cv = &S.pairs[4 * (size_t)k];
        for(int a = 0; a < 4; ++a)
            for(int b = a; b < 4; ++b) S.A.add(foff + cv[a], foff + cv[b], &K[(3 * a) * 12 + 3 * b], 12);
    }
    // values: affine bodies
    for(int bdy = 0; bdy < nbody; ++bdy)
    {
        double Am[12][12], D[144];
        for(int i = 0; i < 12; ++i)
            for(int j = 0; j < 12; ++j) Am[i][j] = normal(p.seed, kStreamBody, 144ull * bdy + 12 * i + j);
        for(int a = 0; a < 12; ++a)
            for(int b = 0; b < 12; ++b)
            {
                double acc = a == b ? p.body_shift : 0.0;
                for(int k = 0; k < 12; ++k) acc += Am[k][a] * Am[k][b];
                D[a * 12 + b] = acc;
            }
        for(int q = 0; q < 4; ++q)
            for(int s = q; s < 4; ++s) S.A.add(4 * bdy + q, 4 * bdy + s, &D[(3 * q) * 12 + 3 * s], 12);
    }
    for(size_t k = 0; k < bpairs.size(); ++k)
    {
        double G[3][12], K[144], negK[144];
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 12; ++j) G[i][j] = normal(p.seed, kStreamBodyPair + 100, 64ull * k + 12 * i + j);
        const double s = p.contact_stiffness * std::exp(p.contact_sigma * normal(p.seed, kStreamBodyPair + 200, k));
        for(int a = 0; a < 12; ++a)
            for(int b = 0; b < 12; ++b)
            {
                double acc = 0.0;
                for(int i = 0; i < 3; ++i) acc += G[i][a] * G[i][b];
                K[a * 12 + b]    = s * acc;
                negK[a * 12 + b] = -s * acc;
            }
        const int A0 = 4 * bpairs[k].a, B0 = 4 * bpairs[k].b;
        for(int q = 0; q < 4; ++q)
            for(int t = q; t < 4; ++t)
            {
                S.A.add(A0 + q, A0 + t, &K[(3 * q) * 12 + 3 * t], 12);
                S.A.add(B0 + q, B0 + t, &K[(3 * q) * 12 + 3 * t], 12);
            }
        for(int q = 0; q < 4; ++q)
            for(int t = 0; t < 4; ++t) S.A.add(A0 + q, B0 + t, &negK[(3 * q) * 12 + 3 * t], 12);
    }
    for(size_t k = 0; k < couplings.size(); ++k)
    {
        double G[3][15], K[225];
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 15; ++j) G[i][j] = normal(p.seed, kStreamCoupling + 100, 64ull * k + 15 * i + j);
        const double s = p.contact_stiffness * std::exp(p.contact_sigma * normal(p.seed, kStreamCoupling + 200, k));
        for(int a = 0; a < 15; ++a)
            for(int b = 0; b < 15; ++b)
            {
                double acc = 0.0;
                for(int i = 0; i < 3; ++i) acc += G[i][a] * G[i][b];
                K[a * 15 + b] = s * acc;
            }
        const int B0 = 4 * couplings[k].body, V = foff + couplings[k].vertex;
        for(int q = 0; q < 4; ++q)
            for(int t = q; t < 4; ++t) S.A.add(B0 + q, B0 + t, &K[(3 * q) * 15 + 3 * t], 15);
        for(int q = 0; q < 4; ++q) S.A.add(B0 + q, V, &K[(3 * q) * 15 + 12], 15);
        S.A.add(V, V, &K[12 * 15 + 12], 15);
    }

    S.rows.resize(nb);
    for(int r = 0; r < block_rows; ++r)
        for(long long k = S.A.row_ptr[r]; k < S.A.row_ptr[r + 1]; ++k) S.rows[k] = r;
    S.rhs.resize(3 * (size_t)block_rows);
    for(size_t i = 0; i < S.rhs.size(); ++i) S.rhs[i] = normal(p.seed, kStreamRhs, i);

    S.sz.n_vertices = nv;
    S.sz.n_tets     = ntet;
    S.sz.n_tris     = ntri;
    S.sz.n_pairs    = npair;
    S.sz.n_bodies   = nbody;
    S.sz.block_rows = block_rows;
    S.sz.fem_offset = foff;
    S.sz.nb         = nb;
}

extern "C" int masx_gen_create(const masx_gen_params* p, masx_gen_system** out)
{
    try
    {
        if(!p || !out) throw std::invalid_argument("null argument");
        if(p->nx < 0 || p->ny < 0 || p->nz < 0 || p->n_bodies < 0 || p->n_contacts < 0)
            throw std::invalid_argument("negative size");
        auto s = std::make_unique<masx_gen_system>();
        s->p   = *p;
        build(*s);
        *out = s.release();
        return 0;
    }
    catch(const std::exception& e)
    {
        g_error = e.what();
        return 1;
    }
}

extern "C" int masx_gen_sizes_of(const masx_gen_system* s, masx_gen_sizes* out)
{
    if(!s || !out) return 1;
    *out = s->sz;
    return 0;
}

extern "C" int masx_gen_copy(const masx_gen_system* s, int field, void* dst)
{
    if(!s || !dst) return 1;
    auto cp = [&](const auto& v) {
        std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        return 0;
    };
    switch(field)
    {
        case MASX_GEN_POSITIONS: return cp(s->pos);
        case MASX_GEN_TETS: return cp(s->tets);
        case MASX_GEN_TRIS: return cp(s->tris);
        case MASX_GEN_PAIRS: return cp(s->pairs);
        case MASX_GEN_VALUES: return cp(s->A.val);
        case MASX_GEN_ROWS: return cp(s->rows);
        case MASX_GEN_COLS: return cp(s->A.col);
        case MASX_GEN_RHS: return cp(s->rhs);
        case MASX_GEN_DM_INV: return cp(s->dminv);
        case MASX_GEN_VOLUME: return cp(s->vol);
        case MASX_GEN_DEFGRAD: return cp(s->defgrad);
        case MASX_GEN_MU: return cp(s->mu);
        case MASX_GEN_LAMBDA: return cp(s->lambda);
        case MASX_GEN_MASS: return cp(s->mass);
        default: g_error = "unknown field"; return 1;
    }
}

extern "C" void masx_gen_destroy(masx_gen_system* s) { delete s; }
