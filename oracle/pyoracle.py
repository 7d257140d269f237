"""ctypes binding of the CPU oracle (oracle/masx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libmasx_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("failure", C.c_int), ("rz0", C.c_double),
                ("rz_last", C.c_double), ("rec_rel_res", C.c_double), ("true_rel_res", C.c_double)]


def _declare(L):
    dp, ip, up, fp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_uint), C.POINTER(C.c_float)
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.argtypes = []
    L.orc_sym_bcoo_spmv.argtypes = [C.c_int, dp, ip, ip, C.c_double, dp, C.c_double, dp, C.c_int]
    L.orc_morton_partition.argtypes = [C.c_int, dp, C.c_int, ip, ip]
    L.orc_mesh_neighbors.argtypes = [C.c_int, C.c_int, ip, C.c_int, ip, up, up, up, C.c_int]
    L.orc_num_levels.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_hierarchy_capacity.argtypes = [C.c_longlong] * 4
    L.orc_hierarchy_capacity.restype = C.c_longlong
    L.orc_mas_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, up, up, up, ip, ip]
    L.orc_mas_create.restype = C.c_void_p
    L.orc_mas_destroy.argtypes = [C.c_void_p]
    L.orc_mas_reorder.argtypes = [C.c_void_p, C.c_int, ip, C.c_int]
    L.orc_mas_levels.argtypes = [C.c_void_p]
    L.orc_mas_total_clusters.argtypes = [C.c_void_p]
    L.orc_mas_get_hierarchy.argtypes = [C.c_void_p, ip, ip, up, ip, ip, ip]
    L.orc_mas_assemble.argtypes = [C.c_void_p, dp, ip, ip, up, C.c_int, C.c_int]
    L.orc_mas_invert.argtypes = [C.c_void_p, C.c_int]
    L.orc_mas_get_local.argtypes = [C.c_void_p, dp]
    L.orc_mas_get_inverse.argtypes = [C.c_void_p, fp]
    L.orc_mas_set_inverse.argtypes = [C.c_void_p, fp]
    L.orc_mas_apply.argtypes = [C.c_void_p, dp, dp]
    L.orc_inverse_gj_pivot.argtypes = [C.c_int, dp, dp]
    L.orc_abd_block_inverse.argtypes = [C.c_int, C.c_int, dp, ip, ip, dp]
    L.orc_diag3_inverse.argtypes = [C.c_int, C.c_int, dp, ip, ip, dp]
    L.orc_pcg.argtypes = [C.c_int, C.c_int, dp, ip, ip, dp, dp, C.c_int, C.c_int, dp, C.c_void_p,
                          C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(PcgStats)]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def sym_spmv(vals, rows, cols, x, a=1.0, b=0.0, y=None):
    vals, rows, cols, x = _f64(vals), _i32(rows), _i32(cols), _f64(x)
    y = np.zeros_like(x) if y is None else _f64(y).copy()
    lib().orc_sym_bcoo_spmv(len(rows), _p(vals, C.c_double), _p(rows, C.c_int), _p(cols, C.c_int),
                            a, _p(x, C.c_double), b, _p(y, C.c_double), len(x))
    return y


def morton_partition(pos, bank=32):
    pos = _f64(pos)
    n = pos.shape[0]
    padded = (n + bank - 1) // bank * bank
    p2r = np.empty(padded, np.int32)
    r2p = np.empty(n, np.int32)
    got = lib().orc_morton_partition(n, _p(pos, C.c_double), bank, _p(p2r, C.c_int), _p(r2p, C.c_int))
    assert got == padded
    return p2r, r2p


def mesh_neighbors(n, tets, tris=None):
    tets = _i32(tets).reshape(-1)
    tris = _i32(np.zeros((0, 3)) if tris is None else tris).reshape(-1)
    start = np.zeros(n, np.uint32)
    num = np.zeros(n, np.uint32)
    cap = 2 * (len(tets) // 4 * 6 + len(tris) // 3 * 3) + 1
    lst = np.zeros(cap, np.uint32)
    got = lib().orc_mesh_neighbors(n, len(tets) // 4, _p(tets, C.c_int), len(tris) // 3, _p(tris, C.c_int),
                                   _p(start, C.c_uint), _p(num, C.c_uint), _p(lst, C.c_uint), cap)
    assert got >= 0
    return start, num, lst[:got].copy()


def num_levels(vert_num, part_map_size, bank=32):
    return lib().orc_num_levels(vert_num, part_map_size, bank)


def hierarchy_capacity(vertices, mapped, levels, bank):
    return lib().orc_hierarchy_capacity(vertices, mapped, levels, bank)


class Mas:
    """Oracle MAS preconditioner (restatement of TraditionalMAS32Preconditioner)."""

    def __init__(self, n_fem, n_slots, start, num, lst, p2r, r2p, bank=32, levels=0):
        self._keep = [_u32(start), _u32(num), _u32(lst) if len(lst) else np.zeros(1, np.uint32),
                      _i32(p2r), _i32(r2p)]
        s, nn, ll, a, b = self._keep
        self.h = lib().orc_mas_create(n_fem, n_slots, bank, levels, _p(s, C.c_uint), _p(nn, C.c_uint),
                                      _p(ll, C.c_uint), _p(a, C.c_int), _p(b, C.c_int))
        if not self.h:
            raise ValueError("orc_mas_create rejected the arguments")
        self.n_fem, self.n_slots, self.bank = n_fem, n_slots, bank

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mas_destroy(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().orc_mas_levels(self.h)

    @property
    def total_clusters(self):
        return lib().orc_mas_total_clusters(self.h)

    def reorder(self, pairs=None, node_offset=0):
        pairs = _i32(np.zeros((0, 4)) if pairs is None else pairs).reshape(-1)
        return lib().orc_mas_reorder(self.h, len(pairs) // 4, _p(pairs, C.c_int) if len(pairs) else None,
                                     node_offset)

    def hierarchy(self):
        L, T = self.levels, self.total_clusters
        lc = np.zeros(L + 1, np.int32)
        lb = np.zeros(L + 1, np.int32)
        fm = np.zeros(self.n_fem, np.uint32)
        pr = np.zeros(self.n_slots // self.bank, np.int32)
        gn = np.zeros(max(T, 1), np.int32)
        ct = np.zeros((self.n_fem, 6), np.int32)
        lib().orc_mas_get_hierarchy(self.h, _p(lc, C.c_int), _p(lb, C.c_int), _p(fm, C.c_uint),
                                    _p(pr, C.c_int), _p(gn, C.c_int), _p(ct, C.c_int))
        return dict(level_count=lc, level_begin=lb, fine_mask=fm, prefix=pr, going_next=gn,
                    coarse_table=ct)

    def assemble(self, vals, rows, cols, indices=None, offset=0):
        vals, rows, cols = _f64(vals), _i32(rows), _i32(cols)
        idx = _u32(indices) if indices is not None else None
        count = len(idx) if idx is not None else len(rows)
        lib().orc_mas_assemble(self.h, _p(vals, C.c_double), _p(rows, C.c_int), _p(cols, C.c_int),
                               _p(idx, C.c_uint), offset, count)

    def invert(self, method=0):
        lib().orc_mas_invert(self.h, method)

    def local(self):
        nblk = self.bank * (self.bank + 1) // 2
        out = np.zeros((self.total_clusters // self.bank, nblk, 9), np.float64)
        lib().orc_mas_get_local(self.h, _p(out, C.c_double))
        return out

    def inverse(self):
        nblk = self.bank * (self.bank + 1) // 2
        out = np.zeros((self.total_clusters // self.bank, nblk, 9), np.float32)
        lib().orc_mas_get_inverse(self.h, _p(out, C.c_float))
        return out

    def set_inverse(self, inv):
        inv = np.ascontiguousarray(inv, dtype=np.float32)
        lib().orc_mas_set_inverse(self.h, _p(inv, C.c_float))

    def apply(self, r):
        r = _f64(r)
        z = np.zeros_like(r)
        lib().orc_mas_apply(self.h, _p(r, C.c_double), _p(z, C.c_double))
        return z


def inverse_gj_pivot(a):
    a = _f64(a)
    n = a.shape[0]
    inv = np.zeros_like(a)
    lib().orc_inverse_gj_pivot(n, _p(a, C.c_double), _p(inv, C.c_double))
    return inv


def abd_block_inverse(nbody, vals, rows, cols):
    vals, rows, cols = _f64(vals), _i32(rows), _i32(cols)
    inv = np.zeros((nbody, 12, 12), np.float64)
    lib().orc_abd_block_inverse(nbody, len(rows), _p(vals, C.c_double), _p(rows, C.c_int),
                                _p(cols, C.c_int), _p(inv, C.c_double))
    return inv


def diag3_inverse(block_rows, vals, rows, cols):
    vals, rows, cols = _f64(vals), _i32(rows), _i32(cols)
    inv = np.zeros((block_rows, 3, 3), np.float64)
    lib().orc_diag3_inverse(block_rows, len(rows), _p(vals, C.c_double), _p(rows, C.c_int),
                            _p(cols, C.c_int), _p(inv, C.c_double))
    return inv


RULE_REF_RTZ = 0
RULE_TRUE_REL = 1


def pcg(block_rows, vals, rows, cols, b, *, global_diag=False, n_abd=0, abd_inv=None, mas=None,
        fem_offset=0, tol=1e-4, rule=RULE_REF_RTZ, max_iter=0):
    vals, rows, cols, b = _f64(vals), _i32(rows), _i32(cols), _f64(b)
    x = np.zeros_like(b)
    st = PcgStats()
    ai = _f64(abd_inv) if abd_inv is not None else None
    lib().orc_pcg(block_rows, len(rows), _p(vals, C.c_double), _p(rows, C.c_int), _p(cols, C.c_int),
                  _p(b, C.c_double), _p(x, C.c_double), int(global_diag), n_abd, _p(ai, C.c_double),
                  mas.h if mas is not None else None, fem_offset, tol, rule, max_iter, C.byref(st))
    return x, dict(iterations=st.iterations, failure=st.failure, rz0=st.rz0, rz_last=st.rz_last,
                   rec_rel_res=st.rec_rel_res, true_rel_res=st.true_rel_res)


def convert_triplets(rows, cols, vals, drop_lower=False):
    """Triplets -> unique blocks sorted by (row, col), duplicates summed in
    input order (Converter::convert, linear_system/utils/converter.cu:137-284:
    stable sort on (row << 32) + col, then a segmented sum); drop_lower keeps
    row <= col only (Converter::ge2sym, :286-336).  Values: (n, 9) col-major.
    Returns (rows, cols, vals) of the unique blocks."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64).reshape(-1, 9)
    keep = np.ones(len(rows), bool) if not drop_lower else rows <= cols
    idx = np.nonzero(keep)[0]
    key = (rows[idx] << 32) + cols[idx]
    order = idx[np.argsort(key, kind="stable")]
    k = (rows[order] << 32) + cols[order]
    if len(order) == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 9))
    head = np.ones(len(order), bool)
    head[1:] = k[1:] != k[:-1]
    starts = np.nonzero(head)[0]
    ends = np.append(starts[1:], len(order))
    out = np.zeros((len(starts), 9))
    for u, (s0, s1) in enumerate(zip(starts, ends)):
        acc = np.zeros(9)
        for t in order[s0:s1]:  # sequential, input order
            acc = acc + vals[t]
        out[u] = acc
    return rows[order[starts]].astype(np.int32), cols[order[starts]].astype(np.int32), out


# ------------------------------------------------- Stable Neo-Hookean (FEM) --
# Element Hessian of the reference's Stable Neo-Hookean tets
# (femEnergy.cu:432-459 first Piola stress P = u F + r (J - 1 - u/r) cof F with
# u = 4 mu / 3, r = lambda + 5 mu / 6 from gl_main.cu:586-590; Hessian made PSD
# by clamping its eigenvalues, femEnergy.cu:1641-1717; element matrix
# vol * PFPX^T H PFPX, femEnergy.cu:1719-1855, without the dt^2 factor: the
# system is M/dt^2 + K).  Restated directly in vec(F) coordinates (column-
# major, index i + 3 j) with a dense 9x9 symmetric eigen-decomposition; the
# library computes the same matrix through the analytic eigensystem instead.

_EPS3 = np.zeros((3, 3, 3))
for _i, _j, _k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    _EPS3[_i, _j, _k] = 1.0
    _EPS3[_i, _k, _j] = -1.0


def snh_rates(mu, lam):
    """(u, r) and their derivatives d(u, r)/d(mu, lambda)."""
    return 4.0 * mu / 3.0, lam + 5.0 * mu / 6.0


def snh_stress(F, mu, lam):
    """First Piola stress (femEnergy.cu:441-459)."""
    u, r = snh_rates(mu, lam)
    J = np.linalg.det(F)
    cof = J * np.linalg.inv(F).T
    return u * F + r * (J - 1.0 - u / r) * cof


def snh_hessian_parts(F):
    """(I9, g g^T, d2J, J) in vec(F) coordinates: H = u I + r g g^T + s d2J,
    s = r (J - 1) - u."""
    J = np.linalg.det(F)
    cof = J * np.linalg.inv(F).T
    g = cof.reshape(-1, order="F")
    d2 = np.einsum("ikm,jln,mn->ijkl", _EPS3, _EPS3, F)  # d2J / dF_ij dF_kl
    D2 = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    D2[i + 3 * j, k + 3 * l] = d2[i, j, k, l]
    return np.eye(9), np.outer(g, g), D2, J


def snh_hessian(F, mu, lam):
    """Unprojected 9x9 Hessian d vec(P) / d vec(F)."""
    I9, gg, D2, J = snh_hessian_parts(F)
    u, r = snh_rates(mu, lam)
    return u * I9 + r * gg + (r * (J - 1.0) - u) * D2


def snh_hessian_dtheta(F):
    """dH/dmu and dH/dlambda (H is linear in (mu, lambda) before projection)."""
    I9, gg, D2, J = snh_hessian_parts(F)
    # u = 4mu/3, r = lambda + 5mu/6, s = r (J-1) - u
    dmu = (4.0 / 3.0) * I9 + (5.0 / 6.0) * gg + ((5.0 / 6.0) * (J - 1.0) - 4.0 / 3.0) * D2
    dlam = gg + (J - 1.0) * D2
    return dmu, dlam


def project_psd(H):
    w, Q = np.linalg.eigh(0.5 * (H + H.T))
    return (Q * np.maximum(w, 0.0)) @ Q.T


def project_psd_derivative(H, dH):
    """Directional derivative of the PSD projection Q max(w, 0) Q^T along dH
    (Daleckii-Krein: Q (Gamma o Q^T dH Q) Q^T with divided differences of
    max(., 0))."""
    w, Q = np.linalg.eigh(0.5 * (H + H.T))
    f = np.maximum(w, 0.0)
    n = len(w)
    G = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            if abs(w[a] - w[b]) > 1e-12 * max(1.0, abs(w[a]), abs(w[b])):
                G[a, b] = (f[a] - f[b]) / (w[a] - w[b])
            else:
                G[a, b] = 1.0 if w[a] > 0 else 0.0
    return Q @ (G * (Q.T @ dH @ Q)) @ Q.T


def shape_gradient(dminv):
    """D (9x12): d vec(F) / d x for a tet, dF_ij / dx_{a,i} = G[a][j] with
    G[0] = -sum of the rows of Dm^-1, G[a] = row a-1 of Dm^-1
    (femEnergy.cu __computePFPX3D)."""
    Dm = np.asarray(dminv, np.float64).reshape(3, 3)
    G = np.vstack([-Dm.sum(axis=0), Dm])
    D = np.zeros((9, 12))
    for a in range(4):
        for i in range(3):
            for j in range(3):
                D[i + 3 * j, 3 * a + i] = G[a, j]
    return D


def element_stiffness(dminv, vol, F, mu, lam):
    """12x12 PSD element matrix vol * D^T proj(H) D (rows: vertex a, axis i)."""
    D = shape_gradient(dminv)
    return vol * D.T @ project_psd(snh_hessian(np.asarray(F).reshape(3, 3), mu, lam)) @ D


def element_param_grad(dminv, vol, F, mu, lam, x12, y12):
    """(-y^T dK/dmu x, -y^T dK/dlambda x) for one tet: the element term of
    dL/dtheta = -lambda^T (dH/dtheta) x of an adjoint solve."""
    F = np.asarray(F).reshape(3, 3)
    D = shape_gradient(dminv)
    H = snh_hessian(F, mu, lam)
    dmu, dlam = snh_hessian_dtheta(F)
    a, b = D @ np.asarray(x12), D @ np.asarray(y12)
    return (-vol * b @ project_psd_derivative(H, dmu) @ a,
            -vol * b @ project_psd_derivative(H, dlam) @ a)


def fem_assemble_dense(n, tets, dminv, vol, F, mu, lam, diag=None):
    """Dense (3n x 3n) sum of element matrices (+ diag on the diagonal)."""
    A = np.zeros((3 * n, 3 * n))
    for t, tv in enumerate(np.asarray(tets).reshape(-1, 4)):
        K = element_stiffness(dminv[t], vol[t], F[t], mu[t], lam[t])
        idx = np.concatenate([np.arange(3 * v, 3 * v + 3) for v in tv])
        A[np.ix_(idx, idx)] += K
    if diag is not None:
        A[np.diag_indices(3 * n)] += np.asarray(diag, np.float64)
    return A
