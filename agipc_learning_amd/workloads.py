"""Synthetic Newton systems for the five benchmark workloads.

Thin ctypes front end of the native generator (csrc/gen/masx_gen.cpp).  Each
system is returned as a dict of numpy arrays in the reference's matrix contract
(upper 3x3 block triplets, Eigen column-major values, block order
[4 rows per affine body][FEM vertices]).

Workloads (sizes follow the task's configuration list):
  fem8      8x8x8 hex grid, 729 vertices, 3,072 tets, constant E=1e5, no contact
  fem40c    40^3 grid, 68,921 vertices, lognormal E (sigma 0.5), 20,000 pairs
  hybrid    2,000 affine bodies, 30,000 body pairs, 20^3 FEM block coupled to 200 bodies
  batch64   64 cubes with U[5k,80k] vertices each, fem40c recipe, 5% of vertices in pairs
  fem60     60^3 grid, 226,981 vertices, 1,296,000 tets, lognormal E
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libmasx_gen.so")
_lib = None


class GenParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("spacing", C.c_double),
                ("young", C.c_double), ("young_sigma", C.c_double), ("poisson", C.c_double),
                ("density", C.c_double), ("dt", C.c_double), ("f_noise", C.c_double),
                ("n_contacts", C.c_int), ("contact_stiffness", C.c_double), ("contact_sigma", C.c_double),
                ("n_bodies", C.c_int), ("n_body_pairs", C.c_int), ("n_coupled_bodies", C.c_int),
                ("body_shift", C.c_double), ("seed", C.c_uint64)]


class GenSizes(C.Structure):
    _fields_ = [("n_vertices", C.c_int), ("n_tets", C.c_int), ("n_tris", C.c_int), ("n_pairs", C.c_int),
                ("n_bodies", C.c_int), ("block_rows", C.c_int), ("fem_offset", C.c_int), ("nb", C.c_longlong)]


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"generator library missing: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        L.masx_gen_default_params.argtypes = [C.POINTER(GenParams)]
        L.masx_gen_create.argtypes = [C.POINTER(GenParams), C.POINTER(C.c_void_p)]
        L.masx_gen_sizes_of.argtypes = [C.c_void_p, C.POINTER(GenSizes)]
        L.masx_gen_copy.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.masx_gen_destroy.argtypes = [C.c_void_p]
        L.masx_gen_last_error.restype = C.c_char_p
        L.masx_gen_element_hessian.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double),
                                               C.c_double, C.c_double, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def default_params(**kw) -> GenParams:
    p = GenParams()
    _load().masx_gen_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_FIELDS = {  # name: (id, dtype, shape function)
    "positions": (0, np.float64, lambda s: (s.n_vertices, 3)),
    "tets": (1, np.int32, lambda s: (s.n_tets, 4)),
    "tris": (2, np.int32, lambda s: (s.n_tris, 3)),
    "pairs": (3, np.int32, lambda s: (s.n_pairs, 4)),
    "vals": (4, np.float64, lambda s: (s.nb, 9)),
    "rows": (5, np.int32, lambda s: (s.nb,)),
    "cols": (6, np.int32, lambda s: (s.nb,)),
    "rhs": (7, np.float64, lambda s: (3 * s.block_rows,)),
    "dm_inv": (8, np.float64, lambda s: (s.n_tets, 9)),
    "volume": (9, np.float64, lambda s: (s.n_tets,)),
    "defgrad": (10, np.float64, lambda s: (s.n_tets, 9)),
    "mu": (11, np.float64, lambda s: (s.n_tets,)),
    "lam": (12, np.float64, lambda s: (s.n_tets,)),
    "mass": (13, np.float64, lambda s: (s.n_vertices,)),
}


def generate(params: GenParams, fields=None) -> dict:
    L = _load()
    h = C.c_void_p()
    if L.masx_gen_create(C.byref(params), C.byref(h)) != 0:
        raise RuntimeError("generator: " + L.masx_gen_last_error().decode())
    try:
        s = GenSizes()
        L.masx_gen_sizes_of(h, C.byref(s))
        out = dict(n_vertices=s.n_vertices, n_bodies=s.n_bodies, block_rows=s.block_rows,
                   fem_offset=s.fem_offset, nb=s.nb, n_pairs=s.n_pairs)
        names = fields if fields is not None else list(_FIELDS)
        for name in names:
            fid, dt, shp = _FIELDS[name]
            a = np.empty(shp(s), dtype=dt)
            if a.size:
                L.masx_gen_copy(h, fid, a.ctypes.data_as(C.c_void_p))
            out[name] = a
        return out
    finally:
        L.masx_gen_destroy(h)


def element_hessian(dm_inv, volume, F, mu, lam):
    L = _load()
    dm = np.ascontiguousarray(dm_inv, np.float64)
    f = np.ascontiguousarray(F, np.float64)
    H = np.zeros(144, np.float64)
    dp = C.POINTER(C.c_double)
    L.masx_gen_element_hessian(dm.ctypes.data_as(dp), float(volume), f.ctypes.data_as(dp), float(mu),
                               float(lam), H.ctypes.data_as(dp))
    return H.reshape(12, 12)


# ---------------------------------------------------------------- recipes --

def fem_grid(n, *, young_sigma=0.0, n_contacts=0, seed=0, spacing=0.0125):
    return generate(default_params(nx=n, ny=n, nz=n, young_sigma=young_sigma, n_contacts=n_contacts,
                                   seed=seed, spacing=spacing))


def fem8(seed=0):
    """8^3 cells, 729 vertices, plain SNH, no contact."""
    return fem_grid(8, seed=seed)


def fem40c(seed=0, n=40, n_contacts=20000):
    """40^3 cells, 68,921 vertices, lognormal E, 20,000 point-triangle pairs."""
    return fem_grid(n, young_sigma=0.5, n_contacts=n_contacts, seed=seed)


def fem60(seed=0, n=60):
    """60^3 cells, 226,981 vertices, 1,296,000 tets, lognormal E."""
    return fem_grid(n, young_sigma=0.5, seed=seed)


def hybrid(seed=0, n_bodies=2000, n_body_pairs=30000, fem_n=20, n_coupled=200):
    """Affine bodies + FEM block; block rows [4 per body][FEM vertices]."""
    return generate(default_params(nx=fem_n, ny=fem_n, nz=fem_n, n_bodies=n_bodies, n_body_pairs=n_body_pairs,
                                   n_coupled_bodies=n_coupled, seed=seed))


def batch64_sizes(seed=0, count=64, lo=5000, hi=80000):
    """Cells per axis of each cube: vertex count U[lo,hi], n = round(cbrt(V)) - 1."""
    rng = np.random.default_rng(seed)
    verts = rng.integers(lo, hi + 1, size=count)
    return [max(2, int(round(float(v) ** (1.0 / 3.0))) - 1) for v in verts]


def batch64(seed=0, count=64, lo=5000, hi=80000, contact_fraction=0.05):
    systems = []
    for i, n in enumerate(batch64_sizes(seed, count, lo, hi)):
        nv = (n + 1) ** 3
        pairs = int(round(contact_fraction * nv / 4))
        systems.append(fem_grid(n, young_sigma=0.5, n_contacts=pairs, seed=seed * 1000 + i + 1))
    return systems


def neighbors(n, tets, tris=None):
    """Vertex neighbor CSR (unique undirected tet/triangle edges), the input
    the MAS hierarchy expects (load_mesh.cpp:934-1040 builds the same sets)."""
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    pairs = [tets[:, [a, b]] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))]
    if tris is not None and len(tris):
        tris = np.asarray(tris, np.int64).reshape(-1, 3)
        pairs += [tris[:, [a, b]] for a, b in ((0, 1), (1, 2), (2, 0))]
    e = np.concatenate(pairs) if pairs else np.zeros((0, 2), np.int64)
    e = np.sort(e, axis=1)
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(e[:, 0] * n + e[:, 1])
    u, v = e // n, e % n
    src = np.concatenate([u, v])
    dst = np.concatenate([v, u])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    num = np.bincount(src, minlength=n).astype(np.uint32)
    start = np.zeros(n, np.uint32)
    start[1:] = np.cumsum(num)[:-1]
    return start, num, dst.astype(np.uint32)
