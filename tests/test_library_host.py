"""CPU-side checks of the HIP library: it loads, exports every symbol the C
header declares, and its host-side MAS hierarchy equals the oracle's bit for
bit.  No device calls."""
import numpy as np
import pytest

from agipc_learning_amd import native
from agipc_learning_amd import workloads as W
from hier_ref import components_from_hierarchy
from test_oracle import _random_graph_case, kat35


def test_library_exports_every_declared_symbol():
    L = native.load()
    names = native.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def _compare(lib_h, orc, levels):
    assert lib_h["levels"] == levels
    o = orc.hierarchy()
    assert lib_h["total"] == orc.total_clusters
    assert np.array_equal(lib_h["level_count"][1:], o["level_count"][1:])
    assert np.array_equal(lib_h["level_begin"], o["level_begin"])
    assert np.array_equal(lib_h["fine_mask"], o["fine_mask"])
    assert np.array_equal(lib_h["prefix"], o["prefix"])
    assert np.array_equal(lib_h["coarse_table"], o["coarse_table"])
    n_fem = len(lib_h["fine_mask"])
    n_slots = int(o["level_begin"][1])
    T = orc.total_clusters
    gl, go = lib_h["going_next"], o["going_next"]
    assert np.array_equal(gl[:n_fem], go[:n_fem])
    assert np.array_equal(gl[n_slots:T], go[n_slots:T])


def test_hierarchy_reference_kat_library(oracle):
    valid, padded, start, num, lst, p2r, r2p = kat35()
    h = native.hierarchy_host(valid, padded, start, num, lst, p2r, r2p)
    assert list(h["prefix"]) == [1, 1]
    assert h["level_count"][1] == 2 and h["level_begin"][1] == padded
    assert all(h["fine_mask"][v] == (0xFFFFFFFF if v < 32 else 0x7) for v in range(valid))
    assert np.array_equal(h["going_next"][:valid], padded + np.array([0 if v < 32 else 1 for v in range(valid)]))


def test_component_lists_from_hierarchy_kat():
    # 35 vertices over two level-0 banks: one component each (all 32 lanes,
    # then lanes 0-2), whose level-1 nodes are slots 64 and 65; those two are
    # separate components of the level-1 bank; the top level lists nothing
    valid, padded, start, num, lst, p2r, r2p = kat35()
    h = native.hierarchy_host(valid, padded, start, num, lst, p2r, r2p)
    ptr, masks, parents = components_from_hierarchy(h, r2p, 32)
    assert h["levels"] == 3 and h["total"] == 128
    assert list(ptr) == [0, 1, 2, 4, 4]
    assert list(masks) == [0xFFFFFFFF, 0x7, 0x1, 0x2]
    assert list(parents) == [64, 65, 96, 97]


@pytest.mark.parametrize("seed,n,extra,npairs,scramble", [
    (0, 70, 10, 0, False), (1, 200, 60, 15, False), (2, 300, 400, 40, True), (3, 1000, 100, 30, False),
    (4, 129, 5, 8, True), (5, 2100, 300, 50, False), (6, 5000, 3000, 200, True)])
def test_library_hierarchy_equals_oracle_random(oracle, seed, n, extra, npairs, scramble):
    rng = np.random.default_rng(seed)
    edges, start, num, lst, p2r, r2p, pairs, n_slots = _random_graph_case(rng, n, extra, npairs, scramble)
    orc = oracle.Mas(n, n_slots, start, num, lst, p2r, r2p)
    orc.reorder(pairs)
    h = native.hierarchy_host(n, n_slots, start, num, lst, p2r, r2p, pairs=pairs)
    _compare(h, orc, orc.levels)


@pytest.mark.parametrize("n,contacts,bank,levels", [(8, 0, 32, 0), (12, 300, 32, 0), (10, 100, 16, 0),
                                                     (16, 500, 32, 4)])
def test_library_hierarchy_equals_oracle_grids(oracle, n, contacts, bank, levels):
    s = W.fem_grid(n, n_contacts=contacts, seed=n)
    nv = s["n_vertices"]
    p2r, r2p = oracle.morton_partition(s["positions"], bank)
    st, nu, li = oracle.mesh_neighbors(nv, s["tets"])
    orc = oracle.Mas(nv, len(p2r), st, nu, li, p2r, r2p, bank=bank, levels=levels)
    orc.reorder(s["pairs"])
    h = native.hierarchy_host(nv, len(p2r), st, nu, li, p2r, r2p, pairs=s["pairs"], bank=bank, levels=levels)
    _compare(h, orc, orc.levels)


@pytest.mark.parametrize("n,bank", [(6, 32), (13, 16), (20, 32)])
def test_library_morton_partition_equals_oracle(oracle, n, bank):
    s = W.fem_grid(n, seed=n)
    p2r, r2p = native.morton_partition(s["positions"], bank)
    q2r, s2p = oracle.morton_partition(s["positions"], bank)
    assert np.array_equal(p2r, q2r) and np.array_equal(r2p, s2p)


def test_mesh_structure_equals_the_generator_structure():
    from agipc_learning_amd import adjoint

    w = W.fem_grid(5, seed=1)
    rows, cols = adjoint.mesh_structure(w["n_vertices"], w["tets"])
    assert np.array_equal(rows, w["rows"]) and np.array_equal(cols, w["cols"])


def test_library_rejects_bad_input_with_an_error():
    with pytest.raises(RuntimeError):
        native.hierarchy_host(10, 5, np.zeros(10), np.zeros(10), np.zeros(1), np.arange(5), np.arange(10))


def _upper_of(rows, cols, perm=None):
    """Upper structure (r <= c, sorted, unique) after renumbering rows by perm."""
    r, c = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    if perm is not None:
        r, c = perm[r], perm[c]
    lo, hi = np.minimum(r, c), np.maximum(r, c)
    key = np.unique(lo * (1 << 32) + hi)
    return (key >> 32).astype(np.int32), (key & 0xffffffff).astype(np.int32)


@pytest.mark.parametrize("order", ["grid", "morton"])
@pytest.mark.parametrize("entries", [1, 2, 4, 8])
def test_spmv_layout_equals_the_full_csr_product(order, entries):
    # the 40^3 + 20,000-contact structure: the layout's product (pairs inside a
    # tile stored once, their transposed products summed by the column row)
    # equals the full both-triangle CSR product; every pair stored once saves
    # exactly one entry
    w = W.fem_grid(40, n_contacts=20000, seed=0)
    n = w["n_vertices"]
    perm = None
    if order == "morton":
        _, r2p = native.morton_partition(w["positions"], 32)
        perm = r2p.astype(np.int64)
    rows, cols = _upper_of(w["rows"], w["cols"], perm)
    err, n_ent, n_once = native.spmv_layout_check(n, rows, cols, entries, seed=3)
    off = int((rows != cols).sum())
    assert err <= 1e-13
    assert n_ent == n + 2 * off - n_once
    assert 0 < n_once < off
    if order == "morton" and entries == 8:
        assert n_once >= 0.55 * off  # Morton tiles keep most pairs local


def test_spmv_layout_long_rows_and_empty_rows():
    # a hub row coupled to every other row (longer than one chunk: a tile of
    # its own, several chunks), rows with only their diagonal, a tridiagonal
    # band, and the hub as the last row too
    n = 3000
    rows = [np.arange(n), np.arange(0, n - 1, 2), np.zeros(n - 1, np.int64), np.arange(n - 1)]
    cols = [np.arange(n), np.arange(1, n, 2), np.arange(1, n), np.full(n - 1, n - 1)]
    r, c = _upper_of(np.concatenate(rows), np.concatenate(cols))
    for entries in (1, 2, 8):
        err, n_ent, n_once = native.spmv_layout_check(n, r, c, entries, seed=entries)
        off = int((r != c).sum())
        assert err <= 1e-13
        assert n_ent == n + 2 * off - n_once


def test_spmv_layout_rejects_bad_structure():
    with pytest.raises(RuntimeError, match="upper structure"):
        native.spmv_layout_check(4, np.array([1, 0]), np.array([0, 0]), 2)
    with pytest.raises(RuntimeError, match="duplicate"):
        native.spmv_layout_check(4, np.array([0, 0]), np.array([1, 1]), 2)
    with pytest.raises(RuntimeError, match="entries per thread"):
        native.spmv_layout_check(4, np.array([0]), np.array([0]), 512)


@pytest.mark.parametrize("entries", [1, 2, 4, 8])
def test_spmv_layout_with_body_couplings(entries):
    # affine bodies lead the rows (4 per body): their couplings to bodies are
    # 12 x 12 blocks read through the full CSR, their other entries (to FEM
    # vertices) 3 x 3 entries of body tiles; the product equals the full CSR
    # product.  Also: a body coupled to many vertices (its tile spans several
    # chunks) and a body coupled to nothing but itself.
    w = W.hybrid(seed=4, n_bodies=60, n_body_pairs=300, fem_n=5, n_coupled=12)
    nb_ = w["n_bodies"]
    rows, cols = _upper_of(w["rows"], w["cols"])
    n = w["block_rows"]
    hub = 4 * 7  # body 7's first row coupled to every FEM vertex
    extra_c = np.arange(4 * nb_, n)
    rows = np.concatenate([rows, np.full(len(extra_c), hub)])
    cols = np.concatenate([cols, extra_c])
    keep = ~((rows >= 4 * 59) & (rows < 4 * 60) & (cols >= 0) & (cols < 4 * 59))  # body 59: self only
    keep &= ~((cols >= 4 * 59) & (cols < 4 * 60) & (rows < 4 * 59))
    rows, cols = _upper_of(rows[keep], cols[keep])
    err, n_ent, n_once = native.spmv_layout_check(n, rows, cols, entries, seed=entries, n_bodies=nb_)
    off = int((rows != cols).sum())
    in_bodies = (rows < 4 * nb_) & (cols < 4 * nb_)
    body_full = int(in_bodies.sum() + (in_bodies & (rows != cols)).sum())
    assert err <= 1e-13
    assert n_ent == n + 2 * off - body_full - n_once
