"""The MAS hierarchy built on the GPU (masx_hier_*) against the oracle and the
library's host build, bit for bit (integer work).

The per-bank component lists have no direct oracle output; they are derived
from the oracle's going_next and fine masks (hier_ref.components_from_hierarchy,
pinned on the known-answer case by a CPU test)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from agipc_learning_amd import native
from agipc_learning_amd import workloads as W
from hier_ref import components_from_hierarchy
from test_oracle import _random_graph_case, kat35

pytestmark = pytest.mark.gpu


def _cu(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def device_build(hier, n_fem, n_slots, start, num, lst, p2r, r2p, pairs=None, bank=32, levels=0, node_offset=0):
    pr = None
    if pairs is not None and len(pairs):
        pr = _cu(np.asarray(pairs, np.int32).reshape(-1, 4), np.int32)
    hier.build(n_fem, n_slots, _cu(start, np.uint32), _cu(num, np.uint32),
               _cu(lst if len(lst) else np.zeros(0), np.uint32), _cu(p2r, np.int32), _cu(r2p, np.int32),
               pairs=pr, node_offset=node_offset, bank=bank, levels=levels)
    return hier.download()


def compare(d, ref, r2p, bank):
    n_fem = len(ref["fine_mask"])
    T = int(ref["total"])
    n_slots = int(ref["level_begin"][1])
    assert d["levels"] == ref["levels"] and d["total"] == T
    assert np.array_equal(d["level_count"][1:], ref["level_count"][1:])
    assert np.array_equal(d["level_begin"], ref["level_begin"])
    assert np.array_equal(d["fine_mask"], ref["fine_mask"])
    assert np.array_equal(d["prefix"], ref["prefix"])
    assert np.array_equal(d["coarse_table"], ref["coarse_table"])
    assert np.array_equal(d["going_next"][:n_fem], ref["going_next"][:n_fem])
    assert np.array_equal(d["going_next"][n_slots:T], ref["going_next"][n_slots:T])
    ptr, masks, parents = components_from_hierarchy(ref, r2p, bank)
    assert np.array_equal(d["comp_ptr"], ptr)
    assert np.array_equal(d["comp_mask"], masks)
    assert np.array_equal(d["comp_parent"], parents)


def oracle_dict(orc):
    o = orc.hierarchy()
    o["levels"], o["total"] = orc.levels, orc.total_clusters
    return o


def test_device_hierarchy_reference_kat(oracle):
    valid, padded, start, num, lst, p2r, r2p = kat35()
    d = device_build(native.DeviceHierarchy(), valid, padded, start, num, lst, p2r, r2p)
    assert list(d["prefix"]) == [1, 1]
    assert d["level_count"][1] == 2 and d["level_begin"][1] == padded
    assert all(d["fine_mask"][v] == (0xFFFFFFFF if v < 32 else 0x7) for v in range(valid))
    # two level-0 banks, one component each, parents = the two level-1 slots
    assert list(d["comp_ptr"][:3]) == [0, 1, 2]
    assert list(d["comp_mask"][:2]) == [0xFFFFFFFF, 0x7]
    assert list(d["comp_parent"][:2]) == [padded, padded + 1]
    orc = oracle.Mas(valid, padded, start, num, lst, p2r, r2p)
    orc.reorder()
    compare(d, oracle_dict(orc), r2p, 32)


@pytest.mark.parametrize("seed,n,extra,npairs,scramble", [
    (0, 70, 10, 0, False), (1, 200, 60, 15, False), (2, 300, 400, 40, True), (3, 1000, 100, 30, False),
    (4, 129, 5, 8, True), (5, 2100, 300, 50, False), (6, 5000, 3000, 200, True)])
def test_device_hierarchy_equals_oracle_random(oracle, seed, n, extra, npairs, scramble):
    rng = np.random.default_rng(seed)
    edges, start, num, lst, p2r, r2p, pairs, n_slots = _random_graph_case(rng, n, extra, npairs, scramble)
    orc = oracle.Mas(n, n_slots, start, num, lst, p2r, r2p)
    orc.reorder(pairs)
    d = device_build(native.DeviceHierarchy(), n, n_slots, start, num, lst, p2r, r2p, pairs=pairs)
    compare(d, oracle_dict(orc), r2p, 32)


@pytest.mark.parametrize("n,contacts,bank,levels", [(8, 0, 32, 0), (12, 300, 32, 0), (10, 100, 16, 0),
                                                     (16, 500, 32, 4), (14, 200, 16, 6)])
def test_device_hierarchy_equals_oracle_grids(oracle, n, contacts, bank, levels):
    s = W.fem_grid(n, n_contacts=contacts, seed=n)
    nv = s["n_vertices"]
    p2r, r2p = oracle.morton_partition(s["positions"], bank)
    st, nu, li = oracle.mesh_neighbors(nv, s["tets"])
    orc = oracle.Mas(nv, len(p2r), st, nu, li, p2r, r2p, bank=bank, levels=levels)
    orc.reorder(s["pairs"])
    d = device_build(native.DeviceHierarchy(), nv, len(p2r), st, nu, li, p2r, r2p, pairs=s["pairs"], bank=bank,
                     levels=levels)
    compare(d, oracle_dict(orc), r2p, bank)


def test_device_hierarchy_node_offset_and_all_pair_kinds(oracle):
    # contacts given in global ids (vertices shifted by node_offset), every
    # sign pattern of the pair encoding, some pairs reaching outside the range
    s = W.fem_grid(10, n_contacts=0, seed=3)
    nv, off = s["n_vertices"], 40
    p2r, r2p = oracle.morton_partition(s["positions"], 32)
    st, nu, li = oracle.mesh_neighbors(nv, s["tets"])
    rng = np.random.default_rng(9)
    pairs = []
    for k in range(400):
        v = (rng.choice(nv + 20, size=4, replace=False) - 10 + off).astype(np.int64)
        kind = k % 6
        if kind == 0:
            p = [v[0], v[1], v[2], v[3]]
        elif kind == 1:
            p = [v[0], v[1], v[2], -v[3] - 1]
        elif kind == 2:
            p = [-v[0] - 1, v[1], -1, -1]
        elif kind == 3:
            p = [-v[0] - 1, -v[1] - 1, -v[2] - 1, -v[3] - 1]
        elif kind == 4:
            p = [-v[0] - 1, v[1], v[2], -1]
        else:
            p = [-v[0] - 1, -v[1] - 1, v[2], -v[3] - 1]
        pairs.append(p)
    pairs = np.array(pairs, np.int32)
    host = native.hierarchy_host(nv, len(p2r), st, nu, li, p2r, r2p, pairs=pairs, node_offset=off)
    d = device_build(native.DeviceHierarchy(), nv, len(p2r), st, nu, li, p2r, r2p, pairs=pairs, node_offset=off)
    compare(d, host, r2p, 32)


def test_device_hierarchy_large_grid_matches_host_and_reuses_scratch(oracle):
    # the 40^3 grid with 20k contacts, then a smaller system on the same object
    hier = native.DeviceHierarchy()
    for n, contacts in ((40, 20000), (12, 300)):
        s = W.fem_grid(n, n_contacts=contacts, seed=n)
        nv = s["n_vertices"]
        p2r, r2p = oracle.morton_partition(s["positions"], 32)
        st, nu, li = oracle.mesh_neighbors(nv, s["tets"])
        host = native.hierarchy_host(nv, len(p2r), st, nu, li, p2r, r2p, pairs=s["pairs"])
        d = device_build(hier, nv, len(p2r), st, nu, li, p2r, r2p, pairs=s["pairs"])
        compare(d, host, r2p, 32)
        arrays = hier.device_arrays()
        assert all(v for v in arrays.values())


def test_device_hierarchy_rejects_bad_input():
    valid, padded, start, num, lst, p2r, r2p = kat35()
    hier = native.DeviceHierarchy()
    bad = np.array(lst, np.uint32)
    bad[3] = valid + 5                     # neighbour id out of range
    with pytest.raises(RuntimeError, match="neighbor id"):
        device_build(hier, valid, padded, start, num, bad, p2r, r2p)
    r2p_bad = np.array(r2p, np.int32)
    r2p_bad[[0, 1]] = r2p_bad[[1, 0]]      # partition maps disagree
    with pytest.raises(RuntimeError, match="disagree"):
        device_build(hier, valid, padded, start, num, lst, p2r, r2p_bad)
    p2r_dup = np.array(p2r, np.int32)
    p2r_dup[valid] = p2r_dup[0]            # one vertex in two slots
    with pytest.raises(RuntimeError, match="disagree"):
        device_build(hier, valid, padded, start, num, lst, p2r_dup, r2p)
    # an edge listed at one end only
    asym = np.array(lst, np.uint32)
    v0 = 0
    nbrs0 = set(int(u) for u in asym[start[v0]:start[v0] + num[v0]])
    w = next(u for u in range(1, valid) if u not in nbrs0)
    asym[start[v0]] = w
    with pytest.raises(RuntimeError, match="not symmetric"):
        device_build(hier, valid, padded, start, num, asym, p2r, r2p)
    # the object is still usable afterwards
    d = device_build(hier, valid, padded, start, num, lst, p2r, r2p)
    assert list(d["prefix"]) == [1, 1]
