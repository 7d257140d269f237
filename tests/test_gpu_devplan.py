"""The structure analysis on the device equals the host analysis bit for bit.

For MAS systems without affine bodies, masx_setup builds the block CSR,
the diagonal sources, the Galerkin maps of every coarse level, the bank entry
lists and the SpMV tile layout on the device (the reference rebuilds the same
information on the GPU each Newton step: ReorderRealtime and
PrepareHessian_bcoo, TraditionalMAS32Preconditioner.cu:2189-2230,
2234-2525).  The host analysis (build_plan, build_spmv_layout) is pinned by
the oracle tests; here every array of the device build is compared with the
host build of the same structure, for every SpMV tiling, and the two solvers
must then solve bit for bit alike.  Structures: grids with and without
contacts, bank 16, a level override, duplicate blocks and blocks given in
the lower triangle, a structure change on the same solver, and a batch of
systems of different sizes."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from agipc_learning_amd import native
from agipc_learning_amd import session
from agipc_learning_amd import workloads as W

pytestmark = pytest.mark.gpu

ARRAYS = native.Solver.PLAN_ARRAYS


def say(msg):
    print(msg, flush=True)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def with_duplicates_and_lower(w, seed):
    """The same matrix with some blocks split into two duplicates (halves) and
    some off-diagonal blocks given as their transpose in the lower triangle."""
    rng = np.random.default_rng(seed)
    rows, cols = w["rows"].copy(), w["cols"].copy()
    vals = w["vals"].reshape(-1, 3, 3).copy()  # column-major 3x3 per block: vals[k][c][r]
    nb = len(rows)
    dup = rng.choice(nb, nb // 20, replace=False)
    half = vals[dup] * 0.5
    vals[dup] = half
    off = np.flatnonzero(rows != cols)
    low = rng.choice(off, len(off) // 10, replace=False)
    rows[low], cols[low] = cols[low].copy(), rows[low].copy()
    vals[low] = vals[low].transpose(0, 2, 1)
    out = dict(w)
    out["rows"] = np.concatenate([rows, w["rows"][dup]]).astype(np.int32)
    out["cols"] = np.concatenate([cols, w["cols"][dup]]).astype(np.int32)
    out["vals"] = np.ascontiguousarray(np.concatenate([vals, half]).reshape(-1, 9))
    return out


def solve(s, w):
    s.set_values(0, dev(w["vals"]))
    s.build("fp32")
    b = session.rhs_of(w)
    x = torch.zeros(len(b), dtype=torch.float64, device="cuda")
    st = s.solve(dev(b), x, rule=1, tol=1e-4)[0]
    torch.cuda.synchronize()
    return x.cpu().numpy(), st


def compare(sd, sh, label, sd_on_device=True):
    assert sd.device_structure_used() == sd_on_device and not sh.device_structure_used()
    assert sd.info() == sh.info()
    for u in (2, 4, 8):
        sd.set_spmv_tiling(u)
        sh.set_spmv_tiling(u)
        for name in ARRAYS:
            if not name.startswith("spmv") and u != 2:
                continue
            a, b = sd.plan_array(name), sh.plan_array(name)
            assert a.shape == b.shape, (label, u, name, a.shape, b.shape)
            assert np.array_equal(a, b), (label, u, name, int(np.flatnonzero(a != b)[0]))
    say(f"{label}: {len(ARRAYS)} arrays equal at 2/4/8 entries per thread "
        f"(nnz {sd.info()['nnz']}, coarse entries {sd.info()['n_gal']})")


CASES = [
    ("8^3", dict(n=8, contacts=0, seed=0), {}),
    ("16^3 + 600 contacts", dict(n=16, contacts=600, seed=3), {}),
    ("12^3 + 300 contacts, bank 16", dict(n=12, contacts=300, seed=4), dict(bank=16)),
    ("14^3 + 200 contacts, 2 levels", dict(n=14, contacts=200, seed=5), dict(levels=2)),
    ("10^3 + 100 contacts, duplicates and lower blocks", dict(n=10, contacts=100, seed=6, dup=True), {}),
]


@pytest.mark.parametrize("label,spec,kw", CASES, ids=[c[0] for c in CASES])
def test_device_structure_equals_host(label, spec, kw):
    w = W.fem_grid(spec["n"], n_contacts=spec["contacts"], seed=spec["seed"], young_sigma=0.5)
    if spec.get("dup"):
        w = with_duplicates_and_lower(w, spec["seed"])
    sh = session.describe(w, **kw)
    sh.set_device_structure(0)
    sh.setup()
    sd = session.describe(w, **kw)
    sd.setup()
    compare(sd, sh, label)
    xd, std = solve(sd, w)
    xh, sth = solve(sh, w)
    assert std["iterations"] == sth["iterations"]
    assert np.array_equal(xd, xh)
    assert std["rel_res_true"] <= 1e-4 * (1 + 1e-9)
    sd.close()
    sh.close()


def test_device_structure_at_the_40_cubed_size_and_after_a_contact_change():
    # the 40^3 grid with 20,000 contacts, then a new contact set on the same
    # solver (buffers reused), each against a fresh host analysis
    w1 = W.fem40c(seed=0)
    w2 = W.fem_grid(40, n_contacts=20000, seed=1, young_sigma=0.5)
    sd = session.describe(w1)
    sd.setup()
    for k, w in enumerate((w1, w2)):
        if k:
            sd.set_structure(0, w["rows"], w["cols"])
            sd.set_contacts(0, w["pairs"])
            sd.setup()
        sh = session.describe(w)
        sh.set_device_structure(0)
        sh.setup()
        compare(sd, sh, f"40^3 + 20k contacts (structure {k})")
        xd, std = solve(sd, w)
        xh, sth = solve(sh, w)
        assert std["iterations"] == sth["iterations"]
        assert np.array_equal(xd, xh)
        sh.close()
    sd.close()


def test_batch_device_structure_equals_host():
    # a batch of systems of different sizes (so different level counts and
    # coarse sizes; one without contacts), as one solver: the device analysis
    # keys the coarse levels by system and ranks each bank's entries, the
    # SpMV tiles stop at every system's end
    ws = [W.fem_grid(n, n_contacts=c, seed=30 + n, young_sigma=0.5)
          for n, c in ((3, 0), (12, 300), (7, 40), (16, 600), (5, 10))]
    sh = session.describe(ws)
    sh.set_device_structure(0)
    sh.setup()
    sd = session.describe(ws)
    sd.setup()
    compare(sd, sh, "batch of 5 systems")
    for s in (sd, sh):
        for i, w in enumerate(ws):
            s.set_values(i, dev(w["vals"]))
        s.build("fp32")
    b = session.rhs_of(ws)
    xs = []
    for s in (sd, sh):
        x = torch.zeros(len(b), dtype=torch.float64, device="cuda")
        st = s.solve(dev(b), x, rule=1, tol=1e-4)
        torch.cuda.synchronize()
        assert all(q["rel_res_true"] <= 1e-4 * (1 + 1e-9) for q in st)
        xs.append((x.cpu().numpy(), [q["iterations"] for q in st]))
    assert xs[0][1] == xs[1][1]
    assert np.array_equal(xs[0][0], xs[1][0])
    sd.close()
    sh.close()


@pytest.mark.parametrize("device_analysis", [1, 0])
def test_structure_from_device_arrays(device_analysis):
    """masx_set_structure_device: a batch whose new contact sets arrive as
    device arrays gives the same plan and solve as host arrays, with the
    device analysis (no host round trip of the structure) and with the host
    analysis (which downloads the device-held structure)."""
    ws = [W.fem_grid(n, n_contacts=c, seed=40 + n, young_sigma=0.5) for n, c in ((10, 200), (7, 40))]
    w2 = [W.fem_grid(n, n_contacts=c, seed=50 + n, young_sigma=0.5) for n, c in ((10, 300), (7, 60))]
    sd = session.describe(ws)
    sd.set_device_structure(device_analysis)
    sd.setup()
    stream = torch.cuda.Stream()
    for i, w in enumerate(w2):
        rows = torch.from_numpy(np.ascontiguousarray(w["rows"], np.int32)).cuda()
        cols = torch.from_numpy(np.ascontiguousarray(w["cols"], np.int32)).cuda()
        torch.cuda.synchronize()
        sd.set_structure_device(i, rows, cols, stream=stream.cuda_stream)
        sd.set_contacts(i, w["pairs"])
    torch.cuda.synchronize()
    sd.setup()
    assert sd.device_structure_used() == bool(device_analysis)
    sh = session.describe(w2)
    sh.set_device_structure(0)
    sh.setup()
    compare(sd, sh, f"device-held structure, device analysis {device_analysis}", bool(device_analysis))
    for s in (sd, sh):
        for i, w in enumerate(w2):
            s.set_values(i, dev(w["vals"]))
        s.build("fp32")
    b = session.rhs_of(w2)
    xs = []
    for s in (sd, sh):
        x = torch.zeros(len(b), dtype=torch.float64, device="cuda")
        st = s.solve(dev(b), x, rule=1, tol=1e-4)
        torch.cuda.synchronize()
        xs.append((x.cpu().numpy(), [q["iterations"] for q in st]))
    assert xs[0][1] == xs[1][1]
    assert np.array_equal(xs[0][0], xs[1][0])
    bad = torch.full((4,), 10**6, dtype=torch.int32, device="cuda")
    sd.set_structure_device(0, bad, bad)
    with pytest.raises(RuntimeError):
        sd.setup()  # out-of-range block indices, found by the analysis
    sd.close()
    sh.close()


def test_bodies_keep_the_host_analysis():
    h = W.hybrid(n_bodies=20, n_body_pairs=60, fem_n=4, n_coupled=5, seed=0)
    s = session.describe(h)
    s.setup()
    assert not s.device_structure_used()
    s.close()


def test_mas_object_takes_a_new_contact_set_on_the_same_solver():
    # the reference's per-step call (setPreconditioner_bcoo with the new
    # structure and contact pairs) on one MAS object: the object keeps its
    # solver and re-analyzes on the device; the apply equals that of a fresh
    # object given the new structure directly, bit for bit
    wa = W.fem_grid(12, n_contacts=200, seed=21, young_sigma=0.5)
    wb = W.fem_grid(12, n_contacts=500, seed=22, young_sigma=0.5)
    nv = wa["n_vertices"]
    p2r, r2p = native.morton_partition(wa["positions"])
    st, nu, li = W.neighbors(nv, wa["tets"])
    R = dev(np.random.default_rng(4).standard_normal(3 * nv))

    def fresh():
        m = native.MasPreconditioner(32)
        m.init_neighbor(nv, 0, len(li), len(p2r))
        m.upload(li, st, nu, p2r, r2p)
        m.init_matrix()
        return m

    def apply(m, w):
        n = len(w["rows"])
        pairs = dev(np.asarray(w["pairs"], np.int32).reshape(-1, 4))
        m.set_preconditioner_bcoo(dev(w["vals"]), dev(w["rows"]), dev(w["cols"]), dev(np.arange(n, dtype=np.int32)),
                                  0, n, pairs=pairs, cp_num=len(pairs))
        Z = torch.zeros_like(R)
        m.preconditioning(R, Z)
        torch.cuda.synchronize()
        return Z.cpu().numpy()

    m = fresh()
    za = apply(m, wa)
    zb = apply(m, wb)    # same object, new contact set
    zb_ref = apply(fresh(), wb)
    za2 = apply(m, wa)   # and back
    assert np.array_equal(zb, zb_ref)
    assert np.array_equal(za2, za)
    assert not np.array_equal(za, zb)
