"""Solve status, failure gates, restarts and inert post-stop work on the GPU.

The diagonal cases restate the device PCG self-test of the reference
(pro/candidate-paper-t57-mas-fused-init/StiffGIPC/agipc/agipc_device_pcg_tests.inl:8-106):
identity preconditioner, diagonal SPD matrices with known solutions, and the
failure gates (non-positive and non-finite curvature, invalid r.z, iteration
caps of 1 and 512, zero entries in b) with the vectors left as the last
complete iteration wrote them.  Tolerances: the reference's 1e-6 rr0 stop
(relative residual 1e-3) and a solution within 2e-3 of b / diag; counts are
exact.  The local cases restate the MAS local diagnostics of
TraditionalMAS32Preconditioner.cu:182-215 through masx_factor_status."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from agipc_learning_amd import native
from agipc_learning_amd import session
from agipc_learning_amd import workloads as W

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def diagonal_system(d, *, global_diag=False):
    """Block-diagonal solver over the scalar diagonal d (length 3 n): one 3x3
    diagonal block per block row, identity preconditioner (or 3x3 Jacobi)."""
    d = np.asarray(d, np.float64)
    n = len(d) // 3
    rows = cols = np.arange(n, dtype=np.int32)
    vals = np.zeros((n, 9))
    vals[:, 0], vals[:, 4], vals[:, 8] = d[0::3], d[1::3], d[2::3]
    s = native.Solver()
    s.add_system(block_rows=n, rows=rows, cols=cols, use_mas=False, global_diag=global_diag)
    s.setup()
    s.set_values(0, dev(vals))
    s.build("fp32")
    return s


def unit_diagonal(count):
    i = np.arange(count)
    return 1.0 + (i % 29) * 0.07, 0.2 + np.sin(0.21 * i)


@pytest.mark.parametrize("count,zero_entry", [(771, False), (258, False), (3, False), (771, True)])
def test_diagonal_spd_converges_to_the_exact_solution(count, zero_entry):
    hd, rhs = unit_diagonal(count)
    if zero_entry:
        rhs[0] = 0.0
    s = diagonal_system(hd)
    x = torch.zeros(count, dtype=torch.float64, device="cuda")
    st = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=1e-3, max_iter=512)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_OK, st
    xr = host(x)
    ref = rhs / hd
    assert np.linalg.norm(rhs - hd * xr) <= 1e-3 * np.linalg.norm(rhs) * (1 + 2e-6)
    assert np.linalg.norm(xr - ref) <= 2e-3 * np.linalg.norm(ref)
    assert 1 <= st["iterations"] <= 29  # 29 distinct eigenvalues


def test_one_eigenvalue_converges_in_one_iteration_and_later_work_is_inert():
    # the reference's batch_tail_convergence: one update converges; the rest
    # of a captured chunk (and the chunks launched ahead) leave x unchanged
    rhs = np.array([1.0, -2.0, 0.5])
    results = []
    for chunk, use_graph in ((2, True), (8, True), (8, False), (16, True)):
        s = diagonal_system(np.full(3, 3.0))
        x = torch.zeros(3, dtype=torch.float64, device="cuda")
        st = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=1e-3, chunk=chunk, use_graph=use_graph)[0]
        assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_OK and st["iterations"] == 1
        results.append(host(x))
    assert np.allclose(results[0], rhs / 3.0, rtol=1e-6, atol=0)
    assert all(np.array_equal(r, results[0]) for r in results)


def test_chunks_launched_after_the_stop_leave_the_result_unchanged():
    # A long solve makes the next solve on the same solver launch many chunks
    # ahead of its first status check; a right-hand side on one eigenvalue
    # stops after one update, so nearly all of them run after the stop.
    i = np.arange(771)
    hd = 10.0 ** (8.0 * i / 770)
    rhs = 0.2 + np.sin(0.21 * i)
    one = np.where(i == 0, 1.0, 0.0)
    s = diagonal_system(hd)
    x = torch.zeros(771, dtype=torch.float64, device="cuda")
    long = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=0.0, max_iter=400)[0]
    assert long["iterations"] == 400
    st = s.solve(dev(one), x, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    fresh = diagonal_system(hd)
    y = torch.zeros_like(x)
    st2 = fresh.solve(dev(one), y, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    assert st["iterations"] == st2["iterations"] == 1
    assert np.array_equal(host(x), host(y))
    assert np.allclose(host(x), one, rtol=1e-6, atol=0)


def test_nonpositive_curvature_stops_before_any_update():
    s = diagonal_system([-2.0, 1.0, 1.0])
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    st = s.solve(dev([1.0, 0.1, 0.1]), x, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_BAD_CURVATURE, st
    assert st["iterations"] == 0 and float(x.abs().max()) == 0.0


def test_nonfinite_curvature_stops_before_any_update():
    hd, rhs = unit_diagonal(6)
    hd[0] = np.nan
    s = diagonal_system(hd)
    x = torch.zeros(6, dtype=torch.float64, device="cuda")
    st = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_BAD_CURVATURE, st
    assert st["iterations"] == 0 and float(x.abs().max()) == 0.0


def test_nonpositive_rz_stops_before_any_update():
    # a 3x3 block-Jacobi preconditioner of a negative definite block makes
    # r.z < 0 at the first step (the reference's nonpositive_rz case)
    s = diagonal_system([-1.0, -2.0, -3.0], global_diag=True)
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    st = s.solve(dev([1.0, 2.0, 3.0]), x, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_BAD_RZ, st
    assert st["iterations"] == 0 and float(x.abs().max()) == 0.0


def test_nonfinite_rhs_is_a_failure_not_a_converged_solve():
    hd, rhs = unit_diagonal(9)
    rhs[4] = np.nan
    s = diagonal_system(hd)
    x = torch.zeros(9, dtype=torch.float64, device="cuda")
    st = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=1e-3)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_NONFINITE, st
    assert st["iterations"] == 0 and float(x.abs().max()) == 0.0


def test_iteration_cap_of_one():
    # two distinct eigenvalues cannot converge in one update
    s = diagonal_system([2.0, 4.0, 7.0])
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    st = s.solve(dev([1.0, 2.0, 3.0]), x, rule=native.RULE_TRUE_REL, tol=1e-3, max_iter=1)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_ITERATION_CAP, st
    assert st["iterations"] == 1 and float(x.abs().max()) > 0.0


@pytest.mark.parametrize("check_every", [0, 50])
def test_iteration_cap_of_512(check_every):
    # condition number 1e8, tolerance 0: only the cap can stop it (with
    # checks, every drift restart keeps the iteration count)
    i = np.arange(771)
    hd = 10.0 ** (8.0 * i / 770)
    rhs = 0.2 + np.sin(0.21 * i)
    s = diagonal_system(hd)
    x = torch.zeros(771, dtype=torch.float64, device="cuda")
    st = s.solve(dev(rhs), x, rule=native.RULE_TRUE_REL, tol=0.0, max_iter=512, check_every=check_every)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_ITERATION_CAP, st
    assert st["iterations"] == 512
    assert np.isfinite(host(x)).all()


def test_ref_rule_cap_reports_the_cap():
    i = np.arange(771)
    s = diagonal_system(10.0 ** (8.0 * i / 770))
    x = torch.zeros(771, dtype=torch.float64, device="cuda")
    st = s.solve(dev(0.2 + np.sin(0.21 * i)), x, rule=native.RULE_REF_RTZ, tol=0.0, max_iter=64)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_ITERATION_CAP, st
    assert st["iterations"] == 64


# --------------------------------------------------- restarts and drift ----

def test_long_solve_without_drift_does_not_restart():
    # no preconditioner on a stiff FEM grid: about 150 iterations (the
    # oracle's count); the FP32 recursion stays within 10 tol ||b|| of
    # b - A x at every check
    w = W.fem_grid(12, young_sigma=1.0, n_contacts=100, seed=21)
    s = session.solver_for(w, use_mas=False, global_diag=False)
    s.build("fp32")
    b = dev(session.rhs_of(w))
    x = torch.zeros_like(b)
    st = s.solve(b, x, rule=native.RULE_TRUE_REL, tol=1e-4, check_every=50)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_OK, st
    assert st["iterations"] > 100, st
    assert st["true_checks"] >= 2 and st["restarts"] == 0 and st["rejected"] == 0, st
    assert st["rel_res_true"] <= 1e-4 * (1 + 1e-6)


def test_tolerance_below_fp32_reach_restarts_and_rejects_candidates():
    # FP32 iterates floor the true residual well above 1e-8 ||b|| while the
    # recursive residual keeps falling: the checks see the drift and restart,
    # and recursive convergence candidates are rejected by the FP64 check
    w = W.fem_grid(10, young_sigma=0.5, n_contacts=100, seed=22)
    s = session.solver_for(w)
    s.build("fp32")
    b = dev(session.rhs_of(w))
    x = torch.zeros_like(b)
    st = s.solve(b, x, rule=native.RULE_TRUE_REL, tol=1e-8, max_iter=400, check_every=50)[0]
    assert st["status"] == native.STATUS_DONE and st["failure"] == native.FAIL_ITERATION_CAP, st
    assert st["restarts"] >= 1 and st["rejected"] >= 1, st
    assert st["rel_res_true"] <= 1e-4, st
    assert np.isfinite(host(x)).all()


# ------------------------------------------------------ buffers and layout --

def x_edges_only(w, start, num, lst):
    """The neighbour CSR restricted to edges along x: every Morton bank then
    falls apart into several components, so the hierarchy has more banks."""
    pos = w["positions"]
    src = np.repeat(np.arange(len(start)), num)
    keep = (pos[src, 1] == pos[lst, 1]) & (pos[src, 2] == pos[lst, 2])
    src, dst = src[keep], lst[keep]
    n2 = np.bincount(src, minlength=len(start)).astype(np.uint32)
    s2 = np.zeros(len(start), np.uint32)
    s2[1:] = np.cumsum(n2)[:-1]
    return s2, n2, dst.astype(np.uint32)


def test_inverse_buffer_grows_with_the_bank_count():
    # the same solver first holds the full mesh topology (fewer MAS banks),
    # then a sparser one (more banks): setup must grow the inverse storage,
    # and the solve must equal a fresh solver's bit for bit
    w = W.fem_grid(10, n_contacts=0, seed=11, young_sigma=0.5)
    full = W.neighbors(w["n_vertices"], w["tets"])
    sparse = x_edges_only(w, *full)
    p2r, r2p = native.morton_partition(w["positions"])

    def make(topo):
        s = native.Solver()
        s.add_system(block_rows=w["block_rows"], rows=w["rows"], cols=w["cols"], use_mas=True)
        s.set_topology(0, *topo, p2r, r2p)
        s.setup()
        return s

    def solve_with(s, prec):
        s.set_values(0, dev(w["vals"]))
        s.build(prec)
        b = session.rhs_of(w)
        x = torch.zeros(len(b), dtype=torch.float64, device="cuda")
        st = s.solve(dev(b), x, rule=1, tol=1e-4)[0]
        return host(x), st

    for prec in ("fp32", "fp16"):
        s = make(full)
        few = s.info()["n_banks"]
        solve_with(s, prec)
        s.set_topology(0, *sparse, p2r, r2p)
        s.setup()
        many = s.info()["n_banks"]
        assert many > few, (few, many)
        got, st = solve_with(s, prec)
        fresh = make(sparse)
        ref, st_ref = solve_with(fresh, prec)
        assert st["iterations"] == st_ref["iterations"] and st["failure"] == 0
        assert np.array_equal(got, ref)
        assert np.array_equal(s.inverse(), fresh.inverse())
        fresh.close()
        s.close()


def test_wrapper_rejects_wrong_dtype_and_size():
    w = W.fem_grid(4)
    s = session.solver_for(w)
    s.build("fp32")
    n = 3 * w["block_rows"]
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        s.solve(torch.zeros(n, dtype=torch.float32, device="cuda"), x)
    with pytest.raises(ValueError):
        s.solve(torch.zeros(n - 3, dtype=torch.float64, device="cuda"), x)
    with pytest.raises(ValueError):
        s.set_values(0, torch.zeros((len(w["rows"]) - 1, 9), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        s.set_values(0, np.zeros((len(w["rows"]), 8)))
    with pytest.raises(ValueError):
        s.apply_preconditioner(x, torch.zeros(n, dtype=torch.float32, device="cuda"))


# ---------------------------------------------------------- local checks ----

def one_bank_system(diag0=1.0, nan=False):
    """32 vertices in one bank (identity partition, a chain of neighbours),
    identity blocks on the diagonal: the bank's local matrix is I, or I with
    entry (0, 0) replaced (TraditionalMAS32Preconditioner.cu:182-215)."""
    n = 32
    rows = cols = np.arange(n, dtype=np.int32)
    vals = np.zeros((n, 9))
    vals[:, 0] = vals[:, 4] = vals[:, 8] = 1.0
    vals[0, 0] = np.nan if nan else diag0
    start = np.array([0] + [1 + 2 * (i - 1) for i in range(1, n)], np.uint32)
    num = np.array([1] + [2] * (n - 2) + [1], np.uint32)
    lst = []
    for i in range(n):
        if i > 0:
            lst.append(i - 1)
        if i < n - 1:
            lst.append(i + 1)
    s = native.Solver()
    s.add_system(block_rows=n, rows=rows, cols=cols, use_mas=True)
    p = np.arange(n, dtype=np.int32)
    s.set_topology(0, start, num, np.array(lst, np.uint32), p, p)
    s.setup()
    s.set_values(0, dev(vals))
    return s


@pytest.mark.parametrize("factor_waves", [1, 2])
@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
def test_local_identity_factors_cleanly(precision, factor_waves):
    # the fine bank's local matrix is I; the one coarse node above it (the
    # chain is one component) aggregates 32 identity blocks: 32 I, padded
    # slots identity (both MAS factor kernels)
    s = one_bank_system()
    s.set_factor_waves(factor_waves)
    s.build(precision)
    assert s.factor_status() == 0
    inv = s.inverse()
    nblk = 32 * 33 // 2
    want = np.zeros((inv.shape[0], nblk, 9), np.float32)
    for r in range(32):
        want[:, 32 * r - r * (r + 1) // 2 + r] = np.eye(3, dtype=np.float32).reshape(9)
    want[1:, 0] /= 32.0
    assert np.array_equal(inv, want)


@pytest.mark.parametrize("factor_waves", [1, 2])
@pytest.mark.parametrize("nan", [False, True])
def test_local_not_spd_or_not_finite_is_counted(nan, factor_waves):
    s = one_bank_system(diag0=-1.0, nan=nan)
    s.set_factor_waves(factor_waves)
    s.build("fp32")
    assert s.factor_status() >= 1
    s2 = one_bank_system(diag0=-1.0, nan=nan)
    s2.set_factor_waves(factor_waves)
    s2.build("fp16")
    assert s2.factor_status() >= 1
    if not nan:
        # floored pivots leave the modified matrix SPD: the stored half inverse
        # is finite and equals the FP32 one to half precision of the bank scale
        a, h = s.inverse(), s2.inverse()
        assert np.isfinite(h).all()
        for ab, hb in zip(a, h):
            assert np.abs(ab - hb).max() <= 2.0 ** -10 * np.abs(ab).max()
