"""Stable Neo-Hookean assembly, its parameter gradient and the adjoint-
differentiated MAS-PCG solve on the GPU, against the FP64 oracle.

Tolerances: element matrices and assembled values are FP64 on both sides and
follow different eigen routes (analytic eigensystem on the GPU, dense eigh in
the oracle): 1e-11 relative to the largest entry.  Per-tet gradients with the
same x, y: 1e-9 relative to the largest.  End to end, with both solves
refined in FP64 to true relative residual 1e-9 (FP32 MAS-PCG corrections),
the gradient is within 1e-6 (2-norm, relative) of the exact FP64 gradient
and within 1e-4 of central finite differences of the loss (h = 1e-4); with
single FP32 solves at 1e-4 it is within 2e-2."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from agipc_learning_amd import adjoint
from agipc_learning_amd import native
from agipc_learning_amd import workloads as W

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def fem_of(w):
    rows, cols = adjoint.mesh_structure(w["n_vertices"], w["tets"])
    return native.FemHessian(w["tets"], rows, cols), rows, cols


def test_fem_assemble_matches_the_generator_values():
    # 8^3 grid, lognormal E: mass/dt^2 + sum K_e, same block order as the
    # generator (whose structure is the mesh's)
    w = W.fem_grid(8, young_sigma=0.5, seed=2)
    fem, rows, cols = fem_of(w)
    assert np.array_equal(rows, w["rows"]) and np.array_equal(cols, w["cols"])
    base = np.zeros((len(rows), 9))
    diag = rows == cols
    for q in (0, 4, 8):
        base[diag, q] = w["mass"][rows[diag]] * 3600.0
    vals = fem.assemble(cu(w["dm_inv"]), cu(w["volume"]), cu(w["defgrad"]), cu(w["mu"]), cu(w["lam"]),
                        base=cu(base))
    got = host(vals)
    assert np.abs(got - w["vals"]).max() <= 1e-11 * np.abs(w["vals"]).max()


def test_fem_assemble_matches_the_oracle_dense_matrix(oracle):
    # far-from-identity deformation gradients (inverted tets included) on a 3^3 grid
    w = W.fem_grid(3, young_sigma=0.5, seed=5)
    rng = np.random.default_rng(1)
    F = w["defgrad"].copy()
    F[::3] = rng.standard_normal((len(F[::3]), 9))
    fem, rows, cols = fem_of(w)
    vals = host(fem.assemble(cu(w["dm_inv"]), cu(w["volume"]), cu(F), cu(w["mu"]), cu(w["lam"])))
    n = w["n_vertices"]
    A = np.zeros((3 * n, 3 * n))
    for r, c, v in zip(rows, cols, vals):
        B = v.reshape(3, 3, order="F")
        A[3 * r:3 * r + 3, 3 * c:3 * c + 3] += B
        if r != c:
            A[3 * c:3 * c + 3, 3 * r:3 * r + 3] += B.T
    ref = oracle.fem_assemble_dense(n, w["tets"], w["dm_inv"], w["volume"], F, w["mu"], w["lam"])
    assert np.abs(A - ref).max() <= 1e-11 * np.abs(ref).max()


def test_fem_assemble_is_deterministic():
    w = W.fem_grid(6, young_sigma=0.5, seed=3)
    fem, rows, cols = fem_of(w)
    args = [cu(w["dm_inv"]), cu(w["volume"]), cu(w["defgrad"]), cu(w["mu"]), cu(w["lam"])]
    a = host(fem.assemble(*args))
    b = host(fem.assemble(*args))
    assert np.array_equal(a, b)


def test_fem_param_grad_matches_the_oracle(oracle):
    w = W.fem_grid(8, young_sigma=0.5, seed=6)
    rng = np.random.default_rng(2)
    F = w["defgrad"].copy()
    F[::7] = rng.standard_normal((len(F[::7]), 9))
    fem, _, _ = fem_of(w)
    x = rng.standard_normal(3 * w["n_vertices"])
    y = rng.standard_normal(3 * w["n_vertices"])
    gm, gl = fem.param_grad(cu(w["dm_inv"]), cu(w["volume"]), cu(F), cu(w["mu"]), cu(w["lam"]), cu(x), cu(y))
    gm, gl = host(gm), host(gl)
    sel = rng.choice(len(gm), 300, replace=False)
    ref = np.zeros((len(sel), 2))
    tets = w["tets"]
    for k, e in enumerate(sel):
        d = np.concatenate([np.arange(3 * v, 3 * v + 3) for v in tets[e]])
        ref[k] = oracle.element_param_grad(w["dm_inv"][e], w["volume"][e], F[e], w["mu"][e], w["lam"][e], x[d], y[d])
    got = np.stack([gm[sel], gl[sel]], 1)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_fem_rejects_a_structure_without_the_tet_blocks():
    w = W.fem_grid(2, seed=0)
    rows, cols = adjoint.mesh_structure(w["n_vertices"], w["tets"])
    keep = np.ones(len(rows), bool)
    keep[np.nonzero(rows != cols)[0][0]] = False
    with pytest.raises(RuntimeError, match="no block"):
        native.FemHessian(w["tets"], rows[keep], cols[keep])


def test_adjoint_gradient_matches_fp64_and_finite_differences(oracle):
    # the config-0 mesh (8^3 grid, 729 vertices, 3,072 tets) with lognormal
    # Lame parameters; L = 1/2 ||x - x*||^2
    import fem_ref

    w = W.fem_grid(8, young_sigma=0.5, seed=0)
    rng = np.random.default_rng(7)
    xstar = rng.standard_normal(3 * w["n_vertices"])
    xr, yr, gm, gl, cf = fem_ref.adjoint_gradient(w, w["mu"], w["lam"], w["rhs"], xstar)
    ref = np.stack([gm, gl], 1)
    sel = rng.choice(len(gm), 100, replace=False)
    fd = fem_ref.fd_gradient(w, w["mu"], w["lam"], w["rhs"], xstar, sel, h=1e-4, cf=cf)
    for tol, bound_x, bound_g, bound_fd in ((1e-9, 1e-7, 1e-6, 1e-4), (1e-4, 1e-2, 2e-2, 2e-2)):
        system = adjoint.ElasticSystem.from_workload(w, tol=tol)
        mu = cu(w["mu"]).requires_grad_(True)
        lam = cu(w["lam"]).requires_grad_(True)
        x = adjoint.solve(mu, lam, system, cu(w["rhs"]))
        loss = 0.5 * ((x - cu(xstar)) ** 2).sum()
        loss.backward()
        assert system.adjoint_stats["rel_res_true"] <= tol * (1 + 1e-6)
        g = np.stack([host(mu.grad), host(lam.grad)], 1)
        assert np.linalg.norm(host(x) - xr) <= bound_x * np.linalg.norm(xr)
        assert np.linalg.norm(g - ref) <= bound_g * np.linalg.norm(ref)
        assert np.linalg.norm(g[sel] - fd) <= bound_fd * np.linalg.norm(fd)


def test_adjoint_backward_rebuilds_after_a_later_forward():
    # two forwards with different parameters, then the first one's backward:
    # the backward must use H of its own parameters
    w = W.fem_grid(5, young_sigma=0.5, seed=1)
    system = adjoint.ElasticSystem.from_workload(w, tol=1e-9)
    b = cu(w["rhs"])
    mu1 = cu(w["mu"]).requires_grad_(True)
    lam1 = cu(w["lam"]).requires_grad_(True)
    x1 = adjoint.solve(mu1, lam1, system, b)
    mu2 = cu(w["mu"] * 2.0).requires_grad_(True)
    lam2 = cu(w["lam"] * 2.0).requires_grad_(True)
    adjoint.solve(mu2, lam2, system, b)
    (0.5 * (x1 ** 2).sum()).backward()
    g_late = host(mu1.grad)
    mu1b = cu(w["mu"]).requires_grad_(True)
    lam1b = cu(w["lam"]).requires_grad_(True)
    x1b = adjoint.solve(mu1b, lam1b, system, b)
    (0.5 * (x1b ** 2).sum()).backward()
    g_fresh = host(mu1b.grad)
    assert np.linalg.norm(g_late - g_fresh) <= 1e-6 * np.linalg.norm(g_fresh)
