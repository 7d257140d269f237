"""Adjoint-differentiated MAS-PCG solve behind a torch.autograd.Function.

The system is H(mu, lambda) x = b with H = M/dt^2 + sum_e K_e(F_e; mu_e,
lambda_e), the Stable Neo-Hookean element matrices of femEnergy.cu (first
Piola stress :441-459, PSD projection :1641-1717, element matrix
:1719-1855) assembled on the GPU (masx_fem_assemble) into the block structure
the solver was set up with.  Forward: assemble, build the MAS preconditioner,
MAS-PCG solve.  Backward, for a loss L(x): H is symmetric, so the adjoint
y = H^-1 dL/dx is one more MAS-PCG solve with the preconditioner already built,
and dL/dtheta_e = -y_e^T (dK_e/dtheta_e) x_e per element comes from
masx_fem_param_grad (the derivative of the PSD projection in closed form);
dL/db = y.
"""
from __future__ import annotations

import numpy as np

from . import native
from . import workloads as W


def mesh_structure(n_vertices, tets):
    """Upper 3x3 block structure of a tet mesh: every vertex's diagonal block
    and every tet edge (min, max), sorted by (row, col)."""
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    pairs = [np.stack([np.arange(n_vertices), np.arange(n_vertices)], 1)]
    for a in range(4):
        for b in range(a + 1, 4):
            pairs.append(np.stack([np.minimum(t[:, a], t[:, b]), np.maximum(t[:, a], t[:, b])], 1))
    key = np.unique(np.concatenate(pairs)[:, 0] * n_vertices + np.concatenate(pairs)[:, 1])
    return (key // n_vertices).astype(np.int32), (key % n_vertices).astype(np.int32)


class ElasticSystem:
    """A Stable Neo-Hookean FEM Newton system with a fixed mesh, solved by
    MAS-PCG on the GPU; mu and lambda (one per tet) are the inputs.

    tets (n_tets, 4) vertex ids; dminv (n_tets, 9) row-major Dm^-1; vol
    (n_tets,); F (n_tets, 9) row-major deformation gradients; mass
    (n_vertices,); positions (n_vertices, 3) for the Morton partition."""

    def __init__(self, tets, dminv, vol, F, mass, dt, positions, *, levels=0, bank=32, precision="fp32",
                 tol=1e-4, rule=native.RULE_TRUE_REL, max_iter=0, inner_tol=1e-4, max_refine=8):
        import torch

        tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        n = len(positions)
        rows, cols = mesh_structure(n, tets)
        self.n_vertices, self.n_tets, self.nb = n, len(tets), len(rows)
        self.precision, self.tol, self.rule, self.max_iter = precision, tol, rule, max_iter
        self.inner_tol, self.max_refine = inner_tol, max_refine
        self.fem = native.FemHessian(tets, rows, cols)
        self.solver = native.Solver()
        self.solver.add_system(block_rows=n, rows=rows, cols=cols, use_mas=True, bank=bank, levels=levels)
        start, num, lst = W.neighbors(n, tets)
        self.solver.set_topology_morton(0, start, num, lst, positions)
        self.solver.setup()
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
        self.dminv = dev(np.asarray(dminv).reshape(-1, 9))
        self.vol = dev(vol)
        self.F = dev(np.asarray(F).reshape(-1, 9))
        base = np.zeros((self.nb, 9))
        diag = rows == cols
        md = np.asarray(mass, np.float64)[rows[diag]] / (dt * dt)
        for q in (0, 4, 8):
            base[diag, q] = md
        self.base = dev(base)
        self.vals = torch.empty((self.nb, 9), dtype=torch.float64, device="cuda")
        self._built_for = None
        self.last_stats = None

    @classmethod
    def from_workload(cls, w, dt=1.0 / 60.0, **kw):
        """From a generated FEM workload (no contacts, no bodies)."""
        if w["n_bodies"] or w["n_pairs"]:
            raise ValueError("the elastic system covers FEM-only workloads")
        return cls(w["tets"], w["dm_inv"], w["volume"], w["defgrad"], w["mass"], dt, w["positions"], **kw)

    @staticmethod
    def _key(mu, lam):
        return (mu.data_ptr(), mu._version, lam.data_ptr(), lam._version)

    def build(self, mu, lam):
        """Assemble H(mu, lambda) and build its preconditioner."""
        self.fem.assemble(self.dminv, self.vol, self.F, mu.detach(), lam.detach(), base=self.base, out=self.vals)
        self.solver.set_values(0, self.vals)
        self.solver.build(self.precision)
        self._built_for = self._key(mu, lam)

    def is_built_for(self, mu, lam):
        return self._built_for == self._key(mu, lam)

    def solve(self, rhs):
        """x = H^-1 rhs to relative residual tol.  Tolerances below inner_tol
        (which FP32 iterates cannot reach: x itself is FP32 inside the PCG)
        use FP64 iterative refinement: x and the residual b - A x stay FP64
        (the library's FP64 SpMV), each correction is an FP32 MAS-PCG solve
        of the current residual to what is still missing (tol over the
        current relative residual, with a margin of 2), but never tighter
        than inner_tol: a tolerance of 1e-5 then costs one full solve and a
        short correction."""
        import torch

        rhs = rhs.detach().to(torch.float64).contiguous()
        if self.tol >= self.inner_tol:
            x = torch.zeros_like(rhs)
            st = self.solver.solve(rhs, x, rule=self.rule, tol=self.tol, max_iter=self.max_iter)[0]
            st["refinements"] = 1
            self.last_stats = st
            if st["failure"] != 0 or st["status"] != native.STATUS_DONE:
                raise RuntimeError(f"MAS-PCG failed: {st}")
            return x
        bnorm = float(torch.linalg.vector_norm(rhs))
        x = torch.zeros_like(rhs)
        r = rhs.clone()
        d = torch.empty_like(rhs)
        iters, rel, st = 0, 1.0, None
        for k in range(self.max_refine):
            d.zero_()
            step_tol = min(0.5, max(self.inner_tol, 0.5 * self.tol / rel)) if rel > 0 else self.inner_tol
            st = self.solver.solve(r, d, rule=native.RULE_TRUE_REL, tol=step_tol, max_iter=self.max_iter)[0]
            if st["failure"] != 0 or st["status"] != native.STATUS_DONE:
                raise RuntimeError(f"MAS-PCG failed in refinement step {k}: {st}")
            iters += st["iterations"]
            x += d
            r.copy_(rhs)
            self.solver.spmv(x, r, a=-1.0, b=1.0)
            rel = float(torch.linalg.vector_norm(r)) / bnorm if bnorm > 0 else 0.0
            if rel <= self.tol:
                break
        st = dict(st, iterations=iters, rel_res_true=rel, refinements=k + 1)
        self.last_stats = st
        if rel > self.tol:
            raise RuntimeError(f"refinement stopped at relative residual {rel:.3e} (target {self.tol:.1e})")
        return x

    def param_grad(self, mu, lam, x, y):
        return self.fem.param_grad(self.dminv, self.vol, self.F, mu.detach(), lam.detach(), x, y)


def _function():
    import torch

    class AdjointSolve(torch.autograd.Function):
        """x = H(mu, lambda)^-1 b; gradients for mu, lambda (per tet) and b."""

        @staticmethod
        def forward(ctx, mu, lam, system, b):
            system.build(mu, lam)
            x = system.solve(b)
            ctx.system = system
            ctx.save_for_backward(mu, lam, x)
            system.forward_stats = system.last_stats
            return x

        @staticmethod
        def backward(ctx, gx):
            mu, lam, x = ctx.saved_tensors
            s = ctx.system
            if not s.is_built_for(mu, lam):
                s.build(mu, lam)
            y = s.solve(gx)
            s.adjoint_stats = s.last_stats
            g_mu, g_lam = s.param_grad(mu, lam, x, y)
            return g_mu, g_lam, None, (y if ctx.needs_input_grad[3] else None)

    return AdjointSolve


_AdjointSolve = None


def solve(mu, lam, system, b):
    """Differentiable x = H(mu, lambda)^-1 b (torch CUDA float64 tensors)."""
    global _AdjointSolve
    if _AdjointSolve is None:
        _AdjointSolve = _function()
    return _AdjointSolve.apply(mu, lam, system, b)
