"""FP64 dense reference for the adjoint-differentiated elastic solve (small
meshes only): assembly from the oracle's element matrices, exact solves, the
adjoint gradient, and central finite differences of the loss.

Loss L = 1/2 ||x - x*||^2 with H(theta) x = b.  A finite difference perturbs
one tet's parameter by a relative step h: H' = H + P dK P^T (P picks the tet's
12 DOFs), so x' = x - H^-1 P (I + dK P^T H^-1 P)^-1 dK P^T x (Woodbury), and
L' - L = -(x - x*).c + 1/2 ||c||^2 with c = x - x' is formed without
cancellation."""
import numpy as np
import scipy.linalg as sl

import pyoracle as orc


def dense_system(w, mu, lam, dt=1.0 / 60.0):
    n = w["n_vertices"]
    diag = np.repeat(np.asarray(w["mass"], np.float64) / (dt * dt), 3)
    return orc.fem_assemble_dense(n, w["tets"], w["dm_inv"], w["volume"], w["defgrad"], mu, lam, diag=diag)


def tet_dofs(tv):
    return np.concatenate([np.arange(3 * v, 3 * v + 3) for v in tv])


def adjoint_gradient(w, mu, lam, b, xstar, dt=1.0 / 60.0):
    """(x, y, g_mu, g_lambda) with exact FP64 solves."""
    H = dense_system(w, mu, lam, dt)
    cf = sl.cho_factor(H)
    x = sl.cho_solve(cf, b)
    y = sl.cho_solve(cf, x - xstar)
    gm = np.zeros(len(mu))
    gl = np.zeros(len(mu))
    for e, tv in enumerate(np.asarray(w["tets"]).reshape(-1, 4)):
        d = tet_dofs(tv)
        gm[e], gl[e] = orc.element_param_grad(w["dm_inv"][e], w["volume"][e], w["defgrad"][e], mu[e], lam[e],
                                              x[d], y[d])
    return x, y, gm, gl, cf


def fd_gradient(w, mu, lam, b, xstar, tets_sel, h=1e-4, cf=None, dt=1.0 / 60.0):
    """Central differences (relative step h) of L for the selected tets:
    array (len(tets_sel), 2) of (dL/dmu, dL/dlambda)."""
    if cf is None:
        cf = sl.cho_factor(dense_system(w, mu, lam, dt))
    x = sl.cho_solve(cf, b)
    r0 = x - xstar
    tets = np.asarray(w["tets"]).reshape(-1, 4)
    out = np.zeros((len(tets_sel), 2))
    for k, e in enumerate(tets_sel):
        d = tet_dofs(tets[e])
        E = np.zeros((len(b), 12))
        E[d, np.arange(12)] = 1.0
        HinvP = sl.cho_solve(cf, E)
        K0 = orc.element_stiffness(w["dm_inv"][e], w["volume"][e], w["defgrad"][e], mu[e], lam[e])
        for col, which in enumerate(("mu", "lam")):
            dl = []
            for sgn in (1.0, -1.0):
                m, l_ = mu[e], lam[e]
                if which == "mu":
                    m = mu[e] * (1.0 + sgn * h)
                else:
                    l_ = lam[e] * (1.0 + sgn * h)
                dK = orc.element_stiffness(w["dm_inv"][e], w["volume"][e], w["defgrad"][e], m, l_) - K0
                M = np.eye(12) + dK @ HinvP[d, :]
                c = HinvP @ np.linalg.solve(M, dK @ x[d])
                dl.append(-r0 @ c + 0.5 * c @ c)
            step = h * (mu[e] if which == "mu" else lam[e])
            out[k, col] = (dl[0] - dl[1]) / (2.0 * step)
    return out
