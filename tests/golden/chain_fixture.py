"""Synthetic SPD block chain with an exact solution.

Restates the recipe of reports/agipc_reproduction/make_coarse_replay_fixture.py
(block mass diag(1,2,3) plus 100*||x_i - R_i x_{i+1}||^2 couplings with
alternating +-0.13 rad z-rotations, reference solution of sines/cosines) and
the padding the replay harness applies (agipc/agipc_coarse_replay.cu:226-237:
pad to a multiple of 32 with identity diagonal blocks, identity partition,
neighbors from the off-diagonal pattern).

The reference holds the outcome of the 257-block instance
(reports/agipc_reproduction/AGIPC_COARSE_MAS32_REPLAY.md:41-54): block-Jacobi
PCG 47 iterations, MAS32 PCG 14 iterations, both to ||r|| <= 1e-3 ||b||.
"""
import numpy as np

REFERENCE_ITERATIONS = {"jacobi": 47, "mas32": 14}
REFERENCE_TRUE_REL_RES = {"jacobi": 9.431245e-4, "mas32": 9.428987e-4}


def chain_system(count=257):
    idx = np.arange(count, dtype=np.float64)
    solution = np.stack([np.sin(0.37 * idx) + 0.1 * idx / count,
                         np.cos(0.23 * idx),
                         np.sin(0.17 * idx + 0.3)], axis=1)
    diag = np.zeros((count, 3, 3))
    diag[:] = np.diag([1.0, 2.0, 3.0])
    off = np.zeros((count - 1, 3, 3))
    for i in range(count - 1):
        ang = -0.13 if i % 2 else 0.13
        c, s = np.cos(ang), np.sin(ang)
        off[i] = -100.0 * np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
        diag[i] += 100.0 * np.eye(3)
        diag[i + 1] += 100.0 * np.eye(3)
    rhs = np.einsum("nij,nj->ni", diag, solution)
    for i in range(count - 1):
        rhs[i] += off[i] @ solution[i + 1]
        rhs[i + 1] += off[i].T @ solution[i]
    blocks = np.concatenate([diag, off])              # row-major (n,3,3)
    rows = np.concatenate([np.arange(count), np.arange(count - 1)]).astype(np.int32)
    cols = np.concatenate([np.arange(count), np.arange(1, count)]).astype(np.int32)
    return blocks, rows, cols, rhs.reshape(-1), solution.reshape(-1)


def padded_chain(count=257, bank=32):
    """The system exactly as the replay harness hands it to MAS32."""
    blocks, rows, cols, rhs, sol = chain_system(count)
    padded = (count + bank - 1) // bank * bank
    extra = padded - count
    blocks = np.concatenate([blocks, np.tile(np.eye(3), (extra, 1, 1))])
    rows = np.concatenate([rows, np.arange(count, padded, dtype=np.int32)])
    cols = np.concatenate([cols, np.arange(count, padded, dtype=np.int32)])
    rhs = np.concatenate([rhs, np.zeros(3 * extra)])
    sol = np.concatenate([sol, np.zeros(3 * extra)])
    # neighbor CSR from the off-diagonal pattern, sorted and unique per vertex
    nbrs = [[] for _ in range(padded)]
    for r, c in zip(rows, cols):
        if r != c:
            nbrs[r].append(int(c))
            nbrs[c].append(int(r))
    start, num, lst = [], [], []
    for v in range(padded):
        nb = sorted(set(nbrs[v]))
        start.append(len(lst))
        num.append(len(nb))
        lst.extend(nb)
    colmajor = blocks.transpose(0, 2, 1).reshape(-1, 9).copy()
    ident = np.arange(padded, dtype=np.int32)
    return dict(vals=colmajor, rows=rows, cols=cols, rhs=rhs, solution=sol, n=padded, real=count,
                start=np.array(start, np.uint32), num=np.array(num, np.uint32),
                list=np.array(lst, np.uint32), part_to_real=ident, real_to_part=ident.copy())
