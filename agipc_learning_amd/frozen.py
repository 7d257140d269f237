"""Frozen linear systems on disk, in the two layouts the reference writes.

* Frozen Newton system (linear_system/linear_system/global_linear_system.cu:
  366-402): ``frozen_A_values.bin`` = Eigen::Matrix3d[nb] (9 doubles per
  block, column-major), ``frozen_A_rows.bin`` / ``frozen_A_cols.bin`` =
  int32[nb] block indices of the upper-triangular BCOO, ``frozen_b.bin`` =
  double[dof].  Raw little-endian arrays, no header.
* Coarse replay sample, format ``agipc_coarse_system_v1``
  (reports/agipc_reproduction/make_coarse_replay_fixture.py:40-52):
  ``coarse_A_values.f64x9.bin`` (column-major 3x3 blocks),
  ``coarse_A_rows.i32.bin``, ``coarse_A_cols.i32.bin``, ``coarse_rhs.f64.bin``,
  optional ``coarse_solution.f64.bin``, and ``metadata.json`` (block count,
  solve tolerance and iteration cap under ``coarse_solve``).

Values are returned as an (nb, 9) float64 array in the same column-major
block order, the layout masx_set_values takes.
"""
from __future__ import annotations

import json
import os

import numpy as np

FROZEN_FILES = ("frozen_A_values.bin", "frozen_A_rows.bin", "frozen_A_cols.bin", "frozen_b.bin")
COARSE_FORMAT = "agipc_coarse_system_v1"


def _read(path, dtype, count=-1):
    a = np.fromfile(path, dtype=dtype, count=count)
    return a


def _check_blocks(vals, rows, cols, rhs):
    nb = len(rows)
    if len(cols) != nb or vals.shape != (nb, 9):
        raise ValueError(f"inconsistent block arrays: values {vals.shape}, rows {len(rows)}, cols {len(cols)}")
    if rhs.size % 3:
        raise ValueError(f"right-hand side length {rhs.size} is not a multiple of 3")
    n = rhs.size // 3
    if nb and (rows.min() < 0 or cols.min() < 0 or rows.max() >= n or cols.max() >= n):
        raise ValueError("block index outside the right-hand side's block rows")
    return n


def read_frozen(directory):
    """Reads a frozen Newton system; returns dict(vals, rows, cols, rhs, block_rows)."""
    p = [os.path.join(directory, f) for f in FROZEN_FILES]
    vals = _read(p[0], "<f8")
    if vals.size % 9:
        raise ValueError(f"{p[0]}: {vals.size} doubles is not a whole number of 3x3 blocks")
    vals = vals.reshape(-1, 9)
    rows, cols = _read(p[1], "<i4"), _read(p[2], "<i4")
    rhs = _read(p[3], "<f8")
    n = _check_blocks(vals, rows, cols, rhs)
    return dict(vals=vals, rows=rows, cols=cols, rhs=rhs, block_rows=n)


def write_frozen(directory, vals, rows, cols, rhs):
    """Writes a frozen Newton system in the reference's layout."""
    vals = np.ascontiguousarray(vals, "<f8").reshape(-1, 9)
    rows, cols = np.ascontiguousarray(rows, "<i4"), np.ascontiguousarray(cols, "<i4")
    rhs = np.ascontiguousarray(rhs, "<f8").reshape(-1)
    _check_blocks(vals, rows, cols, rhs)
    os.makedirs(directory, exist_ok=True)
    for arr, name in zip((vals, rows, cols, rhs), FROZEN_FILES):
        arr.tofile(os.path.join(directory, name))


def read_coarse_sample(directory):
    """Reads an agipc_coarse_system_v1 sample; returns dict(vals, rows, cols,
    rhs, block_rows, solution or None, tol, max_iter, metadata)."""
    meta = json.load(open(os.path.join(directory, "metadata.json")))
    if meta.get("format") != COARSE_FORMAT:
        raise ValueError(f"not a {COARSE_FORMAT} sample: {meta.get('format')!r}")
    j = lambda name: os.path.join(directory, name)  # noqa: E731
    vals = _read(j("coarse_A_values.f64x9.bin"), "<f8").reshape(-1, 9)
    rows, cols = _read(j("coarse_A_rows.i32.bin"), "<i4"), _read(j("coarse_A_cols.i32.bin"), "<i4")
    rhs = _read(j("coarse_rhs.f64.bin"), "<f8")
    n = _check_blocks(vals, rows, cols, rhs)
    if "coarse_block_nodes" in meta and meta["coarse_block_nodes"] != n:
        raise ValueError("metadata block count does not match the right-hand side")
    sol = _read(j("coarse_solution.f64.bin"), "<f8") if os.path.exists(j("coarse_solution.f64.bin")) else None
    solve = meta.get("coarse_solve", {})
    return dict(vals=vals, rows=rows, cols=cols, rhs=rhs, block_rows=n, solution=sol,
                tol=solve.get("relative_tolerance", 1e-3), max_iter=solve.get("max_iterations", 0), metadata=meta)


def write_coarse_sample(directory, vals, rows, cols, rhs, *, solution=None, tol=1e-3, max_iter=None,
                        name=None, category="synthetic"):
    """Writes an agipc_coarse_system_v1 sample (with its manifest)."""
    vals = np.ascontiguousarray(vals, "<f8").reshape(-1, 9)
    rows, cols = np.ascontiguousarray(rows, "<i4"), np.ascontiguousarray(cols, "<i4")
    rhs = np.ascontiguousarray(rhs, "<f8").reshape(-1)
    n = _check_blocks(vals, rows, cols, rhs)
    os.makedirs(directory, exist_ok=True)
    j = lambda f: os.path.join(directory, f)  # noqa: E731
    vals.tofile(j("coarse_A_values.f64x9.bin"))
    rows.tofile(j("coarse_A_rows.i32.bin"))
    cols.tofile(j("coarse_A_cols.i32.bin"))
    rhs.tofile(j("coarse_rhs.f64.bin"))
    if solution is not None:
        np.ascontiguousarray(solution, "<f8").reshape(-1).tofile(j("coarse_solution.f64.bin"))
    meta = {"format": COARSE_FORMAT, "sample_name": name or os.path.basename(os.path.abspath(directory)),
            "category": category, "coarse_block_nodes": n, "coarse_unique_blocks": len(rows),
            "coarse_matrix_half_storage": True,
            "coarse_solve": {"max_iterations": int(max_iter if max_iter is not None else min(512, 3 * n)),
                             "relative_tolerance": float(tol)}}
    with open(j("metadata.json"), "w") as f:
        f.write(json.dumps(meta, indent=2) + "\n")
    with open(j("manifest.json"), "w") as f:
        f.write(json.dumps({"format": "agipc_coarse_samples_v1", "samples": [{**meta, "sample_name": "."}]},
                           indent=2) + "\n")
    return meta
