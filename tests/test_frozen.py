"""Frozen-system files in the reference's two layouts: byte layout, round
trips, and a replay of the 257-block chain written to disk reproducing the
iteration count the reference holds (oracle PCG, CPU)."""
import json
import os

import numpy as np
import pytest

from agipc_learning_amd import frozen
from chain_fixture import REFERENCE_ITERATIONS, chain_system, padded_chain


def test_frozen_byte_layout(tmp_path):
    """One block B = [[1,2,3],[4,5,6],[7,8,9]] is stored column-major
    (Eigen::Matrix3d): 1,4,7,2,5,8,3,6,9; indices int32, rhs float64."""
    B = np.arange(1.0, 10.0).reshape(3, 3)
    frozen.write_frozen(tmp_path, B.T.reshape(1, 9), [0], [0], [1.0, 2.0, 3.0])
    raw = np.fromfile(tmp_path / "frozen_A_values.bin", "<f8")
    assert list(raw) == [1, 4, 7, 2, 5, 8, 3, 6, 9]
    assert os.path.getsize(tmp_path / "frozen_A_rows.bin") == 4
    assert os.path.getsize(tmp_path / "frozen_b.bin") == 24
    d = frozen.read_frozen(tmp_path)
    assert d["block_rows"] == 1 and np.array_equal(d["vals"].reshape(3, 3).T, B)


def test_frozen_round_trip_and_checks(tmp_path):
    d = padded_chain(65)
    frozen.write_frozen(tmp_path, d["vals"], d["rows"], d["cols"], d["rhs"])
    r = frozen.read_frozen(tmp_path)
    assert np.array_equal(r["vals"], d["vals"]) and np.array_equal(r["rows"], d["rows"])
    assert np.array_equal(r["rhs"], d["rhs"]) and r["block_rows"] == d["n"]
    with pytest.raises(ValueError):
        frozen.write_frozen(tmp_path, d["vals"], d["rows"], d["cols"], d["rhs"][:-3])  # index out of range
    np.zeros(10, "<f8").tofile(tmp_path / "frozen_A_values.bin")
    with pytest.raises(ValueError):
        frozen.read_frozen(tmp_path)


def test_coarse_sample_round_trip(tmp_path):
    blocks, rows, cols, rhs, sol = chain_system(257)
    vals = blocks.transpose(0, 2, 1).reshape(-1, 9)
    meta = frozen.write_coarse_sample(tmp_path, vals, rows, cols, rhs, solution=sol, tol=1e-3, name="chain")
    assert meta["coarse_block_nodes"] == 257 and meta["coarse_unique_blocks"] == 513
    assert meta["coarse_solve"] == {"max_iterations": 512, "relative_tolerance": 1e-3}
    m = json.load(open(tmp_path / "manifest.json"))
    assert m["format"] == "agipc_coarse_samples_v1" and m["samples"][0]["sample_name"] == "."
    r = frozen.read_coarse_sample(tmp_path)
    assert np.array_equal(r["vals"], vals) and np.array_equal(r["solution"], sol)
    assert r["tol"] == 1e-3 and r["max_iter"] == 512


def test_replay_from_disk_reproduces_reference_count(tmp_path, oracle):
    d = padded_chain()
    frozen.write_frozen(tmp_path, d["vals"], d["rows"], d["cols"], d["rhs"])
    r = frozen.read_frozen(tmp_path)
    x, st = oracle.pcg(r["block_rows"], r["vals"], r["rows"], r["cols"], r["rhs"], global_diag=True, tol=1e-3,
                       rule=oracle.RULE_TRUE_REL, max_iter=512)
    assert st["iterations"] == REFERENCE_ITERATIONS["jacobi"]
