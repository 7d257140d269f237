"""The allocation-free Morton partition (masx_morton_partition_device_noalloc,
no counterpart in the reference) against the allocating one
(masx_morton_partition_device, itself equal to the host partition and the
oracle bit for bit): identical arrays on calls in a row through the C ABI,
each later call smaller than the one before (the static scratch keeps the
last call's contents), at the size limit and with a 16-slot bank; sizes over
the limit are an error, reported before any kernel runs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from agipc_learning_amd import native

pytestmark = pytest.mark.gpu


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.standard_normal((n, 3)) * [2.0, 0.5, 1.0] + [1.0, -3.0, 7.0]).cuda()


def allocating(pos, bank):
    dm = native.DeviceMorton()
    p2r, r2p = dm.partition(pos, bank)
    torch.cuda.synchronize()
    dm.close()
    return p2r.cpu().numpy(), r2p.cpu().numpy()


def noalloc(pos, bank, sentinel=-7):
    n = pos.shape[0]
    padded = (n + bank - 1) // bank * bank
    # outputs prefilled with a sentinel: every slot must be written
    p2r = torch.full((padded,), sentinel, dtype=torch.int32, device="cuda")
    r2p = torch.full((n,), sentinel, dtype=torch.int32, device="cuda")
    native.morton_partition_noalloc(pos, p2r, r2p, bank)
    torch.cuda.synchronize()
    return p2r.cpu().numpy(), r2p.cpu().numpy()


def check(pos, bank):
    hp, hr = allocating(pos, bank)
    dp, dr = noalloc(pos, bank)
    n = pos.shape[0]
    print(f"noalloc morton: n {n} bank {bank}", flush=True)
    assert np.array_equal(dp, hp)
    assert np.array_equal(dr, hr)
    assert np.array_equal(np.sort(dp[:n]), np.arange(n)) and np.all(dp[n:] == -1)


def test_noalloc_equals_the_allocating_partition_on_calls_in_a_row():
    # a large call fills the static scratch, then smaller calls reuse it
    for n, seed in ((200_000, 0), (70_001, 1), (33, 2)):
        check(cloud(n, seed), 32)


def test_noalloc_at_the_limit_and_bank_16():
    check(cloud(native.MORTON_NOALLOC_MAX, 3), 32)
    check(cloud(5_000, 4), 16)


def test_noalloc_rejects_more_vertices_than_its_scratch_holds():
    n = native.MORTON_NOALLOC_MAX + 1
    pos = cloud(n, 5)
    p2r = torch.full((n + 31,), -7, dtype=torch.int32, device="cuda")
    r2p = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="at most 524288 vertices"):
        native.morton_partition_noalloc(pos, p2r, r2p, 32)
    torch.cuda.synchronize()
    assert bool((p2r == -7).all()) and bool((r2p == -7).all())  # nothing was launched
    # and a valid call after the refusal still works
    check(cloud(1_000, 6), 32)
