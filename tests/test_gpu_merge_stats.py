"""Merge statistics of the last batched call (paco_merge_stats): realised deflation.

Checks independent of the kernels' arithmetic: a problem whose partition has L leaves has
L - 1 merges (leaf counts from GenerateTiling's blob); the pole of z is never deflated, so
k = k_left + k_right + 1 per merge; k <= n; the block-restricted product FLOPs never exceed
the dense ones; a zero matrix deflates every pole but that one (k = 1).
"""
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    import pacosvd
    pacosvd.lib()
    return torch


def leaves(N):
    import pacosvd
    blob = pacosvd.tiling_blob(N)
    return struct.unpack("<4I", blob[:16])[3]


@pytest.mark.parametrize("graph", [False, True])
def test_merge_stats_invariants(torch_dev, graph):
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(77)
    shapes = [(40, 30), (200, 150), (512, 511), (64, 8), (300, 299), (17, 17)]
    As = [torch.from_numpy(rng.standard_normal(sh).astype(np.float32)).cuda() for sh in shapes]
    say("merge stats", shapes, "graph" if graph else "eager")
    pacosvd.svd_batched(As, graph=graph)
    st = pacosvd.merge_stats()
    assert len(st) >= 1
    merges = sum(r["merges"] for r in st)
    assert merges == sum(leaves(N) - 1 for _, N in shapes)
    for r in st:
        assert r["sum_k"] == r["sum_k_left"] + r["sum_k_right"] + r["merges"], r
        assert r["merges"] <= r["sum_k"] <= r["sum_n"], r
        assert 0 < r["block_flops"] <= r["dense_flops"], r


def test_merge_stats_zero_matrix_deflates(torch_dev):
    torch = torch_dev
    import pacosvd
    As = [torch.zeros((100, 90), device="cuda")]
    pacosvd.svd_batched(As)
    st = pacosvd.merge_stats()
    for r in st:
        assert r["sum_k"] == r["merges"], r  # only the pole of z is kept
