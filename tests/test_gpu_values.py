"""Singular values only (paco_svd_values_batched, LAPACK jobz = 'N') on the GPU.

Bars: normwise sigma error max|s - s*| / s*_max <= 1e-5 against FP64 LAPACK dgesdd, and
<= 1e-6 against the full-SVD path's sigma (the same algorithm: the values-only merges
form only the first and last right-vector columns each merge reads, in another order of
FP32 additions).  Graph replays are bitwise equal to the eager launch sequence.
"""
import sys

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

SIG_TOL = 1e-5


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    import pacosvd
    pacosvd.lib()
    return torch


SHAPES = [(1, 1), (2, 1), (3, 2), (5, 4), (9, 8), (40, 30), (64, 64), (100, 1), (129, 97), (300, 200), (512, 512),
          (700, 600), (1100, 1000)]


@pytest.mark.parametrize("kind", ["normal", "clustered", "graded", "zero"])
def test_values_vs_lapack_and_full_path(torch_dev, kind):
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(len(kind))
    mats = []
    for M, N in SHAPES:
        if kind == "normal":
            A = rng.standard_normal((M, N))
        elif kind == "zero":
            A = np.zeros((M, N))
        else:
            s = np.ones(N) + 1e-7 * rng.uniform(size=N) if kind == "clustered" else 10.0 ** (-6.0 * np.arange(N) / max(N - 1, 1))
            Q1 = workloads.haar_np(rng.standard_normal((M, N)))
            Q2 = workloads.haar_np(rng.standard_normal((N, N)))
            A = (Q1 * s[None, :]) @ Q2.T
        mats.append(A.astype(np.float32))
    say("values", kind, [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    S = pacosvd.svd_values(As)
    torch.cuda.synchronize()
    assert pacosvd.last_info(len(As)) == [0] * len(As)
    full = pacosvd.svd_batched(As)
    torch.cuda.synchronize()
    for i, A in enumerate(mats):
        s = S[i].cpu().numpy().astype(np.float64)
        ref = workloads.lapack_sigma(A)
        smax = max(float(ref[0]), 1e-30)
        err = float(np.max(np.abs(s - ref))) / smax if kind != "zero" else float(np.max(np.abs(s)))
        err_full = float(np.max(np.abs(s - full[i][0].cpu().numpy()))) / smax if kind != "zero" else 0.0
        assert np.all(np.diff(s) <= 0), (A.shape, "not descending")
        assert err <= SIG_TOL, (A.shape, err)
        assert err_full <= 1e-6, (A.shape, err_full)


def test_values_graph_replays_bitwise(torch_dev):
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(9)
    shapes = [(int(m), int(rng.integers(8, m + 1))) for m in rng.integers(16, 400, size=40)]
    As = [torch.from_numpy(rng.standard_normal(sh).astype(np.float32)).cuda() for sh in shapes]
    say("values graph", len(As))
    eager = pacosvd.svd_values(As)
    torch.cuda.synchronize()
    S = [torch.full_like(s, -1.0) for s in eager]
    for rep in range(3):  # the first call records the graph, the others replay it
        for x in S:
            x.fill_(-1.0)
        pacosvd.svd_values(As, S=S, graph=True)
        torch.cuda.synchronize()
        for a, b in zip(S, eager):
            assert torch.equal(a, b), rep
