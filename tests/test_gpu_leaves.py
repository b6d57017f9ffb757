"""Divide-and-conquer leaves larger than the reference's 3 columns (one wavefront per leaf,
FP64 one-sided Jacobi) on the GPU.

Bars, against FP64 LAPACK dgesdd of the FP32 input: normwise sigma error and
orthogonality max|U^T U - I|, max|V^T V - I| <= 1e-5, residual ||A - U S Vt||_F /
(||A||_F max(M, N) eps32) <= 10.  A matrix with N < ls is a single leaf, so its sigma and
vectors are the leaf solver's alone (composed with Q and P^T).  Also: at every leaf size
the graph replay and a matrix solved alone give the batched path's bits, the plain path
(other bidiagonalization launch shapes) its sigma within the bar, a graph recorded
at one leaf size is not replayed at another, and entry point (b) reports a tiling blob
made for another leaf size with info = -3 while still solving the problem.
"""
import sys

import numpy as np
import pytest

import workloads
from svdcheck import measures

pytestmark = pytest.mark.gpu

SIG_TOL = 1e-5
ORTH_TOL = 1e-5
RESID_TOL = 10.0


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    import pacosvd
    pacosvd.lib()
    return torch


class leaf_size:
    def __init__(self, ls):
        self.ls = ls

    def __enter__(self):
        import pacosvd
        self.old = pacosvd.set_leaf_size(self.ls)
        assert pacosvd.set_leaf_size(0) == self.ls

    def __exit__(self, *exc):
        import pacosvd
        pacosvd.set_leaf_size(self.old)


def matrices(kind, shapes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for M, N in shapes:
        if kind == "normal":
            A = rng.standard_normal((M, N))
        elif kind == "zero":
            A = np.zeros((M, N))
        elif kind == "rankdef":  # rank N // 2: zero singular values inside every leaf
            A = rng.standard_normal((M, max(N // 2, 1))) @ rng.standard_normal((max(N // 2, 1), N))
        else:
            s = (np.ones(N) + 1e-7 * rng.uniform(size=N) if kind == "clustered"
                 else 10.0 ** (-6.0 * np.arange(N) / max(N - 1, 1)))
            A = (workloads.haar_np(rng.standard_normal((M, N))) * s[None, :]) @ workloads.haar_np(
                rng.standard_normal((N, N))).T
        out.append(A.astype(np.float32))
    return out


def check_all(mats, outs, what):
    for A, out in zip(mats, outs):
        m = measures(A, *(t.cpu().numpy() for t in out))
        if not np.any(A):
            m["sigma_err"] = float(np.max(np.abs(out[0].cpu().numpy())))
        assert m["sigma_err"] <= SIG_TOL, (what, A.shape, m)
        assert m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (what, A.shape, m)
        assert m["resid_ratio"] <= RESID_TOL, (what, A.shape, m)


SINGLE_LEAF = [(1, 1), (2, 2), (3, 2), (5, 4), (9, 7), (16, 16), (20, 17), (31, 31), (40, 31), (64, 24)]
MANY_LEAVES = [(33, 32), (64, 64), (100, 65), (129, 97), (200, 150), (300, 299), (520, 511)]


@pytest.mark.parametrize("ls", [6, 32])
@pytest.mark.parametrize("kind", ["normal", "graded", "clustered", "zero", "rankdef"])
def test_single_leaf_vs_lapack(torch_dev, kind, ls):
    """ls 32: every shape is one leaf (one wavefront); ls 6: leaves of the one-thread kernel."""
    torch = torch_dev
    import pacosvd
    mats = matrices(kind, SINGLE_LEAF, 101 + len(kind))
    say("single leaf", kind, ls, [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    with leaf_size(ls):
        outs = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
        assert pacosvd.last_info(len(As)) == [0] * len(As)
    check_all(mats, outs, "single leaf " + kind)


@pytest.mark.parametrize("ls", [4, 5, 6, 7, 8, 16, 24, 32])
@pytest.mark.parametrize("kind", ["normal", "graded", "clustered", "rankdef"])
def test_leaf_sizes_vs_lapack(torch_dev, ls, kind):
    torch = torch_dev
    import pacosvd
    mats = matrices(kind, MANY_LEAVES, 7 * ls + len(kind))
    say("leaf size", ls, kind)
    As = [torch.from_numpy(m).cuda() for m in mats]
    with leaf_size(ls):
        outs = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
        assert pacosvd.last_info(len(As)) == [0] * len(As)
    check_all(mats, outs, "ls %d %s" % (ls, kind))


@pytest.mark.parametrize("ls", [3, 6, 16, 32])
def test_paths_agree_at_each_leaf_size(torch_dev, ls):
    torch = torch_dev
    import pacosvd
    mats = matrices("normal", [(60, 40), (257, 200), (31, 31), (512, 512), (130, 129)], 5)
    As = [torch.from_numpy(m).cuda() for m in mats]
    say("paths", ls)
    with leaf_size(ls):
        ref = pacosvd.svd_batched(As)
        plain = pacosvd.svd_plain(As)
        graph = pacosvd.svd_batched(As, graph=True)
        alone = pacosvd.svd_batched([As[3]])
        torch.cuda.synchronize()
        for i in range(len(As)):
            for x, z in zip(ref[i], graph[i]):
                assert torch.equal(x, z), (ls, i)
            smax = float(ref[i][0][0])
            assert float((ref[i][0] - plain[i][0]).abs().max()) <= SIG_TOL * smax, (ls, i)
        for x, y in zip(alone[0], ref[3]):
            assert torch.equal(x, y), ls
    check_all(mats, ref, "paths ls %d" % ls)
    check_all(mats, plain, "plain ls %d" % ls)


def test_graph_not_replayed_across_leaf_sizes(torch_dev):
    torch = torch_dev
    import pacosvd
    mats = matrices("normal", [(100, 90), (300, 200)], 17)
    As = [torch.from_numpy(m).cuda() for m in mats]
    outs = pacosvd.alloc_outputs(As)
    res = {}
    for ls in (3, 32, 3, 32):
        with leaf_size(ls):
            pacosvd.svd_batched(As, outs=outs, graph=True)
            torch.cuda.synchronize()
            eager = pacosvd.svd_batched(As)
            torch.cuda.synchronize()
            for o, e in zip(outs, eager):
                for x, y in zip(o, e):
                    assert torch.equal(x, y), ls
            res.setdefault(ls, [t.clone() for o in outs for t in o])
    # the two leaf sizes do not give the same bits (different merge trees)
    assert not all(torch.equal(x, y) for x, y in zip(res[3], res[32]))


def test_entry_b_blob_of_another_leaf_size(torch_dev):
    torch = torch_dev
    import pacosvd
    A = np.random.default_rng(23).standard_normal((150, 120)).astype(np.float32)
    say("entry b blob check 150x120")
    with leaf_size(3):
        blob3 = pacosvd.tiling_blob(120)
    outs = {}
    with leaf_size(32):
        blob32 = pacosvd.tiling_blob(120)
        for name, blob in (("match", blob32), ("other", blob3)):
            r = pacosvd.upper_bidiagonalization(torch.from_numpy(A).cuda())
            d, u, vt = pacosvd.svd_DC(150, 120, r["a"], r["u"], r["vt"], r["d"], r["e"], ws=r["ws"], blob=blob)
            torch.cuda.synchronize()
            outs[name] = (pacosvd.entry_info(r["ws"], 150, 120), d.clone(), u.clone(), vt.clone())
    assert outs["match"][0] == 0 and outs["other"][0] == -3
    for x, y in zip(outs["match"][1:], outs["other"][1:]):
        assert torch.equal(x, y)  # the blob is only checked; the partition is the library's
    m = measures(A, *(t.cpu().numpy() for t in outs["match"][1:]))
    assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, m


def test_values_only_with_large_leaves(torch_dev):
    torch = torch_dev
    import pacosvd
    mats = matrices("graded", [(40, 30), (300, 256), (700, 600)], 29) + matrices("normal", [(33, 33), (512, 100)], 30)
    As = [torch.from_numpy(m).cuda() for m in mats]
    say("values only ls 32")
    with leaf_size(32):
        S = pacosvd.svd_values(As)
        full = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
    for i, A in enumerate(mats):
        ref = workloads.lapack_sigma(A)
        s = S[i].cpu().numpy().astype(np.float64)
        assert float(np.max(np.abs(s - ref))) / float(ref[0]) <= SIG_TOL, A.shape
        assert float(np.max(np.abs(s - full[i][0].cpu().numpy()))) / float(ref[0]) <= 1e-6, A.shape
