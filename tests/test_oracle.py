"""The oracle (FP64 restatement of the reference path) pinned against LAPACK, closed
forms, an independent textbook algorithm and brute force."""
import hashlib
import json
import os

import numpy as np
import pytest
from scipy.linalg import lapack

import oracle
from svdcheck import measures
from golden.make_fixtures import c1_matrix, lanczos_bidiag

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def c1():
    return np.load(os.path.join(GOLD, "c1.npz"))


def test_fixture_matches_generator():
    A = c1_matrix()
    man = json.load(open(os.path.join(GOLD, "manifest.json")))
    assert hashlib.sha256(A.tobytes()).hexdigest() == man["c1_A_sha256"]
    assert np.array_equal(c1()["A"], A)


@pytest.mark.parametrize("x", [[3.0, 4.0], [-3.0, 4.0, 12.0], [0.0, 1.0, -2.0], [5.0], [2.0, 0.0, 0.0],
                               [-1e-3, 1e-3, 2e-3, -5e-4]])
def test_house_closed_form(x):
    x = np.array(x)
    v, beta, tau = oracle.house(x)
    n = len(x)
    if n == 1 or np.all(x[1:] == 0):
        assert tau == 0.0 and beta == x[0]
        return
    hv = np.concatenate([[1.0], v[1:]])
    H = np.eye(n) - tau * np.outer(hv, hv)
    assert np.allclose(H @ H.T, np.eye(n), atol=1e-14)  # orthogonal <=> tau = 2 / v^T v
    y = H @ x
    assert beta >= 0 and abs(beta - np.linalg.norm(x)) < 1e-14 * np.linalg.norm(x)
    assert np.allclose(y, np.concatenate([[beta], np.zeros(n - 1)]), atol=1e-14)
    assert 0.0 <= tau <= 2.0
    assert v[0] == beta


@pytest.mark.parametrize("shape", [(128, 96), (40, 40), (57, 13), (9, 9), (5, 1), (2, 2), (300, 7)])
def test_bidiag_invariants(shape):
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal(shape)
    if shape == (128, 96):
        A = c1()["A"].astype(np.float64)
    r = oracle.bidiag(A)
    M, N = A.shape
    B = np.zeros((M, N))
    B[np.arange(N), np.arange(N)] = r["d"]
    if N > 1:
        B[np.arange(N - 1), np.arange(1, N)] = r["e"]
    an = np.linalg.norm(A)
    assert np.linalg.norm(r["u"] @ B @ r["vt"] - A) <= 1e-13 * an
    assert np.max(np.abs(r["u"].T @ r["u"] - np.eye(M))) < 1e-13
    assert np.max(np.abs(r["vt"] @ r["vt"].T - np.eye(N))) < 1e-13
    # d_i, e_i >= 0 except where the reflector has length 1 (tau = 0, beta = x_1 signed,
    # upper_bidiagonalization.cpp:130-135,158-163): e_{N-2} always, d_{N-1} when M == N
    dchk = r["d"][:-1] if M == N else r["d"]
    assert np.all(dchk >= 0) and np.all(r["e"][:-1] >= 0)
    assert np.allclose(r["vt"][0], np.eye(N)[0])  # P e_1 = e_1
    s = lapack.dgesdd(A, compute_uv=0)[1]
    sb = np.linalg.svd(B[:N], compute_uv=False)
    assert np.max(np.abs(sb - s)) <= 1e-13 * s[0]
    # uniqueness: the independent Lanczos bidiagonalization gives the same d, e
    alpha, beta = lanczos_bidiag(A)
    assert np.max(np.abs(np.abs(r["d"]) - alpha)) <= 1e-11 * an
    if N > 1:
        assert np.max(np.abs(np.abs(r["e"]) - beta)) <= 1e-11 * an
    # stored reflectors: tau = 2 / (v^T v) for every non-trivial column reflector
    a = r["a"]
    for i in range(N):
        if r["tauq"][i] != 0:
            v = np.concatenate([[1.0], a[i + 1:, i]])
            assert abs(r["tauq"][i] - 2.0 / (v @ v)) < 1e-12


def test_bidiag_c1_golden():
    g = c1()
    r = oracle.bidiag(g["A"].astype(np.float64))
    assert np.max(np.abs(np.abs(r["d"]) - g["alpha"])) < 1e-11
    assert np.max(np.abs(np.abs(r["e"]) - g["beta"])) < 1e-11


def test_leaf_partition_matches_reference_rule():
    # svd_tiling.cpp:83-130 for small N, written out by hand
    assert oracle.leaf_partition(3) == [(0, 2), (2, 3)]
    assert oracle.leaf_partition(4) == [(0, 3), (3, 4)]
    assert oracle.leaf_partition(5) == [(0, 3), (3, 5)]
    assert oracle.leaf_partition(6) == [(0, 3), (3, 5), (5, 6)]
    assert oracle.leaf_partition(7) == [(0, 3), (3, 6), (6, 7)]
    assert oracle.leaf_partition(8) == [(0, 3), (3, 6), (6, 8)]
    assert oracle.leaf_partition(1) == [(0, 1)]
    assert oracle.leaf_partition(2) == [(0, 2)]
    for N in range(3, 200):
        parts = oracle.leaf_partition(N)
        assert len(parts) == N // 3 + 1
        assert parts[0][0] == 0 and parts[-1][1] == N
        assert all(parts[i][1] == parts[i + 1][0] for i in range(len(parts) - 1))
        assert all(1 <= e - s <= 3 for s, e in parts)


LEAF_CASES = [
    (2, 3, [1.0, 2.0], [0.5, -0.7]), (2, 3, [0.0, 0.0], [0.0, 0.0]), (2, 3, [0.0, 2.0], [0.0, 1.5]),
    (2, 3, [1.0, 0.0], [2.0, 0.0]), (2, 3, [0.0, 3.0], [1.0, 0.0]), (2, 3, [1e-8, 1.0], [1.0, 1e-8]),
    (2, 2, [3.0, -1.0], [0.5]), (2, 2, [0.0, 0.0], [0.0]), (2, 2, [0.0, 1.0], [2.0]), (2, 2, [1.0, 1.0], [0.0]),
    (1, 2, [3.0], [4.0]), (1, 2, [0.0], [0.0]), (1, 2, [-2.0], [0.0]), (1, 1, [-5.0], []), (1, 1, [0.0], []),
]


@pytest.mark.parametrize("n,m,dd,ee", LEAF_CASES)
def test_leaf_svd_brute_force(n, m, dd, ee):
    B = np.zeros((n, m))
    for r in range(n):
        B[r, r] = dd[r]
        if r + 1 < m:
            B[r, r + 1] = ee[r]
    sig, U, V = oracle.leaf_svd(n, m, dd, ee)
    assert np.allclose(sig, np.linalg.svd(B, compute_uv=False), atol=1e-14)
    assert np.all(np.diff(sig) <= 0) and np.all(sig >= 0)
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-14)
    assert np.allclose(V.T @ V, np.eye(m), atol=1e-14)
    S = np.zeros((n, m))
    S[np.arange(n), np.arange(n)] = sig
    assert np.allclose(U @ S @ V.T, B, atol=1e-14)


def test_secular_vs_dlasd4_fixture():
    g = np.load(os.path.join(GOLD, "secular.npz"))
    i = 0
    while "d%d" % i in g:
        d, z, sig = g["d%d" % i], g["z%d" % i], g["sig%d" % i]
        for j in range(len(d)):
            s, K, mu, it = oracle.secular_root(d, z, j)
            assert it >= 0
            assert abs(s - sig[j]) <= 4e-15 * sig[j], (i, j)
            # interlacing and the returned (K, mu) representation
            assert d[j] < s and (j == len(d) - 1 or s < d[j + 1])
            assert abs(np.sqrt(d[K] ** 2 + mu) - s) <= 1e-15 * s
        i += 1
    assert i == 5


def test_secular_vs_dlasd4_random():
    rng = np.random.default_rng(11)
    for trial in range(40):
        k = int(rng.integers(2, 50))
        d = np.sort(rng.uniform(0, 1, k))
        d[0] = 0
        z = rng.standard_normal(k)
        rho = z @ z
        for j in range(k):
            delta, sig, work, info = lapack.dlasd4(j, d, z / np.sqrt(rho), rho)
            s, K, mu, it = oracle.secular_root(d, z, j)
            assert abs(s - sig) <= 4e-15 * sig
            diff = (d - d[K]) * (d + d[K]) - mu  # d_j^2 - s^2 without cancellation
            assert np.max(np.abs(diff - delta * work) / np.abs(delta * work)) < 1e-9


def _bidiag_matrix(d, e):
    N = len(d)
    B = np.diag(d)
    if N > 1:
        B[np.arange(N - 1), np.arange(1, N)] = e
    return B


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 13, 31, 64, 97, 200])
def test_bdc_random(N):
    rng = np.random.default_rng(N)
    d = rng.standard_normal(N)
    e = rng.standard_normal(max(N - 1, 0))
    B = _bidiag_matrix(d, e)
    s_ref = lapack.dgesdd(B, compute_uv=0)[1]
    s, qt, wt = oracle.bdc(d, e, defl_eps=oracle.DBL_EPS)
    assert np.max(np.abs(s - s_ref)) <= 1e-13 * s_ref[0]
    assert np.max(np.abs(qt @ qt.T - np.eye(N))) < 1e-12
    assert np.max(np.abs(wt @ wt.T - np.eye(N))) < 1e-12
    assert np.linalg.norm(qt.T @ np.diag(s) @ wt - B) <= 1e-13 * max(s_ref[0], 1e-300) * N


@pytest.mark.parametrize("kind", ["clustered", "graded", "zeros", "rank_deficient"])
def test_bdc_deflation_cases(kind):
    rng = np.random.default_rng(5)
    N = 150
    if kind == "clustered":
        # every leaf is [[1, 0, 0], [0, 0.6, 0.8]] (sigma = 1, 1) and the connecting rows
        # are random, so merged poles coincide with non-zero z: deflation by rotation
        N = 151
        d = np.tile([1.0, 0.6, 0.0], N // 3 + 1)[:N]
        e = np.tile([0.0, 0.8, 0.0], N // 3 + 1)[: N - 1]
        d[2::3] = rng.uniform(0.5, 1.0, len(d[2::3]))
        e[2::3] = rng.uniform(0.5, 1.0, len(e[2::3]))
    elif kind == "graded":
        d = 10.0 ** (-6 * np.arange(N) / (N - 1))
        e = 1e-3 * rng.standard_normal(N - 1)
    elif kind == "zeros":
        d = rng.standard_normal(N)
        e = rng.standard_normal(N - 1)
        d[::7] = 0
        e[::5] = 0
    else:
        d = rng.standard_normal(N)
        e = rng.standard_normal(N - 1)
        d[N // 2:] = 0
        e[N // 2:] = 0
    B = _bidiag_matrix(d, e)
    s_ref = lapack.dgesdd(B, compute_uv=0)[1]
    s, qt, wt, stats = oracle.bdc(d, e, defl_eps=oracle.FLT_EPS, with_stats=True)
    # deflation at tol = 8 eps32 |B| perturbs sigma by at most a small multiple of it
    assert np.max(np.abs(s - s_ref)) <= 64 * oracle.FLT_EPS * s_ref[0]
    assert np.max(np.abs(qt @ qt.T - np.eye(N))) < 1e-12
    assert np.max(np.abs(wt @ wt.T - np.eye(N))) < 1e-12
    assert np.linalg.norm(qt.T @ np.diag(s) @ wt - B) <= 64 * oracle.FLT_EPS * s_ref[0] * N
    ndefl = sum(st[3] for st in stats)
    assert ndefl > 0
    if kind == "clustered":
        assert sum(st[2] for st in stats) > 0  # rotations happened


def test_whole_path_c1_golden():
    g = c1()
    A = g["A"].astype(np.float64)
    for eps, tol in [(oracle.DBL_EPS, 1e-13), (oracle.FLT_EPS, 2e-6)]:
        s, U, Vt = oracle.svd(A, eps)
        m = measures(A, s, U, Vt, s_ref=g["sigma"])
        assert m["sigma_err"] <= tol
        assert m["orth_u"] < 1e-12 and m["orth_v"] < 1e-12
        assert m["resid_ratio"] < (1e-4 if eps == oracle.DBL_EPS else 10.0)


def test_svd_DC_matches_bdc_backtransform():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((70, 45))
    r = oracle.bidiag(A)
    s, U, Vt = oracle.svd_DC(r["u"], r["vt"], r["d"], r["e"], oracle.DBL_EPS)
    m = measures(A, s, U, Vt)
    assert m["sigma_err"] < 1e-13 and m["orth_u"] < 1e-12 and m["orth_v"] < 1e-12
    # columns N..M-1 of U keep Q's columns (svd.cpp:1925-1953 touches only the first N)
    assert np.array_equal(U[:, 45:], r["u"][:, 45:])


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("shape", [(128, 96), (57, 13), (200, 200), (300, 7)])
def test_bidiag_f32_yardstick(shape, order):
    """The FP32 run of the bidiagonalization (the yardstick of FP32 forward error used by
    the GPU parity test), in each evaluation order of its inner products, stays within FP32
    forward error of the FP64 oracle, and its own reflectors are consistent (tau in [0, 2],
    tau (1 + |v_tail|^2) = 2).  The orders differ only by rounding: never bit-identical
    over a whole factorization, always within FP32 forward error of each other."""
    rng = np.random.default_rng(shape[0] + shape[1])
    A = rng.uniform(-1, 1, shape).astype(np.float32)
    M, N = shape
    f = oracle.bidiag_f32(A, order)
    if order and N > 8:
        f0 = oracle.bidiag_f32(A, 0)
        assert not np.array_equal(f["a"], f0["a"])
    o = oracle.bidiag(A.astype(np.float64))
    an = np.linalg.norm(A.astype(np.float64), 2)
    assert np.max(np.abs(f["d"] - o["d"])) <= 1e-4 * an
    assert np.max(np.abs(f["e"] - o["e"])) <= 1e-4 * an
    assert np.max(np.abs(f["tauq"] - o["tauq"])) <= 1e-3
    assert np.all((f["tauq"] >= 0) & (f["tauq"] <= 2)) and np.all((f["taup"] >= 0) & (f["taup"] <= 2))
    for i in range(N):
        v = f["a"][i + 1:, i].astype(np.float64)
        if f["tauq"][i] != 0:
            assert abs(f["tauq"][i] * (1 + v @ v) - 2.0) <= 1e-5 * (1 + M)
        if i < N - 1 and f["taup"][i] != 0:
            w = f["a"][i, i + 2:].astype(np.float64)
            assert abs(f["taup"][i] * (1 + w @ w) - 2.0) <= 1e-5 * (1 + N)
    # and the stored reflectors themselves within FP32 forward error of the FP64 ones
    assert np.max(np.abs(f["a"] - o["a"])) <= 1e-3 * max(1.0, np.max(np.abs(o["a"])))
