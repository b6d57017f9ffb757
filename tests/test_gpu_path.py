"""GPU parity of the HIP path (through the C ABI) against the oracle and FP64 LAPACK.

Bars (stated per result): sigma normwise error max|s - s*| / s*_max <= 1e-5, entry-wise
orthogonality max|U^T U - I|, max|V^T V - I| <= 1e-5, residual ratio
||A - U S Vt||_F / (||A||_F max(M, N) eps32) <= 10.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from svdcheck import measures

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SIG_TOL = 1e-5
ORTH_TOL = 1e-5
RESID_TOL = 10.0
BIDIAG_TOL = 5e-4


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    import pacosvd
    pacosvd.lib()
    oracle.build()
    return torch


def np_of(t):
    return t.detach().cpu().numpy()


def assert_entry_a_parity(r, A):
    """Every output of entry point (a) element-wise against the FP64 oracle: d, e, tauq,
    taup and the whole of a (d / e on the band, the column reflectors below it, the row
    reflectors right of it; upper_bidiagonalization.cpp:204,271-282).  Golub-Kahan
    coefficients are forward-sensitive, so the bar is derived from the same algorithm run in
    FP32 on the CPU: max|gpu - fp64| <= 4 yard + 16 eps32 scale, per output, where yard is
    the largest max|fp32 - fp64| over three evaluation orders of the FP32 run's inner
    products (scale: ||A||_2 for d, e; 1 for tau in [0, 2]; max|a| for a).  One order alone
    is a thin sample: on a 1100 x 1000 N(0,1) matrix the FP32 error of d is 0.0098, 0.021
    and 0.034 in the three orders, and 0.045 to 0.105 on row permutations of the matrix
    (which leave d, e unchanged in exact arithmetic)."""
    shape = A.shape
    g = {k: np_of(v).astype(np.float64) for k, v in r.items() if k in ("a", "d", "e", "tauq", "taup")}
    o = oracle.bidiag(A.astype(np.float64))
    fs = [oracle.bidiag_f32(A, order) for order in (0, 1, 2)]
    an = np.linalg.norm(A.astype(np.float64), 2)
    eps = float(np.finfo(np.float32).eps)
    for k, scale in (("d", an), ("e", an), ("tauq", 1.0), ("taup", 1.0), ("a", max(1.0, np.max(np.abs(o["a"]))))):
        if g[k].size == 0 or (k in ("e", "taup") and shape[1] < 2):
            continue
        err = float(np.max(np.abs(g[k] - o[k])))
        yard = max(float(np.max(np.abs(f[k].astype(np.float64) - o[k]))) for f in fs)
        bar = 4.0 * yard + 16.0 * eps * scale
        say(k, "gpu err %.3e fp32 err %.3e bar %.3e" % (err, yard, bar))
        assert err <= bar, (k, err, yard)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("shape", [(64, 64, 64, False, False), (100, 37, 73, False, True), (130, 65, 129, True, False),
                                   (5, 200, 17, True, True), (257, 96, 300, False, False)])
def test_gemm_tile(torch_dev, mode, shape):
    torch = torch_dev
    import pacosvd
    m, n, k, ta, tb = shape
    say("gemm", shape, "mode", mode)
    g = torch.Generator().manual_seed(m * 7 + n)
    A = torch.randn((k, m) if ta else (m, k), generator=g)
    B = torch.randn((n, k) if tb else (k, n), generator=g)
    C0 = torch.randn((m, n), generator=g)
    old = pacosvd.set_precision(mode)
    try:
        C = pacosvd.gemm(A.cuda(), B.cuda(), C0.clone().cuda(), transA=ta, transB=tb, alpha=0.5, beta=-1.0)
        torch.cuda.synchronize()
    finally:
        pacosvd.set_precision(old)
    Ad = (A.T if ta else A).double()
    Bd = (B.T if tb else B).double()
    ref = 0.5 * (Ad @ Bd) - C0.double()
    err = (C.cpu().double() - ref).abs().max().item()
    scale = (Ad.abs() @ Bd.abs()).max().item()
    assert err <= (1e-6 if mode != 1 else 4e-6) * scale, err


def run_entry_points(torch, A_np):
    import pacosvd
    M, N = A_np.shape
    a = torch.from_numpy(A_np.copy()).cuda()
    r = pacosvd.upper_bidiagonalization(a)
    torch.cuda.synchronize()
    bid = {k: np_of(v) for k, v in r.items() if k != "ws"}
    d, u, vt = pacosvd.svd_DC(M, N, r["a"], r["u"], r["vt"], r["d"], r["e"], ws=r["ws"])
    torch.cuda.synchronize()
    return bid, np_of(d), np_of(u), np_of(vt)


def test_c1_entry_points(torch_dev):
    torch = torch_dev
    say("c1 128x96 entry points")
    g = np.load(os.path.join(GOLD, "c1.npz"))
    A = g["A"]
    bid, s, U, Vt = run_entry_points(torch, A)
    orc = oracle.bidiag(A.astype(np.float64))
    smax = g["sigma"][0]
    # entry point (a): d, e are unique (non-negative convention) but forward-sensitive: an
    # FP32 run of the same algorithm on the CPU deviates from FP64 by up to ~1e-4 sigma_max
    # at 512^2, so the element-wise bar is BIDIAG_TOL; the backward error below is tight.
    assert np.max(np.abs(bid["d"] - orc["d"])) <= BIDIAG_TOL * smax
    assert np.max(np.abs(bid["e"] - orc["e"])) <= BIDIAG_TOL * smax
    assert np.max(np.abs(bid["tauq"] - orc["tauq"])) <= 1e-3
    assert np.max(np.abs(bid["u"].T @ bid["u"] - np.eye(128))) <= ORTH_TOL
    assert np.max(np.abs(bid["vt"] @ bid["vt"].T - np.eye(96))) <= ORTH_TOL
    M, N = A.shape
    B = np.zeros((M, N))
    B[np.arange(N), np.arange(N)] = bid["d"]
    B[np.arange(N - 1), np.arange(1, N)] = bid["e"]
    assert np.linalg.norm(bid["u"] @ B @ bid["vt"] - A) <= 1e-5 * np.linalg.norm(A)
    # entry point (b): whole SVD vs golden FP64 dgesdd and the oracle's sigma
    m = measures(A, s, U, Vt, s_ref=g["sigma"])
    say("c1 measures", m)
    assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL
    assert m["resid_ratio"] <= RESID_TOL
    s_orc, _, _ = oracle.svd(A.astype(np.float64))
    assert np.max(np.abs(s - s_orc)) <= SIG_TOL * smax
    assert np.all(np.diff(s) <= 0)


@pytest.mark.parametrize("shape", [(1, 1), (2, 1), (2, 2), (3, 3), (4, 3), (5, 4), (6, 6), (7, 5), (9, 8), (33, 17),
                                   (64, 64), (100, 1), (150, 149), (300, 200), (512, 512)])
def test_entry_points_shapes(torch_dev, shape):
    torch = torch_dev
    say("entry points", shape)
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    A = rng.standard_normal(shape).astype(np.float32)
    _, s, U, Vt = run_entry_points(torch, A)
    m = measures(A, s, U, Vt)
    say("measures", m)
    assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL
    assert m["resid_ratio"] <= RESID_TOL


@pytest.mark.parametrize("path", ["resident", "panel"])
@pytest.mark.parametrize("shape", [(1100, 1000), (2048, 700), (1200, 1023)])
def test_grid_bidiag_vs_oracle(torch_dev, shape, path):
    """Large matrices (M*N >= 2^20) take the chip-wide bidiagonalization: the matrix in the
    registers of the chip (default) or the panel kernels (PACO_GRID_RES=0)."""
    torch = torch_dev
    import pacosvd
    say("grid bidiag", shape, path)
    rng = np.random.default_rng(shape[1])
    A = rng.standard_normal(shape).astype(np.float32)
    a = torch.from_numpy(A.copy()).cuda()
    if path == "panel":
        os.environ["PACO_GRID_RES"] = "0"
    try:
        r = pacosvd.upper_bidiagonalization(a)
        torch.cuda.synchronize()
    finally:
        os.environ.pop("PACO_GRID_RES", None)
    assert pacosvd.entry_info(r["ws"], *shape) == 0
    assert_entry_a_parity(r, A)
    d, e = np_of(r["d"]), np_of(r["e"])
    u, vt = np_of(r["u"]), np_of(r["vt"])
    M, N = shape
    assert np.max(np.abs(u.T @ u - np.eye(M))) <= ORTH_TOL
    assert np.max(np.abs(vt @ vt.T - np.eye(N))) <= ORTH_TOL
    B = np.zeros((M, N))
    B[np.arange(N), np.arange(N)] = d
    B[np.arange(N - 1), np.arange(1, N)] = e
    assert np.linalg.norm(u @ B @ vt - A) <= 1e-5 * np.linalg.norm(A)


@pytest.mark.parametrize("shape", [(1100, 1000), (1200, 1023)])
def test_grid_bidiag_rank_deficient(torch_dev, shape):
    """Repeated columns leave a trailing block of rounding noise (~1e-26) whose squares
    underflow FP32: the reflectors must stay orthogonal (tau v.v = 2)."""
    torch = torch_dev
    import pacosvd
    say("grid bidiag rank-deficient", shape)
    rng = np.random.default_rng(shape[1] + 7)
    A = rng.standard_normal(shape).astype(np.float32)
    A[:, 1::2] = A[:, 0:-1:2]
    r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
    torch.cuda.synchronize()
    u, vt, d, e = np_of(r["u"]), np_of(r["vt"]), np_of(r["d"]), np_of(r["e"])
    M, N = shape
    assert np.max(np.abs(u.T @ u - np.eye(M))) <= ORTH_TOL
    assert np.max(np.abs(vt @ vt.T - np.eye(N))) <= ORTH_TOL
    B = np.zeros((M, N))
    B[np.arange(N), np.arange(N)] = d
    B[np.arange(N - 1), np.arange(1, N)] = e
    assert np.linalg.norm(u @ B @ vt - A) <= 1e-5 * np.linalg.norm(A)


@pytest.mark.parametrize("kind", ["zero", "rank1", "repeated", "clustered", "graded"])
def test_entry_points_special(torch_dev, kind):
    torch = torch_dev
    say("special", kind)
    rng = np.random.default_rng(1)
    M, N = 120, 90
    if kind == "zero":
        A = np.zeros((M, N), np.float32)
    elif kind == "rank1":
        A = np.outer(rng.standard_normal(M), rng.standard_normal(N)).astype(np.float32)
    elif kind == "repeated":  # rank N/2, exact duplicate columns
        A = rng.standard_normal((M, N)).astype(np.float32)
        A[:, 1::2] = A[:, 0:-1:2]
    else:
        if kind == "clustered":
            sig = np.concatenate([1.0 + 1e-7 * rng.uniform(size=60), 10.0 ** rng.uniform(-3, 0, N - 60)])
        else:
            sig = 10.0 ** (-6 * np.arange(N) / (N - 1))
        Q1, _ = np.linalg.qr(rng.standard_normal((M, N)))
        Q2, _ = np.linalg.qr(rng.standard_normal((N, N)))
        A = ((Q1 * sig) @ Q2.T).astype(np.float32)
    _, s, U, Vt = run_entry_points(torch, A)
    m = measures(A, s, U, Vt)
    say("measures", m)
    assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL
    if kind != "zero":
        assert m["resid_ratio"] <= RESID_TOL


def mixed_batch(rng, count, mmax=200):
    shapes = []
    for _ in range(count):
        M = int(rng.integers(16, mmax + 1))
        N = int(rng.integers(8, M + 1))
        shapes.append((M, N))
    return [rng.standard_normal(sh).astype(np.float32) for sh in shapes]


def test_batched_matches_plain_and_alone(torch_dev):
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(3)
    mats = mixed_batch(rng, 24)
    say("batched", [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    outs_b = pacosvd.svd_batched(As)
    outs_p = pacosvd.svd_plain(As)
    torch.cuda.synchronize()
    assert pacosvd.last_info(len(As)) == [0] * len(As)
    for i, A in enumerate(mats):
        sb, Ub, Vb = (np_of(x) for x in outs_b[i])
        sp = np_of(outs_p[i][0])
        mb = measures(A, sb, Ub, Vb)
        assert mb["sigma_err"] <= SIG_TOL and mb["orth_u"] <= ORTH_TOL and mb["orth_v"] <= ORTH_TOL, (i, mb)
        assert mb["resid_ratio"] <= RESID_TOL
        assert np.max(np.abs(sb - sp)) <= SIG_TOL * max(sb[0], 1e-30)
    # alone vs in the mix: a matrix's schedule depends only on its shape -> identical bits
    for i in (0, 5, 11):
        alone = pacosvd.svd_batched([As[i]])
        torch.cuda.synchronize()
        for x, y in zip(alone[0], outs_b[i]):
            assert torch.equal(x, y), i


def test_batched_dc_beside_qp_same_bits(torch_dev):
    """The batched D&C runs on a forked stream beside the formation of Q and P^T (it keeps
    its left secular vectors in a buffer of its own, so it never writes a, which that
    formation reads).  The same kernels run in the same order on each matrix whatever the
    schedule: every output must equal the serial schedule's (PACO_DC_SERIAL=1) and
    PACO_QP_FORK=0's (P's rounds after Q's) bit for bit, with accurate results."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(41)
    mats = mixed_batch(rng, 32, mmax=512)
    say("dc beside qp", [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    forked = pacosvd.svd_batched(As)
    torch.cuda.synchronize()
    assert pacosvd.last_info(len(As)) == [0] * len(As)
    others = []
    # the serial schedule; and the D&C forked with P's rounds after Q's on the caller's
    # stream
    for env in ("PACO_DC_SERIAL", "PACO_QP_FORK"):
        os.environ[env] = "1" if env == "PACO_DC_SERIAL" else "0"
        try:
            others.append(pacosvd.svd_batched(As))
            torch.cuda.synchronize()
        finally:
            del os.environ[env]
    for i, A in enumerate(mats):
        for other in others:
            for x, y in zip(forked[i], other[i]):
                assert torch.equal(x, y), i
        m = measures(A, *(np_of(x) for x in forked[i]))
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (i, m)
        assert m["resid_ratio"] <= RESID_TOL


@pytest.mark.parametrize("variant", ["strip", "rounds", "bf16x3", "fma_strips"])
def test_batched_qp_formation_paths(torch_dev, variant):
    """Q and P^T of a batch by every formation path (one strip launch per round on the
    matrix cores, the extract + two tile GEMMs per round, the BF16x3 GEMMs, the 64-column
    FMA strips): full SVD accuracy; where only the formation differs (same
    bidiagonalization and D&C arithmetic: not BF16x3, which also changes the D&C
    products), U and V^T within 2e-5 of the strip path."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(17)
    mats = mixed_batch(rng, 20, mmax=330)
    mats += [rng.standard_normal(sh).astype(np.float32)
             for sh in [(512, 512), (300, 33), (97, 96), (3, 1), (600, 200), (513, 64)]]
    say("qp formation", variant, [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    ref = pacosvd.svd_batched(As)
    torch.cuda.synchronize()
    env = {"rounds": {"PACO_QP_STRIP": "0"}, "fma_strips": {"PACO_QP_BATCH": "0"}}.get(variant, {})
    os.environ.update(env)
    old = pacosvd.set_precision(1) if variant == "bf16x3" else None
    try:
        out = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
    finally:
        for k in env:
            del os.environ[k]
        if old is not None:
            pacosvd.set_precision(old)
    assert pacosvd.last_info(len(As)) == [0] * len(As)
    for i, A in enumerate(mats):
        s, U, Vt = (np_of(x) for x in out[i])
        m = measures(A, s, U, Vt)
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (i, A.shape, m)
        assert m["resid_ratio"] <= RESID_TOL, (i, A.shape, m)
        if variant == "bf16x3":
            continue
        N = A.shape[1]
        U0, V0 = np_of(ref[i][1]), np_of(ref[i][2])
        assert np.max(np.abs(U[:, :N] - U0[:, :N])) <= 2e-5, (i, A.shape)
        assert np.max(np.abs(Vt - V0)) <= 2e-5, (i, A.shape)


def test_batched_graph_replays(torch_dev):
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(9)
    mats = mixed_batch(rng, 12, mmax=120)
    say("graph", [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    eager = pacosvd.svd_batched(As)
    torch.cuda.synchronize()
    outs = pacosvd.alloc_outputs(As)
    for rep in range(3):
        for o in outs:
            for t in o:
                t.zero_()
        pacosvd.svd_batched(As, outs=outs, graph=True)
        torch.cuda.synchronize()
        _ = (torch.randn(256, 256, device="cuda") @ torch.randn(256, 256, device="cuda")).sum().item()
        for i in range(len(As)):
            for x, y in zip(outs[i], eager[i]):
                assert torch.equal(x, y), (rep, i)


@pytest.mark.parametrize("threads", ["256", "512", "1024"])
def test_bidiag_variants_bitwise_equal(torch_dev, threads):
    """The LDS-resident and the streamed one-workgroup bidiagonalization run the same
    arithmetic in the same order at one workgroup size: their outputs must be identical.
    (The blocked kernel, another order of operations, is kept off for the streamed run.)"""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(21)
    mats = mixed_batch(rng, 16, mmax=180)
    say("variants", threads, [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    runs = []
    for env in ({"PACO_RES_THREADS": threads},
                {"PACO_LDS_RESIDENT_MAX": "0", "PACO_BIDIAG_THREADS": threads, "PACO_BIDIAG_BLOCKED": "0"}):
        os.environ.update(env)
        try:
            runs.append(pacosvd.svd_batched(As))
            torch.cuda.synchronize()
        finally:
            for k in env:
                del os.environ[k]
    for i in range(len(As)):
        for x, y in zip(runs[0][i], runs[1][i]):
            assert torch.equal(x, y), i


def test_bidiag_one_wavefront(torch_dev):
    """LDS-resident matrices of a batch on the one-wavefront kernel (forced for every LDS
    class): full SVD accuracy against LAPACK, agreement with the one-workgroup kernel,
    edge shapes included."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(13)
    shapes = [(16, 8), (17, 9), (64, 64), (65, 65), (128, 96), (129, 128), (200, 150), (300, 100), (40, 1), (2, 1),
              (3, 2), (511, 70), (180, 180), (256, 150)]
    mats = [rng.standard_normal(sh).astype(np.float32) for sh in shapes]
    mats[5][:, 1::2] = mats[5][:, 0:-1:2]  # rank-deficient: repeated columns
    say("one wavefront", shapes)
    As = [torch.from_numpy(m).cuda() for m in mats]
    os.environ["PACO_WAVE_MAX_CLASS"] = "3"
    try:
        wave = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
    finally:
        del os.environ["PACO_WAVE_MAX_CLASS"]
    os.environ["PACO_BIDIAG_WAVE"] = "0"
    try:
        wg = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
    finally:
        del os.environ["PACO_BIDIAG_WAVE"]
    assert pacosvd.last_info(len(As)) == [0] * len(As)
    for i, A in enumerate(mats):
        s, U, Vt = (np_of(x) for x in wave[i])
        m = measures(A, s, U, Vt)
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (shapes[i], m)
        assert m["resid_ratio"] <= RESID_TOL, (shapes[i], m)
        s2 = np_of(wg[i][0])
        assert np.max(np.abs(s - s2)) <= SIG_TOL * max(float(s2[0]), 1e-30), shapes[i]


@pytest.mark.parametrize("threads", ["512", "1024", "blocked"])
def test_bidiag_tiers_bitwise_equal(torch_dev, threads):
    """Rows streamed from global memory or cached in LDS take the same arithmetic in the
    same order: every split of a matrix between the two gives the same bits, including
    the re-packing of the cached rows as the trailing matrix narrows (all of the LDS and
    64 KiB re-pack the larger shapes here; 0 bytes caches no row).  "blocked": the panel
    kernel (1024 threads, PACO_BIDIAG_BLOCKED=1), whose cached rows are not re-packed."""
    blocked = threads == "blocked"
    threads = "1024" if blocked else threads
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(5)
    shapes = [(512, 500), (300, 250), (480, 130), (260, 259), (200, 64), (100, 70), (511, 300), (64, 33),
              (450, 420), (390, 380), (170, 129), (40, 9)]
    mats = [rng.standard_normal(sh).astype(np.float32) for sh in shapes]
    say("tiers", threads, shapes)
    As = [torch.from_numpy(m).cuda() for m in mats]
    # LDS bytes of the streamed launch (None: all of it)
    variants = [None, "0", "65536"]
    results = []
    for lds in variants:
        os.environ["PACO_BIDIAG_THREADS"] = threads
        os.environ["PACO_LDS_RESIDENT_MAX"] = "0"
        if blocked:
            os.environ["PACO_BIDIAG_BLOCKED"] = "1"
        if lds is not None:
            os.environ["PACO_BIDIAG_LDS"] = lds
        try:
            results.append(pacosvd.svd_batched(As))
            torch.cuda.synchronize()
        finally:
            for k in ("PACO_BIDIAG_THREADS", "PACO_LDS_RESIDENT_MAX", "PACO_BIDIAG_LDS", "PACO_BIDIAG_BLOCKED"):
                os.environ.pop(k, None)
    for i, A in enumerate(mats):
        m = measures(A, *(np_of(x) for x in results[0][i]))
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (i, m)
        for v in range(1, len(variants)):
            for x, y in zip(results[0][i], results[v][i]):
                assert torch.equal(x, y), (i, variants[v])


@pytest.mark.parametrize("threads", ["256", "512", "1024"])
def test_bidiag_workgroup_sizes(torch_dev, threads):
    """Matrices too large for LDS run with 256, 512 or 1024 threads (rows partly cached)."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(int(threads))
    mats = [rng.standard_normal(sh).astype(np.float32) for sh in [(300, 250), (512, 512), (480, 130), (260, 259)]]
    say("threads", threads, [m.shape for m in mats])
    As = [torch.from_numpy(m).cuda() for m in mats]
    os.environ["PACO_BIDIAG_THREADS"] = threads
    try:
        outs = pacosvd.svd_batched(As)
        torch.cuda.synchronize()
    finally:
        del os.environ["PACO_BIDIAG_THREADS"]
    for i, A in enumerate(mats):
        m = measures(A, *(np_of(x) for x in outs[i]))
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (i, m)
        assert m["resid_ratio"] <= RESID_TOL


def test_server_alone_vs_mix_and_admission(torch_dev):
    """Continuous batching: every request's result equals the one computed alone."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(17)
    shapes = [(40, 30), (200, 150), (1100, 1000), (64, 64), (300, 120), (17, 9), (600, 550), (128, 96)]
    mats = [rng.standard_normal(sh).astype(np.float32) for sh in shapes]
    say("server", shapes)
    As = [torch.from_numpy(m).cuda() for m in mats]
    outs = pacosvd.alloc_outputs(As)
    srv = pacosvd.Server(2 << 30, max_batch=4, batch_lanes=2, wide_lanes=1, max_M=1100, max_N=1100)
    try:
        for i, (A, (S, U, Vt)) in enumerate(zip(As, outs)):
            assert srv.submit(A, S, U, Vt, i) == 0
        srv.drain()
        done = srv.poll()
        assert sorted(t for t, _ in done) == list(range(len(As)))
        st = srv.stats()
        say("stats", st)
        assert st["completed"] == len(As) and st["wide_requests"] == 2
        # a request that can never be admitted is refused at submission
        big = torch.zeros((2000, 1500), device="cuda")
        o = pacosvd.alloc_outputs([big])[0]
        assert srv.submit(big, *o, 99) == -1
    finally:
        srv.close()
    for i, A in enumerate(mats):
        m = measures(A, *(np_of(x) for x in outs[i]))
        assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (i, m)
    # alone: a fresh server, one request at a time
    srv = pacosvd.Server(2 << 30, max_batch=4, batch_lanes=1, wide_lanes=1, max_M=1100, max_N=1100)
    try:
        for i in (0, 2, 4, 6):
            alone = pacosvd.alloc_outputs([As[i]])[0]
            assert srv.submit(As[i], *alone, 100 + i) == 0
            srv.drain()
            for x, y in zip(alone, outs[i]):
                assert torch.equal(x, y), i
    finally:
        srv.close()


@pytest.mark.parametrize("shape", [(1100, 1000), (2048, 700), (1500, 1500)])
def test_grid_persistent_matches_launch_per_phase(torch_dev, shape):
    """The persistent panel launch (grid barriers) and the launch-per-phase schedule run the
    same tile code with the same reduction order: entry point (a) gives bitwise equal
    outputs; a smaller switch to the one-workgroup finish stays within the parity bars."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    A = rng.standard_normal(shape).astype(np.float32)
    os.environ["PACO_GRID_RES"] = "0"  # the panel kernels throughout
    try:
        grid_panel_schedules(torch, pacosvd, shape, A)
    finally:
        del os.environ["PACO_GRID_RES"]


def grid_panel_schedules(torch, pacosvd, shape, A):
    outs = {}
    for persist in ("1", "0"):
        say("grid", shape, "persistent", persist)
        os.environ["PACO_GRID_PERSIST"] = persist
        try:
            r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
            torch.cuda.synchronize()
        finally:
            del os.environ["PACO_GRID_PERSIST"]
        assert pacosvd.entry_info(r["ws"], *shape) == 0
        outs[persist] = r
    for k in ("a", "u", "vt", "d", "e", "tauq", "taup"):
        assert torch.equal(outs["1"][k], outs["0"][k]), k
    # several tiles per workgroup (fewer participants than tiles), with a workspace full of
    # NaN bytes: only entries the panel defined may be read back
    say("grid", shape, "7 participants, NaN-filled workspace")
    ws = pacosvd.workspace(*shape)
    ws.fill_(255)
    os.environ["PACO_GRID_PARTS"] = "7"
    try:
        r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda(), ws=ws)
        torch.cuda.synchronize()
    finally:
        del os.environ["PACO_GRID_PARTS"]
    for k in ("a", "u", "vt", "d", "e", "tauq", "taup"):
        assert torch.equal(outs["1"][k], r[k]), k
    say("grid", shape, "switch 64")
    os.environ["PACO_GRID_SWITCH"] = "64"
    try:
        _, s, U, Vt = run_entry_points(torch, A)
    finally:
        del os.environ["PACO_GRID_SWITCH"]
    m = measures(A, s, U, Vt)
    say("measures", m)
    assert m["sigma_err"] <= SIG_TOL and m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL
    assert m["resid_ratio"] <= RESID_TOL


@pytest.mark.parametrize("blocked", [False, True])
@pytest.mark.parametrize("shape", [(128, 96), (300, 200), (512, 512), (90, 1), (1100, 1000), (2048, 700),
                                   (1024, 450), (700, 512), (257, 240), (520, 17)])
def test_entry_a_outputs_vs_oracle(torch_dev, shape, blocked):
    """Every output of entry point (a) element-wise against the FP64 oracle: d, e, tauq,
    taup and the whole of a (d / e on the band, the column reflectors below it, the row
    reflectors right of it; upper_bidiagonalization.cpp:204,271-282).  Golub-Kahan
    coefficients are forward-sensitive, so the bar is derived from the same algorithm run in
    FP32 on the CPU: max|gpu - fp64| <= 4 max|fp32 - fp64| + 16 eps32 scale, per output
    (scale: ||A||_2 for d, e; 1 for tau in [0, 2]; max|a| for a)."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(shape[0] * 3 + shape[1])
    A = (rng.uniform(-1, 1, shape) if shape == (128, 96) else rng.standard_normal(shape)).astype(np.float32)
    if shape == (128, 96):
        A = np.load(os.path.join(GOLD, "c1.npz"))["A"]
    say("entry (a) outputs", shape, "blocked" if blocked else "")
    if blocked:  # the panel kernel for the shapes it takes (streamed, N <= 512, M <= 1024)
        os.environ["PACO_BIDIAG_BLOCKED"] = "1"
    try:
        r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
        torch.cuda.synchronize()
    finally:
        os.environ.pop("PACO_BIDIAG_BLOCKED", None)
    assert_entry_a_parity(r, A)


@pytest.mark.parametrize("shape,pmax", [((1100, 1000), 0), ((1100, 1000), 16), ((2048, 700), 0), ((1500, 1500), 0),
                                        ((1500, 1500), 40), ((4100, 300), 0)])
def test_resident_bidiag_vs_oracle(torch_dev, shape, pmax):
    """The bidiagonalization held in the registers of P workgroups (bidiag_res.hip), every
    output of entry point (a) element-wise against the FP64 oracle with the bar of
    test_entry_a_outputs_vs_oracle.  pmax caps the workgroups (PACO_RES_PMAX): with 16
    (1100 x 1000) and 40 (1500 x 1500) the matrix does not fit at first, so panels run until
    the trailing matrix does.  Two runs give the same bits."""
    torch = torch_dev
    import pacosvd
    rng = np.random.default_rng(shape[0] + shape[1] + pmax)
    A = rng.standard_normal(shape).astype(np.float32)
    say("resident bidiag", shape, "pmax", pmax)
    if pmax:
        os.environ["PACO_RES_PMAX"] = str(pmax)
    try:
        runs = []
        for _ in range(2):
            r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
            torch.cuda.synchronize()
            assert pacosvd.entry_info(r["ws"], *shape) == 0
            runs.append(r)
    finally:
        os.environ.pop("PACO_RES_PMAX", None)
    for k in ("a", "d", "e", "tauq", "taup"):
        assert torch.equal(runs[0][k], runs[1][k]), k
    assert_entry_a_parity(runs[0], A)


def test_resident_bidiag_timeout_reports_info(torch_dev):
    """A wait of the resident kernel that exceeds its bound stops the call with info -2 (the
    bound forced to zero from the host: every wait that is not satisfied at once times out;
    every workgroup still exits).  The next call on the same workspace, with the normal
    bound, gives the same bits as a call on a fresh workspace."""
    torch = torch_dev
    import pacosvd
    shape = (1100, 1000)
    A = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    say("resident timeout", shape)
    ref = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
    torch.cuda.synchronize()
    assert pacosvd.entry_info(ref["ws"], *shape) == 0
    os.environ["PACO_RES_TIMEOUT_US"] = "0"
    try:
        r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda())
        torch.cuda.synchronize()
    finally:
        del os.environ["PACO_RES_TIMEOUT_US"]
    assert pacosvd.entry_info(r["ws"], *shape) == -2
    r = pacosvd.upper_bidiagonalization(torch.from_numpy(A.copy()).cuda(), ws=r["ws"])
    torch.cuda.synchronize()
    assert pacosvd.entry_info(r["ws"], *shape) == 0
    for k in ("a", "u", "vt", "d", "e", "tauq", "taup"):
        assert torch.equal(r[k], ref[k]), k
