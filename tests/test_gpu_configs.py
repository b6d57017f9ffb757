"""The benchmark configurations at full size on the GPU, against FP64 LAPACK.

Every result states sigma error (normwise max|s - s*| / s*_max and the largest per-value
relative error), residual ||A - U S Vt||_F / (||A||_F max(M, N) eps32) and orthogonality
max|U^T U - I|, max|V^T V - I|, with s* = LAPACK dgesdd (FP64, jobz = 'N') of the FP32
input.  Bars: normwise sigma error and orthogonality <= 1e-5, residual ratio <= 10.

  C2  4096 x 4096 N(0, 1), seed 1, the two entry points
  C3  8192 x 4096 graded sigma (1 .. 1e-6), seed 2, FP32 matrix cores and BF16x3 split
  C4  the whole 10,000-matrix batch, seed 3: every info word 0, a stratified sample against
      LAPACK, the hipGraph replay bitwise equal to the eager launch sequence
  C5  120 requests of the seed-5 Poisson stream (M up to 4096, clustered sigma among them)
      through the serving loop; each against LAPACK; a subset bitwise equal alone vs mixed
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import workloads
from svdcheck import measures, measures_gpu

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


def check(m, what):
    say(what, m)
    assert m["sigma_err"] <= SIG_TOL, (what, m)
    assert m["orth_u"] <= ORTH_TOL and m["orth_v"] <= ORTH_TOL, (what, m)
    assert m["resid_ratio"] <= RESID_TOL, (what, m)


def entry_points(torch, A):
    import pacosvd
    M, N = A.shape
    a = A.clone()
    r = pacosvd.upper_bidiagonalization(a)
    d, u, vt = pacosvd.svd_DC(M, N, r["a"], r["u"], r["vt"], r["d"], r["e"], ws=r["ws"])
    torch.cuda.synchronize()
    info = pacosvd.entry_info(r["ws"], M, N)
    return d, u, vt, info


@pytest.mark.timeout(600)
def test_c2_plain_4096(torch_dev):
    torch = torch_dev
    say("C2 4096x4096 N(0,1) seed 1")
    A_np = workloads.c2_matrix()
    with ThreadPoolExecutor(1) as ex:
        ref = ex.submit(workloads.lapack_sigma, A_np)
        A = torch.from_numpy(A_np).cuda()
        d, u, vt, info = entry_points(torch, A)
        assert info == 0
        check(measures_gpu(torch, A, d, u, vt, ref.result()), "C2")


@pytest.mark.timeout(600)
def test_c3_precision_8192x4096(torch_dev):
    torch = torch_dev
    import pacosvd
    say("C3 8192x4096 graded seed 2: generating")
    A_np, _ = workloads.c3_matrix()
    with ThreadPoolExecutor(1) as ex:
        ref = ex.submit(workloads.lapack_sigma, A_np)
        A = torch.from_numpy(A_np).cuda()
        res = {}
        for mode, name in [(0, "fp32_mfma"), (1, "bf16x3_split")]:
            say("C3", name)
            old = pacosvd.set_precision(mode)
            try:
                res[name] = entry_points(torch, A)
            finally:
                pacosvd.set_precision(old)
        s_ref = ref.result()
        for name, (d, u, vt, info) in res.items():
            assert info == 0
            check(measures_gpu(torch, A, d, u, vt, s_ref), "C3 " + name)


@pytest.mark.timeout(600)
def test_c4_whole_batch(torch_dev):
    torch = torch_dev
    import pacosvd
    say("C4 10,000 matrices seed 3: generating")
    host, offs, Ms, Ns = workloads.c4_host()
    dev = torch.from_numpy(host).cuda()
    As = [dev[offs[i]:offs[i + 1]].view(int(Ms[i]), int(Ns[i])) for i in range(len(Ms))]

    def outputs():
        S = torch.zeros(int(Ns.sum()), device="cuda")
        U = torch.zeros(int((Ms.astype(np.int64) ** 2).sum()), device="cuda")
        V = torch.zeros(int((Ns.astype(np.int64) ** 2).sum()), device="cuda")
        outs, s, u, v = [], 0, 0, 0
        for M, N in zip(Ms.tolist(), Ns.tolist()):
            outs.append((S[s:s + N], U[u:u + M * M].view(M, M), V[v:v + N * N].view(N, N)))
            s, u, v = s + N, u + M * M, v + N * N
        return (S, U, V), outs

    flat_e, outs_e = outputs()
    flat_g, outs_g = outputs()
    say("C4 eager")
    pacosvd.svd_batched(As, outs=outs_e)
    torch.cuda.synchronize()
    info = pacosvd.last_info(len(As))
    assert len(info) == len(As) and not any(info), [i for i, x in enumerate(info) if x][:10]
    say("C4 graph")
    pacosvd.svd_batched(As, outs=outs_g, graph=True)
    torch.cuda.synchronize()
    info = pacosvd.last_info(len(As))
    assert not any(info)
    for x, y in zip(flat_e, flat_g):
        assert torch.equal(x, y)
    # stratified sample: 48 matrices spread over the size order, plus the largest
    order = np.argsort(Ms.astype(np.int64) * Ns)
    pick = sorted(set(order[np.linspace(0, len(order) - 1, 48).astype(int)].tolist()))
    worst = {}
    for i in pick:
        A = host[offs[i]:offs[i + 1]].reshape(int(Ms[i]), int(Ns[i]))
        m = measures(A, *(t.cpu().numpy() for t in outs_e[i]))
        for k, v in m.items():
            worst[k] = max(worst.get(k, 0.0), v)
    check(worst, "C4 sample of %d" % len(pick))


@pytest.mark.timeout(900)
def test_c5_serving_stream(torch_dev):
    torch = torch_dev
    import time
    import pacosvd
    count = 120
    reqs = workloads.c5_requests(count)
    say("C5 %d requests: max M %d, clustered %d, wide %d" % (
        count, max(r["M"] for r in reqs), sum(r["clustered"] for r in reqs),
        sum(1 for r in reqs if r["N"] > 512 or r["M"] * r["N"] >= 1 << 20)))
    As = workloads.c5_matrices(torch, reqs)
    torch.cuda.synchronize()
    host = [A.cpu().numpy() for A in As]
    with ThreadPoolExecutor(2) as ex:
        refs = [ex.submit(workloads.lapack_sigma, h) for h in host]
        outs = pacosvd.alloc_outputs(As)
        srv = pacosvd.Server(64 << 30, max_batch=512, batch_lanes=2, wide_lanes=4, max_M=4096, max_N=4096)
        try:
            t0 = time.perf_counter()
            for i, r in enumerate(reqs):
                while time.perf_counter() - t0 < r["t"]:
                    pass
                assert srv.submit(As[i], *outs[i], i) == 0
            srv.drain()
            fin = srv.poll_info(4096)
            st = srv.stats()
        finally:
            srv.close()
        say("C5 served", st)
        done = {t: ms for t, ms, _ in fin}
        assert sorted(done) == list(range(count)) and all(ms >= 0 for ms in done.values())
        # every request's info word: no wait timed out, no NaN, every secular root converged
        assert all(info == 0 for _, _, info in fin) and st["nonzero_info"] == 0, [f for f in fin if f[2]]
        worst = {}
        for i in range(count):
            m = measures_gpu(torch, As[i], *outs[i], refs[i].result())
            if reqs[i]["clustered"] or reqs[i]["M"] >= 2048:
                say("C5 request", i, (reqs[i]["M"], reqs[i]["N"]), "clustered" if reqs[i]["clustered"] else "", m)
            for k, v in m.items():
                worst[k] = max(worst.get(k, 0.0), v)
        check(worst, "C5 worst over %d requests" % count)
    # alone: a fresh server of the same configuration (its wide lanes use the same tile
    # grid), one request at a time: bitwise the same as in the mix
    sub = [i for i in range(count) if reqs[i]["clustered"]][:3]
    sub += [i for i in range(count) if reqs[i]["M"] >= 2048 and i not in sub][:3]
    sub += [i for i in range(count) if i not in sub][:4]
    srv = pacosvd.Server(64 << 30, max_batch=512, batch_lanes=2, wide_lanes=4, max_M=4096, max_N=4096)
    try:
        for i in sub:
            alone = pacosvd.alloc_outputs([As[i]])[0]
            assert srv.submit(As[i], *alone, 1000 + i) == 0
            srv.drain()
            for x, y in zip(alone, outs[i]):
                assert torch.equal(x, y), (i, reqs[i]["M"], reqs[i]["N"])
    finally:
        srv.close()
