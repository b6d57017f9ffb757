"""Benchmark: batched SVD throughput on one MI355X vs the plain path and the CPU.

Workload (one "step" = one pass of the batched path over the whole batch):
  10,000 FP32 matrices, M ~ U[16, 512], N ~ U[8, M], entries N(0, 1), numpy
  default_rng(3) (all M drawn first, then all N, then the matrices in order).
Inputs and outputs stay resident in HBM for the timed region.

Printed JSON line (rank 0): value = batched SVDs/s; extra fields: plain-path SVDs/s on
the same stream (the >= 10x target's denominator), per-phase milliseconds from hipEvents
on the launch stream, the bidiagonalization roofline, accuracy vs FP64 LAPACK on a
sample, and the CPU baseline (the oracle, one core, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "paco-svd_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFS = 157.3


def workload_shapes(count=10000, seed=3):
    rng = np.random.default_rng(seed)
    Ms = rng.integers(16, 513, size=count)
    Ns = np.array([rng.integers(8, m + 1) for m in Ms])
    return rng, Ms, Ns


def make_workload(torch, count=10000, seed=3):
    rng, Ms, Ns = workload_shapes(count, seed)
    sizes = Ms * Ns
    offs = np.concatenate([[0], np.cumsum(sizes)])
    host = np.empty(int(offs[-1]), dtype=np.float32)
    for i in range(count):
        host[offs[i]:offs[i + 1]] = rng.standard_normal((int(Ms[i]), int(Ns[i]))).astype(np.float32).ravel()
    dev = torch.from_numpy(host).cuda()
    As = [dev[offs[i]:offs[i + 1]].view(int(Ms[i]), int(Ns[i])) for i in range(count)]
    return host, offs, Ms, Ns, As


def alloc_outputs(torch, Ms, Ns):
    nS = int(Ns.sum())
    nU = int((Ms.astype(np.int64) ** 2).sum())
    nV = int((Ns.astype(np.int64) ** 2).sum())
    S = torch.zeros(nS, dtype=torch.float32, device="cuda")
    U = torch.zeros(nU, dtype=torch.float32, device="cuda")
    V = torch.zeros(nV, dtype=torch.float32, device="cuda")
    outs = []
    s = u = v = 0
    for M, N in zip(Ms.tolist(), Ns.tolist()):
        outs.append((S[s:s + N], U[u:u + M * M].view(M, M), V[v:v + N * N].view(N, N)))
        s += N
        u += M * M
        v += N * N
    return outs


def bidiag_bytes(M, N):
    """Algorithmic HBM bytes of the bidiagonalization's two GEMVs per column:
    4 bytes * sum_{i<N} 2 (M-i)(N-i)."""
    i = np.arange(N, dtype=np.float64)
    return float(8.0 * np.sum((M - i) * (N - i)))


def svd_flops(M, N):
    return 4.0 * M * M * N + 2.0 * M * N * N + 6.0 * N ** 3


def cpu_baseline(host, offs, Ms, Ns, budget_s=12.0):
    """The oracle (FP64 C restatement of the reference path, one thread) on the first
    matrices of the same workload, bounded by budget_s of CPU time."""
    import oracle
    oracle.lib()
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and done < len(Ms):
        A = host[offs[done]:offs[done + 1]].reshape(int(Ms[done]), int(Ns[done])).astype(np.float64)
        oracle.svd(A)
        done += 1
    dt = time.perf_counter() - t0
    # LAPACK dgesdd (all cores) on the same sample, for reference
    from scipy.linalg import lapack
    t1 = time.perf_counter()
    for i in range(done):
        A = host[offs[i]:offs[i + 1]].reshape(int(Ms[i]), int(Ns[i])).astype(np.float64)
        lapack.dgesdd(A, full_matrices=1)
    dt2 = time.perf_counter() - t1
    return dict(value=done / dt, unit="SVDs/s", cores=1, kind="port",
                sample="first %d matrices of the workload (oracle FP64 bidiag + D&C, 1 thread, %.1f s)" % (done, dt),
                lapack_dgesdd_svds_per_s=done / dt2, lapack_threads=os.cpu_count())


def accuracy_sample(torch, host, offs, Ms, Ns, outs, n=40):
    from svdcheck import measures
    idx = np.linspace(0, len(Ms) - 1, n).astype(int)
    worst = dict(sigma_err=0.0, resid_ratio=0.0, orth_u=0.0, orth_v=0.0)
    for i in idx:
        A = host[offs[i]:offs[i + 1]].reshape(int(Ms[i]), int(Ns[i]))
        S, U, Vt = (t.cpu().numpy() for t in outs[i])
        m = measures(A, S, U, Vt)
        for k in worst:
            worst[k] = max(worst[k], m[k])
    worst["matrices_checked"] = int(n)
    return worst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=10000)
    ap.add_argument("--plain-budget", type=float, default=20.0, help="seconds for the plain-path measurement")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only run the batched step (warmup + steps), for counter collection")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: independent SVDs shard by matrix across GPUs with no "
              "collective (see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import pacosvd
    pacosvd.lib()
    print("[bench] generating workload", flush=True)
    host, offs, Ms, Ns, As = make_workload(torch, args.count)
    outs = alloc_outputs(torch, Ms, Ns)
    torch.cuda.synchronize()

    # ---- batched path (eager, with phase events), warmup + timed steps
    pacosvd.timing_enable(True)
    for _ in range(args.warmup):
        pacosvd.svd_batched(As, outs=outs)
    torch.cuda.synchronize()
    print("[bench] warmup done", flush=True)
    phase_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pacosvd.svd_batched(As, outs=outs)
        phase_ms.append(pacosvd.timing_read())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    info = pacosvd.last_info(len(As))
    ms_step = 1000.0 * dt / args.steps
    svds_s = len(As) * args.steps / dt
    ph = {k: float(np.mean([p[k] for p in phase_ms])) for k in phase_ms[0]}
    print("[bench] batched %.1f ms/step, phases %s" % (ms_step, ph), flush=True)

    if args.profile_batched:
        print(json.dumps({"profile_batched_ms_per_step": ms_step, "phase_ms": ph}), flush=True)
        return

    # ---- batched path replayed from a hipGraph
    pacosvd.svd_batched(As, outs=outs, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pacosvd.svd_batched(As, outs=outs, graph=True)
    torch.cuda.synchronize()
    dt_g = time.perf_counter() - t0
    svds_s_graph = len(As) * args.steps / dt_g
    print("[bench] graph %.1f ms/step" % (1000.0 * dt_g / args.steps), flush=True)

    acc = accuracy_sample(torch, host, offs, Ms, Ns, outs)
    print("[bench] accuracy %s" % acc, flush=True)

    # ---- plain path on the same stream (first matrices, bounded time)
    plain_outs = alloc_outputs(torch, Ms[:1], Ns[:1])
    pacosvd.svd_plain(As[:4])
    torch.cuda.synchronize()
    done = 0
    t0 = time.perf_counter()
    while done < len(As) and time.perf_counter() - t0 < args.plain_budget:
        chunk = As[done:done + 250]
        pacosvd.svd_plain(chunk)
        torch.cuda.synchronize()
        done += len(chunk)
    dt_p = max(time.perf_counter() - t0, 1e-9)
    plain_svds_s = max(done, 1) / dt_p
    del plain_outs
    print("[bench] plain %d matrices %.2f s" % (done, dt_p), flush=True)

    # ---- roofline of the bidiagonalization kernel (one launch over all matrices)
    bb = sum(bidiag_bytes(int(m), int(n)) for m, n in zip(Ms, Ns))
    t_bid = ph["bidiag"] / 1000.0
    achieved = bb / t_bid / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, kernel="bidiag_wg_kernel",
                algorithmic_bytes_per_launch=bb, launch_ms=round(ph["bidiag"], 3))
    flops = sum(svd_flops(int(m), int(n)) for m, n in zip(Ms, Ns))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(host, offs, Ms, Ns, args.cpu_budget)
        print("[bench] cpu %s" % cpu, flush=True)

    best = max(svds_s, svds_s_graph)
    line = {
        "metric": "SVDs/sec + p50/p99 latency, served/batched vs plain, at sigma rel err <=1e-5",
        "value": round(best, 1),
        "unit": "SVDs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * len(As) / best, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: %d FP32 SVDs (S, U, Vt), M~U[16,512], N~U[8,M], N(0,1), seed 3"
                               % len(As), "matrices": len(As), "path": "paco_svd_batched (one launch per phase)"},
        "batched_eager_svds_per_s": round(svds_s, 1),
        "batched_graph_svds_per_s": round(svds_s_graph, 1),
        "plain_svds_per_s": round(plain_svds_s, 1),
        "plain_matrices_timed": done,
        "speedup_vs_plain": round(best / plain_svds_s, 2),
        "latency_ms": {"p50": round(ms_step, 2), "p99": round(ms_step, 2),
                       "note": "batched: every SVD of a step completes with the step"},
        "phase_ms": {k: round(v, 3) for k, v in ph.items()},
        "achieved_tflops": round(flops / (1000.0 * len(As) / best / 1000.0) / 1e12 / 1.0, 3),
        "accuracy_vs_fp64_lapack": acc,
        "info_nonzero": int(sum(1 for x in info if x != 0)),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
