"""Autotuner of the engine's launch shapes (run on the target GPU).

Sweeps, each timed with the library's hipEvent phase timer, and writes the winners to
autotune_table.json next to libpacosvd.so, which the library reads at first use (the
PACO_* environment variables of each parameter still take precedence):

  "MxN": threads          workgroup size of the one-workgroup bidiagonalization of streamed
                          matrices (too large for LDS) per (M, N) power-of-two bucket:
                          256, 512 or 1024 threads (PACO_BIDIAG_THREADS); bidiagonalization
                          phase of a batch drawn inside the bucket
  "res_threads:c": t      workgroup size of the LDS-resident launches of LDS class c
                          (PACO_RES_THREADS); bidiagonalization phase of the resident
                          matrices of the 10,000-matrix batch
  "leaf_size": ls         columns per D&C leaf (paco_set_leaf_size); D&C phase of the
                          10,000-matrix batch run after Q and P^T (PACO_DC_SERIAL=1; the
                          4096 x 4096 matrix's D&C phase is reported beside it)
  "grid_tiles:MxN": g     tiles per side of the chip-wide bidiagonalization (PACO_GRID) and
  "grid_switch:MxN": s    the columns it leaves to the one-workgroup kernel
                          (PACO_GRID_SWITCH); bidiagonalization phase of one matrix

A candidate replaces the library's built-in default only when it is more than 2 % faster
(the run-to-run spread of these phases is about 1 %).

Usage: python paco-svd_amd/autotune.py [--count 128] [--out path] [--only threads,res,leaf,grid]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))

CANDIDATES = (256, 512, 1024)
LDS_BYTES = 160 * 1024
LEAF_SIZES = (3, 4, 5, 6, 8, 12, 16, 24, 32)
GRID_SHAPES = ((1024, 1024), (2048, 2048), (4096, 4096), (8192, 4096))
GRID_TILES = (8, 12, 16, 20, 24)
GRID_SWITCH = (256, 384, 512)


MARGIN = 0.02


def say(*a):
    print(*a, flush=True)


def pick(times, default):
    """The fastest candidate, unless the default is within MARGIN of it."""
    best = min(times, key=times.get)
    return default if default in times and times[default] <= times[best] * (1 + MARGIN) else best


def resident(M, N, nt=256):
    vec = M + 5 * N + (nt // 64) * N + nt + 8
    return 4 * (M * (N | 1) + vec) <= LDS_BYTES


def buckets():
    out = []
    for mb in (64, 128, 256, 512, 1024):
        for nb in (8, 16, 32, 64, 128, 256, 512):
            if nb <= mb and not resident(mb, nb):
                out.append((mb, nb))
    return out


class env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update({k: str(v) for k, v in self.kv.items()})

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def phase(torch, pacosvd, fn, name, reps=2):
    fn()  # warm
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        fn()
        torch.cuda.synchronize()
        best = min(best, pacosvd.timing_read()[name])
    return best


def tune_threads(torch, pacosvd, count, rng, table, log):
    for mb, nb in buckets():
        shapes = []
        for _ in range(count):
            M = int(rng.integers(mb // 2 + 1, mb + 1))
            N = int(rng.integers(max(1, nb // 2 + 1), min(nb, M) + 1))
            shapes.append((M, N))
        As = [torch.randn(sh, device="cuda") for sh in shapes]
        outs = pacosvd.alloc_outputs(As)
        times = {}
        for nt in CANDIDATES:
            with env(PACO_BIDIAG_THREADS=nt):
                times[nt] = phase(torch, pacosvd, lambda: pacosvd.svd_batched(As, outs=outs), "bidiag")
        choice = pick(times, 1024)
        table["%dx%d" % (mb, nb)] = choice
        log.append("threads %4dx%-4d %s -> %d" % (mb, nb, {k: round(v, 3) for k, v in times.items()}, choice))
        say(log[-1])
        del As, outs
        torch.cuda.empty_cache()


def c4_batch(torch):
    import workloads
    host, offs, Ms, Ns = workloads.c4_host()
    dev = torch.from_numpy(host).cuda()
    return [dev[offs[i]:offs[i + 1]].view(int(Ms[i]), int(Ns[i])) for i in range(len(Ms))]


def tune_res(torch, pacosvd, table, log, As):
    res = [A for A in As if resident(*A.shape)]
    outs = pacosvd.alloc_outputs(res)
    base = [512, 512, 1024, 1024]
    for cl in range(4):
        times = {}
        for nt in CANDIDATES:
            v = list(base)
            v[cl] = nt
            with env(PACO_RES_THREADS=",".join(map(str, v))):
                times[nt] = phase(torch, pacosvd, lambda: pacosvd.svd_batched(res, outs=outs), "bidiag")
        choice = pick(times, base[cl])
        base[cl] = choice
        table["res_threads:%d" % cl] = choice
        log.append("res_threads class %d (%d resident matrices) %s -> %d" % (
            cl, len(res), {k: round(v, 3) for k, v in times.items()}, choice))
        say(log[-1])


def tune_leaf(torch, pacosvd, table, log, As):
    import workloads
    outs = pacosvd.alloc_outputs(As)
    A2 = torch.from_numpy(workloads.c2_matrix()).cuda()

    def single():
        M, N = A2.shape
        r = pacosvd.upper_bidiagonalization(A2.clone())
        pacosvd.svd_DC(M, N, r["a"], r["u"], r["vt"], r["d"], r["e"], ws=r["ws"])

    times, times2 = {}, {}
    old = pacosvd.set_leaf_size(0)
    # the batched D&C alone: beside the formation of Q and P^T (the default schedule) its
    # duration also depends on how much of the chip it takes from that formation
    os.environ["PACO_DC_SERIAL"] = "1"
    try:
        for ls in LEAF_SIZES:
            pacosvd.set_leaf_size(ls)
            times[ls] = phase(torch, pacosvd, lambda: pacosvd.svd_batched(As, outs=outs), "dc")
            times2[ls] = phase(torch, pacosvd, single, "dc", reps=1)
            log.append("leaf_size %2d: D&C of the 10,000-matrix batch %.3f ms, of the 4096 x 4096 matrix %.3f ms"
                       % (ls, times[ls], times2[ls]))
            say(log[-1])
    finally:
        pacosvd.set_leaf_size(old)
        del os.environ["PACO_DC_SERIAL"]
    choice = pick(times, 6)
    table["leaf_size"] = choice
    log.append("leaf_size -> %d" % choice)
    say(log[-1])
    del outs


def tune_grid(torch, pacosvd, table, log):
    for M, N in GRID_SHAPES:
        A = torch.randn((M, N), device="cuda")
        ws = pacosvd.workspace(M, N)
        times = {}
        for g in GRID_TILES:
            for sw in GRID_SWITCH:
                with env(PACO_GRID=g, PACO_GRID_SWITCH=sw):
                    times[(g, sw)] = phase(torch, pacosvd,
                                           lambda: pacosvd.upper_bidiagonalization(A.clone(), ws=ws), "bidiag", reps=1)
        g, sw = pick(times, (16, 384))
        table["grid_tiles:%dx%d" % (M, N)] = g
        table["grid_switch:%dx%d" % (M, N)] = sw
        log.append("grid %dx%d: %s -> tiles %d, switch %d" % (
            M, N, ", ".join("%d/%d %.1f" % (k[0], k[1], v) for k, v in sorted(times.items())), g, sw))
        say(log[-1])
        del A, ws
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=128)
    ap.add_argument("--out", default=os.path.join(HERE, "autotune_table.json"))
    ap.add_argument("--log", default=None, help="also write the measured times here")
    ap.add_argument("--only", default="threads,res,leaf,grid")
    args = ap.parse_args()
    import torch
    import pacosvd
    pacosvd.lib()
    pacosvd.timing_enable(True)
    rng = np.random.default_rng(11)
    table, log = {}, []
    t0 = time.time()
    parts = args.only.split(",")
    if "threads" in parts:
        tune_threads(torch, pacosvd, args.count, rng, table, log)
    if "res" in parts or "leaf" in parts:
        As = c4_batch(torch)
        if "res" in parts:
            tune_res(torch, pacosvd, table, log, As)
        if "leaf" in parts:
            tune_leaf(torch, pacosvd, table, log, As)
        del As
        torch.cuda.empty_cache()
    if "grid" in parts:
        tune_grid(torch, pacosvd, table, log)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    say("wrote", args.out, "in %.0f s" % (time.time() - t0))
    if args.log:
        with open(args.log, "w") as f:
            f.write("\n".join(log) + "\n")


if __name__ == "__main__":
    main()
