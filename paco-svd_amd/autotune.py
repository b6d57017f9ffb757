"""Launch-shape autotuner of the batched bidiagonalization (run on the target GPU).

For every (M, N) power-of-two bucket whose matrices are too large to sit whole in LDS,
a batch of random matrices drawn inside the bucket is run through paco_svd_batched once
per candidate workgroup size (256, 512, 1024 threads; forced with PACO_BIDIAG_THREADS),
the bidiagonalization phase is timed with the library's hipEvent phase timer, and the
fastest size is written to autotune_table.json next to libpacosvd.so, which the
library reads at first use ("MxN": threads).

Usage: python paco-svd_amd/autotune.py [--count 256] [--out path]
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

CANDIDATES = (256, 512, 1024)
LDS_BYTES = 160 * 1024


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


def tune(count, seed=11):
    import torch
    import pacosvd
    pacosvd.lib()
    pacosvd.timing_enable(True)
    rng = np.random.default_rng(seed)
    table = {}
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
            os.environ["PACO_BIDIAG_THREADS"] = str(nt)
            try:
                pacosvd.svd_batched(As, outs=outs)  # warm
                torch.cuda.synchronize()
                best = float("inf")
                for _ in range(2):
                    pacosvd.svd_batched(As, outs=outs)
                    best = min(best, pacosvd.timing_read()["bidiag"])
                times[nt] = best
            finally:
                del os.environ["PACO_BIDIAG_THREADS"]
        choice = min(times, key=times.get)
        table["%dx%d" % (mb, nb)] = choice
        print("bucket %4dx%-4d  %s  -> %d" % (mb, nb, {k: round(v, 3) for k, v in times.items()}, choice),
              flush=True)
        del As, outs
        torch.cuda.empty_cache()
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--out", default=os.path.join(HERE, "autotune_table.json"))
    args = ap.parse_args()
    table = tune(args.count)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
