"""Allocation-free plain path (paco_svd_plain_noalloc) against paco_svd_plain.

The variant keeps the jobs' scratch in the library's fixed device array instead of
memory it allocates; its kernels are those of the plain path, so the outputs must be
bitwise equal.  Two calls in a row, the second with smaller matrices (the array keeps the
first call's contents), then a job beyond the array's size, which is refused before
anything runs.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_noalloc_matches_plain_twice_then_refuses_too_big():
    import torch
    import pacosvd
    pacosvd.lib()
    rng = np.random.default_rng(77)
    calls = [[(300, 200), (129, 97), (64, 64), (1100, 1000)], [(100, 50), (33, 17), (8, 8)]]
    for shapes in calls:
        say("noalloc", shapes)
        As = [torch.from_numpy(rng.standard_normal(sh).astype(np.float32)).cuda() for sh in shapes]
        ref = pacosvd.svd_plain(As)
        torch.cuda.synchronize()
        ref_info = pacosvd.last_info(len(As))
        rc, outs = pacosvd.svd_plain_noalloc(As)
        torch.cuda.synchronize()
        assert rc == 0, rc
        assert pacosvd.last_info(len(As)) == ref_info == [0] * len(As)
        for i, (x, y) in enumerate(zip(outs, ref)):
            for a, b in zip(x, y):
                assert torch.equal(a, b), (shapes[i],)
    # a job whose scratch exceeds the fixed array: refused, outputs untouched
    say("noalloc too big 2048x2048")
    A = torch.zeros((2048, 2048), device="cuda")
    S = torch.full((2048,), 7.0, device="cuda")
    U = torch.zeros(1, device="cuda")
    job = pacosvd.SvdJob(2048, 2048, A.data_ptr(), S.data_ptr(), U.data_ptr(), U.data_ptr())
    rc = pacosvd.lib().paco_svd_plain_noalloc(ctypes.byref(job), 1, pacosvd._stream())
    torch.cuda.synchronize()
    assert rc == -4
    assert torch.all(S == 7.0)
