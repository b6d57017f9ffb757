"""LAPACK-style front end paco_sgesdd on the GPU: column-major input with a padded
leading dimension, m >= n and m < n, jobz 'A', 'S', 'N', against FP64 LAPACK dgesdd.

Bars: sigma normwise error <= 1e-5; orthogonality of the returned U columns and VT rows
<= 1e-5; residual ||A - U S VT||_F / (||A||_F max(m, n) eps32) <= 10 (jobz 'A' and 'S').
The input buffer is unchanged.
"""
import sys

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu
FLT_EPS = float(np.finfo(np.float32).eps)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def colmajor_buffer(torch, A, ld):
    """A (m x n) into a flat column-major device buffer with leading dimension ld."""
    m, n = A.shape
    buf = np.full((n, ld), np.nan, dtype=np.float32)  # padding rows hold NaN: never read
    buf[:, :m] = A.T
    return torch.from_numpy(buf.ravel()).cuda()


def from_colmajor(t, rows, cols, ld):
    return t.cpu().numpy().reshape(cols, ld)[:, :rows].T.astype(np.float64)


@pytest.mark.parametrize("shape", [(60, 40), (40, 60), (1, 5), (5, 1), (300, 300), (257, 513), (1100, 900), (700, 1300)])
@pytest.mark.parametrize("jobz", ["A", "S", "N"])
def test_sgesdd(shape, jobz):
    import torch
    import pacosvd
    pacosvd.lib()
    m, n = shape
    k = min(m, n)
    rng = np.random.default_rng(m * 31 + n)
    A = rng.standard_normal((m, n)).astype(np.float32)
    lda, ldu, ldvt = m + 3, m + 1, (n if jobz == "A" else k) + 2
    say("sgesdd", shape, jobz)
    Ad = colmajor_buffer(torch, A, lda)
    A0 = Ad.clone()
    S = torch.zeros(k, device="cuda")
    ucols, vrows = (m, n) if jobz == "A" else (k, k)
    U = torch.zeros(ucols * ldu, device="cuda")
    VT = torch.zeros(n * ldvt, device="cuda")
    rc = pacosvd.sgesdd(jobz, m, n, Ad, lda, S, U, ldu, VT, ldvt)
    torch.cuda.synchronize()
    assert rc == 0
    assert pacosvd.last_info(1) == [0]
    assert torch.equal(torch.nan_to_num(Ad, nan=7.0), torch.nan_to_num(A0, nan=7.0))
    s = S.cpu().numpy().astype(np.float64)
    ref = workloads.lapack_sigma(A)
    assert np.max(np.abs(s - ref)) <= 1e-5 * ref[0], np.max(np.abs(s - ref)) / ref[0]
    if jobz == "N":
        return
    Un = from_colmajor(U, m, ucols, ldu)
    Vn = from_colmajor(VT, vrows, n, ldvt)
    assert np.max(np.abs(Un.T @ Un - np.eye(ucols))) <= 1e-5
    assert np.max(np.abs(Vn @ Vn.T - np.eye(vrows))) <= 1e-5
    R = A.astype(np.float64) - (Un[:, :k] * s[None, :]) @ Vn[:k, :]
    resid = np.linalg.norm(R) / (np.linalg.norm(A) * max(m, n) * FLT_EPS)
    assert resid <= 10.0, resid


def test_sgesdd_rejects_bad_arguments():
    import torch
    import pacosvd
    pacosvd.lib()
    A = torch.zeros(100, device="cuda")
    S = torch.zeros(10, device="cuda")
    assert pacosvd.sgesdd("O", 10, 10, A, 10, S, A, 10, A, 10) == -1
    assert pacosvd.sgesdd("A", 10, 10, A, 9, S, A, 10, A, 10) == -1   # lda < m
    assert pacosvd.sgesdd("A", 10, 5, A, 10, S, A, 9, A, 5) == -1     # ldu < m
    assert pacosvd.sgesdd("S", 5, 10, A, 5, S, A, 5, A, 4) == -1      # ldvt < min(m, n)
