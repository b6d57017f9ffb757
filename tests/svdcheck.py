"""Accuracy measures stated for every result (verify_svd.py:35-130 restated in FP64)."""
import numpy as np

FLT_EPS = float(np.finfo(np.float32).eps)


def lapack_sigma(A):
    from scipy.linalg import lapack
    return lapack.dgesdd(np.asarray(A, dtype=np.float64), compute_uv=0)[1]


def measures(A, s, U, Vt, s_ref=None):
    """sigma error (normwise), residual ratio ||A - U S Vt||_F / (||A||_F max(M,N) eps32),
    orthogonality max|U^T U - I|, max|V^T V - I| (entry-wise, as verify_svd.py:44-48)."""
    A = np.asarray(A, dtype=np.float64)
    M, N = A.shape
    s = np.asarray(s, dtype=np.float64)
    U = np.asarray(U, dtype=np.float64)
    Vt = np.asarray(Vt, dtype=np.float64)
    if s_ref is None:
        s_ref = lapack_sigma(A)
    smax = max(float(s_ref[0]), 1e-300)
    sig_err = float(np.max(np.abs(s - s_ref)) / smax)
    R = A - (U[:, :N] * s[None, :]) @ Vt
    an = max(np.linalg.norm(A), 1e-300)
    resid = float(np.linalg.norm(R) / (an * max(M, N) * FLT_EPS))
    orth_u = float(np.max(np.abs(U.T @ U - np.eye(M))))
    orth_v = float(np.max(np.abs(Vt @ Vt.T - np.eye(N))))
    return dict(sigma_err=sig_err, resid_ratio=resid, orth_u=orth_u, orth_v=orth_v)


def measures_gpu(torch, A, s, U, Vt, s_ref):
    """The same measures computed on the GPU in FP64 (A, s, U, Vt device tensors, s_ref a
    host FP64 array from LAPACK); adds the largest per-value relative sigma error."""
    A64 = A.double()
    M, N = A64.shape
    s64 = s.double()
    U64 = U.double()
    V64 = Vt.double()
    ref = torch.from_numpy(np.asarray(s_ref, dtype=np.float64)).to(A.device)
    smax = max(float(ref[0]), 1e-300)
    sig_err = float((s64 - ref).abs().max()) / smax
    per = float(((s64 - ref).abs() / ref.clamp_min(1e-300)).max())
    R = A64 - (U64[:, :N] * s64[None, :]) @ V64
    an = max(float(torch.linalg.norm(A64)), 1e-300)
    resid = float(torch.linalg.norm(R)) / (an * max(M, N) * FLT_EPS)
    eye_m = torch.eye(M, dtype=torch.float64, device=A.device)
    eye_n = torch.eye(N, dtype=torch.float64, device=A.device)
    orth_u = float((U64.T @ U64 - eye_m).abs().max())
    orth_v = float((V64 @ V64.T - eye_n).abs().max())
    return dict(sigma_err=sig_err, sigma_rel_err_per_value_max=per, resid_ratio=resid, orth_u=orth_u, orth_v=orth_v)
