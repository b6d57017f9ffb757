"""Writes the committed golden fixtures (run in a container with scipy's LAPACK).

The reference ships no golden vectors (its inputs come from an unseeded Fortran RNG and
its LAPACK goldens are never compared element-wise), so these come from FP64 LAPACK:
  c1.npz       the 128 x 96 uniform[-1,1] seed-0 matrix of the reference's CPU-sized
               configuration, its singular values from dgesdd (FP64), and the
               bidiagonal (alpha, beta) of an FP64 Golub-Kahan-Lanczos run started at
               e_1 with full reorthogonalisation (the bidiagonal is unique with d, e >= 0)
  secular.npz  secular-equation cases (d, z) and LAPACK dlasd4 roots for each i
Inputs are regenerated the same way by the tests, which also check the stored hashes.
Usage: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def c1_matrix():
    return np.random.default_rng(0).uniform(-1.0, 1.0, (128, 96)).astype(np.float32)


def lanczos_bidiag(A):
    """FP64 Golub-Kahan-Lanczos with v_1 = e_1 and full reorthogonalisation."""
    A = np.asarray(A, dtype=np.float64)
    M, N = A.shape
    U = np.zeros((M, N))
    V = np.zeros((N, N))
    alpha = np.zeros(N)
    beta = np.zeros(max(N - 1, 0))
    v = np.zeros(N)
    v[0] = 1.0
    V[:, 0] = v
    u = A @ v
    for j in range(N):
        if j > 0:
            u = A @ V[:, j] - beta[j - 1] * U[:, j - 1]
        for _ in range(2):
            u -= U[:, :j] @ (U[:, :j].T @ u)
        alpha[j] = np.linalg.norm(u)
        U[:, j] = u / alpha[j]
        if j == N - 1:
            break
        w = A.T @ U[:, j] - alpha[j] * V[:, j]
        for _ in range(2):
            w -= V[:, : j + 1] @ (V[:, : j + 1].T @ w)
        beta[j] = np.linalg.norm(w)
        V[:, j + 1] = w / beta[j]
    return alpha, beta


def secular_cases():
    rng = np.random.default_rng(7)
    cases = []
    for k, kind in [(5, "plain"), (17, "plain"), (40, "clustered"), (33, "tiny_z"), (64, "plain")]:
        d = np.sort(rng.uniform(0.0, 1.0, k))
        d[0] = 0.0
        if kind == "clustered":
            d[1:] = np.sort(0.5 + 1e-6 * rng.uniform(0.0, 1.0, k - 1))
        z = rng.standard_normal(k)
        if kind == "tiny_z":
            z *= 10.0 ** rng.uniform(-7, 0, k)
        rho = float(z @ z)
        zn = z / np.sqrt(rho)
        sig = np.array([lapack.dlasd4(i, d, zn, rho)[1] for i in range(k)])
        cases.append((d, z, sig))
    return cases


def main():
    A = c1_matrix()
    s = lapack.dgesdd(A.astype(np.float64), compute_uv=0)[1]
    alpha, beta = lanczos_bidiag(A)
    np.savez(os.path.join(HERE, "c1.npz"), A=A, sigma=s, alpha=alpha, beta=beta)
    cases = secular_cases()
    arrs = {}
    for i, (d, z, sig) in enumerate(cases):
        arrs["d%d" % i] = d
        arrs["z%d" % i] = z
        arrs["sig%d" % i] = sig
    np.savez(os.path.join(HERE, "secular.npz"), **arrs)
    manifest = {"c1_A_sha256": hashlib.sha256(A.tobytes()).hexdigest(), "secular_cases": len(cases)}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(manifest)


if __name__ == "__main__":
    main()
