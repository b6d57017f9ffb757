"""Synthetic inputs of the five benchmark configurations (shared by tests and bench.py).

Every generator is deterministic: numpy default_rng(seed) (PCG64), draws in the stated
order, generated in float64 and cast to float32 unless noted.
  C1  128 x 96, U[-1, 1], seed 0
  C2  4096 x 4096, N(0, 1), seed 1
  C3  8192 x 4096, A = Q1 diag(s) Q2^T, Haar Q1 (8192 x 4096) and Q2 (4096 x 4096) from
      the QR of Gaussian matrices with the sign of diag(R) folded in, s_i = 10^(-6 i / (N-1)),
      seed 2 (G1 drawn first, then G2)
  C4  10,000 matrices, M ~ U[16, 512], N ~ U[8, M], N(0, 1), seed 3 (all M, then all N,
      then the matrices in order)
  C5  Poisson arrivals at 2000 req/s, M log-uniform in [32, 4096], M / N ~ U[1, 4], 10%
      with clustered sigma (min(1000, N) values 1 + 1e-7 u), the rest log-uniform in
      [1e-3, 1], seed 5; matrices Q1 diag(sigma) Q2^T with Haar factors generated on the GPU
"""
import numpy as np


def c1_matrix():
    return np.random.default_rng(0).uniform(-1.0, 1.0, (128, 96)).astype(np.float32)


def c2_matrix(M=4096, N=4096, seed=1):
    return np.random.default_rng(seed).standard_normal((M, N)).astype(np.float32)


def haar_np(G):
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))[None, :]


def c3_matrix(M=8192, N=4096, seed=2):
    """Graded sigma_i = 10^(-6 i / (N - 1)); returns (A float32, designed sigma)."""
    rng = np.random.default_rng(seed)
    G1 = rng.standard_normal((M, N))
    G2 = rng.standard_normal((N, N))
    Q1 = haar_np(G1)
    del G1
    Q2 = haar_np(G2)
    del G2
    s = 10.0 ** (-6.0 * np.arange(N) / (N - 1))
    A = ((Q1 * s[None, :]) @ Q2.T).astype(np.float32)
    return A, s


def c4_shapes(count=10000, seed=3):
    rng = np.random.default_rng(seed)
    Ms = rng.integers(16, 513, size=count)
    Ns = np.array([rng.integers(8, m + 1) for m in Ms])
    return rng, Ms, Ns


def c4_host(count=10000, seed=3):
    """All matrices of the batch in one flat float32 array: (host, offsets, Ms, Ns)."""
    rng, Ms, Ns = c4_shapes(count, seed)
    sizes = Ms * Ns
    offs = np.concatenate([[0], np.cumsum(sizes)])
    host = np.empty(int(offs[-1]), dtype=np.float32)
    for i in range(count):
        host[offs[i]:offs[i + 1]] = rng.standard_normal((int(Ms[i]), int(Ns[i]))).astype(np.float32).ravel()
    return host, offs, Ms, Ns


def c5_requests(count, rate=2000.0, seed=5):
    """Request stream: arrival time t (s), M, N, clustered flag, designed sigma (descending)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        M = int(round(np.exp(rng.uniform(np.log(32), np.log(4096)))))
        a = rng.uniform(1.0, 4.0)
        N = max(1, int(round(M / a)))
        clustered = bool(rng.uniform() < 0.1)
        sig = 10.0 ** rng.uniform(-3.0, 0.0, N)
        if clustered:
            k = min(1000, N)
            sig[:k] = 1.0 + 1e-7 * rng.uniform(size=k)
        reqs.append(dict(t=t, M=M, N=N, clustered=clustered, sigma=np.sort(sig)[::-1].copy()))
    return reqs


def haar_matrix_gpu(torch, M, N, sigma, seed):
    """A = Q1 diag(sigma) Q2^T with Haar factors (GPU QR of Gaussian matrices, the sign of
    diag(R) folded in), torch generator seed `seed`, FP32."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    G1 = torch.randn((M, N), generator=g, device="cuda")
    G2 = torch.randn((N, N), generator=g, device="cuda")
    Q1, R1 = torch.linalg.qr(G1)
    Q2, R2 = torch.linalg.qr(G2)
    Q1 = Q1 * torch.sign(torch.diagonal(R1))[None, :]
    Q2 = Q2 * torch.sign(torch.diagonal(R2))[None, :]
    s = torch.from_numpy(sigma.astype(np.float32)).cuda()
    return ((Q1 * s[None, :]) @ Q2.T.contiguous()).contiguous()


def c5_matrices(torch, reqs):
    return [haar_matrix_gpu(torch, r["M"], r["N"], r["sigma"], 500000 + i) for i, r in enumerate(reqs)]


def lapack_sigma(A):
    """FP64 LAPACK dgesdd singular values (jobz = 'N') of a float32 / float64 matrix."""
    from scipy.linalg import lapack
    s = lapack.dgesdd(np.asarray(A, dtype=np.float64), compute_uv=0)
    return s[1]
