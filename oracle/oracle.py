"""TEST INFRASTRUCTURE ONLY -- ctypes front end of the oracle (paco_oracle.c).

The oracle is an FP64 CPU restatement of the reference's path (Golub-Kahan
bidiagonalization, upper_bidiagonalization.cpp; bidiagonal divide and conquer,
svd.cpp).  It is the checker: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FLT_EPS = float(np.finfo(np.float32).eps)
DBL_EPS = float(np.finfo(np.float64).eps)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class MergeStat(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("k", ctypes.c_int), ("nrot", ctypes.c_int), ("ndefl", ctypes.c_int)]


def build():
    """Compile liboracle.so with gcc (no GPU, no reference sources needed)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_house.argtypes = [ctypes.c_int, _dp, ctypes.c_long, _dp, _dp]
        L.orc_bidiag.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 7
        _fp = ctypes.POINTER(ctypes.c_float)
        L.orc_bidiag_f32.argtypes = [ctypes.c_int, ctypes.c_int] + [_fp] * 5
        L.orc_bidiag_f32_order.argtypes = [ctypes.c_int, ctypes.c_int] + [_fp] * 5 + [ctypes.c_int]
        L.orc_leaf_partition.argtypes = [ctypes.c_int, _ip, _ip]
        L.orc_leaf_partition.restype = ctypes.c_int
        L.orc_leaf_svd.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_secular_root.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _ip, _dp]
        L.orc_secular_root.restype = ctypes.c_int
        L.orc_bdc.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.POINTER(MergeStat)]
        L.orc_bdc.restype = ctypes.c_int
        L.orc_svd_DC.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_svd.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double]
        _lib = L
    return _lib


def _p(x):
    return x.ctypes.data_as(_dp)


def house(x):
    """Reference-convention Householder of a 1-D vector; returns (v_stored, beta, tau)."""
    x = np.array(x, dtype=np.float64)
    beta = ctypes.c_double()
    tau = ctypes.c_double()
    lib().orc_house(len(x), _p(x), 1, ctypes.byref(beta), ctypes.byref(tau))
    return x, beta.value, tau.value


def bidiag_f32(A, order=0):
    """The bidiagonalization in FP32 (the yardstick of FP32 forward error): dict(a, d, e,
    tauq, taup) in the reference layout.  order selects the evaluation order of the
    reflector inner products (0 forward, 1 backward, 2 eight interleaved partial sums)."""
    a = np.ascontiguousarray(A, dtype=np.float32).copy()
    M, N = a.shape
    d = np.zeros(N, np.float32)
    e = np.zeros(max(N - 1, 1), np.float32)
    tq = np.zeros(N, np.float32)
    tp = np.zeros(max(N - 1, 1), np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib().orc_bidiag_f32_order(M, N, a.ctypes.data_as(fp), d.ctypes.data_as(fp), e.ctypes.data_as(fp),
                               tq.ctypes.data_as(fp), tp.ctypes.data_as(fp), int(order))
    return dict(a=a, d=d, e=e[: N - 1], tauq=tq, taup=tp[: N - 1])


def bidiag(A):
    """Golub-Kahan upper bidiagonalization of A (M x N, M >= N), reference layout.

    Returns dict(a, u, vt, d, e, tauq, taup) with A = u @ B @ vt."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    M, N = A.shape
    a = A.copy()
    u = np.zeros((M, M))
    vt = np.zeros((N, N))
    d = np.zeros(N)
    e = np.zeros(max(N - 1, 1))
    tq = np.zeros(N)
    tp = np.zeros(max(N - 1, 1))
    lib().orc_bidiag(M, N, _p(a), _p(u), _p(vt), _p(d), _p(e), _p(tq), _p(tp))
    return dict(a=a, u=u, vt=vt, d=d, e=e[: N - 1], tauq=tq, taup=tp[: N - 1])


def leaf_partition(N):
    s = (ctypes.c_int * (N // 3 + 2))()
    t = (ctypes.c_int * (N // 3 + 2))()
    cnt = lib().orc_leaf_partition(N, s, t)
    return [(s[i], t[i]) for i in range(cnt)]


def leaf_svd(n, m, dd, ee):
    dd = np.array(list(dd) + [0.0] * (2 - len(dd)), dtype=np.float64)
    ee = np.array(list(ee) + [0.0] * (2 - len(ee)), dtype=np.float64)
    sig = np.zeros(2)
    U = np.zeros(4)
    V = np.zeros(9)
    lib().orc_leaf_svd(n, m, _p(dd), _p(ee), _p(sig), _p(U), _p(V))
    return sig[:n], U[: n * n].reshape(n, n), V[: m * m].reshape(m, m)


def secular_root(d, z, i):
    """Root i of 1 + sum z^2/(d^2 - s^2) = 0; returns (sigma, K, mu, iters)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    s = ctypes.c_double()
    K = ctypes.c_int()
    mu = ctypes.c_double()
    it = lib().orc_secular_root(len(d), _p(d), _p(z), i, ctypes.byref(s), ctypes.byref(K), ctypes.byref(mu))
    return s.value, K.value, mu.value, it


def bdc(d, e, defl_eps=FLT_EPS, with_stats=False):
    """D&C SVD of the square upper bidiagonal (d, e): returns (sigma, qt, wt[, stats])."""
    d = np.array(d, dtype=np.float64)
    N = len(d)
    e = np.array(list(e) + [0.0], dtype=np.float64)
    qt = np.zeros((N, N))
    wt = np.zeros((N, N))
    stats = (MergeStat * max(N, 1))()
    nm = lib().orc_bdc(N, _p(d), _p(e), _p(qt), _p(wt), defl_eps, stats)
    if with_stats:
        return d, qt, wt, [(s.n, s.k, s.nrot, s.ndefl) for s in stats[:nm]]
    return d, qt, wt


def svd_DC(u, vt, d, e, defl_eps=FLT_EPS):
    """svd_DC restated: consumes bidiag outputs, returns (sigma, U, Vt)."""
    u = np.array(u, dtype=np.float64)
    vt = np.array(vt, dtype=np.float64)
    d = np.array(d, dtype=np.float64)
    M, N = u.shape[0], vt.shape[0]
    e = np.array(list(e) + [0.0], dtype=np.float64)
    lib().orc_svd_DC(M, N, _p(u), _p(vt), _p(d), _p(e), defl_eps)
    return d, u, vt


def svd(A, defl_eps=FLT_EPS):
    """Whole path on one matrix: returns (sigma, U (M x M), Vt (N x N))."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    M, N = A.shape
    s = np.zeros(N)
    U = np.zeros((M, M))
    Vt = np.zeros((N, N))
    lib().orc_svd(M, N, _p(A), _p(s), _p(U), _p(Vt), defl_eps)
    return s, U, Vt
