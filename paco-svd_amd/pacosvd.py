"""ctypes front end of libpacosvd.so for the Python harness (tests, bench, smoke).

The product is the C ABI in include/paco_svd.h; this module only passes torch device
pointers and the current HIP stream through it.  It never falls back to a CPU
implementation: if the library is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpacosvd.so")

EXPORTS = [
    "paco_upper_bidiagonalization", "paco_svd_DC", "getTilingSize", "GenerateTiling",
    "paco_workspace_size", "paco_set_precision", "paco_svd_plain", "paco_svd_batched",
    "paco_svd_batched_graph", "paco_svd_values_batched", "paco_sgesdd", "paco_svd_plain_noalloc", "paco_batch_workspace_size", "paco_last_info", "paco_release", "paco_gemm_f32",
    "paco_timing_enable", "paco_timing_read", "paco_info_offset", "paco_set_secular_maxiter",
    "paco_set_leaf_size", "paco_leaf_size", "paco_merge_stats",
    "paco_server_create", "paco_server_submit", "paco_server_poll", "paco_server_poll_info", "paco_server_cancel", "paco_server_drain",
    "paco_server_stats", "paco_server_destroy",
]
PHASES = ["copy_in", "bidiag", "form_qp", "dc", "backtransform"]

_lib = None
_vp = ctypes.c_void_p


class SvdJob(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("A", _vp), ("S", _vp), ("U", _vp), ("Vt", _vp)]


def lib():
    """Load libpacosvd.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libpacosvd.so not built: run python paco-svd_amd/build.py (%s)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.paco_upper_bidiagonalization.argtypes = [ctypes.c_uint, _vp, ctypes.c_int, ctypes.c_int] + [_vp] * 8
    L.paco_svd_DC.argtypes = [ctypes.c_uint, _vp, ctypes.c_int, ctypes.c_int] + [_vp] * 10
    L.getTilingSize.argtypes = [ctypes.c_int, ctypes.c_int]
    L.getTilingSize.restype = ctypes.c_int
    L.GenerateTiling.argtypes = [ctypes.c_int, ctypes.c_int, _vp]
    L.GenerateTiling.restype = None
    L.paco_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int]
    L.paco_workspace_size.restype = ctypes.c_size_t
    L.paco_set_precision.argtypes = [ctypes.c_int]
    L.paco_svd_plain.argtypes = [ctypes.POINTER(SvdJob), ctypes.c_int, _vp]
    L.paco_svd_batched.argtypes = [ctypes.POINTER(SvdJob), ctypes.c_int, _vp]
    L.paco_svd_plain_noalloc.argtypes = [ctypes.POINTER(SvdJob), ctypes.c_int, _vp]
    L.paco_svd_values_batched.argtypes = [ctypes.POINTER(SvdJob), ctypes.c_int, ctypes.c_int, _vp]
    L.paco_sgesdd.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp,
                              ctypes.c_int, _vp]
    L.paco_svd_batched_graph.argtypes = [ctypes.POINTER(SvdJob), ctypes.c_int, _vp]
    L.paco_batch_workspace_size.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.paco_batch_workspace_size.restype = ctypes.c_size_t
    L.paco_last_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, _vp]
    L.paco_info_offset.argtypes = [ctypes.c_int, ctypes.c_int]
    L.paco_info_offset.restype = ctypes.c_size_t
    L.paco_set_secular_maxiter.argtypes = [ctypes.c_int]
    L.paco_set_leaf_size.argtypes = [ctypes.c_int]
    L.paco_merge_stats.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, _vp]
    L.paco_leaf_size.argtypes = []
    L.paco_gemm_f32.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_float, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, _vp, ctypes.c_int]
    L.paco_timing_enable.argtypes = [ctypes.c_int]
    L.paco_timing_read.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    L.paco_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]
    L.paco_server_create.restype = _vp
    L.paco_server_submit.argtypes = [_vp, ctypes.POINTER(SvdJob), ctypes.c_uint64]
    L.paco_server_poll.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.paco_server_poll_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.paco_server_cancel.argtypes = [_vp, ctypes.c_uint64]
    L.paco_server_drain.argtypes = [_vp]
    L.paco_server_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.paco_server_destroy.argtypes = [_vp]
    L.paco_server_destroy.restype = None
    L.paco_release.argtypes = []
    L.paco_release.restype = None
    _lib = L
    return L


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with code %d" % (what, rc))


def tiling_blob(N, svd_block_num=1):
    size = lib().getTilingSize(N, svd_block_num)
    buf = (ctypes.c_uint8 * size)()
    lib().GenerateTiling(N, svd_block_num, buf)
    return bytes(buf)


def workspace(M, N, device="cuda"):
    import torch
    return torch.zeros(max(1, lib().paco_workspace_size(M, N)), dtype=torch.uint8, device=device)


def upper_bidiagonalization(a, ws=None):
    """Entry point (a) on device tensor a (M x N fp32, overwritten).  Returns dict of
    device tensors: a, u, vt, d, e, tauq, taup."""
    import torch
    M, N = a.shape
    dev = a.device
    u = torch.zeros((M, M), dtype=torch.float32, device=dev)
    vt = torch.zeros((N, N), dtype=torch.float32, device=dev)
    d = torch.zeros(N, dtype=torch.float32, device=dev)
    e = torch.zeros(max(N - 1, 1), dtype=torch.float32, device=dev)
    tq = torch.zeros(N, dtype=torch.float32, device=dev)
    tp = torch.zeros(max(N - 1, 1), dtype=torch.float32, device=dev)
    if ws is None:
        ws = workspace(M, N, dev)
    rc = lib().paco_upper_bidiagonalization(40, _stream(), M, N, _ptr(a), _ptr(u), _ptr(vt), _ptr(d), _ptr(e),
                                            _ptr(tq), _ptr(tp), _ptr(ws))
    _check(rc, "paco_upper_bidiagonalization")
    return dict(a=a, u=u, vt=vt, d=d, e=e[: N - 1], tauq=tq, taup=tp[: N - 1], ws=ws)


def svd_DC(M, N, a, u, vt, d, e, ws=None, use_tiling=True, blob=None):
    """Entry point (b): d <- sigma, u, vt <- singular vectors (in place).  The tiling blob
    is GenerateTiling's for N unless blob (bytes) is given."""
    import torch
    dev = u.device
    qt = torch.zeros(2 * N * N, dtype=torch.float32, device=dev)
    wt = torch.zeros(2 * N * N, dtype=torch.float32, device=dev)
    idx = torch.zeros(16 * N, dtype=torch.int32, device=dev)
    if ws is None:
        ws = workspace(M, N, dev)
    tiling = None
    if use_tiling:
        blob = tiling_blob(N) if blob is None else blob
        tiling = torch.frombuffer(bytearray(blob), dtype=torch.uint8).to(dev)
    if e.numel() < max(N - 1, 1):
        e = torch.zeros(1, dtype=torch.float32, device=dev)
    rc = lib().paco_svd_DC(1, _stream(), M, N, _ptr(a), _ptr(u), _ptr(vt), _ptr(d), _ptr(e), _ptr(qt), _ptr(wt),
                           _ptr(idx), _ptr(ws), _ptr(tiling))
    _check(rc, "paco_svd_DC")
    return d, u, vt


def _jobs(As, outs):
    jobs = (SvdJob * len(As))()
    for i, (A, (S, U, Vt)) in enumerate(zip(As, outs)):
        jobs[i].M, jobs[i].N = A.shape
        jobs[i].A = A.data_ptr()
        jobs[i].S = S.data_ptr()
        jobs[i].U = U.data_ptr()
        jobs[i].Vt = Vt.data_ptr()
    return jobs


def alloc_outputs(As):
    import torch
    outs = []
    for A in As:
        M, N = A.shape
        outs.append((torch.zeros(N, dtype=torch.float32, device=A.device),
                     torch.zeros((M, M), dtype=torch.float32, device=A.device),
                     torch.zeros((N, N), dtype=torch.float32, device=A.device)))
    return outs


def svd_plain(As, outs=None):
    """Plain path: each matrix through the two entry points.  Returns [(S, U, Vt)]."""
    outs = outs if outs is not None else alloc_outputs(As)
    _check(lib().paco_svd_plain(_jobs(As, outs), len(As), _stream()), "paco_svd_plain")
    return outs


def svd_plain_noalloc(As, outs=None):
    """Allocation-free plain path (scratch in the library's fixed device array).  Returns
    the library's return code and [(S, U, Vt)]."""
    outs = outs if outs is not None else alloc_outputs(As)
    rc = lib().paco_svd_plain_noalloc(_jobs(As, outs), len(As), _stream())
    return rc, outs


def sgesdd(jobz, m, n, A, lda, S, U=None, ldu=1, VT=None, ldvt=1):
    """LAPACK-style SVD of a column-major m x n matrix held in device tensor A (flat, leading
    dimension lda); S, U, VT are flat device tensors (column-major).  Returns the code."""
    return lib().paco_sgesdd(jobz.encode() if isinstance(jobz, str) else jobz, m, n, _ptr(A), lda, _ptr(S), _ptr(U),
                             ldu, _ptr(VT), ldvt, _stream())


def svd_values(As, S=None, graph=False):
    """Singular values only of every matrix (batched): returns [S]."""
    import torch
    S = S if S is not None else [torch.zeros(A.shape[1], dtype=torch.float32, device=A.device) for A in As]
    jobs = (SvdJob * len(As))()
    for i, (A, s) in enumerate(zip(As, S)):
        jobs[i].M, jobs[i].N = A.shape
        jobs[i].A = A.data_ptr()
        jobs[i].S = s.data_ptr()
        jobs[i].U = None
        jobs[i].Vt = None
    _check(lib().paco_svd_values_batched(jobs, len(As), 1 if graph else 0, _stream()), "paco_svd_values_batched")
    return S


def svd_batched(As, outs=None, graph=False):
    """Batched path (one launch per phase over all matrices)."""
    outs = outs if outs is not None else alloc_outputs(As)
    svd_batched_jobs(_jobs(As, outs), len(As), graph)
    return outs


def prepare_jobs(As, outs):
    """The job descriptor array of matrices As and outputs outs (device pointers), built
    once for repeated calls of svd_batched_jobs on the same buffers."""
    return _jobs(As, outs)


def svd_batched_jobs(jobs, count, graph=False):
    """Batched path on a prepared job array (prepare_jobs)."""
    fn = lib().paco_svd_batched_graph if graph else lib().paco_svd_batched
    _check(fn(jobs, count, _stream()), "paco_svd_batched")


def last_info(count):
    buf = (ctypes.c_int * count)()
    n = lib().paco_last_info(buf, count, _stream())
    return list(buf[:n])


def gemm(A, B, C=None, transA=False, transB=False, alpha=1.0, beta=0.0):
    """C = alpha op(A) op(B) + beta C with the engine's GEMM tile (row-major fp32)."""
    import torch
    m = A.shape[1] if transA else A.shape[0]
    k = A.shape[0] if transA else A.shape[1]
    n = B.shape[0] if transB else B.shape[1]
    if C is None:
        C = torch.zeros((m, n), dtype=torch.float32, device=A.device)
    rc = lib().paco_gemm_f32(_stream(), int(transA), int(transB), m, n, k, alpha, _ptr(A), A.stride(0), _ptr(B),
                             B.stride(0), beta, _ptr(C), C.stride(0))
    _check(rc, "paco_gemm_f32")
    return C


def set_precision(mode):
    return lib().paco_set_precision(mode)


MERGE_STATS = ("merges", "sum_n", "sum_k", "sum_k_left", "sum_k_right", "mixed_rotations", "dense_flops",
               "block_flops")


def merge_stats(max_rounds=32):
    """Per merge round of the last batched call on the current stream: dicts of
    MERGE_STATS (realised k of deflation, flops of the merge products)."""
    buf = (ctypes.c_double * (len(MERGE_STATS) * max_rounds))()
    n = lib().paco_merge_stats(buf, max_rounds, _stream())
    _check(min(n, 0), "paco_merge_stats")
    return [dict(zip(MERGE_STATS, buf[len(MERGE_STATS) * r:len(MERGE_STATS) * (r + 1)])) for r in range(n)]


def set_leaf_size(ls):
    """Columns per divide-and-conquer leaf (3..32) for later calls; returns the previous
    size (values outside 3..32 only query)."""
    return lib().paco_set_leaf_size(ls)


def set_secular_maxiter(maxit):
    """Iteration limit of the secular solver per root; returns the previous one."""
    return lib().paco_set_secular_maxiter(maxit)


def entry_info(ws, M, N):
    """Info word the two entry points wrote into workspace tensor ws (synchronises)."""
    import torch
    off = lib().paco_info_offset(M, N)
    return int(ws[off:off + 4].view(torch.int32).item())


def timing_enable(on=True):
    _check(lib().paco_timing_enable(1 if on else 0), "paco_timing_enable")


def timing_read():
    """Per-phase milliseconds of the last batched call (dict), via hipEvents on its stream."""
    buf = (ctypes.c_float * len(PHASES))()
    n = lib().paco_timing_read(buf, len(PHASES))
    return {PHASES[i]: float(buf[i]) for i in range(n)}


SERVER_STATS = ["completed", "p50_ms", "p99_ms", "mean_ms", "svds_per_s", "batches", "mean_batch", "max_lanes_busy",
                "wide_requests", "failed", "nonzero_info", "retried", "cancelled"]


class Server:
    """Continuous-batching SVD server (paco_server_*).  Requests reference device
    tensors that must stay alive until they are reported finished."""

    def __init__(self, hbm_budget_bytes, max_batch=256, batch_lanes=2, wide_lanes=1, max_M=4096, max_N=4096):
        self.h = lib().paco_server_create(hbm_budget_bytes, max_batch, batch_lanes, wide_lanes, max_M, max_N)
        if not self.h:
            raise RuntimeError("paco_server_create failed")

    def submit(self, A, S, U, Vt, tag):
        job = SvdJob(A.shape[0], A.shape[1], A.data_ptr(), S.data_ptr(), U.data_ptr(), Vt.data_ptr())
        return lib().paco_server_submit(self.h, ctypes.byref(job), tag)

    def poll(self, max_n=4096):
        tags = (ctypes.c_uint64 * max_n)()
        lat = (ctypes.c_double * max_n)()
        n = lib().paco_server_poll(self.h, tags, lat, max_n)
        return [(int(tags[i]), float(lat[i])) for i in range(n)]

    def poll_info(self, max_n=4096):
        """Finished requests as (tag, latency ms, info word)."""
        tags = (ctypes.c_uint64 * max_n)()
        lat = (ctypes.c_double * max_n)()
        info = (ctypes.c_int * max_n)()
        n = lib().paco_server_poll_info(self.h, tags, lat, info, max_n)
        return [(int(tags[i]), float(lat[i]), int(info[i])) for i in range(n)]

    def cancel(self, tag):
        """Withdraw a queued request (paco_server_cancel): True when it was withdrawn,
        False when no queued request has that tag (it then completes as usual)."""
        rc = lib().paco_server_cancel(self.h, tag)
        if rc not in (0, 1):
            _check(rc, "paco_server_cancel")
        return rc == 0

    def drain(self):
        _check(lib().paco_server_drain(self.h), "paco_server_drain")

    def stats(self):
        buf = (ctypes.c_double * len(SERVER_STATS))()
        n = lib().paco_server_stats(self.h, buf, len(SERVER_STATS))
        return {SERVER_STATS[i]: float(buf[i]) for i in range(n)}

    def close(self):
        if self.h:
            lib().paco_server_destroy(self.h)
            self.h = None
