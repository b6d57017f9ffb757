/*
 * paco_svd.h -- C ABI of the MI355X (gfx950) SVD engine, libpacosvd.so.
 *
 * Drop-in for the two device entry points of DEAVEN-MITCH/paco-svd and their
 * global-memory buffer layouts (row-major FP32, M >= N; sizes as main.cpp:33-43):
 *   a[M*N] u[M*M] vt[N*N] d[N] e[N-1] tauq[N] taup[N-1] qt[2*N*N] wt[2*N*N] idx[16*N].
 * All pointers passed to the entry points are device pointers; every entry point is
 * stream-ordered and asynchronous (no host synchronisation inside).  Return value: 0 on
 * success, otherwise a hipError_t code (launch failure) or one of the PACO_E* codes below
 * (argument error, checked before anything is launched).
 *
 * Scratch and reentrancy.  Scratch the library owns (the workspace of the two entry
 * points when the caller passes NULL, the plain path's scratch, the batched path's plan)
 * is kept per stream: calls on different streams never share it, calls on one stream are
 * ordered by the stream.  Such a buffer is allocated on the first call on a stream and
 * grown when a later call needs more; hipMalloc / hipFree synchronise the device, so a
 * call that grows a buffer is neither asynchronous nor capturable.  The two entry points
 * with a caller workspace allocate nothing and may be captured in a hipGraph;
 * paco_svd_batched_graph records its own graph.
 *
 * Deviation from the reference: the reference passes a 40*32-byte user workspace
 * (main.cpp:46-48).  These entry points need the workspace sized by
 * paco_workspace_size(M, N) (block reflector factors, secular vectors, merge states).
 *
 * Status.  Each problem has a LAPACK-style info word: 0 ok, -1 a NaN or infinite entry
 * in the bidiagonal handed to the divide and conquer (the problem is then solved as the
 * zero matrix), -2 a grid barrier of the chip-wide bidiagonalization timed out (results
 * invalid), -3 the tiling blob given to paco_svd_DC does not describe the leaf partition
 * of N at the current leaf size (it was ignored and that partition used: the results are
 * valid), > 0 the number of secular roots that reached the iteration limit.  The entry
 * points keep it in their workspace at paco_info_offset(M, N): entry point (a) resets
 * it, (b) adds to it.  The whole-SVD paths report it through paco_last_info.
 *
 * Size limits: N <= PACO_MAX_N; the bidiagonalization keeps one column of M floats and a
 * few rows in the LDS of one workgroup (M + 21 min(N, 384) + 1100 floats <= 40960, i.e.
 * M up to about 31,000 for wide matrices, about 40,000 for narrow ones); larger problems
 * return PACO_EINVAL.
 */
#ifndef PACO_SVD_H
#define PACO_SVD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PACO_EINVAL (-1)   /* bad argument (M < N, N < 1, null pointer, size over the limits) */
#define PACO_ENOMEM (-2)   /* an internal buffer could not be allocated */
#define PACO_EINJECTED (-3) /* serving loop: failure injected by PACO_SERVE_FAIL_EVERY (tests) */
#define PACO_ETOOBIG (-4)   /* paco_svd_plain_noalloc: a job's scratch exceeds the fixed array */
#define PACO_MAX_N 7680    /* D&C deflation stage keeps one merge in LDS (20 bytes per row) */
#define PACO_LEAF_DEFAULT 6 /* columns per D&C leaf unless paco_set_leaf_size says otherwise */

/* Entry point (a): Golub-Kahan upper bidiagonalization, A = Q * B * P^T.
 * Replaces the kernel upper_bidiagonalization (upper_bidiagonalization.cpp:967-999),
 * launched at main.cpp:126-127 as ACLRT_LAUNCH_KERNEL(upper_bidiagonalization)
 * (40, stream, M, N, A, U, Vt, D, E, Tauq, Taup, workspace).
 * Out: a holds d/e on the band and the Householder vectors below / right of it
 * (unit heads implicit), d[N], e[N-1], tauq[N], taup[N-1], u = Q (M x M),
 * vt = P^T (N x N).  blockDim is accepted and ignored (the launch shape is chosen
 * by the library). */
int paco_upper_bidiagonalization(unsigned blockDim, void *stream, int M, int N, float *a, float *u, float *vt,
                                 float *d, float *e, float *tauq, float *taup, void *workspace);

/* Entry point (b): bidiagonal divide and conquer, B = U_B * S * V_B^T, followed by the
 * back-transforms u[:, :N] <- u[:, :N] U_B and vt <- V_B^T vt.
 * Replaces the kernel svd_DC (svd.cpp:2105-2118), launched at main.cpp:157-158 as
 * ACLRT_LAUNCH_KERNEL(svd_DC)(1, stream, M, N, A, U, Vt, D, E, Qt, Wt, idx, workspace,
 * Tiling).  In: the outputs of entry point (a).  Out: d = sigma (descending),
 * u = left singular vectors (columns N..M-1 keep Q's), vt = right singular vectors
 * (rows).  a, qt, wt and idx are scratch.  tiling_dev is the device copy of the blob
 * written by GenerateTiling or NULL.  The leaves are always the partition of N at the
 * current leaf size (paco_set_leaf_size); a blob is checked against it on the device and
 * info = -3 reports one that differs (see Status). */
int paco_svd_DC(unsigned blockDim, void *stream, int M, int N, float *a, float *u, float *vt, float *d, float *e,
                float *qt, float *wt, uint32_t *idx, void *workspace, const void *tiling_dev);

/* Host tiling (svd_tiling.cpp:67-80 getTilingSize, :83-130 GenerateTiling): size in
 * bytes of the D&C tiling blob for N columns, and the blob itself: a 64-byte header
 * (uint32 magic "PACO", version 2, N, leaf count, leaf offset 64, svdBlockNum, leaf size)
 * and {uint32 start_col, end_col} per leaf, for the current leaf size (with leaf size 3,
 * the reference's partition). */
int getTilingSize(int N, int svdBlockNum);
void GenerateTiling(int N, int svdBlockNum, uint8_t *tiling_host);

/* Bytes of device workspace the two entry points need for an M x N problem.
 * No counterpart in the reference (which needs only 40*32 bytes for its barriers). */
size_t paco_workspace_size(int M, int N);

/* Arithmetic of the GEMMs (merge products, back-transforms):
 * 0 = FP32 on the matrix cores (default), 1 = BF16x3 compensated split,
 * 2 = FP32 vector FMA.  FP64 is used for secular roots, deflation and Loewner z in
 * every mode.  Returns the previous mode.  No counterpart in the reference. */
int paco_set_precision(int mode);

/* ---- whole SVDs (no counterpart in the reference; they chain (a) and (b)) ---- */

/* One SVD job: device pointers.  A (M x N, read-only), S (N), U (M x M), Vt (N x N). */
typedef struct {
    int M, N;
    const float *A;
    float *S;
    float *U;
    float *Vt;
} paco_svd_job;

/* Plain path: every job runs through the two entry points, one matrix per launch
 * sequence (the baseline the batched path is measured against). */
int paco_svd_plain(const paco_svd_job *jobs, int count, void *stream);

/* LAPACK-style front end (the natural caller of the two entry points): the SVD of one
 * column-major m x n matrix A (leading dimension lda >= m), as LAPACK sgesdd with device
 * pointers: S (min(m, n) values, descending); jobz 'A': U (m x m, ldu >= m) and VT
 * (n x n, ldvt >= n); 'S': the first min(m, n) columns of U and rows of VT (ldvt >=
 * min(m, n)); 'N': S only (U, VT unused).  Any m, n: the reference rejects m < n
 * (main.cpp:25-29); here m < n runs on A^T, whose row-major form is the input itself.  A is
 * not modified ('O' is not offered).  Scratch is the stream's library-owned buffer (see
 * Scratch above); the info word goes to paco_last_info.  Returns 0, PACO_EINVAL for a bad
 * argument, or a hipError_t code.  No counterpart in the reference. */
int paco_sgesdd(char jobz, int m, int n, const float *A, int lda, float *S, float *U, int ldu, float *VT, int ldvt,
                void *stream);

/* Allocation-free variant of paco_svd_plain.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, hosts that forbid hipMalloc).  The jobs' scratch and info words live in a fixed
 * device array of 64 MiB that the library holds on every GPU it runs on until the
 * process ends; the call allocates nothing (no hipMalloc / hipFree).  Limits: at most 4096
 * jobs per call, and each job's scratch (its a, e, tau, D&C bases and the entry points'
 * workspace) within 64 MiB - 16 KiB: e.g. 1400 x 1400 needs 58.6 MiB, 2048 x 1024 40.1 MiB,
 * 4096 x 256 19.9 MiB (square matrices up to about 1450 x 1450); larger jobs return
 * PACO_ETOOBIG before anything is launched.  Every
 * call on one GPU shares that GPU's array, which keeps the last call's contents: calls
 * that overlap on one GPU (several streams or threads) race.  Results are bitwise those of
 * paco_svd_plain; paco_last_info reports the jobs' info words. */
int paco_svd_plain_noalloc(const paco_svd_job *jobs, int count, void *stream);

/* Batched path: all jobs in one sequence of launches, each phase one launch over
 * every matrix (tile shapes per size bucket from the autotune table).  jobs is a host
 * array.  Scratch comes from a library pool sized by paco_batch_workspace_size. */
int paco_svd_batched(const paco_svd_job *jobs, int count, void *stream);

/* Same as paco_svd_batched, but the sequence is recorded into a hipGraph on first use
 * of a batch signature (shapes and pointers) and replayed afterwards. */
int paco_svd_batched_graph(const paco_svd_job *jobs, int count, void *stream);

/* Singular values only (LAPACK jobz = 'N'), batched: S of every job, U and Vt ignored
 * (may be NULL).  Q and P^T are not formed, the divide and conquer keeps no left vectors
 * and only the first and last columns of each node's right vectors (all a merge reads),
 * and there is no back-transform.  graph != 0 records / replays a hipGraph per batch
 * signature as paco_svd_batched_graph does.  No counterpart in the reference (svd_DC always
 * forms the vectors, svd.cpp:1911-1981). */
int paco_svd_values_batched(const paco_svd_job *jobs, int count, int graph, void *stream);

/* Device bytes the batched path needs for these shapes (scratch + plan). */
size_t paco_batch_workspace_size(const int *M, const int *N, int count);

/* Info word per job (in call order) of the last paco_svd_plain / paco_svd_batched /
 * paco_svd_batched_graph call on this thread (see Status above).  Copies min(count, that
 * call's count) values (synchronises the stream); returns the count copied, 0 after
 * paco_release. */
int paco_last_info(int *info_host, int count, void *stream);

/* Byte offset of the int info word inside a workspace of paco_workspace_size(M, N)
 * bytes (written by the two entry points).  No counterpart in the reference. */
size_t paco_info_offset(int M, int N);

/* Merges of the last paco_svd_batched / paco_svd_batched_graph call on this stream,
 * summed per round (merge level) into out[PACO_MERGE_STATS * r + q], r < the returned
 * count (at most max_rounds): q = 0 merges, 1 sum of n (merged rows), 2 sum of k (poles
 * kept by deflation), 3 / 4 of them from the left / right child, 5 deflation rotations
 * that mixed a left and a right pole, 6 FLOPs of the two merge products in dense form
 * 2 k^2 (n + m), 7 FLOPs of the same products restricted to their nonzero blocks
 * ([U1 0; 0 U2] structure).  Synchronises the stream.  Returns the rounds written (0
 * when no batched call is on record), PACO_EINVAL for bad arguments.  No counterpart in
 * the reference (it logs nothing). */
#define PACO_MERGE_STATS 8
int paco_merge_stats(double *out, int max_rounds, void *stream);

/* Columns per leaf of the divide and conquer, 3..32 (default PACO_LEAF_DEFAULT).  Leaves
 * of 3 columns are the reference's partition (svd_tiling.cpp:83-130, solved in closed form
 * at svd.cpp:1596-1864); larger leaves are solved on chip by FP64 one-sided Jacobi (one
 * thread per leaf up to 6 columns, one wavefront per leaf above) and remove about
 * log2(ls / 3) merge rounds.  Applies to calls made after it
 * (and to GenerateTiling / getTilingSize, whose blob paco_svd_DC checks against it).
 * Returns the previous size; values outside 3..32 only query.  No counterpart in the
 * reference. */
int paco_set_leaf_size(int ls);
int paco_leaf_size(void);

/* Iteration limit of the secular solver per root (default 200; a root that reaches it
 * adds 1 to its problem's info).  Returns the previous limit; values < 1 only query.
 * No counterpart in the reference (svd.cpp:1144-1156 iterates without a check). */
int paco_set_secular_maxiter(int maxit);

/* C = alpha * op(A) * op(B) + beta * C on one stream with the engine's GEMM tile
 * (precision mode of paco_set_precision); row-major, op = transpose when trans != 0.
 * No counterpart in the reference (its products go through the Ascend Matmul API,
 * svd.cpp:78-80); exported so the tile can be checked on its own. */
int paco_gemm_f32(void *stream, int transA, int transB, int m, int n, int k, float alpha, const float *A, int lda,
                  const float *B, int ldb, float beta, float *C, int ldc);

/* Phase timing of the batched path with hipEvents recorded on the caller's stream
 * (phases: 0 copy-in, 1 bidiagonalization, 2 Q/P^T formation, 3 D&C leaves + merges,
 * 4 back-transforms).  With full vectors the batched D&C runs on a forked stream beside
 * the Q/P^T formation: phase 3 is then its own duration, overlapping phase 2, and phase 4
 * starts when both are done.  paco_timing_enable(1) turns recording on for later
 * paco_svd_batched calls; paco_timing_read synchronises on the last call's events and
 * writes up to max durations in milliseconds; returns the number written.
 * No counterpart in the reference (it has no timers, main.cpp:1). */
int paco_timing_enable(int on);
int paco_timing_read(float *ms, int max);

/* ---- serving loop (no counterpart in the reference; main.cpp runs one matrix) ----
 * Continuous batching of SVD requests whose buffers stay resident on the device.
 * hbm_budget_bytes bounds the library scratch of admitted requests (their footprint
 * is charged at admission and released at completion); narrow requests are packed in
 * arrival order into batches of at most max_batch on batch_lanes streams, wide ones
 * (N > 512 or M*N >= 2^20) run one at a time on wide_lanes streams through the two
 * entry points; lanes overlap on the device.  max_M, max_N size the wide lanes.
 * Returns an opaque handle or NULL. */
void *paco_server_create(size_t hbm_budget_bytes, int max_batch, int batch_lanes, int wide_lanes, int max_M,
                         int max_N);
/* Queue one request (thread-safe, returns at once).  tag is returned by poll.
 * PACO_EINVAL when the request could never be admitted (bad shape, over the budget). */
int paco_server_submit(void *server, const paco_svd_job *job, uint64_t tag);
/* Copy up to max finished requests (tag, latency in ms from submit to observed
 * completion; -1 for a request whose enqueue failed) and forget them; returns the count. */
int paco_server_poll(void *server, uint64_t *tags, double *latency_ms, int max);
/* paco_server_poll plus each request's info word (LAPACK-style, as paco_last_info: 0 ok,
 * -1 a NaN / Inf in the bidiagonal, -2 a wait of the chip-wide bidiagonalization timed
 * out, > 0 secular roots at the iteration limit; 0 for a failed enqueue).  A request
 * with a nonzero word finished, but its results are not to be trusted as they stand.
 * No counterpart in the reference (it has no server). */
int paco_server_poll_info(void *server, uint64_t *tags, double *latency_ms, int *info, int max);
/* Withdraw a queued request that no lane has taken yet: it is never enqueued, its output
 * buffers are left untouched and its tag never comes back from poll.  Returns 0 when
 * the request was withdrawn, 1 when no queued request has that tag (already admitted,
 * finished, being run again after a timed-out wait, or never submitted: it then
 * completes as usual), PACO_EINVAL for a NULL server.  With several queued requests of
 * one tag, the oldest is withdrawn.  No counterpart in the reference (it has no server). */
int paco_server_cancel(void *server, uint64_t tag);
/* Block until every queued request has finished (or failed); returns the first error or 0. */
int paco_server_drain(void *server);
/* Statistics since creation: [completed, p50 ms, p99 ms, mean ms, SVDs/s over first
 * submit .. last completion, batches, mean batch size, most lanes busy at once,
 * wide requests, failed requests, requests with a nonzero info word, wide requests run
 * again, requests withdrawn by paco_server_cancel]; returns the count written (at most
 * 13).  A wide request whose chip-wide bidiagonalization timed out (-2: its workgroups
 * could not all be resident beside the other lanes' work) runs once more, without
 * kernels whose workgroups wait on each other, before it is reported. */
int paco_server_stats(void *server, double *out, int max);
/* Drain and release the server. */
void paco_server_destroy(void *server);

/* Release the library's pools (device memory, pinned staging, graphs). */
void paco_release(void);

#ifdef __cplusplus
}
#endif

#endif /* PACO_SVD_H */
