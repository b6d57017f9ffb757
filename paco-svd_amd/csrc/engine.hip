// Host orchestration of the gfx950 SVD engine and its C ABI (include/paco_svd.h).
//
// Plain entry points (paco_upper_bidiagonalization / paco_svd_DC) take one problem by
// value: every schedule (leaves, merge rounds, tiles) is derived on the device from
// (M, N), so no host plan is uploaded and the calls are pure stream-ordered launches
// (main.cpp:126-160 launches the two reference kernels the same way).  The batched
// path uploads one small plan (problem descriptors + per-problem prefix arrays) and
// runs each phase as one launch over every matrix.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/paco_svd.h"
#include "lane.hpp"
#include "paco_kernels.hpp"

namespace paco {

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// tiny kernels of the engine
// ---------------------------------------------------------------------------
__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void copy_int_kernel(int *dst, const int *src)
{
    if (threadIdx.x == 0) *dst = *src;
}

// Input guard of the divide and conquer, one workgroup per problem: a bidiagonal with a
// NaN or infinite entry (non-finite input, or overflow in the bidiagonalization) sets
// info = -1 and is replaced by zeros, so the deflation's sorting and compaction only
// ever see finite poles.  The reference checks nothing (svd.cpp:134-195).
// Before the D&C of each problem: a NaN / infinite entry of d or e sets info = -1 and the
// problem is solved as the zero matrix.  With a tiling blob (entry point (b), one problem),
// the blob must describe the leaf partition of N with leaves of ls columns (header: magic
// "PACO", N, leaf count, leaf offset 64, leaf size; then {start, end} per leaf); otherwise
// info = -3.  The leaves are always taken from that partition, never from the blob, so a
// blob made for another N or by another tool cannot steer writes.
__global__ __launch_bounds__(256) void dc_guard_kernel(ProbList pl, const uint32_t *blob, int ls)
{
    const ProbDesc P = pl.get(blockIdx.x);
    const int N = P.N;
    __shared__ int bad, blob_bad;
    if (threadIdx.x == 0) {
        bad = 0;
        blob_bad = 0;
        if (blob) {
            const int L = leaf_count(N, ls);
            blob_bad = blob[0] != 0x4F434150u || blob[2] != (uint32_t)N || blob[3] != (uint32_t)L || blob[4] != 64u ||
                       blob[6] != (uint32_t)ls;
        }
    }
    __syncthreads();
    if (blob && !blob_bad) { // the header vouches for leaf_count(N, ls) entries
        int bb = 0;
        for (int l = threadIdx.x; l < leaf_count(N, ls); l += 256)
            bb |= blob[16 + 2 * l] != (uint32_t)leaf_start(N, l, ls) || blob[17 + 2 * l] != (uint32_t)leaf_start(N, l + 1, ls);
        if (bb) blob_bad = 1;
    }
    int b = 0;
    for (int i = threadIdx.x; i < N; i += 256) b |= !isfinite(P.d[i]);
    for (int i = threadIdx.x; i < N - 1; i += 256) b |= !isfinite(P.e[i]);
    if (b) bad = 1;
    __syncthreads();
    if (!bad) {
        if (threadIdx.x == 0 && blob_bad) *P.info = -3;
        return;
    }
    for (int i = threadIdx.x; i < N; i += 256) P.d[i] = 0.f;
    for (int i = threadIdx.x; i < N - 1; i += 256) P.e[i] = 0.f;
    if (threadIdx.x == 0) *P.info = -1;
}

// a <- A (row-major M x N) for every problem: rows [ROW_GROUP g, ROW_GROUP (g + 1)) of
// every input into its working copy
__global__ __launch_bounds__(256) void copy_in_kernel(ProbList pl, const float *const *src, const int *row_off)
{
    const int pid = find_task(row_off, pl.count, blockIdx.x);
    const int g = blockIdx.x - off_at(row_off, pid);
    const ProbDesc P = pl.get(pid);
    const int r0 = g * BATCH_ROW_GROUP, nr = min(BATCH_ROW_GROUP, P.M - r0);
    if (nr <= 0) return;
    // the rows of a group are contiguous in both (leading dimension N)
    copy_block(P.a + (size_t)r0 * P.N, src[pid] + (size_t)r0 * P.N, nr * P.N);
}

// ---------------------------------------------------------------------------
// single-problem workspace
// ---------------------------------------------------------------------------
struct WsLayout {
    size_t vmt, tq, tp, info, states, grid, qp, total;
};

// matrices that take the chip-wide bidiagonalization and the GEMM formation of Q, P
static bool large_matrix(int M, int N) { return N > BIDIAG_WG_MAXN || (size_t)M * N >= (size_t)1024 * 1024; }

// merges of the D&C tree: most with the smallest leaves (scratch is sized for those)
static int total_merges(int N) { return leaf_count(N, LEAF_MIN) - 1; }

static bool use_grid_bidiag(int M, int N);

static WsLayout ws_layout(int M, int N)
{
    WsLayout L;
    size_t o = 0;
    L.vmt = o;
    o = align_up(o + sizeof(float) * (size_t)N * N, 256);
    L.tq = o;
    o = align_up(o + sizeof(float) * 1024 * (size_t)cdiv(N, 32), 256);
    L.tp = o;
    o = align_up(o + sizeof(float) * 1024 * (size_t)cdiv(std::max(N - 1, 1), 32), 256);
    L.info = o;
    o = align_up(o + 64, 256);
    L.states = o;
    o = align_up(o + sizeof(MergeState) * (size_t)std::max(total_merges(N), 1), 256);
    L.grid = o; // the chip-wide bidiagonalization's buffers, for the shapes that take it
    if (use_grid_bidiag(M, N)) o = align_up(o + bidiag_grid_ws_bytes(M, N), 256);
    L.qp = o;
    if (large_matrix(M, N)) o = align_up(o + qp_gemm_ws_bytes(M, N), 256);
    L.total = o;
    return L;
}


static std::mutex g_mu;
// bumped by paco_release: per-thread pointers into freed pools compare against it
static unsigned g_release_gen = 0;

// Device buffers the library owns on behalf of one stream: callers on different streams
// never share scratch, and calls on one stream are ordered by the stream itself.  A buffer
// grows when a call needs more (hipFree / hipMalloc, which synchronise the device);
// calls that fit the current sizes allocate nothing.
struct BatchRes;
static void batch_res_destroy(BatchRes &r);
struct DevBuf {
    void *p = nullptr;
    size_t size = 0;
    void *get(size_t bytes)
    {
        if (bytes > size) {
            if (p) (void)hipFree(p);
            p = nullptr;
            size = 0;
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            size = bytes;
        }
        return p;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        size = 0;
    }
};
struct StreamRes;
static StreamRes *stream_res(hipStream_t s);

static void *internal_ws(hipStream_t s, size_t bytes);

// iteration limit of the secular solver (paco_set_secular_maxiter)
static int g_secular_maxit = 200;
// columns per D&C leaf (paco_set_leaf_size; at first use the autotuner's table, else
// PACO_LEAF_DEFAULT); GenerateTiling describes this partition
static std::atomic<int> g_leaf_size{PACO_LEAF_DEFAULT};
static int leaf_size_now()
{
    static std::once_flag once;
    std::call_once(once, [] { g_leaf_size = tuned_leaf_size(PACO_LEAF_DEFAULT); });
    return g_leaf_size.load();
}

static bool use_grid_bidiag(int M, int N)
{
    const char *env = getenv("PACO_BIDIAG");
    if (env && !strcmp(env, "wg") && N <= BIDIAG_WG_MAXN) return false;
    if (env && !strcmp(env, "grid")) return true;
    return large_matrix(M, N);
}

static ProbDesc bidiag_desc(int M, int N, float *a, float *u, float *vt, float *d, float *e, float *tauq,
                            float *taup, char *ws, const WsLayout &L)
{
    ProbDesc P;
    memset(&P, 0, sizeof(P));
    P.M = M;
    P.N = N;
    P.a = a;
    P.lsv = a;
    P.u = u;
    P.vt = vt;
    P.d = d;
    P.e = e;
    P.tauq = tauq;
    P.taup = taup;
    P.vmt = reinterpret_cast<float *>(ws + L.vmt);
    P.tq = reinterpret_cast<float *>(ws + L.tq);
    P.tp = reinterpret_cast<float *>(ws + L.tp);
    P.info = reinterpret_cast<int *>(ws + L.info);
    return P;
}

// Phase events (batched path outside graph capture, and the two entry points).
static const int kPhases = 5;
static bool g_timing = false;
static hipEvent_t g_tev[kPhases + 1] = {};
static bool g_tev_recorded = false;

// start and end of the D&C when it runs on its own stream beside Q and P^T (batched path)
static hipEvent_t g_tev_dc[2] = {};
static bool g_tev_dc_on = false;

static void tmark(hipStream_t s, int i, bool on)
{
    if (on && i == 0) g_tev_dc_on = false;
    if (on && g_tev[i]) (void)hipEventRecord(g_tev[i], s);
}
static void tmark_dc(hipStream_t s, int i, bool on)
{
    if (on && g_tev_dc[i]) {
        (void)hipEventRecord(g_tev_dc[i], s);
        g_tev_dc_on = true;
    }
}

// Launch shape of the one-workgroup bidiagonalization for an M x N matrix: matrices
// that fit LDS whole run with 256 threads and exactly their LDS; larger ones with the
// larger workgroup and the whole LDS budget (bottom rows cached).  The tuned table
// (autotune.cpp) overrides the defaults per size bucket.
struct BidiagChoice {
    int nt;
    size_t lds;
    bool resident;
};
// Streamed matrices at 1024 threads take the blocked kernel (bidiag_blk.hip) when
// PACO_BIDIAG_BLOCKED=1.  It writes the trailing matrix once per panel instead of once per
// column, but issues more vector instructions than the one-pass-per-column kernel, which
// is bound by instruction issue rather than memory traffic, so the latter stays the default.
static bool blocked_bidiag(int M, int N, int nt)
{
    const char *env = getenv("PACO_BIDIAG_BLOCKED");
    return env && env[0] == '1' && nt == 1024 && bidiag_blk_supported(M, N);
}

static BidiagChoice bidiag_choice(int M, int N)
{
    BidiagChoice c;
    size_t lim = bidiag_wg_lds_limit();
    const char *env = getenv("PACO_LDS_RESIDENT_MAX");
    const size_t res_lim = env ? std::min(lim, (size_t)strtoull(env, nullptr, 10)) : lim;
    const size_t lr = bidiag_wg_lds_resident_bytes(M, N, 256);
    if (lr <= res_lim) {
        c.nt = 256;
        c.lds = lr;
        c.resident = true;
        return c;
    }
    c.nt = tuned_bidiag_threads(M, N);
    // PACO_BIDIAG_LDS caps the LDS of the streamed variant (tests of the
    // tier-independence of the results)
    const char *envl = getenv("PACO_BIDIAG_LDS");
    c.lds = envl ? std::min(lim, (size_t)strtoull(envl, nullptr, 10)) : lim;
    if (c.lds < bidiag_wg_lds_bytes(M, N, c.nt)) c.lds = bidiag_wg_lds_bytes(M, N, c.nt);
    c.resident = false;
    return c;
}

// Whether the kernels hold an M x N bidiagonalization: the one-workgroup kernel keeps
// cbuf[M] and its row vectors in LDS (bidiag_wg_lds_bytes), also for the last columns of
// the chip-wide path.  Checked before anything is launched.
static bool bidiag_fits(int M, int N)
{
    if (use_grid_bidiag(M, N)) return bidiag_grid_supported(M, N);
    return bidiag_choice(M, N).lds <= bidiag_wg_lds_limit();
}

// Q and P^T from the reflectors (GetUVt, upper_bidiagonalization.cpp:774-893), blocked WY;
// idq / idp count groups of id_rows rows, tpair groups of tp_per tile pairs.
static hipError_t form_q_pt(ProbList pl, const int *idq, int nidq, const int *idp, int nidp, const int *bq, int nbq,
                            const int *bp, int nbp, const int *sq, int nsq, const int *sp, int nsp, const int *tpair,
                            int ntp, hipStream_t s, int sw, int id_rows, int tp_per)
{
    hipError_t e;
    if ((e = launch_identity(pl, idq, nidq, 0, s, id_rows)) != hipSuccess) return e;
    if ((e = launch_identity(pl, idp, nidp, 1, s, id_rows)) != hipSuccess) return e;
    if ((e = launch_tfactor(pl, bq, nbq, 0, s)) != hipSuccess) return e;
    if ((e = launch_tfactor(pl, bp, nbp, 1, s)) != hipSuccess) return e;
    if ((e = launch_wy_apply(pl, sq, nsq, 0, s, sw)) != hipSuccess) return e;
    if ((e = launch_wy_apply(pl, sp, nsp, 1, s, sw)) != hipSuccess) return e;
    return launch_transpose_vt(pl, tpair, ntp, s, tp_per);
}

// vonly: singular values only (Q and P^T are not formed; u, vt are not touched)
// grid: tile grid of the chip-wide kernel (0: default)
// no_wait: the chip-wide path without kernels whose workgroups wait on each other
static int bidiag_single(int M, int N, float *a, float *u, float *vt, float *d, float *e, float *tauq, float *taup,
                         void *workspace, hipStream_t s, bool vonly = false, int grid = 0, bool no_wait = false)
{
    const WsLayout L = ws_layout(M, N);
    char *ws = static_cast<char *>(workspace ? workspace : internal_ws(s, L.total));
    if (!ws) return PACO_ENOMEM;
    ProbList pl;
    pl.ptr = nullptr;
    pl.count = 1;
    pl.single = bidiag_desc(M, N, a, u, vt, d, e, tauq, taup, ws, L);
    hipError_t err;
    // entry point (a) starts the problem's info word; (b) adds to it
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, pl.single.info, 1);
    const bool tm = g_timing;
    if (tm) {
        g_tev_recorded = true;
        tmark(s, 0, tm);
        tmark(s, 1, tm);
    }
    if (use_grid_bidiag(M, N)) {
        err = launch_bidiag_grid(pl.single, ws + L.grid, s, grid, no_wait);
    } else {
        const BidiagChoice c = bidiag_choice(M, N);
        if (!c.resident && blocked_bidiag(M, N, c.nt))
            err = launch_bidiag_blk(pl, nullptr, 1, std::max(c.lds, bidiag_blk_lds_bytes(M, N)), s);
        else
            err = launch_bidiag_wg(pl, nullptr, 1, c.lds, s, c.nt);
    }
    if (err != hipSuccess) return (int)err;
    tmark(s, 2, tm);
    if (vonly) {
        tmark(s, 3, tm);
        return 0;
    }
    const int nt = cdiv(N, 32);
    if (large_matrix(M, N)) {
        // blocks of 32 reflectors applied to the whole trailing part of Q, P as GEMMs
        if ((err = launch_identity(pl, nullptr, cdiv(M, ROW_GROUP), 0, s, ROW_GROUP)) != hipSuccess) return (int)err;
        if ((err = launch_identity(pl, nullptr, cdiv(N, ROW_GROUP), 1, s, ROW_GROUP)) != hipSuccess) return (int)err;
        if ((err = launch_tfactor(pl, nullptr, cdiv(N, 32), 0, s)) != hipSuccess) return (int)err;
        if ((err = launch_tfactor(pl, nullptr, cdiv(std::max(N - 1, 0), 32), 1, s)) != hipSuccess) return (int)err;
        if ((err = launch_qp_gemm(pl.single, 0, ws + L.qp, s)) != hipSuccess) return (int)err;
        if ((err = launch_qp_gemm(pl.single, 1, ws + L.qp, s)) != hipSuccess) return (int)err;
        err = launch_transpose_vt(pl, nullptr, nt * (nt + 1) / 2, s, 1);
        tmark(s, 3, tm);
        return (int)err;
    }
    // a single matrix: 16-column strips unless 64-column ones already give every CU a strip
    const int sw = cdiv(M, 64) >= 512 ? 64 : 16;
    err = form_q_pt(pl, nullptr, cdiv(M, ROW_GROUP), nullptr, cdiv(N, ROW_GROUP), nullptr, cdiv(N, 32), nullptr, cdiv(std::max(N - 1, 0), 32), nullptr,
                    cdiv(M, sw), nullptr, cdiv(N, sw), nullptr, nt * (nt + 1) / 2, s, sw, ROW_GROUP, 1);
    tmark(s, 3, tm);
    return (int)err;
}

// D&C levels + back-transform for a problem list (single: offsets null).
struct DcOffsets {
    const int *leaf_off = nullptr;
    int nleaf = 0;
    int ls = LEAF_MIN; // columns per leaf
    std::vector<const int *> task_off; // per round
    std::vector<const int *> task_prob; // per round (empty: search task_off)
    std::vector<int> ntask, maxn;
    std::vector<MergeState *> states;
    const int *bt_tu = nullptr, *bt_tv = nullptr, *bt_ru = nullptr, *bt_rv = nullptr;
    bool vonly = false; // singular values only (no left vectors, two right-vector columns per node)
    int ntu = 0, ntv = 0, nru = 0, nrv = 0;
    int bt_rows = ROW_GROUP; // rows per group of the row offsets
};

static hipError_t run_dc_levels(ProbList pl, const DcOffsets &o, hipStream_t s)
{
    hipError_t e = launch_dc_leaves(pl, o.leaf_off, o.nleaf, o.ls, s);
    if (e != hipSuccess) return e;
    for (size_t r = 1; r <= o.ntask.size(); ++r) {
        LevelArgs la;
        la.task_off = o.task_off[r - 1];
        la.task_prob = o.task_prob.empty() ? nullptr : o.task_prob[r - 1];
        la.states = o.states[r - 1];
        la.ntask = o.ntask[r - 1];
        la.r = (int)r;
        // every merge of the round has at most min(ls 2^r, N) rows and columns: the
        // launches give each merge that many threads / tiles
        la.maxn = o.maxn[r - 1];
        la.stride = std::max(1, std::min(o.ls << r, la.maxn));
        la.ls = o.ls;
        la.maxit = g_secular_maxit;
        la.vonly = o.vonly ? 1 : 0;
        if ((e = launch_dc_level(pl, la, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

static hipError_t run_dc_backtransform(ProbList pl, const DcOffsets &o, hipStream_t s)
{
    return launch_dc_backtransform(pl, o.bt_tu, o.ntu, o.bt_tv, o.ntv, o.bt_ru, o.nru, o.bt_rv, o.nrv, o.bt_rows, s);
}

static hipError_t run_dc(ProbList pl, const DcOffsets &o, hipStream_t s)
{
    hipError_t e = run_dc_levels(pl, o, s);
    if (e != hipSuccess) return e;
    return run_dc_backtransform(pl, o, s);
}

static int dc_single(int M, int N, float *a, float *u, float *vt, float *d, float *e, float *qt, float *wt,
                     uint32_t *idx, void *workspace, const void *tiling_dev, hipStream_t s, bool vonly = false)
{
    const WsLayout L = ws_layout(M, N);
    char *ws = static_cast<char *>(workspace ? workspace : internal_ws(s, L.total));
    if (!ws) return PACO_ENOMEM;
    ProbList pl;
    pl.ptr = nullptr;
    pl.count = 1;
    ProbDesc &P = pl.single;
    P = bidiag_desc(M, N, a, u, vt, d, e, nullptr, nullptr, ws, L);
    P.qt = qt;
    P.wt = wt;
    P.idx = idx;
    const int ls = leaf_size_now();
    hipLaunchKernelGGL(dc_guard_kernel, dim3(1), dim3(256), 0, s, pl, static_cast<const uint32_t *>(tiling_dev), ls);
    DcOffsets o;
    o.vonly = vonly;
    o.ls = ls;
    o.nleaf = leaf_count(N, ls);
    MergeState *st = reinterpret_cast<MergeState *>(ws + L.states);
    const int R = rounds_of(N, ls);
    for (int r = 1; r <= R; ++r) {
        o.task_off.push_back(nullptr);
        o.ntask.push_back(merges_in_round(N, r, ls));
        o.maxn.push_back(std::min(ls << r, N));
        o.states.push_back(st);
        st += merges_in_round(N, r, ls);
    }
    o.ntu = cdiv(M, 64) * cdiv(N, 64);
    o.ntv = cdiv(N, 64) * cdiv(N, 64);
    o.nru = cdiv(M, ROW_GROUP);
    o.nrv = cdiv(N, ROW_GROUP);
    const bool tm = g_timing;
    if (tm) {
        g_tev_recorded = true;
        tmark(s, 3, tm);
    }
    hipError_t err = run_dc_levels(pl, o, s);
    if (err != hipSuccess) return (int)err;
    tmark(s, 4, tm);
    if (!vonly) err = run_dc_backtransform(pl, o, s);
    tmark(s, 5, tm);
    return (int)err;
}

// ---------------------------------------------------------------------------
// batched path
// ---------------------------------------------------------------------------
struct BatchScratch {
    size_t a, e, tauq, taup, qt, wt, idx, vmt, lsv, tq, tp, states, qp, total;
};

static BatchScratch scratch_layout(int M, int N)
{
    BatchScratch B;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t r = o;
        o = align_up(o + bytes, 256);
        return r;
    };
    B.a = take(sizeof(float) * (size_t)M * N);
    B.e = take(sizeof(float) * (size_t)std::max(N - 1, 1));
    B.tauq = take(sizeof(float) * (size_t)N);
    B.taup = take(sizeof(float) * (size_t)std::max(N - 1, 1));
    B.qt = take(sizeof(float) * 2 * (size_t)N * N);
    B.wt = take(sizeof(float) * 2 * (size_t)N * N);
    B.idx = take(sizeof(uint32_t) * 16 * (size_t)N);
    B.vmt = take(sizeof(float) * (size_t)N * N);
    B.lsv = take(sizeof(float) * (size_t)N * N);
    B.tq = take(sizeof(float) * 1024 * (size_t)cdiv(N, 32));
    B.tp = take(sizeof(float) * 1024 * (size_t)cdiv(std::max(N - 1, 1), 32));
    B.states = take(sizeof(MergeState) * (size_t)std::max(total_merges(N), 1));
    B.qp = take(sizeof(float) * qpb_scratch_floats(M, N));
    B.total = o;
    return B;
}

// LDS classes (0..3: <= 1/8, 1/4, 1/2, all of the LDS) whose resident matrices take the
// one-wavefront kernel by default
static constexpr int kWaveMaxClassDefault = -1;
// workgroup size of the LDS-resident launches per LDS class: the two small classes hold up
// to 8 and 4 matrices per CU's LDS, and at 512 threads two of them share a CU, so the
// per-column section that every wavefront repeats runs on 8 wavefronts per matrix, not 16
static constexpr int kResThreadsDefault[4] = {512, 512, 1024, 1024};

// Host-side plan of a batch: a byte blob (uploaded once) + launch sizes.
struct BatchPlan {
    std::vector<uint8_t> blob;
    size_t scratch_bytes = 0;
    int count = 0;
    // offsets into blob (bytes)
    size_t o_probs = 0, o_src = 0, o_info = 0;
    // streamed bidiagonalization: one launch per workgroup size (256, 512, 1024)
    size_t o_order_nt[3] = {0, 0, 0}, lds_nt[3] = {0, 0, 0};
    int n_nt[3] = {0, 0, 0};
    // streamed matrices of the blocked kernel (bidiag_blk.hip); LDS beyond its vectors
    // caches bottom rows
    size_t o_order_blk = 0, lds_blk = 0;
    int n_blk = 0;
    // one wavefront per LDS-resident matrix: one launch per LDS class (the dynamic LDS
    // of a launch is its largest member's, which sets the occupancy of all)
    // and per chunk count C = ceil(N / 64) (the kernel is specialised on it)
    static constexpr int kWaveClasses = 4, kWaveGroups = kWaveClasses * BIDIAG_WAVE_MAXC;
    size_t o_order_wave[kWaveGroups] = {}, lds_wave[kWaveGroups] = {};
    int n_wave[kWaveGroups] = {};
    // one workgroup per LDS-resident matrix not taken by the wave kernel: one launch per
    // LDS class
    size_t o_order_res[kWaveClasses] = {}, lds_res[kWaveClasses] = {};
    int n_res[kWaveClasses] = {};
    int res_threads[kWaveClasses] = {1024, 1024, 1024, 1024};
    size_t o_rowin = 0, o_idq = 0, o_idp = 0, o_bq = 0, o_bp = 0, o_sq = 0, o_sp = 0, o_tp = 0;
    size_t o_leaf = 0, o_tu = 0, o_tv = 0, o_ru = 0, o_rv = 0;
    std::vector<size_t> o_task, o_tprob;
    int nrowin = 0, nidq = 0, nidp = 0, nbq = 0, nbp = 0, nsq = 0, nsp = 0, ntp = 0, nleaf = 0;
    int ntu = 0, ntv = 0, nru = 0, nrv = 0;
    std::vector<int> ntask, maxn;
    std::vector<size_t> state_off; // per round, in MergeState units from the states base
    // Q, P by rounds of grouped GEMMs: per side and round, offsets of the extraction
    // blocks, W tiles and X tiles over the problems; per-problem scratch pointers at o_qp
    size_t o_qp = 0;
    std::vector<size_t> o_qext[2], o_qg1[2], o_qg3[2];
    std::vector<int> n_qext[2], n_qg1[2], n_qg3[2];
    bool fits = true;
    bool vonly = false; // singular values only: no Q, P^T, back-transforms or left vectors
    int ls = LEAF_MIN;  // columns per D&C leaf
};

template <class T>
static size_t blob_put(std::vector<uint8_t> &b, const std::vector<T> &v)
{
    size_t o = align_up(b.size(), 256);
    b.resize(o + sizeof(T) * v.size());
    if (!v.empty()) memcpy(b.data() + o, v.data(), sizeof(T) * v.size());
    return o;
}

// Prefix of per-problem counts: returns offsets (count+1 entries) and the total.
template <class F>
static std::vector<int> prefix(int count, F f, int &total)
{
    std::vector<int> off(count + 1);
    int acc = 0;
    for (int i = 0; i < count; ++i) {
        off[i] = acc;
        acc += f(i);
    }
    off[count] = acc;
    total = acc;
    return off;
}

// prefix() followed by the problem of every task (read by task_of)
template <class F>
static std::vector<int> prefix_table(int count, F f, int &total)
{
    std::vector<int> v = prefix(count, f, total);
    v.resize((size_t)count + 1 + total);
    for (int i = 0; i < count; ++i)
        for (int q = v[i]; q < v[i + 1]; ++q) v[(size_t)count + 1 + q] = i;
    return v;
}

static double svd_cost(int M, int N) { return 4.0 * M * M * N + 2.0 * M * N * N + 6.0 * (double)N * N * N; }

static void build_plan(const paco_svd_job *jobs, int count, BatchPlan &bp, std::vector<size_t> &scr_off)
{
    bp.ls = leaf_size_now();
    bp.count = count;
    scr_off.resize(count);
    size_t so = 0;
    int maxN = 0;
    for (int i = 0; i < count; ++i) {
        scr_off[i] = so;
        so += scratch_layout(jobs[i].M, jobs[i].N).total;
        maxN = std::max(maxN, jobs[i].N);
    }
    bp.scratch_bytes = so;
    bp.fits = maxN <= PACO_MAX_N;
    auto &b = bp.blob;
    b.clear();
    b.resize(0);
    // problem descriptors are patched with device addresses at upload time
    std::vector<ProbDesc> probs(count);
    bp.o_probs = blob_put(b, probs);
    std::vector<const float *> src(count);
    for (int i = 0; i < count; ++i) src[i] = jobs[i].A;
    bp.o_src = blob_put(b, src);
    // bidiagonalization variant per matrix: LDS-resident when the matrix fits, else
    // streamed; each list in longest-first order
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return svd_cost(jobs[x].M, jobs[x].N) > svd_cost(jobs[y].M, jobs[y].N);
    });
    // a matrix's launch shape (and so its arithmetic) depends on its shape only
    std::vector<int> og[3], ob, orr[BatchPlan::kWaveClasses], ow[BatchPlan::kWaveGroups];
    // LDS class of a resident matrix: 0..3 for <= 1/8, 1/4, 1/2, all of the limit
    auto lds_class = [](size_t bytes, size_t lim) {
        int cl = 0;
        while (cl < BatchPlan::kWaveClasses - 1 && bytes > (lim >> (BatchPlan::kWaveClasses - 1 - cl))) ++cl;
        return cl;
    };
    // PACO_BIDIAG_WAVE=0: no one-wavefront kernel; PACO_WAVE_MAX_CLASS: the LDS classes
    // (0..3) it takes (larger resident matrices run one workgroup each, whose several
    // wavefronts share the CU the matrix's LDS occupies)
    const char *wenv = getenv("PACO_BIDIAG_WAVE");
    const char *cenv = getenv("PACO_WAVE_MAX_CLASS");
    const int wave_max_class = (wenv && wenv[0] == '0') ? -1 : cenv ? atoi(cenv) : kWaveMaxClassDefault;
    // PACO_RES_THREADS: workgroup size of the resident launches, one value for all LDS
    // classes or four comma-separated ones (classes <= 1/8, 1/4, 1/2, all of the LDS);
    // 512,512,1024,1024 measured best (bidiagonalization 59.3 ms against 60.1 at 1024 for
    // all classes; one value for all: 62.1 at 1024, 63.7 at 512, 66.7 at 256)
    for (int cl = 0; cl < BatchPlan::kWaveClasses; ++cl) bp.res_threads[cl] = tuned_res_threads(cl, kResThreadsDefault[cl]);
    if (const char *tenv = getenv("PACO_RES_THREADS")) {
        const char *p = tenv;
        for (int cl = 0; cl < BatchPlan::kWaveClasses; ++cl) {
            const int v = atoi(p);
            bp.res_threads[cl] = (v == 256 || v == 1024) ? v : 512;
            const char *comma = strchr(p, ',');
            if (comma) p = comma + 1; // a single value applies to the remaining classes
        }
    }
    const size_t wave_lim = bidiag_wave_lds_limit();
    for (int i : order) {
        const size_t wl = bidiag_wave_lds(jobs[i].M, jobs[i].N);
        if (jobs[i].N <= 64 * BIDIAG_WAVE_MAXC && wl <= wave_lim && lds_class(wl, wave_lim) <= wave_max_class) {
            const int g = lds_class(wl, wave_lim) * BIDIAG_WAVE_MAXC + (cdiv(jobs[i].N, 64) - 1);
            ow[g].push_back(i);
            bp.lds_wave[g] = std::max(bp.lds_wave[g], wl);
            continue;
        }
        const BidiagChoice c = bidiag_choice(jobs[i].M, jobs[i].N);
        if (c.resident) {
            // class from the 256-thread footprint, then that class's workgroup size
            const int cl = lds_class(c.lds, bidiag_wg_lds_limit()), nt = bp.res_threads[cl];
            const size_t lr = bidiag_wg_lds_resident_bytes(jobs[i].M, jobs[i].N, nt);
            if (lr <= bidiag_wg_lds_limit()) {
                orr[cl].push_back(i);
                bp.lds_res[cl] = std::max(bp.lds_res[cl], lr);
                continue;
            }
            // too large for LDS at the wider workgroup: the streamed launch of that size
            const int g = nt >= 1024 ? 2 : nt >= 512 ? 1 : 0;
            og[g].push_back(i);
            bp.lds_nt[g] = std::max(bp.lds_nt[g], bidiag_wg_lds_limit());
        } else if (blocked_bidiag(jobs[i].M, jobs[i].N, c.nt)) {
            ob.push_back(i);
            bp.lds_blk = std::max(bp.lds_blk, std::max(c.lds, bidiag_blk_lds_bytes(jobs[i].M, jobs[i].N)));
        } else {
            const int g = c.nt >= 1024 ? 2 : c.nt >= 512 ? 1 : 0;
            og[g].push_back(i);
            bp.lds_nt[g] = std::max(bp.lds_nt[g], c.lds);
        }
    }
    for (int g = 0; g < 3; ++g) {
        bp.n_nt[g] = (int)og[g].size();
        bp.o_order_nt[g] = blob_put(b, og[g]);
    }
    bp.n_blk = (int)ob.size();
    bp.o_order_blk = blob_put(b, ob);
    for (int cl = 0; cl < BatchPlan::kWaveClasses; ++cl) {
        bp.n_res[cl] = (int)orr[cl].size();
        bp.o_order_res[cl] = blob_put(b, orr[cl]);
    }
    for (int g = 0; g < BatchPlan::kWaveGroups; ++g) {
        bp.n_wave[g] = (int)ow[g].size();
        bp.o_order_wave[g] = blob_put(b, ow[g]);
    }
    bp.o_rowin = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].M, BATCH_ROW_GROUP); }, bp.nrowin));
    bp.o_idq = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].M, BATCH_ROW_GROUP); }, bp.nidq));
    bp.o_idp = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].N, BATCH_ROW_GROUP); }, bp.nidp));
    {
        std::vector<float *> qpp(count, nullptr); // device addresses patched at upload
        bp.o_qp = blob_put(b, qpp);
        for (int side = 0; side < 2; ++side) {
            int rounds = 0;
            for (int i = 0; i < count; ++i)
                rounds = std::max(rounds, cdiv(side == 0 ? jobs[i].N : std::max(jobs[i].N - 1, 0), 32));
            bp.o_qext[side].assign(rounds, 0);
            bp.o_qg1[side].assign(rounds, 0);
            bp.o_qg3[side].assign(rounds, 0);
            bp.n_qext[side].assign(rounds, 0);
            bp.n_qg1[side].assign(rounds, 0);
            bp.n_qg3[side].assign(rounds, 0);
            for (int r = 0; r < rounds; ++r) {
                auto lr = [&](int i) {
                    const int L = side == 0 ? jobs[i].M : jobs[i].N;
                    const int nb = cdiv(side == 0 ? jobs[i].N : std::max(jobs[i].N - 1, 0), 32);
                    const int b = nb - 1 - r;
                    return b < 0 ? 0 : L - (32 * b + side);
                };
                bp.o_qext[side][r] = blob_put(b, prefix(count, [&](int i) { return cdiv(lr(i), 8); }, bp.n_qext[side][r]));
                bp.o_qg1[side][r] = blob_put(b, prefix(count, [&](int i) { return cdiv(lr(i), 64); }, bp.n_qg1[side][r]));
                bp.o_qg3[side][r] =
                    blob_put(b, prefix(count, [&](int i) { return cdiv(lr(i), 64) * cdiv(lr(i), 64); }, bp.n_qg3[side][r]));
            }
        }
    }
    // with the block -> problem table that tfactor_kernel reads
    bp.o_bq = blob_put(b, prefix_table(count, [&](int i) { return cdiv(jobs[i].N, 32); }, bp.nbq));
    bp.o_bp = blob_put(b, prefix_table(count, [&](int i) { return cdiv(std::max(jobs[i].N - 1, 0), 32); }, bp.nbp));
    bp.o_sq = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].M, 64); }, bp.nsq));
    bp.o_sp = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].N, 64); }, bp.nsp));
    bp.o_tp = blob_put(b, prefix(count, [&](int i) {
        const int nt = cdiv(jobs[i].N, 32);
        return cdiv(nt * (nt + 1) / 2, TP_GROUP);
    }, bp.ntp));
    const int ls = bp.ls;
    bp.o_leaf = blob_put(b, prefix(count, [&](int i) { return leaf_count(jobs[i].N, ls); }, bp.nleaf));
    int R = 0;
    for (int i = 0; i < count; ++i) R = std::max(R, rounds_of(jobs[i].N, ls));
    bp.o_task.clear();
    bp.o_tprob.clear();
    bp.ntask.clear();
    bp.maxn.clear();
    bp.state_off.clear();
    size_t sacc = 0;
    for (int r = 1; r <= R; ++r) {
        int nt = 0;
        bp.o_task.push_back(blob_put(b, prefix(count, [&](int i) { return merges_in_round(jobs[i].N, r, ls); }, nt)));
        std::vector<int> tprob;
        tprob.reserve(nt);
        for (int i = 0; i < count; ++i)
            for (int q = merges_in_round(jobs[i].N, r, ls); q > 0; --q) tprob.push_back(i);
        bp.o_tprob.push_back(blob_put(b, tprob));
        bp.ntask.push_back(nt);
        int mx = 0;
        for (int i = 0; i < count; ++i)
            if (merges_in_round(jobs[i].N, r, ls) > 0) mx = std::max(mx, std::min(ls << r, jobs[i].N));
        bp.maxn.push_back(mx);
        bp.state_off.push_back(sacc);
        sacc += nt;
    }
    bp.o_tu = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].M, 64) * cdiv(jobs[i].N, 64); }, bp.ntu));
    bp.o_tv = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].N, 64) * cdiv(jobs[i].N, 64); }, bp.ntv));
    bp.o_ru = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].M, BATCH_ROW_GROUP); }, bp.nru));
    bp.o_rv = blob_put(b, prefix(count, [&](int i) { return cdiv(jobs[i].N, BATCH_ROW_GROUP); }, bp.nrv));
    std::vector<MergeState> states(std::max<size_t>(sacc, 1));
    bp.o_info = blob_put(b, std::vector<int>(count, 0));
    // merge states live after the blob in device memory (not uploaded)
    b.resize(align_up(b.size(), 256));
}

// Device-side resources of a batch: plan buffer + scratch + states.
struct BatchRes {
    void *dev = nullptr;
    size_t dev_size = 0;
    void *pinned = nullptr;
    size_t pinned_size = 0;
    hipEvent_t ev = nullptr;
    hipStream_t side = nullptr;      // forked stream for the concurrent bidiagonalization
    hipEvent_t evf = nullptr, evj = nullptr;
    hipStream_t dcs = nullptr;       // forked stream of the D&C (concurrent with Q and P^T)
    hipEvent_t evdf = nullptr, evdj = nullptr;
};

// Frees everything a BatchRes holds (device and pinned memory, the forked stream, events).
static void batch_res_destroy(BatchRes &r)
{
    if (r.dev) (void)hipFree(r.dev);
    if (r.pinned) (void)hipHostFree(r.pinned);
    if (r.ev) (void)hipEventDestroy(r.ev);
    if (r.evf) (void)hipEventDestroy(r.evf);
    if (r.evj) (void)hipEventDestroy(r.evj);
    if (r.side) (void)hipStreamDestroy(r.side);
    if (r.evdf) (void)hipEventDestroy(r.evdf);
    if (r.evdj) (void)hipEventDestroy(r.evdj);
    if (r.dcs) (void)hipStreamDestroy(r.dcs);
    r = BatchRes();
}

static hipError_t ensure(BatchRes &r, size_t dev_bytes, size_t pin_bytes)
{
    if (dev_bytes > r.dev_size) {
        if (r.dev) (void)hipFree(r.dev);
        r.dev = nullptr;
        r.dev_size = 0;
        hipError_t e = hipMalloc(&r.dev, dev_bytes);
        if (e != hipSuccess) return e;
        r.dev_size = dev_bytes;
    }
    if (pin_bytes > r.pinned_size) {
        if (r.pinned) (void)hipHostFree(r.pinned);
        r.pinned = nullptr;
        r.pinned_size = 0;
        hipError_t e = hipHostMalloc(&r.pinned, pin_bytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        r.pinned_size = pin_bytes;
    }
    hipError_t e = hipSuccess;
    if (!r.ev && (e = hipEventCreateWithFlags(&r.ev, hipEventDisableTiming)) != hipSuccess) return e;
    if (!r.side && (e = hipStreamCreateWithFlags(&r.side, hipStreamNonBlocking)) != hipSuccess) return e;
    if (!r.evf && (e = hipEventCreateWithFlags(&r.evf, hipEventDisableTiming)) != hipSuccess) return e;
    if (!r.evj && (e = hipEventCreateWithFlags(&r.evj, hipEventDisableTiming)) != hipSuccess) return e;
    if (!r.dcs && (e = hipStreamCreateWithFlags(&r.dcs, hipStreamNonBlocking)) != hipSuccess) return e;
    if (!r.evdf && (e = hipEventCreateWithFlags(&r.evdf, hipEventDisableTiming)) != hipSuccess) return e;
    if (!r.evdj && (e = hipEventCreateWithFlags(&r.evdj, hipEventDisableTiming)) != hipSuccess) return e;
    return hipSuccess;
}

// Prepared batch: device layout fixed, descriptors patched, blob staged in pinned memory.
struct Prepared {
    size_t o_states = 0;
    int *info = nullptr;   // per-job info inside the plan buffer (zeroed by the launch sequence)
    bool own_info = true;  // false: the jobs' info words live in a caller array (already zeroed)
};

// Lay out device memory [blob | states | scratch...], patch descriptors and stage the
// blob in pinned memory.  Allocates (outside any stream capture).  With info_ext, job i
// reports into info_ext[info_idx[i]] instead of the plan's own array.
static hipError_t prepare_batch(const paco_svd_job *jobs, int count, BatchPlan &bp, BatchRes &res, Prepared &pr,
                                int *info_ext = nullptr, const int *info_idx = nullptr, bool vonly = false)
{
    std::vector<size_t> scr_off;
    build_plan(jobs, count, bp, scr_off);
    bp.vonly = vonly;
    if (!bp.fits) return hipErrorInvalidValue;
    size_t nstates = 0;
    for (int v : bp.ntask) nstates += v;
    pr.o_states = align_up(bp.blob.size(), 256);
    const size_t o_scratch = align_up(pr.o_states + sizeof(MergeState) * std::max<size_t>(nstates, 1), 256);
    const size_t total = o_scratch + bp.scratch_bytes;
    // the previous upload from the pinned buffer must be complete before overwriting it
    if (res.ev) (void)hipEventSynchronize(res.ev);
    hipError_t e = ensure(res, total, bp.blob.size());
    if (e != hipSuccess) return e;
    char *dev = static_cast<char *>(res.dev);
    ProbDesc *probs = reinterpret_cast<ProbDesc *>(bp.blob.data() + bp.o_probs);
    pr.info = reinterpret_cast<int *>(dev + bp.o_info);
    for (int i = 0; i < count; ++i) {
        const BatchScratch L = scratch_layout(jobs[i].M, jobs[i].N);
        char *base = dev + o_scratch + scr_off[i];
        ProbDesc &P = probs[i];
        memset(&P, 0, sizeof(P));
        P.M = jobs[i].M;
        P.N = jobs[i].N;
        P.a = reinterpret_cast<float *>(base + L.a);
        P.u = jobs[i].U;
        P.vt = jobs[i].Vt;
        P.d = jobs[i].S;
        P.e = reinterpret_cast<float *>(base + L.e);
        P.tauq = reinterpret_cast<float *>(base + L.tauq);
        P.taup = reinterpret_cast<float *>(base + L.taup);
        P.qt = reinterpret_cast<float *>(base + L.qt);
        P.wt = reinterpret_cast<float *>(base + L.wt);
        P.idx = reinterpret_cast<uint32_t *>(base + L.idx);
        P.vmt = reinterpret_cast<float *>(base + L.vmt);
        P.lsv = reinterpret_cast<float *>(base + L.lsv);
        P.tq = reinterpret_cast<float *>(base + L.tq);
        P.tp = reinterpret_cast<float *>(base + L.tp);
        P.info = info_ext ? info_ext + info_idx[i] : pr.info + i;
        reinterpret_cast<float **>(bp.blob.data() + bp.o_qp)[i] = reinterpret_cast<float *>(base + L.qp);
    }
    if (info_ext) pr.own_info = false;
    memcpy(res.pinned, bp.blob.data(), bp.blob.size());
    return hipSuccess;
}

// Upload + the whole launch sequence (capturable: no allocation, no host sync).
static hipError_t enqueue_batch(const BatchPlan &bp, BatchRes &res, const Prepared &pr, hipStream_t s, bool record)
{
    const bool tm = record && g_timing;
    if (tm) g_tev_recorded = true;
    const int count = bp.count;
    char *dev = static_cast<char *>(res.dev);
    hipError_t e;
    if ((e = hipMemcpyAsync(dev, res.pinned, bp.blob.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if (record && (e = hipEventRecord(res.ev, s)) != hipSuccess) return e;
    auto I = [&](size_t o) { return reinterpret_cast<const int *>(dev + o); };
    ProbList pl;
    pl.ptr = reinterpret_cast<const ProbDesc *>(dev + bp.o_probs);
    pl.count = count;
    memset(&pl.single, 0, sizeof(pl.single));
    tmark(s, 0, tm);
    if (pr.own_info) hipLaunchKernelGGL(zero_ints_kernel, dim3(cdiv(count, 256)), dim3(256), 0, s, pr.info, count);
    hipLaunchKernelGGL(copy_in_kernel, dim3(bp.nrowin), dim3(256), 0, s, pl,
                       reinterpret_cast<const float *const *>(dev + bp.o_src), I(bp.o_rowin));
    tmark(s, 1, tm);
    // the LDS-resident matrices (one wavefront each, or one workgroup) run on a forked
    // stream, concurrently with the streamed ones
    int n_side = 0;
    for (int cl = 0; cl < BatchPlan::kWaveClasses; ++cl) n_side += bp.n_res[cl];
    for (int g = 0; g < BatchPlan::kWaveGroups; ++g) n_side += bp.n_wave[g];
    // PACO_BIDIAG_SERIAL=1: every bidiagonalization launch on the caller's stream (schedule
    // comparisons)
    const char *serial_env = getenv("PACO_BIDIAG_SERIAL");
    const bool serial = serial_env && serial_env[0] == '1';
    hipStream_t side = serial ? s : res.side;
    if (n_side > 0) {
        if (!serial) {
            if ((e = hipEventRecord(res.evf, s)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(res.side, res.evf, 0)) != hipSuccess) return e;
        }
        for (int g = BatchPlan::kWaveGroups - 1; g >= 0; --g)
            if ((e = launch_bidiag_wave(pl, I(bp.o_order_wave[g]), bp.n_wave[g], g % BIDIAG_WAVE_MAXC + 1,
                                        bp.lds_wave[g], side)) != hipSuccess)
                return e;
        for (int cl = BatchPlan::kWaveClasses - 1; cl >= 0; --cl)
            if ((e = launch_bidiag_wg(pl, I(bp.o_order_res[cl]), bp.n_res[cl], bp.lds_res[cl], side,
                                      bp.res_threads[cl])) != hipSuccess)
                return e;
        if (!serial && (e = hipEventRecord(res.evj, res.side)) != hipSuccess) return e;
    }
    if ((e = launch_bidiag_blk(pl, I(bp.o_order_blk), bp.n_blk, bp.lds_blk, s)) != hipSuccess) return e;
    for (int g = 2; g >= 0; --g)
        if ((e = launch_bidiag_wg(pl, I(bp.o_order_nt[g]), bp.n_nt[g], bp.lds_nt[g], s, 256 << g)) != hipSuccess)
            return e;
    if (n_side > 0 && !serial && (e = hipStreamWaitEvent(s, res.evj, 0)) != hipSuccess) return e;
    tmark(s, 2, tm);
    DcOffsets o;
    o.vonly = bp.vonly;
    o.ls = bp.ls;
    o.leaf_off = I(bp.o_leaf);
    o.nleaf = bp.nleaf;
    MergeState *st = reinterpret_cast<MergeState *>(dev + pr.o_states);
    for (size_t r = 0; r < bp.ntask.size(); ++r) {
        o.task_off.push_back(I(bp.o_task[r]));
        o.task_prob.push_back(I(bp.o_tprob[r]));
        o.ntask.push_back(bp.ntask[r]);
        o.maxn.push_back(bp.maxn[r]);
        o.states.push_back(st + bp.state_off[r]);
    }
    o.bt_tu = I(bp.o_tu);
    o.ntu = bp.ntu;
    o.bt_tv = I(bp.o_tv);
    o.ntv = bp.ntv;
    o.bt_ru = I(bp.o_ru);
    o.bt_rows = BATCH_ROW_GROUP;
    o.nru = bp.nru;
    o.bt_rv = I(bp.o_rv);
    o.nrv = bp.nrv;
    // The D&C reads d and e and writes only its own buffers (lsv, vmt, qt, wt, idx, the
    // merge states, d, info); the formation of Q and P^T reads a, tauq and taup and writes u,
    // vt and the T factors.  With full vectors the D&C runs on a forked stream beside it.
    // PACO_DC_SERIAL=1: the D&C after Q and P^T on the caller's stream (schedule comparisons).
    const char *dser = getenv("PACO_DC_SERIAL");
    const bool dc_fork = !bp.vonly && !(dser && dser[0] == '1');
    auto run_dc = [&](hipStream_t sd) -> hipError_t {
        hipLaunchKernelGGL(dc_guard_kernel, dim3(count), dim3(256), 0, sd, pl, nullptr, bp.ls);
        return run_dc_levels(pl, o, sd);
    };
    if (dc_fork) {
        if ((e = hipEventRecord(res.evdf, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(res.dcs, res.evdf, 0)) != hipSuccess) return e;
        tmark_dc(res.dcs, 0, tm);
        if ((e = run_dc(res.dcs)) != hipSuccess) return e;
        tmark_dc(res.dcs, 1, tm);
        if ((e = hipEventRecord(res.evdj, res.dcs)) != hipSuccess) return e;
    }
    const char *qenv = getenv("PACO_QP_BATCH");
    if (bp.vonly) {
        // singular values only: Q and P^T are not formed
    } else if (qenv && qenv[0] == '0') { // strips of 64 columns, FMA
        if ((e = form_q_pt(pl, I(bp.o_idq), bp.nidq, I(bp.o_idp), bp.nidp, I(bp.o_bq), bp.nbq, I(bp.o_bp), bp.nbp,
                           I(bp.o_sq), bp.nsq, I(bp.o_sp), bp.nsp, I(bp.o_tp), bp.ntp, s, 64, BATCH_ROW_GROUP,
                           TP_GROUP)) != hipSuccess)
            return e;
    } else { // rounds of grouped GEMMs on the matrix cores
        float *const *qp = reinterpret_cast<float *const *>(dev + bp.o_qp);
        // P's identity, T factors and rounds on the forked stream, concurrently with Q's
        // (PACO_QP_FORK=0: after them)
        const char *qf = getenv("PACO_QP_FORK");
        const bool qp_fork = !(qf && qf[0] == '0');
        if ((e = hipEventRecord(res.evf, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(res.side, res.evf, 0)) != hipSuccess) return e;
        for (int side = 1; side >= 0; --side) {
            hipStream_t st = side == 1 && qp_fork ? res.side : s;
            if ((e = launch_identity(pl, I(side == 0 ? bp.o_idq : bp.o_idp), side == 0 ? bp.nidq : bp.nidp, side, st,
                                     BATCH_ROW_GROUP)) != hipSuccess)
                return e;
            if ((e = launch_tfactor(pl, I(side == 0 ? bp.o_bq : bp.o_bp), side == 0 ? bp.nbq : bp.nbp, side, st)) !=
                hipSuccess)
                return e;
            for (size_t r = 0; r < bp.o_qext[side].size(); ++r)
                if ((e = launch_qpb_round(pl, qp, side, (int)r, I(bp.o_qext[side][r]), bp.n_qext[side][r],
                                          I(bp.o_qg1[side][r]), bp.n_qg1[side][r], I(bp.o_qg3[side][r]),
                                          bp.n_qg3[side][r], st)) != hipSuccess)
                    return e;
        }
        if ((e = hipEventRecord(res.evj, res.side)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s, res.evj, 0)) != hipSuccess) return e;
        if ((e = launch_transpose_vt(pl, I(bp.o_tp), bp.ntp, s, TP_GROUP)) != hipSuccess) return e;
    }
    tmark(s, 3, tm);
    if (dc_fork) {
        if ((e = hipStreamWaitEvent(s, res.evdj, 0)) != hipSuccess) return e;
    } else if ((e = run_dc(s)) != hipSuccess) {
        return e;
    }
    tmark(s, 4, tm);
    if (!bp.vonly) e = run_dc_backtransform(pl, o, s);
    tmark(s, 5, tm);
    return e;
}

// info words of the last batched or plain call on this thread (valid while t_last_gen ==
// g_release_gen: paco_release and graph eviction free the arrays)
static thread_local int *t_last_info = nullptr;
static thread_local int t_last_count = 0;
static thread_local unsigned t_last_gen = 0;

static void set_last_info(int *info, int count)
{
    t_last_info = info;
    t_last_count = count;
    t_last_gen = g_release_gen;
}

// Merge states of the last batched call on a stream (paco_merge_stats): device address,
// merges per round and their offsets; valid while the release generation is unchanged.
struct LastMerges {
    const MergeState *dev = nullptr;
    std::vector<int> ntask;
    std::vector<size_t> off;
    uint64_t gen = 0;
};

struct StreamRes {
    DevBuf ws;      // workspace of the entry points when the caller passes none
    DevBuf plain;   // scratch of paco_svd_plain and of the wide jobs of a batched call
    DevBuf info;    // per-job info of a plain call or of a batched call with wide jobs
    BatchRes batch; // plan, scratch and forked stream of the batched path
    LastMerges merges;
};
static std::unordered_map<hipStream_t, StreamRes *> g_streams; // guarded by g_mu

static StreamRes *stream_res(hipStream_t s)
{
    StreamRes *&r = g_streams[s];
    if (!r) r = new StreamRes();
    return r;
}

static void *internal_ws(hipStream_t s, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return stream_res(s)->ws.get(bytes);
}

struct GraphEntry {
    BatchRes res;
    BatchPlan plan;
    hipGraphExec_t exec = nullptr;
    int *info = nullptr;
    size_t o_states = 0;
};

static void note_merges(hipStream_t s, const void *dev, size_t o_states, const BatchPlan &bp)
{
    LastMerges &L = stream_res(s)->merges;
    L.dev = reinterpret_cast<const MergeState *>(static_cast<const char *>(dev) + o_states);
    L.ntask = bp.ntask;
    L.off = bp.state_off;
    L.gen = g_release_gen;
}
static std::unordered_map<uint64_t, GraphEntry *> g_graphs;

static uint64_t batch_signature(const paco_svd_job *jobs, int count)
{
    uint64_t h = 1469598103934665603ull ^ (uint64_t)count;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    for (int i = 0; i < count; ++i) {
        mix((uint64_t)jobs[i].M);
        mix((uint64_t)jobs[i].N);
        mix((uint64_t)(uintptr_t)jobs[i].A);
        mix((uint64_t)(uintptr_t)jobs[i].S);
        mix((uint64_t)(uintptr_t)jobs[i].U);
        mix((uint64_t)(uintptr_t)jobs[i].Vt);
    }
    mix((uint64_t)gemm_get_mode());
    mix((uint64_t)g_secular_maxit);
    mix((uint64_t)leaf_size_now());
    return h;
}

static bool job_ok(const paco_svd_job &j, bool vonly = false)
{
    return j.M >= j.N && j.N >= 1 && j.A && j.S && (vonly || (j.U && j.Vt)) && j.N <= PACO_MAX_N &&
           bidiag_fits(j.M, j.N);
}

// ---------------------------------------------------------------------------
// lanes for the serving loop (lane.hpp)
// ---------------------------------------------------------------------------
struct Lane {
    BatchRes res;
    size_t capacity = 0;
    int max_jobs = 0;
    int *info = nullptr; // per-job info words of the last enqueued batch (device)
    int count = 0;
};

size_t job_footprint(int M, int N)
{
    return scratch_layout(M, N).total + sizeof(ProbDesc) + 64 * sizeof(int) +
           sizeof(MergeState) * (size_t)std::max(total_merges(N), 1) + 4096;
}

bool job_is_wide(int M, int N) { return use_grid_bidiag(M, N); }

Lane *lane_create(size_t capacity, int max_jobs)
{
    Lane *l = new Lane();
    l->capacity = capacity;
    l->max_jobs = max_jobs;
    // plan blob bound: descriptors + per-job prefix entries of every array (+ rounds)
    const size_t pin = (size_t)max_jobs * (sizeof(ProbDesc) + 64 * sizeof(int)) + (1 << 20);
    if (ensure(l->res, capacity + pin, pin) != hipSuccess) {
        delete l;
        return nullptr;
    }
    return l;
}

void lane_destroy(Lane *l)
{
    if (!l) return;
    batch_res_destroy(l->res);
    delete l;
}

int lane_enqueue(Lane *l, const paco_svd_job *jobs, int count, hipStream_t s)
{
    if (count <= 0) return 0;
    if (count > l->max_jobs) return PACO_EINVAL;
    BatchPlan bp;
    Prepared pr;
    hipError_t e = prepare_batch(jobs, count, bp, l->res, pr);
    if (e != hipSuccess) return (e == hipErrorInvalidValue) ? PACO_EINVAL : (int)e;
    e = enqueue_batch(bp, l->res, pr, s, true);
    l->info = pr.info;
    l->count = count;
    return (int)e;
}

int lane_info(Lane *l, int *out, int count)
{
    if (!l->info || count > l->count) return PACO_EINVAL;
    return hipMemcpy(out, l->info, sizeof(int) * (size_t)count, hipMemcpyDeviceToHost) == hipSuccess ? 0
                                                                                                         : PACO_EINVAL;
}

size_t wide_scratch_bytes(int M, int N) { return scratch_layout(M, N).total + ws_layout(M, N).total; }

bool job_supported(int M, int N) { return M >= N && N >= 1 && N <= PACO_MAX_N && bidiag_fits(M, N); }

int wide_info(const paco_svd_job &job, const void *scratch, int *out)
{
    const char *ws = static_cast<const char *>(scratch) + scratch_layout(job.M, job.N).total;
    return hipMemcpy(out, ws + ws_layout(job.M, job.N).info, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess
               ? 0
               : PACO_EINVAL;
}

int wide_enqueue(const paco_svd_job &job, void *scratch, hipStream_t s, int grid, bool no_wait)
{
    const int M = job.M, N = job.N;
    const BatchScratch L = scratch_layout(M, N);
    char *base = static_cast<char *>(scratch);
    char *ws = base + L.total;
    float *a = reinterpret_cast<float *>(base + L.a);
    hipError_t e = hipMemcpyAsync(a, job.A, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    int rc = bidiag_single(M, N, a, job.U, job.Vt, job.S, reinterpret_cast<float *>(base + L.e),
                           reinterpret_cast<float *>(base + L.tauq), reinterpret_cast<float *>(base + L.taup), ws, s,
                           false, grid, no_wait);
    if (rc) return rc;
    return dc_single(M, N, a, job.U, job.Vt, job.S, reinterpret_cast<float *>(base + L.e),
                     reinterpret_cast<float *>(base + L.qt), reinterpret_cast<float *>(base + L.wt),
                     reinterpret_cast<uint32_t *>(base + L.idx), ws, nullptr, s);
}

// Runs jobs through the two entry points, one after the other on stream s (the plain
// path); job i reports its info word into info[idx ? idx[i] : i].
static size_t plain_job_bytes(int M, int N) { return scratch_layout(M, N).total + ws_layout(M, N).total; }

// The plain sequence of jobs on stream s with one job's scratch at buf (reused by every
// job, in stream order).
static int run_plain_in(const paco_svd_job *jobs, int count, hipStream_t s, int *info, const int *idx, char *buf,
                        bool vonly = false)
{
    for (int i = 0; i < count; ++i) {
        const int M = jobs[i].M, N = jobs[i].N;
        const BatchScratch L = scratch_layout(M, N);
        char *ws = buf + L.total;
        float *a = reinterpret_cast<float *>(buf + L.a);
        hipError_t e = hipMemcpyAsync(a, jobs[i].A, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return (int)e;
        int rc = bidiag_single(M, N, a, jobs[i].U, jobs[i].Vt, jobs[i].S, reinterpret_cast<float *>(buf + L.e),
                               reinterpret_cast<float *>(buf + L.tauq), reinterpret_cast<float *>(buf + L.taup), ws, s,
                               vonly);
        if (rc) return rc;
        rc = dc_single(M, N, a, jobs[i].U, jobs[i].Vt, jobs[i].S, reinterpret_cast<float *>(buf + L.e),
                       reinterpret_cast<float *>(buf + L.qt), reinterpret_cast<float *>(buf + L.wt),
                       reinterpret_cast<uint32_t *>(buf + L.idx), ws, nullptr, s, vonly);
        if (rc) return rc;
        hipLaunchKernelGGL(copy_int_kernel, dim3(1), dim3(64), 0, s, info + (idx ? idx[i] : i),
                           reinterpret_cast<const int *>(ws + ws_layout(M, N).info));
    }
    return (int)hipGetLastError();
}

static int run_plain(const paco_svd_job *jobs, int count, hipStream_t s, int *info, const int *idx,
                     bool vonly = false)
{
    size_t need = 0;
    for (int i = 0; i < count; ++i) need = std::max(need, plain_job_bytes(jobs[i].M, jobs[i].N));
    char *buf;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        buf = static_cast<char *>(stream_res(s)->plain.get(need));
    }
    if (!buf) return PACO_ENOMEM;
    return run_plain_in(jobs, count, s, info, idx, buf, vonly);
}

// A zeroed per-job info array of the stream (count entries).
static int *stream_info(hipStream_t s, int count)
{
    int *info;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        info = static_cast<int *>(stream_res(s)->info.get(sizeof(int) * (size_t)std::max(count, 1)));
    }
    if (info) hipLaunchKernelGGL(zero_ints_kernel, dim3(cdiv(count, 256)), dim3(256), 0, s, info, count);
    return info;
}

// ---------------------------------------------------------------------------
// LAPACK-style front end (paco_sgesdd): column-major input and outputs with leading
// dimensions, jobz = 'A' | 'S' | 'N', any M, N.
// ---------------------------------------------------------------------------
// dst[r * ldd + c] = src[r + c * lds] for r < rows, c < cols (a column-major matrix into a
// row-major one, or back): 32 x 32 tiles through LDS, both sides coalesced
__global__ __launch_bounds__(256) void transpose_copy_kernel(const float *src, int lds, float *dst, int ldd, int rows,
                                                              int cols)
{
    __shared__ float tile[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) { // read: consecutive r along src's columns
        const int r = r0 + tx, c = c0 + k;
        if (r < rows && c < cols) tile[k][tx] = src[(size_t)c * lds + r];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) { // write: consecutive c along dst's rows
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) dst[(size_t)r * ldd + c] = tile[tx][k];
    }
}

// dst[r * ldd + c] = src[r * lds + c] for r < rows, c < cols
__global__ __launch_bounds__(256) void strided_copy_kernel(const float *src, int lds, float *dst, int ldd, int rows,
                                                            int cols)
{
    const int r = blockIdx.y;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < cols; c += 256 * gridDim.x)
        if (r < rows) dst[(size_t)r * ldd + c] = src[(size_t)r * lds + c];
}

static hipError_t transpose_copy(const float *src, int lds, float *dst, int ldd, int rows, int cols, hipStream_t s)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_copy_kernel, dim3(cdiv(cols, 32), cdiv(rows, 32)), dim3(256), 0, s, src, lds, dst, ldd,
                       rows, cols);
    return hipGetLastError();
}
static hipError_t strided_copy(const float *src, int lds, float *dst, int ldd, int rows, int cols, hipStream_t s)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(strided_copy_kernel, dim3(std::min(cdiv(cols, 256), 16), rows), dim3(256), 0, s, src, lds, dst,
                       ldd, rows, cols);
    return hipGetLastError();
}

static int sgesdd_impl(char jobz, int m, int n, const float *A, int lda, float *S, float *U, int ldu, float *VT,
                       int ldvt, hipStream_t s)
{
    const bool all = jobz == 'A' || jobz == 'a', some = jobz == 'S' || jobz == 's', none = jobz == 'N' || jobz == 'n';
    if (!all && !some && !none) return PACO_EINVAL;
    if (m < 1 || n < 1 || !A || !S || lda < m) return PACO_EINVAL;
    const int k = std::min(m, n);
    if (!none) {
        if (!U || !VT || ldu < m || ldvt < (all ? n : k)) return PACO_EINVAL;
    }
    // the engine's problem: row-major Mr x Nr with Mr >= Nr.  m >= n: B = A (a transpose
    // of the column-major input); m < n: B = A^T, whose row-major form is the input itself.
    const bool tall = m >= n;
    const int Mr = tall ? m : n, Nr = tall ? n : m;
    if (Nr > PACO_MAX_N || !bidiag_fits(Mr, Nr)) return PACO_EINVAL;
    const BatchScratch L = scratch_layout(Mr, Nr);
    const size_t o_u = align_up(plain_job_bytes(Mr, Nr), 256);
    const size_t o_vt = align_up(o_u + (none ? 0 : sizeof(float) * (size_t)Mr * Mr), 256);
    const size_t need = align_up(o_vt + (none ? 0 : sizeof(float) * (size_t)Nr * Nr), 256);
    char *buf;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        buf = static_cast<char *>(stream_res(s)->plain.get(need));
    }
    if (!buf) return PACO_ENOMEM;
    char *ws = buf + L.total;
    float *a = reinterpret_cast<float *>(buf + L.a);
    float *u = none ? nullptr : reinterpret_cast<float *>(buf + o_u);
    float *vt = none ? nullptr : reinterpret_cast<float *>(buf + o_vt);
    hipError_t e = tall ? transpose_copy(A, lda, a, n, m, n, s) : strided_copy(A, lda, a, m, n, m, s);
    if (e != hipSuccess) return (int)e;
    int rc = bidiag_single(Mr, Nr, a, u, vt, S, reinterpret_cast<float *>(buf + L.e),
                           reinterpret_cast<float *>(buf + L.tauq), reinterpret_cast<float *>(buf + L.taup), ws, s, none);
    if (rc) return rc;
    rc = dc_single(Mr, Nr, a, u, vt, S, reinterpret_cast<float *>(buf + L.e), reinterpret_cast<float *>(buf + L.qt),
                   reinterpret_cast<float *>(buf + L.wt), reinterpret_cast<uint32_t *>(buf + L.idx), ws, nullptr, s,
                   none);
    if (rc) return rc;
    int *info = stream_info(s, 1);
    if (!info) return PACO_ENOMEM;
    hipLaunchKernelGGL(copy_int_kernel, dim3(1), dim3(64), 0, s, info,
                       reinterpret_cast<const int *>(ws + ws_layout(Mr, Nr).info));
    if (!none) {
        const int ucols = all ? m : k, vrows = all ? n : k;
        if (tall) {
            // U (m x ucols) = u[:, :ucols], VT (n x n) = vt: column-major copies of row-major
            if ((e = transpose_copy(u, Mr, U, ldu, ucols, m, s)) != hipSuccess) return (int)e;
            if ((e = transpose_copy(vt, Nr, VT, ldvt, n, n, s)) != hipSuccess) return (int)e;
        } else {
            // A = V_B S U_B^T: U = V_B, whose column-major form is vt row by row; VT = U_B^T,
            // whose column-major form is u row by row (its first vrows entries)
            if ((e = strided_copy(vt, Nr, U, ldu, m, m, s)) != hipSuccess) return (int)e;
            if ((e = strided_copy(u, Mr, VT, ldvt, n, vrows, s)) != hipSuccess) return (int)e;
        }
    }
    set_last_info(info, 1);
    return 0;
}

// ---------------------------------------------------------------------------
// allocation-free plain path (paco_svd_plain_noalloc): the scratch of the jobs and their
// info words live in a fixed device array of the library, so a call allocates nothing.
// Every call on one device shares that device's array: calls that overlap on one device
// race.  Layout: [info words of up to NOALLOC_MAX_JOBS jobs | one job's scratch].
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_JOBS = 4096;
constexpr size_t NOALLOC_INFO_BYTES = sizeof(int) * NOALLOC_MAX_JOBS;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static char *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of the current device's scratch array: looked up once per device (the
// array's name is a host-side stand-in, never a device address).
static int noalloc_scratch(char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return PACO_EINVAL;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return PACO_ENOMEM;
        g_noalloc_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}

static int svd_plain_noalloc(const paco_svd_job *jobs, int count, hipStream_t s)
{
    if (count > NOALLOC_MAX_JOBS) return PACO_ETOOBIG;
    for (int i = 0; i < count; ++i) {
        if (!job_ok(jobs[i])) return PACO_EINVAL;
        if (plain_job_bytes(jobs[i].M, jobs[i].N) > NOALLOC_SCRATCH_BYTES - NOALLOC_INFO_BYTES) return PACO_ETOOBIG;
    }
    char *base = nullptr;
    const int rc0 = noalloc_scratch(&base);
    if (rc0) return rc0;
    int *info = reinterpret_cast<int *>(base);
    // the array keeps the previous call's contents: the info words start from zero (a
    // fill kernel, so the call can be captured); the rest is written before it is read
    hipLaunchKernelGGL(zero_ints_kernel, dim3(cdiv(count, 256)), dim3(256), 0, s, info, count);
    const int rc = run_plain_in(jobs, count, s, info, nullptr, base + NOALLOC_INFO_BYTES);
    if (rc) return rc;
    set_last_info(info, count);
    return 0;
}

} // namespace paco

using namespace paco;

extern "C" {

int paco_upper_bidiagonalization(unsigned blockDim, void *stream, int M, int N, float *a, float *u, float *vt,
                                 float *d, float *e, float *tauq, float *taup, void *workspace)
{
    (void)blockDim;
    if (M < N || N < 1 || !a || !u || !vt || !d || !tauq || (N > 1 && (!e || !taup))) return PACO_EINVAL;
    if (!bidiag_fits(M, N)) return PACO_EINVAL;
    return bidiag_single(M, N, a, u, vt, d, e, tauq, taup, workspace, static_cast<hipStream_t>(stream));
}

int paco_svd_DC(unsigned blockDim, void *stream, int M, int N, float *a, float *u, float *vt, float *d, float *e,
                float *qt, float *wt, uint32_t *idx, void *workspace, const void *tiling_dev)
{
    (void)blockDim;
    if (M < N || N < 1 || N > PACO_MAX_N || !a || !u || !vt || !d || !qt || !wt || !idx || (N > 1 && !e))
        return PACO_EINVAL;
    return dc_single(M, N, a, u, vt, d, e, qt, wt, idx, workspace, tiling_dev, static_cast<hipStream_t>(stream));
}

size_t paco_workspace_size(int M, int N)
{
    if (M < N || N < 1) return 0;
    return ws_layout(M, N).total;
}

int paco_set_precision(int mode)
{
    const int old = gemm_get_mode();
    if (mode >= 0 && mode <= 2) gemm_set_mode(mode);
    return old;
}

int paco_sgesdd(char jobz, int m, int n, const float *A, int lda, float *S, float *U, int ldu, float *VT, int ldvt,
                void *stream)
{
    return sgesdd_impl(jobz, m, n, A, lda, S, U, ldu, VT, ldvt, static_cast<hipStream_t>(stream));
}

int paco_svd_plain_noalloc(const paco_svd_job *jobs, int count, void *stream)
{
    if (count <= 0) return 0;
    return svd_plain_noalloc(jobs, count, static_cast<hipStream_t>(stream));
}

int paco_svd_plain(const paco_svd_job *jobs, int count, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (count <= 0) return 0;
    for (int i = 0; i < count; ++i)
        if (!job_ok(jobs[i])) return PACO_EINVAL;
    int *info = stream_info(s, count);
    if (!info) return PACO_ENOMEM;
    const int rc = run_plain(jobs, count, s, info, nullptr);
    if (rc) return rc;
    set_last_info(info, count);
    return 0;
}

// Jobs wider than the one-workgroup bidiagonalization go through the chip-wide path;
// narrow / wide hold the jobs' indices in the call.
static bool split_wide(const paco_svd_job *jobs, int count, std::vector<int> &narrow, std::vector<int> &wide)
{
    for (int i = 0; i < count; ++i) (jobs[i].N > BIDIAG_WG_MAXN ? wide : narrow).push_back(i);
    return !wide.empty();
}

// The batched launch sequence of jobs on stream s.  With info_ext, job i reports into
// info_ext[info_idx[i]]; otherwise into the plan's own array, which becomes the thread's
// last info.
static int batched_narrow(const paco_svd_job *jobs, int count, hipStream_t s, int *info_ext, const int *info_idx,
                          bool vonly)
{
    std::lock_guard<std::mutex> lk(g_mu);
    BatchRes &res = stream_res(s)->batch;
    BatchPlan bp;
    Prepared pr;
    hipError_t e = prepare_batch(jobs, count, bp, res, pr, info_ext, info_idx, vonly);
    if (e == hipErrorInvalidValue && !bp.fits) return PACO_EINVAL;
    if (e != hipSuccess) return (int)e;
    e = enqueue_batch(bp, res, pr, s, true);
    if (e != hipSuccess) return (int)e;
    if (!info_ext) set_last_info(pr.info, count);
    note_merges(s, res.dev, pr.o_states, bp);
    return 0;
}

static void graph_entry_destroy(GraphEntry *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    batch_res_destroy(g->res);
    delete g;
}

// The narrow jobs recorded into a hipGraph per batch signature (shapes, pointers,
// precision, iteration limit, values-only) on first use, replayed afterwards.
static int batched_graph(const paco_svd_job *jobs, int count, hipStream_t s, bool vonly)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const uint64_t sig = batch_signature(jobs, count) ^ (vonly ? 0x9e3779b97f4a7c15ull : 0ull);
    auto it = g_graphs.find(sig);
    GraphEntry *g = nullptr;
    if (it == g_graphs.end()) {
        if (g_graphs.size() >= 8) {
            // the evicted graphs may still run on any stream
            (void)hipDeviceSynchronize();
            for (auto &kv : g_graphs) graph_entry_destroy(kv.second);
            g_graphs.clear();
            ++g_release_gen;
        }
        g = new GraphEntry();
        Prepared pr;
        hipError_t e = prepare_batch(jobs, count, g->plan, g->res, pr, nullptr, nullptr, vonly);
        if (e != hipSuccess) {
            graph_entry_destroy(g);
            return (e == hipErrorInvalidValue) ? PACO_EINVAL : (int)e;
        }
        g->info = pr.info;
        g->o_states = pr.o_states;
        // record on a library stream (the caller's may be the legacy null stream, which
        // cannot be captured); the graph is launched on the caller's stream below
        static hipStream_t cap = nullptr;
        if (!cap && (e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking)) != hipSuccess) {
            graph_entry_destroy(g);
            return (int)e;
        }
        hipGraph_t graph;
        e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            graph_entry_destroy(g);
            return (int)e;
        }
        e = enqueue_batch(g->plan, g->res, pr, cap, false);
        hipError_t e2 = hipStreamEndCapture(cap, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (e2 == hipSuccess) (void)hipGraphDestroy(graph);
            graph_entry_destroy(g);
            return (int)(e != hipSuccess ? e : e2);
        }
        e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            graph_entry_destroy(g);
            return (int)e;
        }
        g_graphs[sig] = g;
    } else {
        g = it->second;
    }
    hipError_t e = hipGraphLaunch(g->exec, s);
    if (e != hipSuccess) return (int)e;
    set_last_info(g->info, count);
    note_merges(s, g->res.dev, g->o_states, g->plan);
    return 0;
}

// Whole SVDs (vonly: singular values only) of a batch: narrow jobs as one launch sequence
// (eager or a replayed graph), wide jobs through the entry points; info in call order.
static int batched_impl(const paco_svd_job *jobs, int count, hipStream_t s, bool graph, bool vonly)
{
    if (count <= 0) return 0;
    for (int i = 0; i < count; ++i)
        if (!job_ok(jobs[i], vonly)) return PACO_EINVAL;
    std::vector<int> ni, wi;
    if (!split_wide(jobs, count, ni, wi))
        return graph ? batched_graph(jobs, count, s, vonly) : batched_narrow(jobs, count, s, nullptr, nullptr, vonly);
    // mixed: one info array in call order; the narrow jobs as one batch, the wide ones
    // through the entry points
    int *info = stream_info(s, count);
    if (!info) return PACO_ENOMEM;
    std::vector<paco_svd_job> nj, wj;
    for (int i : ni) nj.push_back(jobs[i]);
    for (int i : wi) wj.push_back(jobs[i]);
    int rc = nj.empty() ? 0 : batched_narrow(nj.data(), (int)nj.size(), s, info, ni.data(), vonly);
    if (rc) return rc;
    rc = run_plain(wj.data(), (int)wj.size(), s, info, wi.data(), vonly);
    if (rc) return rc;
    set_last_info(info, count);
    return 0;
}

int paco_svd_batched(const paco_svd_job *jobs, int count, void *stream)
{
    return batched_impl(jobs, count, static_cast<hipStream_t>(stream), false, false);
}

int paco_svd_batched_graph(const paco_svd_job *jobs, int count, void *stream)
{
    return batched_impl(jobs, count, static_cast<hipStream_t>(stream), true, false);
}

int paco_svd_values_batched(const paco_svd_job *jobs, int count, int graph, void *stream)
{
    return batched_impl(jobs, count, static_cast<hipStream_t>(stream), graph != 0, true);
}

size_t paco_batch_workspace_size(const int *M, const int *N, int count)
{
    size_t t = 0;
    for (int i = 0; i < count; ++i) t += scratch_layout(M[i], N[i]).total;
    return t + (size_t)count * (sizeof(ProbDesc) + 64 * sizeof(int)) + (1 << 20);
}

int paco_last_info(int *info_host, int count, void *stream)
{
    if (!t_last_info || !info_host || t_last_gen != g_release_gen) return 0;
    const int n = std::min(count, t_last_count);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(info_host, t_last_info, sizeof(int) * n, hipMemcpyDeviceToHost, s) != hipSuccess) return 0;
    if (hipStreamSynchronize(s) != hipSuccess) return 0;
    return n;
}

int paco_gemm_f32(void *stream, int transA, int transB, int m, int n, int k, float alpha, const float *A, int lda,
                  const float *B, int ldb, float beta, float *C, int ldc)
{
    if (m < 0 || n < 0 || k < 0 || !A || !B || !C) return PACO_EINVAL;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    GemmDesc g;
    memset(&g, 0, sizeof(g));
    g.A = A;
    g.B = B;
    g.C = C;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.m = m;
    g.n = n;
    g.k = k;
    g.transA = transA;
    g.transB = transB;
    g.alpha = alpha;
    g.beta = beta;
    g.tile_off = 0;
    g.tiles_n = cdiv(n, 64);
    const int tiles = cdiv(m, 64) * g.tiles_n;
    static GemmDesc *dg = nullptr;
    static hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!dg && hipMalloc(&dg, sizeof(GemmDesc)) != hipSuccess) return PACO_ENOMEM;
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return PACO_ENOMEM;
        // the descriptor buffer is reused: the previous product must have consumed it
        (void)hipEventSynchronize(ev);
        hipError_t e = hipMemcpyAsync(dg, &g, sizeof(g), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return (int)e;
        e = launch_gemm_grouped(dg, 1, tiles, s);
        if (e != hipSuccess) return (int)e;
        (void)hipEventRecord(ev, s);
        (void)hipStreamSynchronize(s);
    }
    return 0;
}

int paco_timing_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (on) {
        for (int i = 0; i <= kPhases; ++i)
            if (!g_tev[i] && hipEventCreate(&g_tev[i]) != hipSuccess) return PACO_ENOMEM;
        for (int i = 0; i < 2; ++i)
            if (!g_tev_dc[i] && hipEventCreate(&g_tev_dc[i]) != hipSuccess) return PACO_ENOMEM;
    }
    g_timing = on != 0;
    return 0;
}

int paco_timing_read(float *ms, int max)
{
    if (!g_tev_recorded || !ms) return 0;
    if (hipEventSynchronize(g_tev[kPhases]) != hipSuccess) return 0;
    int n = 0;
    for (int i = 0; i < kPhases && i < max; ++i) {
        float t = 0.f;
        // the D&C beside Q and P^T: its own duration (phases 2 and 3 then overlap)
        hipEvent_t a = g_tev[i], b = g_tev[i + 1];
        if (i == 3 && g_tev_dc_on) {
            a = g_tev_dc[0];
            b = g_tev_dc[1];
        }
        if (hipEventElapsedTime(&t, a, b) != hipSuccess) t = -1.f;
        ms[n++] = t;
    }
    return n;
}

void paco_release(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipDeviceSynchronize();
    for (auto &kv : g_graphs) graph_entry_destroy(kv.second);
    g_graphs.clear();
    for (auto &kv : g_streams) {
        batch_res_destroy(kv.second->batch);
        kv.second->ws.release();
        kv.second->plain.release();
        kv.second->info.release();
        delete kv.second;
    }
    g_streams.clear();
    ++g_release_gen;
}

int paco_merge_stats(double *out, int max_rounds, void *stream)
{
    if (!out || max_rounds < 1) return PACO_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    LastMerges L;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        L = stream_res(s)->merges;
        if (!L.dev || L.gen != g_release_gen) return 0;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return PACO_EINVAL;
    size_t total = 0;
    for (size_t r = 0; r < L.ntask.size(); ++r) total = std::max(total, L.off[r] + (size_t)L.ntask[r]);
    std::vector<MergeState> st(std::max<size_t>(total, 1));
    if (total && hipMemcpy(st.data(), L.dev, sizeof(MergeState) * total, hipMemcpyDeviceToHost) != hipSuccess)
        return PACO_EINVAL;
    const int R = std::min((int)L.ntask.size(), max_rounds);
    for (int r = 0; r < R; ++r) {
        double *o = out + (size_t)PACO_MERGE_STATS * r;
        for (int q = 0; q < PACO_MERGE_STATS; ++q) o[q] = 0.0;
        for (int t = 0; t < L.ntask[r]; ++t) {
            const MergeState &S = st[L.off[r] + t];
            const double k = S.k;
            o[0] += 1;
            o[1] += S.n;
            o[2] += S.k;
            o[3] += S.kl;
            o[4] += S.kr;
            o[5] += S.km;
            o[6] += 2.0 * k * k * (S.n + S.m);
            o[7] += 2.0 * k * ((double)(S.kl + S.km) * S.n1 + (double)(S.kr + S.km) * S.n2 + 1.0) +
                    2.0 * k * ((double)(S.kl + S.km + 1) * S.m1 + (double)(S.kr + S.km + 1) * S.m2);
        }
    }
    return R;
}

int paco_set_leaf_size(int ls)
{
    const int cur = leaf_size_now();
    if (ls >= LEAF_MIN && ls <= LEAF_MAX) return g_leaf_size.exchange(ls);
    return cur;
}

int paco_leaf_size(void) { return leaf_size_now(); }

int paco_set_secular_maxiter(int maxit)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const int old = g_secular_maxit;
    if (maxit >= 1) g_secular_maxit = maxit;
    return old;
}

size_t paco_info_offset(int M, int N)
{
    if (M < N || N < 1) return 0;
    return ws_layout(M, N).info;
}

} // extern "C"
