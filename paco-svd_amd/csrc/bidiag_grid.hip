// Golub-Kahan bidiagonalization of one large matrix on the whole chip.
//
// Reference: upper_bidiagonalization.cpp Process :83-120 (one reflector per column,
// rank-1 updates of the whole trailing matrix on 40 cores, 4 grid barriers per column).
// Here: a blocked panel of NB columns (LAPACK dlabrd structure).  Inside a panel the
// trailing matrix is only read (two GEMVs per column) and the panel's reflectors are
// applied once per panel as a rank-2*NB update on the matrix cores.  The trailing
// matrix of a panel is cut into a GR x GC grid of tiles.  The Householder scaling of
// each column / row reflector is deferred (partial sums are taken against the
// unscaled vector and scaled on reduction), so every column needs two grid-wide
// synchronisation points:
//   column phase (column i): finish x of column i-1 from its GEMV partials (every tile
//     reduces only its own rows), column-i update, norm partials, GEMV A^T v partials
//   row phase (column i): column-i reflector, y for own columns, row-i update, norm
//     partials, GEMV A u partials
// One persistent launch per panel runs both phases of all NB columns, separated by grid
// barriers.  Its workgroups register when they start; the first to register closes the
// registration once every workgroup of the grid has joined (or after a short wait, when
// other work holds part of the chip) and the participants share the tiles, so two such
// launches on different streams never wait for each other's workgroups.  Every tile's
// partial sums sit in their own slot and are reduced in tile order, so the results do
// not depend on how many workgroups took part.  A barrier wait is bounded: on a timeout
// the launch (and every later panel of the call) stops and the problem's info is -2.
// Below the switch column the remaining trailing matrix is finished by the
// one-workgroup kernel (bidiag_wg_kernel) on the sub-matrix.
#include <algorithm>
#include <cstdlib>

#include "gemm_tile.hpp"
#include "paco_kernels.hpp"

namespace paco {

// threads of a panel workgroup: one row (and one column) of its tile each, 256 VGPRs
#define GBT 512
constexpr int NB = 32;
// workgroup grid over the trailing matrix of a panel: gr x gc tiles (runtime, at most
// GMAX x GMAX; the workspace holds GMAX partial sums per row / column)
constexpr int GMAX = 64;
constexpr int SWITCH_DEFAULT = 384;
// columns of a tile at most (its Y and U columns sit in LDS: 2 NB TW_MAX floats)
constexpr int TW_MAX = 512;
static inline size_t panel_dyn_floats_host(int th, int tw) { return 2 * (size_t)NB * tw + th + tw; }
constexpr unsigned REG_CLOSED = 1u << 30;
constexpr unsigned NPARTS_READY = 1u << 30;

// Synchronisation words of the persistent panel launches (in the workspace).
struct GridSync {
    unsigned reg;    // registrations so far | REG_CLOSED once closed
    unsigned nparts; // participants | NPARTS_READY (written by the first registrant)
    unsigned count;  // arrivals at the current barrier
    unsigned gen;    // barriers completed
    unsigned abort;  // set when a wait timed out: every later wait of the call returns at once
    unsigned pad[59];
};

// participants of a panel launch at most (one arrival flag each)
constexpr int MAX_PARTS = 1024;
constexpr int FLAG_STRIDE = 16; // unsigned words between two flags (one 64-byte line each)

struct GridWs {
    float *X, *Y, *wcol, *wrow, *g1p, *g2p, *t1p, *t2p, *t3p, *t4p;
    double *ssp, *rsp; // FP64 partial sums of squares (see house_params_d)
    GridSync *sync;
    unsigned *flags; // [MAX_PARTS * FLAG_STRIDE] arrival flags of the grid barrier
};

static size_t carve(int M, int N, char *base, GridWs *W)
{
    size_t o = 0;
    auto take = [&](size_t nfl) {
        float *p = base ? reinterpret_cast<float *>(base + o) : nullptr;
        o += (nfl * sizeof(float) + 255) / 256 * 256;
        return p;
    };
    GridWs w;
    w.X = take((size_t)M * NB);
    w.Y = take((size_t)N * NB);
    w.wcol = take(M);
    w.wrow = take(N);
    w.g1p = take((size_t)GMAX * N);
    w.g2p = take((size_t)GMAX * M);
    w.ssp = reinterpret_cast<double *>(take(2 * GMAX));
    w.rsp = reinterpret_cast<double *>(take(2 * GMAX));
    w.t1p = take(GMAX * NB);
    w.t2p = take(GMAX * NB);
    w.t3p = take(GMAX * NB);
    w.t4p = take(GMAX * NB);
    w.sync = reinterpret_cast<GridSync *>(take(sizeof(GridSync) / sizeof(float)));
    w.flags = reinterpret_cast<unsigned *>(take((size_t)MAX_PARTS * FLAG_STRIDE));
    if (W) *W = w;
    return o;
}

// the panel kernels' part, then the resident kernel's exchange buffers
size_t bidiag_grid_ws_bytes(int M, int N) { return carve(M, N, nullptr, nullptr) + bidiag_res_ws_bytes(N); }

struct PanelArgs {
    ProbDesc P;
    GridWs W;
    int p, th, tw;
    int gr, gc; // tile grid
    long long timeout; // bound of one barrier wait, in wall-clock ticks
};

// Data one workgroup writes during a panel and another reads after a barrier (the
// reflectors in A, X, Y, the reflector columns / rows and every partial sum) moves through
// device-coherent accesses (agent-scope atomic loads and stores: they bypass the per-XCD
// L2), so a barrier needs no cache write-back or invalidation.  The trailing matrix
// itself does not change during a panel and is read through the caches.
__device__ inline float ldc(const float *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double ldc(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc(float *p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void stc(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// p[0] + p[stride] + ... + p[(n - 1) stride] through device-coherent loads, added in that
// order, with 16 loads in flight (a loop of dependent loads would wait for each one)
template <typename T>
__device__ __attribute__((always_inline)) inline T ldc_sum(const T *p, size_t stride, int n)
{
    T s = 0;
    for (int q0 = 0; q0 < n; q0 += 16) {
        T v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = ldc(p + (size_t)min(q0 + k, n - 1) * stride);
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (q0 + k < n) s += v[k];
    }
    return s;
}

// column reflector t of the panel (unit at row p+t), entry at row r
__device__ inline float Vel(const float *A, int N, int p, int t, int r)
{
    const int c = p + t;
    return r > c ? ldc(A + (size_t)r * N + c) : (r == c ? 1.f : 0.f);
}
// row reflector t of the panel (row p+t, unit at column p+t+1), entry at column col
__device__ inline float Uel(const float *A, int N, int p, int t, int col)
{
    const int u = p + t + 1;
    return col > u ? ldc(A + (size_t)(p + t) * N + col) : (col == u ? 1.f : 0.f);
}

// Per-tile state of a panel.  Row role: thread tid owns row r0 + tid (th <= GBT) and
// keeps V[row][t] and X[row][t] of the panel's reflectors in registers, and c (column i
// at its row) from the column phase to the row phase.  Column role: thread tid < tw owns
// column c0 + tid; Y[col][t] and U[t][col] of its column sit in LDS (Yl, Ul), w (row i
// at its column) stays in a register from the row phase to the next column phase.  When
// a workgroup takes more than one tile (fewer participants than tiles, or one launch per
// phase), the state is reloaded from the global copies at the start of every phase: the
// same values, so the results are the same bit for bit.
struct RowState {
    float v[NB], x[NB];
    float c, w;
};

// every entry zero: threads without a row or column multiply their (unused) state by zero
// in the block reductions, which must not see stale non-finite register contents
__device__ __attribute__((always_inline)) inline void row_state_zero(RowState &R)
{
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        R.v[t] = 0.f;
        R.x[t] = 0.f;
    }
    R.c = R.w = 0.f;
}

// LDS of the panel launch: the dynamic part holds Yl[NB][tw], Ul[NB][tw], wc[th], wl[tw].
struct PhaseLds {
    union {
        float4 red4[GBT];   // GEMV1 column sums across the wavefronts
        struct {            // partial sums of all tiles, staged for an in-order reduction
            float pa[GMAX][NB], pb[GMAX][NB];
            double pd[GMAX];
        } st;
    } u;
    float redk[GBT / 64][2 * NB];
    float crossa[NB], crossb[NB];               // row i / column i entries of the panel vectors
    float ta[NB + 1], tb[NB + 1];               // combined coefficients of the current phase
    double red[GBT / 64];
};

// a[j] = v for a runtime j (a select per entry keeps the array in registers)
__device__ __attribute__((always_inline)) inline void set_at(float (&a)[NB], int j, float v)
{
#pragma unroll
    for (int t = 0; t < NB; ++t) a[t] = (t == j) ? v : a[t];
}

// value of a[t] for a runtime t < NB (a select chain keeps the array in registers)
__device__ __attribute__((always_inline)) inline float get_at(const float (&a)[NB], int t)
{
    float v = 0.f;
#pragma unroll
    for (int q = 0; q < NB; ++q) v = (q == t) ? a[q] : v;
    return v;
}

// Reload of the panel state of one row (V[row][t], X[row][t], t < nv / t < nx) and one
// column (Y[col][t], U[t][col], t < ny / t < nu) from the global copies: every load is
// issued unconditionally from a valid address and the entries the panel has not defined
// are selected away (V is 1 / 0 at and above its head, X exists below row p + t, Y right
// of column p + t, U is 1 / 0 at and left of its head), so all loads are in flight at once.
__device__ __attribute__((always_inline)) inline void reload_row(const float *A, const GridWs &W, int N, int p, int row,
                                                                int nv, int nx, RowState &R)
{
    float v[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = ldc(A + (size_t)row * N + p + t);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        const int c = p + t;
        R.v[t] = t < nv ? (row > c ? v[t] : (row == c ? 1.f : 0.f)) : 0.f;
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = ldc(W.X + (size_t)row * NB + t);
#pragma unroll
    for (int t = 0; t < NB; ++t) R.x[t] = (t < nx && row > p + t) ? v[t] : 0.f;
}
__device__ __attribute__((always_inline)) inline void reload_col(const float *A, const GridWs &W, int N, int p, int col,
                                                                int ny, int nu, float *Yc, float *Uc, int tw)
{
    float y[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) y[t] = ldc(W.Y + (size_t)col * NB + t);
#pragma unroll
    for (int t = 0; t < NB; ++t) Yc[t * tw] = (t < ny && col > p + t) ? y[t] : 0.f;
#pragma unroll
    for (int t = 0; t < NB; ++t) y[t] = ldc(A + (size_t)(p + t) * N + col);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        const int h = p + t + 1; // head of row reflector t
        Uc[t * tw] = t < nu ? (col > h ? y[t] : (col == h ? 1.f : 0.f)) : 0.f;
    }
}

// Column phase of tile (rb, cb) for panel column j (j == NB: only the finish of x of
// column NB - 1 and the row reflector of row p + NB - 1).
__device__ __attribute__((always_inline)) inline void gb_col_phase(const PanelArgs &a, int j, int rb, int cb, float *dyn, PhaseLds &S, RowState &R,
                             bool resident)
{
    const int GC = a.gc;
    const bool finish_only = j == NB;
    const ProbDesc &P = a.P;
    const GridWs &W = a.W;
    const int M = P.M, N = P.N, p = a.p, i = p + j, ip = i - 1, jp = j - 1;
    float *A = P.a;
    const int r0 = p + rb * a.th, r1 = min(M, r0 + a.th);
    const int c0 = p + cb * a.tw, c1 = min(N, c0 + a.tw);
    const int tw = a.tw;
    float *Yl = dyn, *Ul = dyn + NB * tw, *wc = dyn + 2 * NB * tw;
    const int tid = threadIdx.x;
    const int row = r0 + tid, col = c0 + tid;
    const bool rowv = row < r1 && tid < a.th, colv = tid < tw && col < c1;

    // ---- one round of loads, spread over the threads: every tile's partial sums into LDS
    // (reduced in tile order below), the row i / column i entries, this thread's own row
    float x12 = 0.f;
    if (j > 0) {
        for (int q = tid; q < GC * NB; q += GBT) {
            S.u.st.pa[q / NB][q % NB] = ldc(W.t3p + q);
            S.u.st.pb[q / NB][q % NB] = ldc(W.t4p + q);
        }
        if (tid < GC) S.u.st.pd[tid] = ldc(W.rsp + tid);
        if (tid == GBT - 1) x12 = ldc(W.wrow + ip + 1);
    }
    if (tid < NB) {
        const int t = tid;
        // Y[i][t] (t < j) and U[t][i] (t < j; the unit head for t = j - 1)
        S.crossa[t] = t < j ? ldc(W.Y + (size_t)i * NB + t) : 0.f;
        S.crossb[t] = t < j ? Uel(A, N, p, t, i) : 0.f;
    }
    float g2 = 0.f, ari = 0.f;
    const bool live = rowv && row >= i;
    if (live) {
        if (j > 0) g2 = ldc_sum(W.g2p + row, (size_t)M, GC);
        if (!finish_only) ari = A[(size_t)row * N + i];
        else ari = A[(size_t)row * N + ip + 1];
    }
    if (!resident) {
        // (the workspace holds stale values where the panel has not written: selected away)
        if (rowv) reload_row(A, W, N, p, row, j, jp, R);
        if (colv) {
            reload_col(A, W, N, p, col, j, jp, Yl + tid, Ul + tid, tw);
            R.w = (j > 0 && col > ip) ? ldc(W.wrow + col) : 0.f;
        }
    }
    if (j > 0 && tid == GBT - 1) S.ta[NB] = x12;
    __syncthreads();
    if (j > 0) {
        // every thread forms the row reflector's parameters from the staged sums (the same
        // sums in the same order: the same values everywhere)
        double sig2 = 0.0;
        for (int q = 0; q < GC; ++q) sig2 += S.u.st.pd[q];
        float beta2, tau2, inv2;
        house_params_d(S.ta[NB], sig2, &beta2, &tau2, &inv2);
        if (tid < NB) {
            const int t = tid;
            float s3 = 0.f, s4 = 0.f;
            for (int q = 0; q < GC; ++q) {
                s3 += S.u.st.pa[q][t];
                s4 += S.u.st.pb[q][t];
            }
            S.ta[t] = t <= jp ? S.crossa[t] + inv2 * s3 : 0.f; // t3: Y^T u
            S.tb[t] = t < jp ? S.crossb[t] + inv2 * s4 : 0.f;  // t4: U u
        }
        __syncthreads();
        // ---- finish the row reflector of row ip and x of column ip
        if (live) {
            float acc = ari + inv2 * g2;
#pragma unroll
            for (int t = 0; t < NB; ++t) acc -= R.v[t] * S.ta[t];
#pragma unroll
            for (int t = 0; t < NB; ++t) acc -= R.x[t] * S.tb[t];
            const float x = tau2 * acc;
            set_at(R.x, jp, x);
            if (cb == 0) stc(W.X + (size_t)row * NB + jp, x);
        }
        // the row reflector u_{ip} for own columns; into A (row ip) by the first tile row
        if (colv && col >= ip + 1) {
            const float u = (col == ip + 1) ? 1.f : (tau2 != 0.f) ? R.w * inv2 : R.w;
            Ul[jp * tw + tid] = u;
            if (rb == 0)
                stc(A + (size_t)ip * N + col, col == ip + 1 ? (finish_only ? 1.f : beta2) : u);
        } else if (colv) {
            Ul[jp * tw + tid] = 0.f;
        }
        if (rb == 0 && cb == 0 && tid == 0) {
            P.e[ip] = beta2;
            P.taup[ip] = tau2;
        }
    }
    if (finish_only) {
        __syncthreads();
        return;
    }

    // ---- column i for own rows >= i (redundant over cb)
    double part = 0.0;
    if (live) {
        float val = ari;
#pragma unroll
        for (int t = 0; t < NB; ++t) val -= R.v[t] * S.crossa[t] + R.x[t] * S.crossb[t];
        wc[tid] = val;
        R.c = val;
        if (cb == 0) stc(W.wcol + row, val);
        if (row > i) part = (double)val * val;
    } else if (rowv) {
        wc[tid] = 0.f;
    }
    // one reduction for the sum of squares and this tile's share of t1p[rb][t] =
    // sum_{row > i} V[row][t] c[row], t2p[rb][t] = sum X[row][t] c[row] (t < j): the 2 NB
    // entries are shared out over the GC tiles of the row block
    const int K1 = (2 * NB + GC - 1) / GC, e01 = cb * K1, ke1 = max(0, min(K1, 2 * NB - e01));
    {
        const int lane = tid & 63, wv = tid >> 6;
        const float cw = (rowv && row > i) ? R.c : 0.f;
        const double ws = wave_sum(part);
        if (lane == 0) S.red[wv] = ws;
        for (int k = 0; k < ke1; ++k) {
            const int e = e01 + k;
            const float v = wave_sum_dpp((e < NB ? get_at(R.v, e) : get_at(R.x, e - NB)) * cw);
            if (lane == 0) S.redk[wv][k] = v;
        }
    }
    __syncthreads(); // (also publishes wc)
    if (tid == 0) {
        double ss = 0.0;
#pragma unroll
        for (int w = 0; w < GBT / 64; ++w) ss += S.red[w];
        if (cb == 0) stc(W.ssp + rb, ss);
    }
    if (tid < ke1) {
        float tot = 0.f;
#pragma unroll
        for (int q = 0; q < GBT / 64; ++q) tot += S.redk[q][tid];
        const int e = e01 + tid, t = e % NB;
        stc((e < NB ? W.t1p : W.t2p) + rb * NB + t, t < j ? tot : 0.f);
    }
    // ---- GEMV1 partials: g1p[rb][col] = sum_{row > i} wc[row] A[row][col], col > i
    const int rs = max(r0, i + 1);
    if ((N & 3) == 0 && rs < r1) {
        // lanes over float4 column groups (64 per pass), the 8 wavefronts over interleaved
        // rows in batches of 8 loads in flight; addresses clamped into the tile, invalid
        // rows weighted 0, invalid columns never stored
        const int lane = tid & 63, wv = tid >> 6;
        const int ca = max(c0, i + 1) & ~3, ncg = (c1 - ca + 3) >> 2;
        for (int cg0 = 0; cg0 < ncg; cg0 += 64) {
            const int cc = ca + 4 * min(cg0 + lane, ncg - 1);
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int rb0 = rs + wv; rb0 < r1; rb0 += 8 * (GBT / 64)) {
                float4 v[8];
                float wk[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int rr = rb0 + k * (GBT / 64);
                    v[k] = *reinterpret_cast<const float4 *>(A + (size_t)min(rr, r1 - 1) * N + cc);
                    wk[k] = rr < r1 ? wc[rr - r0] : 0.f;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    acc.x += wk[k] * v[k].x;
                    acc.y += wk[k] * v[k].y;
                    acc.z += wk[k] * v[k].z;
                    acc.w += wk[k] * v[k].w;
                }
            }
            S.u.red4[wv * 64 + lane] = acc;
            __syncthreads();
            if (wv == 0 && cg0 + lane < ncg) {
                float4 t = S.u.red4[lane];
#pragma unroll
                for (int q = 1; q < GBT / 64; ++q) {
                    const float4 u = S.u.red4[q * 64 + lane];
                    t.x += u.x;
                    t.y += u.y;
                    t.z += u.z;
                    t.w += u.w;
                }
                const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) // own columns only (the aligned start may reach into the previous tile)
                    if (cc + e > i && cc + e >= c0 && cc + e < c1) stc(W.g1p + (size_t)rb * N + cc + e, tv[e]);
            }
            __syncthreads();
        }
    } else if (rs < r1) {
        for (int cc = max(c0, i + 1) + tid; cc < c1; cc += GBT) {
            float acc = 0.f;
#pragma unroll 8
            for (int rr = rs; rr < r1; ++rr) acc += wc[rr - r0] * A[(size_t)rr * N + cc];
            stc(W.g1p + (size_t)rb * N + cc, acc);
        }
    } else {
        for (int cc = max(c0, i + 1) + tid; cc < c1; cc += GBT) stc(W.g1p + (size_t)rb * N + cc, 0.f);
    }
    __syncthreads();
}

// Row phase of tile (rb, cb) for panel column j.
__device__ __attribute__((always_inline)) inline void gb_row_phase(const PanelArgs &a, int j, int rb, int cb, float *dyn, PhaseLds &S, RowState &R,
                             bool resident)
{
    const int GR = a.gr;
    const ProbDesc &P = a.P;
    const GridWs &W = a.W;
    const int M = P.M, N = P.N, p = a.p, i = p + j;
    float *A = P.a;
    const int r0 = p + rb * a.th, r1 = min(M, r0 + a.th);
    const int c0 = p + cb * a.tw, c1 = min(N, c0 + a.tw);
    const int tw = a.tw;
    float *Yl = dyn, *Ul = dyn + NB * tw, *wl = dyn + 2 * NB * tw + a.th;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row = r0 + tid, col = c0 + tid;
    const bool rowv = row < r1 && tid < a.th, colv = tid < tw && col < c1;

    // ---- one round of loads, spread over the threads (as in the column phase)
    float x1 = 0.f;
    for (int q = tid; q < GR * NB; q += GBT) {
        S.u.st.pa[q / NB][q % NB] = ldc(W.t1p + q);
        S.u.st.pb[q / NB][q % NB] = ldc(W.t2p + q);
    }
    if (tid < GR) S.u.st.pd[tid] = ldc(W.ssp + tid);
    if (tid == GBT - 1) x1 = ldc(W.wcol + i);
    if (tid < NB) {
        const int t = tid;
        // V[i][t] and X[i][t], t < j
        S.crossa[t] = t < j ? Vel(A, N, p, t, i) : 0.f;
        S.crossb[t] = t < j ? ldc(W.X + (size_t)i * NB + t) : 0.f;
    }
    float g1 = 0.f, aic = 0.f;
    const bool clive = colv && col > i;
    if (clive) {
        g1 = ldc_sum(W.g1p + col, (size_t)N, GR);
        aic = A[(size_t)i * N + col];
    }
    if (!resident) {
        if (rowv) {
            reload_row(A, W, N, p, row, j, j, R);
            R.c = row >= i ? ldc(W.wcol + row) : 0.f;
        }
        if (colv) reload_col(A, W, N, p, col, j, j, Yl + tid, Ul + tid, tw);
    }
    if (tid == GBT - 1) S.ta[NB] = x1;
    __syncthreads();
    // every thread forms the column reflector's parameters from the staged sums
    double sig = 0.0;
    for (int q = 0; q < GR; ++q) sig += S.u.st.pd[q];
    float beta, tau, inv;
    house_params_d(S.ta[NB], sig, &beta, &tau, &inv);
    float ta_t = 0.f, tb_t = 0.f;
    if (tid < NB) {
        const int t = tid;
        float s1 = 0.f, s2 = 0.f;
        for (int q = 0; q < GR; ++q) {
            s1 += S.u.st.pa[q][t];
            s2 += S.u.st.pb[q][t];
        }
        ta_t = t < j ? S.crossa[t] + inv * s1 : 0.f; // V^T v
        tb_t = t < j ? S.crossb[t] + inv * s2 : 0.f; // X^T v
    }
    // ---- column reflector of column i: V[row][j] for own rows, into A by the first tile column
    if (rowv) {
        float v = 0.f;
        if (row > i) {
            v = (tau != 0.f) ? R.c * inv : R.c;
            if (cb == 0) stc(A + (size_t)row * N + i, v);
        } else if (row == i) {
            v = 1.f;
        }
        set_at(R.v, j, v);
    }
    if (rb == 0 && cb == 0 && tid == 0) {
        stc(A + (size_t)i * N + i, beta);
        P.d[i] = beta;
        P.tauq[i] = tau;
    }
    if (i == N - 1) {
        __syncthreads();
        return;
    }
    __syncthreads(); // every thread has read S.ta[NB]
    if (tid < NB) {
        S.ta[tid] = ta_t;
        S.tb[tid] = tb_t;
    }
    __syncthreads();
    // ---- y for own columns, row-i update
    double part = 0.0;
    if (clive) {
        float acc = aic + inv * g1;
        for (int t = 0; t < j; ++t) acc -= Yl[t * tw + tid] * S.ta[t] + Ul[t * tw + tid] * S.tb[t];
        const float y = tau * acc;
        Yl[j * tw + tid] = y;
        if (rb == 0) stc(W.Y + (size_t)col * NB + j, y);
        float w = aic - y;
        for (int t = 0; t < j; ++t) w -= Yl[t * tw + tid] * S.crossa[t] + Ul[t * tw + tid] * S.crossb[t];
        wl[tid] = w;
        R.w = w;
        if (rb == 0) stc(W.wrow + col, w);
        if (col > i + 1) part = (double)w * w;
    } else if (colv) {
        Yl[j * tw + tid] = 0.f;
        wl[tid] = 0.f;
    }
    // one reduction for the sum of squares and this tile's share of t3p[cb][t] =
    // sum_{col > i+1} Y[col][t] w (t <= j), t4p[cb][t] = sum U[t][col] w (t < j): the 2 NB
    // entries are shared out over the GR tiles of the column block
    const int K3 = (2 * NB + GR - 1) / GR, e03 = rb * K3, ke3 = max(0, min(K3, 2 * NB - e03));
    {
        const float ww = (colv && col > i + 1) ? R.w : 0.f;
        const double ws = wave_sum(part);
        if (lane == 0) S.red[wv] = ws;
        for (int k = 0; k < ke3; ++k) {
            const int e = e03 + k;
            const float v = wave_sum_dpp(colv ? (e < NB ? Yl[e * tw + tid] : Ul[(e - NB) * tw + tid]) * ww : 0.f);
            if (lane == 0) S.redk[wv][k] = v;
        }
    }
    __syncthreads(); // (also publishes wl, Yl)
    if (tid == 0) {
        double ss = 0.0;
#pragma unroll
        for (int w = 0; w < GBT / 64; ++w) ss += S.red[w];
        if (rb == 0) stc(W.rsp + cb, ss);
    }
    if (tid < ke3) {
        float tot = 0.f;
#pragma unroll
        for (int q = 0; q < GBT / 64; ++q) tot += S.redk[q][tid];
        const int e = e03 + tid;
        if (e < NB)
            stc(W.t3p + cb * NB + e, e <= j ? tot : 0.f);
        else
            stc(W.t4p + cb * NB + e - NB, e - NB < j ? tot : 0.f);
    }
    // ---- GEMV2 partials: g2p[cb][row] = sum_{col > i+1} A[row][col] wl[col], row > i
    const int cs = max(c0, i + 2);
    const int rstart = max(r0, i + 1);
    if ((N & 3) == 0 && cs < c1) {
        // a wavefront takes 8 consecutive rows at a time, its lanes float4 column groups;
        // the 8 row sums are reduced across the lanes by a transposing butterfly
        const int ca = cs & ~3, ncg = (c1 - ca + 3) >> 2;
        for (int rb0 = rstart + 8 * wv; rb0 < r1; rb0 += 8 * (GBT / 64)) {
            float acc[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = 0.f;
            for (int cg0 = 0; cg0 < ncg; cg0 += 64) {
                const int cgi = cg0 + lane, cc = ca + 4 * min(cgi, ncg - 1);
                float w4[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int col = cc + e;
                    w4[e] = (cgi < ncg && col >= cs && col < c1) ? wl[col - c0] : 0.f;
                }
                float4 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    v[k] = *reinterpret_cast<const float4 *>(A + (size_t)min(rb0 + k, r1 - 1) * N + cc);
#pragma unroll
                for (int k = 0; k < 8; ++k) acc[k] += v[k].x * w4[0] + v[k].y * w4[1] + v[k].z * w4[2] + v[k].w * w4[3];
            }
            // lanes 8 g + k hold, after the butterfly over each group of 8 lanes, the group's
            // sum for row k; three exchanges add the eight groups
#pragma unroll
            for (int l = 2; l >= 0; --l) {
                const int sh = 1 << l;
                const bool up = (lane & sh) != 0;
#pragma unroll
                for (int k = 0; k < sh; ++k) {
                    const float lo = acc[k], hi = acc[k + sh];
                    acc[k] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, sh, 64);
                }
            }
            float tot = acc[0];
            tot += __shfl_xor(tot, 8, 64);
            tot += __shfl_xor(tot, 16, 64);
            tot += __shfl_xor(tot, 32, 64);
            if (lane < 8 && rb0 + lane < r1) stc(W.g2p + (size_t)cb * M + rb0 + lane, tot);
        }
    } else if (cs < c1) {
        // eight rows per wavefront in flight; lanes over the tile's columns
        for (int rr = rstart + 8 * wv; rr < r1; rr += 8 * (GBT / 64)) {
            float acc[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] = 0.f;
            for (int cc = cs + lane; cc < c1; cc += 64) {
                const float w = wl[cc - c0];
                float v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = (rr + q < r1) ? A[(size_t)(rr + q) * N + cc] : 0.f;
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] += v[q] * w;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] = wave_sum(acc[q]);
            if (lane < 8 && rr + lane < r1) {
                float av = acc[0];
#pragma unroll
                for (int q = 1; q < 8; ++q)
                    if (lane == q) av = acc[q];
                stc(W.g2p + (size_t)cb * M + rr + lane, av);
            }
        }
    } else {
        for (int rr = rstart + tid; rr < r1; rr += GBT) stc(W.g2p + (size_t)cb * M + rr, 0.f);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// grid synchronisation of the persistent panel launch
// ---------------------------------------------------------------------------
__device__ inline unsigned ld_acq(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_rlx(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Marks the call as failed: later waits return at once, the problem's info becomes -2.
__device__ inline void gb_abort(GridSync *gs, int *info)
{
    __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(info, -2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Registration: returns this workgroup's rank (or -1: registration already closed, the
// workgroup leaves) and the number of participants in *parts.
__device__ int gb_register(GridSync *gs, int *info, long long timeout, int *parts)
{
    __shared__ int s_rank, s_parts;
    if (threadIdx.x == 0) {
        int rank = -1, np = 0;
        const unsigned old = __hip_atomic_fetch_add(&gs->reg, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (!(old & REG_CLOSED) && !ld_rlx(&gs->abort)) {
            rank = (int)old;
            const long long t0 = wall_clock64();
            if (rank == 0) {
                // close once every workgroup of the grid has joined, or after a short wait
                // (about 20 us) when part of the chip is busy with other work
                const unsigned G = gridDim.x;
                while ((ld_rlx(&gs->reg) & ~REG_CLOSED) < G && wall_clock64() - t0 < timeout / 100000)
                    __builtin_amdgcn_s_sleep(1);
                const unsigned n =
                    __hip_atomic_fetch_or(&gs->reg, REG_CLOSED, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) & ~REG_CLOSED;
                np = (int)n;
                __hip_atomic_store(&gs->nparts, n | NPARTS_READY, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned v;
                while (!((v = ld_rlx(&gs->nparts)) & NPARTS_READY)) {
                    if (wall_clock64() - t0 > timeout) {
                        gb_abort(gs, info);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                np = (v & NPARTS_READY) ? (int)(v & ~NPARTS_READY) : 0;
                // registered after the count was taken: not a participant
                if (rank >= np) rank = -1;
            }
        }
        s_rank = rank;
        s_parts = np;
    }
    __syncthreads();
    *parts = s_parts;
    return s_rank;
}

// Barrier b (0, 1, ... within the launch) over the np participants; false when the call
// has been aborted.  Every participant stores b + 1 into its own flag (no contended
// atomic); the workgroup of rank 0 polls all flags with one thread per flag, then
// publishes the barrier count; the others wait for that count.  Data crossing the
// barrier moves through device-coherent accesses, so only the stores' completion is
// awaited before arriving.
__device__ bool gb_barrier(GridSync *gs, unsigned *flags, int rank, unsigned np, unsigned b, int *info,
                           long long timeout)
{
    __shared__ int s_ok, s_done;
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        s_ok = 1;
        __hip_atomic_store(flags + (size_t)rank * FLAG_STRIDE, b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (rank == 0) {
        const long long t0 = wall_clock64();
        for (;;) {
            if (threadIdx.x == 0) s_done = 1;
            __syncthreads();
            for (unsigned q = threadIdx.x; q < np; q += GBT)
                if (ld_rlx(flags + (size_t)q * FLAG_STRIDE) != b + 1) s_done = 0;
            __syncthreads();
            if (s_done) break;
            if (wall_clock64() - t0 > timeout || ld_rlx(&gs->abort)) {
                if (threadIdx.x == 0) {
                    gb_abort(gs, info);
                    s_ok = 0;
                }
                break;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && s_ok)
            __hip_atomic_store(&gs->gen, b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        const long long t0 = wall_clock64();
        while (ld_rlx(&gs->gen) != b + 1) {
            if (ld_rlx(&gs->abort)) {
                s_ok = 0;
                break;
            }
            if (wall_clock64() - t0 > timeout) {
                gb_abort(gs, info);
                s_ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return s_ok != 0;
}

// Resets the synchronisation words and the arrival flags before a panel launch (first:
// also the abort flag, at the start of the call).  A fill kernel, not a memset, so that
// the launch sequence can be captured in a graph.
__global__ void gb_sync_reset_kernel(GridSync *gs, unsigned *flags, int first)
{
    for (int q = threadIdx.x; q < MAX_PARTS; q += blockDim.x) flags[(size_t)q * FLAG_STRIDE] = 0;
    if (threadIdx.x == 0) {
        gs->reg = 0;
        gs->nparts = 0;
        gs->count = 0;
        gs->gen = 0;
        if (first) gs->abort = 0;
    }
}


// One panel: both phases of its NB columns, then the finish of the last column.  With as
// many participants as tiles, each participant keeps its tile's panel state (registers,
// LDS) for the whole launch.
__global__ __launch_bounds__(GBT) void gb_panel_kernel(PanelArgs a)
{
    extern __shared__ float dyn[];
    __shared__ PhaseLds S;
    GridSync *gs = a.W.sync;
    int np;
    const int rank = gb_register(gs, a.P.info, a.timeout, &np);
    if (rank < 0) return;
    const int ntiles = a.gr * a.gc;
    const bool resident = np == ntiles;
    RowState R;
    row_state_zero(R);
    for (size_t q = threadIdx.x; q < 2 * (size_t)NB * a.tw; q += GBT) dyn[q] = 0.f;
    __syncthreads();
    unsigned b = 0;
    for (int j = 0; j < NB; ++j) {
        for (int t = rank; t < ntiles; t += np) gb_col_phase(a, j, t / a.gc, t % a.gc, dyn, S, R, resident);
        if (!gb_barrier(gs, a.W.flags, rank, (unsigned)np, b++, a.P.info, a.timeout)) return;
        for (int t = rank; t < ntiles; t += np) gb_row_phase(a, j, t / a.gc, t % a.gc, dyn, S, R, resident);
        if (!gb_barrier(gs, a.W.flags, rank, (unsigned)np, b++, a.P.info, a.timeout)) return;
    }
    for (int t = rank; t < ntiles; t += np) gb_col_phase(a, NB, t / a.gc, t % a.gc, dyn, S, R, resident);
}

// The same phases as one launch each (two launches per column), one tile per workgroup;
// PACO_GRID_PERSIST=0 selects it (comparisons of the two schedules).
__global__ __launch_bounds__(GBT) void gb_col_kernel(PanelArgs a, int j)
{
    extern __shared__ float dyn[];
    __shared__ PhaseLds S;
    if (a.W.sync->abort) return;
    RowState R;
    row_state_zero(R);
    gb_col_phase(a, j, blockIdx.x / a.gc, blockIdx.x % a.gc, dyn, S, R, false);
}
__global__ __launch_bounds__(GBT) void gb_row_kernel(PanelArgs a, int j)
{
    extern __shared__ float dyn[];
    __shared__ PhaseLds S;
    if (a.W.sync->abort) return;
    RowState R;
    row_state_zero(R);
    gb_row_phase(a, j, blockIdx.x / a.gc, blockIdx.x % a.gc, dyn, S, R, false);
}

__global__ void gb_restore_kernel(float *A, int N, int q, const float *e)
{
    if (threadIdx.x == 0) A[(size_t)(q - 1) * N + q] = e[q - 1];
}

// trailing update A[q:, q:] -= V Y^T + X U (q = p + NB), two products on the matrix cores
template <int MODE>
__global__ __launch_bounds__(GT) void gb_trail_kernel(TileCtx base, int tiles_n)
{
    TileCtx t = base;
    t.m0 = (blockIdx.x / tiles_n) * 64;
    t.n0 = (blockIdx.x % tiles_n) * 64;
    tile_gemm<MODE, 16>(t);
}

template <int MODE>
static void launch_trail(const ProbDesc &P, const GridWs &W, int q, int p, hipStream_t s)
{
    const int M = P.M, N = P.N, m = M - q, n = N - q;
    const int tn = (n + 63) / 64, tiles = ((m + 63) / 64) * tn;
    TileCtx t;
    t.brow = t.crow = nullptr;
    t.m = m;
    t.n = n;
    t.K = NB;
    t.alpha = -1.f;
    t.beta = 1.f;
    t.C = P.a + (size_t)q * N + q;
    t.ldc = N;
    t.m0 = t.n0 = 0;
    // V (rows q.., columns p..p+NB-1 of A) times Y^T (rows q.. of Y)
    t.A = P.a + (size_t)q * N + p;
    t.lda = N;
    t.transA = 0;
    t.B = W.Y + (size_t)q * NB;
    t.ldb = NB;
    t.transB = 1;
    hipLaunchKernelGGL(gb_trail_kernel<MODE>, dim3(tiles), dim3(GT), 0, s, t, tn);
    // X (rows q..) times U (rows p..p+NB-1 of A, columns q..)
    t.A = W.X + (size_t)q * NB;
    t.lda = NB;
    t.B = P.a + (size_t)p * N + q;
    t.ldb = N;
    t.transB = 0;
    hipLaunchKernelGGL(gb_trail_kernel<MODE>, dim3(tiles), dim3(GT), 0, s, t, tn);
}

// columns left to the one-workgroup kernel: PACO_GRID_SWITCH, else the autotuner's table
// for the (M, N) bucket, else SWITCH_DEFAULT (clamped to NB .. 512)
static int grid_switch(int M, int N)
{
    const char *e = getenv("PACO_GRID_SWITCH");
    const int v = e ? atoi(e) : tuned_grid_switch(M, N, SWITCH_DEFAULT);
    return v < NB ? NB : v > 512 ? 512 : v;
}

// first column of the one-workgroup finish: panels of NB columns while more than the
// switch count of columns remain
static int grid_tail_start(int M, int N)
{
    const int sw = grid_switch(M, N);
    int p = 0;
    while (N - p > sw) p += NB;
    return p;
}

// Workgroups asked of the resident kernel (bidiag_res.hip): grid^2 for a tile grid the
// caller gives (a share of the chip), else PACO_RES_P, else 128 (the kernel takes more when
// the matrix needs more registers).  Its time per column is bound by the latency of the
// two exchanges, which more workgroups do not shorten: 128 and 256 measure within 5 % of
// each other from 1100 x 1000 to 4096 x 1024, and 128 is 12 % faster at 2048 x 2048.
static int res_request(int grid)
{
    if (grid > 0) return grid * grid;
    const char *e = getenv("PACO_RES_P");
    return e ? std::max(1, atoi(e)) : 128;
}

// First panel boundary from which the trailing matrix fits the registers of the chip
// (the resident kernel finishes it); -1: it never does, or PACO_GRID_RES=0.
static int res_start(int M, int N, int preq)
{
    const char *e = getenv("PACO_GRID_RES");
    if (e && e[0] == '0') return -1;
    for (int p = 0; N - p >= 2; p += NB)
        if (bidiag_res_workgroups(M - p, N - p, preq) > 0) return p;
    return -1;
}

bool bidiag_grid_supported(int M, int N)
{
    if (res_start(M, N, res_request(0)) >= 0) return true;
    const int p = grid_tail_start(M, N);
    return bidiag_wg_lds_bytes(M - p, N - p, 1024) <= bidiag_wg_lds_limit();
}

hipError_t launch_bidiag_grid(const ProbDesc &P, void *ws, hipStream_t s, int grid, bool no_wait)
{
    const int M = P.M, N = P.N;
    GridWs W;
    const size_t panel_ws = carve(M, N, static_cast<char *>(ws), &W);
    // panels up to p_end, then the resident kernel (from p_res) or the one-workgroup kernel;
    // no_wait: no kernel whose workgroups wait on each other (panels two launches per
    // column, then the one-workgroup kernel), for a solve that timed out beside other work
    const int preq = res_request(grid);
    const int p_res = no_wait ? -1 : res_start(M, N, preq);
    const int p_end = p_res >= 0 ? p_res : grid_tail_start(M, N);
    if (p_res < 0 && bidiag_wg_lds_bytes(M - p_end, N - p_end, 1024) > bidiag_wg_lds_limit())
        return hipErrorInvalidValue;
    // tile grid: the caller's, else PACO_GRID, else the autotuner's table, else 16 (4 .. GMAX);
    // PACO_GRID_PERSIST=0: two launches per column
    const char *genv = getenv("PACO_GRID");
    int G = grid > 0 ? grid : genv ? atoi(genv) : tuned_grid_tiles(M, N, 16);
    G = G < 4 ? 4 : G > GMAX ? GMAX : G;
    const char *penv = getenv("PACO_GRID_PERSIST");
    const bool persist = !no_wait && !(penv && penv[0] == '0');
    static int cus = -1;
    static long long timeout = 0;
    static size_t maxdyn = 0;
    if (cus < 0) {
        int dev = 0, khz = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        timeout = 2000LL * khz; // 2 s per barrier wait
        maxdyn = allow_dynamic_lds((const void *)gb_panel_kernel);
        maxdyn = std::min(maxdyn, allow_dynamic_lds((const void *)gb_col_kernel));
        maxdyn = std::min(maxdyn, allow_dynamic_lds((const void *)gb_row_kernel));
    }
    // at most GBT rows and TW_MAX columns per tile (one row / column per thread)
    const int GR = std::min(GMAX, std::max(G, (M + GBT - 1) / GBT));
    const int GC = std::min(GMAX, std::max(G, (N + TW_MAX - 1) / TW_MAX));
    if (p_end > 0) hipLaunchKernelGGL(gb_sync_reset_kernel, dim3(1), dim3(256), 0, s, W.sync, W.flags, 1);
    int p = 0;
    while (p < p_end) {
        PanelArgs a;
        a.P = P;
        a.W = W;
        a.p = p;
        a.gr = GR;
        a.gc = GC;
        a.th = (M - p + GR - 1) / GR;
        a.tw = (N - p + GC - 1) / GC;
        a.timeout = timeout;
        if (a.th > GBT || a.tw > TW_MAX) return hipErrorInvalidValue;
        const size_t lds = sizeof(float) * panel_dyn_floats_host(a.th, a.tw);
        if (lds > maxdyn) return hipErrorInvalidValue;
        if (persist) {
            if (p > 0) hipLaunchKernelGGL(gb_sync_reset_kernel, dim3(1), dim3(256), 0, s, W.sync, W.flags, 0);
            // PACO_GRID_PARTS caps the workgroups (tests of the several-tiles-per-workgroup path)
            const char *penv2 = getenv("PACO_GRID_PARTS");
            const int cap = penv2 ? std::max(1, atoi(penv2)) : MAX_PARTS;
            const int grid = std::min(std::min(std::min(GR * GC, std::max(cus, 1)), MAX_PARTS), cap);
            hipLaunchKernelGGL(gb_panel_kernel, dim3(grid), dim3(GBT), lds, s, a);
        } else {
            for (int j = 0; j < NB; ++j) {
                hipLaunchKernelGGL(gb_col_kernel, dim3(GR * GC), dim3(GBT), lds, s, a, j);
                hipLaunchKernelGGL(gb_row_kernel, dim3(GR * GC), dim3(GBT), lds, s, a, j);
            }
            hipLaunchKernelGGL(gb_col_kernel, dim3(GR * GC), dim3(GBT), lds, s, a, NB);
        }
        const int mode = gemm_get_mode();
        if (mode == 2)
            launch_trail<2>(P, W, p + NB, p, s);
        else if (mode == 1)
            launch_trail<1>(P, W, p + NB, p, s);
        else
            launch_trail<0>(P, W, p + NB, p, s);
        hipLaunchKernelGGL(gb_restore_kernel, dim3(1), dim3(64), 0, s, P.a, N, p + NB, (const float *)P.e);
        p += NB;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // finish the trailing (M-p) x (N-p) matrix in the registers of the chip, or with one
    // workgroup
    ProbList pl;
    pl.ptr = nullptr;
    pl.count = 1;
    pl.single = P;
    pl.single.M = M - p;
    pl.single.N = N - p;
    pl.single.ld = N;
    pl.single.a = P.a + (size_t)p * N + p;
    pl.single.d = P.d + p;
    pl.single.e = P.e + p;
    pl.single.tauq = P.tauq + p;
    pl.single.taup = P.taup + p;
    if (p_res >= 0) return launch_bidiag_res(pl.single, static_cast<char *>(ws) + panel_ws, s, preq);
    return launch_bidiag_wg(pl, nullptr, 1, bidiag_wg_lds_limit(), s, 1024);
}

} // namespace paco
