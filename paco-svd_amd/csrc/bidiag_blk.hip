// Blocked one-workgroup Golub-Kahan bidiagonalization on gfx950 for matrices whose
// rows do not all fit the LDS (N <= 512, M <= 1024).
//
// Reference: upper_bidiagonalization.cpp (Process :83-120, Householder :182-283, rank-1
// applications :426-563).  The reference applies both reflectors of every column to the
// whole trailing matrix; bidiag_wg_kernel (bidiag.hip) does the same in one fused pass,
// reading and writing every trailing row once per column.  Here the columns are taken in
// panels of BK_NB (the dlabrd structure): inside a panel the trailing matrix is only read,
// and the applications of the panel's 2 BK_NB reflectors are deferred as
//   A <- A - U Y^T - X V^T          (U, V: the reflectors, Y, X: their products with A)
// which the pass of the panel's last column applies while it streams every row anyway.
// Per column one pass over the trailing rows r > i computes, from the row as it was at
// the start of the panel,
//   X[r, i]   = taup_i (A[r, :] v_i - U[r, :] s1 - X[r, :] s2)           (s1 = Y^T v_i, s2 = V v_i)
//   c_r       = A[r, i+1] - U[r, :] Y[i+1, :] - X[r, :] V[:, i+1]       (column i+1, updated)
//   pw       += c_r A[r, i+2:]                                          (the next column's GEMV)
// so the trailing matrix is read once per column and written once per panel, against
// read and written once per column without panels.
//
// Opt-in (PACO_BIDIAG_BLOCKED=1).  On the streamed matrices of a 10,000-matrix batch it
// issues a fifth of the per-column kernel's global writes but 30% more vector
// instructions (the panel update reloads BK_NB values of Y and V per element from LDS),
// and that kernel is bound by instruction issue, not by memory traffic: 61 against 48 ms.
//
// Layout of the per-row panel data: row r belongs to wavefront r % NW, which holds
// U[r, :], X[r, :] and c_r in lane r / NW (registers); the rows a wavefront streams are its
// own, visited in increasing order whatever the LDS split.  The products that combine
// rows (U^T u, X^T u for the next Y column; Y^T v, V v for the next X column) are formed
// as per-wavefront partial sums reduced in a fixed order by the last wavefront to arrive,
// which also forms the reflector (FP64 sums of squares, as house_params_d).  Two
// workgroup barriers per column.
//
// LDS: FP64 partial sums, per-wavefront partials of the small products, the reflector
// parameters and small vectors, rowbuf[N] (row i+1 as read by the pass), rbuf[N] (row i
// after its update), Y and V^T of the panel [BK_NB][BK_LD] each, the pass's per-wavefront
// partials pwr[NW][BK_LD], then the cached rows [split, M) of A at leading dimension N | 1.
#include <algorithm>

#include "paco_common.hpp"
#include "paco_kernels.hpp"

namespace paco {

typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(1))) float glb_f32;

constexpr int BK_NT = 1024, BK_NW = BK_NT / 64;
// row stride of the LDS vectors indexed by column (Y, V panels, pwr): a compile-time
// constant, so the panel index and wavefront index are immediate offsets of one address
constexpr int BK_LD = BIDIAG_WG_MAXN;
constexpr int BK_PARTS = 2 * BK_NB;
// prm slots: column reflector (beta, tau, 1/v1), row reflector (beta, tau, 1/v1), the
// diagonal entry c_i of the column being formed, the arrival counter; then the small vectors
enum {
    BK_BETA = 0,
    BK_TAU = 1,
    BK_INV = 2,
    BK_BETA2 = 3,
    BK_TAU2 = 4,
    BK_INV2 = 5,
    BK_CDIAG = 6,
    BK_CNT = 7,
    BK_T1 = 8,
    BK_T2 = BK_T1 + BK_NB,
    BK_S1 = BK_T2 + BK_NB,
    BK_S2 = BK_S1 + BK_NB,
    BK_UI = BK_S2 + BK_NB,
    BK_XI = BK_UI + BK_NB,
    BK_PRM = BK_XI + BK_NB
};

struct BkLayout {
    size_t part, prm, rowbuf, rbuf, yp, vp, pwr, vec; // float offsets
};
// FP64 slots at the start: [NW] pass sums of squares, [NW] row sums of squares, x1 of
// the row reflector, one spare
__host__ __device__ inline BkLayout bk_layout(int N)
{
    BkLayout L;
    size_t o = 2 * (2 * BK_NW + 2);
    L.part = o;
    o += (size_t)BK_NW * BK_PARTS;
    L.prm = o;
    o += BK_PRM;
    (void)N;
    L.rowbuf = o;
    o += BK_LD;
    L.rbuf = o;
    o += BK_LD;
    L.yp = o;
    o += (size_t)BK_NB * BK_LD;
    L.vp = o;
    o += (size_t)BK_NB * BK_LD;
    L.pwr = o;
    o += (size_t)BK_NW * BK_LD;
    L.vec = (o + 3) & ~(size_t)3;
    return L;
}

size_t bidiag_blk_lds_bytes(int M, int N)
{
    (void)M;
    return sizeof(float) * bk_layout(N).vec;
}

// first LDS-cached row for lds_bytes of dynamic LDS (M: none)
__host__ __device__ inline int bk_split_row(int M, int N, size_t lds_bytes)
{
    const size_t vec = sizeof(float) * bk_layout(N).vec;
    if (lds_bytes <= vec) return M;
    const size_t rows = (lds_bytes - vec) / (sizeof(float) * (size_t)(N | 1));
    return rows >= (size_t)M ? 0 : M - (int)rows;
}

bool bidiag_blk_supported(int M, int N)
{
    return M >= N && N >= 1 && N <= BIDIAG_WG_MAXN && M <= BIDIAG_BLK_MAXM &&
           bidiag_blk_lds_bytes(M, N) <= bidiag_wg_lds_limit();
}

// arr[idx] = v for a runtime idx (selects: the array stays in registers)
__device__ inline void bk_set(float (&arr)[BK_NB], int idx, float v)
{
#pragma unroll
    for (int k = 0; k < BK_NB; ++k) arr[k] = (k == idx) ? v : arr[k];
}

__device__ inline float bk_readlane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// row r of the matrix whose rows [split, M) are cached in L (leading dimension ldl)
__device__ inline void bk_store(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int split, int r, int j, float v)
{
    if (r < split)
        G[(size_t)r * ldg + j] = v;
    else
        L[(size_t)(r - split) * ldl + j] = v;
}

// Pairwise sum of p[0], p[stride], ..., p[(NW - 1) stride] in a fixed order.
template <int NW, typename T, typename P>
__device__ inline T bk_tree(const P *p, int stride)
{
    T v[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) v[w] = p[(size_t)w * stride];
#pragma unroll
    for (int h = NW / 2; h > 0; h >>= 1)
#pragma unroll
        for (int w = 0; w < h; ++w) v[w] += v[w + h];
    return v[0];
}

// True in the wavefront that arrives last (see arrive_last in bidiag.hip); it resets the
// counter, whose next use is after the barrier that follows.
__device__ inline bool bk_arrive_last(unsigned *cnt, int lane)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    unsigned old = 0;
    if (lane == 0) old = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    old = __builtin_amdgcn_readfirstlane(old);
    if (old != BK_NW - 1) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane == 0) *cnt = 0;
    return true;
}

enum { BK_GATHER = 0, BK_PASS = 1, BK_UPDATE = 2 };

// rows per group (two groups in flight: the next is loaded while this one is used);
// larger groups in the read-only passes measured slower (register spills)
template <int MODE, int C>
struct BkRows {
    static constexpr int RQ = C <= 2 ? 4 : C <= 4 ? 2 : 1;
};

// Loads rows r_q = w + NW (t + q), q < RQ, columns [k, N), into x (chunk c of a lane is
// column k + lane + 64 c); rows at or past M are replaced by row M - 1 (never used).
template <int C, int RQ>
__device__ inline void bk_load(const glb_f32 *G, int ldg, const lds_f32 *L, int ldl, int split, int M, int N, int k,
                               int w, int t, int lane, float (&x)[RQ][C])
{
    const bool tail = k + lane + 64 * (C - 1) < N;
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        int r = w + BK_NW * (t + q);
        r = r < M ? r : M - 1;
        if (r < split) {
            const glb_f32 *row = G + (size_t)r * ldg + k + lane;
#pragma unroll
            for (int c = 0; c < C - 1; ++c) x[q][c] = row[64 * c];
            x[q][C - 1] = tail ? row[64 * (C - 1)] : 0.f;
        } else {
            const lds_f32 *row = L + (size_t)(r - split) * ldl + k + lane;
#pragma unroll
            for (int c = 0; c < C - 1; ++c) x[q][c] = row[64 * c];
            x[q][C - 1] = tail ? row[64 * (C - 1)] : 0.f;
        }
    }
}

// One pass over the trailing rows r >= k = i + 1 of this wavefront (see the file comment).
// MODE: BK_GATHER (i = -1: gathers column 0 only), BK_PASS (read-only), BK_UPDATE (the
// panel's last column: also applies the panel to every row and writes it back).
// Outputs: X[r, jj] and c_r in the owner lanes, row k in rowbuf, the per-wavefront
// partials pw (pwr), sum_{r>k} c_r^2 (returned, lane 0), and for BK_PASS the partials of
// sum_{r>k} c_r U[r, kk], c_r X[r, kk] (kk <= jj) in part.
template <int MODE, int C>
__device__ __attribute__((always_inline)) inline double
bk_pass_c(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int split, int M, int N, int i, int jj, int nbp, float tau2,
          float inv2, const lds_f32 *rbuf, const lds_f32 *Yp, const lds_f32 *Vp, lds_f32 *rowbuf, lds_f32 *pwr,
          float *part, float *prm, float (&Uown)[BK_NB], float (&Xown)[BK_NB], float &cown, float corr, float pre)
{
    constexpr int RQ = BkRows<MODE, C>::RQ;
    // the wavefront index is uniform: scalar registers for row indices and branches
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), k = i + 1;
    const bool tail = k + lane + 64 * (C - 1) < N;
    float vr[C], pwp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = k + lane + 64 * c;
        pwp[c] = 0.f;
        vr[c] = 0.f;
        if (MODE != BK_GATHER && j < N) vr[c] = j == k ? 1.f : (tau2 != 0.f ? rbuf[j] * inv2 : rbuf[j]);
    }
    double ss = 0.0;
    int t = k > w ? (k - w + BK_NW - 1) / BK_NW : 0; // first row of this wavefront: w + NW t >= k
    float xa[RQ][C], xb[RQ][C];
    if (w + BK_NW * t < M) bk_load<C, RQ>(G, ldg, L, ldl, split, M, N, k, w, t, lane, xa);
    for (; w + BK_NW * t < M; t += RQ) {
        if (w + BK_NW * (t + RQ) < M) bk_load<C, RQ>(G, ldg, L, ldl, split, M, N, k, w, t + RQ, lane, xb);
        float acc[RQ];
        if (MODE != BK_GATHER) {
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                float a = 0.f;
#pragma unroll
                for (int c = 0; c < C; ++c) a += xa[q][c] * vr[c];
                acc[q] = a;
            }
#pragma unroll
            for (int q = 0; q < RQ; ++q) acc[q] = wave_sum_dpp(acc[q]);
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int tq = t + q;
                if (w + BK_NW * tq >= M) continue;
                const float xr = tau2 * (acc[q] - bk_readlane(corr, tq));
                float cur = Xown[0];
#pragma unroll
                for (int kk = 1; kk < BK_NB; ++kk) cur = kk == jj ? Xown[kk] : cur;
                bk_set(Xown, jj, lane == tq ? xr : cur);
                acc[q] = xr;
            }
        }
        if (MODE == BK_UPDATE) {
            // x[r, j] -= sum_kk U[r, kk] Y[j, kk] + X[r, kk] V[kk, j], j >= k
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk) {
                if (kk >= nbp) break;
                float yk[C], vk[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int j = k + lane + 64 * c;
                    const bool on = c < C - 1 || tail;
                    yk[c] = on ? Yp[(size_t)(kk) * BK_LD + j] : 0.f;
                    vk[c] = kk == jj ? vr[c] : (on ? Vp[(size_t)(kk) * BK_LD + j] : 0.f);
                }
#pragma unroll
                for (int q = 0; q < RQ; ++q) {
                    const int tq = t + q < 64 ? t + q : 63;
                    const float uq = bk_readlane(Uown[kk], tq), xq = bk_readlane(Xown[kk], tq);
#pragma unroll
                    for (int c = 0; c < C; ++c) xa[q][c] -= uq * yk[c] + xq * vk[c];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            const int tq = t + q, r = w + BK_NW * tq;
            if (r >= M) continue;
            float cv = bk_readlane(xa[q][0], 0); // column k of the row
            if (MODE == BK_PASS) cv -= bk_readlane(pre, tq) + acc[q];
            cown = lane == tq ? cv : cown;
            if (MODE == BK_UPDATE) {
                if (r < split) {
                    glb_f32 *row = G + (size_t)r * ldg + k + lane;
#pragma unroll
                    for (int c = 0; c < C - 1; ++c) row[64 * c] = xa[q][c];
                    if (tail) row[64 * (C - 1)] = xa[q][C - 1];
                } else {
                    lds_f32 *row = L + (size_t)(r - split) * ldl + k + lane;
#pragma unroll
                    for (int c = 0; c < C - 1; ++c) row[64 * c] = xa[q][c];
                    if (tail) row[64 * (C - 1)] = xa[q][C - 1];
                }
            }
            if (r == k) {
                lds_f32 *rb = rowbuf + k + lane;
#pragma unroll
                for (int c = 0; c < C - 1; ++c) rb[64 * c] = xa[q][c];
                if (tail) rb[64 * (C - 1)] = xa[q][C - 1];
                if (lane == 0) {
                    prm[BK_CDIAG] = cv;
                    if (MODE == BK_PASS) {
#pragma unroll
                        for (int kk = 0; kk < BK_NB; ++kk)
                            if (kk <= jj) {
                                prm[BK_UI + kk] = bk_readlane(Uown[kk], tq);
                                prm[BK_XI + kk] = bk_readlane(Xown[kk], tq);
                            }
                    }
                }
            } else {
                if (lane == 0) ss += (double)cv * cv;
#pragma unroll
                for (int c = 0; c < C; ++c) pwp[c] += cv * xa[q][c];
            }
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q)
#pragma unroll
            for (int c = 0; c < C; ++c) xa[q][c] = xb[q][c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = k + lane + 64 * c;
        if (c < C - 1 || j < N) pwr[(size_t)w * BK_LD + j] = pwp[c];
    }
    if (MODE == BK_PASS) {
        // partials of sum_{r > k} c_r U[r, kk] and c_r X[r, kk], kk <= jj (owner lanes)
        const int orow = w + BK_NW * lane;
        const bool mine = orow > k && orow < M;
#pragma unroll
        for (int kk = 0; kk < BK_NB; ++kk) {
            if (kk > jj) break;
            const float su = wave_sum_dpp(mine ? cown * Uown[kk] : 0.f);
            const float sx = wave_sum_dpp(mine ? cown * Xown[kk] : 0.f);
            if (lane == 0) {
                part[w * BK_PARTS + kk] = su;
                part[w * BK_PARTS + BK_NB + kk] = sx;
            }
        }
    }
    return ss;
}

// State of the column loop that survives from one column to the next.
struct BkState {
    double ss;   // lane 0: sum of c_r^2 over the rows below the next diagonal (this wavefront)
    float cown;  // c_r of the owned row for the next column
    int p, nbp;  // first column and width of the current panel
};

struct BkCtx {
    ProbDesc P;
    glb_f32 *G;
    lds_f32 *L, *rowbuf, *rbuf, *Yp, *Vp, *pwr;
    double *dred;
    float *part, *prm;
    unsigned *wcnt;
    int M, N, ldg, ldl, split, orow;
};

// Columns i, i + 1, ... while the pass of column i spans C chunks per lane (and i < N - 1):
// the column and row reflectors of column i, Y[:, i] and the pass.  Returns the next
// column; N when the last column is done.  Every chunk count has its own loop, so what
// the compiler keeps in registers across columns is that variant's only.
template <int C>
__device__ __attribute__((always_inline)) inline int bk_columns(const BkCtx &X, int i, BkState &S,
                                                                float (&Uown)[BK_NB], float (&Xown)[BK_NB])
{
    constexpr int NW = BK_NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int M = X.M, N = X.N;
    for (;; ++i) {
        const int jj = i - S.p;
        if (lane == 0) X.dred[wv] = S.ss;
        // ---- column reflector of column i; t1 = U^T u, t2 = X^T u over rows >= i
        if (bk_arrive_last(X.wcnt, lane)) {
            const double sg = bk_tree<NW, double>(X.dred, 1);
            float b, t, v;
            house_params_d(X.prm[BK_CDIAG], sg, &b, &t, &v);
            if (lane < BK_NB && lane < jj) {
                const float su = bk_tree<NW, float>(X.part + lane, BK_PARTS);
                const float sx = bk_tree<NW, float>(X.part + BK_NB + lane, BK_PARTS);
                X.prm[BK_T1 + lane] = X.prm[BK_UI + lane] + v * su;
                X.prm[BK_T2 + lane] = X.prm[BK_XI + lane] + v * sx;
            }
            if (lane == 0) {
                X.prm[BK_BETA] = b;
                X.prm[BK_TAU] = t;
                X.prm[BK_INV] = v;
            }
        }
        __syncthreads();
        const float beta = X.prm[BK_BETA], tau = X.prm[BK_TAU], inv = X.prm[BK_INV];
        {
            float u = 0.f;
            if (X.orow > i && X.orow < M) {
                u = tau != 0.f ? S.cown * inv : S.cown;
                bk_store(X.G, X.ldg, X.L, X.ldl, X.split, X.orow, i, u);
            } else if (X.orow == i) {
                u = 1.f;
            }
            bk_set(Uown, jj, u);
        }
        if (tid == 0) {
            bk_store(X.G, X.ldg, X.L, X.ldl, X.split, i, i, beta);
            X.P.d[i] = beta;
            X.P.tauq[i] = tau;
        }
        if (i == N - 1) return N;
        // ---- Y[:, jj] and row i updated by the panel so far (thread tid: column i + 1 + tid)
        const int j = i + 1 + tid;
        float rowv = 0.f;
        double sig = 0.0;
        float sy[BK_NB], sv[BK_NB];
#pragma unroll
        for (int kk = 0; kk < BK_NB; ++kk) {
            sy[kk] = 0.f;
            sv[kk] = 0.f;
        }
        if (j < N) {
            const lds_f32 *yj = X.Yp + j, *vj = X.Vp + j;
            const float a = bk_tree<NW, float>(X.pwr + j, BK_LD);
            float y = X.rowbuf[j] + inv * a;
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk)
                if (kk < jj) y -= yj[kk * BK_LD] * X.prm[BK_T1 + kk] + vj[kk * BK_LD] * X.prm[BK_T2 + kk];
            y *= tau;
            X.Yp[(size_t)jj * BK_LD + j] = y;
            rowv = X.rowbuf[j] - y;
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk)
                if (kk < jj) rowv -= yj[kk * BK_LD] * X.prm[BK_UI + kk] + vj[kk * BK_LD] * X.prm[BK_XI + kk];
            X.rbuf[j] = rowv;
            if (j == i + 1) {
                X.dred[2 * NW] = (double)rowv;
            } else {
                sig = (double)rowv * rowv;
#pragma unroll
                for (int kk = 0; kk < BK_NB; ++kk) {
                    if (kk < jj) sy[kk] = yj[kk * BK_LD] * rowv;
                    if (kk < jj) sv[kk] = vj[kk * BK_LD] * rowv;
                }
                sy[0] = jj == 0 ? y * rowv : sy[0];
#pragma unroll
                for (int kk = 1; kk < BK_NB; ++kk) sy[kk] = kk == jj ? y * rowv : sy[kk];
            }
        }
        // a wavefront without columns (its first is i + 1 + 64 wv) contributes zeros
        if (i + 1 + 64 * wv < N) {
            sig = wave_sum_dpp_d(sig);
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk) {
                if (kk > jj) break;
                sy[kk] = wave_sum_dpp(sy[kk]);
                sv[kk] = wave_sum_dpp(sv[kk]);
            }
        }
        if (lane == 0) {
            X.dred[NW + wv] = sig;
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk)
                if (kk <= jj) {
                    X.part[wv * BK_PARTS + kk] = sy[kk];
                    X.part[wv * BK_PARTS + BK_NB + kk] = sv[kk];
                }
        }
        // ---- row reflector of row i; s1 = Y^T v, s2 = V v over columns > i
        if (bk_arrive_last(X.wcnt, lane)) {
            const double sigma2 = bk_tree<NW, double>(X.dred + NW, 1);
            float b, t, v;
            house_params_d(X.dred[2 * NW], sigma2, &b, &t, &v);
            if (lane < BK_NB && lane <= jj) {
                const float s_y = bk_tree<NW, float>(X.part + lane, BK_PARTS);
                X.prm[BK_S1 + lane] = X.Yp[(size_t)(lane) * BK_LD + i + 1] + v * s_y;
                if (lane < jj) {
                    const float s_v = bk_tree<NW, float>(X.part + BK_NB + lane, BK_PARTS);
                    X.prm[BK_S2 + lane] = X.Vp[(size_t)(lane) * BK_LD + i + 1] + v * s_v;
                }
            }
            if (lane == 0) {
                X.prm[BK_BETA2] = b;
                X.prm[BK_TAU2] = t;
                X.prm[BK_INV2] = v;
            }
        }
        __syncthreads();
        const float beta2 = X.prm[BK_BETA2], tau2 = X.prm[BK_TAU2], inv2 = X.prm[BK_INV2];
        if (j < N) {
            const float v = j == i + 1 ? 1.f : (tau2 != 0.f ? rowv * inv2 : rowv);
            X.Vp[(size_t)jj * BK_LD + j] = v;
            bk_store(X.G, X.ldg, X.L, X.ldl, X.split, i, j, j == i + 1 ? beta2 : v);
        }
        if (tid == 0) {
            X.P.e[i] = beta2;
            X.P.taup[i] = tau2;
        }
        // owner lanes: U[r, :] s1 + X[r, :] s2 and the panel's part of column i + 1
        float corr = 0.f, pre = 0.f;
        {
            const lds_f32 *yn = X.Yp + i + 1, *vn = X.Vp + i + 1;
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk) {
                if (kk <= jj) {
                    corr += Uown[kk] * X.prm[BK_S1 + kk];
                    pre += Uown[kk] * yn[kk * BK_LD];
                }
                if (kk < jj) {
                    corr += Xown[kk] * X.prm[BK_S2 + kk];
                    pre += Xown[kk] * vn[kk * BK_LD];
                }
            }
        }
        if (jj == S.nbp - 1) {
            S.ss = bk_pass_c<BK_UPDATE, C>(X.G, X.ldg, X.L, X.ldl, X.split, M, N, i, jj, S.nbp, tau2, inv2, X.rbuf,
                                           X.Yp, X.Vp, X.rowbuf, X.pwr, X.part, X.prm, Uown, Xown, S.cown, corr, pre);
            S.p += S.nbp;
            S.nbp = min(BK_NB, N - S.p);
#pragma unroll
            for (int kk = 0; kk < BK_NB; ++kk) {
                Uown[kk] = 0.f;
                Xown[kk] = 0.f;
            }
        } else {
            S.ss = bk_pass_c<BK_PASS, C>(X.G, X.ldg, X.L, X.ldl, X.split, M, N, i, jj, S.nbp, tau2, inv2, X.rbuf, X.Yp,
                                         X.Vp, X.rowbuf, X.pwr, X.part, X.prm, Uown, Xown, S.cown, corr, pre);
        }
        // the next column's pass spans fewer chunks: the next loop takes it
        if (((N - i - 2 + 63) >> 6) != C) return i + 1;
    }
}

__global__ __launch_bounds__(BK_NT) __attribute__((amdgpu_waves_per_eu(4))) void bidiag_blk_kernel(
    ProbList pl, const int *order, unsigned lds_bytes)
{
    constexpr int NW = BK_NW;
    const int pid = order ? order[blockIdx.x] : blockIdx.x;
    BkCtx X;
    X.P = pl.get(pid);
    // uniform over the workgroup: scalar registers
    X.M = __builtin_amdgcn_readfirstlane(X.P.M);
    X.N = __builtin_amdgcn_readfirstlane(X.P.N);
    X.ldg = __builtin_amdgcn_readfirstlane(X.P.ld ? X.P.ld : X.P.N);
    const int M = X.M, N = X.N;
    extern __shared__ float smem[];
    const BkLayout Lo = bk_layout(N);
    X.dred = reinterpret_cast<double *>(smem);
    X.part = smem + Lo.part;
    X.prm = smem + Lo.prm;
    X.wcnt = reinterpret_cast<unsigned *>(X.prm + BK_CNT);
    lds_f32 *lsm = (lds_f32 *)smem;
    X.rowbuf = lsm + Lo.rowbuf;
    X.rbuf = lsm + Lo.rbuf;
    X.Yp = lsm + Lo.yp;
    X.Vp = lsm + Lo.vp;
    X.pwr = lsm + Lo.pwr;
    X.L = lsm + Lo.vec;
    X.split = bk_split_row(M, N, lds_bytes);
    X.ldl = N | 1;
    X.G = (glb_f32 *)X.P.a;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int r = X.split + wv; r < M; r += NW)
        for (int c = lane; c < N; c += 64) X.L[(size_t)(r - X.split) * X.ldl + c] = X.G[(size_t)r * X.ldg + c];
    if (tid == 0) *X.wcnt = 0;
    // panel data of the row this lane owns: row orow = wv + NW lane
    X.orow = wv + NW * lane;
    float Uown[BK_NB], Xown[BK_NB];
#pragma unroll
    for (int kk = 0; kk < BK_NB; ++kk) {
        Uown[kk] = 0.f;
        Xown[kk] = 0.f;
    }
    BkState S;
    S.cown = 0.f;
    S.p = 0;
    S.nbp = min(BK_NB, N);
    __syncthreads();
#define PACO_BK_GATHER(CC)                                                                                         \
    S.ss = bk_pass_c<BK_GATHER, CC>(X.G, X.ldg, X.L, X.ldl, X.split, M, N, -1, 0, 0, 0.f, 1.f, X.rbuf, X.Yp, X.Vp,    \
                                    X.rowbuf, X.pwr, X.part, X.prm, Uown, Xown, S.cown, 0.f, 0.f)
    switch ((N + 63) >> 6) {
    case 1: PACO_BK_GATHER(1); break;
    case 2: PACO_BK_GATHER(2); break;
    case 3: PACO_BK_GATHER(3); break;
    case 4: PACO_BK_GATHER(4); break;
    case 5: PACO_BK_GATHER(5); break;
    case 6: PACO_BK_GATHER(6); break;
    case 7: PACO_BK_GATHER(7); break;
    default: PACO_BK_GATHER(8); break;
    }
#undef PACO_BK_GATHER
    for (int i = 0; i < N;) {
        // chunks of the pass of column i (columns i + 1 .. N - 1); 0 only for i = N - 1
        switch ((N - i - 1 + 63) >> 6) {
        case 0:
        case 1: i = bk_columns<1>(X, i, S, Uown, Xown); break;
        case 2: i = bk_columns<2>(X, i, S, Uown, Xown); break;
        case 3: i = bk_columns<3>(X, i, S, Uown, Xown); break;
        case 4: i = bk_columns<4>(X, i, S, Uown, Xown); break;
        case 5: i = bk_columns<5>(X, i, S, Uown, Xown); break;
        case 6: i = bk_columns<6>(X, i, S, Uown, Xown); break;
        case 7: i = bk_columns<7>(X, i, S, Uown, Xown); break;
        default: i = bk_columns<8>(X, i, S, Uown, Xown); break;
        }
    }
    __syncthreads();
    for (int r = X.split + wv; r < M; r += NW)
        for (int c = lane; c < N; c += 64) X.G[(size_t)r * X.ldg + c] = X.L[(size_t)(r - X.split) * X.ldl + c];
}

hipError_t launch_bidiag_blk(ProbList pl, const int *order, int count, size_t lds, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    static size_t maxdyn = allow_dynamic_lds((const void *)bidiag_blk_kernel);
    if (lds > maxdyn || lds > bidiag_wg_lds_limit()) return hipErrorInvalidValue;
    if (!pl.ptr && (!bidiag_blk_supported(pl.single.M, pl.single.N) ||
                    lds < bidiag_blk_lds_bytes(pl.single.M, pl.single.N)))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bidiag_blk_kernel, dim3(count), dim3(BK_NT), lds, s, pl, order, (unsigned)lds);
    return hipGetLastError();
}

} // namespace paco
