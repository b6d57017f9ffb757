// Golub-Kahan upper bidiagonalization and formation of Q, P^T on gfx950.
//
// Reference: upper_bidiagonalization.cpp (Process :83-120, Householder :182-283,
// rank-1 applications :426-563, GetUVt :774-893).  The reference streams one
// reflector at a time through 40 vector cores with 4 grid barriers per column; here
//  * bidiag_wg_kernel: one workgroup owns one matrix and fuses, per column, the
//    application of both reflectors with the gathering of the next column's GEMV data,
//    so the trailing matrix is read and written once per column and no grid barrier is
//    needed.  Used for the batched path and small matrices.
//  * Q and P are formed by blocked WY (32 reflectors per block): T factors per block
//    (tfactor_kernel), applied by the strip kernels of qp_gemm.hip or, with FMA, by
//    wy_apply_kernel (each workgroup owns a column strip of the target and applies all
//    blocks to it), replacing the reference's N rank-1 passes with barriers.
#include <algorithm>

#include "paco_common.hpp"
#include "paco_kernels.hpp"

typedef float tf_floatx16 __attribute__((ext_vector_type(16)));

namespace paco {

#define BT 256

// ---------------------------------------------------------------------------
// One workgroup of NT threads per matrix (N <= 512).  Each column costs one streaming
// pass over the trailing rows: while a wavefront holds an updated row in registers it
// also accumulates the next column's Householder data -- that column's entries c_r,
// sum c_r^2, and sum_r c_r A'[r, :] -- so the next left GEMV needs no pass of its own
// (its scaling by 1/v_1 is applied after the reduction):
//   w_{i+1} = tau (A'[i+1, :] + (1/v_1) sum_{r > i+1} c_r A'[r, :]).
// Rows [split, M) are kept in LDS (all rows when the matrix fits), the rest is
// streamed; every reduction runs in an order that does not depend on the split.
// Sums of squares of the Householder vectors are taken in FP64 (see house_params_d).
// LDS: FP64 partial sums [2 NW + 2], cbuf[M] wbuf[N] ubuf[N] rowbuf[N] (spare N) pwr[NW][N]
//      (+ spare), then the cached rows (columns [c0, N), leading dimension (N - c0) | 1).
// ---------------------------------------------------------------------------
// LDS- and global-typed pointers: accesses through them compile to LDS (ds_*) and
// global (global_*) instructions, never to FLAT ones that must wait on both counters.
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(1))) float glb_f32;

constexpr int WG_MAXN = 512; // 8 chunks of 64 columns per lane
// the first LDS-cached row is a multiple of this (every group size of the pass divides it)
constexpr int WG_SPLIT_ALIGN = 8;
// re-pack the cached rows once max(PACO_COMPACT_MIN, rows / PACO_COMPACT_DIV) more fit
#ifndef PACO_COMPACT_MIN
#define PACO_COMPACT_MIN 16
#define PACO_COMPACT_DIV 8
#endif
#ifndef RQ1024_2
#define RQ1024_2 4
#define RQ1024_4 2
#define RQ1024_8 1
#endif

// (the first 2 (2 nt / 64 + 2) floats hold FP64 partial sums and a broadcast slot)
__host__ __device__ inline size_t bidiag_vec_floats(int M, int N, int nt)
{
    return (size_t)M + 5 * (size_t)N + (size_t)(nt / 64) * N + nt + 8 + 2 * (size_t)(2 * (nt / 64) + 2);
}

// rows of A that fit in lds_bytes of dynamic LDS next to the vectors
__host__ __device__ inline int bidiag_lds_rows(int M, int N, int nt, size_t lds_bytes)
{
    const size_t vec = sizeof(float) * bidiag_vec_floats(M, N, nt);
    if (lds_bytes <= vec) return 0;
    const size_t r = (lds_bytes - vec) / (sizeof(float) * (size_t)(N | 1));
    return r > (size_t)M ? M : (int)r;
}

// first LDS-cached row: M - (rows that fit), rounded up to WG_SPLIT_ALIGN (M: none)
__host__ __device__ inline int bidiag_split_row(int M, int N, int nt, size_t lds_bytes)
{
    const int s = M - bidiag_lds_rows(M, N, nt, lds_bytes);
    if (s == 0) return 0;
    const int a = (s + WG_SPLIT_ALIGN - 1) / WG_SPLIT_ALIGN * WG_SPLIT_ALIGN;
    return a < M ? a : M;
}

// rows per group of the fused pass (one DPP reduction in flight per row; the next
// group is loaded while this one is processed), chosen so that the two groups stay in
// registers: every workgroup size runs at 4 wavefronts per SIMD, 128 VGPRs per lane
template <int NT, int C>
struct PassRows {
    static constexpr int RQ = C <= 2 ? RQ1024_2 : C <= 4 ? RQ1024_4 : RQ1024_8;
};

// Chunk c of a lane holds column k + lane + 64 c; with C = ceil((N - k) / 64) only the
// last chunk can reach past column N - 1, so every other chunk is read and written
// without a test, at immediate offsets from one address per row.
//
// Loads the group of RQ rows starting at row r0 into x.  base points at column cb of row
// rbase; rows outside [k, lim) are replaced by the nearest row inside (read, never written
// back).
template <int C, int RQ, typename T>
__device__ inline void pass_load(const T *base, int ld, int rbase, int cb, int r0, int k, int lim, int N, int lane,
                                 float (&x)[RQ][C])
{
    const bool tail = k + lane + 64 * (C - 1) < N;
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        int rr = r0 + q;
        rr = rr < k ? k : rr;
        rr = rr >= lim ? lim - 1 : rr;
        const T *row = base + (size_t)(rr - rbase) * ld + (k - cb) + lane;
#pragma unroll
        for (int c = 0; c < C - 1; ++c) x[q][c] = row[64 * c];
        x[q][C - 1] = tail ? row[64 * (C - 1)] : 0.f;
    }
}

// One group of the fused pass over rows r > i (i = -1: read-only initial pass).  Row update
//   t = A[r, j] - v_r w_j, x_r = tau2 t.u, A'[r, j] = t_j - x_r u_j   (j > i), A'[r, i] = v_r
// then, for column k = i + 1: row k -> rowbuf; rows r > k: cbuf[r] = A'[r, k], sum of
// squares (FP64, lane 0) and the pw partials of this lane (the entry of column k itself
// is never read).  Columns past N - 1 hold zeros (w, u are zero there).
template <int C, int RQ, typename T>
__device__ inline void pass_group(T *base, int ld, int rbase, int cb, int r0, int k, int M, int N, int lane, bool upd,
                                  float tau, float inv, float tau2, const float (&wr)[C], const float (&ur)[C],
                                  float (&x)[RQ][C], lds_f32 *cbuf, lds_f32 *rowbuf, float (&pwp)[C])
{
    const bool tail = k + lane + 64 * (C - 1) < N;
    if (upd) {
        float acc[RQ];
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            const int rr = r0 + q;
            const bool on = rr >= k && rr < M;
            const float cr = cbuf[rr < k ? k : (rr >= M ? M - 1 : rr)];
            const float vr = (on && tau != 0.f) ? cr * inv : 0.f;
            float a = 0.f;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                x[q][c] -= vr * wr[c];
                a += x[q][c] * ur[c];
            }
            acc[q] = a;
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q) acc[q] = wave_sum_dpp(acc[q]) * tau2;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            const int rr = r0 + q;
#pragma unroll
            for (int c = 0; c < C; ++c) x[q][c] -= acc[q] * ur[c];
            if (rr >= k && rr < M) {
                T *row = base + (size_t)(rr - rbase) * ld + (k - cb) + lane;
#pragma unroll
                for (int c = 0; c < C - 1; ++c) row[64 * c] = x[q][c];
                if (tail) row[64 * (C - 1)] = x[q][C - 1];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        const int rr = r0 + q;
        if (rr < k || rr >= M) continue;
        const float cv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x[q][0]), 0));
        if (rr == k) {
            lds_f32 *rb = rowbuf + k + lane;
#pragma unroll
            for (int c = 0; c < C - 1; ++c) rb[64 * c] = x[q][c];
            if (tail) rb[64 * (C - 1)] = x[q][C - 1];
        } else {
            if (lane == 0) cbuf[rr] = cv;
#pragma unroll
            for (int c = 0; c < C; ++c) pwp[c] += cv * x[q][c];
        }
    }
}

// Pairwise sum of the NW values p[0], p[stride], ..., p[(NW - 1) stride] (NW a power of two),
// in a fixed order, so every thread that sums the same values gets the same result.
template <int NW, typename T, typename P>
__device__ inline T tree_sum(const P *p, int stride)
{
    T v[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w, p += stride) v[w] = *p;
#pragma unroll
    for (int h = NW / 2; h > 0; h >>= 1)
#pragma unroll
        for (int w = 0; w < h; ++w) v[w] += v[w + h];
    return v[0];
}

// True in the wavefront that arrives last at this point of the workgroup (every one of the
// NW wavefronts passes it once, then a barrier follows); that wavefront also resets the LDS
// counter, which is next used after the barrier.  Writes to LDS made before the call by any
// wavefront are visible to the last one after it.
template <int NW>
__device__ inline bool arrive_last(unsigned *cnt, int lane)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    unsigned old = 0;
    if (lane == 0) old = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    old = __builtin_amdgcn_readfirstlane(old);
    if (old != NW - 1) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane == 0) *cnt = 0;
    return true;
}

// Sum of squares (FP64) of the LDS floats p[lo, hi) by one wavefront: lane l takes entries
// lo + l, lo + l + 64, ..., then a wave sum (all 64 lanes active); the order depends on
// lo and hi alone.
__device__ inline double wave_sumsq_lds(const lds_f32 *p, int lo, int hi, int lane)
{
    double s = 0.0;
    for (int r = lo + lane; r < hi; r += 64) {
        const double c = p[r];
        s += c * c;
    }
    return wave_sum_dpp_d(s);
}

// Fused pass over rows r > i with C = ceil((N - i - 1) / 64) chunks per row.  Rows are
// taken in groups of RQ at multiples of RQ, group g by wavefront g % NW, so that every
// wavefront visits its rows in increasing order whatever the LDS split (the reductions
// do not depend on it) and no group straddles the split (a multiple of RQ below M).
// Wave partials of pw go to pwr[wave]; column k's entries c_r go to cbuf (their sum of
// squares is formed from there by the wavefront that arrives last).
template <int NT, int C>
__device__ __attribute__((always_inline)) inline void fused_pass_c(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int c0, int split, int M, int N, int i,
                                      float tau, float inv, float tau2, float inv2, lds_f32 *cbuf,
                                      const lds_f32 *wbuf, const lds_f32 *ybuf, lds_f32 *rowbuf, lds_f32 *pwr)
{
    constexpr int NW = NT / 64, RQ = PassRows<NT, C>::RQ;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, k = i + 1;
    const bool upd = i >= 0;
    float wr[C], ur[C], pwp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = k + lane + 64 * c;
        wr[c] = (upd && j < N) ? wbuf[j] : 0.f;
        // u of the row reflector from y = row i - w: unit at column k, y / v1 beyond
        ur[c] = !(upd && j < N) ? 0.f : j == k ? 1.f : (tau2 != 0.f) ? ybuf[j] * inv2 : ybuf[j];
        pwp[c] = 0.f;
    }
    const int g0 = k / RQ;
    const int gf = g0 + ((wv - g0 % NW) % NW + NW) % NW; // first group of this wavefront
    int g = gf;
    float xa[RQ][C], xb[RQ][C];
    // rows in global memory: [k, split)
    if (RQ * g < split) pass_load<C, RQ>(G, ldg, 0, 0, RQ * g, k, split, N, lane, xa);
    for (; RQ * g < split; g += NW) {
        if (RQ * (g + NW) < split) pass_load<C, RQ>(G, ldg, 0, 0, RQ * (g + NW), k, split, N, lane, xb);
        pass_group<C, RQ>(G, ldg, 0, 0, RQ * g, k, M, N, lane, upd, tau, inv, tau2, wr, ur, xa, cbuf, rowbuf, pwp);
#pragma unroll
        for (int q = 0; q < RQ; ++q)
#pragma unroll
            for (int c = 0; c < C; ++c) xa[q][c] = xb[q][c];
    }
    // rows in LDS: [max(k, split), M)
    if (RQ * g < M) pass_load<C, RQ>(L, ldl, split, c0, RQ * g, k, M, N, lane, xa);
    for (; RQ * g < M; g += NW) {
        if (RQ * (g + NW) < M) pass_load<C, RQ>(L, ldl, split, c0, RQ * (g + NW), k, M, N, lane, xb);
        pass_group<C, RQ>(L, ldl, split, c0, RQ * g, k, M, N, lane, upd, tau, inv, tau2, wr, ur, xa, cbuf, rowbuf,
                          pwp);
#pragma unroll
        for (int q = 0; q < RQ; ++q)
#pragma unroll
            for (int c = 0; c < C; ++c) xa[q][c] = xb[q][c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = k + lane + 64 * c;
        if (c < C - 1 || j < N) pwr[(size_t)wv * N + j] = pwp[c];
    }
}

// ---------------------------------------------------------------------------
// Packed form of the pass for trailing widths <= 256: a row takes LW = 16 or 32 lanes
// (lane = slot LW + sub), so one wave instruction covers R = 64 / LW rows and the
// per-row work (reflector coefficient, row address, reduction, broadcast of c_r) is paid
// once per R rows instead of once per row.  Lane (slot, sub) holds columns
// k + sub + LW c, c < CP.  A group is RQP instructions of R rows: rows r0 + q R + slot.
// Group sizes R RQP divide WG_SPLIT_ALIGN, so groups never straddle the LDS split and
// every lane visits its rows in increasing order whatever the split.
// ---------------------------------------------------------------------------
// sum over the LW lanes of a row, the same bits in every lane of it
template <int LW>
__device__ inline float row_sum_dpp(float v)
{
#define PACO_ROW_DPP(v, ctrl) __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false))
    v += PACO_ROW_DPP(v, 0xB1);  // quad_perm [1,0,3,2]
    v += PACO_ROW_DPP(v, 0x4E);  // quad_perm [2,3,0,1]
    v += PACO_ROW_DPP(v, 0x141); // row_half_mirror
    v += PACO_ROW_DPP(v, 0x140); // row_mirror: every lane holds its 16-lane sum
#undef PACO_ROW_DPP
    if (LW == 32) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401f)); // lane ^ 16
    return v;
}

// the value of sub-lane 0 of this lane's row
template <int LW>
__device__ inline float row_first(float v)
{
    // swizzle bitmask mode within 32 lanes: source lane = lane & and_mask
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), LW == 16 ? 0x10 : 0x00));
}

template <int LW, int CP, int RQP, typename T>
__device__ inline void passp_load(const T *base, int ld, int rbase, int cb, int r0, int k, int lim, int N, int sub,
                                  int slot, float (&x)[RQP][CP])
{
    constexpr int R = 64 / LW;
    const bool tail = k + sub + LW * (CP - 1) < N;
#pragma unroll
    for (int q = 0; q < RQP; ++q) {
        int rr = r0 + q * R + slot;
        rr = rr < k ? k : rr;
        rr = rr >= lim ? lim - 1 : rr;
        const T *row = base + (size_t)(rr - rbase) * ld + (k - cb) + sub;
#pragma unroll
        for (int c = 0; c < CP - 1; ++c) x[q][c] = row[LW * c];
        x[q][CP - 1] = tail ? row[LW * (CP - 1)] : 0.f;
    }
}

template <int LW, int CP, int RQP, typename T>
__device__ inline void passp_group(T *base, int ld, int rbase, int cb, int r0, int k, int M, int N, int sub, int slot,
                                   bool upd, float tau, float inv, float tau2, const float (&wr)[CP],
                                   const float (&ur)[CP], float (&x)[RQP][CP], lds_f32 *cbuf, lds_f32 *rowbuf,
                                   float (&pwp)[CP])
{
    constexpr int R = 64 / LW;
    const bool tail = k + sub + LW * (CP - 1) < N;
    if (upd) {
        float acc[RQP];
#pragma unroll
        for (int q = 0; q < RQP; ++q) {
            const int rr = r0 + q * R + slot;
            const bool on = rr >= k && rr < M;
            const float cr = cbuf[rr < k ? k : (rr >= M ? M - 1 : rr)];
            const float vr = (on && tau != 0.f) ? cr * inv : 0.f;
            float a = 0.f;
#pragma unroll
            for (int c = 0; c < CP; ++c) {
                x[q][c] -= vr * wr[c];
                a += x[q][c] * ur[c];
            }
            acc[q] = a;
        }
#pragma unroll
        for (int q = 0; q < RQP; ++q) acc[q] = row_sum_dpp<LW>(acc[q]) * tau2;
#pragma unroll
        for (int q = 0; q < RQP; ++q) {
            const int rr = r0 + q * R + slot;
#pragma unroll
            for (int c = 0; c < CP; ++c) x[q][c] -= acc[q] * ur[c];
            if (rr >= k && rr < M) {
                T *row = base + (size_t)(rr - rbase) * ld + (k - cb) + sub;
#pragma unroll
                for (int c = 0; c < CP - 1; ++c) row[LW * c] = x[q][c];
                if (tail) row[LW * (CP - 1)] = x[q][CP - 1];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RQP; ++q) {
        const int rr = r0 + q * R + slot;
        const float cv = row_first<LW>(x[q][0]);
        if (rr == k) {
            lds_f32 *rb = rowbuf + k + sub;
#pragma unroll
            for (int c = 0; c < CP - 1; ++c) rb[LW * c] = x[q][c];
            if (tail) rb[LW * (CP - 1)] = x[q][CP - 1];
        } else if (rr > k && rr < M) {
            if (sub == 0) cbuf[rr] = cv;
#pragma unroll
            for (int c = 0; c < CP; ++c) pwp[c] += cv * x[q][c];
        }
    }
}

// Packed fused pass over rows r > i; the same contract as fused_pass_c.  Groups of
// RG = R RQP rows at multiples of RG, group g by wavefront g % NW.  The R slots' partials
// of pw are added (slot pairs, then pairs of pairs) before the wave's partial goes to pwr.
template <int NT, int LW, int CP, int RQP>
__device__ __attribute__((always_inline)) inline void fused_pass_p(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int c0,
                                                                     int split, int M, int N, int i, float tau,
                                                                     float inv, float tau2, float inv2, lds_f32 *cbuf,
                                                                     const lds_f32 *wbuf, const lds_f32 *ybuf,
                                                                     lds_f32 *rowbuf, lds_f32 *pwr)
{
    constexpr int NW = NT / 64, R = 64 / LW, RG = R * RQP;
    static_assert(WG_SPLIT_ALIGN % RG == 0, "a group must not straddle the LDS split");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, k = i + 1;
    const int sub = lane & (LW - 1), slot = lane / LW;
    const bool upd = i >= 0;
    float wr[CP], ur[CP], pwp[CP];
#pragma unroll
    for (int c = 0; c < CP; ++c) {
        const int j = k + sub + LW * c;
        wr[c] = (upd && j < N) ? wbuf[j] : 0.f;
        ur[c] = !(upd && j < N) ? 0.f : j == k ? 1.f : (tau2 != 0.f) ? ybuf[j] * inv2 : ybuf[j];
        pwp[c] = 0.f;
    }
    const int g0 = k / RG;
    const int gf = g0 + ((wv - g0 % NW) % NW + NW) % NW;
    int g = gf;
    float xa[RQP][CP], xb[RQP][CP];
    if (RG * g < split) passp_load<LW, CP, RQP>(G, ldg, 0, 0, RG * g, k, split, N, sub, slot, xa);
    for (; RG * g < split; g += NW) {
        if (RG * (g + NW) < split) passp_load<LW, CP, RQP>(G, ldg, 0, 0, RG * (g + NW), k, split, N, sub, slot, xb);
        passp_group<LW, CP, RQP>(G, ldg, 0, 0, RG * g, k, M, N, sub, slot, upd, tau, inv, tau2, wr, ur, xa, cbuf,
                                 rowbuf, pwp);
#pragma unroll
        for (int q = 0; q < RQP; ++q)
#pragma unroll
            for (int c = 0; c < CP; ++c) xa[q][c] = xb[q][c];
    }
    if (RG * g < M) passp_load<LW, CP, RQP>(L, ldl, split, c0, RG * g, k, M, N, sub, slot, xa);
    for (; RG * g < M; g += NW) {
        if (RG * (g + NW) < M) passp_load<LW, CP, RQP>(L, ldl, split, c0, RG * (g + NW), k, M, N, sub, slot, xb);
        passp_group<LW, CP, RQP>(L, ldl, split, c0, RG * g, k, M, N, sub, slot, upd, tau, inv, tau2, wr, ur, xa, cbuf,
                                 rowbuf, pwp);
#pragma unroll
        for (int q = 0; q < RQP; ++q)
#pragma unroll
            for (int c = 0; c < CP; ++c) xa[q][c] = xb[q][c];
    }
#pragma unroll
    for (int c = 0; c < CP; ++c) {
        float v = pwp[c];
        if (LW == 16) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401f)); // slot ^ 1
        v += __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, __float_as_int(v)));  // slot ^ (R / 2)
        const int j = k + sub + LW * c;
        if (slot == 0 && (c < CP - 1 || j < N)) pwr[(size_t)wv * N + j] = v;
    }
}

// rows per slot per group of the packed pass: with two, the kernels spill registers (128
// VGPRs per lane at 4 wavefronts per SIMD)
#ifndef RQP16_4
#define RQP16_4 1 // 16 lanes per row, CP <= 4
#define RQP16_8 1 // 16 lanes per row, CP 5..8
#define RQP32_8 1 // 32 lanes per row, CP 5..8
#endif

template <int NT>
__device__ __attribute__((always_inline)) inline void fused_pass(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int c0, int split, int M, int N, int i,
                                    float tau, float inv, float tau2, float inv2, lds_f32 *cbuf,
                                    const lds_f32 *wbuf, const lds_f32 *ybuf, lds_f32 *rowbuf, lds_f32 *pwr)
{
#define PACO_PASS(CC)                                                                                              \
    return fused_pass_c<NT, CC>(G, ldg, L, ldl, c0, split, M, N, i, tau, inv, tau2, inv2, cbuf, wbuf, ybuf, rowbuf, \
                                pwr)
#define PACO_PASSP(LW, CP, RQP)                                                                                   \
    return fused_pass_p<NT, LW, CP, RQP>(G, ldg, L, ldl, c0, split, M, N, i, tau, inv, tau2, inv2, cbuf, wbuf, ybuf, \
                                         rowbuf, pwr)
    const int w = N - i - 1; // trailing width, columns [i + 1, N)
    if (w <= 128) {          // 16 lanes per row
        switch ((w + 15) >> 4) {
        case 1: PACO_PASSP(16, 1, RQP16_4);
        case 2: PACO_PASSP(16, 2, RQP16_4);
        case 3: PACO_PASSP(16, 3, RQP16_4);
        case 4: PACO_PASSP(16, 4, RQP16_4);
        case 5: PACO_PASSP(16, 5, RQP16_8);
        case 6: PACO_PASSP(16, 6, RQP16_8);
        case 7: PACO_PASSP(16, 7, RQP16_8);
        default: PACO_PASSP(16, 8, RQP16_8);
        }
    }
    if (w <= 256) { // 32 lanes per row
        switch ((w + 31) >> 5) {
        case 5: PACO_PASSP(32, 5, RQP32_8);
        case 6: PACO_PASSP(32, 6, RQP32_8);
        case 7: PACO_PASSP(32, 7, RQP32_8);
        default: PACO_PASSP(32, 8, RQP32_8);
        }
    }
    switch ((w + 63) >> 6) {
    case 5: PACO_PASS(5);
    case 6: PACO_PASS(6);
    case 7: PACO_PASS(7);
    default: PACO_PASS(8);
    }
#undef PACO_PASSP
#undef PACO_PASS
}

// element (r, j) of the matrix whose rows [split, M) are in LDS from column c0 on (j >= c0)
__device__ inline void store_elem(glb_f32 *G, int ldg, lds_f32 *L, int ldl, int c0, int split, int r, int j, float v)
{
    if (r < split)
        G[(size_t)r * ldg + j] = v;
    else
        L[(size_t)(r - split) * ldl + (j - c0)] = v;
}

// 4 wavefronts per SIMD whatever NT: up to 4 workgroups of 256 threads or 2 of 512 share
// a CU when their LDS allows it
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void bidiag_wg_kernel(ProbList pl,
                                                                                               const int *order,
                                                                                               unsigned lds_bytes)
{
    constexpr int NW = NT / 64;
    const int pid = order ? order[blockIdx.x] : blockIdx.x;
    const ProbDesc P = pl.get(pid);
    const int M = P.M, N = P.N, ldg = P.ld ? P.ld : P.N;
    extern __shared__ float smem[];
    lds_f32 *lsm = (lds_f32 *)smem;
    // FP64: [NW] sum-of-squares partials of the pass, [NW] of the row reflector, [1] x1 of
    // the row reflector
    double *dred = reinterpret_cast<double *>(smem);
    lds_f32 *sm = lsm + 2 * (2 * NW + 2);
    lds_f32 *cbuf = sm, *wbuf = cbuf + M, *ubuf = wbuf + N, *rowbuf = ubuf + N;
    lds_f32 *pwr = rowbuf + 2 * N;
    // reflector parameters published by the last-arriving wavefront: [0..2] beta, tau, 1/v1
    // of the column reflector, [4..6] of the row reflector, [7] the arrival counter
    const size_t poff = 2 * (2 * NW + 2) + (size_t)M + 4 * (size_t)N + (size_t)NW * N;
    float *prm = smem + poff;
    unsigned *wcnt = reinterpret_cast<unsigned *>(smem + poff + 7);
    // rows [split, M) of A, columns [c0, N), leading dimension ldl, after all the vectors
    const size_t vecf = bidiag_vec_floats(M, N, NT);
    lds_f32 *L = lsm + vecf;
    const int cap = lds_bytes / sizeof(float) > vecf ? (int)(lds_bytes / sizeof(float) - vecf) : 0;
    int split = bidiag_split_row(M, N, NT, lds_bytes), ldl = N | 1, c0 = 0;
    glb_f32 *G = (glb_f32 *)P.a;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int r = split + wv; r < M; r += NW)
        for (int c = lane; c < N; c += 64) L[(size_t)(r - split) * ldl + c] = G[(size_t)r * ldg + c];
    if (tid == 0) *wcnt = 0;
    __syncthreads();

    // pass i = -1 is the read-only gathering pass of column 0; pass i >= 0 applies the
    // reflectors of column i to rows > i and gathers column i + 1.  Per column two
    // barriers: after the pass; after the row reflector's partial sums (and y, w).
    // The wavefront that reaches a barrier last forms the reflector's FP64 parameters from
    // the partial sums (always in the same order) and publishes them in LDS: the other
    // wavefronts, which wait at the barrier anyway, do not repeat the FP64 section.
    constexpr int JPT = (WG_MAXN + NT - 1) / NT; // columns of row i per thread
    float tau = 0.f, inv = 1.f, tau2 = 0.f, inv2 = 1.f;
    for (int i = -1;;) {
        // As the trailing matrix narrows, more of its rows fit in LDS at leading dimension
        // (N - i) | 1: once at least max(16, 1/8 of them) more rows fit, the cached rows go
        // back to A and rows [split', M), columns [i, N) are cached instead (split' a
        // multiple of WG_SPLIT_ALIGN, never above the first live row's multiple).  The pass
        // visits rows in an order that does not depend on the split, so the results do not
        // change; only fewer rows are streamed.  The decision depends on (i, split) alone,
        // so every thread takes it alike.
        if (i >= 0 && split > i + 1) {
            const int ldl2 = (N - i) | 1, rows2 = min(M, cap / ldl2);
            const int lo = (i + 1) / WG_SPLIT_ALIGN * WG_SPLIT_ALIGN;
            const int s2 = max(lo, (M - rows2 + WG_SPLIT_ALIGN - 1) / WG_SPLIT_ALIGN * WG_SPLIT_ALIGN);
            if (split - s2 >= max(PACO_COMPACT_MIN, rows2 / PACO_COMPACT_DIV)) {
                // row i and the entries stored since the last barrier are visible to all
                __syncthreads();
                for (int r = split + wv; r < M; r += NW)
                    for (int c = c0 + lane; c < N; c += 64)
                        G[(size_t)r * ldg + c] = L[(size_t)(r - split) * ldl + (c - c0)];
                __syncthreads();
                split = s2;
                ldl = ldl2;
                c0 = i;
                for (int r = split + wv; r < M; r += NW)
                    for (int c = c0 + lane; c < N; c += 64)
                        L[(size_t)(r - split) * ldl + (c - c0)] = G[(size_t)r * ldg + c];
                __syncthreads();
            }
        }
        fused_pass<NT>(G, ldg, L, ldl, c0, split, M, N, i, tau, inv, tau2, inv2, cbuf, wbuf, ubuf, rowbuf, pwr);
        ++i;
        // ---- column reflector of column i: the last-arriving wavefront sums c_r^2 over the
        // rows r > i from cbuf (every wavefront wrote its rows' entries before arriving)
        if (arrive_last<NW>(wcnt, lane)) {
            const double sg = wave_sumsq_lds(cbuf, i + 1, M, lane);
            float b, t, v;
            house_params_d(rowbuf[i], sg, &b, &t, &v);
            if (lane == 0) {
                prm[0] = b;
                prm[1] = t;
                prm[2] = v;
            }
        }
        __syncthreads();
        const float beta = prm[0];
        tau = prm[1];
        inv = prm[2];
        if (tid == 0) {
            store_elem(G, ldg, L, ldl, c0, split, i, i, beta);
            P.d[i] = beta;
            P.tauq[i] = tau;
        }
        // the column reflector's vector into column i below the diagonal (read by no pass)
        for (int r = i + 1 + tid; r < M; r += NT)
            store_elem(G, ldg, L, ldl, c0, split, r, i, (tau != 0.f) ? cbuf[r] * inv : cbuf[r]);
        if (i == N - 1) break;
        // ---- w = tau (row i + (1/v1) pw) and y = row i - w on this thread's columns
        float yv[JPT];
#pragma unroll
        for (int t = 0; t < JPT; ++t) {
            const int j = i + 1 + tid + t * NT;
            float y = 0.f;
            if (j < N) {
                const float a = tree_sum<NW, float>(pwr + j, N);
                const float wj = (tau != 0.f) ? tau * (rowbuf[j] + inv * a) : 0.f;
                wbuf[j] = wj;
                y = rowbuf[j] - wj;
                ubuf[j] = y;
                if (j == i + 1) dred[2 * NW] = (double)y;
            }
            yv[t] = y;
        }
        // ---- row reflector of row i: the last-arriving wavefront sums y_j^2 over j > i + 1
        // from ubuf
        if (arrive_last<NW>(wcnt, lane)) {
            const double sigma2 = wave_sumsq_lds(ubuf, i + 2, N, lane);
            float b, t, v;
            house_params_d(dred[2 * NW], sigma2, &b, &t, &v);
            if (lane == 0) {
                prm[4] = b;
                prm[5] = t;
                prm[6] = v;
            }
        }
        __syncthreads();
        const float beta2 = prm[4];
        tau2 = prm[5];
        inv2 = prm[6];
#pragma unroll
        for (int t = 0; t < JPT; ++t) {
            const int j = i + 1 + tid + t * NT;
            if (j < N) store_elem(G, ldg, L, ldl, c0, split, i, j, j == i + 1 ? beta2 : (tau2 != 0.f) ? yv[t] * inv2 : yv[t]);
        }
        if (tid == 0) {
            P.e[i] = beta2;
            P.taup[i] = tau2;
        }
        // no barrier before the next pass: it reads y (ubuf) and w (wbuf), published before
        // the barrier above, and forms u itself; this section reads nothing the pass writes
    }
    __syncthreads();
    for (int r = split + wv; r < M; r += NW)
        for (int c = c0 + lane; c < N; c += 64) G[(size_t)r * ldg + c] = L[(size_t)(r - split) * ldl + (c - c0)];
}

// ---------------------------------------------------------------------------
// One wavefront per matrix, the whole matrix in LDS (batched path, matrices whose rows
// and one column vector fit the LDS of one workgroup).  With a single wavefront no
// workgroup barrier is needed: row k, the reflector vectors w and u and the partial
// sums of the next left GEMV live in registers (a lane holds the columns lane + 64 c,
// c < 8), every reduction is a DPP wave sum, and the rows of the trailing matrix are
// visited four at a time so that four reductions are in flight.  Per column k:
//   (x1, sum c_r^2)          -> column reflector (beta, tau, 1/v1)
//   w = tau (row_k + (1/v1) sum_r c_r A'[r, :]),  y = row_k - w -> row reflector u
//   rows r > k: t = A'[r, :] - v_r w, A''[r, :] = t - tau2 (t.u) u, A''[r, k] = v_r,
//   and column k + 1 gathered on the way (c_r, sum c_r^2, sum c_r A''[r, :]).
// LDS: rows [M][N|1], cbuf[M].
// ---------------------------------------------------------------------------
constexpr int WV_R = 4;

__host__ __device__ inline size_t bidiag_wave_lds_bytes(int M, int N)
{
    return sizeof(float) * ((size_t)M * (size_t)(N | 1) + (size_t)M);
}

// value of the lane that holds column j (uniform j) of a row held as x[c]
template <int C>
__device__ inline float col_of(const float (&x)[C], int j)
{
    float v = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (c == (j >> 6)) v = x[c];
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j & 63));
}

// C = ceil(N / 64) chunks of 64 columns per row
template <int C>
__device__ inline void bidiag_wave_body(const ProbDesc &P, float *sm)
{
    const int M = uni(P.M), N = uni(P.N), ldg = uni(P.ld ? P.ld : P.N);
    const int lane = threadIdx.x & 63, ldl = N | 1;
    float *A = P.a;
    float *Ls = sm, *cbuf = sm + (size_t)M * ldl;
    for (int r = 0; r < M; ++r)
        for (int j = lane; j < N; j += 64) Ls[(size_t)r * ldl + j] = A[(size_t)r * ldg + j];
    __builtin_amdgcn_wave_barrier();

    // lane columns: only the last chunk can reach past column N - 1
    const bool last_ok = lane + 64 * (C - 1) < N;
    float rowk[C], pwp[C], w[C], u[C];
    double ss = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        rowk[c] = 0.f;
        pwp[c] = 0.f;
        w[c] = 0.f;
        u[c] = 0.f;
    }
    float tau = 0.f, inv = 1.f, tau2 = 0.f;
    // k = -1: gathering pass of column 0; k >= 0: rows > k take the reflectors of
    // column k and column k + 1 is gathered
    for (int k = -1;;) {
        const int kn = k + 1;
        ss = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) pwp[c] = 0.f;
        const int c0 = k < 0 ? 0 : (k >> 6); // chunks left of c0 are final
        for (int r0 = kn; r0 < M; r0 += WV_R) {
            float x[WV_R][C];
#pragma unroll
            for (int q = 0; q < WV_R; ++q) {
                const int r = r0 + q < M ? r0 + q : M - 1; // a row past the end re-reads row M - 1
                const float *row = Ls + (size_t)r * ldl;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    float v = 0.f;
                    if (c >= c0 && (c < C - 1 || last_ok)) v = row[lane + 64 * c];
                    x[q][c] = v;
                }
            }
            if (k >= 0) {
                float acc[WV_R], colv[WV_R];
#pragma unroll
                for (int q = 0; q < WV_R; ++q) {
                    const float cr = cbuf[r0 + q < M ? r0 + q : M - 1];
                    const float vr = (tau != 0.f) ? cr * inv : 0.f;
                    colv[q] = (tau != 0.f) ? vr : cr;
                    float a = 0.f;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        x[q][c] -= vr * w[c];
                        a += x[q][c] * u[c];
                    }
                    acc[q] = a;
                }
#pragma unroll
                for (int q = 0; q < WV_R; ++q) acc[q] = wave_sum_dpp(acc[q]) * tau2;
                const bool lk = lane == (k & 63);
#pragma unroll
                for (int q = 0; q < WV_R; ++q)
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        x[q][c] -= acc[q] * u[c];
                        if (c == (k >> 6) && lk) x[q][c] = colv[q];
                    }
#pragma unroll
                for (int q = 0; q < WV_R; ++q) {
                    if (r0 + q >= M) break;
                    float *row = Ls + (size_t)(r0 + q) * ldl;
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (c >= c0 && (c < C - 1 || last_ok)) row[lane + 64 * c] = x[q][c];
                }
            }
#pragma unroll
            for (int q = 0; q < WV_R; ++q) {
                const int r = r0 + q;
                if (r >= M) break;
                const float cv = col_of<C>(x[q], kn);
                if (r == kn) {
#pragma unroll
                    for (int c = 0; c < C; ++c) rowk[c] = x[q][c];
                } else {
                    if (lane == 0) cbuf[r] = cv;
                    ss += (double)cv * cv;
#pragma unroll
                    for (int c = 0; c < C; ++c) pwp[c] += (lane + 64 * c > kn) ? cv * x[q][c] : 0.f;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        k = kn;
        // ---- column reflector of column k (row k in rowk, column entries in cbuf)
        float beta;
        house_params_d(col_of<C>(rowk, k), ss, &beta, &tau, &inv);
        if (lane == 0) {
            P.d[k] = beta;
            P.tauq[k] = tau;
        }
        if (k == N - 1) {
            if (lane == 0) Ls[(size_t)k * ldl + k] = beta;
            for (int r = k + 1 + lane; r < M; r += 64)
                Ls[(size_t)r * ldl + k] = (tau != 0.f) ? cbuf[r] * inv : cbuf[r];
            break;
        }
        // ---- w, the updated row k and its row reflector
        double part = 0.0;
        float y[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int j = lane + 64 * c;
            const bool on = j > k && (c < C - 1 || last_ok);
            w[c] = (on && tau != 0.f) ? tau * (rowk[c] + inv * pwp[c]) : 0.f;
            y[c] = on ? rowk[c] - w[c] : 0.f;
            part += (j > k + 1) ? (double)y[c] * y[c] : 0.0;
        }
        const double sigma2 = wave_sum(part);
        float beta2, inv2;
        house_params_d(col_of<C>(y, k + 1), sigma2, &beta2, &tau2, &inv2);
        float *rowkp = Ls + (size_t)k * ldl;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int j = lane + 64 * c;
            const bool in = c < C - 1 || last_ok;
            u[c] = j == k + 1 ? 1.f : (j > k + 1 && in) ? ((tau2 != 0.f) ? y[c] * inv2 : y[c]) : 0.f;
            if (j >= k && in) rowkp[j] = j == k ? beta : j == k + 1 ? beta2 : u[c];
        }
        if (lane == 0) {
            P.e[k] = beta2;
            P.taup[k] = tau2;
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int r = 0; r < M; ++r)
        for (int j = lane; j < N; j += 64) A[(size_t)r * ldg + j] = Ls[(size_t)r * ldl + j];
}

template <int C>
__global__ __launch_bounds__(64) void bidiag_wave_kernel(ProbList pl, const int *order)
{
    const int pid = order ? order[blockIdx.x] : blockIdx.x;
    const ProbDesc P = pl.get(pid);
    extern __shared__ float sm[];
    bidiag_wave_body<C>(P, sm);
}

// ---------------------------------------------------------------------------
// Reflector accessor for blocked formation.  side 0 (Q): reflector i acts on rows
// i..M-1, unit at row i, tail A[r][i].  side 1 (P): reflector i (0..N-2) acts on
// coordinates i+1..N-1, unit at i+1, tail A[i][c].  Length L = M or N.
// ---------------------------------------------------------------------------
__device__ inline float refl_elem(const ProbDesc &P, int side, int i, int r)
{
    const int unit = i + side;
    if (r < unit) return 0.f;
    if (r == unit) return 1.f;
    return side == 0 ? P.a[(size_t)r * P.N + i] : P.a[(size_t)i * P.N + r];
}
__device__ inline int refl_count(const ProbDesc &P, int side) { return side == 0 ? P.N : P.N - 1; }
__device__ inline int refl_len(const ProbDesc &P, int side) { return side == 0 ? P.M : P.N; }
__device__ inline float refl_tau(const ProbDesc &P, int side, int i) { return side == 0 ? P.tauq[i] : P.taup[i]; }

// T factor (forward, columnwise; H_b ... H_{b+nb-1} = I - V T V^T) for every block of
// 32 reflectors.  One workgroup per (problem, block); grid mapping through offsets that
// carry the block -> problem table (task_of).
__global__ __launch_bounds__(BT) void tfactor_kernel(ProbList pl, const int *blk_off, int side)
{
    const int pid = task_of(blk_off, pl.count, blockIdx.x);
    const int b = blockIdx.x - off_at(blk_off, pid);
    const ProbDesc P = pl.get(pid);
    const int nref = refl_count(P, side), L = refl_len(P, side);
    const int i0 = b * 32, nb = min(32, nref - i0);
    if (nb <= 0) return;
    __shared__ float Vs[64][33];
    __shared__ float Gp[3][32][33];
    __shared__ float G[32][33];
    __shared__ float T[32][33];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l32 = lane & 31, lh = lane >> 5;
    // Gram matrix G = V^T V on the matrix cores: wavefront w takes rows 16 w .. 16 w + 15
    // of every 64-row chunk (the A and B operands are the same rows of V)
    tf_floatx16 acc;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.f;
    // the next chunk's eight entries per thread are loaded while this chunk's products run
    float vn[8];
    auto load_chunk = [&](int r0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int q = tid + e * BT, rr = q / 32, t = q % 32, r = r0 + rr;
            vn[e] = (t < nb && r < L) ? refl_elem(P, side, i0 + t, r) : 0.f;
        }
    };
    load_chunk(i0 + side);
    for (int r0 = i0 + side; r0 < L; r0 += 64) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int q = tid + e * BT;
            Vs[q / 32][q % 32] = vn[e];
        }
        __syncthreads();
        if (r0 + 64 < L) load_chunk(r0 + 64);
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const float v = Vs[16 * w + 2 * kk + lh][l32];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v, v, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[j] holds G[8 (j / 4) + 4 lh + j % 4][l32] of this wavefront's rows
    if (w > 0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) Gp[w - 1][(j >> 2) * 8 + lh * 4 + (j & 3)][l32] = acc[j];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int r = (j >> 2) * 8 + lh * 4 + (j & 3);
            G[r][l32] = ((acc[j] + Gp[0][r][l32]) + Gp[1][r][l32]) + Gp[2][r][l32];
        }
        __builtin_amdgcn_wave_barrier();
        // T[0:t, t] = -tau_t * T[0:t, 0:t] * G[0:t, t]: row i of the product needs only row i
        // of T, so lane i < 32 keeps its row in registers and forms it alone (G read as
        // broadcasts; LDS accesses of one wavefront complete in order).  Entries left of the
        // diagonal are zero, so the sums over q >= i add the same terms in the same order
        // as sums started at q = i.
        const float taul = lane < nb ? refl_tau(P, side, i0 + lane) : 0.f;
        float tr[32];
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            const float tt = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(taul), t));
            float s = 0.f;
#pragma unroll
            for (int q = 0; q < t; ++q) s += tr[q] * G[q][t];
            tr[t] = t >= nb ? 0.f : (lane < t ? -tt * s : (lane == t ? tt : 0.f));
        }
        if (lane < 32) {
#pragma unroll
            for (int t = 0; t < 32; ++t) T[lane][t] = tr[t];
        }
    }
    __syncthreads();
    float *out = (side == 0 ? P.tq : P.tp) + (size_t)b * 1024;
    for (int q = tid; q < 1024; q += BT) out[q] = T[q / 32][q % 32];
}

// X <- (H_0 ... H_{n-1}) X, X initially the identity: X = Q (side 0, X = u, M x M) or
// X = P (side 1, X = vt, N x N, transposed afterwards).  One workgroup per strip of SW
// columns (SW = 64 for batches; 16 for a single large matrix so every CU has a strip),
// 256 / SW row groups; blocks applied in descending order.
template <int SW>
__global__ __launch_bounds__(BT) void wy_apply_kernel(ProbList pl, const int *strip_off, int side)
{
    constexpr int G = BT / SW;
    const int pid = find_task(strip_off, pl.count, blockIdx.x);
    const int strip = blockIdx.x - off_at(strip_off, pid);
    const ProbDesc P = pl.get(pid);
    const int L = refl_len(P, side), nref = refl_count(P, side);
    float *X = side == 0 ? P.u : P.vt;
    const int ldx = L;
    const int cl = threadIdx.x % SW, grp = threadIdx.x / SW;
    const int col = strip * SW + cl;
    const bool colok = col < L;
    __shared__ float Vc[64][33];
    __shared__ float Ts[32][33];
    __shared__ float Wr[G][32][SW + 1];
    const float *Tbase = side == 0 ? P.tq : P.tp;
    const int nblk = (nref + 31) / 32;
    for (int b = nblk - 1; b >= 0; --b) {
        const int i0 = b * 32, nb = min(32, nref - i0);
        const int rfirst = i0 + side;
        if (strip * SW + SW - 1 < i0 + side) continue; // every column of the strip is untouched
        for (int q = threadIdx.x; q < 1024; q += BT) Ts[q / 32][q % 32] = Tbase[(size_t)b * 1024 + q];
        float w[32];
#pragma unroll
        for (int t = 0; t < 32; ++t) w[t] = 0.f;
        for (int r0 = rfirst; r0 < L; r0 += 64) {
            __syncthreads();
            for (int q = threadIdx.x; q < 64 * 32; q += BT) {
                const int rr = q / 32, t = q % 32, r = r0 + rr;
                Vc[rr][t] = (t < nb && r < L) ? refl_elem(P, side, i0 + t, r) : 0.f;
            }
            __syncthreads();
            if (colok) {
                for (int rr = grp; rr < 64 && r0 + rr < L; rr += G) {
                    const float x = X[(size_t)(r0 + rr) * ldx + col];
#pragma unroll
                    for (int t = 0; t < 32; ++t) w[t] += Vc[rr][t] * x;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 32; ++t) Wr[grp][t][cl] = w[t];
        __syncthreads();
        // W = sum over row groups; W2 = T W (T upper triangular)
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            float a = 0.f;
#pragma unroll
            for (int g = 0; g < G; ++g) a += Wr[g][t][cl];
            w[t] = a;
        }
        float w2[32];
#pragma unroll
        for (int s2 = 0; s2 < 32; ++s2) {
            float acc = 0.f;
#pragma unroll
            for (int t = 0; t < 32; ++t)
                if (t >= s2) acc += Ts[s2][t] * w[t];
            w2[s2] = acc;
        }
        for (int r0 = rfirst; r0 < L; r0 += 64) {
            __syncthreads();
            for (int q = threadIdx.x; q < 64 * 32; q += BT) {
                const int rr = q / 32, t = q % 32, r = r0 + rr;
                Vc[rr][t] = (t < nb && r < L) ? refl_elem(P, side, i0 + t, r) : 0.f;
            }
            __syncthreads();
            if (colok) {
                for (int rr = grp; rr < 64 && r0 + rr < L; rr += G) {
                    float acc = 0.f;
#pragma unroll
                    for (int t = 0; t < 32; ++t) acc += Vc[rr][t] * w2[t];
                    X[(size_t)(r0 + rr) * ldx + col] -= acc;
                }
            }
        }
        __syncthreads();
    }
}

// X = identity (L x L), grid over (problem, group of ROW_GROUP rows) through offsets.
// Rows [r0, r0 + rows) of X = I (L x L, row-major): one contiguous range of floats, written
// in order by the whole workgroup; entry q is 1 where q is a multiple of L + 1.
__global__ __launch_bounds__(BT) void identity_kernel(ProbList pl, const int *grp_off, int side, int rows)
{
    const int pid = find_task(grp_off, pl.count, blockIdx.x);
    const int g = blockIdx.x - off_at(grp_off, pid);
    const ProbDesc P = pl.get(pid);
    const int L = side == 0 ? P.M : P.N;
    float *X = side == 0 ? P.u : P.vt;
    const int r0 = g * rows;
    if (r0 >= L) return;
    const size_t hi = (size_t)min(r0 + rows, L) * L;
    const unsigned L1 = (unsigned)L + 1, step = BT % L1;
    size_t q = (size_t)r0 * L + threadIdx.x;
    unsigned m = (unsigned)(q % L1);
    for (; q < hi; q += BT) {
        gmem(X)[q] = m == 0 ? 1.f : 0.f;
        m += step;
        m = m >= L1 ? m - L1 : m;
    }
}

// In-place transpose of vt (N x N) by 32x32 tile pairs; grid over (problem, tile pair).
// A workgroup swaps `per` consecutive tile pairs (pair_off counts groups of `per` pairs).
__global__ __launch_bounds__(BT) void transpose_vt_kernel(ProbList pl, const int *pair_off, int per)
{
    const int pid = find_task(pair_off, pl.count, blockIdx.x);
    const int g = blockIdx.x - off_at(pair_off, pid);
    const ProbDesc P = pl.get(pid);
    const int N = P.N, nt = (N + 31) / 32, qend = min(g * per + per, nt * (nt + 1) / 2);
    __shared__ float t1[32][33], t2[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 8 rows per pass
    float *X = P.vt;
    for (int q0 = g * per; q0 < qend; ++q0) {
        int q = q0, bi = 0;
        while (q >= nt - bi) {
            q -= nt - bi;
            ++bi;
        }
        const int bj = bi + q;
        __syncthreads(); // the previous pair's reads of the tiles are done
        for (int y = ty; y < 32; y += 8) {
            int r = bi * 32 + y, c = bj * 32 + tx;
            t1[y][tx] = (r < N && c < N) ? X[(size_t)r * N + c] : 0.f;
            r = bj * 32 + y;
            c = bi * 32 + tx;
            t2[y][tx] = (r < N && c < N) ? X[(size_t)r * N + c] : 0.f;
        }
        __syncthreads();
        for (int y = ty; y < 32; y += 8) {
            int r = bi * 32 + y, c = bj * 32 + tx;
            if (r < N && c < N) X[(size_t)r * N + c] = t2[tx][y];
            r = bj * 32 + y;
            c = bi * 32 + tx;
            if (r < N && c < N) X[(size_t)r * N + c] = t1[tx][y];
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t bidiag_wg_lds_bytes(int M, int N, int nt) { return sizeof(float) * bidiag_vec_floats(M, N, nt); }
size_t bidiag_wg_lds_resident_bytes(int M, int N, int nt)
{
    return sizeof(float) * ((size_t)M * (size_t)(N | 1)) + bidiag_wg_lds_bytes(M, N, nt);
}
size_t bidiag_wg_lds_limit()
{
    static size_t lim = 0;
    if (!lim) {
        const size_t a = allow_dynamic_lds((const void *)bidiag_wg_kernel<256>);
        const size_t b = allow_dynamic_lds((const void *)bidiag_wg_kernel<512>);
        const size_t c = allow_dynamic_lds((const void *)bidiag_wg_kernel<1024>);
        lim = a < b ? (a < c ? a : c) : (b < c ? b : c);
    }
    return lim;
}

hipError_t launch_bidiag_wg(ProbList pl, const int *order, int count, size_t lds, hipStream_t s, int nt)
{
    if (count <= 0) return hipSuccess;
    if (lds > bidiag_wg_lds_limit()) return hipErrorInvalidValue;
    if (!pl.ptr && pl.single.N > WG_MAXN) return hipErrorInvalidValue;
    // the per-column vectors (cbuf[M], row vectors) must fit the LDS the launch gets
    if (!pl.ptr && lds < bidiag_wg_lds_bytes(pl.single.M, pl.single.N, nt)) return hipErrorInvalidValue;
    const unsigned lb = (unsigned)lds;
    if (nt == 1024)
        hipLaunchKernelGGL(bidiag_wg_kernel<1024>, dim3(count), dim3(1024), lds, s, pl, order, lb);
    else if (nt == 512)
        hipLaunchKernelGGL(bidiag_wg_kernel<512>, dim3(count), dim3(512), lds, s, pl, order, lb);
    else
        hipLaunchKernelGGL(bidiag_wg_kernel<256>, dim3(count), dim3(256), lds, s, pl, order, lb);
    return hipGetLastError();
}

size_t bidiag_wave_lds(int M, int N) { return bidiag_wave_lds_bytes(M, N); }
using WaveKernel = void (*)(ProbList, const int *);
// a matrix fits the LDS of one workgroup only if N <= 200 (N <= M): C <= 4
static WaveKernel wave_kernel_of(int C)
{
    switch (C) {
    case 1: return bidiag_wave_kernel<1>;
    case 2: return bidiag_wave_kernel<2>;
    case 3: return bidiag_wave_kernel<3>;
    default: return bidiag_wave_kernel<4>;
    }
}
size_t bidiag_wave_lds_limit()
{
    static size_t lim = 0;
    if (!lim) {
        lim = (size_t)-1;
        for (int c = 1; c <= BIDIAG_WAVE_MAXC; ++c)
            lim = std::min(lim, allow_dynamic_lds((const void *)wave_kernel_of(c)));
    }
    return lim;
}
// matrices of one launch share N's chunk count C = ceil(N / 64)
hipError_t launch_bidiag_wave(ProbList pl, const int *order, int count, int C, size_t lds, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    if (C < 1 || C > BIDIAG_WAVE_MAXC || lds > bidiag_wave_lds_limit()) return hipErrorInvalidValue;
    hipLaunchKernelGGL(wave_kernel_of(C), dim3(count), dim3(64), lds, s, pl, order);
    return hipGetLastError();
}

hipError_t launch_tfactor(ProbList pl, const int *blk_off, int nblocks, int side, hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(tfactor_kernel, dim3(nblocks), dim3(BT), 0, s, pl, blk_off, side);
    return hipGetLastError();
}

hipError_t launch_wy_apply(ProbList pl, const int *strip_off, int nstrips, int side, hipStream_t s, int sw)
{
    if (nstrips <= 0) return hipSuccess;
    if (sw == 16)
        hipLaunchKernelGGL(wy_apply_kernel<16>, dim3(nstrips), dim3(BT), 0, s, pl, strip_off, side);
    else
        hipLaunchKernelGGL(wy_apply_kernel<64>, dim3(nstrips), dim3(BT), 0, s, pl, strip_off, side);
    return hipGetLastError();
}

hipError_t launch_identity(ProbList pl, const int *grp_off, int ngroups, int side, hipStream_t s, int rows)
{
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(identity_kernel, dim3(ngroups), dim3(BT), 0, s, pl, grp_off, side, rows);
    return hipGetLastError();
}

hipError_t launch_transpose_vt(ProbList pl, const int *pair_off, int ngroups, hipStream_t s, int per)
{
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_vt_kernel, dim3(ngroups), dim3(BT), 0, s, pl, pair_off, per);
    return hipGetLastError();
}

} // namespace paco
