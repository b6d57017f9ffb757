// Q and P of one large matrix from its reflectors on the matrix cores (GetUVt,
// upper_bidiagonalization.cpp:774-893).  X = I, then for the blocks of 32 reflectors in
// descending order (rows and columns R0.. of X, Lr = L - R0 of each):
//   V, VT = V T           (qp_extract_kernel: V with its unit diagonal explicit)
//   W_s = V^T X[R0:, R0:] (split over the row range: S partial products in parallel)
//   W = sum_s W_s          (qp_sum_kernel)
//   X[R0:, R0:] -= VT W    (one GEMM, K = 32)
// The columns of X left of R0 are identity columns that the block does not touch.  The
// GEMMs are the 64 x 64 tiles of gemm_tile.hpp (precision setting honoured); their
// descriptors travel as kernel arguments, so nothing is uploaded.
#include <cstdlib>
#include <cstring>

#include "gemm_tile.hpp"
#include "paco_kernels.hpp"

namespace paco {

static inline int cdiv_host(int a, int b) { return (a + b - 1) / b; }

// reflector i of side 0 (Q: unit at row i, tail A[r][i]) or 1 (P: unit at i+1, tail A[i][r])
__device__ inline float qp_refl(const ProbDesc &P, int side, int i, int r)
{
    const int unit = i + side;
    if (r < unit) return 0.f;
    if (r == unit) return 1.f;
    return side == 0 ? P.a[(size_t)r * P.N + i] : P.a[(size_t)i * P.N + r];
}

// V (Lr x 32) and VT = V T (Lr x 32) of block b; eight rows per workgroup
__global__ __launch_bounds__(256) void qp_extract_kernel(ProbDesc P, int side, int b, const float *T, float *V,
                                                         float *VT)
{
    __shared__ float Ts[32][33];
    __shared__ float Vr[8][33];
    const int L = side == 0 ? P.M : P.N, nref = side == 0 ? P.N : P.N - 1;
    const int i0 = 32 * b, nb = min(32, nref - i0), R0 = i0 + side;
    const int tid = threadIdx.x, t = tid & 31, rb = tid >> 5;
    for (int q = tid; q < 1024; q += 256) Ts[q >> 5][q & 31] = T[(size_t)b * 1024 + q];
    const int rr = blockIdx.x * 8 + rb, r = R0 + rr;
    const bool ok = r < L;
    const float v = (ok && t < nb) ? qp_refl(P, side, i0 + t, r) : 0.f;
    Vr[rb][t] = v;
    __syncthreads();
    float acc = 0.f;
    for (int s = 0; s <= t; ++s) acc += Vr[rb][s] * Ts[s][t];
    if (ok) {
        V[(size_t)rr * 32 + t] = v;
        VT[(size_t)rr * 32 + t] = acc;
    }
}

// one GEMM; blockIdx.y = split s of the K range (A, B advance by s * kchunk rows of K,
// C by s * cstride)
template <int MODE>
__global__ __launch_bounds__(GT) void qp_gemm_kernel(GemmDesc G, int kchunk, size_t cstride)
{
    TileCtx t;
    const int s = blockIdx.y, k0 = s * kchunk;
    t.A = G.A + (G.transA ? (size_t)k0 * G.lda : (size_t)k0);
    t.B = G.B + (G.transB ? (size_t)k0 : (size_t)k0 * G.ldb);
    t.C = G.C + (size_t)s * cstride;
    t.brow = nullptr;
    t.crow = nullptr;
    t.lda = G.lda;
    t.ldb = G.ldb;
    t.ldc = G.ldc;
    t.m = G.m;
    t.n = G.n;
    t.K = min(kchunk, G.k - k0);
    t.transA = G.transA;
    t.transB = G.transB;
    t.alpha = G.alpha;
    t.beta = G.beta;
    t.m0 = (blockIdx.x / G.tiles_n) * 64;
    t.n0 = (blockIdx.x % G.tiles_n) * 64;
    if (t.m0 >= t.m || t.n0 >= t.n || t.K <= 0) return;
    tile_gemm<MODE, 16>(t);
}

// W[q] = sum_s Wp[s][q], q < count
__global__ __launch_bounds__(256) void qp_sum_kernel(const float *Wp, int S, size_t stride, float *W, int count)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= count) return;
    float a = 0.f;
    for (int s = 0; s < S; ++s) a += Wp[(size_t)s * stride + q];
    W[q] = a;
}

static hipError_t qp_gemm(const GemmDesc &G, int splits, int kchunk, size_t cstride, hipStream_t st)
{
    const int tiles = cdiv_host(G.m, 64) * G.tiles_n;
    const dim3 grid(tiles, splits);
    const int mode = gemm_get_mode();
    if (mode == 1)
        hipLaunchKernelGGL(qp_gemm_kernel<1>, grid, dim3(GT), 0, st, G, kchunk, cstride);
    else if (mode == 2)
        hipLaunchKernelGGL(qp_gemm_kernel<2>, grid, dim3(GT), 0, st, G, kchunk, cstride);
    else
        hipLaunchKernelGGL(qp_gemm_kernel<0>, grid, dim3(GT), 0, st, G, kchunk, cstride);
    return hipGetLastError();
}

size_t qp_gemm_ws_bytes(int M, int N)
{
    const size_t L = (size_t)(M > N ? M : N);
    return sizeof(float) * (2 * 32 * L + (size_t)(QP_MAX_SPLITS + 1) * 32 * L) + 256;
}

// X (side 0: u = Q, M x M; side 1: vt = P, N x N) for one matrix whose X is already the
// identity and whose T factors are in P.tq / P.tp.
hipError_t launch_qp_gemm(const ProbDesc &P, int side, char *ws, hipStream_t st)
{
    const int L = side == 0 ? P.M : P.N, nref = side == 0 ? P.N : P.N - 1;
    if (nref <= 0) return hipSuccess;
    float *X = side == 0 ? P.u : P.vt;
    const float *T = side == 0 ? P.tq : P.tp;
    const size_t Lmax = (size_t)(P.M > P.N ? P.M : P.N);
    float *V = reinterpret_cast<float *>(ws), *VT = V + 32 * Lmax, *W = VT + 32 * Lmax, *Wp = W + 32 * Lmax;
    hipError_t e;
    for (int b = (nref + 31) / 32 - 1; b >= 0; --b) {
        const int R0 = 32 * b + side, Lr = L - R0;
        hipLaunchKernelGGL(qp_extract_kernel, dim3(cdiv_host(Lr, 8)), dim3(256), 0, st, P, side, b, T, V, VT);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // W_s = V^T X[R0:, R0:]: 32 x Lr, K = Lr split so that about 512 workgroups run
        GemmDesc g;
        std::memset(&g, 0, sizeof(g));
        g.A = V;
        g.lda = 32;
        g.transA = 1;
        g.B = X + (size_t)R0 * L + R0;
        g.ldb = L;
        g.C = Wp;
        g.ldc = Lr;
        g.m = 32;
        g.n = Lr;
        g.k = Lr;
        g.alpha = 1.f;
        g.beta = 0.f;
        g.tiles_n = cdiv_host(Lr, 64);
        int S = 512 / g.tiles_n;
        S = S < 1 ? 1 : S > QP_MAX_SPLITS ? QP_MAX_SPLITS : S;
        const int kchunk = ((cdiv_host(Lr, S) + 15) / 16) * 16;
        S = cdiv_host(Lr, kchunk);
        if ((e = qp_gemm(g, S, kchunk, (size_t)32 * Lr, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(qp_sum_kernel, dim3(cdiv_host(32 * Lr, 256)), dim3(256), 0, st, Wp, S, (size_t)32 * Lr, W,
                           32 * Lr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // X[R0:, R0:] -= VT W
        GemmDesc h;
        std::memset(&h, 0, sizeof(h));
        h.A = VT;
        h.lda = 32;
        h.transA = 0;
        h.B = W;
        h.ldb = Lr;
        h.C = X + (size_t)R0 * L + R0;
        h.ldc = L;
        h.m = Lr;
        h.n = Lr;
        h.k = 32;
        h.alpha = -1.f;
        h.beta = 1.f;
        h.tiles_n = cdiv_host(Lr, 64);
        if ((e = qp_gemm(h, 1, 32, 0, st)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Batched: round r applies, for every problem with more than r blocks, its block
// nblk - 1 - r; the work of a round is dealt through per-problem offsets (off) and each
// problem has a scratch area qp[pid] of 2 x 96 max(M, N) floats (V, VT, W per side).
// ---------------------------------------------------------------------------
struct QpbRound {
    int R0, Lr, L, b;
    float *X, *V, *VT, *W;
    const float *T;
};
__device__ inline QpbRound qpb_round(const ProbDesc &P, float *qp, int side, int round)
{
    QpbRound q;
    q.L = side == 0 ? P.M : P.N;
    const int nref = side == 0 ? P.N : P.N - 1;
    q.b = (nref + 31) / 32 - 1 - round;
    q.R0 = 32 * q.b + side;
    q.Lr = q.L - q.R0;
    q.X = side == 0 ? P.u : P.vt;
    const size_t Lmax = (size_t)(P.M > P.N ? P.M : P.N);
    // each side has its own half of the area: the two sides' rounds may run concurrently
    q.V = qp + (size_t)side * 96 * Lmax;
    q.VT = q.V + 32 * Lmax;
    q.W = q.V + 64 * Lmax;
    q.T = side == 0 ? P.tq : P.tp;
    return q;
}

__global__ __launch_bounds__(256) void qpb_extract_kernel(ProbList pl, float *const *qp, const int *off, int side,
                                                          int round)
{
    __shared__ float Ts[32][33];
    __shared__ float Vr[8][33];
    const int pid = find_task(off, pl.count, blockIdx.x);
    const int blk = blockIdx.x - off_at(off, pid);
    const ProbDesc P = pl.get(pid);
    const QpbRound q = qpb_round(P, qp[pid], side, round);
    const int nref = side == 0 ? P.N : P.N - 1, i0 = 32 * q.b, nb = min(32, nref - i0);
    const int tid = threadIdx.x, t = tid & 31, rb = tid >> 5;
    for (int e = tid; e < 1024; e += 256) Ts[e >> 5][e & 31] = q.T[(size_t)q.b * 1024 + e];
    const int rr = blk * 8 + rb, r = q.R0 + rr;
    const bool ok = rr < q.Lr;
    const float v = (ok && t < nb) ? qp_refl(P, side, i0 + t, r) : 0.f;
    Vr[rb][t] = v;
    __syncthreads();
    float acc = 0.f;
    for (int s2 = 0; s2 <= t; ++s2) acc += Vr[rb][s2] * Ts[s2][t];
    if (ok) {
        q.V[(size_t)rr * 32 + t] = v;
        q.VT[(size_t)rr * 32 + t] = acc;
    }
}

// W (32 x Lr) = V^T X[R0:, R0:]; one 64-column tile of W per workgroup
template <int MODE>
__global__ __launch_bounds__(GT) void qpb_gemm1_kernel(ProbList pl, float *const *qp, const int *off, int side,
                                                       int round)
{
    const int pid = find_task(off, pl.count, blockIdx.x);
    const int tile = blockIdx.x - off_at(off, pid);
    const ProbDesc P = pl.get(pid);
    const QpbRound q = qpb_round(P, qp[pid], side, round);
    TileCtx t;
    t.A = q.V;
    t.lda = 32;
    t.transA = 1;
    t.B = q.X + (size_t)q.R0 * q.L + q.R0;
    t.ldb = q.L;
    t.transB = 0;
    t.C = q.W;
    t.ldc = q.Lr;
    t.brow = t.crow = nullptr;
    t.m = 32;
    t.n = q.Lr;
    t.K = q.Lr;
    t.alpha = 1.f;
    t.beta = 0.f;
    t.m0 = 0;
    t.n0 = tile * 64;
    if (t.n0 >= t.n) return;
    tile_gemm<MODE>(t);
}

// X[R0:, R0:] -= VT W; one 64 x 64 tile per workgroup
template <int MODE>
__global__ __launch_bounds__(GT) void qpb_gemm3_kernel(ProbList pl, float *const *qp, const int *off, int side,
                                                       int round)
{
    const int pid = find_task(off, pl.count, blockIdx.x);
    const int tile = blockIdx.x - off_at(off, pid);
    const ProbDesc P = pl.get(pid);
    const QpbRound q = qpb_round(P, qp[pid], side, round);
    const int tn = (q.Lr + 63) / 64;
    TileCtx t;
    t.A = q.VT;
    t.lda = 32;
    t.transA = 0;
    t.B = q.W;
    t.ldb = q.Lr;
    t.transB = 0;
    t.C = q.X + (size_t)q.R0 * q.L + q.R0;
    t.ldc = q.L;
    t.brow = t.crow = nullptr;
    t.m = q.Lr;
    t.n = q.Lr;
    t.K = 32;
    t.alpha = -1.f;
    t.beta = 1.f;
    t.m0 = (tile / tn) * 64;
    t.n0 = (tile % tn) * 64;
    if (t.m0 >= t.m || t.n0 >= t.n) return;
    tile_gemm<MODE>(t);
}

// ---------------------------------------------------------------------------
// Batched, FP32 matrix cores: one round of one side in a single launch.  A workgroup owns
// a 64-column strip [c0, c0 + 64) of X[R0:, R0:] (the strips of a round touch disjoint
// columns, so they are independent):
//   W  = V^T X[R0:, strip]   (K = Lr in 64-row chunks; wavefront (h, kh) takes column
//                             half h and row half kh of every chunk, the kh halves are
//                             added through LDS)
//   W' = T W                 (T upper triangular: H_b ... H_{b+nb-1} = I - V T V^T)
//   X[R0:, strip] -= V W'
// V is read from the reflectors kept in A (its unit diagonal and the zeros above made
// explicit), staged per 64-row chunk through LDS in a coalesced order for either side.
// ---------------------------------------------------------------------------
// This thread's eight elements of the 64-row chunk of V at row base (Q: a row of V is a
// row segment of A; P: a column of V is a row segment of A; both read coalesced), and
// their place in the LDS tile.
__device__ inline void qpb_load_v(float (&v)[8], const float *A, int N, int side, int i0, int nb, int base, int L)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int q = tid + 256 * e;
        const int i = side == 0 ? (q & 31) : (q >> 6), rr = side == 0 ? (q >> 5) : (q & 63);
        const int r = base + rr, g = i0 + i, unit = g + side;
        float x = 0.f;
        if (i < nb && r < L && r >= unit)
            x = r == unit ? 1.f : (side == 0 ? gmem(A)[(size_t)r * N + g] : gmem(A)[(size_t)g * N + r]);
        v[e] = x;
    }
}
__device__ inline void qpb_store_v(float (*Vs)[33], const float (&v)[8], int side)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int q = tid + 256 * e;
        const int i = side == 0 ? (q & 31) : (q >> 6), rr = side == 0 ? (q >> 5) : (q & 63);
        Vs[rr][i] = v[e];
    }
}

// the lane's 16 entries of column cx of X in the chunk at row base (rows base + r0 + 2 kk)
__device__ inline void qpb_load_x(float (&x)[16], const float *X, int L, int cx, bool cok, int base, int r0)
{
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        const int r = base + r0 + 2 * kk;
        x[kk] = (cok && r < L) ? gmem(X)[(size_t)r * L + cx] : 0.f;
    }
}

// Block b of one side applied to the 64 columns [c0, c0 + 64) of X[R0:, :] (columns
// left of R0 are identity columns the block leaves unchanged; their W columns are zero).
// Every thread of the workgroup calls it; it starts and ends with the workgroup's
// accesses to X ordered by a barrier.
__device__ inline void qpb_strip_block(const ProbDesc &P, int side, int b, int c0, float (*Vs)[33], float (*Ts)[33],
                                       float (*Wp)[68], float (*W2)[68])
{
    const int N = P.N, L = side == 0 ? P.M : P.N, nref = side == 0 ? P.N : P.N - 1;
    const int i0 = 32 * b, nb = min(32, nref - i0), R0 = i0 + side;
    float *X = side == 0 ? P.u : P.vt;
    const float *T = (side == 0 ? P.tq : P.tp) + (size_t)b * 1024;
    const float *A = P.a;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = w & 1, kh = w >> 1;
    const int l32 = lane & 31, lh = lane >> 5;
    const int cx = c0 + 32 * h + l32; // this lane's column of X
    const bool cok = cx < L;
    // the chunk's loads are issued one chunk ahead: V (staged through LDS) and X (after
    // the current chunk's products)
    float vn[8], xn[16];
    qpb_load_v(vn, A, N, side, i0, nb, R0, L);
    qpb_load_x(xn, X, L, cx, cok, R0, 32 * kh + lh);
    for (int q = tid; q < 1024; q += 256) Ts[q >> 5][q & 31] = gmem(T)[q];
    floatx16 acc;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.f;
    for (int base = R0; base < L; base += 64) {
        __syncthreads();
        qpb_store_v(Vs, vn, side);
        if (base + 64 < L) qpb_load_v(vn, A, N, side, i0, nb, base + 64, L);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Vs[32 * kh + 2 * kk + lh][l32], xn[kk], acc, 0, 0, 0);
        // the next chunk's X entries, in flight over the next V store and barrier
        if (base + 64 < L) qpb_load_x(xn, X, L, cx, cok, base + 64, 32 * kh + lh);
    }
    // first chunk of the update, loaded while W is reduced; the update's X entries of a
    // lane are rows base + 32 kh + 8 (j / 4) + 4 lh + j % 4 (the accumulator layout)
    qpb_load_v(vn, A, N, side, i0, nb, R0, L);
    // W = sum of the two row halves
    if (kh == 1) {
#pragma unroll
        for (int j = 0; j < 16; ++j) Wp[(j >> 2) * 8 + lh * 4 + (j & 3)][32 * h + l32] = acc[j];
    }
    __syncthreads();
    if (kh == 0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            float *p = &Wp[(j >> 2) * 8 + lh * 4 + (j & 3)][32 * h + l32];
            *p = acc[j] + *p;
        }
    }
    __syncthreads();
    if (kh == 0) { // W' = T W, column half h
        floatx16 a2;
#pragma unroll
        for (int j = 0; j < 16; ++j) a2[j] = 0.f;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
            a2 = __builtin_amdgcn_mfma_f32_32x32x2f32(Ts[l32][2 * kk + lh], Wp[2 * kk + lh][32 * h + l32], a2, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 16; ++j) W2[(j >> 2) * 8 + lh * 4 + (j & 3)][32 * h + l32] = a2[j];
    }
    for (int base = R0; base < L; base += 64) {
        __syncthreads();
        qpb_store_v(Vs, vn, side);
        if (base + 64 < L) qpb_load_v(vn, A, N, side, i0, nb, base + 64, L);
        float xo[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int r = base + 32 * kh + (j >> 2) * 8 + lh * 4 + (j & 3);
            xo[j] = (cok && r < L) ? gmem(X)[(size_t)r * L + cx] : 0.f;
        }
        __syncthreads();
        floatx16 a3;
#pragma unroll
        for (int j = 0; j < 16; ++j) a3[j] = 0.f;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
            a3 = __builtin_amdgcn_mfma_f32_32x32x2f32(Vs[32 * kh + l32][2 * kk + lh], W2[2 * kk + lh][32 * h + l32], a3, 0,
                                                      0, 0);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int r = base + 32 * kh + (j >> 2) * 8 + lh * 4 + (j & 3);
            if (cok && r < L) gmem(X)[(size_t)r * L + cx] = xo[j] - a3[j];
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void qpb_strip_kernel(ProbList pl, const int *off, int side, int round)
{
    __shared__ float Vs[64][33];
    __shared__ float Ts[32][33];
    __shared__ float Wp[32][68];
    __shared__ float W2[32][68];
    const int pid = find_task(off, pl.count, blockIdx.x);
    const int strip = blockIdx.x - off_at(off, pid);
    const ProbDesc P = pl.get(pid);
    const int L = side == 0 ? P.M : P.N, nref = side == 0 ? P.N : P.N - 1;
    const int b = (nref + 31) / 32 - 1 - round;
    if (b < 0) return;
    const int c0 = 32 * b + side + 64 * strip;
    if (c0 >= L) return;
    qpb_strip_block(P, side, b, c0, Vs, Ts, Wp, W2);
}

size_t qpb_scratch_floats(int M, int N) { return 2 * 96 * (size_t)(M > N ? M : N); }

// one round of one side: extraction blocks, W tiles and X tiles through their offsets
hipError_t launch_qpb_round(ProbList pl, float *const *qp, int side, int round, const int *ext_off, int next,
                            const int *g1_off, int ng1, const int *g3_off, int ng3, hipStream_t st)
{
    if (next <= 0) return hipSuccess;
    const int mode = gemm_get_mode();
    const char *senv = getenv("PACO_QP_STRIP");
    if (mode == 0 && !(senv && senv[0] == '0')) { // FP32 matrix cores: one strip launch per round
        if (ng1 > 0) hipLaunchKernelGGL(qpb_strip_kernel, dim3(ng1), dim3(256), 0, st, pl, g1_off, side, round);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(qpb_extract_kernel, dim3(next), dim3(256), 0, st, pl, qp, ext_off, side, round);
    if (mode == 1) {
        hipLaunchKernelGGL(qpb_gemm1_kernel<1>, dim3(ng1), dim3(GT), 0, st, pl, qp, g1_off, side, round);
        hipLaunchKernelGGL(qpb_gemm3_kernel<1>, dim3(ng3), dim3(GT), 0, st, pl, qp, g3_off, side, round);
    } else if (mode == 2) {
        hipLaunchKernelGGL(qpb_gemm1_kernel<2>, dim3(ng1), dim3(GT), 0, st, pl, qp, g1_off, side, round);
        hipLaunchKernelGGL(qpb_gemm3_kernel<2>, dim3(ng3), dim3(GT), 0, st, pl, qp, g3_off, side, round);
    } else {
        hipLaunchKernelGGL(qpb_gemm1_kernel<0>, dim3(ng1), dim3(GT), 0, st, pl, qp, g1_off, side, round);
        hipLaunchKernelGGL(qpb_gemm3_kernel<0>, dim3(ng3), dim3(GT), 0, st, pl, qp, g3_off, side, round);
    }
    return hipGetLastError();
}

} // namespace paco
