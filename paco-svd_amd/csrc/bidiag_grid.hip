// Golub-Kahan bidiagonalization of one large matrix on the whole chip.
//
// Reference: upper_bidiagonalization.cpp Process :83-120 (one reflector per column,
// rank-1 updates of the whole trailing matrix on 40 cores, 4 grid barriers per column).
// Here: a blocked panel of NB columns (LAPACK dlabrd structure).  Inside a panel the
// trailing matrix is only read (two GEMVs per column) and the panel's reflectors are
// applied once per panel as a rank-2*NB update on the matrix cores.  The workgroups
// form a GR x GC grid of tiles over the trailing matrix.  The Householder scaling of
// each column / row reflector is deferred (partial sums are taken against the
// unscaled vector and scaled on reduction), so every column needs two grid-wide
// synchronisation points, realised as two launches:
//   col kernel (column i): finish x of column i-1 from its GEMV partials (every tile
//     reduces only its own rows), column-i update, norm partials, GEMV A^T v partials
//   row kernel (column i): column-i reflector, y for own columns, row-i update, norm
//     partials, GEMV A u partials
// Below SWITCH columns the remaining trailing matrix is finished by the one-workgroup
// kernel (bidiag_wg_kernel) on the sub-matrix.
#include "gemm_tile.hpp"
#include "paco_kernels.hpp"

namespace paco {

#define GBT 256
constexpr int NB = 32;
// workgroup grid over the trailing matrix of a panel: gr x gc tiles (runtime, at most
// GMAX x GMAX; the workspace holds GMAX partial sums per row / column)
constexpr int GMAX = 32;
constexpr int SWITCH = 384;

struct GridWs {
    float *X, *Y, *wcol, *wrow, *g1p, *g2p, *t1p, *t2p, *t3p, *t4p;
    double *ssp, *rsp; // FP64 partial sums of squares (see house_params_d)
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
    if (W) *W = w;
    return o;
}

size_t bidiag_grid_ws_bytes(int M, int N) { return carve(M, N, nullptr, nullptr); }

struct PanelArgs {
    ProbDesc P;
    GridWs W;
    int p, j, th, tw, finish_only;
    int gr, gc; // tile grid
};

// column reflector t of the panel (unit at row p+t), entry at row r
__device__ inline float Vel(const float *A, int N, int p, int t, int r)
{
    const int c = p + t;
    return r > c ? A[(size_t)r * N + c] : (r == c ? 1.f : 0.f);
}
// row reflector t of the panel (row p+t, unit at column p+t+1), entry at column col
__device__ inline float Uel(const float *A, int N, int p, int t, int col)
{
    const int u = p + t + 1;
    return col > u ? A[(size_t)(p + t) * N + col] : (col == u ? 1.f : 0.f);
}

__device__ inline double gb_block_sum(double v, double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < GBT / 64; ++w) r += red[w];
    return r;
}

// dynamic LDS: xs[th] (x of the previous column), wc[th] (updated column i)
__global__ __launch_bounds__(GBT) void gb_col_kernel(PanelArgs a)
{
    const int GR = a.gr, GC = a.gc;
    const int rb = blockIdx.x / GC, cb = blockIdx.x % GC;
    const ProbDesc &P = a.P;
    const GridWs &W = a.W;
    const int M = P.M, N = P.N, p = a.p, j = a.j, i = p + j;
    float *A = P.a;
    const int r0 = p + rb * a.th, r1 = min(M, r0 + a.th);
    const int c0 = p + cb * a.tw, c1 = min(N, c0 + a.tw);
    extern __shared__ float sm[];
    float *xs = sm, *wc = sm + a.th;
    __shared__ float t3s[NB], t4s[NB], ys[NB], us[NB];
    __shared__ double red[GBT / 64];
    __shared__ float s_tau2, s_inv2, s_beta2;
    const int tid = threadIdx.x;

    if (j > 0) {
        // ---- finish the row reflector and x of column ip = i - 1
        const int ip = i - 1, jp = j - 1;
        if (tid == 0) {
            double sigma2 = 0.0;
            for (int q = 0; q < GC; ++q) sigma2 += W.rsp[q];
            float beta2, tau2, inv2;
            house_params_d(W.wrow[ip + 1], sigma2, &beta2, &tau2, &inv2);
            s_tau2 = tau2;
            s_inv2 = inv2;
            s_beta2 = beta2;
        }
        __syncthreads();
        const float tau2 = s_tau2, inv2 = s_inv2;
        if (tid < NB) {
            const int t = tid;
            float s3 = 0.f, s4 = 0.f;
            for (int q = 0; q < GC; ++q) {
                s3 += W.t3p[q * NB + t];
                s4 += W.t4p[q * NB + t];
            }
            t3s[t] = (t <= jp) ? W.Y[(size_t)(ip + 1) * NB + t] + inv2 * s3 : 0.f;
            t4s[t] = (t < jp) ? Uel(A, N, p, t, ip + 1) + inv2 * s4 : 0.f;
        }
        __syncthreads();
        for (int row = max(r0, i) + tid; row < r1; row += GBT) {
            float g2 = 0.f;
            for (int q = 0; q < GC; ++q) g2 += W.g2p[(size_t)q * M + row];
            float acc = A[(size_t)row * N + ip + 1] + inv2 * g2;
#pragma unroll 8
            for (int t = 0; t <= jp; ++t) acc -= Vel(A, N, p, t, row) * t3s[t];
#pragma unroll 8
            for (int t = 0; t < jp; ++t) acc -= W.X[(size_t)row * NB + t] * t4s[t];
            const float x = tau2 * acc;
            xs[row - r0] = x;
            if (cb == 0) W.X[(size_t)row * NB + jp] = x;
        }
        // the row reflector u_{ip} into A (row ip), its head e_{ip}
        if (rb == 0) {
            for (int col = max(c0, ip + 1) + tid; col < c1; col += GBT) {
                if (col == ip + 1) // unit head while the trailing update reads row ip as U
                    A[(size_t)ip * N + col] = a.finish_only ? 1.f : s_beta2;
                else
                    A[(size_t)ip * N + col] = (tau2 != 0.f) ? W.wrow[col] * inv2 : W.wrow[col];
            }
            if (cb == 0 && tid == 0) {
                P.e[ip] = s_beta2;
                P.taup[ip] = tau2;
            }
        }
        __syncthreads();
    }
    if (a.finish_only) return;

    // ---- column i update for rows >= i of this block (redundant over cb)
    if (tid < NB) {
        ys[tid] = (tid < j) ? W.Y[(size_t)i * NB + tid] : 0.f;
        us[tid] = (tid < j) ? Uel(A, N, p, tid, i) : 0.f;
    }
    __syncthreads();
    double part = 0.0;
    for (int row = max(r0, i) + tid; row < r1; row += GBT) {
        float val = A[(size_t)row * N + i];
#pragma unroll 8
        for (int t = 0; t < j; ++t) {
            const float xv = (t == j - 1) ? xs[row - r0] : W.X[(size_t)row * NB + t];
            val -= Vel(A, N, p, t, row) * ys[t] + xv * us[t];
        }
        wc[row - r0] = val;
        if (cb == 0) W.wcol[row] = val;
        if (row > i) part += (double)val * val;
    }
    const double ss = gb_block_sum(part, red);
    if (cb == 0 && tid == 0) W.ssp[rb] = ss;
    // ---- GEMV1 partials: g1p[rb][col] = sum_{row > i} wc[row] A[row][col], col > i
    const int rs = max(r0, i + 1);
    if ((N & 3) == 0) {
        // float4 column groups x row groups, 8 loads in flight per thread, LDS reduction
        __shared__ float4 red4[GBT];
        const int ca = max(c0, i + 1) & ~3, ncg = (c1 - ca + 3) >> 2;
        int cg = 1;
        while (cg < ncg && cg < GBT) cg <<= 1;
        const int RG = GBT / cg, g = tid / cg, c = tid % cg;
        for (int cgb = 0; cgb < ncg; cgb += cg) {
            const int col = ca + 4 * (cgb + c);
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            if (cgb + c < ncg) {
                int row = rs + g;
                for (; row + 7 * RG < r1; row += 8 * RG) {
                    float4 v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) v[k] = *reinterpret_cast<const float4 *>(A + (size_t)(row + k * RG) * N + col);
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const float w = wc[row + k * RG - r0];
                        acc.x += w * v[k].x;
                        acc.y += w * v[k].y;
                        acc.z += w * v[k].z;
                        acc.w += w * v[k].w;
                    }
                }
                for (; row < r1; row += RG) {
                    const float4 v = *reinterpret_cast<const float4 *>(A + (size_t)row * N + col);
                    const float w = wc[row - r0];
                    acc.x += w * v.x;
                    acc.y += w * v.y;
                    acc.z += w * v.z;
                    acc.w += w * v.w;
                }
            }
            __syncthreads();
            red4[tid] = acc;
            __syncthreads();
            if (g == 0 && cgb + c < ncg) {
                float4 t = red4[c];
                for (int q = 1; q < RG; ++q) {
                    const float4 u = red4[q * cg + c];
                    t.x += u.x;
                    t.y += u.y;
                    t.z += u.z;
                    t.w += u.w;
                }
                const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) // own columns only (the aligned start may reach into the previous tile)
                    if (col + e > i && col + e >= c0 && col + e < c1) W.g1p[(size_t)rb * N + col + e] = tv[e];
            }
        }
    } else {
        for (int col = max(c0, i + 1) + tid; col < c1; col += GBT) {
            float acc = 0.f;
#pragma unroll 8
            for (int row = rs; row < r1; ++row) acc += wc[row - r0] * A[(size_t)row * N + col];
            W.g1p[(size_t)rb * N + col] = acc;
        }
    }
    // t1p[rb][t] = sum_{row > i} V[row][t] wc[row], t2p[rb][t] = sum X[row][t] wc[row], t < j;
    // 32 column lanes x 8 row groups, reduced through LDS
    if (cb == 0) {
        __shared__ float r1s[8][NB + 1], r2s[8][NB + 1];
        const int t = tid % NB, g = tid / NB;
        float s1 = 0.f, s2 = 0.f;
        if (t < j) {
#pragma unroll 4
            for (int row = rs + g; row < r1; row += 8) {
                const float w = wc[row - r0];
                s1 += Vel(A, N, p, t, row) * w;
                s2 += ((t == j - 1) ? xs[row - r0] : W.X[(size_t)row * NB + t]) * w;
            }
        }
        r1s[g][t] = s1;
        r2s[g][t] = s2;
        __syncthreads();
        if (tid < NB) {
            float a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                a1 += r1s[q][tid];
                a2 += r2s[q][tid];
            }
            W.t1p[rb * NB + tid] = a1;
            W.t2p[rb * NB + tid] = a2;
        }
    }
}

// dynamic LDS: yl[tw] (y of column i for own columns), wl[tw] (updated row i)
__global__ __launch_bounds__(GBT) void gb_row_kernel(PanelArgs a)
{
    const int GR = a.gr, GC = a.gc;
    const int rb = blockIdx.x / GC, cb = blockIdx.x % GC;
    const ProbDesc &P = a.P;
    const GridWs &W = a.W;
    const int M = P.M, N = P.N, p = a.p, j = a.j, i = p + j;
    float *A = P.a;
    const int r0 = p + rb * a.th, r1 = min(M, r0 + a.th);
    const int c0 = p + cb * a.tw, c1 = min(N, c0 + a.tw);
    extern __shared__ float sm[];
    float *yl = sm, *wl = sm + a.tw;
    __shared__ float t1s[NB], t2s[NB], xi[NB], vi[NB];
    __shared__ double red[GBT / 64];
    __shared__ float s_tau, s_inv, s_beta;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // ---- column reflector of column i
    if (tid == 0) {
        double sigma = 0.0;
        for (int q = 0; q < GR; ++q) sigma += W.ssp[q];
        float beta, tau, inv;
        house_params_d(W.wcol[i], sigma, &beta, &tau, &inv);
        s_tau = tau;
        s_inv = inv;
        s_beta = beta;
    }
    __syncthreads();
    const float tau = s_tau, inv = s_inv;
    if (cb == 0) {
        for (int row = max(r0, i + 1) + tid; row < r1; row += GBT)
            A[(size_t)row * N + i] = (tau != 0.f) ? W.wcol[row] * inv : W.wcol[row];
        if (rb == 0 && tid == 0) {
            A[(size_t)i * N + i] = s_beta;
            P.d[i] = s_beta;
            P.tauq[i] = tau;
        }
    }
    if (i == N - 1) return;

    if (tid < NB) {
        const int t = tid;
        float s1 = 0.f, s2 = 0.f;
        for (int q = 0; q < GR; ++q) {
            s1 += W.t1p[q * NB + t];
            s2 += W.t2p[q * NB + t];
        }
        t1s[t] = (t < j) ? Vel(A, N, p, t, i) + inv * s1 : 0.f;
        t2s[t] = (t < j) ? W.X[(size_t)i * NB + t] + inv * s2 : 0.f;
        xi[t] = (t < j) ? W.X[(size_t)i * NB + t] : 0.f;
        vi[t] = (t < j) ? Vel(A, N, p, t, i) : (t == j ? 1.f : 0.f);
    }
    __syncthreads();
    // ---- y for own columns, row-i update
    double part = 0.0;
    for (int col = max(c0, i + 1) + tid; col < c1; col += GBT) {
        float g1 = 0.f;
        for (int q = 0; q < GR; ++q) g1 += W.g1p[(size_t)q * N + col];
        float acc = A[(size_t)i * N + col] + inv * g1;
#pragma unroll 8
        for (int t = 0; t < j; ++t) acc -= W.Y[(size_t)col * NB + t] * t1s[t] + Uel(A, N, p, t, col) * t2s[t];
        const float y = tau * acc;
        yl[col - c0] = y;
        if (rb == 0) W.Y[(size_t)col * NB + j] = y;
        float w = A[(size_t)i * N + col] - y * vi[j];
#pragma unroll 8
        for (int t = 0; t < j; ++t) w -= W.Y[(size_t)col * NB + t] * vi[t] + Uel(A, N, p, t, col) * xi[t];
        wl[col - c0] = w;
        if (rb == 0) W.wrow[col] = w;
        if (col > i + 1) part += (double)w * w;
    }
    const double ss = gb_block_sum(part, red);
    if (rb == 0 && tid == 0) W.rsp[cb] = ss;
    // ---- GEMV2 partials: g2p[cb][row] = sum_{col > i+1} A[row][col] wl[col], row > i
    const int cs = max(c0, i + 2);
    // eight rows per wavefront in flight; lanes over the tile's columns
    for (int row = max(r0, i + 1) + 8 * wv; row < r1; row += 8 * (GBT / 64)) {
        float acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = 0.f;
        for (int col = cs + lane; col < c1; col += 64) {
            const float w = wl[col - c0];
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = (row + q < r1) ? A[(size_t)(row + q) * N + col] : 0.f;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] += v[q] * w;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = wave_sum(acc[q]);
        if (lane < 8 && row + lane < r1) {
            float a = acc[0];
#pragma unroll
            for (int q = 1; q < 8; ++q)
                if (lane == q) a = acc[q];
            W.g2p[(size_t)cb * M + row + lane] = a;
        }
    }
    // t3p[cb][t] = sum_{col > i+1} Y[col][t] w (t <= j), t4p[cb][t] = sum U[t][col] w (t < j);
    // 32 lanes over t x 8 column groups, reduced through LDS
    if (rb == 0) {
        __shared__ float r3s[8][NB + 1], r4s[8][NB + 1];
        const int t = tid % NB, g = tid / NB;
        float s3 = 0.f, s4 = 0.f;
#pragma unroll 4
        for (int col = cs + g; col < c1; col += 8) {
            const float w = wl[col - c0];
            if (t < j) {
                s3 += W.Y[(size_t)col * NB + t] * w;
                s4 += Uel(A, N, p, t, col) * w;
            } else if (t == j) {
                s3 += yl[col - c0] * w;
            }
        }
        r3s[g][t] = s3;
        r4s[g][t] = s4;
        __syncthreads();
        if (tid < NB) {
            float a3 = 0.f, a4 = 0.f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                a3 += r3s[q][tid];
                a4 += r4s[q][tid];
            }
            W.t3p[cb * NB + tid] = a3;
            W.t4p[cb * NB + tid] = a4;
        }
    }
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
    tile_gemm<MODE>(t);
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

// first column of the one-workgroup finish: panels of NB columns while more than SWITCH
// columns remain
static int grid_tail_start(int N)
{
    int p = 0;
    while (N - p > SWITCH) p += NB;
    return p;
}

bool bidiag_grid_supported(int M, int N)
{
    const int p = grid_tail_start(N);
    return bidiag_wg_lds_bytes(M - p, N - p, 1024) <= bidiag_wg_lds_limit();
}

hipError_t launch_bidiag_grid(const ProbDesc &P, void *ws, hipStream_t s)
{
    const int M = P.M, N = P.N;
    GridWs W;
    carve(M, N, static_cast<char *>(ws), &W);
    // tile grid: PACO_GRID overrides (8 .. GMAX)
    const char *genv = getenv("PACO_GRID");
    int G = genv ? atoi(genv) : 16;
    G = G < 8 ? 8 : G > GMAX ? GMAX : G;
    const int GR = G, GC = G;
    int p = 0;
    while (N - p > SWITCH) {
        PanelArgs a;
        a.P = P;
        a.W = W;
        a.p = p;
        a.gr = GR;
        a.gc = GC;
        a.th = (M - p + GR - 1) / GR;
        a.tw = (N - p + GC - 1) / GC;
        const size_t lds_col = sizeof(float) * 2 * (size_t)a.th;
        const size_t lds_row = sizeof(float) * 2 * (size_t)a.tw;
        static size_t maxc = allow_dynamic_lds((const void *)gb_col_kernel);
        static size_t maxr = allow_dynamic_lds((const void *)gb_row_kernel);
        if (lds_col > maxc || lds_row > maxr) return hipErrorInvalidValue;
        for (int j = 0; j < NB; ++j) {
            a.j = j;
            a.finish_only = 0;
            hipLaunchKernelGGL(gb_col_kernel, dim3(GR * GC), dim3(GBT), lds_col, s, a);
            hipLaunchKernelGGL(gb_row_kernel, dim3(GR * GC), dim3(GBT), lds_row, s, a);
        }
        a.j = NB;
        a.finish_only = 1;
        hipLaunchKernelGGL(gb_col_kernel, dim3(GR * GC), dim3(GBT), lds_col, s, a);
        const int mode = gemm_get_mode();
        if (mode == 2)
            launch_trail<2>(P, W, p + NB, p, s);
        else
            launch_trail<0>(P, W, p + NB, p, s);
        hipLaunchKernelGGL(gb_restore_kernel, dim3(1), dim3(64), 0, s, P.a, N, p + NB, (const float *)P.e);
        p += NB;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // finish the trailing (M-p) x (N-p) matrix with one workgroup
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
    if (bidiag_wg_lds_bytes(M - p, N - p, 1024) > bidiag_wg_lds_limit()) return hipErrorInvalidValue;
    return launch_bidiag_wg(pl, nullptr, 1, bidiag_wg_lds_limit(), s, 1024);
}

} // namespace paco
