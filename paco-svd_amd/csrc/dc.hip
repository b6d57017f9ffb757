// Bidiagonal divide and conquer on gfx950.
//
// Reference: svd.cpp (BDC::Process :134-195, MergeSubMatrix :225-323, Deflation
// :325-975, SecularEquationSolver :977-1336, MergeSubMatrix_step2 :1337-1594, leaves
// :1596-1864).  The reference runs every merge serially on one vector core.  Here the
// merge tree is processed level by level; every phase of a level is one launch over
// all merges of that level (and, in the batched path, of every matrix):
//   leaves    FP64 one-sided Jacobi: one thread per leaf of <= 3 columns, one
//             wavefront per larger leaf (up to LEAF_MAX columns)
//   deflate   one workgroup per merge: parallel merge-path ranks, scan, compaction
//   gather    one thread per basis column: gathered bases + deflation rotations
//   secular   one wavefront per root, FP64 safeguarded rational iteration
//   lowner    one thread per pole: FP64 Gu-Eisenstat z and output positions
//   vectors   one wavefront per root: FP64 secular vectors, stored FP32
//   gemm      grouped MFMA GEMM (gemm.hip) for left and right vectors
//   finalize  deflated rows, null vector and sigma written in descending order
// Node storage (as svd.cpp:87-133): d[c0..] sigma of the node (descending), qt rows
// c0.. = left vectors, wt rows c0.. = right vectors, confined to the node's diagonal
// block; qt+N*N and wt+N*N hold the gathered bases of the merge in flight.
#include <float.h>

#include "gemm_tile.hpp"
#include "paco_kernels.hpp"

namespace paco {

#define DT 256
#define SRC_POLE0 0x7fffffff

// ---------------------------------------------------------------------------
// leaves (replaces compute_2x3/2x2/1x2/1x1_svd, svd.cpp:1627-1864)
// ---------------------------------------------------------------------------
__device__ void leaf_svd(int n, int m, const double *dd, const double *ee, double *sigma, double *U, double *V)
{
    double W[2][3] = {{0, 0, 0}, {0, 0, 0}}, J[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int r = 0; r < n; r++) {
        W[r][r] = dd[r];
        if (r + 1 < m) W[r][r + 1] = ee[r];
    }
    for (int sweep = 0; sweep < 30; sweep++) {
        int rotated = 0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                double app = 0, aqq = 0, apq = 0;
                for (int r = 0; r < n; r++) {
                    app += W[r][p] * W[r][p];
                    aqq += W[r][q] * W[r][q];
                    apq += W[r][p] * W[r][q];
                }
                if (apq == 0.0 || fabs(apq) <= 1e-17 * sqrt(app * aqq)) continue;
                rotated = 1;
                double zeta = (aqq - app) / (2.0 * apq);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < n; r++) {
                    double wp = W[r][p], wq = W[r][q];
                    W[r][p] = c * wp - s * wq;
                    W[r][q] = s * wp + c * wq;
                }
                for (int r = 0; r < m; r++) {
                    double jp = J[r][p], jq = J[r][q];
                    J[r][p] = c * jp - s * jq;
                    J[r][q] = s * jp + c * jq;
                }
            }
        if (!rotated) break;
    }
    double nrm[3] = {0, 0, 0};
    int ord[3] = {0, 1, 2};
    for (int c = 0; c < m; c++) {
        double s = 0;
        for (int r = 0; r < n; r++) s += W[r][c] * W[r][c];
        nrm[c] = sqrt(s);
    }
    for (int a = 0; a < m; a++)
        for (int b = a + 1; b < m; b++)
            if (nrm[ord[b]] > nrm[ord[a]]) {
                int t = ord[a];
                ord[a] = ord[b];
                ord[b] = t;
            }
    for (int i = 0; i < m; i++)
        for (int r = 0; r < m; r++) V[r * m + i] = J[r][ord[i]];
    for (int i = 0; i < n; i++) {
        double s = nrm[ord[i]];
        sigma[i] = s;
        if (s > 0.0) {
            for (int r = 0; r < n; r++) U[r * n + i] = W[r][ord[i]] / s;
        } else if (i == 0) {
            for (int r = 0; r < n; r++) U[r * n + i] = (r == 0) ? 1.0 : 0.0;
        } else {
            U[0 * n + 1] = -U[1 * n + 0];
            U[1 * n + 1] = U[0 * n + 0];
        }
    }
}

// one thread per leaf of at most 3 columns
__global__ __launch_bounds__(DT) void dc_leaf_kernel(ProbList pl, const int *leaf_off, int nleaf, int ls)
{
    const int l = blockIdx.x * DT + threadIdx.x;
    if (l >= nleaf) return;
    const int prob = find_task(leaf_off, pl.count, l);
    const int ll = l - off_at(leaf_off, prob);
    const ProbDesc P = pl.get(prob);
    const int N = P.N, c0 = leaf_start(N, ll, ls), c1 = leaf_start(N, ll + 1, ls);
    const int m = c1 - c0, n = (c1 == N) ? m : m - 1;
    if (m < 1 || m > 3 || n < 1) return;
    double dd[2] = {0, 0}, ee[2] = {0, 0}, sig[2], U[4], V[9];
    for (int r = 0; r < n; r++) {
        dd[r] = P.d[c0 + r];
        if (r + 1 < m) ee[r] = P.e[c0 + r];
    }
    leaf_svd(n, m, dd, ee, sig, U, V);
    for (int i = 0; i < n; i++) {
        P.d[c0 + i] = (float)sig[i];
        for (int r = 0; r < n; r++) P.qt[(size_t)(c0 + i) * N + c0 + r] = (float)U[r * n + i];
    }
    for (int i = 0; i < m; i++)
        for (int r = 0; r < m; r++) P.wt[(size_t)(c0 + i) * N + c0 + r] = (float)V[r * m + i];
}

// ---------------------------------------------------------------------------
// Leaves of up to LEAF_THREAD_MAX columns: one thread per leaf, the block W and the
// rotations V in registers (FP64, compile-time indices), cyclic-by-rows one-sided Jacobi
// with the same rotation, threshold and ordering rules as the wavefront kernel below.
// ---------------------------------------------------------------------------
constexpr int LEAF_THREAD_MAX = 6;
// Jacobi of the leaves: sweep limit, rotation threshold |w_p . w_q| / (||w_p|| ||w_q||),
// sigma below which a left vector is rebuilt (relative), column norm treated as zero
// (relative to ||B||_F), LDS column stride of the wavefront kernel (doubles)
constexpr int LEAF_SWEEPS = 30;
constexpr double LEAF_TOL = 1e-14;
constexpr double LEAF_SMALL = 1e-8;
constexpr double LEAF_ZERO = 1e-15;
constexpr int LEAF_LP = LEAF_MAX + 1;

template <int MP>
__global__ __launch_bounds__(DT) void dc_leaf_thread_kernel(ProbList pl, const int *leaf_off, int nleaf, int ls)
{
    const int l = blockIdx.x * DT + threadIdx.x;
    if (l >= nleaf) return;
    const int prob = find_task(leaf_off, pl.count, l);
    const int ll = l - off_at(leaf_off, prob);
    const ProbDesc P = pl.get(prob);
    const int N = P.N, c0 = leaf_start(N, ll, ls), c1 = leaf_start(N, ll + 1, ls);
    const int m = c1 - c0, n = (c1 == N) ? m : m - 1;
    if (m < 1 || m > MP || n < 1) return;
    double W[MP][MP], V[MP][MP]; // W[c][r]: column c; V[c][r]
    double fro2 = 0.0;
#pragma unroll
    for (int c = 0; c < MP; ++c)
#pragma unroll
        for (int r = 0; r < MP; ++r) {
            V[c][r] = (c == r && c < m) ? 1.0 : 0.0;
            double w = 0.0;
            if (r < n && c == r) w = P.d[c0 + r];
            if (r < n && c == r + 1 && c < m) w = P.e[c0 + r];
            W[c][r] = w;
            fro2 += w * w;
        }
    const double zero2 = LEAF_ZERO * LEAF_ZERO * fro2;
    for (int sweep = 0; sweep < LEAF_SWEEPS; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < MP - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < MP; ++q) {
                double app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
                for (int r = 0; r < MP; ++r) {
                    app += W[p][r] * W[p][r];
                    aqq += W[q][r] * W[q][r];
                    apq += W[p][r] * W[q][r];
                }
                if (app > zero2 && aqq > zero2 && apq != 0.0 && fabs(apq) > LEAF_TOL * sqrt(app) * sqrt(aqq)) {
                    const double zeta = (aqq - app) / (2.0 * apq);
                    const double t = fabs(zeta) > 1e150
                                         ? 0.5 / zeta
                                         : (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
                    for (int r = 0; r < MP; ++r) {
                        const double wp = W[p][r], wq = W[q][r], vp = V[p][r], vq = V[q][r];
                        W[p][r] = cs * wp - sn * wq;
                        W[q][r] = sn * wp + cs * wq;
                        V[p][r] = cs * vp - sn * vq;
                        V[q][r] = sn * vp + cs * vq;
                    }
                    rotated = true;
                }
            }
        if (!rotated) break;
    }
    // sigma_c = ||w_c||; rank of each column (descending, ties by column)
    double sg[MP];
    int rk[MP];
#pragma unroll
    for (int c = 0; c < MP; ++c) {
        double s2 = 0.0;
#pragma unroll
        for (int r = 0; r < MP; ++r) s2 += W[c][r] * W[c][r];
        sg[c] = c < m ? sqrt(s2) : -1.0; // padding columns rank last
    }
    double smax = 0.0;
#pragma unroll
    for (int c = 0; c < MP; ++c) {
        int k = 0;
#pragma unroll
        for (int c2 = 0; c2 < MP; ++c2) k += (sg[c2] > sg[c]) || (sg[c2] == sg[c] && c2 < c);
        rk[c] = k;
        smax = fmax(smax, sg[c]);
    }
    const double small = LEAF_SMALL * smax;
    // left vectors: w_c / sigma_c; the small ones (in rank order) orthogonalised against
    // the others, or replaced by the first unit vector that is not in their span
#pragma unroll
    for (int c = 0; c < MP; ++c)
        if (sg[c] > small)
#pragma unroll
            for (int r = 0; r < MP; ++r) W[c][r] /= sg[c];
    for (int i = 0; i < n; ++i) {
#pragma unroll
        for (int c = 0; c < MP; ++c) {
            if (rk[c] != i || sg[c] > small) continue;
            double x[MP];
#pragma unroll
            for (int r = 0; r < MP; ++r) x[r] = sg[c] > 0.0 ? W[c][r] / sg[c] : (r == i ? 1.0 : 0.0);
            for (int a = 0; a <= n; ++a) {
                if (a > 0)
#pragma unroll
                    for (int r = 0; r < MP; ++r) x[r] = (r == a - 1) ? 1.0 : 0.0;
#pragma unroll
                for (int pass = 0; pass < 2; ++pass)
#pragma unroll
                    for (int c2 = 0; c2 < MP; ++c2) {
                        if (c2 == c || rk[c2] >= n || (rk[c2] > i && sg[c2] <= small)) continue;
                        double dot = 0.0;
#pragma unroll
                        for (int r = 0; r < MP; ++r) dot += x[r] * W[c2][r];
#pragma unroll
                        for (int r = 0; r < MP; ++r) x[r] -= dot * W[c2][r];
                    }
                double nr = 0.0;
#pragma unroll
                for (int r = 0; r < MP; ++r) nr += x[r] * x[r];
                nr = sqrt(nr);
                if (nr > 1e-3) {
#pragma unroll
                    for (int r = 0; r < MP; ++r) x[r] /= nr;
                    break;
                }
            }
#pragma unroll
            for (int r = 0; r < MP; ++r) W[c][r] = x[r];
        }
    }
#pragma unroll
    for (int c = 0; c < MP; ++c) {
        if (c >= m) continue;
        const int i = rk[c];
        if (i < n) {
            P.d[c0 + i] = (float)sg[c];
#pragma unroll
            for (int r = 0; r < MP; ++r)
                if (r < n) P.qt[(size_t)(c0 + i) * N + c0 + r] = (float)W[c][r];
        }
#pragma unroll
        for (int r = 0; r < MP; ++r)
            if (r < m) P.wt[(size_t)(c0 + i) * N + c0 + r] = (float)V[c][r];
    }
}

// ---------------------------------------------------------------------------
// Leaves of up to LEAF_MAX columns: one wavefront per leaf runs FP64 one-sided (Hestenes)
// Jacobi on the n x m bidiagonal block W (n = m - 1 rows, m for the last leaf).  The
// m2 = m rounded up to even columns are paired by the round-robin (circle) schedule,
// m2 / 2 disjoint pairs per step and every pair once in a sweep of m2 - 1 steps; pair k
// takes lanes k, k + 16, k + 32, k + 48, each summing and rotating every 4th row.  A
// pair rotates (Rutishauser's formulas) when |w_p . w_q| > tol ||w_p|| ||w_q|| and both
// columns are above 1e-15 ||B||_F (a column below that is rounding noise, whose direction
// never settles: an n x (n + 1) leaf has one); sweeps repeat until no pair rotates.  Then sigma_c = ||w_c||, in descending order; the right
// vectors are the accumulated rotations V; the left vectors are w_c / sigma_c, except
// where sigma_c <= 1e-8 sigma_max: those are orthogonalised (twice, in order) against
// the others and, when nothing is left, replaced by unit vectors that are, so the left
// basis is orthonormal also for rank-deficient leaves.
// LDS: W and V column-major, column c at c * LP (LP = LEAF_MAX + 1 doubles against bank
// conflicts between the pairs' columns).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void dc_leaf_wave_kernel(ProbList pl, const int *leaf_off, int ls)
{
    __shared__ double W[LEAF_MAX * LEAF_LP], V[LEAF_MAX * LEAF_LP];
    __shared__ double sg[LEAF_MAX];
    __shared__ int col_of[LEAF_MAX];
    const int l = blockIdx.x, lane = threadIdx.x;
    const int prob = find_task(leaf_off, pl.count, l);
    const int ll = l - off_at(leaf_off, prob);
    const ProbDesc P = pl.get(prob);
    const int N = P.N, c0 = leaf_start(N, ll, ls), c1 = leaf_start(N, ll + 1, ls);
    const int m = c1 - c0, n = (c1 == N) ? m : m - 1;
    if (m < 1 || m > LEAF_MAX || n < 1) return; // uniform over the workgroup
    for (int i = lane; i < LEAF_MAX * LEAF_LP; i += 64) {
        W[i] = 0.0;
        V[i] = 0.0;
    }
    __syncthreads();
    if (lane < m) V[lane * LEAF_LP + lane] = 1.0;
    if (lane < n) {
        W[lane * LEAF_LP + lane] = (double)P.d[c0 + lane];
        if (lane + 1 < m) W[(lane + 1) * LEAF_LP + lane] = (double)P.e[c0 + lane];
    }
    __syncthreads();
    const int m2 = (m + 1) & ~1, np = m2 >> 1, k = lane & 15, sub = lane >> 4;
    double fro2 = 0.0; // ||B||_F^2, the same in every lane
    for (int r = 0; r < n; ++r) {
        const double a = P.d[c0 + r], b = r + 1 < m ? P.e[c0 + r] : 0.f;
        fro2 += a * a + b * b;
    }
    const double zero2 = LEAF_ZERO * LEAF_ZERO * fro2;
    for (int sweep = 0; sweep < LEAF_SWEEPS; ++sweep) {
        int rotated = 0;
        for (int st = 0; st < m2 - 1; ++st) {
            const int p = k == 0 ? st : (st + k) % (m2 - 1);
            const int q = k == 0 ? m2 - 1 : (st - k + m2 - 1) % (m2 - 1);
            const bool act = k < np && p < m && q < m;
            double wp[LEAF_MAX / 4], wq[LEAF_MAX / 4], app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
            for (int j = 0; j < LEAF_MAX / 4; ++j) {
                const int r = sub + 4 * j;
                wp[j] = (act && r < n) ? W[p * LEAF_LP + r] : 0.0;
                wq[j] = (act && r < n) ? W[q * LEAF_LP + r] : 0.0;
                app += wp[j] * wp[j];
                aqq += wq[j] * wq[j];
                apq += wp[j] * wq[j];
            }
            // the four lanes of the pair: the same sums in every one of them
            app += __shfl_xor(app, 16);
            aqq += __shfl_xor(aqq, 16);
            apq += __shfl_xor(apq, 16);
            app += __shfl_xor(app, 32);
            aqq += __shfl_xor(aqq, 32);
            apq += __shfl_xor(apq, 32);
            if (act && app > zero2 && aqq > zero2 && apq != 0.0 && fabs(apq) > LEAF_TOL * sqrt(app) * sqrt(aqq)) {
                const double zeta = (aqq - app) / (2.0 * apq);
                const double t = fabs(zeta) > 1e150 ? 0.5 / zeta
                                                    : (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
#pragma unroll
                for (int j = 0; j < LEAF_MAX / 4; ++j) {
                    const int r = sub + 4 * j;
                    if (r < n) {
                        W[p * LEAF_LP + r] = c * wp[j] - s * wq[j];
                        W[q * LEAF_LP + r] = s * wp[j] + c * wq[j];
                    }
                    if (r < m) {
                        const double vp = V[p * LEAF_LP + r], vq = V[q * LEAF_LP + r];
                        V[p * LEAF_LP + r] = c * vp - s * vq;
                        V[q * LEAF_LP + r] = s * vp + c * vq;
                    }
                }
                rotated = 1;
            }
            __syncthreads();
        }
        if (!__syncthreads_or(rotated)) break;
    }
    // sigma_c = ||w_c||; col_of[i]: the column of rank i (descending, ties by column)
    if (lane < m) {
        double s2 = 0.0;
        for (int r = 0; r < n; ++r) s2 += W[lane * LEAF_LP + r] * W[lane * LEAF_LP + r];
        sg[lane] = sqrt(s2);
    }
    __syncthreads();
    if (lane < m) {
        int rk = 0;
        for (int c = 0; c < m; ++c) rk += (sg[c] > sg[lane]) || (sg[c] == sg[lane] && c < lane);
        col_of[rk] = lane;
    }
    __syncthreads();
    const double small = LEAF_SMALL * sg[col_of[0]];
    for (int x = lane; x < m * m; x += 64) {
        const int i = x / m, r = x - i * m;
        P.wt[(size_t)(c0 + i) * N + c0 + r] = (float)V[col_of[i] * LEAF_LP + r];
    }
    if (lane < n) P.d[c0 + lane] = (float)sg[col_of[lane]];
    for (int x = lane; x < n * n; x += 64) {
        const int i = x / n, r = x - i * n, c = col_of[i];
        if (sg[c] > small) W[c * LEAF_LP + r] /= sg[c];
    }
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        const int c = col_of[i];
        if (sg[c] > small) continue; // uniform: sg and col_of are in LDS
        double x = 0.0;
        if (lane < n) x = sg[c] > 0.0 ? W[c * LEAF_LP + lane] / sg[c] : (lane == i ? 1.0 : 0.0);
        // attempt 0: the column itself; attempt a > 0: the unit vector e_{a-1}
        for (int a = 0; a <= n; ++a) {
            if (a > 0) x = (lane == a - 1) ? 1.0 : 0.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < n; ++j) {
                    const int cj = col_of[j];
                    if (j == i || (j > i && sg[cj] <= small)) continue; // not (yet) in the basis
                    const double u = lane < n ? W[cj * LEAF_LP + lane] : 0.0;
                    x -= wave_sum_dpp_d(x * u) * u;
                }
            const double nr = sqrt(wave_sum_dpp_d(x * x));
            if (nr > 1e-3) {
                x /= nr;
                break;
            }
        }
        __syncthreads();
        if (lane < n) W[c * LEAF_LP + lane] = x;
        __syncthreads();
    }
    for (int x = lane; x < n * n; x += 64) {
        const int i = x / n, r = x - i * n;
        P.qt[(size_t)(c0 + i) * N + c0 + r] = (float)W[col_of[i] * LEAF_LP + r];
    }
}

// Merge t of a level: owning problem and columns (device-derived schedule).
struct MergeCtx {
    int prob, c0, c1, c2;
    MergeState *S;
};
__device__ inline MergeCtx merge_ctx(const ProbList &pl, const LevelArgs &la, int t)
{
    MergeCtx c;
    c.prob = la.task_prob ? la.task_prob[t] : find_task(la.task_off, pl.count, t);
    const int j = t - off_at(la.task_off, c.prob);
    merge_cols(pl.get(c.prob).N, la.r, j, la.ls, c.c0, c.c1, c.c2);
    c.S = la.states + t;
    return c;
}
__host__ __device__ inline int strips_per_task(const LevelArgs &la) { return (la.stride + DT - 1) / DT; }

// Column-parallel merge kernels (gather, finalize): a merge takes TPM = the power of two
// >= its column count threads when that is <= DT (DT / TPM merges per workgroup), else
// strips of DT columns.  Thread -> (merge t, first column x, column step).  With singular
// values only, a merge has two columns: the first and the last of its right vectors.
__host__ __device__ inline int col_count(const LevelArgs &la) { return la.vonly ? 2 : la.stride; }
__host__ __device__ inline int merge_tpm(const LevelArgs &la)
{
    int t = 1;
    while (t < col_count(la)) t <<= 1;
    return t < DT ? t : DT;
}
__host__ __device__ inline int strips_of(const LevelArgs &la) { return (col_count(la) + DT - 1) / DT; }
__host__ inline int column_blocks(const LevelArgs &la)
{
    return col_count(la) <= DT ? (la.ntask + DT / merge_tpm(la) - 1) / (DT / merge_tpm(la)) : la.ntask * strips_of(la);
}
struct ColumnSlot {
    int t, x, step, first; // first: this slot's threads hold columns 0..
};
__device__ inline ColumnSlot column_slot(const LevelArgs &la)
{
    ColumnSlot c;
    if (col_count(la) <= DT) {
        const int tpm = merge_tpm(la);
        c.t = blockIdx.x * (DT / tpm) + (int)threadIdx.x / tpm;
        c.x = threadIdx.x % tpm;
        c.step = tpm;
        c.first = 1;
    } else {
        const int spt = strips_of(la), strip = blockIdx.x % spt;
        c.t = blockIdx.x / spt;
        c.x = strip * DT + threadIdx.x;
        c.step = DT;
        c.first = strip == 0;
    }
    return c;
}

// ---------------------------------------------------------------------------
// deflate: one workgroup per merge (Deflation, svd.cpp:325-975; dlasd2 structure)
// dynamic LDS: pd[n] (double), pz[n] (double), flag[n] (int), scan[DT+1] (int)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DT) void dc_deflate_kernel(ProbList pl, LevelArgs la)
{
    const MergeCtx T = merge_ctx(pl, la, blockIdx.x);
    const ProbDesc P = pl.get(T.prob);
    const int N = P.N, c0 = T.c0, c1 = T.c1, c2 = T.c2;
    const int n1 = c1 - c0 - 1, m1 = c1 - c0, m2 = c2 - c1, n2 = (c2 == N) ? m2 : m2 - 1;
    const int n = n1 + 1 + n2, m = m1 + m2, sq = m - n;
    const int tid = threadIdx.x;
    extern __shared__ double dsm[];
    double *pd = dsm, *pz = pd + n;
    int *flag = reinterpret_cast<int *>(pz + n);
    int *scan = flag + n;
    __shared__ double s_tol;
    __shared__ int s_nrot;

    const float *d = P.d, *wt = P.wt;
    const double alpha = P.d[c1 - 1], beta = P.e[c1 - 1];
    double orgnrm = fmax(fabs(alpha), fabs(beta));
    double dmax = 0.0;
    if (n1 > 0) dmax = fmax(dmax, (double)d[c0]);
    if (n2 > 0) dmax = fmax(dmax, (double)d[c1]);
    orgnrm = fmax(orgnrm, dmax);
    const double scale = orgnrm > 0.0 ? orgnrm : 1.0;
    const double al = alpha / scale, be = beta / scale;
    Arena Ar = make_arena(P.idx, c0, n);

    // merged ascending order of the two descending lists (ties: left first)
    for (int p = tid; p < n1; p += DT) {
        const float x = d[c0 + p];
        int lo = 0, hi = n2;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (d[c1 + mid] < x)
                hi = mid;
            else
                lo = mid + 1;
        }
        const int t = 1 + (n1 - 1 - p) + (n2 - lo);
        pd[t] = (double)x / scale;
        pz[t] = al * (double)wt[(size_t)(c0 + p) * N + (c1 - 1)];
        Ar.src[t] = p;
    }
    for (int q = tid; q < n2; q += DT) {
        const float x = d[c1 + q];
        int lo = 0, hi = n1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (d[c0 + mid] <= x)
                hi = mid;
            else
                lo = mid + 1;
        }
        const int t = 1 + (n2 - 1 - q) + (n1 - lo);
        pd[t] = (double)x / scale;
        pz[t] = be * (double)wt[(size_t)(c1 + q) * N + c1];
        Ar.src[t] = -1 - q;
    }
    if (tid == 0) {
        const double tol = 8.0 * (double)FLT_EPSILON * fmax(fmax(fabs(al), fabs(be)), dmax / scale);
        const double zn1 = al * (double)wt[(size_t)(c0 + n1) * N + (c1 - 1)];
        double nc = 1.0, ns = 0.0;
        if (sq) {
            const double zn2 = be * (double)wt[(size_t)(c1 + n2) * N + c1];
            double r = hypot(zn1, zn2);
            if (r <= tol) {
                r = tol;
            } else {
                nc = zn1 / r;
                ns = zn2 / r;
            }
            pz[0] = r;
        } else {
            pz[0] = fabs(zn1) <= tol ? tol : zn1;
        }
        pd[0] = 0.0;
        flag[0] = 0;
        Ar.src[0] = SRC_POLE0;
        s_tol = tol;
        MergeState &S = *T.S;
        S.scale = scale;
        S.tol = tol;
        S.nc = nc;
        S.ns = ns;
        S.n = n;
        S.m = m;
        S.n1 = n1;
        S.m1 = m1;
        S.n2 = n2;
        S.m2 = m2;
        S.sq = sq;
    }
    __syncthreads();
    // sequential deflation scan (svd.cpp:637-742): small z, then close poles by rotation
    if (tid == 0) {
        const double tol = s_tol;
        int jp = -1, nrot = 0;
        for (int j = 1; j < n; ++j) {
            if (fabs(pz[j]) <= tol) {
                flag[j] = 1;
                continue;
            }
            if (jp >= 0 && fabs(pd[j] - pd[jp]) <= tol) {
                const double tau = hypot(pz[jp], pz[j]);
                const double c = pz[j] / tau, s = pz[jp] / tau;
                pz[j] = tau;
                pz[jp] = 0.0;
                flag[jp] = 1;
                Ar.rpair[nrot] = jp | (j << 16);
                Ar.rc[nrot] = (float)c;
                Ar.rs[nrot] = (float)s;
                ++nrot;
            } else if (jp >= 0) {
                flag[jp] = 0;
            }
            jp = j;
        }
        if (jp >= 0) flag[jp] = 0;
        s_nrot = nrot;
    }
    __syncthreads();
    // compaction: kept poles (ascending) then deflated poles (ascending)
    const int chunk = (n + DT - 1) / DT, j0 = min(n, tid * chunk), j1 = min(n, j0 + chunk);
    int cnt = 0;
    for (int j = j0; j < j1; ++j) cnt += (flag[j] == 0);
    scan[tid] = cnt;
    __syncthreads();
    if (tid == 0) { // only the first ceil(n / chunk) threads hold columns
        int acc = 0;
        const int nch = (n + chunk - 1) / chunk;
        for (int t = 0; t < nch; ++t) {
            int v = scan[t];
            scan[t] = acc;
            acc += v;
        }
        scan[DT] = acc;
    }
    __syncthreads();
    const int k = scan[DT];
    int kb = scan[tid];
    for (int j = j0; j < j1; ++j) {
        if (flag[j] == 0) {
            Ar.order[kb] = j;
            Ar.dk[kb] = pd[j];
            Ar.zk[kb] = pz[j];
            ++kb;
        } else {
            const int t = k + (j - kb);
            Ar.order[t] = j;
            Ar.dk[t] = pd[j];
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (k > 1 && Ar.dk[1] <= 0.5 * s_tol) Ar.dk[1] = 0.5 * s_tol;
        T.S->k = k;
        T.S->nrot = s_nrot;
        // where the kept poles come from (the merge products' block structure)
        int kl = 0, kr = 0, km = 0;
        for (int t = 0; t < k; ++t) {
            const int s = Ar.src[Ar.order[t]];
            kl += s != SRC_POLE0 && s >= 0;
            kr += s < 0;
        }
        for (int r = 0; r < s_nrot; ++r) {
            const int pr = Ar.rpair[r], sa = Ar.src[pr & 0xffff], sb = Ar.src[pr >> 16];
            km += sa != SRC_POLE0 && sb != SRC_POLE0 && ((sa >= 0) != (sb >= 0));
        }
        T.S->kl = kl;
        T.S->kr = kr;
        T.S->km = km;
        // support of each sorted pole's basis row: 1 left child, 2 right child, 3 both (the
        // pole of z; a row that absorbed the other child's row by a rotation, in scan order)
        int *ty = flag; // free after the compaction
        for (int j = 0; j < n; ++j) {
            const int s = Ar.src[j];
            ty[j] = s == SRC_POLE0 ? 3 : s >= 0 ? 1 : 2;
        }
        for (int r = 0; r < s_nrot; ++r) {
            const int pr = Ar.rpair[r];
            ty[pr >> 16] |= ty[pr & 0xffff];
        }
        int gl = 0, gb = 0;
        for (int t = 0; t < k; ++t) {
            const int y = ty[Ar.order[t]];
            gl += y == 1;
            gb += y == 3;
        }
        int pl_ = 0, pb = gl, pr_ = gl + gb;
        for (int t = 0; t < k; ++t) {
            const int j = Ar.order[t], y = ty[j];
            const int p = y == 1 ? pl_++ : y == 3 ? pb++ : pr_++;
            Ar.pinv[t] = p;
            Ar.bperm[p] = j;
        }
        T.S->gl = gl;
        T.S->gb = gb;
    }
}

// ---------------------------------------------------------------------------
// gather: bases of the merged problem in sorted-pole order + deflation rotations.
// One thread per column x of the merged basis; strips of DT columns.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DT) void dc_gather_kernel(ProbList pl, LevelArgs la)
{
    const ColumnSlot cs = column_slot(la);
    if (cs.t >= la.ntask) return;
    const MergeCtx T = merge_ctx(pl, la, cs.t);
    const ProbDesc P = pl.get(T.prob);
    const MergeState S = *T.S;
    const int N = P.N, c0 = T.c0, c1 = T.c1;
    const int n = S.n, m = S.m, n1 = S.n1, m1 = S.m1, n2 = S.n2, m2 = S.m2;
    // singular values only: slot 0 takes the first column, slot 1 the last
    const int x = la.vonly ? (cs.x == 0 ? 0 : (cs.x == 1 && m > 1 ? m - 1 : m)) : cs.x;
    if (x >= m) return;
    const bool vonly = la.vonly != 0;
    const Arena Ar = make_arena(P.idx, c0, n);
    const float *qt = P.qt, *wt = P.wt;
    float *st = P.qt + (size_t)N * N, *gt = P.wt + (size_t)N * N;
    const float nc = (float)S.nc, ns = (float)S.ns;
    const float v1n = (x < m1) ? wt[(size_t)(c0 + n1) * N + c0 + x] : 0.f;
    const float v2n = (S.sq && x >= m1) ? wt[(size_t)(c1 + n2) * N + c1 + (x - m1)] : 0.f;
    // entries of row t of the gathered bases (the sources are the first halves of qt, wt,
    // the destinations the second halves: four rows are read before they are written)
    auto fetch = [&](int t, float &qv, float &wv) {
        const int s = Ar.src[t];
        const bool pole = s == SRC_POLE0, left = !pole && s >= 0;
        const int q = -1 - s;
        const bool qok = !pole && (left ? x < n1 : (x > n1 && x < n));
        const bool wok = !pole && (left ? x < m1 : x >= m1);
        qv = (qok && !vonly) ? qt[left ? (size_t)(c0 + s) * N + c0 + x : (size_t)(c1 + q) * N + c1 + (x - n1 - 1)] : 0.f;
        wv = wok ? wt[left ? (size_t)(c0 + s) * N + c0 + x : (size_t)(c1 + q) * N + c1 + (x - m1)] : 0.f;
        if (pole) {
            qv = (x == n1) ? 1.f : 0.f;
            wv = S.sq ? (x < m1 ? nc * v1n : ns * v2n) : v1n;
        }
    };
    auto put = [&](int t, float qv, float wv) {
        if (x < n && !vonly) st[(size_t)(c0 + t) * N + c0 + x] = qv;
        gt[(size_t)(c0 + t) * N + c0 + x] = wv;
    };
    int t = 0;
    for (; t + 4 <= n; t += 4) {
        float qv[4], wv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) fetch(t + u, qv[u], wv[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) put(t + u, qv[u], wv[u]);
    }
    for (; t < n; ++t) {
        float qv, wv;
        fetch(t, qv, wv);
        put(t, qv, wv);
    }
    if (S.sq) gt[(size_t)(c0 + n) * N + c0 + x] = (x < m1) ? -ns * v1n : nc * v2n;
    for (int r = 0; r < S.nrot; ++r) {
        const int pr = Ar.rpair[r], jp = pr & 0xffff, j = pr >> 16;
        const float c = Ar.rc[r], s = Ar.rs[r];
        if (x < n && !vonly) {
            float *a = st + (size_t)(c0 + jp) * N + c0 + x, *b = st + (size_t)(c0 + j) * N + c0 + x;
            const float av = *a, bv = *b;
            *a = c * av - s * bv;
            *b = s * av + c * bv;
        }
        float *a = gt + (size_t)(c0 + jp) * N + c0 + x, *b = gt + (size_t)(c0 + j) * N + c0 + x;
        const float av = *a, bv = *b;
        *a = c * av - s * bv;
        *b = s * av + c * bv;
    }
}

// ---------------------------------------------------------------------------
// secular roots: one wavefront per root (replaces SecularEquationSolver,
// svd.cpp:977-1336): FP64, origin chosen by the midpoint test, two-pole rational
// model, bracket with bisection fallback, convergence flag into info.
// ---------------------------------------------------------------------------
// sum over the G lanes of an aligned lane group (G = 64: the wavefront)
template <int G>
__device__ inline double group_sum(double v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int G>
__global__ __launch_bounds__(DT) void dc_secular_kernel(ProbList pl, LevelArgs la)
{
    const int gw = (blockIdx.x * DT + threadIdx.x) / G, lane = threadIdx.x % G;
    if (gw >= la.ntask * la.stride) return;
    const int i = gw % la.stride;
    const MergeCtx T = merge_ctx(pl, la, gw / la.stride);
    const MergeState &S = *T.S;
    const int k = S.k;
    if (i >= k) return;
    const ProbDesc P = pl.get(T.prob);
    const Arena Ar = make_arena(P.idx, T.c0, S.n);
    const double *d = Ar.dk, *z = Ar.zk;
    const double eps = DBL_EPSILON;
    if (k == 1) {
        if (lane == 0) {
            Ar.Kx[0] = 0;
            Ar.mu[0] = z[0] * z[0];
        }
        return;
    }
    int K;
    double lo, hi;
    if (i == k - 1) {
        double zz = 0;
        for (int j = lane; j < k; j += G) zz += z[j] * z[j];
        zz = group_sum<G>(zz);
        K = k - 1;
        lo = 0.0;
        hi = zz;
    } else {
        const double dl = d[i], dr = d[i + 1], om = 0.5 * (dl + dr);
        double f = 0;
        for (int j = lane; j < k; j += G) f += z[j] * z[j] / ((d[j] - om) * (d[j] + om));
        f = 1.0 + group_sum<G>(f);
        if (f > 0) {
            K = i;
            lo = 0.0;
            hi = (om - dl) * (om + dl);
        } else {
            K = i + 1;
            lo = (om - dr) * (om + dr);
            hi = 0.0;
        }
    }
    const double dK = d[K];
    double mu = 0.5 * (lo + hi);
    int it;
    for (it = 1; it <= la.maxit; ++it) {
        double psi = 0, dpsi = 0, phi = 0, dphi = 0, erretm = 0;
        for (int j = lane; j < k; j += G) {
            const double t = (d[j] - dK) * (d[j] + dK) - mu;
            const double q = z[j] / t, term = z[j] * q;
            if (j <= i) {
                psi += term;
                dpsi += q * q;
            } else {
                phi += term;
                dphi += q * q;
            }
            erretm += fabs(term);
        }
        psi = group_sum<G>(psi);
        dpsi = group_sum<G>(dpsi);
        phi = group_sum<G>(phi);
        dphi = group_sum<G>(dphi);
        erretm = group_sum<G>(erretm);
        const double w = 1.0 + psi + phi;
        if (w == 0.0 || fabs(w) <= 8.0 * eps * (1.0 + erretm)) break;
        if (w < 0)
            lo = mu;
        else
            hi = mu;
        const double a1 = (d[i] - dK) * (d[i] + dK) - mu;
        double next;
        if (i == k - 1) {
            const double s = a1 * a1 * dpsi, c = w - s / a1;
            next = (c > 0) ? mu + a1 + s / c : 0.5 * (lo + hi);
        } else {
            const double a2 = (d[i + 1] - dK) * (d[i + 1] + dK) - mu;
            const double s = a1 * a1 * dpsi, t = a2 * a2 * dphi;
            const double c = w - s / a1 - t / a2;
            const double B = c * (a1 + a2) + s + t, C = a1 * a2 * w;
            double eta;
            if (c == 0.0) {
                eta = C / B;
            } else {
                double disc = B * B - 4.0 * c * C;
                if (disc < 0) disc = 0;
                const double sqd = sqrt(disc);
                const double qq = 0.5 * (B + (B >= 0 ? sqd : -sqd));
                const double r1 = qq / c, r2 = (qq != 0.0) ? C / qq : 0.0;
                eta = (r1 > a1 && r1 < a2) ? r1 : r2;
            }
            next = mu + eta;
        }
        if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        const double step = fabs(next - mu);
        mu = next;
        if (step <= 2.0 * eps * fabs(mu) || (hi - lo) <= 4.0 * eps * fmax(fabs(lo), fabs(hi))) break;
    }
    if (lane == 0) {
        Ar.Kx[i] = K;
        Ar.mu[i] = mu;
        if (it > la.maxit) info_count_failure(P.info);
    }
}

// Same iteration, one thread per root with serial pole loops: used for the small merges
// of the low levels (k <= 96), where a wavefront per root would leave most lanes idle.
__device__ inline int secular_serial(int k, const double *d, const double *z, int i, int maxit, int &Kout,
                                     double &muout)
{
    const double eps = DBL_EPSILON;
    if (k == 1) {
        Kout = 0;
        muout = z[0] * z[0];
        return 0;
    }
    int K;
    double lo, hi;
    if (i == k - 1) {
        double zz = 0;
        for (int j = 0; j < k; ++j) zz += z[j] * z[j];
        K = k - 1;
        lo = 0.0;
        hi = zz;
    } else {
        const double dl = d[i], dr = d[i + 1], om = 0.5 * (dl + dr);
        double f = 1.0;
        for (int j = 0; j < k; ++j) f += z[j] * z[j] / ((d[j] - om) * (d[j] + om));
        if (f > 0) {
            K = i;
            lo = 0.0;
            hi = (om - dl) * (om + dl);
        } else {
            K = i + 1;
            lo = (om - dr) * (om + dr);
            hi = 0.0;
        }
    }
    const double dK = d[K];
    double mu = 0.5 * (lo + hi);
    int it;
    for (it = 1; it <= maxit; ++it) {
        double psi = 0, dpsi = 0, phi = 0, dphi = 0, erretm = 0;
        for (int j = 0; j < k; ++j) {
            const double t = (d[j] - dK) * (d[j] + dK) - mu;
            const double q = z[j] / t, term = z[j] * q;
            if (j <= i) {
                psi += term;
                dpsi += q * q;
            } else {
                phi += term;
                dphi += q * q;
            }
            erretm += fabs(term);
        }
        const double w = 1.0 + psi + phi;
        if (w == 0.0 || fabs(w) <= 8.0 * eps * (1.0 + erretm)) break;
        if (w < 0)
            lo = mu;
        else
            hi = mu;
        const double a1 = (d[i] - dK) * (d[i] + dK) - mu;
        double next;
        if (i == k - 1) {
            const double s = a1 * a1 * dpsi, c = w - s / a1;
            next = (c > 0) ? mu + a1 + s / c : 0.5 * (lo + hi);
        } else {
            const double a2 = (d[i + 1] - dK) * (d[i + 1] + dK) - mu;
            const double s = a1 * a1 * dpsi, t = a2 * a2 * dphi;
            const double c = w - s / a1 - t / a2;
            const double B = c * (a1 + a2) + s + t, C = a1 * a2 * w;
            double eta;
            if (c == 0.0) {
                eta = C / B;
            } else {
                double disc = B * B - 4.0 * c * C;
                if (disc < 0) disc = 0;
                const double sqd = sqrt(disc);
                const double qq = 0.5 * (B + (B >= 0 ? sqd : -sqd));
                const double r1 = qq / c, r2 = (qq != 0.0) ? C / qq : 0.0;
                eta = (r1 > a1 && r1 < a2) ? r1 : r2;
            }
            next = mu + eta;
        }
        if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        const double step = fabs(next - mu);
        mu = next;
        if (step <= 2.0 * eps * fabs(mu) || (hi - lo) <= 4.0 * eps * fmax(fabs(lo), fabs(hi))) break;
    }
    Kout = K;
    muout = mu;
    return it;
}

__global__ __launch_bounds__(DT) void dc_secular_thread_kernel(ProbList pl, LevelArgs la)
{
    const int g = blockIdx.x * DT + threadIdx.x;
    if (g >= la.ntask * la.stride) return;
    const int i = g % la.stride;
    const MergeCtx T = merge_ctx(pl, la, g / la.stride);
    const MergeState &S = *T.S;
    const int k = S.k;
    if (i >= k) return;
    const ProbDesc P = pl.get(T.prob);
    const Arena Ar = make_arena(P.idx, T.c0, S.n);
    int K;
    double mu;
    const int it = secular_serial(k, Ar.dk, Ar.zk, i, la.maxit, K, mu);
    Ar.Kx[i] = K;
    Ar.mu[i] = mu;
    if (it > la.maxit) info_count_failure(P.info);
}

// ---------------------------------------------------------------------------
// Loewner z (svd.cpp:1424-1438, in FP64) and output positions; one thread per pole.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DT) void dc_lowner_kernel(ProbList pl, LevelArgs la)
{
    const int g = blockIdx.x * DT + threadIdx.x;
    if (g >= la.ntask * la.stride) return;
    const int j = g % la.stride;
    const MergeCtx T = merge_ctx(pl, la, g / la.stride);
    const MergeState &S = *T.S;
    const int k = S.k, n = S.n;
    if (j >= n) return;
    const ProbDesc P = pl.get(T.prob);
    const Arena Ar = make_arena(P.idx, T.c0, n);
    const double *d = Ar.dk;
    int rank;
    if (j < k) {
        const double dj = d[j];
        double prod;
        {
            const int K = Ar.Kx[k - 1];
            prod = Ar.mu[k - 1] - (dj - d[K]) * (dj + d[K]);
        }
        for (int i = 0; i < j; ++i) {
            const int K = Ar.Kx[i];
            const double df = (dj - d[K]) * (dj + d[K]) - Ar.mu[i];
            prod *= df / ((dj - d[i]) * (dj + d[i]));
        }
        for (int i = j; i < k - 1; ++i) {
            const int K = Ar.Kx[i];
            const double df = (dj - d[K]) * (dj + d[K]) - Ar.mu[i];
            prod *= df / ((dj - d[i + 1]) * (dj + d[i + 1]));
        }
        const double zj = Ar.zk[j];
        Ar.zk[j] = copysign(sqrt(fabs(prod)), zj);
        // rank of root j among deflated poles (those with d^2 < sigma_j^2)
        const int K = Ar.Kx[j];
        const double s2 = d[K] * d[K] + Ar.mu[j];
        int lo = k, hi = n;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (d[mid] * d[mid] < s2)
                lo = mid + 1;
            else
                hi = mid;
        }
        rank = j + (lo - k);
    } else {
        const double dd2 = d[j] * d[j];
        int lo = 0, hi = k;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            const int K = Ar.Kx[mid];
            if (d[K] * d[K] + Ar.mu[mid] <= dd2)
                lo = mid + 1;
            else
                hi = mid;
        }
        rank = (j - k) + lo;
    }
    Ar.pos[j] = n - 1 - rank;
}

// ---------------------------------------------------------------------------
// secular vectors (svd.cpp:1440-1496): a group of G lanes per root i (G = 64, a
// wavefront, for the large merges; 16, 8 or 4 for small ones, several roots per
// wavefront); FP64, stored FP32.  UMt (left, in a) and VMt (right, in vmt), row i =
// coefficients of root i.
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(DT) void dc_vectors_kernel(ProbList pl, LevelArgs la)
{
    const int gw = (blockIdx.x * DT + threadIdx.x) / G, lane = threadIdx.x % G;
    if (gw >= la.ntask * la.stride) return;
    const int i = gw % la.stride;
    const MergeCtx T = merge_ctx(pl, la, gw / la.stride);
    const MergeState &S = *T.S;
    const int k = S.k;
    if (i >= k) return;
    const ProbDesc P = pl.get(T.prob);
    const int N = P.N, c0 = T.c0;
    const Arena Ar = make_arena(P.idx, c0, S.n);
    const double *d = Ar.dk, *z = Ar.zk;
    float *um = P.lsv + (size_t)(c0 + i) * N + c0;
    float *vm = P.vmt + (size_t)(c0 + i) * N + c0;
    const bool left = !la.vonly;
    // coefficient column of pole j: grouped by the support of its basis row (pinv)
    if (k == 1 && z[0] == 0.0) {
        if (lane == 0) {
            if (left) um[Ar.pinv[0]] = 1.f;
            vm[Ar.pinv[0]] = 1.f;
        }
        return;
    }
    const int K = Ar.Kx[i];
    const double dK = d[K], mu = Ar.mu[i];
    double su = 0, sv = 0;
    for (int j = lane; j < k; j += G) {
        const double df = (d[j] - dK) * (d[j] + dK) - mu;
        const double v = z[j] / df, u = (j == 0) ? -1.0 : d[j] * z[j] / df;
        su += u * u;
        sv += v * v;
    }
    su = group_sum<G>(su);
    sv = group_sum<G>(sv);
    const double nu = 1.0 / sqrt(su), nv = 1.0 / sqrt(sv);
    for (int j = lane; j < k; j += G) {
        const double df = (d[j] - dK) * (d[j] + dK) - mu;
        const double v = z[j] / df, u = (j == 0) ? -1.0 : d[j] * z[j] / df;
        const int p = Ar.pinv[j];
        if (left) um[p] = (float)(u * nu);
        vm[p] = (float)(v * nv);
    }
}

// ---------------------------------------------------------------------------
// finalize: deflated rows, the null vector, and sigma in descending order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DT) void dc_finalize_kernel(ProbList pl, LevelArgs la)
{
    const ColumnSlot cs = column_slot(la);
    if (cs.t >= la.ntask) return;
    const MergeCtx T = merge_ctx(pl, la, cs.t);
    const ProbDesc P = pl.get(T.prob);
    const MergeState S = *T.S;
    const int N = P.N, c0 = T.c0, n = S.n, m = S.m, k = S.k;
    const Arena Ar = make_arena(P.idx, c0, n);
    const float *st = P.qt + (size_t)N * N, *gt = P.wt + (size_t)N * N;
    const bool vonly = la.vonly != 0;
    const int x = vonly ? (cs.x == 0 ? 0 : (cs.x == 1 && m > 1 ? m - 1 : m)) : cs.x;
    if (x < m) {
        for (int t = k; t < n; ++t) {
            const int src = Ar.order[t], dst = Ar.pos[t];
            if (x < n && !vonly) P.qt[(size_t)(c0 + dst) * N + c0 + x] = st[(size_t)(c0 + src) * N + c0 + x];
            P.wt[(size_t)(c0 + dst) * N + c0 + x] = gt[(size_t)(c0 + src) * N + c0 + x];
        }
        if (S.sq) P.wt[(size_t)(c0 + n) * N + c0 + x] = gt[(size_t)(c0 + n) * N + c0 + x];
    }
    if (cs.first) {
        for (int t = cs.x; t < n; t += cs.step) {
            double s;
            if (t < k) {
                const int K = Ar.Kx[t];
                s = sqrt(Ar.dk[K] * Ar.dk[K] + Ar.mu[t]);
            } else {
                s = Ar.dk[t];
            }
            P.d[c0 + Ar.pos[t]] = (float)(s * S.scale);
        }
    }
}

// ---------------------------------------------------------------------------
// merge products (svd.cpp:1504-1516): rows pos[i] of the node block <-
//   sum_p coef[i][p] * basis[bperm[p]], i, p < k, one 64x64 tile per workgroup (the
//   coefficient columns grouped by the support of their basis rows, see the deflation).
// side 0: coef = left secular vectors (in lsv), basis = gathered left bases, out qt;
// side 1: coef = right secular vectors (vmt), basis = gathered right bases, out wt.
// ---------------------------------------------------------------------------
// Rounds of at least MERGE_XCD_MIN merges: workgroup b runs on XCD b % 8 (workgroups are
// handed to the 8 XCDs in turn) and takes merge (b / 8 / tiles) 8 + b % 8 and that merge's
// tile (b / 8) % tiles, so all tiles of a merge share one XCD's L2 (its coefficient and
// basis blocks are read by every tile row and column) while consecutive merges still
// alternate between XCDs; the grid holds ceil(ntask / 8) 8 merges and the ones past ntask
// exit.  Rounds of fewer merges (one large matrix: down to a single merge of thousands of
// tiles) take tile b % tiles of merge b / tiles, spread over every XCD.
constexpr int MERGE_XCD_MIN = 64;
__host__ __device__ inline bool merge_xcd_grouped(const LevelArgs &la) { return la.ntask >= MERGE_XCD_MIN; }

template <int MODE, int TKV>
__global__ __launch_bounds__(GT) void dc_merge_gemm_kernel(ProbList pl, LevelArgs la, int side)
{
    const int tpt = (la.stride + 63) / 64, tiles = tpt * tpt;
    const int b = blockIdx.x;
    int task, tile;
    if (merge_xcd_grouped(la)) {
        task = ((b >> 3) / tiles) * 8 + (b & 7);
        tile = (b >> 3) % tiles;
    } else {
        task = b / tiles;
        tile = b % tiles;
    }
    if (task >= la.ntask) return;
    const MergeCtx T = merge_ctx(pl, la, task);
    const MergeState &S = *T.S;
    const ProbDesc P = pl.get(T.prob);
    const int N = P.N, c0 = T.c0, k = S.k;
    const Arena Ar = make_arena(P.idx, c0, S.n);
    TileCtx t;
    const size_t base = (size_t)c0 * N + c0;
    t.A = (side == 0 ? P.lsv : P.vmt) + base;
    t.B = (side == 0 ? P.qt : P.wt) + (size_t)N * N + base;
    t.C = (side == 0 ? P.qt : P.wt) + base;
    t.brow = Ar.bperm;
    t.crow = Ar.pos;
    t.lda = t.ldb = t.ldc = N;
    t.m = k;
    t.K = k;
    t.n = side == 0 ? S.n : S.m;
    t.transA = t.transB = 0;
    t.alpha = 1.f;
    t.beta = 0.f;
    t.m0 = (tile / tpt) * 64;
    t.n0 = (tile % tpt) * 64;
    if (t.m0 >= t.m || t.n0 >= t.n) return;
    // a column tile inside one child's block needs only the coefficient columns whose basis
    // rows have support there: left block [0, n1) of U / [0, m1) of W, right block
    // [n1 + 1, n) / [m1, m) (U's column n1 belongs to the pole of z alone)
    const int lend = side == 0 ? S.n1 : S.m1, rbeg = side == 0 ? S.n1 + 1 : S.m1;
    const int n1t = min(t.n0 + 64, t.n);
    if (n1t <= lend) {
        t.K = S.gl + S.gb;
    } else if (t.n0 >= rbeg) {
        t.A += S.gl;
        t.brow += S.gl;
        t.K = k - S.gl;
    }
    tile_gemm<MODE, TKV>(t);
}

// Right-vector product of a merge with singular values only: the first and last columns
// of rows pos[i] of the node block <- sum_p VMt[i][p] * gathered[bperm[p]] (i, p < k), one
// thread per output entry, FP32 in the order j = 0, 1, ...
__global__ __launch_bounds__(DT) void dc_merge_cols_kernel(ProbList pl, LevelArgs la)
{
    const int g = blockIdx.x * DT + threadIdx.x;
    const int per = 2 * la.stride;
    if (g >= la.ntask * per) return;
    const int w = g % per, i = w >> 1, side = w & 1;
    const MergeCtx T = merge_ctx(pl, la, g / per);
    const MergeState &S = *T.S;
    const int k = S.k, m = S.m;
    if (i >= k || (side == 1 && m == 1)) return;
    const ProbDesc P = pl.get(T.prob);
    const int N = P.N, c0 = T.c0, x = side ? m - 1 : 0;
    const Arena Ar = make_arena(P.idx, c0, S.n);
    const float *vm = P.vmt + (size_t)(c0 + i) * N + c0;
    const float *gt = P.wt + (size_t)N * N + (size_t)c0 * N + c0 + x;
    float acc = 0.f;
    for (int p = 0; p < k; ++p) acc += vm[p] * gt[(size_t)Ar.bperm[p] * N];
    P.wt[(size_t)(c0 + Ar.pos[i]) * N + c0 + x] = acc;
}

// back-transforms (updateUVt, svd.cpp:1911-1981):
//   side 0: a (M x N) <- u[:, :N] * qt^T        side 1: qt+N*N (N x N) <- wt * vt
template <int MODE>
__global__ __launch_bounds__(GT) void dc_bt_gemm_kernel(ProbList pl, const int *tile_off, int side)
{
    const int pid = find_task(tile_off, pl.count, blockIdx.x);
    const int tile = blockIdx.x - off_at(tile_off, pid);
    const ProbDesc P = pl.get(pid);
    const int M = P.M, N = P.N, tn = (N + 63) / 64;
    TileCtx t;
    if (side == 0) {
        t.A = P.u;
        t.lda = M;
        t.B = P.qt;
        t.ldb = N;
        t.transB = 1;
        t.C = P.a;
        t.ldc = N;
        t.m = M;
    } else {
        t.A = P.wt;
        t.lda = N;
        t.B = P.vt;
        t.ldb = N;
        t.transB = 0;
        t.C = P.qt + (size_t)N * N;
        t.ldc = N;
        t.m = N;
    }
    t.transA = 0;
    t.brow = t.crow = nullptr;
    t.n = N;
    t.K = N;
    t.alpha = 1.f;
    t.beta = 0.f;
    t.m0 = (tile / tn) * 64;
    t.n0 = (tile % tn) * 64;
    if (t.m0 >= t.m) return;
    tile_gemm<MODE, 16>(t);
}

// copy of the back-transform results into u[:, :N] (side 0) and vt (side 1)
// rows [rows g, rows (g + 1)) of the back-transformed vectors into U (its first N columns)
// or Vt
__global__ __launch_bounds__(DT) void dc_bt_copy_kernel(ProbList pl, const int *row_off, int side, int rows)
{
    const int pid = find_task(row_off, pl.count, blockIdx.x);
    const int g = blockIdx.x - off_at(row_off, pid);
    const ProbDesc P = pl.get(pid);
    const int M = P.M, N = P.N, r0 = g * rows;
    const int nr = min(rows, (side == 0 ? M : N) - r0);
    if (nr <= 0) return;
    if (side == 0) { // the group's rows are contiguous in a (leading dimension N): element q of
                     // the group is row q / N, column q % N, tracked incrementally
        const float *src = P.a + (size_t)r0 * N;
        float *dst = P.u + (size_t)r0 * M;
        const int n = nr * N, dr = DT / N, dc = DT % N;
        int r = (int)threadIdx.x / N, c = (int)threadIdx.x % N;
        for (int q = threadIdx.x; q < n; q += DT) {
            gmem(dst)[(size_t)r * M + c] = gmem(src)[q];
            r += dr;
            c += dc;
            if (c >= N) {
                c -= N;
                ++r;
            }
        }
    } else {
        copy_block(P.vt + (size_t)r0 * N, P.qt + (size_t)N * N + (size_t)r0 * N, nr * N);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_dc_leaves(ProbList pl, const int *leaf_off, int nleaf, int ls, hipStream_t s)
{
    if (nleaf <= 0) return hipSuccess;
    if (ls < LEAF_MIN || ls > LEAF_MAX) return hipErrorInvalidValue;
    if (ls <= 3)
        hipLaunchKernelGGL(dc_leaf_kernel, dim3((nleaf + DT - 1) / DT), dim3(DT), 0, s, pl, leaf_off, nleaf, ls);
    else if (ls <= LEAF_THREAD_MAX)
        hipLaunchKernelGGL(dc_leaf_thread_kernel<LEAF_THREAD_MAX>, dim3((nleaf + DT - 1) / DT), dim3(DT), 0, s, pl,
                           leaf_off, nleaf, ls);
    else
        hipLaunchKernelGGL(dc_leaf_wave_kernel, dim3(nleaf), dim3(64), 0, s, pl, leaf_off, ls);
    return hipGetLastError();
}

template <int MODE>
static void launch_merge_gemms(ProbList pl, LevelArgs la, hipStream_t s)
{
    const int tpt = (la.stride + 63) / 64;
    // grouped by XCD: whole groups of 8 merges (one per XCD)
    const int grid = (merge_xcd_grouped(la) ? (la.ntask + 7) / 8 * 8 : la.ntask) * tpt * tpt;
    // Merges of up to 512 columns: K steps of 16, 76 VGPRs against 128 at steps of 32, so
    // six workgroups share a CU instead of four and hide more of the operand loads' latency
    // (D&C of the 10,000-matrix batch 31.8 -> 29.6 ms).  Larger merges (one large matrix:
    // few tiles, K up to N) keep steps of 32 (at 16 the 4096^2 D&C took 27.3 ms, not 17.3).
    if (la.stride <= 512) {
        hipLaunchKernelGGL((dc_merge_gemm_kernel<MODE, 16>), dim3(grid), dim3(GT), 0, s, pl, la, 0);
        hipLaunchKernelGGL((dc_merge_gemm_kernel<MODE, 16>), dim3(grid), dim3(GT), 0, s, pl, la, 1);
    } else {
        hipLaunchKernelGGL((dc_merge_gemm_kernel<MODE, 32>), dim3(grid), dim3(GT), 0, s, pl, la, 0);
        hipLaunchKernelGGL((dc_merge_gemm_kernel<MODE, 32>), dim3(grid), dim3(GT), 0, s, pl, la, 1);
    }
}

hipError_t launch_dc_level(ProbList pl, LevelArgs la, hipStream_t s)
{
    static size_t maxdyn = allow_dynamic_lds((const void *)dc_deflate_kernel);
    if (la.ntask <= 0) return hipSuccess;
    const size_t lds = (size_t)la.maxn * (2 * sizeof(double) + sizeof(int)) + (DT + 1) * sizeof(int);
    if (lds > maxdyn) return hipErrorInvalidValue;
    const int rows = la.ntask * la.stride;
    const int wave_blocks = (rows + DT / 64 - 1) / (DT / 64);
    hipLaunchKernelGGL(dc_deflate_kernel, dim3(la.ntask), dim3(DT), lds, s, pl, la);
    hipLaunchKernelGGL(dc_gather_kernel, dim3(column_blocks(la)), dim3(DT), 0, s, pl, la);
    if (la.stride <= 96)
        hipLaunchKernelGGL(dc_secular_thread_kernel, dim3((rows + DT - 1) / DT), dim3(DT), 0, s, pl, la);
    else if (la.stride <= 192) // lanes per root: 16, 32, then a wavefront
        hipLaunchKernelGGL(dc_secular_kernel<16>, dim3((int)(((size_t)rows * 16 + DT - 1) / DT)), dim3(DT), 0, s, pl, la);
    else if (la.stride <= 384)
        hipLaunchKernelGGL(dc_secular_kernel<32>, dim3((int)(((size_t)rows * 32 + DT - 1) / DT)), dim3(DT), 0, s, pl, la);
    else
        hipLaunchKernelGGL(dc_secular_kernel<64>, dim3(wave_blocks), dim3(DT), 0, s, pl, la);
    hipLaunchKernelGGL(dc_lowner_kernel, dim3((rows + DT - 1) / DT), dim3(DT), 0, s, pl, la);
    // lanes per root: a wavefront for merges above 192 columns, fewer below
    const int vg = la.stride <= 12 ? 4 : la.stride <= 48 ? 8 : la.stride <= 192 ? 16 : 64;
    const int vblocks = (int)(((size_t)rows * vg + DT - 1) / DT);
    if (vg == 4)
        hipLaunchKernelGGL(dc_vectors_kernel<4>, dim3(vblocks), dim3(DT), 0, s, pl, la);
    else if (vg == 8)
        hipLaunchKernelGGL(dc_vectors_kernel<8>, dim3(vblocks), dim3(DT), 0, s, pl, la);
    else if (vg == 16)
        hipLaunchKernelGGL(dc_vectors_kernel<16>, dim3(vblocks), dim3(DT), 0, s, pl, la);
    else
        hipLaunchKernelGGL(dc_vectors_kernel<64>, dim3(vblocks), dim3(DT), 0, s, pl, la);
    const int mode = gemm_get_mode();
    if (la.vonly)
        hipLaunchKernelGGL(dc_merge_cols_kernel, dim3((int)(((size_t)la.ntask * 2 * la.stride + DT - 1) / DT)), dim3(DT),
                           0, s, pl, la);
    else if (mode == 1)
        launch_merge_gemms<1>(pl, la, s);
    else if (mode == 2)
        launch_merge_gemms<2>(pl, la, s);
    else
        launch_merge_gemms<0>(pl, la, s);
    hipLaunchKernelGGL(dc_finalize_kernel, dim3(column_blocks(la)), dim3(DT), 0, s, pl, la);
    return hipGetLastError();
}

template <int MODE>
static void launch_bt(ProbList pl, const int *tou, int ntu, const int *tov, int ntv, hipStream_t s)
{
    if (ntu > 0) hipLaunchKernelGGL(dc_bt_gemm_kernel<MODE>, dim3(ntu), dim3(GT), 0, s, pl, tou, 0);
    if (ntv > 0) hipLaunchKernelGGL(dc_bt_gemm_kernel<MODE>, dim3(ntv), dim3(GT), 0, s, pl, tov, 1);
}

hipError_t launch_dc_backtransform(ProbList pl, const int *tile_off_u, int ntiles_u, const int *tile_off_v,
                                   int ntiles_v, const int *row_off_u, int nrows_u, const int *row_off_v,
                                   int nrows_v, int rows, hipStream_t s)
{
    const int mode = gemm_get_mode();
    if (mode == 1)
        launch_bt<1>(pl, tile_off_u, ntiles_u, tile_off_v, ntiles_v, s);
    else if (mode == 2)
        launch_bt<2>(pl, tile_off_u, ntiles_u, tile_off_v, ntiles_v, s);
    else
        launch_bt<0>(pl, tile_off_u, ntiles_u, tile_off_v, ntiles_v, s);
    if (nrows_u > 0) hipLaunchKernelGGL(dc_bt_copy_kernel, dim3(nrows_u), dim3(DT), 0, s, pl, row_off_u, 0, rows);
    if (nrows_v > 0) hipLaunchKernelGGL(dc_bt_copy_kernel, dim3(nrows_v), dim3(DT), 0, s, pl, row_off_v, 1, rows);
    return hipGetLastError();
}

} // namespace paco
