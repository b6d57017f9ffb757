/*
 * paco_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement, in plain C and IEEE double, of the reference's SVD path:
 *   (a) Golub-Kahan upper bidiagonalization   (upper_bidiagonalization.cpp)
 *   (b) bidiagonal divide-and-conquer SVD       (svd.cpp, svd_tiling.cpp)
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this file's library.  The product (libpacosvd.so) never links or calls it.
 *
 * Layouts follow the reference's global-memory contract (main.cpp:33-48):
 * row-major, M >= N, a[M*N], u[M*M], vt[N*N], d[N], e[N-1], tauq[N], taup[N-1].
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_house            closed form: H x = mu e1, mu >= 0, H orthogonal, tau in [0,2]
 *   orc_bidiag           invariants vs FP64 LAPACK: U, Vt orthogonal; U B Vt = A;
 *                        sigma(B) = dgesdd(A); d_i, e_i >= 0 (uniqueness of Golub-Kahan
 *                        with this sign convention); brute-force 2x2 / 3x2 cases
 *   orc_bidiag_f32       the FP32 yardstick (three evaluation orders of its inner products):
 *                        each within FP32 forward error of orc_bidiag, tau in [0, 2],
 *                        tau = 2 / |v|^2 for every stored vector
 *   orc_leaf_svd         brute force vs numpy.linalg.svd on every 1x1/1x2/2x2/2x3 case,
 *                        including rank-0/1 and zero-pattern branches
 *   orc_secular_root     scipy.linalg.lapack.dlasd4 per root (LAPACK's own solver)
 *   orc_bdc / orc_svd_DC sigma vs dgesdd, residual, orthogonality; deflation
 *                        definitions; golden fixture of the 128x96 seed-0 matrix
 * No function here is "parity unpinned".
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* (a) bidiagonalization                                                      */
/* ------------------------------------------------------------------------- */

/* Householder reflector with the reference's sign convention
 * (upper_bidiagonalization.cpp:182-218 ComputeHouseholder, :220-250 ComputeSigma,
 * :251-283 ComputeHouseVec; length<=1 shortcut :130-135 / :158-163):
 *   sigma = sum_{k>=1} x_k^2; sigma == 0 -> tau = 0, beta = x_0 (signed);
 *   mu = sqrt(x_0^2 + sigma) (always >= 0); v_0 = x_0 - mu if x_0 <= 0 else
 *   -sigma/(x_0 + mu) (Parlett); tau = 2 v_0^2 / (sigma + v_0^2);
 *   x[1:] <- x[1:] / v_0 (implicit unit head), x[0] <- mu = beta. */
ORC_EXPORT void orc_house(int n, double *x, long inc, double *beta, double *tau)
{
    if (n <= 1) {
        *tau = 0.0;
        *beta = x[0];
        return;
    }
    double x0 = x[0], sigma = 0.0;
    for (int k = 1; k < n; k++) sigma += x[k * inc] * x[k * inc];
    if (sigma == 0.0) {
        *tau = 0.0;
        *beta = x0;
        return;
    }
    double mu = sqrt(x0 * x0 + sigma);
    double v0 = (x0 <= 0.0) ? x0 - mu : -sigma / (x0 + mu);
    *tau = 2.0 * v0 * v0 / (sigma + v0 * v0);
    double inv = 1.0 / v0;
    for (int k = 1; k < n; k++) x[k * inc] *= inv;
    x[0] = mu;
    *beta = mu;
}

/* Apply H = I - tau v v^T (v_0 = 1, v_k = hv[k*hinc]) to the vector y (stride yinc). */
static void apply_reflector(int n, const double *hv, long hinc, double tau, double *y, long yinc)
{
    double c = y[0];
    for (int k = 1; k < n; k++) c += hv[k * hinc] * y[k * yinc];
    c *= tau;
    y[0] -= c;
    for (int k = 1; k < n; k++) y[k * yinc] -= c * hv[k * hinc];
}

/* Golub-Kahan upper bidiagonalization, unblocked, exactly the reference's order of
 * operations (Process, upper_bidiagonalization.cpp:83-120): column reflector i,
 * applied to columns i+1..N-1 (ApplyColumnTransformV2 :426-475); row reflector i,
 * applied to rows i+1..M-1 (ApplyRowTransformV2 :476-527).  Then u = Q = H_0..H_{N-1}
 * and vt = P^T = G_{N-2}..G_0 by backward accumulation from the identity
 * (GetUVt :774-893, initUV :894-933).  On exit A = u * B * vt. */
ORC_EXPORT void orc_bidiag(int M, int N, double *a, double *u, double *vt, double *d, double *e,
                           double *tauq, double *taup)
{
    for (int i = 0; i < N; i++) {
        orc_house(M - i, a + (long)i * N + i, N, d + i, tauq + i);
        if (tauq[i] != 0.0)
            for (int j = i + 1; j < N; j++)
                apply_reflector(M - i, a + (long)i * N + i, N, tauq[i], a + (long)i * N + j, N);
        if (i < N - 1) {
            orc_house(N - i - 1, a + (long)i * N + i + 1, 1, e + i, taup + i);
            if (taup[i] != 0.0)
                for (int r = i + 1; r < M; r++)
                    apply_reflector(N - i - 1, a + (long)i * N + i + 1, 1, taup[i], a + (long)r * N + i + 1, 1);
        }
    }
    memset(u, 0, sizeof(double) * (size_t)M * M);
    memset(vt, 0, sizeof(double) * (size_t)N * N);
    for (int i = 0; i < M; i++) u[(long)i * M + i] = 1.0;
    for (int i = 0; i < N; i++) vt[(long)i * N + i] = 1.0;
    for (int i = N - 1; i >= 0; i--) {
        if (tauq[i] == 0.0) continue;
        for (int j = i; j < M; j++)
            apply_reflector(M - i, a + (long)i * N + i, N, tauq[i], u + (long)i * M + j, M);
    }
    for (int i = N - 2; i >= 0; i--) {
        if (taup[i] == 0.0) continue;
        for (int j = i; j < N; j++)
            apply_reflector(N - i - 1, a + (long)i * N + i + 1, 1, taup[i], vt + (long)j * N + i + 1, 1);
    }
}

/* The same bidiagonalization in FP32 (vectors, reflector applications, stored outputs),
 * with each reflector's sum of squares and its parameters in FP64 as the GPU forms them.
 * It is the yardstick of the FP32 forward error: Golub-Kahan coefficients are
 * forward-sensitive, so the GPU's outputs are compared with the FP64 oracle within a
 * margin of what an FP32 run of the same algorithm deviates.  Only the reflectors, d, e,
 * tauq, taup are formed (no u, vt). */
static void house_f32(int n, float *x, long inc, float *beta, float *tau)
{
    if (n <= 1) {
        *tau = 0.f;
        *beta = x[0];
        return;
    }
    const double x0 = x[0];
    double sigma = 0.0;
    for (int k = 1; k < n; k++) sigma += (double)x[k * inc] * x[k * inc];
    if (sigma == 0.0) {
        *tau = 0.f;
        *beta = (float)x0;
        return;
    }
    const double mu = sqrt(x0 * x0 + sigma);
    const double v0 = (x0 <= 0.0) ? x0 - mu : -sigma / (x0 + mu);
    *tau = (float)(2.0 * v0 * v0 / (sigma + v0 * v0));
    const float inv = (float)(1.0 / v0);
    for (int k = 1; k < n; k++) x[k * inc] *= inv;
    x[0] = (float)mu;
    *beta = (float)mu;
}

/* The inner product v^T y of a reflector application in one of three FP32 evaluation
 * orders, all exact-arithmetic equal: 0 forward, 1 backward, 2 eight interleaved partial
 * sums added pairwise at the end (the order of a vector unit).  Their spread samples how
 * far FP32 rounding alone moves the forward-sensitive outputs. */
static void apply_reflector_f32(int n, const float *hv, long hinc, float tau, float *y, long yinc, int order)
{
    float c;
    if (order == 1) {
        c = 0.f;
        for (int k = n - 1; k >= 1; k--) c += hv[k * hinc] * y[k * yinc];
        c += y[0];
    } else if (order == 2) {
        float p[8] = {y[0], 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 1; k < n; k++) p[k & 7] += hv[k * hinc] * y[k * yinc];
        c = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    } else {
        c = y[0];
        for (int k = 1; k < n; k++) c += hv[k * hinc] * y[k * yinc];
    }
    c *= tau;
    y[0] -= c;
    for (int k = 1; k < n; k++) y[k * yinc] -= c * hv[k * hinc];
}

ORC_EXPORT void orc_bidiag_f32_order(int M, int N, float *a, float *d, float *e, float *tauq, float *taup, int order)
{
    for (int i = 0; i < N; i++) {
        house_f32(M - i, a + (long)i * N + i, N, d + i, tauq + i);
        if (tauq[i] != 0.f)
            for (int j = i + 1; j < N; j++)
                apply_reflector_f32(M - i, a + (long)i * N + i, N, tauq[i], a + (long)i * N + j, N, order);
        if (i < N - 1) {
            house_f32(N - i - 1, a + (long)i * N + i + 1, 1, e + i, taup + i);
            if (taup[i] != 0.f)
                for (int r = i + 1; r < M; r++)
                    apply_reflector_f32(N - i - 1, a + (long)i * N + i + 1, 1, taup[i], a + (long)r * N + i + 1, 1,
                                        order);
        }
    }
}

ORC_EXPORT void orc_bidiag_f32(int M, int N, float *a, float *d, float *e, float *tauq, float *taup)
{
    orc_bidiag_f32_order(M, N, a, d, e, tauq, taup, 0);
}

/* ------------------------------------------------------------------------- */
/* (b) divide and conquer                                                     */
/* ------------------------------------------------------------------------- */

/* Leaf partition of the N columns (svd_tiling.cpp:83-130 GenerateTiling): blocks of 3
 * columns; the tail is [2x2 square], [1x1], or [1x2]+[1x1] by N mod 3.  A block
 * [c0,c1) is a (c1-c0-1) x (c1-c0) upper-bidiagonal piece, square when c1 == N.
 * N <= 2 is a single leaf (the reference special-cases it, svd.cpp:140-159).
 * Returns the number of leaves; starts/ends must hold N/3+1 entries. */
ORC_EXPORT int orc_leaf_partition(int N, int *starts, int *ends)
{
    if (N <= 2) {
        starts[0] = 0;
        ends[0] = N;
        return 1;
    }
    int cnt = N / 3 + 1, rem = N % 3;
    for (int i = 0; i < cnt - 2; i++) {
        starts[i] = 3 * i;
        ends[i] = 3 * i + 3;
    }
    int i = cnt - 2;
    starts[i] = 3 * i;
    ends[i] = 3 * i + (rem == 0 ? 2 : 3);
    i = cnt - 1;
    if (rem == 0) {
        starts[i] = 3 * i - 1;
        ends[i] = 3 * i;
    } else {
        starts[i] = 3 * i;
        ends[i] = 3 * i + rem;
    }
    return cnt;
}

/* SVD of a leaf block B (n x m, n <= 2, m in {n, n+1}, upper bidiagonal with
 * B[r][r] = dd[r], B[r][r+1] = ee[r]) by one-sided Jacobi on the columns of B.
 * Replaces the closed forms of svd.cpp:1627-1864 (compute_2x3/2x2/1x2/1x1_svd) with
 * a method that covers their rank-0/1 branches uniformly.  Outputs sigma[n]
 * descending and >= 0, U (n x n, column i = left vector i) and V (m x m, column
 * i = right vector i; column n is the null vector when m = n+1). */
ORC_EXPORT void orc_leaf_svd(int n, int m, const double *dd, const double *ee, double *sigma,
                             double *U, double *V)
{
    double W[2][3] = {{0}}, J[3][3] = {{0}};
    for (int r = 0; r < n; r++) {
        W[r][r] = dd[r];
        if (r + 1 < m) W[r][r + 1] = ee[r];
    }
    for (int c = 0; c < m; c++) J[c][c] = 1.0;
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
    double nrm[3];
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
            for (int r = 0; r < n; r++) U[r * n + i] = (r == 0);
        } else { /* i == 1, n == 2: perpendicular completion */
            U[0 * n + 1] = -U[1 * n + 0];
            U[1 * n + 1] = U[0 * n + 0];
        }
    }
}

/* Safeguarded root finder for the secular equation
 *   f(s) = 1 + sum_j z_j^2 / (d_j^2 - s^2) = 0,   0 = d_0 < d_1 < ... < d_{k-1},
 * root i lying in (d_i, d_{i+1}) (or (d_{k-1}, sqrt(d_{k-1}^2 + |z|^2)] for the last).
 * Plays the role of SecularEquationSolver (svd.cpp:977-1336): same midpoint test to
 * pick the origin d_K (svd.cpp:1032-1063), same two-pole rational model; added here:
 * a bracket kept on every step with bisection fallback, FP64 throughout, and a
 * convergence flag.  Returns s^2 = d_K^2 + mu, with (K, mu) so that differences
 * d_j^2 - s^2 = (d_j - d_K)(d_j + d_K) - mu are formed without cancellation.
 * Return value: iterations used, or -1 when the iteration limit was reached. */
ORC_EXPORT int orc_secular_root(int k, const double *d, const double *z, int i, double *sig,
                                int *Kout, double *muout)
{
    const double eps = DBL_EPSILON;
    if (k == 1) {
        *Kout = 0;
        *muout = z[0] * z[0];
        *sig = sqrt(d[0] * d[0] + z[0] * z[0]);
        return 0;
    }
    int K;
    double lo, hi;
    if (i == k - 1) {
        double zz = 0;
        for (int j = 0; j < k; j++) zz += z[j] * z[j];
        K = k - 1;
        lo = 0.0;
        hi = zz;
    } else {
        double dl = d[i], dr = d[i + 1];
        double om = 0.5 * (dl + dr);
        double f = 1.0;
        for (int j = 0; j < k; j++) f += z[j] * z[j] / ((d[j] - om) * (d[j] + om));
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
    for (it = 1; it <= 200; it++) {
        double psi = 0, dpsi = 0, phi = 0, dphi = 0, erretm = 0;
        for (int j = 0; j < k; j++) {
            double t = (d[j] - dK) * (d[j] + dK) - mu;
            double q = z[j] / t, term = z[j] * q;
            if (j <= i) {
                psi += term;
                dpsi += q * q;
            } else {
                phi += term;
                dphi += q * q;
            }
            erretm += fabs(term);
        }
        double w = 1.0 + psi + phi;
        if (w == 0.0 || fabs(w) <= 8.0 * eps * (1.0 + erretm)) break;
        if (w < 0)
            lo = mu;
        else
            hi = mu;
        double a1 = (d[i] - dK) * (d[i] + dK) - mu;
        double next;
        if (i == k - 1) {
            /* f(x) ~ c + s/(a1 - eta): all poles lumped into the nearest one */
            double s = a1 * a1 * dpsi, c = w - s / a1;
            next = (c > 0) ? mu + a1 + s / c : 0.5 * (lo + hi);
        } else {
            double a2 = (d[i + 1] - dK) * (d[i + 1] + dK) - mu;
            double s = a1 * a1 * dpsi, t = a2 * a2 * dphi;
            double c = w - s / a1 - t / a2;
            /* c eta^2 - B eta + a1 a2 w = 0, root in (a1, a2) */
            double B = c * (a1 + a2) + s + t, C = a1 * a2 * w;
            double eta;
            if (c == 0.0) {
                eta = C / B;
            } else {
                double disc = B * B - 4.0 * c * C;
                if (disc < 0) disc = 0;
                double sq = sqrt(disc);
                double qq = -0.5 * (-B + (B >= 0 ? -sq : sq));
                double r1 = qq / c, r2 = (qq != 0.0) ? C / qq : 0.0;
                eta = (r1 > a1 && r1 < a2) ? r1 : r2;
            }
            next = mu + eta;
        }
        if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        double step = fabs(next - mu);
        mu = next;
        if (step <= 2.0 * eps * fabs(mu) || (hi - lo) <= 4.0 * eps * fmax(fabs(lo), fabs(hi))) break;
    }
    *Kout = K;
    *muout = mu;
    *sig = sqrt(dK * dK + mu);
    return it > 200 ? -1 : it;
}

/* Per-merge statistics, for tests and logging. */
typedef struct {
    int n, k, nrot, ndefl;
} orc_merge_stat;

/* Merge of the left node [c0,c1) and right node [c1,c2) (MergeSubMatrix svd.cpp:225-323
 * driving Deflation :325-975 and MergeSubMatrix_step2 :1337-1594).  Node storage,
 * as in the reference (svd.cpp:87-133): d[c0..] singular values, qt rows c0.. are
 * left vectors (row i = vector i), wt rows c0.. are right vectors, each confined to
 * the node's diagonal block; nodes are kept sorted by descending sigma.
 * defl_eps is the unit roundoff used by the deflation tolerance
 * tol = 8 eps max(|alpha|,|beta|,max d) (svd.cpp:637-638; the reference uses FLT_EPSILON). */
static void merge_nodes(int N, int c0, int c1, int c2, double *d, const double *e, double *qt,
                        double *wt, double defl_eps, orc_merge_stat *st)
{
    const int n1 = c1 - c0 - 1, m1 = c1 - c0;
    const int m2 = c2 - c1, n2 = (c2 == N) ? m2 : m2 - 1;
    const int n = n1 + 1 + n2, m = m1 + m2, sq = m - n;
    double alpha = d[c1 - 1], beta = e[c1 - 1];

    double orgnrm = fmax(fabs(alpha), fabs(beta));
    for (int p = 0; p < n1; p++) orgnrm = fmax(orgnrm, fabs(d[c0 + p]));
    for (int q = 0; q < n2; q++) orgnrm = fmax(orgnrm, fabs(d[c1 + q]));

    /* basis vectors of the merged problem, in merged-local coordinates:
     * Bq (n x n) rows = left basis, Bw (m x m) rows = right basis. */
    int np = n - 1; /* number of poles besides pole 0 */
    double *pd = malloc(sizeof(double) * (n + 1)), *pz = malloc(sizeof(double) * (n + 1));
    int *src = malloc(sizeof(int) * (n + 1)); /* >=0: left index p; <0: -(q+1) right index */
    double *Bq = calloc((size_t)n * n, sizeof(double)), *Bw = calloc((size_t)m * m, sizeof(double));
    double *nullw = calloc((size_t)m, sizeof(double));
    double scale = orgnrm > 0 ? orgnrm : 1.0;
    double al = alpha / scale, be = beta / scale;

    /* merge the two descending lists into ascending order (DLAMRG-like, svd.cpp:527-585) */
    {
        int a = n1 - 1, b = n2 - 1, t = 1;
        while (a >= 0 || b >= 0) {
            int takeLeft;
            if (a < 0)
                takeLeft = 0;
            else if (b < 0)
                takeLeft = 1;
            else
                takeLeft = d[c0 + a] <= d[c1 + b];
            if (takeLeft) {
                src[t] = a;
                pd[t] = d[c0 + a] / scale;
                pz[t] = al * wt[(long)(c0 + a) * N + (c1 - 1)];
                a--;
            } else {
                src[t] = -(b + 1);
                pd[t] = d[c1 + b] / scale;
                pz[t] = be * wt[(long)(c1 + b) * N + c1];
                b--;
            }
            t++;
        }
    }
    for (int t = 1; t <= np; t++) {
        int s = src[t];
        if (s >= 0) {
            for (int c = 0; c < n1; c++) Bq[(long)t * n + c] = qt[(long)(c0 + s) * N + c0 + c];
            for (int c = 0; c < m1; c++) Bw[(long)t * m + c] = wt[(long)(c0 + s) * N + c0 + c];
        } else {
            int q = -s - 1;
            for (int c = 0; c < n2; c++) Bq[(long)t * n + n1 + 1 + c] = qt[(long)(c1 + q) * N + c1 + c];
            for (int c = 0; c < m2; c++) Bw[(long)t * m + m1 + c] = wt[(long)(c1 + q) * N + c1 + c];
        }
    }
    double tol = 8.0 * defl_eps * fmax(fmax(fabs(al), fabs(be)), np > 0 ? pd[np] : 0.0);

    /* pole 0: connecting row e_{n1} on the left, null vectors combined on the right */
    Bq[n1] = 1.0;
    {
        double zn1 = al * wt[(long)(c0 + n1) * N + (c1 - 1)];
        const double *v1n = wt + (long)(c0 + n1) * N + c0;
        if (sq) {
            double zn2 = be * wt[(long)(c1 + n2) * N + c1];
            const double *v2n = wt + (long)(c1 + n2) * N + c1;
            double r = hypot(zn1, zn2), c, s;
            if (r <= tol) {
                c = 1.0;
                s = 0.0;
                r = tol;
            } else {
                c = zn1 / r;
                s = zn2 / r;
            }
            pz[0] = r;
            for (int x = 0; x < m1; x++) {
                Bw[x] = c * v1n[x];
                nullw[x] = -s * v1n[x];
            }
            for (int x = 0; x < m2; x++) {
                Bw[m1 + x] = s * v2n[x];
                nullw[m1 + x] = c * v2n[x];
            }
        } else {
            pz[0] = fabs(zn1) <= tol ? tol : zn1;
            for (int x = 0; x < m1; x++) Bw[x] = v1n[x];
        }
        pd[0] = 0.0;
    }

    /* deflation scan (dlasd2-style; svd.cpp:637-742) */
    int *kept = malloc(sizeof(int) * (n + 1)), *defl = malloc(sizeof(int) * (n + 1));
    int k = 0, nd = 0, nrot = 0, jp = -1;
    kept[k++] = 0;
    for (int j = 1; j <= np; j++) {
        if (fabs(pz[j]) <= tol) {
            defl[nd++] = j;
            continue;
        }
        if (jp >= 0 && fabs(pd[j] - pd[jp]) <= tol) {
            double tau = hypot(pz[jp], pz[j]);
            double c = pz[j] / tau, s = pz[jp] / tau;
            pz[j] = tau;
            pz[jp] = 0.0;
            for (int x = 0; x < n; x++) {
                double bp = Bq[(long)jp * n + x], bj = Bq[(long)j * n + x];
                Bq[(long)jp * n + x] = c * bp - s * bj;
                Bq[(long)j * n + x] = s * bp + c * bj;
            }
            for (int x = 0; x < m; x++) {
                double bp = Bw[(long)jp * m + x], bj = Bw[(long)j * m + x];
                Bw[(long)jp * m + x] = c * bp - s * bj;
                Bw[(long)j * m + x] = s * bp + c * bj;
            }
            defl[nd++] = jp;
            nrot++;
        } else if (jp >= 0) {
            kept[k++] = jp;
        }
        jp = j;
    }
    if (jp >= 0) kept[k++] = jp;
    /* deflated list must be ascending in d for the final merge: sort (insertion) */
    for (int a = 1; a < nd; a++) {
        int v = defl[a], b = a - 1;
        while (b >= 0 && pd[defl[b]] > pd[v]) {
            defl[b + 1] = defl[b];
            b--;
        }
        defl[b + 1] = v;
    }

    double *dk = malloc(sizeof(double) * k), *zk = malloc(sizeof(double) * k);
    for (int t = 0; t < k; t++) {
        dk[t] = pd[kept[t]];
        zk[t] = pz[kept[t]];
    }
    if (k > 1 && dk[1] <= 0.5 * tol) dk[1] = 0.5 * tol;

    /* secular roots, Loewner z, secular vectors */
    double *sg = malloc(sizeof(double) * k), *mu = malloc(sizeof(double) * k);
    int *Kx = malloc(sizeof(int) * k);
    double *UM = malloc(sizeof(double) * (size_t)k * k), *VM = malloc(sizeof(double) * (size_t)k * k);
    if (k == 1 && zk[0] == 0.0) {
        sg[0] = 0.0;
        UM[0] = 1.0;
        VM[0] = 1.0;
    } else {
        for (int i = 0; i < k; i++) orc_secular_root(k, dk, zk, i, sg + i, Kx + i, mu + i);
        double *zh = malloc(sizeof(double) * k);
        for (int j = 0; j < k; j++) {
            /* Gu-Eisenstat / Loewner recomputation (svd.cpp:1424-1438) */
            double dj = dk[j];
#define DIFF(jj, ii) ((dk[jj] - dk[Kx[ii]]) * (dk[jj] + dk[Kx[ii]]) - mu[ii])
            double prod = -DIFF(j, k - 1);
            for (int i = 0; i < j; i++) prod *= DIFF(j, i) / ((dj - dk[i]) * (dj + dk[i]));
            for (int i = j; i < k - 1; i++) prod *= DIFF(j, i) / ((dj - dk[i + 1]) * (dj + dk[i + 1]));
            zh[j] = copysign(sqrt(fabs(prod)), zk[j]);
        }
        for (int i = 0; i < k; i++) {
            double nu = 0, nv = 0;
            for (int j = 0; j < k; j++) {
                double df = DIFF(j, i);
                double v = zh[j] / df;
                double uu = (j == 0) ? -1.0 : dk[j] * zh[j] / df;
                VM[(long)j * k + i] = v;
                UM[(long)j * k + i] = uu;
                nu += uu * uu;
                nv += v * v;
            }
            nu = 1.0 / sqrt(nu);
            nv = 1.0 / sqrt(nv);
            for (int j = 0; j < k; j++) {
                UM[(long)j * k + i] *= nu;
                VM[(long)j * k + i] *= nv;
            }
        }
#undef DIFF
        free(zh);
    }

    /* output order: merge ascending roots with ascending deflated d, then reverse */
    double *outq = calloc((size_t)n * n, sizeof(double)), *outw = calloc((size_t)m * m, sizeof(double));
    double *outs = malloc(sizeof(double) * n);
    {
        int a = 0, b = 0, pos = n - 1;
        while (a < k || b < nd) {
            int takeRoot;
            if (a >= k)
                takeRoot = 0;
            else if (b >= nd)
                takeRoot = 1;
            else
                takeRoot = sg[a] <= pd[defl[b]];
            if (takeRoot) {
                outs[pos] = sg[a];
                for (int j = 0; j < k; j++) {
                    double cu = UM[(long)j * k + a], cv = VM[(long)j * k + a];
                    const double *bq = Bq + (long)kept[j] * n, *bw = Bw + (long)kept[j] * m;
                    for (int x = 0; x < n; x++) outq[(long)pos * n + x] += cu * bq[x];
                    for (int x = 0; x < m; x++) outw[(long)pos * m + x] += cv * bw[x];
                }
                a++;
            } else {
                int j = defl[b];
                outs[pos] = pd[j];
                memcpy(outq + (long)pos * n, Bq + (long)j * n, sizeof(double) * n);
                memcpy(outw + (long)pos * m, Bw + (long)j * m, sizeof(double) * m);
                b++;
            }
            pos--;
        }
    }
    if (sq) memcpy(outw + (long)n * m, nullw, sizeof(double) * m);
    for (int r = 0; r < n; r++) {
        d[c0 + r] = outs[r] * scale;
        memcpy(qt + (long)(c0 + r) * N + c0, outq + (long)r * n, sizeof(double) * n);
    }
    for (int r = 0; r < m; r++) memcpy(wt + (long)(c0 + r) * N + c0, outw + (long)r * m, sizeof(double) * m);
    if (st) {
        st->n = n;
        st->k = k;
        st->nrot = nrot;
        st->ndefl = nd;
    }
    free(pd); free(pz); free(src); free(Bq); free(Bw); free(nullw); free(kept); free(defl);
    free(dk); free(zk); free(sg); free(mu); free(Kx); free(UM); free(VM);
    free(outq); free(outw); free(outs);
}

/* Divide and conquer SVD of the N x N upper bidiagonal (d, e) (BDC::Process,
 * svd.cpp:134-195): leaves of the reference partition, then rounds of left-to-right
 * pairwise merges with the odd tail carried (svd.cpp:168-189).  On exit d = sigma
 * (descending), qt = U_B^T, wt = V_B^T (row i = singular vector i), B = qt^T S wt.
 * stats (optional) receives one record per merge, in merge order; returns merge count. */
ORC_EXPORT int orc_bdc(int N, double *d, const double *e, double *qt, double *wt, double defl_eps,
                       orc_merge_stat *stats)
{
    int cnt_max = N / 3 + 2;
    int *s = malloc(sizeof(int) * cnt_max), *t = malloc(sizeof(int) * cnt_max);
    int cnt = orc_leaf_partition(N, s, t);
    memset(qt, 0, sizeof(double) * (size_t)N * N);
    memset(wt, 0, sizeof(double) * (size_t)N * N);
    for (int l = 0; l < cnt; l++) {
        int c0 = s[l], c1 = t[l], m = c1 - c0, n = (c1 == N) ? m : m - 1;
        double dd[2] = {0}, ee[2] = {0}, sig[2], U[4], V[9];
        for (int r = 0; r < n; r++) {
            dd[r] = d[c0 + r];
            if (r + 1 < m) ee[r] = e[c0 + r];
        }
        orc_leaf_svd(n, m, dd, ee, sig, U, V);
        for (int i = 0; i < n; i++) {
            d[c0 + i] = sig[i];
            for (int r = 0; r < n; r++) qt[(long)(c0 + i) * N + c0 + r] = U[r * n + i];
        }
        for (int i = 0; i < m; i++)
            for (int r = 0; r < m; r++) wt[(long)(c0 + i) * N + c0 + r] = V[r * m + i];
    }
    int nm = 0;
    while (cnt > 1) {
        int w = 0, i = 0;
        for (; i + 1 < cnt; i += 2) {
            merge_nodes(N, s[i], t[i], t[i + 1], d, e, qt, wt, defl_eps, stats ? stats + nm : NULL);
            nm++;
            s[w] = s[i];
            t[w] = t[i + 1];
            w++;
        }
        if (i < cnt) {
            s[w] = s[i];
            t[w] = t[i];
            w++;
        }
        cnt = w;
    }
    free(s);
    free(t);
    return nm;
}

/* svd_DC restated (svd.cpp:2105-2118 -> Process :134-195 -> updateUVt :1911-1981):
 * inputs are the outputs of orc_bidiag (u = Q, vt = P^T, d, e); on exit d = sigma
 * (descending), u[:, :N] = Q[:, :N] U_B (columns N..M-1 keep Q's), vt = V_B^T P^T. */
ORC_EXPORT void orc_svd_DC(int M, int N, double *u, double *vt, double *d, const double *e,
                           double defl_eps)
{
    double *qt = malloc(sizeof(double) * (size_t)N * N), *wt = malloc(sizeof(double) * (size_t)N * N);
    orc_bdc(N, d, e, qt, wt, defl_eps, NULL);
    double *row = malloc(sizeof(double) * N);
    for (int r = 0; r < M; r++) {
        for (int j = 0; j < N; j++) {
            double acc = 0;
            for (int c = 0; c < N; c++) acc += u[(long)r * M + c] * qt[(long)j * N + c];
            row[j] = acc;
        }
        memcpy(u + (long)r * M, row, sizeof(double) * N);
    }
    double *tmp = malloc(sizeof(double) * (size_t)N * N);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double acc = 0;
            for (int c = 0; c < N; c++) acc += wt[(long)i * N + c] * vt[(long)c * N + j];
            tmp[(long)i * N + j] = acc;
        }
    memcpy(vt, tmp, sizeof(double) * (size_t)N * N);
    free(tmp);
    free(row);
    free(qt);
    free(wt);
}

/* Whole path on one matrix: bidiagonalization then D&C (main.cpp:126-160). */
ORC_EXPORT void orc_svd(int M, int N, const double *A, double *sigma, double *U, double *Vt,
                        double defl_eps)
{
    double *a = malloc(sizeof(double) * (size_t)M * N);
    double *e = malloc(sizeof(double) * (N > 1 ? N - 1 : 1));
    double *tq = malloc(sizeof(double) * N), *tp = malloc(sizeof(double) * (N > 1 ? N - 1 : 1));
    memcpy(a, A, sizeof(double) * (size_t)M * N);
    orc_bidiag(M, N, a, U, Vt, sigma, e, tq, tp);
    orc_svd_DC(M, N, U, Vt, sigma, e, defl_eps);
    free(a);
    free(e);
    free(tq);
    free(tp);
}
