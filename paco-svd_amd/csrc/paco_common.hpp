// Shared device/host definitions of the MI355X (gfx950) SVD engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PACO_WAVE 64

namespace paco {

// Global-memory view of a pointer.  Accesses through a generic pointer compile to FLAT
// instructions, which count in both the vector-memory and the LDS counters, so a wait for
// an LDS access also waits for every FLAT load in flight: a load issued ahead of an LDS
// stage to hide its latency is then waited for at once.  Every buffer of the problem
// descriptors is device memory; through this view its accesses compile to global_* ones.
template <typename T>
__device__ inline __attribute__((address_space(1))) T *gmem(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// One SVD problem: the reference's global-memory buffers (main.cpp:33-48) plus
// the engine's scratch.  All row-major FP32, M >= N.
struct ProbDesc {
    int M, N;
    float *a;      // M*N   in: A; out of bidiag: reflectors; back-transform scratch
    float *u;      // M*M   Q, then left singular vectors
    float *vt;     // N*N   P^T, then right singular vectors (rows)
    float *d;      // N     diagonal, then sigma (descending)
    float *e;      // N-1   superdiagonal
    float *tauq;   // N
    float *taup;   // N-1
    float *qt;     // 2*N*N left vectors of D&C nodes | gathered bases
    float *wt;     // 2*N*N right vectors of D&C nodes | gathered bases
    uint32_t *idx; // 16*N  per-column arena of the merges
    float *vmt;    // N*N   right secular vectors of the merges
    float *lsv;    // N*N   left secular vectors of the merges (leading dimension N): the top
                   //       N x N block of a on the single-matrix paths; a block of their own
                   //       in the batched scratch, so that the D&C never writes a, which
                   //       the formation of Q and P^T reads, and the two run concurrently
    float *tq;     // ceil(N/32)*1024 block reflector factors T (Q side)
    float *tp;     // ceil(N/32)*1024 block reflector factors T (P side)
    int *info;     // 1     LAPACK-style status: 0 ok, <0 bad input, >0 no convergence
    int ld;        // row stride of a for the bidiagonalization (0: N)
};

struct MergeState {
    double scale;  // orgnrm
    double tol;    // deflation tolerance (scaled units)
    double nc, ns; // rotation combining the two null vectors
    int k;         // non-deflated count
    int nrot;      // deflation rotations
    int n, m;      // merged rows / columns
    int n1, m1, n2, m2;
    int sq;
    int kl, kr;    // non-deflated poles from the left / right child (the pole of z aside)
    int km;        // deflation rotations that mixed a left and a right pole
    // coefficient columns of the merge products grouped by the basis rows' support:
    // [0, gl) left child only, [gl, gl + gb) both (the pole of z, rows mixed by a
    // rotation), [gl + gb, k) right child only
    int gl, gb;
};

static_assert(sizeof(MergeState) % 8 == 0, "MergeState alignment");

// Grouped GEMM descriptor: C[crow(i)] = alpha * op(A)[i,:] * op(B)[:, j] (+ beta C)
// with an optional gathered row index on B (row kk of B is brow[kk]) and a
// scattered row index on C.  m and K may be read on the device (dyn), which is
// how the merge GEMMs learn k after deflation.
struct GemmDesc {
    const float *A;
    const float *B;
    float *C;
    const int *brow;
    const int *crow;
    const int *dyn_mk; // if set: m = K = *dyn_mk (bounded by m, k below)
    int lda, ldb, ldc;
    int m, n, k;
    int transA, transB;
    float alpha, beta;
    int tile_off;      // prefix over descs of tiles
    int tiles_n;       // tiles along n
};

// Arena carving of the idx buffer for the merge covering [c0, c0+n): all offsets in
// 32-bit words from idx + 16*c0.  Doubles are 8-byte aligned because c0 and n are
// multiplied by 16 and 2 respectively.
struct Arena {
    double *dk;   // [n] compacted poles (kept first, then deflated), scaled
    double *zk;   // [n] compacted z for kept poles; Loewner z after that phase
    double *mu;   // [n] per root: sigma^2 = dk[K]^2 + mu
    int *Kx;      // [n] per root origin pole
    int *src;     // [n] per sorted pole: basis source (see dc.hip)
    int *order;   // [n] compacted index -> sorted pole index
    int *pos;     // [n] compacted index -> output row (descending sigma)
    int *rpair;   // [n] rotation t: (jp | j << 16)
    float *rc;    // [n] rotation cosine
    float *rs;    // [n] rotation sine
    int *pinv;    // [k] kept pole t -> its column in the secular-vector coefficients
    int *bperm;   // [k] coefficient column p -> sorted pole index (basis row of the products)
};

__host__ __device__ inline Arena make_arena(uint32_t *idx, int c0, int n)
{
    uint32_t *base = idx + 16 * (size_t)c0;
    Arena A;
    A.dk = reinterpret_cast<double *>(base);
    A.zk = reinterpret_cast<double *>(base + 2 * (size_t)n);
    A.mu = reinterpret_cast<double *>(base + 4 * (size_t)n);
    A.Kx = reinterpret_cast<int *>(base + 6 * (size_t)n);
    A.src = reinterpret_cast<int *>(base + 7 * (size_t)n);
    A.order = reinterpret_cast<int *>(base + 8 * (size_t)n);
    A.pos = reinterpret_cast<int *>(base + 9 * (size_t)n);
    A.rpair = reinterpret_cast<int *>(base + 10 * (size_t)n);
    A.rc = reinterpret_cast<float *>(base + 11 * (size_t)n);
    A.rs = reinterpret_cast<float *>(base + 12 * (size_t)n);
    A.pinv = reinterpret_cast<int *>(base + 13 * (size_t)n);
    A.bperm = reinterpret_cast<int *>(base + 14 * (size_t)n);
    return A;
}

// A descriptor read through a pointer that the kernel could also write is loaded with
// vector loads: the compiler then treats its fields as per-lane values.  Every lane
// reads the same descriptor, so the first lane's copy is taken into scalar registers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T>
__device__ inline T *uni(T *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}
__device__ inline ProbDesc uniform_desc(const ProbDesc &P)
{
    ProbDesc U;
    U.M = uni(P.M);
    U.N = uni(P.N);
    U.a = uni(P.a);
    U.u = uni(P.u);
    U.vt = uni(P.vt);
    U.d = uni(P.d);
    U.e = uni(P.e);
    U.tauq = uni(P.tauq);
    U.taup = uni(P.taup);
    U.qt = uni(P.qt);
    U.wt = uni(P.wt);
    U.idx = uni(P.idx);
    U.vmt = uni(P.vmt);
    U.lsv = uni(P.lsv);
    U.tq = uni(P.tq);
    U.tp = uni(P.tp);
    U.info = uni(P.info);
    U.ld = uni(P.ld);
    return U;
}

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Sum over the 64 lanes with DPP moves: inside each row of 16 lanes (quad swaps, then the
// half-row and row mirrors: every lane of row q holds its row's sum R_q), then across the
// rows with the row broadcasts (rows 1 and 3 add lane 15 of the row before, rows 2 and 3
// add lane 31), which leaves (R3 + R2) + (R1 + R0) in lane 63: 7 vector instructions, no
// LDS round trip.  Every lane returns the same value, the order of additions is fixed
// (and gives the bits of (R0 + R1) + (R2 + R3)).  All 64 lanes must be active.  The
// cross-row adds are written as DPP-modified adds, whose lanes outside the row mask keep
// their value (two wait states before a DPP read of a just-written register).
#define PACO_DPP(v, ctrl) __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false))
__device__ inline float wave_sum_dpp(float v)
{
    v += PACO_DPP(v, 0xB1);  // quad_perm [1,0,3,2]
    v += PACO_DPP(v, 0x4E);  // quad_perm [2,3,0,1]
    v += PACO_DPP(v, 0x141); // row_half_mirror
    v += PACO_DPP(v, 0x140); // row_mirror
    asm volatile("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\ts_nop 1"
                 : "+v"(v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// FP64 form of wave_sum_dpp: the same lane moves applied to both 32-bit halves
#define PACO_DPP_D(v, ctrl)                                                                                      \
    __hiloint2double(__builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false),                  \
                     __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false))
__device__ inline double wave_sum_dpp_d(double v)
{
    v += PACO_DPP_D(v, 0xB1);  // quad_perm [1,0,3,2]
    v += PACO_DPP_D(v, 0x4E);  // quad_perm [2,3,0,1]
    v += PACO_DPP_D(v, 0x141); // row_half_mirror
    v += PACO_DPP_D(v, 0x140); // row_mirror
    const int hi = __double2hiint(v), lo = __double2loint(v);
    const double r0 = __hiloint2double(__builtin_amdgcn_readlane(hi, 0), __builtin_amdgcn_readlane(lo, 0));
    const double r1 = __hiloint2double(__builtin_amdgcn_readlane(hi, 16), __builtin_amdgcn_readlane(lo, 16));
    const double r2 = __hiloint2double(__builtin_amdgcn_readlane(hi, 32), __builtin_amdgcn_readlane(lo, 32));
    const double r3 = __hiloint2double(__builtin_amdgcn_readlane(hi, 48), __builtin_amdgcn_readlane(lo, 48));
    return (r0 + r1) + (r2 + r3);
}
#undef PACO_DPP_D
#undef PACO_DPP
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Householder parameters of a vector [x1; c] with sigma = |c|^2, in the reference's sign
// convention (upper_bidiagonalization.cpp:182-218): beta = +mu, v1 = x1 - mu (x1 <= 0) or
// -sigma / (x1 + mu), tau = 2 v1^2 / (sigma + v1^2), inv = 1 / v1; sigma == 0 gives
// tau = 0, beta = x1.  sigma is summed and the parameters are formed in FP64: squares of
// FP32 entries below ~1e-19 underflow FP32 (denormal sigma, a few significant bits), which
// would leave tau inconsistent with the stored vector c / v1 (H = I - tau v v^T no longer
// orthogonal).  When tau rounds to 0 in FP32 or 1/v1 overflows it, the reflector is the
// identity (tau = 0, beta = x1), so no stored vector entry can be infinite.
//
// The operands of the square root and the divisions are nonzero normal doubles (a nonzero
// sum of at most 2^12 squares of FP32 values lies in [2^-298, 2^268]; v1 < 0 always), so
// they are formed from the hardware estimates without the range scaling and special cases
// of the general routines: 1/b by two Newton steps on v_rcp_f64, a/b by a residual
// correction of a * (1/b), sqrt(s) by Goldschmidt's iteration on v_rsq_f64 with two residual
// corrections.  Each is within an ulp or two of FP64, far below the FP32 rounding of the
// results.
__device__ inline double rcp_normal_d(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    return fma(r, fma(-b, r, 1.0), r);
}
__device__ inline double div_normal_d(double a, double b)
{
    const double r = rcp_normal_d(b), q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ inline double sqrt_normal_d(double s)
{
    const double y = __builtin_amdgcn_rsq(s);
    double g = s * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, s), h, g);
    return fma(fma(-g, g, s), h, g);
}
__device__ inline void house_params_d(double x1, double sigma, float *beta, float *tau, float *inv)
{
    if (sigma > 0.0) {
        const double mu = sqrt_normal_d(fma(x1, x1, sigma));
        const double v1 = (x1 <= 0.0) ? x1 - mu : -div_normal_d(sigma, x1 + mu);
        const double v1sq = v1 * v1;
        const float t = (float)div_normal_d(2.0 * v1sq, sigma + v1sq), iv = (float)rcp_normal_d(v1);
        if (t != 0.f && fabsf(iv) <= 3.0e38f) {
            *tau = t;
            *inv = iv;
            *beta = (float)mu;
            return;
        }
    }
    *tau = 0.f;
    *beta = (float)x1;
    *inv = 1.f;
}

// Binary search: index i of the group with offs[i] <= x < offs[i+1] (offs sorted).
// A single group needs no offsets array (offs may be null).
__host__ __device__ inline int find_task(const int *offs, int ntask, int x)
{
    if (ntask <= 1) return 0;
    int lo = 0, hi = ntask - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__host__ __device__ inline int off_at(const int *offs, int i) { return offs ? offs[i] : 0; }

// Count one non-converged secular root in a problem's status word unless the word already
// holds a negative code (bad input), which is kept as it is.
__device__ inline void info_count_failure(int *info)
{
    if (atomicAdd(info, 1) < 0) atomicSub(info, 1);
}

// Problem of task x for offsets that the batch plan writes as count + 1 prefix entries
// followed by one entry per task (its problem): one load instead of a binary search.  A
// single problem needs neither (offs may be null).
__host__ __device__ inline int task_of(const int *offs, int count, int x)
{
    return (count <= 1 || !offs) ? 0 : offs[count + 1 + x];
}

// Leaf partition of N columns with leaves of ls columns (svd_tiling.cpp:83-130 is the
// case ls = 3): L = N/ls + 1 leaves for N >= ls, leaf l starts at ls l; the last starts
// one column early when N % ls == 0.  N < ls is one leaf.  Leaf l covers
// [leaf_start(l), leaf_start(l+1)) with leaf_start(L) = N.  Every leaf but the last is
// an (m - 1) x m upper bidiagonal block (m <= ls columns), the last is square (m < ls).
constexpr int LEAF_MIN = 3, LEAF_MAX = 32;
__host__ __device__ inline int leaf_count(int N, int ls) { return N < ls ? 1 : N / ls + 1; }
__host__ __device__ inline int leaf_start(int N, int l, int ls)
{
    const int L = leaf_count(N, ls);
    if (l >= L) return N;
    if (L == 1) return 0;
    if (l <= L - 2) return ls * l;
    return ls * l - (N % ls == 0 ? 1 : 0);
}
// Merges of round r >= 1 (pairs of nodes of round r-1, left to right, odd tail carried;
// svd.cpp:168-189): node j of round r covers leaves [j 2^r, min((j+1) 2^r, L)); it is
// a merge when its right half is not empty.  A node of round r spans at most ls 2^r
// columns.
__host__ __device__ inline int merges_in_round(int N, int r, int ls)
{
    const int L = leaf_count(N, ls), span = 1 << r, half = span >> 1;
    return L > half ? (L - half + span - 1) / span : 0;
}
__host__ __device__ inline int rounds_of(int N, int ls)
{
    int L = leaf_count(N, ls), r = 0;
    while ((1 << r) < L) ++r;
    return r;
}
__host__ __device__ inline void merge_cols(int N, int r, int j, int ls, int &c0, int &c1, int &c2)
{
    const int L = leaf_count(N, ls), span = 1 << r, half = span >> 1;
    const int l0 = j * span, lm = l0 + half, l1 = (l0 + span < L) ? l0 + span : L;
    c0 = leaf_start(N, l0, ls);
    c1 = leaf_start(N, lm, ls);
    c2 = leaf_start(N, l1, ls);
}

} // namespace paco

#define PACO_CHECK(x)                                                                              \
    do {                                                                                           \
        hipError_t _e = (x);                                                                       \
        if (_e != hipSuccess) return (int)_e;                                                      \
    } while (0)
