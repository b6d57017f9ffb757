// Golub-Kahan bidiagonalization of one large matrix kept in the registers of the chip.
//
// Reference: upper_bidiagonalization.cpp Process :83-120 (one left and one right reflector
// per column, each applied to the whole trailing matrix as a rank-1 update, four grid
// barriers per column on 40 vector cores).
//
// Here the matrix does not leave the chip between columns.  P workgroups of RT threads
// (one per CU) each hold a share of the rows in VGPRs: workgroup q owns rows q, q + P,
// q + 2P, ... and thread t of it columns t VC, ..., t VC + VC - 1 of each of those rows.  The matrix
// is read once and written once (row r when column r is done, rows >= N at the end); a
// row's registers are zeroed once it is written, so rows that are done (and the padding
// rows beyond M) are zero and every update leaves them zero without a test per row.
// Column j costs one pass over the registers and two exchanges between the workgroups:
//   pass j (every workgroup, its rows r > j): apply the left reflector (A -= tau_q v w^T),
//     form x = A u by a workgroup reduction, apply the right reflector (A -= tau_p x u^T),
//     then gather column j+1's data: c_r = A[r, j+1], the partial sum of c_r^2 and the
//     partial GEMV sums g[c] = sum_{r > j+1} c_r A[r, c] (the reflector's 1/v_1 scaling is
//     deferred to the reduction)
//   exchange 1: every workgroup publishes its partial sums, and the owner of row j+1 that
//     row; workgroup q reduces the columns of its slice (positions [q S, q S + S) of the
//     exchange buffers) over the P partials
//     in a fixed order, forms the left reflector's parameters (every workgroup forms them
//     from the same sums in the same order, so they agree bit for bit) and the slice's
//     entries of w = A^T v and of row j+1 after the left reflector
//   exchange 2: every workgroup reads w and the updated row for its columns and forms the
//     right reflector (its sum of squares reduced in the same fixed order everywhere)
// An exchanged value carries its column number in the upper half of a 64-bit word that is
// written and read with single-copy-atomic device-coherent accesses, so a reader polls the
// data itself (no barrier, no fence).  Two buffers alternate between columns: a buffer is
// rewritten two columns later, after each of its readers has published something that
// depends on having read it.  Every wait is bounded; on a timeout the call stops with
// info -2.  All P workgroups must be resident at once: the host keeps the workgroups of
// the launches in flight within the CU count of the device.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "paco_kernels.hpp"

namespace paco {

constexpr int RT = 512;          // threads per workgroup
constexpr int RWV = RT / 64;     // wavefronts per workgroup
constexpr int RES_PMAX = 256;    // workgroups at most
constexpr int RES_NMAX = 8 * RT; // widest matrix (8 columns per thread)

// columns per thread for N columns (a power of two); the exchange buffers hold RT of these
// per row, so every thread stores and loads its columns without a bounds test
__host__ __device__ inline int res_vc(int N)
{
    const int need = (N + RT - 1) / RT;
    return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
}

// The exchange buffers index column c = t VC + k (thread t's k-th column) at position
// k RT + t, so the VC stores or loads of a thread are each coalesced across the lanes of a
// wavefront; a reduction slice is a range of positions.
__host__ __device__ inline int res_pos(int c, int VC) { return (c % VC) * RT + c / VC; }
__host__ __device__ inline int res_col(int pos, int VC) { return (pos % RT) * VC + pos / RT; }

struct ResWs {
    uint64_t *part;  // [2][P][NP] partial GEMV sums (NP = RT * res_vc(N))
    uint64_t *sig;   // [2][P][2] partial sums of squares (FP64: low and high words)
    uint64_t *row;   // [2][NP] the next row, published by its owner
    uint64_t *gw;    // [2][NP] w = A^T v
    uint64_t *ga;    // [2][NP] the row after the left reflector
    unsigned *abort; // set on a timeout: every wait of the call returns at once
};

static size_t res_carve(int N, char *base, ResWs *W)
{
    const size_t NP = (size_t)RT * res_vc(N);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o += (bytes + 255) / 256 * 256;
        return p;
    };
    ResWs w;
    w.part = reinterpret_cast<uint64_t *>(take(8 * 2 * (size_t)RES_PMAX * NP));
    w.sig = reinterpret_cast<uint64_t *>(take(8 * 2 * (size_t)RES_PMAX * 2));
    w.row = reinterpret_cast<uint64_t *>(take(8 * 2 * NP));
    w.gw = reinterpret_cast<uint64_t *>(take(8 * 2 * NP));
    w.ga = reinterpret_cast<uint64_t *>(take(8 * 2 * NP));
    w.abort = reinterpret_cast<unsigned *>(take(256));
    if (W) *W = w;
    return o;
}

// (the launch on a trailing sub-matrix of fewer columns carves less from the same buffer)
size_t bidiag_res_ws_bytes(int N) { return res_carve(N, nullptr, nullptr); }

struct ResArgs {
    float *A;
    float *d, *e, *tauq, *taup;
    int *info;
    ResWs W;
    int M, N, ld, P;
    int S;    // positions per reduction slice (ceil(NP / P))
    int tpc;  // threads per slice column (a power of two, at most 64)
    int nq;   // partials per thread of a slice column (ceil(P / tpc))
    long long timeout; // bound of one wait, in wall-clock ticks
};

__device__ inline uint64_t rld(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void rst(uint64_t *p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint64_t tgu(unsigned bits, unsigned tag) { return ((uint64_t)tag << 32) | bits; }
__device__ inline uint64_t tgf(float v, unsigned tag) { return tgu(__float_as_uint(v), tag); }
__device__ inline float tval(uint64_t w) { return __uint_as_float((unsigned)w); }
__device__ inline bool tag_ok(uint64_t w, unsigned tag) { return (unsigned)(w >> 32) == tag; }

// Loads x[m] = p[m stride] for lo <= m < hi (the others read as 0 with the tag) and polls
// until every one carries tag; false once the call is aborted, by a timeout here or in
// another workgroup.
template <int K>
__device__ __attribute__((always_inline)) inline bool poll(const uint64_t *p, size_t stride, int lo, int hi,
                                                           unsigned tag, uint64_t (&x)[K], const ResArgs &a,
                                                           long long t0)
{
#pragma unroll
    for (int m = 0; m < K; ++m) x[m] = (m >= lo && m < hi) ? rld(p + m * stride) : tgu(0u, tag);
    for (;;) {
        bool ready = true;
#pragma unroll
        for (int m = 0; m < K; ++m) ready = ready && tag_ok(x[m], tag);
        if (ready) return true;
        if (__hip_atomic_load(a.W.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (wall_clock64() - t0 > a.timeout) {
            __hip_atomic_store(a.W.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.info, -2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int m = 0; m < K; ++m)
            if (!tag_ok(x[m], tag)) x[m] = rld(p + m * stride);
    }
}

// Sums over the 64 lanes of the VR values xs[] (VR a multiple of 8): groups of eight rows
// are reduced by a transposing butterfly (after three exchanges lane l holds row l % 8 of
// its group of eight lanes), then across the eight groups; lane k < 8 stores row g 8 + k.
// The order of the additions is fixed.
template <int VR>
__device__ __attribute__((always_inline)) inline void wave_rows_sum(const float (&xs)[VR], float *out)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < VR / 8; ++g) {
        float acc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = xs[8 * g + k];
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
        if (lane < 8) out[8 * g + lane] = tot;
    }
}

// a[k] for a uniform runtime k < K: a tree of selects on the bits of k (K a power of two),
// so the array stays in registers and no per-entry comparison is needed
template <int K>
__device__ __attribute__((always_inline)) inline float pick(const float *a, int k)
{
    if constexpr (K == 1) {
        return a[0];
    } else {
        const float lo = pick<K / 2>(a, k), hi = pick<K / 2>(a + K / 2, k);
        return (k & (K / 2)) ? hi : lo;
    }
}
// v[l][k] for a uniform runtime l (rows L0 .. L0 + CNT - 1, CNT a power of two)
template <int L0, int CNT, int VR, int VC>
__device__ __attribute__((always_inline)) inline float pick_row(const float (&v)[VR][VC], int k, int l)
{
    if constexpr (CNT == 1) {
        return v[L0][k];
    } else {
        const float lo = pick_row<L0, CNT / 2>(v, k, l), hi = pick_row<L0 + CNT / 2, CNT / 2>(v, k, l);
        return (l & (CNT / 2)) ? hi : lo;
    }
}

template <int VR, int VC>
__global__ __launch_bounds__(RT) void bidiag_res_kernel(ResArgs a)
{
    constexpr int NP = RT * VC; // row stride of the exchange buffers
    __shared__ float cb[2][VR];           // column entries c_r of the local rows
    __shared__ float xw[RWV][VR];         // per-wavefront sums of x = A u
    __shared__ float xf[VR];              // tau_p x of the local rows
    __shared__ double sd1[RWV], sd2[RWV]; // per-wavefront FP64 sums
    __shared__ float bc[2];               // broadcast entries
    const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int M = a.M, N = a.N, P = a.P;
    const size_t ld = a.ld;
    const int c0 = tid * VC; // first of the thread's columns

    float v[VR][VC]; // v[lr][k] = A[lr P + q][c0 + k]; zero on rows beyond M
#pragma unroll
    for (int lr = 0; lr < VR; ++lr) {
        const int r = lr * P + q;
        const float *Ar = a.A + (size_t)min(r, M - 1) * ld;
#pragma unroll
        for (int k = 0; k < VC; ++k) v[lr][k] = (r < M && c0 + k < N) ? Ar[c0 + k] : 0.f;
    }

    // Column jj's data after the pass of column jj - 1 (qj = jj % P, lj = jj / P: the owner
    // of row jj and its local row): c_r into LDS, the partial sum of squares, the partial
    // GEMV sums, and row jj from its owner.  Entries of columns <= jj are stored too and
    // never read.
    auto publish = [&](int jj, int qj, int lj) {
        const int buf = jj & 1;
        const unsigned tag = (unsigned)jj + 1;
        if (tid == jj / VC) {
            // c_r = A[r, jj] of the local rows (zero on the rows that are done); the row jj
            // itself is the reflector's head and takes no part in the sums
            const int kj = jj % VC;
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) cb[buf][lr] = pick<VC>(v[lr], kj);
            if (q == qj) cb[buf][lj] = 0.f;
            double ss = 0.0;
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) {
                const double c = (double)cb[buf][lr];
                ss = fma(c, c, ss);
            }
            const unsigned long long bits = (unsigned long long)__double_as_longlong(ss);
            uint64_t *sp = a.W.sig + ((size_t)buf * P + q) * 2;
            rst(sp, tgu((unsigned)bits, tag));
            rst(sp + 1, tgu((unsigned)(bits >> 32), tag));
        }
        __syncthreads();
        float cr[VR];
#pragma unroll
        for (int lr = 0; lr < VR; ++lr) cr[lr] = cb[buf][lr];
        uint64_t *pp = a.W.part + ((size_t)buf * P + q) * NP + tid;
#pragma unroll
        for (int k = 0; k < VC; ++k) {
            float g = 0.f;
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) g = fmaf(cr[lr], v[lr][k], g);
            if (c0 + k > jj) rst(pp + k * RT, tgf(g, tag)); // (only the trailing columns are read)
        }
        if (q == qj) {
            uint64_t *rp = a.W.row + (size_t)buf * NP + tid;
#pragma unroll
            for (int k = 0; k < VC; ++k) rst(rp + k * RT, tgf(pick_row<0, VR>(v, k, lj), tag));
        }
    };

    publish(0, 0, 0);
    float w[VC], u[VC];
    float beta = 0.f, tau = 0.f, inv = 1.f; // left reflector of the current column
    int qj = 0, lj = 0;                     // owner of row jj, its local row
    for (int jj = 0; jj < N; ++jj) {
        const int buf = jj & 1;
        const unsigned tag = (unsigned)jj + 1;
        // ---- exchange 1: reduce the slice's columns, left reflector of column jj
        {
            const long long t0 = wall_clock64();
            bool ok = true;
            const int cc = tid / a.tpc, s = tid % a.tpc;
            const int pos = q * a.S + cc, c = res_col(pos, VC);
            const bool live = cc < a.S && pos < NP && c > jj && c < N;
            float g = 0.f;
            if (live) {
                // partials of workgroups s + tpc m, m < nq, eight at a time, in that order
                const size_t st = (size_t)a.tpc * NP;
                const uint64_t *pc = a.W.part + ((size_t)buf * P + s) * NP + pos;
                const int nsrc = (P - s + a.tpc - 1) / a.tpc; // sources of this thread
                for (int m0 = 0; m0 < nsrc && ok; m0 += 8) {
                    uint64_t x[8];
                    ok = poll<8>(pc + m0 * st, st, 0, min(8, nsrc - m0), tag, x, a, t0);
#pragma unroll
                    for (int m = 0; m < 8; ++m) g += tval(x[m]);
                }
            }
            for (int o = 1; o < a.tpc; o <<= 1) g += __shfl_xor(g, o, 64);
            double sp = 0.0;
            if (tid < P) {
                uint64_t x[2];
                ok = ok && poll<2>(a.W.sig + ((size_t)buf * P + tid) * 2, 1, 0, 2, tag, x, a, t0);
                sp = __longlong_as_double((long long)(((uint64_t)(unsigned)x[1] << 32) | (unsigned)x[0]));
            }
            float rowc = 0.f;
            const uint64_t *rp = a.W.row + (size_t)buf * NP;
            if (live && s == 0) {
                uint64_t x[1];
                ok = ok && poll<1>(rp + pos, 0, 0, 1, tag, x, a, t0);
                rowc = tval(x[0]);
            }
            if (tid == RT - 1) {
                uint64_t x[1];
                ok = ok && poll<1>(rp + res_pos(jj, VC), 0, 0, 1, tag, x, a, t0);
                bc[0] = tval(x[0]);
            }
            const double ws = wave_sum_dpp_d(sp);
            if (lane == 0) sd1[wv] = ws;
            if (__syncthreads_or(!ok)) return;
            double sig = 0.0;
#pragma unroll
            for (int q2 = 0; q2 < RWV; ++q2) sig += sd1[q2];
            house_params_d((double)bc[0], sig, &beta, &tau, &inv);
            if (live && s == 0) {
                const float wc = rowc + inv * g;
                rst(a.W.gw + (size_t)buf * NP + pos, tgf(wc, tag));
                rst(a.W.ga + (size_t)buf * NP + pos, tgf(rowc - tau * wc, tag));
            }
        }
        // ---- exchange 2: w and row jj after the left reflector (zero outside the trailing
        // columns), right reflector of row jj
        float beta2 = 0.f, tau2 = 0.f;
        {
            const long long t0 = wall_clock64();
            const int klo = max(0, min(VC, jj + 1 - c0)), khi = max(klo, min(VC, N - c0));
            uint64_t xw_[VC], xa_[VC];
            bool ok = poll<VC>(a.W.gw + (size_t)buf * NP + tid, RT, klo, khi, tag, xw_, a, t0);
            ok = ok && poll<VC>(a.W.ga + (size_t)buf * NP + tid, RT, klo, khi, tag, xa_, a, t0);
            float ap[VC];
            double sp = 0.0;
#pragma unroll
            for (int k = 0; k < VC; ++k) {
                w[k] = tval(xw_[k]);
                ap[k] = tval(xa_[k]);
                const double t = c0 + k == jj + 1 ? 0.0 : (double)ap[k];
                sp = fma(t, t, sp);
            }
            if (tid == (jj + 1) / VC) bc[1] = pick<VC>(ap, (jj + 1) % VC);
            const double ws = wave_sum_dpp_d(sp);
            if (lane == 0) sd2[wv] = ws;
            if (__syncthreads_or(!ok)) return;
            float inv2 = 1.f;
            if (jj + 1 < N) {
                double sig2 = 0.0;
#pragma unroll
                for (int q2 = 0; q2 < RWV; ++q2) sig2 += sd2[q2];
                house_params_d((double)bc[1], sig2, &beta2, &tau2, &inv2);
            }
#pragma unroll
            for (int k = 0; k < VC; ++k) u[k] = c0 + k == jj + 1 ? 1.f : ap[k] * inv2;
        }
        // ---- outputs of column jj; row jj is final (its registers are zeroed)
        if (q == 0 && tid == 0) {
            a.d[jj] = beta;
            a.tauq[jj] = tau;
            if (jj + 1 < N) {
                a.e[jj] = beta2;
                a.taup[jj] = tau2;
            }
        }
        if (q == qj) {
            float *Ar = a.A + (size_t)jj * ld + c0;
#pragma unroll
            for (int k = 0; k < VC; ++k) {
                const int c = c0 + k;
                const float x = c < jj ? pick_row<0, VR>(v, k, lj) : c == jj ? beta : c == jj + 1 ? beta2 : u[k];
                if (c < N) Ar[k] = x;
            }
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) {
                const bool z = lr == lj;
#pragma unroll
                for (int k = 0; k < VC; ++k) v[lr][k] = z ? 0.f : v[lr][k];
            }
        }
        if (jj + 1 == N) break;
        // ---- pass: both reflectors on the rows r > jj (the others are zero)
        {
            float vr[VR];
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) vr[lr] = cb[buf][lr] * inv; // 0 on rows r <= jj
            if (wv == (jj / VC) >> 6) {
                // the stored reflector entries of column jj (its owner thread)
#pragma unroll
                for (int k = 0; k < VC; ++k) {
                    const bool own = c0 + k == jj;
#pragma unroll
                    for (int lr = 0; lr < VR; ++lr) v[lr][k] = own ? vr[lr] : v[lr][k];
                }
            }
            float xs[VR];
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) {
                const float tv = tau * vr[lr];
                float x = 0.f;
#pragma unroll
                for (int k = 0; k < VC; ++k) {
                    v[lr][k] = fmaf(-tv, w[k], v[lr][k]);
                    x = fmaf(v[lr][k], u[k], x);
                }
                xs[lr] = x;
            }
            wave_rows_sum<VR>(xs, &xw[wv][0]);
            __syncthreads();
            if (tid < VR) {
                float x = 0.f;
#pragma unroll
                for (int q2 = 0; q2 < RWV; ++q2) x += xw[q2][tid];
                xf[tid] = tau2 * x;
            }
            __syncthreads();
#pragma unroll
            for (int lr = 0; lr < VR; ++lr) {
                const float tx = xf[lr];
#pragma unroll
                for (int k = 0; k < VC; ++k) v[lr][k] = fmaf(-tx, u[k], v[lr][k]);
            }
        }
        if (++qj == P) {
            qj = 0;
            ++lj;
        }
        publish(jj + 1, qj, lj);
    }
    // column N-1's reflector entries below row N-1, then the rows >= N
    {
        const int jj = N - 1, buf = jj & 1;
#pragma unroll
        for (int lr = 0; lr < VR; ++lr) {
            const int r = lr * P + q;
            if (r >= N && r < M) {
                const float vr = cb[buf][lr] * inv;
                float *Ar = a.A + (size_t)r * ld + c0;
#pragma unroll
                for (int k = 0; k < VC; ++k) {
                    const int c = c0 + k;
                    if (c < N) Ar[k] = c == jj ? vr : v[lr][k];
                }
            }
        }
    }
}

// zeroes the exchange buffers (a fill kernel, so the sequence can be captured in a graph)
__global__ void res_reset_kernel(uint4 *p, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct ResShape {
    int VR, VC, P;
};

static int device_cus()
{
    static int cus = -1;
    if (cus < 0) {
        int dev = 0, n = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 1;
        cus = n;
    }
    return cus;
}

// Register shape and workgroup count for an M x N matrix: VC columns per thread (N <=
// RT VC), VR rows per workgroup (M <= P VR), at most 128 matrix entries per thread.
// preq workgroups are asked for (fewer when fewer rows suffice); more when the matrix does
// not fit in preq workgroups' registers.  false: no shape holds the matrix on this chip.
static bool res_pick(int M, int N, int preq, ResShape &sh)
{
    if (N < 2 || N > RES_NMAX || M < N) return false;
    const int need = (N + RT - 1) / RT;
    const int vc = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
    // PACO_RES_PMAX caps the workgroups (tests of the panels-then-resident composition)
    const char *ecap = getenv("PACO_RES_PMAX");
    const int pmax = std::min(std::min(RES_PMAX, device_cus()), ecap ? std::max(1, atoi(ecap)) : RES_PMAX);
    const int P0 = std::max(1, std::min(preq, pmax));
    static const int vrs[3] = {8, 16, 32};
    // the most workgroups up to the asked-for count whose registers hold the matrix ...
    for (int vr : vrs) {
        if (vr * vc > 128) break;
        if (vr * vc == 64 && vc == 8) continue; // (M <= 8 P < N cannot occur: M >= N)
        const int p = (M + vr - 1) / vr;       // workgroups that vr rows each need
        if (p > P0 || (RT * vc + p - 1) / p > RT) continue; // (a reduction slice fits one workgroup)
        sh.VR = vr;
        sh.VC = vc;
        sh.P = std::min(P0, p);
        return true;
    }
    // ... else the fewest that do (the largest register share per workgroup)
    for (int i = 2; i >= 0; --i) {
        const int vr = vrs[i];
        if (vr * vc > 128 || (vr * vc == 64 && vc == 8)) continue;
        const int p = (M + vr - 1) / vr;
        if (p > pmax || (RT * vc + p - 1) / p > RT) continue;
        sh.VR = vr;
        sh.VC = vc;
        sh.P = p;
        return true;
    }
    return false;
}

// Workgroups of the resident launches in flight per device: a launch waits (on its stream)
// for the oldest ones until its own fit the CU count, so every launch's workgroups can be
// resident together.  Launches recorded into a graph are not tracked.
struct ResFlight {
    hipEvent_t ev;
    int P, dev;
};
static std::mutex g_res_mu;
static std::vector<ResFlight> g_res_flights;
static std::vector<hipEvent_t> g_res_spare;

static hipError_t res_admit(hipStream_t s, int dev, int P, int cus, bool *track)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) cs = hipStreamCaptureStatusNone;
    *track = cs == hipStreamCaptureStatusNone;
    if (!*track) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_res_mu);
    int used = 0;
    for (size_t i = 0; i < g_res_flights.size();) {
        ResFlight &f = g_res_flights[i];
        if (f.dev == dev && hipEventQuery(f.ev) == hipSuccess) {
            g_res_spare.push_back(f.ev);
            g_res_flights.erase(g_res_flights.begin() + i);
            continue;
        }
        if (f.dev == dev) used += f.P;
        ++i;
    }
    for (size_t i = 0; i < g_res_flights.size() && used + P > cus;) {
        ResFlight &f = g_res_flights[i];
        if (f.dev != dev) {
            ++i;
            continue;
        }
        hipError_t e = hipStreamWaitEvent(s, f.ev, 0);
        if (e != hipSuccess) return e;
        used -= f.P;
        g_res_spare.push_back(f.ev);
        g_res_flights.erase(g_res_flights.begin() + i);
    }
    return hipSuccess;
}

static hipError_t res_track(hipStream_t s, int dev, int P)
{
    std::lock_guard<std::mutex> lk(g_res_mu);
    hipEvent_t ev = nullptr;
    if (!g_res_spare.empty()) {
        ev = g_res_spare.back();
        g_res_spare.pop_back();
    } else {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) {
        g_res_spare.push_back(ev);
        return e;
    }
    g_res_flights.push_back({ev, P, dev});
    return hipSuccess;
}

template <int VR, int VC>
static void res_launch(const ResArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((bidiag_res_kernel<VR, VC>), dim3(a.P), dim3(RT), 0, s, a);
}

// Workgroups the resident kernel would take for an M x N matrix when preq are asked for
// (0: none fit).
int bidiag_res_workgroups(int M, int N, int preq)
{
    ResShape sh;
    return res_pick(M, N, preq, sh) ? sh.P : 0;
}

hipError_t launch_bidiag_res(const ProbDesc &P, void *ws, hipStream_t s, int preq)
{
    ResShape sh;
    if (!res_pick(P.M, P.N, preq > 0 ? preq : RES_PMAX, sh)) return hipErrorInvalidValue;
    static long long khz = 0;
    if (!khz) {
        int dev = 0, k = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&k, hipDeviceAttributeWallClockRate, dev) != hipSuccess || k <= 0) k = 100000;
        khz = k;
    }
    // 2 s per wait; PACO_RES_TIMEOUT_US sets it (tests of the timeout path)
    const char *et = getenv("PACO_RES_TIMEOUT_US");
    const long long timeout = et ? std::max(0LL, atoll(et)) * khz / 1000 : 2000LL * khz;
    ResArgs a;
    a.A = P.a;
    a.d = P.d;
    a.e = P.e;
    a.tauq = P.tauq;
    a.taup = P.taup;
    a.info = P.info;
    res_carve(P.N, static_cast<char *>(ws), &a.W);
    a.M = P.M;
    a.N = P.N;
    a.ld = P.ld ? P.ld : P.N;
    a.P = sh.P;
    a.S = (RT * res_vc(P.N) + sh.P - 1) / sh.P;
    a.tpc = 64;
    while (a.tpc > 1 && a.tpc * a.S > RT) a.tpc >>= 1;
    a.nq = (sh.P + a.tpc - 1) / a.tpc;
    a.timeout = timeout;
    const size_t bytes = bidiag_res_ws_bytes(P.N);
    hipLaunchKernelGGL(res_reset_kernel, dim3(256), dim3(256), 0, s, reinterpret_cast<uint4 *>(ws), bytes / 16);
    int dev = 0;
    (void)hipGetDevice(&dev);
    bool track = false;
    hipError_t e = res_admit(s, dev, sh.P, device_cus(), &track);
    if (e != hipSuccess) return e;
switch (sh.VR * 100 + sh.VC) {
    case 801: res_launch<8, 1>(a, s); break;
    case 802: res_launch<8, 2>(a, s); break;
    case 804: res_launch<8, 4>(a, s); break;
    case 1601: res_launch<16, 1>(a, s); break;
    case 1602: res_launch<16, 2>(a, s); break;
    case 1604: res_launch<16, 4>(a, s); break;
    case 1608: res_launch<16, 8>(a, s); break;
    case 3201: res_launch<32, 1>(a, s); break;
    case 3202: res_launch<32, 2>(a, s); break;
    case 3204: res_launch<32, 4>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return track ? res_track(s, dev, sh.P) : hipSuccess;
}

} // namespace paco
