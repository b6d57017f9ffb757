// Kernel argument structs and host-side launchers of the gfx950 kernels.
#pragma once
#include "paco_common.hpp"

namespace paco {

// Problems of a launch: either one problem passed by value (plain entry points, no
// device descriptor array needed) or a device array (batched path).
struct ProbList {
    const ProbDesc *ptr;
    ProbDesc single;
    int count;
    __host__ __device__ inline ProbDesc get(int i) const { return ptr ? ptr[i] : single; }
};

// One D&C level (merge round r) over every problem of the launch.  Merge t belongs to
// problem p with task_off[p] <= t < task_off[p+1] (task_off may be null for one
// problem); its columns follow from (N, r, t - task_off[p]).  task_prob[t] = p when
// given (one load instead of a binary search over task_off per workgroup).  stride bounds
// the rows and columns of every merge of the round (min(3 * 2^r, largest N of the launch)).
struct LevelArgs {
    const int *task_off;
    const int *task_prob; // [ntask] or null
    MergeState *states; // [ntask]
    int ntask, r, stride;
    int maxn; // largest merged row count of the level (sizes the deflation LDS)
    int maxit; // iteration limit of the secular solver per root (a root that reaches it adds 1 to info)
    int vonly; // singular values only: nodes keep the first and last columns of their right
               // vectors (all a merge reads), no left vectors
    int ls;    // leaf size of the partition (columns per leaf, LEAF_MIN..LEAF_MAX)
};

struct CopyDesc {
    const float *src;
    float *dst;
    int lds, ldd, rows, cols;
    int row_off; // prefix over descs of rows
};

// Raise a kernel's dynamic-LDS limit to what the device allows beyond the kernel's
// static LDS; returns that limit (bytes).  Clears the runtime's last error.
inline size_t allow_dynamic_lds(const void *func)
{
    int dev = 0, optin = 0, basic = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev);
    (void)hipDeviceGetAttribute(&basic, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const int maxb = optin > basic ? optin : basic;
    hipFuncAttributes fa;
    size_t stat = 0;
    if (hipFuncGetAttributes(&fa, func) == hipSuccess) stat = fa.sharedSizeBytes;
    const size_t dyn = maxb > (int)stat ? (size_t)maxb - stat : 0;
    (void)hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    (void)hipGetLastError();
    return dyn;
}

// autotune.cpp: the autotuner's table (fallback when the table has no entry)
int tuned_bidiag_threads(int M, int N);
int tuned_res_threads(int cls, int fallback);
int tuned_leaf_size(int fallback);
int tuned_grid_tiles(int M, int N, int fallback);
int tuned_grid_switch(int M, int N, int fallback);

// bidiag.hip
// one-workgroup bidiagonalization with nt threads; lds bytes of dynamic LDS hold the
// vectors and as many bottom rows of each matrix as fit
size_t bidiag_wg_lds_bytes(int M, int N, int nt = 256);          // vectors only
size_t bidiag_wg_lds_resident_bytes(int M, int N, int nt = 256); // whole matrix in LDS
size_t bidiag_wg_lds_limit();
constexpr int BIDIAG_WG_MAXN = 512; // widest matrix of the one-workgroup kernel
hipError_t launch_bidiag_wg(ProbList pl, const int *order, int count, size_t lds, hipStream_t s, int nt = 256);
// bidiag_blk.hip: blocked one-workgroup bidiagonalization (panels of BK_NB columns, the
// trailing matrix written once per panel) for streamed matrices, N <= 512, M <= 1024
constexpr int BK_NB = 8;
constexpr int BIDIAG_BLK_MAXM = 1024;
size_t bidiag_blk_lds_bytes(int M, int N); // vectors only; the rest of the LDS caches rows
bool bidiag_blk_supported(int M, int N);
hipError_t launch_bidiag_blk(ProbList pl, const int *order, int count, size_t lds, hipStream_t s);
// one wavefront per matrix, whole matrix in LDS (batched path), N <= 64 * BIDIAG_WAVE_MAXC
constexpr int BIDIAG_WAVE_MAXC = 4;
size_t bidiag_wave_lds(int M, int N);
size_t bidiag_wave_lds_limit();
hipError_t launch_bidiag_wave(ProbList pl, const int *order, int count, int C, size_t lds, hipStream_t s);
hipError_t launch_tfactor(ProbList pl, const int *blk_off, int nblocks, int side, hipStream_t s);
// strips of sw columns (64 or 16); strip_off counts strips of that width
hipError_t launch_wy_apply(ProbList pl, const int *strip_off, int nstrips, int side, hipStream_t s, int sw = 64);
// Row-wise kernels (identity, copies) take ROW_GROUP rows per workgroup: one workgroup
// per row made the launch bound by workgroup dispatch.  Their offsets count groups.
constexpr int ROW_GROUP = 8;
// batches: groups of 32 rows (a matrix of the batch is small, so 8-row groups made
// hundreds of thousands of workgroups of a few KiB each)
constexpr int BATCH_ROW_GROUP = 32;
static_assert(ROW_GROUP * 32 == 256, "row-wise kernels: 32 threads per row of a 256-thread workgroup");
// n contiguous floats by one workgroup of 256 threads: 16-byte vectors when both ends are
// aligned, else single floats
__device__ inline void copy_block(float *dst, const float *src, int n)
{
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
        const int n4 = n >> 2;
        float4 *d4 = reinterpret_cast<float4 *>(dst);
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
        for (int q = threadIdx.x; q < n4; q += 256) d4[q] = s4[q];
        for (int q = 4 * n4 + threadIdx.x; q < n; q += 256) dst[q] = src[q];
    } else {
        for (int q = threadIdx.x; q < n; q += 256) dst[q] = src[q];
    }
}
// X = I; grp_off counts groups of `rows` rows per problem (BATCH_ROW_GROUP for batches;
// ROW_GROUP for one large matrix)
hipError_t launch_identity(ProbList pl, const int *grp_off, int ngroups, int side, hipStream_t s, int rows);
// in-place transpose of vt by 32 x 32 tile pairs; pair_off counts groups of `per` pairs
// (TP_GROUP for batches, 1 for one large matrix)
constexpr int TP_GROUP = 4;
hipError_t launch_transpose_vt(ProbList pl, const int *pair_off, int ngroups, hipStream_t s, int per);

// bidiag_grid.hip (large single matrices)
size_t bidiag_grid_ws_bytes(int M, int N);
// whether the LDS of the chip-wide path's kernels holds an M x N problem
bool bidiag_grid_supported(int M, int N);
// grid: tiles per side of the panel kernel's tile grid (0: the default, 16, or PACO_GRID);
// no_wait: only kernels whose workgroups never wait on each other (panels launched phase
// by phase, then the one-workgroup kernel)
hipError_t launch_bidiag_grid(const ProbDesc &P, void *ws, hipStream_t s, int grid = 0, bool no_wait = false);
// bidiag_res.hip: the whole matrix in the registers of P workgroups (one per CU)
size_t bidiag_res_ws_bytes(int N);
// workgroups the kernel takes for an M x N matrix when preq are asked for (0: it does not
// fit the registers of the chip)
int bidiag_res_workgroups(int M, int N, int preq);
hipError_t launch_bidiag_res(const ProbDesc &P, void *ws, hipStream_t s, int preq);

// dc.hip
// leaves of ls columns (leaf_off: prefix of the leaf counts over the problems, or null)
hipError_t launch_dc_leaves(ProbList pl, const int *leaf_off, int nleaf, int ls, hipStream_t s);
hipError_t launch_dc_level(ProbList pl, LevelArgs la, hipStream_t s);
// back-transform: U[:, :N] <- U[:, :N] qt^T and Vt <- wt Vt for every problem; the row
// offsets count groups of `rows` rows
hipError_t launch_dc_backtransform(ProbList pl, const int *tile_off_u, int ntiles_u, const int *tile_off_v,
                                   int ntiles_v, const int *row_off_u, int nrows_u, const int *row_off_v,
                                   int nrows_v, int rows, hipStream_t s);

// qp_gemm.hip: Q, P of one large matrix by blocked GEMMs (X already the identity, T factors
// formed)
constexpr int QP_MAX_SPLITS = 16;
size_t qp_gemm_ws_bytes(int M, int N);
hipError_t launch_qp_gemm(const ProbDesc &P, int side, char *ws, hipStream_t st);
// batched: per-problem scratch qp[pid] of qpb_scratch_floats(M, N) floats; round r applies
// block nblk - 1 - r of every problem that has it
size_t qpb_scratch_floats(int M, int N);
hipError_t launch_qpb_round(ProbList pl, float *const *qp, int side, int round, const int *ext_off, int next,
                            const int *g1_off, int ng1, const int *g3_off, int ng3, hipStream_t st);

// gemm.hip
hipError_t launch_gemm_grouped(const GemmDesc *descs, int ndesc, int ntiles, hipStream_t s);
hipError_t launch_copy2d(const CopyDesc *descs, int ndesc, int nrows, hipStream_t s);
void gemm_set_mode(int mode);
int gemm_get_mode();

} // namespace paco
