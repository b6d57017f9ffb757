// Grouped GEMM over explicit descriptors (tests, blocked bidiagonalization updates).
// The merge products and back-transforms of the D&C derive their operands on the
// device instead (dc.hip); all share the tile of gemm_tile.hpp.
#include "gemm_tile.hpp"
#include "paco_kernels.hpp"

namespace paco {

static int g_gemm_mode = 0;
void gemm_set_mode(int mode) { g_gemm_mode = mode; }
int gemm_get_mode() { return g_gemm_mode; }

template <int MODE>
__global__ __launch_bounds__(GT) void gemm_grouped_kernel(const GemmDesc *descs, int ndesc)
{
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (descs[mid].tile_off <= (int)blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const GemmDesc &G = descs[lo];
    const int tile = blockIdx.x - G.tile_off;
    TileCtx t;
    t.A = G.A;
    t.B = G.B;
    t.C = G.C;
    t.brow = G.brow;
    t.crow = G.crow;
    t.lda = G.lda;
    t.ldb = G.ldb;
    t.ldc = G.ldc;
    t.m = G.m;
    t.n = G.n;
    t.K = G.k;
    t.transA = G.transA;
    t.transB = G.transB;
    t.alpha = G.alpha;
    t.beta = G.beta;
    if (G.dyn_mk) {
        const int kk = *G.dyn_mk;
        t.m = min(t.m, kk);
        t.K = min(t.K, kk);
    }
    t.m0 = (tile / G.tiles_n) * 64;
    t.n0 = (tile % G.tiles_n) * 64;
    if (t.m0 >= t.m || t.n0 >= t.n) return;
    tile_gemm<MODE>(t);
}

__global__ __launch_bounds__(GT) void copy2d_kernel(const CopyDesc *descs, int ndesc)
{
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (descs[mid].row_off <= (int)blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const CopyDesc &D = descs[lo];
    const int r = blockIdx.x - D.row_off;
    if (r >= D.rows) return;
    for (int c = threadIdx.x; c < D.cols; c += GT) D.dst[(size_t)r * D.ldd + c] = D.src[(size_t)r * D.lds + c];
}

hipError_t launch_gemm_grouped(const GemmDesc *descs, int ndesc, int ntiles, hipStream_t s)
{
    if (ndesc <= 0 || ntiles <= 0) return hipSuccess;
    if (g_gemm_mode == 1)
        hipLaunchKernelGGL(gemm_grouped_kernel<1>, dim3(ntiles), dim3(GT), 0, s, descs, ndesc);
    else if (g_gemm_mode == 2)
        hipLaunchKernelGGL(gemm_grouped_kernel<2>, dim3(ntiles), dim3(GT), 0, s, descs, ndesc);
    else
        hipLaunchKernelGGL(gemm_grouped_kernel<0>, dim3(ntiles), dim3(GT), 0, s, descs, ndesc);
    return hipGetLastError();
}

hipError_t launch_copy2d(const CopyDesc *descs, int ndesc, int nrows, hipStream_t s)
{
    if (ndesc <= 0 || nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy2d_kernel, dim3(nrows), dim3(GT), 0, s, descs, ndesc);
    return hipGetLastError();
}

} // namespace paco
