// One 64x64 output tile of C = alpha * op(A) * op(B) (+ beta C) on the gfx950 matrix
// cores; shared by the grouped GEMM (gemm.hip), the merge products and the
// back-transforms (dc.hip).
//
// Workgroup of 256 threads = four wavefronts, each owning a 32x32 sub-tile.
// K advances in steps of TK (32, or 16 for short K) through LDS-staged 64xTK (A) and
// TKx64 (B) tiles; the operands of the next step are loaded into registers while the
// current step's products run.
//  mode 0: v_mfma_f32_32x32x2_f32 (FP32 matrix cores)
//  mode 1: BF16x3 compensated: operands split x = h + m + l in bf16, six products
//          with v_mfma_f32_32x32x8_bf16 accumulated in FP32 (smallest terms first)
//  mode 2: plain FP32 FMA (reference path for debugging)
#pragma once
#include "paco_common.hpp"

namespace paco {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef short shortx4 __attribute__((ext_vector_type(4)));

#define GT 256

struct TileCtx {
    const float *A, *B;
    float *C;
    const int *brow, *crow;
    int lda, ldb, ldc, m, n, K, transA, transB;
    float alpha, beta;
    int m0, n0;
};

// elements of each operand per thread and K step
template <int TK>
struct TileEl {
    static constexpr int n = (64 * TK) / GT;
};

// operands of the K step at k0 into registers (zero outside the matrices)
template <int TK>
__device__ inline void tile_fetch(const TileCtx &t, int k0, float (&ra)[TileEl<TK>::n], float (&rb)[TileEl<TK>::n])
{
    constexpr int TILE_EL = TileEl<TK>::n;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < TILE_EL; ++e) {
        const int q = tid + e * GT;
        int r, kk;
        if (t.transA) {
            r = q % 64;
            kk = q / 64;
        } else {
            kk = q % TK;
            r = q / TK;
        }
        const int gr = t.m0 + r, gk = k0 + kk;
        float v = 0.f;
        if (gr < t.m && gk < t.K) v = t.transA ? gmem(t.A)[(size_t)gk * t.lda + gr] : gmem(t.A)[(size_t)gr * t.lda + gk];
        ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < TILE_EL; ++e) {
        const int q = tid + e * GT;
        int c, kk;
        if (t.transB) {
            kk = q % TK;
            c = q / TK;
        } else {
            c = q % 64;
            kk = q / 64;
        }
        const int gc = t.n0 + c, gk = k0 + kk;
        float v = 0.f;
        if (gc < t.n && gk < t.K) {
            const int krow = t.brow ? gmem(t.brow)[gk] : gk;
            v = t.transB ? gmem(t.B)[(size_t)gc * t.ldb + krow] : gmem(t.B)[(size_t)krow * t.ldb + gc];
        }
        rb[e] = v;
    }
}

// registers of tile_fetch into the LDS tiles As[k][row], Bs[k][col]
template <int TK>
__device__ inline void tile_put(const TileCtx &t, const float (&ra)[TileEl<TK>::n], const float (&rb)[TileEl<TK>::n],
                                float (*As)[64 + 4], float (*Bs)[64 + 4])
{
    constexpr int TILE_EL = TileEl<TK>::n;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < TILE_EL; ++e) {
        const int q = tid + e * GT;
        if (t.transA)
            As[q / 64][q % 64] = ra[e];
        else
            As[q % TK][q / TK] = ra[e];
        if (t.transB)
            Bs[q % TK][q / TK] = rb[e];
        else
            Bs[q / 64][q % 64] = rb[e];
    }
}

__device__ inline void tile_store(const TileCtx &t, int r, int c, float v)
{
    const int gr = t.m0 + r, gc = t.n0 + c;
    if (gr < t.m && gc < t.n) {
        const int cr = t.crow ? gmem(t.crow)[gr] : gr;
        __attribute__((address_space(1))) float *p = gmem(t.C) + (size_t)cr * t.ldc + gc;
        float o = t.alpha * v;
        if (t.beta != 0.f) o += t.beta * *p;
        *p = o;
    }
}

__device__ inline unsigned short bf16_bits(float x)
{
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf16_float(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

__device__ inline void split3(float x, unsigned short &h, unsigned short &m, unsigned short &l)
{
    h = bf16_bits(x);
    const float r1 = x - bf16_float(h);
    m = bf16_bits(r1);
    const float r2 = r1 - bf16_float(m);
    l = bf16_bits(r2);
}

// Whole tile: loads, products, stores.  Every thread of the workgroup must call it.
template <int MODE, int TK = 32>
__device__ inline void tile_gemm(const TileCtx &t)
{
    __shared__ float As[TK][64 + 4];
    __shared__ float Bs[TK][64 + 4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    float ra[TileEl<TK>::n], rb[TileEl<TK>::n];
    tile_fetch<TK>(t, 0, ra, rb);
    if (MODE == 2) {
        const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
        float acc[4][4] = {};
        for (int k0 = 0; k0 < t.K; k0 += TK) {
            tile_put<TK>(t, ra, rb, As, Bs);
            __syncthreads();
            if (k0 + TK < t.K) tile_fetch<TK>(t, k0 + TK, ra, rb);
#pragma unroll
            for (int kk = 0; kk < TK; ++kk)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] += As[kk][ty * 4 + i] * Bs[kk][tx * 4 + j];
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) tile_store(t, ty * 4 + i, tx * 4 + j, acc[i][j]);
        return;
    }
    floatx16 acc;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.f;
    for (int k0 = 0; k0 < t.K; k0 += TK) {
        tile_put<TK>(t, ra, rb, As, Bs);
        __syncthreads();
        if (k0 + TK < t.K) tile_fetch<TK>(t, k0 + TK, ra, rb);
        if (MODE == 0) {
#pragma unroll
            for (int kk = 0; kk < TK; kk += 2) {
                const float a = As[kk + (lane >> 5)][wm * 32 + (lane & 31)];
                const float b = Bs[kk + (lane >> 5)][wn * 32 + (lane & 31)];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < TK; kk += 8) {
                shortx4 ah, am, al, bh, bm, bl;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int kq = kk + 4 * (lane >> 5) + q;
                    unsigned short h, mm, l;
                    split3(As[kq][wm * 32 + (lane & 31)], h, mm, l);
                    ah[q] = (short)h;
                    am[q] = (short)mm;
                    al[q] = (short)l;
                    split3(Bs[kq][wn * 32 + (lane & 31)], h, mm, l);
                    bh[q] = (short)h;
                    bm[q] = (short)mm;
                    bl[q] = (short)l;
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(am, bm, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al, bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(am, bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bm, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int r = wm * 32 + (j >> 2) * 8 + (lane >> 5) * 4 + (j & 3);
        const int c = wn * 32 + (lane & 31);
        tile_store(t, r, c, acc[j]);
    }
}

} // namespace paco
