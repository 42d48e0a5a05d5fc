// Complex FP64 products on the CDNA matrix cores (v_mfma_f64_16x16x4f64).
//
// One wavefront forms a 32 x 32 block of C = X Y (complex, FP64) as 2 x 2
// output blocks of 16 x 16, real and imaginary parts in separate
// accumulators, four real MFMAs per k-step of 4:
//   Cr += Xr Yr - Xi Yi,  Ci += Xr Yi + Xi Yr.
// Operand layout of v_mfma_f64_16x16x4f64 (lane l = 0..63):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l, register v holds D[l / 16 + 4 v][l % 16]
// so lane l owns, for output block (bi, bj) and v = 0..3, the element
//   row = 16 bi + l / 16 + 4 v,  column = 16 bj + l % 16
// of the 32 x 32 block.
#pragma once
#include "wbmi_common.h"

typedef double wbmi_v4d __attribute__((ext_vector_type(4)));

struct MfmaBlock {
    wbmi_v4d re[2][2], im[2][2];
    __device__ inline void zero() {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                re[bi][bj] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
                im[bi][bj] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
            }
    }
    __device__ inline cplx<double> get(int bi, int bj, int v) const { return mk(re[bi][bj][v], im[bi][bj][v]); }
};

// row / column (within the 32 x 32 block) of the lane's element (bi, bj, v)
__device__ inline int mfma_row(int bi, int v) { return 16 * bi + (int)(threadIdx.x & 63) / 16 + 4 * v; }
__device__ inline int mfma_col(int bj) { return 16 * bj + (int)(threadIdx.x & 15); }

// C += X Y over k < nk for the block at rows r0.., columns c0.. of the full
// matrices; X(r, k) and Y(k, c) are called only with r, c < nr, nc and k < nk
// (other operand entries are zero).
template <class FX, class FY>
__device__ __attribute__((always_inline)) void mfma_block(MfmaBlock& C, int nr, int nc, int nk, int r0, int c0, FX X,
                                                          FY Y) {
    const int li = threadIdx.x & 15, lk = (threadIdx.x & 63) >> 4;
    bool rok[2], cok[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        rok[b] = r0 + 16 * b + li < nr;
        cok[b] = c0 + 16 * b + li < nc;
    }
#pragma unroll 2
    for (int k0 = 0; k0 < nk; k0 += 4) {
        const int k = k0 + lk;
        const bool kok = k < nk;
        cplx<double> a[2], b[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            a[q] = (kok && rok[q]) ? X(r0 + 16 * q + li, k) : mk(0.0, 0.0);
            b[q] = (kok && cok[q]) ? Y(k, c0 + 16 * q + li) : mk(0.0, 0.0);
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                C.re[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi].re, b[bj].re, C.re[bi][bj], 0, 0, 0);
                C.re[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[bi].im, b[bj].im, C.re[bi][bj], 0, 0, 0);
                C.im[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi].re, b[bj].im, C.im[bi][bj], 0, 0, 0);
                C.im[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi].im, b[bj].re, C.im[bi][bj], 0, 0, 0);
            }
    }
}

// One 16 x 32 half of such a block: rows r0 .. r0 + 15 (a single output row
// block), columns c0 .. c0 + 31; lane l owns row l / 16 + 4 v, column 16 bj + l % 16.
struct MfmaHalf {
    wbmi_v4d re[2], im[2];
    __device__ inline void zero() {
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            re[bj] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
            im[bj] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
        }
    }
    __device__ inline cplx<double> get(int bj, int v) const { return mk(re[bj][v], im[bj][v]); }
};

__device__ inline int mfma_hrow(int v) { return (int)(threadIdx.x & 63) / 16 + 4 * v; }

template <class FX, class FY>
__device__ __attribute__((always_inline)) void mfma_half(MfmaHalf& C, int nr, int nc, int nk, int r0, int c0, FX X, FY Y) {
    const int li = threadIdx.x & 15, lk = (threadIdx.x & 63) >> 4;
    const bool rok = r0 + li < nr;
    bool cok[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) cok[b] = c0 + 16 * b + li < nc;
#pragma unroll 2
    for (int k0 = 0; k0 < nk; k0 += 4) {
        const int k = k0 + lk;
        const bool kok = k < nk;
        const cplx<double> a = (kok && rok) ? X(r0 + li, k) : mk(0.0, 0.0);
        cplx<double> b[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) b[q] = (kok && cok[q]) ? Y(k, c0 + 16 * q + li) : mk(0.0, 0.0);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            C.re[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b[bj].re, C.re[bj], 0, 0, 0);
            C.im[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b[bj].im, C.im[bj], 0, 0, 0);
        }
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            C.re[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b[bj].im, C.re[bj], 0, 0, 0);
            C.im[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b[bj].re, C.im[bj], 0, 0, 0);
        }
    }
}

