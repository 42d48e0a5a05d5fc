// Gauge rotation Xbar = U^+ X U of libwbmi (data_K/data_K.py:130-132).
//
// One 256-thread workgroup per k-point.  U stays in LDS for all channels of
// the k-point; each channel is loaded once from HBM with coalesced reads into
// LDS (Hermitised there when flagged), T = X U is formed in LDS and
// Xbar = U^+ T is written straight back from registers.  Both products use
// TS x TS register tiles on a 16 x 16 thread grid (TS = 2 for nw <= 32,
// 4 for nw <= 64).  LDS rows are padded by one element.
#include "wbmi_common.h"

#define ROT_THREADS 256

template <typename T, int TS>
__device__ inline void rot_channel(int nw, int ld, const cplx<T>* sU, const cplx<T>* sX, int xld, cplx<T>* sT,
                                   bool herm, cplx<T>* Xc) {
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    const int r0 = TS * tr, c0 = TS * tc;
    const bool active = r0 < nw && c0 < nw;
    cplx<T> acc[TS][TS];
    // T = X U
    if (active) {
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j) acc[i][j] = mk((T)0, (T)0);
        for (int m = 0; m < nw; ++m) {
            cplx<T> x[TS], u[TS];
#pragma unroll
            for (int i = 0; i < TS; ++i) {
                const int r = r0 + i < nw ? r0 + i : nw - 1;
                cplx<T> v = sX[r * xld + m];
                if (herm) {
                    const cplx<T> w = sX[m * xld + r];
                    v = mk((T)0.5 * (v.re + w.re), (T)0.5 * (v.im - w.im));
                }
                x[i] = v;
            }
#pragma unroll
            for (int j = 0; j < TS; ++j) u[j] = sU[m * ld + (c0 + j < nw ? c0 + j : nw - 1)];
#pragma unroll
            for (int i = 0; i < TS; ++i)
#pragma unroll
                for (int j = 0; j < TS; ++j) acc[i][j] += x[i] * u[j];
        }
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j)
                if (r0 + i < nw && c0 + j < nw) sT[(r0 + i) * ld + c0 + j] = acc[i][j];
    }
    __syncthreads();
    // Xbar = U^+ T
    if (active) {
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j) acc[i][j] = mk((T)0, (T)0);
        for (int m = 0; m < nw; ++m) {
            cplx<T> u[TS], t[TS];
#pragma unroll
            for (int i = 0; i < TS; ++i) u[i] = sU[m * ld + (r0 + i < nw ? r0 + i : nw - 1)];
#pragma unroll
            for (int j = 0; j < TS; ++j) t[j] = sT[m * ld + (c0 + j < nw ? c0 + j : nw - 1)];
#pragma unroll
            for (int i = 0; i < TS; ++i)
#pragma unroll
                for (int j = 0; j < TS; ++j) acc[i][j] += cmul(u[i], t[j]);
        }
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j)
                if (r0 + i < nw && c0 + j < nw) Xc[(r0 + i) * nw + c0 + j] = acc[i][j];
    }
}

// stage_x = 0: X is read from HBM (three padded nw x nw tiles would not fit in LDS)
template <typename T>
__global__ void __launch_bounds__(ROT_THREADS) k_rotate(int nw, int nch, const cplx<T>* __restrict__ U,
                                                        cplx<T>* X, uint64_t herm_lo, uint64_t herm_hi,
                                                        uint64_t skip_lo, uint64_t skip_hi, int stage_x) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int ld = nw + 1;
    cplx<T>* sU = (cplx<T>*)smem;
    cplx<T>* sT = sU + nw * ld;
    cplx<T>* sX = sT + nw * ld;
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    const int nw2 = nw * nw;
    const cplx<T>* Uk = U + k * nw2;
    for (int t = tid; t < nw2; t += ROT_THREADS) sU[(t / nw) * ld + t % nw] = Uk[t];
    for (int c = 0; c < nch; ++c) {
        const bool skip = c < 64 ? (skip_lo >> c) & 1 : (skip_hi >> (c - 64)) & 1;
        if (skip) continue;
        const bool herm = c < 64 ? (herm_lo >> c) & 1 : (herm_hi >> (c - 64)) & 1;
        cplx<T>* Xc = X + (k * nch + c) * nw2;
        if (stage_x)
            for (int t = tid; t < nw2; t += ROT_THREADS) sX[(t / nw) * ld + t % nw] = Xc[t];
        __syncthreads();
        const cplx<T>* xs = stage_x ? sX : Xc;
        const int xld = stage_x ? ld : nw;
        if (nw <= 32)
            rot_channel<T, 2>(nw, ld, sU, xs, xld, sT, herm, Xc);
        else
            rot_channel<T, 4>(nw, ld, sU, xs, xld, sT, herm, Xc);
        __syncthreads();
    }
}

extern "C" int wbmi_rotate(wbmi_ctx* ctx, int nk, int nw, int nch, const void* U_dev, void* X_dev,
                           const int32_t* herm_host, const int32_t* skip_host, int prec) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > 64 || nch <= 0 || nch > 128 || !U_dev || !X_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad arguments (nw <= 64, nch <= 128)");
    uint64_t h[2] = {0, 0}, s[2] = {0, 0};
    for (int c = 0; c < nch; ++c) {
        if (herm_host && herm_host[c]) h[c / 64] |= 1ull << (c % 64);
        if (skip_host && skip_host[c]) s[c / 64] |= 1ull << (c % 64);
    }
    if (prec == WBMI_F64) {
        size_t tile = sizeof(cplx<double>) * nw * (nw + 1);
        int stage_x = 3 * tile <= 160 * 1024;
        size_t lds = (stage_x ? 3 : 2) * tile;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate<double>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate<double>, dim3(nk), dim3(ROT_THREADS), lds, ctx->stream, nw, nch,
                           (const cplx<double>*)U_dev, (cplx<double>*)X_dev, h[0], h[1], s[0], s[1], stage_x);
    } else if (prec == WBMI_F32) {
        size_t tile = sizeof(cplx<float>) * nw * (nw + 1);
        int stage_x = 3 * tile <= 160 * 1024;
        size_t lds = (stage_x ? 3 : 2) * tile;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate<float>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate<float>, dim3(nk), dim3(ROT_THREADS), lds, ctx->stream, nw, nch,
                           (const cplx<float>*)U_dev, (cplx<float>*)X_dev, h[0], h[1], s[0], s[1], stage_x);
    } else {
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad precision");
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
