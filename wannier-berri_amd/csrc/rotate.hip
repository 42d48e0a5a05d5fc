// Gauge rotation Xbar = U^+ X U of libwbmi (data_K/data_K.py:130-132).
//
// One 256-thread workgroup per k-point; U and the intermediate T = X U stay
// in LDS while the workgroup walks over the channels of the per-k stack.
// Each thread owns a strip of one row (JT threads per row) so that every X
// element it loads from HBM feeds nw/JT multiply-adds.
#include "wbmi_common.h"

#define ROT_THREADS 256
#define ROT_JT 8

template <typename T>
__global__ void __launch_bounds__(ROT_THREADS) k_rotate(int nw, int nch, const cplx<T>* __restrict__ U,
                                                        cplx<T>* __restrict__ X, uint64_t herm_lo, uint64_t herm_hi,
                                                        uint64_t skip_lo, uint64_t skip_hi) {
    extern __shared__ char smem[];
    cplx<T>* sU = (cplx<T>*)smem;
    cplx<T>* sT = sU + nw * nw;
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    const int nw2 = nw * nw;
    const cplx<T>* Uk = U + k * nw2;
    for (int t = tid; t < nw2; t += ROT_THREADS) sU[t] = Uk[t];
    __syncthreads();
    const int jt = tid % ROT_JT;
    const int i0 = tid / ROT_JT;
    const int istep = ROT_THREADS / ROT_JT;
    for (int c = 0; c < nch; ++c) {
        bool skip = c < 64 ? (skip_lo >> c) & 1 : (skip_hi >> (c - 64)) & 1;
        if (skip) continue;
        bool herm = c < 64 ? (herm_lo >> c) & 1 : (herm_hi >> (c - 64)) & 1;
        cplx<T>* Xc = X + (k * nch + c) * nw2;
        // T = X U
        for (int i = i0; i < nw; i += istep) {
            cplx<T> acc[64 / ROT_JT];
#pragma unroll
            for (int q = 0; q < 64 / ROT_JT; ++q) acc[q] = mk((T)0, (T)0);
            for (int m = 0; m < nw; ++m) {
                cplx<T> x = Xc[i * nw + m];
                if (herm) {
                    cplx<T> y = Xc[m * nw + i];
                    x = mk((T)0.5 * (x.re + y.re), (T)0.5 * (x.im - y.im));
                }
#pragma unroll
                for (int q = 0; q < 64 / ROT_JT; ++q) {
                    int j = jt + q * ROT_JT;
                    if (j < nw) acc[q] += x * sU[m * nw + j];
                }
            }
#pragma unroll
            for (int q = 0; q < 64 / ROT_JT; ++q) {
                int j = jt + q * ROT_JT;
                if (j < nw) sT[i * nw + j] = acc[q];
            }
        }
        __syncthreads();
        // Xbar = U^+ T
        for (int i = i0; i < nw; i += istep) {
            for (int j = jt; j < nw; j += ROT_JT) {
                cplx<T> acc = mk((T)0, (T)0);
                for (int m = 0; m < nw; ++m) acc += cmul(sU[m * nw + i], sT[m * nw + j]);
                Xc[i * nw + j] = acc;
            }
        }
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
        size_t lds = 2 * sizeof(cplx<double>) * nw * nw;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate<double>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate<double>, dim3(nk), dim3(ROT_THREADS), lds, ctx->stream, nw, nch,
                           (const cplx<double>*)U_dev, (cplx<double>*)X_dev, h[0], h[1], s[0], s[1]);
    } else if (prec == WBMI_F32) {
        size_t lds = 2 * sizeof(cplx<float>) * nw * nw;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate<float>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate<float>, dim3(nk), dim3(ROT_THREADS), lds, ctx->stream, nw, nch,
                           (const cplx<float>*)U_dev, (cplx<float>*)X_dev, h[0], h[1], s[0], s[1]);
    } else {
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad precision");
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
