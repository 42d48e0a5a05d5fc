// Gauge rotation Xbar = U^+ X U of libwbmi (data_K/data_K.py:130-132).
//
// One 256-thread workgroup per k-point.  U stays in LDS for all channels of
// the k-point; each channel is loaded once from HBM with coalesced reads into
// LDS (Hermitised there when flagged), T = X U is formed in LDS and
// Xbar = U^+ T is written straight back from registers.  Both products use
// TS x TS register tiles on a 16 x 16 thread grid (TS = 2 for nw <= 32,
// 4 for nw <= 64).  LDS rows are padded by one element.
//
// Complex64 storage: channels are rotated with FP64 arithmetic on the
// complex64 operands (only the stored result is rounded).  FP32 accumulation
// would leave a normwise error ~eps32 |X| |U| in every element, which the
// 1/(E_n - E_m)^2 of the Berry curvature amplifies for small elements between
// close bands; without packed math FP32 and FP64 FMAs issue at the same rate.
#include "mfma_c64.h"
#include "wbmi_common.h"

#define ROT_THREADS 256

// T: storage type of U and X; TA: arithmetic type (and type of the LDS buffer sT)
template <typename T, typename TA, int TS>
__device__ inline void rot_channel(int nw, int ld, const cplx<T>* sU, const cplx<T>* sX, int xld, cplx<TA>* sT,
                                   bool herm, cplx<T>* Xc) {
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    const int r0 = TS * tr, c0 = TS * tc;
    const bool active = r0 < nw && c0 < nw;
    cplx<TA> acc[TS][TS];
    // T = X U
    if (active) {
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j) acc[i][j] = mk((TA)0, (TA)0);
        for (int m = 0; m < nw; ++m) {
            cplx<TA> x[TS], u[TS];
#pragma unroll
            for (int i = 0; i < TS; ++i) {
                const int r = r0 + i < nw ? r0 + i : nw - 1;
                const cplx<T> v = sX[r * xld + m];
                x[i] = mk((TA)v.re, (TA)v.im);
                if (herm) {
                    const cplx<T> w = sX[m * xld + r];
                    x[i] = mk((TA)0.5 * (x[i].re + (TA)w.re), (TA)0.5 * (x[i].im - (TA)w.im));
                }
            }
#pragma unroll
            for (int j = 0; j < TS; ++j) {
                const cplx<T> v = sU[m * ld + (c0 + j < nw ? c0 + j : nw - 1)];
                u[j] = mk((TA)v.re, (TA)v.im);
            }
#pragma unroll
            for (int i = 0; i < TS; ++i)
#pragma unroll
                for (int j = 0; j < TS; ++j) cfma(acc[i][j], x[i], u[j]);
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
            for (int j = 0; j < TS; ++j) acc[i][j] = mk((TA)0, (TA)0);
        for (int m = 0; m < nw; ++m) {
            cplx<TA> u[TS], t[TS];
#pragma unroll
            for (int i = 0; i < TS; ++i) {
                const cplx<T> v = sU[m * ld + (r0 + i < nw ? r0 + i : nw - 1)];
                u[i] = mk((TA)v.re, (TA)v.im);
            }
#pragma unroll
            for (int j = 0; j < TS; ++j) t[j] = sT[m * ld + (c0 + j < nw ? c0 + j : nw - 1)];
#pragma unroll
            for (int i = 0; i < TS; ++i)
#pragma unroll
                for (int j = 0; j < TS; ++j) cfma_cj(acc[i][j], u[i], t[j]);
        }
#pragma unroll
        for (int i = 0; i < TS; ++i)
#pragma unroll
            for (int j = 0; j < TS; ++j)
                if (r0 + i < nw && c0 + j < nw) Xc[(r0 + i) * nw + c0 + j] = mk((T)acc[i][j].re, (T)acc[i][j].im);
    }
}

// stage_x = 0: X is read from HBM (the LDS tiles would not fit).
// sT holds cplx<double> for T = float so that flagged channels keep T in FP64.
template <typename T>
__global__ void __launch_bounds__(ROT_THREADS) k_rotate(int nw, int nch, const cplx<T>* __restrict__ U,
                                                        cplx<T>* X, uint64_t herm_lo, uint64_t herm_hi,
                                                        uint64_t skip_lo, uint64_t skip_hi, int stage_x) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int ld = nw + 1;
    cplx<T>* sU = (cplx<T>*)smem;
    cplx<T>* sX = sU + nw * ld;
    char* sT = (char*)(sX + (stage_x ? nw * ld : 0));
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
        if (sizeof(T) == 4) {
            if (nw <= 32)
                rot_channel<T, double, 2>(nw, ld, sU, xs, xld, (cplx<double>*)sT, herm, Xc);
            else
                rot_channel<T, double, 4>(nw, ld, sU, xs, xld, (cplx<double>*)sT, herm, Xc);
        } else {
            if (nw <= 32)
                rot_channel<T, T, 2>(nw, ld, sU, xs, xld, (cplx<T>*)sT, herm, Xc);
            else
                rot_channel<T, T, 4>(nw, ld, sU, xs, xld, (cplx<T>*)sT, herm, Xc);
        }
        __syncthreads();
    }
}

// nw <= 32: each wave rotates its own channels (ROTW_WAVES per round, one per
// wave) on the matrix cores (FP64 MFMA, mfma_c64.h).  U and the
// tiles are held in FP64 whatever the storage type (without packed math FP32
// and FP64 FMAs issue at the same rate, so complex64 channels are rotated
// with FP64 arithmetic at no cost).  A wave owns one LDS tile: X is staged
// there, T = X U is formed in the MFMA accumulators (Hermitising X on the fly
// when flagged), overwrites the tile after a barrier, and Xbar = U^+ T is
// stored from the accumulators.  One tile per wave lets eight waves share the CU's LDS.
// Barriers stay in uniform control flow.
#define ROTW_WAVES 8

__device__ inline bool bit_of(uint64_t lo, uint64_t hi, int c) { return c < 64 ? (lo >> c) & 1 : (hi >> (c - 64)) & 1; }

__host__ __device__ inline size_t rotw_tile_bytes(int nw) { return sizeof(cplx<double>) * nw * (nw + 1); }

template <typename T>
__global__ void __launch_bounds__(64 * ROTW_WAVES) k_rotate_w(int nw, int nch, const cplx<T>* __restrict__ U,
                                                             cplx<T>* X, uint64_t herm_lo, uint64_t herm_hi,
                                                             uint64_t skip_lo, uint64_t skip_hi) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int ld = nw + 1, nw2 = nw * nw;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    cplx<double>* sU = (cplx<double>*)smem;
    cplx<double>* tile = sU + (1 + wave) * nw * ld;
    const int64_t k = blockIdx.x;
    const cplx<T>* Uk = U + k * nw2;
    for (int t = tid; t < nw2; t += 64 * ROTW_WAVES) {
        const cplx<T> u = Uk[t];
        sU[(t / nw) * ld + t % nw] = mk((double)u.re, (double)u.im);
    }
    int nrot = 0;
    for (int c = 0; c < nch; ++c) nrot += !bit_of(skip_lo, skip_hi, c);
    // channel of rotation index `want` (scan of the skip mask)
    auto channel = [&](int want) {
        int seen = 0;
        for (int c = 0; c < nch; ++c) {
            if (bit_of(skip_lo, skip_hi, c)) continue;
            if (seen == want) return c;
            ++seen;
        }
        return -1;
    };
    __syncthreads();  // sU staged
    // From here on every wave works on its own tile: the phases of a channel are
    // ordered by wave-level fences only, so the waves drift apart and one wave's
    // loads and stores overlap the other waves' products on the matrix cores.
    for (int base = 0; base < nrot; base += ROTW_WAVES) {
        const int c = base + wave < nrot ? channel(base + wave) : -1;
        if (c < 0) break;
        cplx<T>* Xc = X + (k * nch + c) * nw2;
        const bool herm = bit_of(herm_lo, herm_hi, c);
        {
            // all of the lane's loads in flight before the first LDS store (nw2 <= 1024 = 16 x 64)
            cplx<T> buf[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int t = lane + 64 * u;
                if (t < nw2) buf[u] = Xc[t];
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int t = lane + 64 * u;
                if (t < nw2) tile[(t / nw) * ld + t % nw] = mk((double)buf[u].re, (double)buf[u].im);
            }
        }
        wave_lds_sync();
        MfmaBlock acc;
        acc.zero();
        {
            // T = X U, X = 0.5 (X + X^+) when flagged (matrix cores, FP64)
            if (herm)
                mfma_block(acc, nw, nw, nw, 0, 0,
                           [&](int r, int m) {
                               const cplx<double> x = tile[r * ld + m], w = tile[m * ld + r];
                               return mk(0.5 * (x.re + w.re), 0.5 * (x.im - w.im));
                           },
                           [&](int m, int cc) { return sU[m * ld + cc]; });
            else
                mfma_block(acc, nw, nw, nw, 0, 0, [&](int r, int m) { return tile[r * ld + m]; },
                           [&](int m, int cc) { return sU[m * ld + cc]; });
        }
        wave_lds_sync();  // every read of X is done before T overwrites the tile
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mfma_row(bi, v), cc = mfma_col(bj);
                    if (r < nw && cc < nw) tile[r * ld + cc] = acc.get(bi, bj, v);
                }
        wave_lds_sync();
        {
            // Xbar = U^+ T
            acc.zero();
            mfma_block(acc, nw, nw, nw, 0, 0, [&](int r, int m) { return conj(sU[m * ld + r]); },
                       [&](int m, int cc) { return tile[m * ld + cc]; });
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = mfma_row(bi, v), cc = mfma_col(bj);
                        if (r < nw && cc < nw) {
                            const cplx<double> x = acc.get(bi, bj, v);
                            Xc[r * nw + cc] = mk((T)x.re, (T)x.im);
                        }
                    }
        }
        wave_lds_sync();  // every read of T is done before the next channel overwrites the tile
    }
}

// 64 < nw <= 128: one 512-thread workgroup per k, channels in turn.  Both
// products are 32 x 32 wave tasks on the FP64 matrix cores with operands read
// from global memory (L2): T = X U into a per-k FP64 workspace, then
// Xbar = U^+ T over the channel in place.
#define ROTG_THREADS 512
template <typename T>
__global__ void __launch_bounds__(ROTG_THREADS) k_rotate_g(int nw, int nch, const cplx<T>* __restrict__ U, cplx<T>* X,
                                                           uint64_t herm_lo, uint64_t herm_hi, uint64_t skip_lo,
                                                           uint64_t skip_hi, cplx<double>* __restrict__ work) {
    const int n = nw, n2 = nw * nw, wave = threadIdx.x >> 6, nwave = ROTG_THREADS / 64;
    const int nb = (n + 31) / 32;
    const int64_t k = blockIdx.x;
    const cplx<T>* Uk = U + k * n2;
    cplx<double>* Tk = work + k * (int64_t)n2;
    auto u = [&](int r, int c) {
        const cplx<T> v = Uk[r * n + c];
        return mk((double)v.re, (double)v.im);
    };
    for (int c = 0; c < nch; ++c) {
        if (bit_of(skip_lo, skip_hi, c)) continue;
        const bool herm = bit_of(herm_lo, herm_hi, c);
        cplx<T>* Xc = X + (k * nch + c) * n2;
        auto x = [&](int r, int m) {
            const cplx<T> v = Xc[r * n + m];
            cplx<double> a = mk((double)v.re, (double)v.im);
            if (herm) {
                const cplx<T> w = Xc[m * n + r];
                a = mk(0.5 * (a.re + (double)w.re), 0.5 * (a.im - (double)w.im));
            }
            return a;
        };
        for (int t = wave; t < nb * nb; t += nwave) {
            const int rb = t / nb, cb = t % nb;
            MfmaBlock C;
            C.zero();
            mfma_block(C, n, n, n, rb * 32, cb * 32, x, u);
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = rb * 32 + mfma_row(bi, v), cc = cb * 32 + mfma_col(bj);
                        if (r < n && cc < n) Tk[r * n + cc] = C.get(bi, bj, v);
                    }
        }
        __syncthreads();  // T complete; every read of X done
        for (int t = wave; t < nb * nb; t += nwave) {
            const int rb = t / nb, cb = t % nb;
            MfmaBlock C;
            C.zero();
            mfma_block(C, n, n, n, rb * 32, cb * 32, [&](int r, int m) { return conj(u(m, r)); },
                       [&](int m, int cc) { return Tk[m * n + cc]; });
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = rb * 32 + mfma_row(bi, v), cc = cb * 32 + mfma_col(bj);
                        if (r < n && cc < n) {
                            const cplx<double> val = C.get(bi, bj, v);
                            Xc[r * n + cc] = mk((T)val.re, (T)val.im);
                        }
                    }
        }
        __syncthreads();  // the workspace is reused by the next channel
    }
}

static size_t rotate_work_bytes(int nk, int nw) { return nw > 64 ? sizeof(cplx<double>) * (size_t)nw * nw * nk : 0; }

extern "C" size_t wbmi_rotate_work_bytes(int nk, int nw, int prec) {
    (void)prec;
    return nk > 0 && nw > 0 ? rotate_work_bytes(nk, nw) : 0;
}

template <typename T>
static int launch_rotate(wbmi_ctx* ctx, int nk, int nw, int nch, const void* U_dev, void* X_dev, const uint64_t* h,
                         const uint64_t* s, void* work, size_t work_bytes) {
    if (nw > 64) {
        if (!work || work_bytes < rotate_work_bytes(nk, nw))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: nw = %d needs %zu bytes of workspace", nw,
                             rotate_work_bytes(nk, nw));
        hipLaunchKernelGGL(k_rotate_g<T>, dim3(nk), dim3(ROTG_THREADS), 0, ctx->stream, nw, nch, (const cplx<T>*)U_dev,
                           (cplx<T>*)X_dev, h[0], h[1], s[0], s[1], (cplx<double>*)work);
        WBMI_LAUNCH_CHECK(ctx);
        return WBMI_OK;
    }
    if (nw <= 32) {
        const size_t lds = (1 + ROTW_WAVES) * rotw_tile_bytes(nw);
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate_w<T>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate_w<T>, dim3(nk), dim3(64 * ROTW_WAVES), lds, ctx->stream, nw, nch,
                           (const cplx<T>*)U_dev, (cplx<T>*)X_dev, h[0], h[1], s[0], s[1]);
        WBMI_LAUNCH_CHECK(ctx);
        return WBMI_OK;
    }
    const size_t tile = sizeof(cplx<T>) * nw * (nw + 1);
    const size_t ttile = sizeof(cplx<double>) * nw * (nw + 1);  // T = X U (FP64 for flagged channels)
    const int stage_x = 2 * tile + ttile <= 160 * 1024;
    const size_t lds = (stage_x ? 2 : 1) * tile + ttile;
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)lds));
    hipLaunchKernelGGL(k_rotate<T>, dim3(nk), dim3(ROT_THREADS), lds, ctx->stream, nw, nch, (const cplx<T>*)U_dev,
                       (cplx<T>*)X_dev, h[0], h[1], s[0], s[1], stage_x);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" int wbmi_rotate(wbmi_ctx* ctx, int nk, int nw, int nch, const void* U_dev, void* X_dev,
                           const int32_t* herm_host, const int32_t* skip_host, int prec, void* work_dev,
                           size_t work_bytes) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > 128 || nch <= 0 || nch > 128 || !U_dev || !X_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad arguments (nw <= 128, nch <= 128)");
    uint64_t h[2] = {0, 0}, s[2] = {0, 0};
    for (int c = 0; c < nch; ++c) {
        if (herm_host && herm_host[c]) h[c / 64] |= 1ull << (c % 64);
        if (skip_host && skip_host[c]) s[c / 64] |= 1ull << (c % 64);
    }
    if (prec == WBMI_F64) return launch_rotate<double>(ctx, nk, nw, nch, U_dev, X_dev, h, s, work_dev, work_bytes);
    if (prec == WBMI_F32) return launch_rotate<float>(ctx, nk, nw, nch, U_dev, X_dev, h, s, work_dev, work_bytes);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad precision");
}
