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

__host__ __device__ inline bool bit_of(uint64_t lo, uint64_t hi, int c) { return c < 64 ? (lo >> c) & 1 : (hi >> (c - 64)) & 1; }

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
    const int nwave = (int)(blockDim.x >> 6);
    for (int t = tid; t < nw2; t += (int)blockDim.x) {
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
    for (int base = 0; base < nrot; base += nwave) {
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

// Split-half rotation (complex64 storage, nw <= 32): Xbar = U^+ X U with both products
// on the fp16 matrix cores (v_mfma_f32_16x16x16_f16).  Each real operand is scaled by a
// power of two and split into two fp16 numbers, x = hi + lo (hi = fp16(x), lo = fp16(x - hi):
// 22 significant bits), and a product is Ah Bh + Ah Bl + Al Bh accumulated in FP32; the
// dropped Al Bl is 2^-22 of the terms.  So a product carries FP32-accumulation rounding,
// eps32 sqrt(nw) |X| |U| per element, where the FP64-arithmetic path carries only the
// rounding of the stored result.  Used for every channel but the velocity dH/dk (whose
// elements between close bands enter the Berry curvature through 1/(E_n - E_m)^2).
//
// One wave per (k, channel), no LDS and no barriers: X is loaded from HBM straight into
// the A-operand registers, U (split once per k) serves as the B operand of T = X U and,
// conjugated, as the A operand of U^+ T, and the accumulators of T are the B operand of
// the second product.  That last step relies on the output layout of the 16 x 16 MFMA:
// lane l, register v holds D[RH_ROW(l / 16, v)][l % 16], so the k index of every operand
// is enumerated in that same order (any order common to A and B gives the same sum).
typedef _Float16 wbmi_h4 __attribute__((ext_vector_type(4)));
typedef float wbmi_f4 __attribute__((ext_vector_type(4)));
#define ROTH_WAVES 4
#define RH_ROW(g, v) (4 * (g) + (v))

struct SplitH {
    wbmi_h4 hi, lo;
};

__device__ inline SplitH split_h(const float* x, float s) {
    SplitH r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float v = x[j] * s;
        const _Float16 h = (_Float16)v;
        r.hi[j] = h;
        r.lo[j] = (_Float16)(v - (float)h);
    }
    return r;
}

__device__ inline wbmi_h4 neg_h(wbmi_h4 a) { return -a; }

// acc += (Ah + Al)(Bh + Bl) without Al Bl
__device__ inline wbmi_f4 mfma3(const SplitH& a, const SplitH& b, wbmi_f4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, acc, 0, 0, 0);
    return acc;
}

__device__ inline SplitH neg_split(const SplitH& a) { return SplitH{neg_h(a.hi), neg_h(a.lo)}; }

// (4 waves per SIMD: the accumulators in VGPRs rather than AGPRs, 128 registers in all)
__global__ void __launch_bounds__(64 * ROTH_WAVES, 4) k_rotate_h(int nw, int nch, const cplx<float>* __restrict__ U,
                                                             cplx<float>* X, uint64_t herm_lo, uint64_t herm_hi,
                                                             uint64_t skip_lo, uint64_t skip_hi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t k = blockIdx.x;
    const int nw2 = nw * nw;
    const cplx<float>* Uk = U + k * nw2;
    // U as the B operand of X U: block (ks, bj) holds U[16 ks + RH_ROW(g, j)][16 bj + li]; the same
    // registers are the A operand of U^+ T for output row block bj (conjugated)
    const float su = 16384.f;
    SplitH ur[2][2], ui[2][2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            float re[4], im[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = 16 * ks + RH_ROW(g, j), c = 16 * bj + li;
                const cplx<float> u = (m < nw && c < nw) ? Uk[m * nw + c] : mk(0.f, 0.f);
                re[j] = u.re;
                im[j] = u.im;
            }
            ur[ks][bj] = split_h(re, su);
            ui[ks][bj] = split_h(im, su);
        }
    // the waves take the channels to rotate in turn (rotation index q = wave, wave + ROTH_WAVES, ...)
    int q = -1;
    for (int ch = 0; ch < nch; ++ch) {
        if (bit_of(skip_lo, skip_hi, ch)) continue;
        if (++q % ROTH_WAVES != wave) continue;
        cplx<float>* Xc = X + (k * nch + ch) * nw2;
        const bool herm = bit_of(herm_lo, herm_hi, ch);
        // X as the A operand: block (bi, ks) holds X[16 bi + li][16 ks + RH_ROW(g, j)], four
        // consecutive elements of a row (at nw = 32 two 16-byte loads: the row offset is a
        // multiple of 4 elements and a channel starts 8 KB-aligned)
        float xre[2][2][4], xim[2][2][4];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const int r = 16 * bi + li, m0 = 16 * ks + RH_ROW(g, 0);
                if (nw == 32) {
                    const float4* p = (const float4*)(Xc + r * 32 + m0);
                    const float4 a = p[0], b = p[1];
                    xre[bi][ks][0] = a.x; xim[bi][ks][0] = a.y; xre[bi][ks][1] = a.z; xim[bi][ks][1] = a.w;
                    xre[bi][ks][2] = b.x; xim[bi][ks][2] = b.y; xre[bi][ks][3] = b.z; xim[bi][ks][3] = b.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const cplx<float> x = (r < nw && m0 + j < nw) ? Xc[r * nw + m0 + j] : mk(0.f, 0.f);
                        xre[bi][ks][j] = x.re;
                        xim[bi][ks][j] = x.im;
                    }
                }
                if (herm) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = m0 + j;
                        const cplx<float> w = (r < nw && m < nw) ? Xc[m * nw + r] : mk(0.f, 0.f);
                        xre[bi][ks][j] = 0.5f * (xre[bi][ks][j] + w.re);
                        xim[bi][ks][j] = 0.5f * (xim[bi][ks][j] - w.im);
                    }
                }
            }
        float mx = 0.f;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int j = 0; j < 4; ++j) mx = fmaxf(mx, fmaxf(fabsf(xre[bi][ks][j]), fabsf(xim[bi][ks][j])));
        for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        // scale: max |x| s < 2^14 (fp16 holds up to 65504)
        int e = 0;
        frexpf(mx, &e);
        const float sx = mx > 0.f ? ldexpf(1.f, 14 - e) : 1.f;
        SplitH xr[2][2], xi[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                xr[bi][ks] = split_h(xre[bi][ks], sx);
                xi[bi][ks] = split_h(xim[bi][ks], sx);
            }
        // T = X U:  Tr = Xr Ur - Xi Ui,  Ti = Xr Ui + Xi Ur   (scaled by sx su)
        wbmi_f4 tr[2][2], ti[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                wbmi_f4 ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    ar = mfma3(xr[bi][ks], ur[ks][bj], ar);
                    ar = mfma3(neg_split(xi[bi][ks]), ui[ks][bj], ar);
                    ai = mfma3(xr[bi][ks], ui[ks][bj], ai);
                    ai = mfma3(xi[bi][ks], ur[ks][bj], ai);
                }
                tr[bi][bj] = ar;
                ti[bi][bj] = ai;
            }
        // T as the B operand of U^+ T: block (ks = bi, bj), element j = v; scale st = sx / 8
        // (|T| <= sqrt(nw) max|X| < 8 max|X|), i.e. the accumulators times 1 / (8 su)
        SplitH bre[2][2], bim[2][2];
        const float tsc = 1.f / (8.f * su);
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                float vr[4], vi[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    vr[v] = tr[bi][bj][v];
                    vi[v] = ti[bi][bj][v];
                }
                bre[bi][bj] = split_h(vr, tsc);
                bim[bi][bj] = split_h(vi, tsc);
            }
        // Xbar = U^+ T: rows bo; A = conj(U)^T: Ar = Ur^T, Ai = -Ui^T
        //   Xr = Ur^T Tr + Ui^T Ti,  Xi = Ur^T Ti - Ui^T Tr   (scaled by su sx / 8)
        const float osc = 8.f / (su * sx);
#pragma unroll
        for (int bo = 0; bo < 2; ++bo)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                wbmi_f4 ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    ar = mfma3(ur[ks][bo], bre[ks][bj], ar);
                    ar = mfma3(ui[ks][bo], bim[ks][bj], ar);
                    ai = mfma3(ur[ks][bo], bim[ks][bj], ai);
                    ai = mfma3(neg_split(ui[ks][bo]), bre[ks][bj], ai);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * bo + RH_ROW(g, v), cc = 16 * bj + li;
                    if (r < nw && cc < nw) Xc[r * nw + cc] = mk(ar[v] * osc, ai[v] * osc);
                }
            }
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
        // one wave (and one LDS tile) per channel to rotate, at most ROTW_WAVES: a call with few
        // channels (the mixed form's velocity) runs small workgroups, several per CU
        int nrot = 0;
        for (int c = 0; c < nch; ++c) nrot += !bit_of(s[0], s[1], c);
        const int waves = nrot < ROTW_WAVES ? (nrot > 0 ? nrot : 1) : ROTW_WAVES;
        const size_t lds = (1 + waves) * rotw_tile_bytes(nw);
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_rotate_w<T>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_rotate_w<T>, dim3(nk), dim3(64 * waves), lds, ctx->stream, nw, nch,
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
    if (prec == WBMI_F32_SPLIT) {
        if (nw > 32)  // above one 32 x 32 block per operand: the FP64-arithmetic path
            return launch_rotate<float>(ctx, nk, nw, nch, U_dev, X_dev, h, s, work_dev, work_bytes);
        hipLaunchKernelGGL(k_rotate_h, dim3(nk), dim3(64 * ROTH_WAVES), 0, ctx->stream, nw, nch,
                           (const cplx<float>*)U_dev, (cplx<float>*)X_dev, h[0], h[1], s[0], s[1]);
        WBMI_LAUNCH_CHECK(ctx);
        return WBMI_OK;
    }
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_rotate: bad precision");
}
