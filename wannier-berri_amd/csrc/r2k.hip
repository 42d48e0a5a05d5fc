// R -> k transforms of libwbmi.
//
//  * wbmi_build_channels: R-space channel stack (derivatives, curl) once per system.
//  * wbmi_r2k_*: the K-block transform on an NKFFT grid as three separable DFT
//    passes over the R bounding box.  Each pass is a small complex GEMM on the
//    FP64 matrix cores that streams the [A][L][B] input once and writes
//    [A][N][B]; lanes run along the contiguous B index so loads and stores
//    are coalesced.
//  * wbmi_r2k_klist: phase-matrix transform for arbitrary k-points (FP64 MFMA GEMM).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "mfma_c64.h"
#include "wbmi_common.h"

// ---------------------------------------------------------------------------
// channel stack
// ---------------------------------------------------------------------------

#define WBMI_MAX_CH 128

struct ChanSpec {
    int8_t s[WBMI_MAX_CH * 8];  // per channel: 2 terms x (sign, a, d1, d2)
};

__global__ void k_build_channels(int nR, int nw, int nch, int nbase, const double* __restrict__ cR,
                                 const cplx<double>* __restrict__ base, ChanSpec spec,
                                 cplx<double>* __restrict__ out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nw2 = (int64_t)nw * nw;
    int64_t total = (int64_t)nR * nch * nw2;
    if (idx >= total) return;
    int64_t mn = idx % nw2;
    int ch = (int)((idx / nw2) % nch);
    int64_t R = idx / (nw2 * nch);
    int64_t e = R * nw2 + mn;  // element (R, m, n)
    const double* r = cR + e * 3;
    cplx<double> acc = mk(0.0, 0.0);
    for (int t = 0; t < 2; ++t) {
        const int8_t* q = spec.s + ch * 8 + t * 4;
        int sg = q[0];
        if (sg == 0) continue;
        cplx<double> x = base[e * nbase + q[1]];
        for (int k = 2; k < 4; ++k) {
            int d = q[k];
            if (d >= 0) x = r[d] * times_i(x);
        }
        if (sg > 0) acc += x; else acc -= x;
    }
    out[idx] = acc;
}

extern "C" int wbmi_build_channels(wbmi_ctx* ctx, int nR, int nw, const double* cRshift_dev,
                                   const void* base_dev, int nbase, const int32_t* spec_host, int nch,
                                   void* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nR <= 0 || nw <= 0 || nbase <= 0 || nch <= 0 || nch > WBMI_MAX_CH || !cRshift_dev || !base_dev ||
        !spec_host || !out_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_build_channels: bad arguments (nch <= %d)", WBMI_MAX_CH);
    ChanSpec spec;
    for (int c = 0; c < nch; ++c)
        for (int t = 0; t < 2; ++t) {
            const int32_t* q = spec_host + (c * 2 + t) * 4;
            if (q[0] < -1 || q[0] > 1 || (q[0] != 0 && (q[1] < 0 || q[1] >= nbase)) || q[2] < -1 || q[2] > 2 ||
                q[3] < -1 || q[3] > 2)
                return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_build_channels: bad spec for channel %d", c);
            for (int k = 0; k < 4; ++k) spec.s[c * 8 + t * 4 + k] = (int8_t)q[k];
        }
    int64_t total = (int64_t)nR * nch * nw * nw;
    hipLaunchKernelGGL(k_build_channels, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, ctx->stream, nR, nw,
                       nch, nbase, cRshift_dev, (const cplx<double>*)base_dev, spec, (cplx<double>*)out_dev);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// ---------------------------------------------------------------------------
// separable DFT on the NKFFT grid
// ---------------------------------------------------------------------------

struct wbmi_r2k_plan {
    int nR;
    int N[3];
    int L[3];        // box extent per axis
    int off[3];      // R value of box index 0 (non-aliased axes)
    int alias[3];    // 1 when the R extent exceeds NKFFT on this axis
    int any_alias;
    int64_t nelem;
    int64_t nslot;
    int32_t* slot_start;  // device CSR: R indices per box slot
    int32_t* slot_list;
    int32_t* iR;          // device copy of iRvec
    int tab;              // stride of one axis table
};

// box[s][e] = sum_{R in slot s} X[R][e] exp(2 pi i sum_{aliased axes} dK_i R_i)
__global__ void k_box_gather(int64_t nslot, int64_t nelem, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ list, const int32_t* __restrict__ iR,
                             const cplx<double>* __restrict__ X, double dk0, double dk1, double dk2,
                             cplx<double>* __restrict__ box) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s = blockIdx.y + (int64_t)blockIdx.z * gridDim.y;
    if (e >= nelem || s >= nslot) return;
    cplx<double> acc = mk(0.0, 0.0);
    for (int q = start[s]; q < start[s + 1]; ++q) {
        int r = list[q];
        double ph = dk0 * iR[3 * r] + dk1 * iR[3 * r + 1] + dk2 * iR[3 * r + 2];
        ph -= floor(ph);
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        acc += mk(cs, sn) * X[(int64_t)r * nelem + e];
    }
    box[s * nelem + e] = acc;
}

// phase[ax][j*L + b] = exp(2 pi i ((j*(b+off)) mod N / N + dKpass*(b+off)))
__global__ void k_pass_phases(int N0, int N1, int N2, int L0, int L1, int L2, int o0, int o1, int o2, double d0,
                              double d1, double d2, int tab, cplx<double>* __restrict__ phase) {
    int ax = blockIdx.y;
    int N = ax == 0 ? N0 : ax == 1 ? N1 : N2;
    int L = ax == 0 ? L0 : ax == 1 ? L1 : L2;
    int off = ax == 0 ? o0 : ax == 1 ? o1 : o2;
    double dk = ax == 0 ? d0 : ax == 1 ? d1 : d2;
    for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < N * L; t += blockDim.x * gridDim.x) {
        int j = t / L, b = t % L;
        long long R = (long long)b + off;
        long long jr = ((long long)j * R) % N;
        if (jr < 0) jr += N;
        double ph = (double)jr / N;
        double f = dk * (double)R;
        f -= floor(f);
        ph += f;
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        phase[ax * tab + t] = mk(cs, sn);
    }
}

// out[a][j][b] = sum_l P[j][l] in[a][l][b] on the FP64 matrix cores (v_mfma_f64_16x16x4f64:
// A[l%16][l/16], B[l/16][l%16], D[l/16 + 4v][l%16]).  One wave per 16 consecutive b of one a:
// the B operand is in[a][l][b] (16 b per row, 4 rows l per step), every 16-row block of j is an
// accumulator triple (the 3M form, cmac3) fed from the phases in LDS (the A operand), so each input element is
// read once and each output element written once.  Rows j come in chunks of 16 NB (NB = 1, 2, 4
// accumulator triples, chosen from N at launch); K = l in chunks of DM_LCHUNK, whose loads are
// issued before their products.
// Complex multiply-accumulate with three real products on the matrix cores (the 3M form):
// s += Pr Xr, t += Pi Xi, u += (Pr + Pi)(Xr + Xi); at the end re = s - t, im = u - s - t.
__device__ inline void cmac3(double pr, double pi, double xr, double xi, wbmi_v4d& s, wbmi_v4d& t, wbmi_v4d& u) {
    s = __builtin_amdgcn_mfma_f64_16x16x4f64(pr, xr, s, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f64_16x16x4f64(pi, xi, t, 0, 0, 0);
    u = __builtin_amdgcn_mfma_f64_16x16x4f64(pr + pi, xr + xi, u, 0, 0, 0);
}

#define DM_WAVES 4
#define DM_LCHUNK 12

// (minimum waves per SIMD per variant: the accumulators stay in VGPRs and nothing spills)
template <typename Tin, typename Tout, int NB, bool PAIR>
__global__ void __launch_bounds__(64 * DM_WAVES, (PAIR ? (NB == 1 ? 3 : NB == 2 ? 2 : 1) : (NB == 1 ? 8 : NB == 2 ? 5 : 2))) k_dft_mfma(const cplx<Tin>* __restrict__ in, cplx<Tout>* __restrict__ out,
                                                          const cplx<double>* __restrict__ P, int64_t A, int L, int N,
                                                          int64_t B) {
    // PAIR (complex64 in and out, B even): a lane owns the adjacent columns b, b + 1 (16-byte
    // loads and stores) as two MFMA column tiles, so a wave covers 32 columns
    static_assert(!PAIR || (sizeof(Tin) == 4 && sizeof(Tout) == 4), "column pairs are for complex64");
    constexpr int CW = PAIR ? 2 : 1;
    extern __shared__ cplx<double> sP[];  // [N][L]
    for (int t = threadIdx.x; t < N * L; t += blockDim.x) sP[t] = P[t];
    __syncthreads();
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    const int64_t tiles_b = (B + 16 * CW - 1) / (16 * CW);
    const int64_t tile = (int64_t)blockIdx.x * DM_WAVES + (threadIdx.x >> 6);
    if (tile >= A * tiles_b) return;  // whole wave; no barrier follows
    const int64_t a = tile / tiles_b, b = (tile % tiles_b) * 16 * CW + CW * li;
    const bool bok = b < B;  // PAIR: B is even, so b + 1 < B as well
    const cplx<Tin>* src = in + a * L * B + b;
    cplx<Tout>* dst = out + a * N * B + b;
    for (int j0 = 0; j0 < N; j0 += 16 * NB) {
        wbmi_v4d as[CW][NB], at[CW][NB], au[CW][NB];
#pragma unroll
        for (int c = 0; c < CW; ++c)
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                as[c][q] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
                at[c][q] = as[c][q];
                au[c][q] = as[c][q];
            }
        const int nb = min(16 * NB, N - j0);
        for (int l0 = 0; l0 < L; l0 += DM_LCHUNK) {
            // the chunk's inputs in their storage type (converted at their products)
            cplx<Tin> x[DM_LCHUNK / 4][CW];
#pragma unroll
            for (int ks = 0; ks < DM_LCHUNK / 4; ++ks) {
                const int l = l0 + 4 * ks + g;
#pragma unroll
                for (int c = 0; c < CW; ++c) x[ks][c] = mk((Tin)0, (Tin)0);
                if (l < L && bok) {
                    if constexpr (PAIR) {
                        const float4 v = *(const float4*)(src + (int64_t)l * B);
                        x[ks][0] = mk((Tin)v.x, (Tin)v.y);
                        x[ks][CW - 1] = mk((Tin)v.z, (Tin)v.w);
                    } else {
                        x[ks][0] = src[(int64_t)l * B];
                    }
                }
            }
#pragma unroll
            for (int ks = 0; ks < DM_LCHUNK / 4; ++ks) {
                const int l = l0 + 4 * ks + g;
                if (l0 + 4 * ks >= L) break;
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    if (16 * q >= nb) break;
                    const int j = j0 + 16 * q + li;
                    const cplx<double> ph = (j < N && l < L) ? sP[j * L + l] : mk(0.0, 0.0);
#pragma unroll
                    for (int c = 0; c < CW; ++c)
                        cmac3(ph.re, ph.im, (double)x[ks][c].re, (double)x[ks][c].im, as[c][q], at[c][q], au[c][q]);
                }
            }
        }
        if (bok) {
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                if (16 * q >= nb) break;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = j0 + 16 * q + g + 4 * v;
                    if (j >= N) continue;
                    const double r0 = as[0][q][v] - at[0][q][v], i0 = au[0][q][v] - as[0][q][v] - at[0][q][v];
                    if constexpr (PAIR) {
                        const double r1 = as[CW - 1][q][v] - at[CW - 1][q][v];
                        const double i1 = au[CW - 1][q][v] - as[CW - 1][q][v] - at[CW - 1][q][v];
                        *(float4*)(dst + (int64_t)j * B) = make_float4((float)r0, (float)i0, (float)r1, (float)i1);
                    } else {
                        dst[(int64_t)j * B] = mk((Tout)r0, (Tout)i0);
                    }
                }
            }
        }
    }
}

// The y and x passes in one kernel: out[j0][j1][c] = sum_{a0,a1} P0[j0][a0] P1[j1][a1] in[a0][a1][c].
// A workgroup owns 16 CW consecutive columns c (CW = 2: complex64 column pairs, as in k_dft_mfma).
// Its waves first form T[a0][j1][c] (the y pass, one task per a0 and 16 rows j1) into LDS as
// complex128, then the x pass reads T from LDS as the MFMA B operand (one task per j1 and 16 rows
// j0), so the y-pass output never goes to HBM: per column the kernel reads L0 L1 and writes N0 N1
// elements.  LDS: P0 [N0][L0], P1 [N1][L1], T [L0][N1p][16 CW] (N1p = N1 rounded up to 16).
#ifndef DM2_WAVES
#define DM2_WAVES 8
#endif
#ifndef DM2_NT
#define DM2_NT 1
#endif
#ifndef DM2_MIXED_PAIR
#define DM2_MIXED_PAIR 1
#endif
#ifndef DM2_FUSE
#define DM2_FUSE 1
#endif

template <typename Tin, typename Tout, bool PAIR>
__global__ void __launch_bounds__(64 * DM2_WAVES, 2) k_dft2_mfma(const cplx<Tin>* __restrict__ in, cplx<Tout>* __restrict__ out,
                                                            const cplx<double>* __restrict__ P0,
                                                            const cplx<double>* __restrict__ P1, int L0, int L1, int N0,
                                                            int N1, int64_t C) {
    static_assert(!PAIR || (sizeof(Tin) == 4 && sizeof(Tout) == 4), "column pairs are for complex64");
    constexpr int CW = PAIR ? 2 : 1;
    constexpr int COLS = 16 * CW;
    // the slab in the storage type of the form (complex64 in the mixed form, as the pass outputs it
    // replaces), so that with one column per lane it fits beside a scan workgroup on a CU
    using TT = typename std::conditional<sizeof(Tin) == 4, float, double>::type;
    extern __shared__ cplx<double> lds2[];
    const int N1p = (N1 + 15) & ~15;
    cplx<double>* sP0 = lds2;                          // [N0][L0]
    cplx<double>* sP1 = sP0 + N0 * L0;                 // [N1][L1]
    cplx<TT>* sT = (cplx<TT>*)(sP1 + N1 * L1);         // [L0][N1p][COLS]
    for (int t = threadIdx.x; t < N0 * L0; t += blockDim.x) sP0[t] = P0[t];
    for (int t = threadIdx.x; t < N1 * L1; t += blockDim.x) sP1[t] = P1[t];
    __syncthreads();
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * COLS;
    const int64_t b = c0 + CW * li;
    const bool bok = b < C;  // PAIR: C is even, so b + 1 < C as well

    // y pass: task (a0, 16-row block of j1)
    const int nj1 = N1p / 16;
    for (int task = wv; task < L0 * nj1; task += DM2_WAVES) {
        const int a0 = task / nj1, j1b = 16 * (task % nj1);
        const cplx<Tin>* src = in + (int64_t)a0 * L1 * C + b;
        wbmi_v4d as[CW], at[CW], au[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            as[c] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
            at[c] = as[c];
            au[c] = as[c];
        }
        for (int l0 = 0; l0 < L1; l0 += DM_LCHUNK) {
            cplx<Tin> x[DM_LCHUNK / 4][CW];
#pragma unroll
            for (int ks = 0; ks < DM_LCHUNK / 4; ++ks) {
                const int l = l0 + 4 * ks + g;
#pragma unroll
                for (int c = 0; c < CW; ++c) x[ks][c] = mk((Tin)0, (Tin)0);
                if (l < L1 && bok) {
                    if constexpr (PAIR) {
                        const float4 v = *(const float4*)(src + (int64_t)l * C);
                        x[ks][0] = mk((Tin)v.x, (Tin)v.y);
                        x[ks][CW - 1] = mk((Tin)v.z, (Tin)v.w);
                    } else {
                        x[ks][0] = src[(int64_t)l * C];
                    }
                }
            }
#pragma unroll
            for (int ks = 0; ks < DM_LCHUNK / 4; ++ks) {
                if (l0 + 4 * ks >= L1) break;
                const int l = l0 + 4 * ks + g, j = j1b + li;
                const cplx<double> ph = (j < N1 && l < L1) ? sP1[j * L1 + l] : mk(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < CW; ++c)
                    cmac3(ph.re, ph.im, (double)x[ks][c].re, (double)x[ks][c].im, as[c], at[c], au[c]);
            }
        }
        // rows j1b + g + 4v of T (rows up to N1p exist; those >= N1 hold zeros and are not read)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int c = 0; c < CW; ++c)
                sT[((int64_t)a0 * N1p + j1b + g + 4 * v) * COLS + CW * li + c] =
                    mk((TT)(as[c][v] - at[c][v]), (TT)(au[c][v] - as[c][v] - at[c][v]));
    }
    __syncthreads();

    // x pass: task (j1, 16-row block of j0); the B operand T[a0][j1][col] comes from LDS
    const int nj0 = (N0 + 15) / 16;
    for (int task = wv; task < N1 * nj0; task += DM2_WAVES) {
        const int j1 = task / nj0, j0b = 16 * (task % nj0);
        wbmi_v4d as[CW], at[CW], au[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            as[c] = wbmi_v4d{0.0, 0.0, 0.0, 0.0};
            at[c] = as[c];
            au[c] = as[c];
        }
        for (int a = 0; a < L0; a += 4) {
            const int a0 = a + g, j = j0b + li;
            const cplx<double> ph = (j < N0 && a0 < L0) ? sP0[j * L0 + a0] : mk(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const cplx<TT> x = a0 < L0 ? sT[((int64_t)a0 * N1p + j1) * COLS + CW * li + c] : mk((TT)0, (TT)0);
                cmac3(ph.re, ph.im, (double)x.re, (double)x.im, as[c], at[c], au[c]);
            }
        }
        if (bok) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j0 = j0b + g + 4 * v;
                if (j0 >= N0) continue;
                cplx<Tout>* dst = out + ((int64_t)j0 * N1 + j1) * C + b;
                const double r0 = as[0][v] - at[0][v], i0 = au[0][v] - as[0][v] - at[0][v];
                if constexpr (PAIR) {
                    const double r1 = as[CW - 1][v] - at[CW - 1][v], i1 = au[CW - 1][v] - as[CW - 1][v] - at[CW - 1][v];
#if DM2_NT
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    __builtin_nontemporal_store(f4{(float)r0, (float)i0, (float)r1, (float)i1}, (f4*)dst);
#else
                    *(float4*)dst = make_float4((float)r0, (float)i0, (float)r1, (float)i1);
#endif
                } else {
#if DM2_NT
                    typedef Tout t2 __attribute__((ext_vector_type(2)));
                    __builtin_nontemporal_store(t2{(Tout)r0, (Tout)i0}, (t2*)dst);
#else
                    *dst = mk((Tout)r0, (Tout)i0);
#endif
                }
            }
        }
    }
}

// LDS bytes of k_dft2_mfma (tsz: bytes of a slab element, 16 or 8)
static size_t dft2_lds(int L0, int L1, int N0, int N1, int cw, size_t tsz) {
    const size_t N1p = (size_t)((N1 + 15) & ~15);
    return sizeof(cplx<double>) * ((size_t)N0 * L0 + (size_t)N1 * L1) + tsz * (size_t)L0 * N1p * 16 * cw;
}

template <typename Tin, typename Tout, bool PAIR>
static int launch_pass2(wbmi_ctx* ctx, const void* in, void* out, const cplx<double>* P0, const cplx<double>* P1,
                        int L0, int L1, int N0, int N1, int64_t C) {
    constexpr int CW = PAIR ? 2 : 1;
    const size_t lds = dft2_lds(L0, L1, N0, N1, CW, 2 * sizeof(Tin));
    const int64_t blocks = ceil_div(C, 16 * CW);
    if (lds > 160 * 1024 || blocks > 0x7fffffff)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_exec: fused pass of %d x %d -> %d x %d too large", L0, L1, N0, N1);
    if (lds > 64 * 1024) {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_dft2_mfma<Tin, Tout, PAIR>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL((k_dft2_mfma<Tin, Tout, PAIR>), dim3((unsigned)blocks), dim3(64 * DM2_WAVES), lds, ctx->stream,
                       (const cplx<Tin>*)in, (cplx<Tout>*)out, P0, P1, L0, L1, N0, N1, C);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

static void grid_2d(int64_t A, dim3& g) {
    // split A over gridDim.y * gridDim.z (y <= 65535)
    int64_t gy = A < 65535 ? A : 65535;
    int64_t gz = ceil_div(A, gy);
    g.y = (unsigned)gy;
    g.z = (unsigned)gz;
}

template <typename Tin, typename Tout, int NB, bool PAIR>
static int launch_pass_nb(wbmi_ctx* ctx, const void* in, void* out, const cplx<double>* P, int64_t A, int L, int N,
                          int64_t B) {
    constexpr int CW = PAIR ? 2 : 1;  // columns per lane
    const int64_t tiles = A * ceil_div(B, 16 * CW);
    const size_t lds = sizeof(cplx<double>) * (size_t)N * L;
    if (lds > 160 * 1024 || ceil_div(tiles, DM_WAVES) > 0x7fffffff)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_exec: pass of %lld x %d x %d x %lld too large", (long long)A, L, N,
                         (long long)B);
    if (lds > 64 * 1024) {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_dft_mfma<Tin, Tout, NB, PAIR>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL((k_dft_mfma<Tin, Tout, NB, PAIR>), dim3((unsigned)ceil_div(tiles, DM_WAVES)), dim3(64 * DM_WAVES), lds,
                       ctx->stream, (const cplx<Tin>*)in, (cplx<Tout>*)out, P, A, L, N, B);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

template <typename Tin, typename Tout, bool PAIR>
static int launch_pass_p(wbmi_ctx* ctx, const void* in, void* out, const cplx<double>* P, int64_t A, int L, int N,
                         int64_t B) {
    if (N <= 16) return launch_pass_nb<Tin, Tout, 1, PAIR>(ctx, in, out, P, A, L, N, B);
    if (N <= 32) return launch_pass_nb<Tin, Tout, 2, PAIR>(ctx, in, out, P, A, L, N, B);
    return launch_pass_nb<Tin, Tout, 4, PAIR>(ctx, in, out, P, A, L, N, B);
}

template <typename Tin, typename Tout>
static int launch_pass(wbmi_ctx* ctx, const void* in, void* out, const cplx<double>* P, int64_t A, int L, int N,
                       int64_t B) {
    // complex64 -> complex64 with an even column count and 16-byte aligned arrays (so every row
    // is 16-byte aligned): column pairs
    if constexpr (sizeof(Tin) == 4 && sizeof(Tout) == 4)
        if ((B & 1) == 0 && (((uintptr_t)in | (uintptr_t)out) & 15) == 0)
            return launch_pass_p<Tin, Tout, true>(ctx, in, out, P, A, L, N, B);
    return launch_pass_p<Tin, Tout, false>(ctx, in, out, P, A, L, N, B);
}

extern "C" {

int wbmi_r2k_plan_create(wbmi_ctx* ctx, int nR, const int32_t* iRvec_host, const int32_t NKFFT[3], int64_t nelem,
                         wbmi_r2k_plan** out) {
    if (!ctx || !out) return WBMI_ERR_ARG;
    *out = nullptr;
    if (nR <= 0 || !iRvec_host || nelem <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_plan_create: bad arguments");
    wbmi_r2k_plan* p = new wbmi_r2k_plan();
    p->nR = nR;
    p->nelem = nelem;
    p->any_alias = 0;
    for (int i = 0; i < 3; ++i) {
        if (NKFFT[i] <= 0) {
            delete p;
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_plan_create: NKFFT must be positive");
        }
        int lo = iRvec_host[i], hi = iRvec_host[i];
        for (int r = 0; r < nR; ++r) {
            lo = std::min(lo, iRvec_host[3 * r + i]);
            hi = std::max(hi, iRvec_host[3 * r + i]);
        }
        p->N[i] = NKFFT[i];
        int ext = hi - lo + 1;
        p->alias[i] = ext > NKFFT[i];
        p->L[i] = p->alias[i] ? NKFFT[i] : ext;
        p->off[i] = p->alias[i] ? 0 : lo;
        p->any_alias |= p->alias[i];
    }
    p->nslot = (int64_t)p->L[0] * p->L[1] * p->L[2];
    std::vector<std::vector<int32_t>> lists(p->nslot);
    for (int r = 0; r < nR; ++r) {
        int64_t s = 0;
        for (int i = 0; i < 3; ++i) {
            int R = iRvec_host[3 * r + i];
            int b = p->alias[i] ? ((R % p->N[i]) + p->N[i]) % p->N[i] : R - p->off[i];
            s = s * p->L[i] + b;
        }
        lists[s].push_back(r);
    }
    std::vector<int32_t> start(p->nslot + 1, 0), flat;
    flat.reserve(nR);
    for (int64_t s = 0; s < p->nslot; ++s) {
        start[s] = (int32_t)flat.size();
        flat.insert(flat.end(), lists[s].begin(), lists[s].end());
    }
    start[p->nslot] = (int32_t)flat.size();
    int NM = std::max(p->N[0], std::max(p->N[1], p->N[2]));
    int LM = std::max(p->L[0], std::max(p->L[1], p->L[2]));
    p->tab = NM * LM;
    hipError_t e = hipMalloc(&p->slot_start, sizeof(int32_t) * (p->nslot + 1));
    if (e == hipSuccess) e = hipMalloc(&p->slot_list, sizeof(int32_t) * nR);
    if (e == hipSuccess) e = hipMalloc(&p->iR, sizeof(int32_t) * 3 * nR);
    if (e == hipSuccess) e = hipMemcpy(p->slot_start, start.data(), sizeof(int32_t) * (p->nslot + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->slot_list, flat.data(), sizeof(int32_t) * nR, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->iR, iRvec_host, sizeof(int32_t) * 3 * nR, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        wbmi_r2k_plan_destroy(p);
        return wbmi_fail(ctx, WBMI_ERR_HIP, "wbmi_r2k_plan_create: %s", hipGetErrorString(e));
    }
    *out = p;
    return WBMI_OK;
}

int wbmi_r2k_plan_destroy(wbmi_r2k_plan* p) {
    if (!p) return WBMI_OK;
    if (p->slot_start) (void)hipFree(p->slot_start);
    if (p->slot_list) (void)hipFree(p->slot_list);
    if (p->iR) (void)hipFree(p->iR);
    delete p;
    return WBMI_OK;
}

// bytes at the head of the workspace for the pass phase tables [3][tab] (256-byte multiple):
// per call, so executions of one plan on different streams share no state
static size_t r2k_phase_bytes(const wbmi_r2k_plan* p) {
    return (sizeof(cplx<double>) * 3 * (size_t)p->tab + 255) & ~(size_t)255;
}

int wbmi_r2k_plan_sizes(const wbmi_r2k_plan* p, size_t* box_bytes, size_t* work_bytes) {
    if (!p) return WBMI_ERR_ARG;
    if (box_bytes) *box_bytes = sizeof(cplx<double>) * (size_t)p->nslot * p->nelem;
    // the three pass phase tables, then two intermediate pass outputs, sized for complex128
    // (also enough for complex64)
    size_t w1 = (size_t)p->L[0] * p->L[1] * p->N[2];
    size_t w2 = (size_t)p->L[0] * p->N[1] * p->N[2];
    if (work_bytes) *work_bytes = r2k_phase_bytes(p) + sizeof(cplx<double>) * (w1 + w2) * p->nelem;
    return WBMI_OK;
}

int wbmi_r2k_box(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* XR_dev, void* box_dev) {
    if (!ctx || !p || !XR_dev || !box_dev) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_box: null argument");
    dim3 g((unsigned)ceil_div(p->nelem, 256));
    grid_2d(p->nslot, g);
    hipLaunchKernelGGL(k_box_gather, g, dim3(256), 0, ctx->stream, p->nslot, p->nelem, p->slot_start, p->slot_list,
                       p->iR, (const cplx<double>*)XR_dev, 0.0, 0.0, 0.0, (cplx<double>*)box_dev);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// pass phase tables of the shift dK into the head of the workspace
static int r2k_phases(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const double dK[3], void* work_dev) {
    double dpass[3];
    for (int i = 0; i < 3; ++i) dpass[i] = p->alias[i] ? 0.0 : dK[i];
    hipLaunchKernelGGL(k_pass_phases, dim3(4, 3), dim3(256), 0, ctx->stream, p->N[0], p->N[1], p->N[2], p->L[0],
                       p->L[1], p->L[2], p->off[0], p->off[1], p->off[2], dpass[0], dpass[1], dpass[2], p->tab,
                       (cplx<double>*)work_dev);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

static size_t r2k_w1_elems(const wbmi_r2k_plan* p) { return (size_t)p->L[0] * p->L[1] * p->N[2] * p->nelem; }

// z pass: box [L0 L1][L2][E] (complex128) -> w1 [L0 L1][N2][E] in the form's type
static int r2k_z(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* box_dev, const cplx<double>* phase, void* w1,
                 int prec) {
    const int64_t A = (int64_t)p->L[0] * p->L[1];
    if (prec == WBMI_F64) return launch_pass<double, double>(ctx, box_dev, w1, phase + 2 * p->tab, A, p->L[2], p->N[2], p->nelem);
    return launch_pass<double, float>(ctx, box_dev, w1, phase + 2 * p->tab, A, p->L[2], p->N[2], p->nelem);
}

// y and x passes: w1 [L0][L1][N2 E] -> out [N0][N1][N2 E], fused through LDS when the
// intermediate fits, else as two passes through w2 ([L0][N1][N2 E])
static int r2k_yx(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* w1, const cplx<double>* phase, void* w2,
                  void* out_dev, int prec) {
    const int L0 = p->L[0], L1 = p->L[1], N0 = p->N[0], N1 = p->N[1];
    const int64_t C = (int64_t)p->N[2] * p->nelem;
    const bool pair = DM2_MIXED_PAIR && prec != WBMI_F64 && (C & 1) == 0 &&
                      (((uintptr_t)w1 | (uintptr_t)out_dev) & 15) == 0;
    const bool fuse = DM2_FUSE && dft2_lds(L0, L1, N0, N1, pair ? 2 : 1, prec == WBMI_F64 ? 16 : 8) <= 160 * 1024;
    int rc;
    if (prec == WBMI_F64) {
        if (fuse) return launch_pass2<double, double, false>(ctx, w1, out_dev, phase, phase + p->tab, L0, L1, N0, N1, C);
        rc = launch_pass<double, double>(ctx, w1, w2, phase + p->tab, L0, L1, N1, C);
        if (rc) return rc;
        return launch_pass<double, double>(ctx, w2, out_dev, phase, 1, L0, N0, (int64_t)N1 * C);
    }
    if (fuse) {
        if (pair) return launch_pass2<float, float, true>(ctx, w1, out_dev, phase, phase + p->tab, L0, L1, N0, N1, C);
        return launch_pass2<float, float, false>(ctx, w1, out_dev, phase, phase + p->tab, L0, L1, N0, N1, C);
    }
    rc = launch_pass<float, float>(ctx, w1, w2, phase + p->tab, L0, L1, N1, C);
    if (rc) return rc;
    return launch_pass<float, float>(ctx, w2, out_dev, phase, 1, L0, N0, (int64_t)N1 * C);
}

int wbmi_r2k_exec(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* XR_dev, void* box_dev, const double dK[3],
                  void* work_dev, void* out_dev, int prec) {
    if (!ctx || !p || !box_dev || !work_dev || !out_dev || !dK)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_exec: null argument");
    if (prec != WBMI_F64 && prec != WBMI_F32) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_exec: bad precision");
    if (p->any_alias) {
        if (!XR_dev) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_exec: aliased plan needs XR_dev");
        dim3 g((unsigned)ceil_div(p->nelem, 256));
        grid_2d(p->nslot, g);
        hipLaunchKernelGGL(k_box_gather, g, dim3(256), 0, ctx->stream, p->nslot, p->nelem, p->slot_start,
                           p->slot_list, p->iR, (const cplx<double>*)XR_dev, p->alias[0] ? dK[0] : 0.0,
                           p->alias[1] ? dK[1] : 0.0, p->alias[2] ? dK[2] : 0.0, (cplx<double>*)box_dev);
        WBMI_LAUNCH_CHECK(ctx);
    }
    int rc = r2k_phases(ctx, p, dK, work_dev);
    if (rc) return rc;
    const cplx<double>* phase = (const cplx<double>*)work_dev;
    const size_t esz = prec == WBMI_F64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
    char* w1 = (char*)work_dev + r2k_phase_bytes(p);
    char* w2 = w1 + esz * r2k_w1_elems(p);
    rc = r2k_z(ctx, p, box_dev, phase, w1, prec);
    if (rc) return rc;
    return r2k_yx(ctx, p, w1, phase, w2, out_dev, prec);
}

int wbmi_r2k_w1_bytes(const wbmi_r2k_plan* p, int prec, size_t* bytes) {
    if (!p || !bytes || (prec != WBMI_F64 && prec != WBMI_F32)) return WBMI_ERR_ARG;
    *bytes = (prec == WBMI_F64 ? sizeof(cplx<double>) : sizeof(cplx<float>)) * r2k_w1_elems(p);
    return WBMI_OK;
}

int wbmi_r2k_zpass(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* box_dev, const double dK[3], void* work_dev,
                   void* w1_dev, int prec) {
    if (!ctx || !p || !box_dev || !work_dev || !w1_dev || !dK)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_zpass: null argument");
    if (prec != WBMI_F64 && prec != WBMI_F32) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_zpass: bad precision");
    if (p->any_alias)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_zpass: the plan aliases R (its box depends on dK); use wbmi_r2k_exec");
    int rc = r2k_phases(ctx, p, dK, work_dev);
    if (rc) return rc;
    return r2k_z(ctx, p, box_dev, (const cplx<double>*)work_dev, w1_dev, prec);
}

int wbmi_r2k_yx(wbmi_ctx* ctx, const wbmi_r2k_plan* p, const void* w1_dev, const double dK[3], void* work_dev,
                void* out_dev, int prec) {
    if (!ctx || !p || !w1_dev || !work_dev || !out_dev || !dK)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_yx: null argument");
    if (prec != WBMI_F64 && prec != WBMI_F32) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_yx: bad precision");
    if (p->any_alias)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_yx: the plan aliases R (its box depends on dK); use wbmi_r2k_exec");
    int rc = r2k_phases(ctx, p, dK, work_dev);
    if (rc) return rc;
    const size_t esz = prec == WBMI_F64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
    char* w2 = (char*)work_dev + r2k_phase_bytes(p) + esz * r2k_w1_elems(p);
    return r2k_yx(ctx, p, w1_dev, (const cplx<double>*)work_dev, w2, out_dev, prec);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// arbitrary k-list: out[k][e] = sum_R exp(2 pi i k.R) X[R][e]
//
// A complex GEMM (M = nk, N = nelem, K = nR) on the FP64 matrix cores.  A
// workgroup owns 64 k x 128 elements (8 waves of 32 x 32, mfma_c64.h); the
// phase factors of its 64 k-points are formed in LDS, KLM_RC R-vectors at a
// time, and the X operand is read straight from HBM (each 16-element row
// segment is 256 contiguous bytes).  blockIdx.x runs over k-tiles, so the
// workgroups resident on one XCD at a time share their X columns
// (nR x 128 x 16 B) through that XCD's L2.
// ---------------------------------------------------------------------------

#define KLM_RC 32

// Workgroup tiles (waves along k x waves along the elements), chosen per nw by the tuning
// table of the context (wbmi_tune_set, WBMI_KNOB_KLIST_TILE); index 0 is the default.
static const int klm_tiles[WBMI_KLIST_NTILES][2] = {{2, 4}, {4, 2}, {1, 8}, {2, 2}, {1, 4}};

template <typename Tout, int WK, int WE>
__global__ void __launch_bounds__(64 * WK * WE) k_r2k_klist(int nR, const int32_t* __restrict__ iR, int64_t nelem,
                                                           const cplx<double>* __restrict__ X, int nk,
                                                           const double* __restrict__ kl,
                                                           cplx<Tout>* __restrict__ out) {
    constexpr int TK = 32 * WK, TE = 32 * WE, NT = 64 * WK * WE;
    __shared__ cplx<double> ph[TK][KLM_RC + 1];
    const int wave = threadIdx.x >> 6;
    const int k0 = blockIdx.x * TK;
    const int64_t e0 = (int64_t)blockIdx.y * TE;
    const int kw = k0 + 32 * (wave / WE);          // first k of this wave's tile
    const int64_t ew = e0 + 32 * (wave % WE);      // first element of this wave's tile
    const int nrow = nk - kw < 32 ? nk - kw : 32;
    const int ncol = nelem - ew < 32 ? (int)(nelem - ew) : 32;
    MfmaBlock C;
    C.zero();
    for (int r0 = 0; r0 < nR; r0 += KLM_RC) {
        __syncthreads();
        for (int t = threadIdx.x; t < TK * KLM_RC; t += NT) {
            const int q = t / KLM_RC, rr = t % KLM_RC;
            const int k = k0 + q, r = r0 + rr;
            cplx<double> v = mk(0.0, 0.0);
            if (k < nk && r < nR) {
                double f = kl[3 * k] * iR[3 * r] + kl[3 * k + 1] * iR[3 * r + 1] + kl[3 * k + 2] * iR[3 * r + 2];
                f -= floor(f);
                double sn, cs;
                sincospi(2.0 * f, &sn, &cs);
                v = mk(cs, sn);
            }
            ph[q][rr] = v;
        }
        __syncthreads();
        if (nrow > 0 && ncol > 0) {
            const int rc = nR - r0 < KLM_RC ? nR - r0 : KLM_RC;
            const int prow = 32 * (wave / WE);
            const cplx<double>* Xr = X + (int64_t)r0 * nelem + ew;
            mfma_block(
                C, nrow, ncol, rc, 0, 0, [&](int r, int k) { return ph[prow + r][k]; },
                [&](int k, int c) { return Xr[(int64_t)k * nelem + c]; });
        }
    }
    if (nrow <= 0 || ncol <= 0) return;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = mfma_row(bi, v), c = mfma_col(bj);
                if (r < nrow && c < ncol) {
                    const cplx<double> z = C.get(bi, bj, v);
                    out[(int64_t)(kw + r) * nelem + ew + c] = mk((Tout)z.re, (Tout)z.im);
                }
            }
}

template <typename Tout, int WK, int WE>
static int launch_klist(wbmi_ctx* ctx, int nR, const int32_t* iR, int64_t nelem, const void* XR, int nk,
                        const double* kl, void* out) {
    constexpr int TK = 32 * WK, TE = 32 * WE;
    if (ceil_div(nelem, TE) > 65535)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_klist: more than %d elements per k-point", 65535 * TE);
    dim3 g((unsigned)ceil_div(nk, TK), (unsigned)ceil_div(nelem, TE));
    hipLaunchKernelGGL((k_r2k_klist<Tout, WK, WE>), g, dim3(64 * WK * WE), 0, ctx->stream, nR, iR, nelem,
                       (const cplx<double>*)XR, nk, kl, (cplx<Tout>*)out);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

template <typename Tout>
static int launch_klist_tile(wbmi_ctx* ctx, int tile, int nR, const int32_t* iR, int64_t nelem, const void* XR,
                             int nk, const double* kl, void* out) {
    switch (tile) {
        case 1: return launch_klist<Tout, 4, 2>(ctx, nR, iR, nelem, XR, nk, kl, out);
        case 2: return launch_klist<Tout, 1, 8>(ctx, nR, iR, nelem, XR, nk, kl, out);
        case 3: return launch_klist<Tout, 2, 2>(ctx, nR, iR, nelem, XR, nk, kl, out);
        case 4: return launch_klist<Tout, 1, 4>(ctx, nR, iR, nelem, XR, nk, kl, out);
        default: return launch_klist<Tout, 2, 4>(ctx, nR, iR, nelem, XR, nk, kl, out);
    }
}

extern "C" int wbmi_r2k_klist_tile(wbmi_ctx* ctx, int nR, const int32_t* iRvec_dev, int64_t nelem,
                                   const void* XR_dev, int nk, const double* klist_dev, void* out_dev, int prec,
                                   int tile) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nR <= 0 || nelem <= 0 || nk <= 0 || !iRvec_dev || !XR_dev || !klist_dev || !out_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_klist: bad arguments");
    if (tile < 0 || tile >= WBMI_KLIST_NTILES)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_klist_tile: tile must be 0..%d", WBMI_KLIST_NTILES - 1);
    if (prec == WBMI_F64)
        return launch_klist_tile<double>(ctx, tile, nR, iRvec_dev, nelem, XR_dev, nk, klist_dev, out_dev);
    if (prec == WBMI_F32)
        return launch_klist_tile<float>(ctx, tile, nR, iRvec_dev, nelem, XR_dev, nk, klist_dev, out_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_r2k_klist: bad precision");
}

extern "C" int wbmi_r2k_klist(wbmi_ctx* ctx, int nR, const int32_t* iRvec_dev, int64_t nelem, const void* XR_dev,
                              int nk, const double* klist_dev, void* out_dev, int prec) {
    return wbmi_r2k_klist_tile(ctx, nR, iRvec_dev, nelem, XR_dev, nk, klist_dev, out_dev, prec, 0);
}

extern "C" int wbmi_r2k_klist_tile_shape(int tile, int* k_rows, int* elements) {
    if (tile < 0 || tile >= WBMI_KLIST_NTILES || !k_rows || !elements) return WBMI_ERR_ARG;
    *k_rows = 32 * klm_tiles[tile][0];
    *elements = 32 * klm_tiles[tile][1];
    return WBMI_OK;
}
