// Per-band tables of channel diagonals (calculators/tabulate.py: Velocity, Spin).
//
// For Matrix_ln formulas whose traces over a band group are sums of diagonal
// elements (formula/covariant.py:321-336: the velocity, whose external term
// i A (E_m - E_n) vanishes on the diagonal, and the spin), the table value of
// band n is the mean of Re X^c_n'n' over the group of n (tabulate.py:40-46).
#include <cmath>

#include "wbmi_common.h"

#define TD_THREADS 64
#define TD_NWMAX 256

template <typename T>
__global__ void __launch_bounds__(TD_THREADS) k_tab_diag(int nw, int nch, int c0, int c1, int c2,
                                                         double degen_thresh, int degen_Kramers,
                                                         const double* __restrict__ Eall,
                                                         const cplx<T>* __restrict__ Xall, double* __restrict__ out) {
    __shared__ double sE[TD_NWMAX];
    __shared__ int sGb[TD_NWMAX + 1], sLab[TD_NWMAX];
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    for (int i = tid; i < nw; i += TD_THREADS) sE[i] = Eall[k * nw + i];
    __syncthreads();
    if (tid == 0) {
        // band groups (grid/tetrahedron.py:132-136)
        int nb = 0;
        sGb[nb++] = 0;
        for (int i = 1; i < nw; ++i)
            if (sE[i] - sE[i - 1] > degen_thresh) sGb[nb++] = i;
        sGb[nb++] = nw;
        if (degen_Kramers) {
            int m = 0;
            for (int i = 0; i < nb; ++i)
                if (sGb[i] % 2 == 0) sGb[m++] = sGb[i];
            nb = m;
        }
        for (int g = 0; g + 1 < nb; ++g)
            for (int n = sGb[g]; n < sGb[g + 1]; ++n) sLab[n] = g;
    }
    __syncthreads();
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + k * nch * nw2;
    const int ch[3] = {c0, c1, c2};
    for (int t = tid; t < 3 * nw; t += TD_THREADS) {
        const int n = t / 3, c = t % 3, g = sLab[n];
        double s = 0.0;
        for (int m = sGb[g]; m < sGb[g + 1]; ++m) s += (double)Xk[ch[c] * nw2 + (int64_t)m * nw + m].re;
        out[(k * nw + n) * 3 + c] = s / (sGb[g + 1] - sGb[g]);
    }
}

extern "C" int wbmi_tabulate_diag(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                  int prec, const int32_t* ch3_host, double degen_thresh, int degen_Kramers,
                                  double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !ch3_host || !out_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_diag: bad arguments");
    if (nw > TD_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_tabulate_diag: nw <= %d", TD_NWMAX);
    for (int c = 0; c < 3; ++c)
        if (ch3_host[c] < 0 || ch3_host[c] >= nch)
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_diag: channel %d out of range", ch3_host[c]);
    if (prec == WBMI_F64)
        hipLaunchKernelGGL(k_tab_diag<double>, dim3(nk), dim3(TD_THREADS), 0, ctx->stream, nw, nch, ch3_host[0],
                           ch3_host[1], ch3_host[2], degen_thresh, degen_Kramers, E_dev,
                           (const cplx<double>*)Xbar_dev, out_dev);
    else if (prec == WBMI_F32)
        hipLaunchKernelGGL(k_tab_diag<float>, dim3(nk), dim3(TD_THREADS), 0, ctx->stream, nw, nch, ch3_host[0],
                           ch3_host[1], ch3_host[2], degen_thresh, degen_Kramers, E_dev,
                           (const cplx<float>*)Xbar_dev, out_dev);
    else
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_diag: bad precision");
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
