// Per-band tables of channel diagonals (calculators/tabulate.py: Velocity, Spin) and of the
// inverse effective mass (InvMass).
//
// For Matrix_ln formulas whose traces over a band group are sums of diagonal
// elements (formula/covariant.py:321-336: the velocity, whose external term
// i A (E_m - E_n) vanishes on the diagonal, and the spin), the table value of
// band n is the mean of Re X^c_n'n' over the group of n (tabulate.py:40-46).
#include <cmath>

#include "wbmi_common.h"

#define TD_THREADS 64
#define TD_NWMAX 256

// channel slots of V (3) and W = Xbar('Ham', 2) (9, row-major)
struct wbmi_tab_invmass_ch {
    int V[3];
    int W[9];
};

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
        // a band in no group (the odd last band under degen_Kramers) keeps label -1: an out band
        // of every group, tabulated as 0 (tabulate.py:20-48 fills only the groups' bands)
        for (int n = 0; n < nw; ++n) sLab[n] = -1;
        for (int g = 0; g + 1 < nb; ++g)
            for (int n = sGb[g]; n < sGb[g + 1]; ++n) sLab[n] = g;
    }
    __syncthreads();
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + k * nch * nw2;
    const int ch[3] = {c0, c1, c2};
    for (int t = tid; t < 3 * nw; t += TD_THREADS) {
        const int n = t / 3, c = t % 3, g = sLab[n];
        if (g < 0) {
            out[(k * nw + n) * 3 + c] = 0.0;
            continue;
        }
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

// Per-band table of the inverse effective mass (formula/elementary.py:28-33, the generalised
// derivative of the velocity, formula.py:95-118): for band n of group G,
//   m^{bd}_n = Re [ W^{bd}_nn + sum_{l not in G} (V^b_nl D^d_ln - D^d_nl V^b_ln) ],
// D^d_xy = -V^d_xy / (E_x - E_y) (0 below 1e-7); the table holds the mean over G.
template <typename T>
__global__ void __launch_bounds__(TD_THREADS) k_tab_invmass(int nw, int nch, wbmi_tab_invmass_ch ch,
                                                            double degen_thresh, int degen_Kramers,
                                                            const double* __restrict__ Eall,
                                                            const cplx<T>* __restrict__ Xall, double* __restrict__ out) {
    __shared__ double sE[TD_NWMAX];
    __shared__ int sGb[TD_NWMAX + 1], sLab[TD_NWMAX];
    __shared__ double sv[TD_NWMAX * 9];
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    for (int i = tid; i < nw; i += TD_THREADS) sE[i] = Eall[k * nw + i];
    __syncthreads();
    if (tid == 0) {
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
        // a band in no group (the odd last band under degen_Kramers) keeps label -1: an out band
        // of every group, tabulated as 0 (tabulate.py:20-48 fills only the groups' bands)
        for (int n = 0; n < nw; ++n) sLab[n] = -1;
        for (int g = 0; g + 1 < nb; ++g)
            for (int n = sGb[g]; n < sGb[g + 1]; ++n) sLab[n] = g;
    }
    __syncthreads();
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + k * nch * nw2;
    auto ld = [&](int c, int i, int j) {
        const cplx<T> x = Xk[(int64_t)c * nw2 + (int64_t)i * nw + j];
        return mk((double)x.re, (double)x.im);
    };
    auto Dd = [&](int d, int i, int j) {
        const double de = sE[i] - sE[j];
        const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        return -m * ld(ch.V[d], i, j);
    };
    for (int t = tid; t < 9 * nw; t += TD_THREADS) {
        const int n = t / 9, b = (t % 9) / 3, d = t % 3;
        cplx<double> v = ld(ch.W[b * 3 + d], n, n);
        for (int l = 0; l < nw; ++l) {
            if (sLab[l] == sLab[n]) continue;
            v += ld(ch.V[b], n, l) * Dd(d, l, n) - Dd(d, n, l) * ld(ch.V[b], l, n);
        }
        sv[n * 9 + b * 3 + d] = v.re;
    }
    __syncthreads();
    for (int t = tid; t < 9 * nw; t += TD_THREADS) {
        const int n = t / 9, c = t % 9, g = sLab[n];
        if (g < 0) {
            out[(k * nw + n) * 9 + c] = 0.0;
            continue;
        }
        double s = 0.0;
        for (int m = sGb[g]; m < sGb[g + 1]; ++m) s += sv[m * 9 + c];
        out[(k * nw + n) * 9 + c] = s / (sGb[g + 1] - sGb[g]);
    }
}

extern "C" int wbmi_tabulate_invmass(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                     const void* Xbar_dev, int prec, const int32_t* chV3_host,
                                     const int32_t* chW9_host, double degen_thresh, int degen_Kramers,
                                     double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !chV3_host || !chW9_host || !out_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_invmass: bad arguments");
    if (nw > TD_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_tabulate_invmass: nw <= %d", TD_NWMAX);
    wbmi_tab_invmass_ch ch;
    for (int c = 0; c < 3; ++c) ch.V[c] = chV3_host[c];
    for (int c = 0; c < 9; ++c) ch.W[c] = chW9_host[c];
    for (int c = 0; c < 9; ++c)
        if ((c < 3 && (ch.V[c] < 0 || ch.V[c] >= nch)) || ch.W[c] < 0 || ch.W[c] >= nch)
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_invmass: channel out of range");
    if (prec == WBMI_F64)
        hipLaunchKernelGGL(k_tab_invmass<double>, dim3(nk), dim3(TD_THREADS), 0, ctx->stream, nw, nch, ch,
                           degen_thresh, degen_Kramers, E_dev, (const cplx<double>*)Xbar_dev, out_dev);
    else if (prec == WBMI_F32)
        hipLaunchKernelGGL(k_tab_invmass<float>, dim3(nk), dim3(TD_THREADS), 0, ctx->stream, nw, nch, ch,
                           degen_thresh, degen_Kramers, E_dev, (const cplx<float>*)Xbar_dev, out_dev);
    else
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_invmass: bad precision");
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
