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

// Third k-derivative of the band energy per band (formula/covariant.py:126-150 Der3E, with
// InvMass, DerDcov, DerWln of formula/elementary.py:28-70): for band m of group G
// (inn = G, out = the rest), c_m^{abc} = Re Der3E.nn[m, m, a, b, c] =
//   W'^{abc}_mm + sum_{l out} [ -D^c_ml W^{ab}_lm + W^{ab}_ml D^c_lm
//       + dV.nl^{ac}_ml D^b_lm + V^a_ml dD.ln^{bc}_lm - dD.nl^{bc}_ml V^a_lm - D^b_ml dV.ln^{ac}_lm ],
//   dV.ln^{ac}_lm = W^{ac}_lm - sum_{m' in} D^c_lm' V^a_m'm + sum_{l' out} V^a_ll' D^c_l'm,
//   dV.nl^{ac}_ml = W^{ac}_ml - sum_{l' out} D^c_ml' V^a_l'l + sum_{m' in} V^a_mm' D^c_m'l,
//   dD.ln^{bc}_lm = -e_lm [W^{bc}_lm + sum_{p out} (V^b_lp D^c_pm + V^c_lp D^b_pm)
//                          - sum_{q in} (D^b_lq V^c_qm + D^c_lq V^b_qm)],
//   dD.nl^{bc}_ml = -e_ml [W^{bc}_ml + sum_{p in} (V^b_mp D^c_pl + V^c_mp D^b_pl)
//                          - sum_{q out} (D^b_mq V^c_ql + D^c_mq V^b_ql)],
// D = D_H, D^c_xy = V^c_xy / (E_y - E_x) and e_xy = 1 / (E_x - E_y) (both 0 below 1e-7).
// One workgroup per k: V and D (complex128) in LDS, W and W' read from the channels; one
// thread per (band, component).  mode 0: the group mean of c (Der3E tabulator), mode 1: c_m
// itself, the band's share of its group's trace (NLDrude_FermiSea, binned by the scan).
#define D3_THREADS 256
#define D3_NWMAX 32

struct wbmi_der3e_ch {
    int V[3];
    int W[9];
    int W3[27];
};

template <typename T>
__global__ void __launch_bounds__(D3_THREADS) k_der3e(int nw, int nch, wbmi_der3e_ch ch, double degen_thresh,
                                                      int degen_Kramers, int mode, const double* __restrict__ Eall,
                                                      const cplx<T>* __restrict__ Xall, double* __restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x, nw2 = nw * nw;
    cplx<double>* sV = (cplx<double>*)smem;  // [3][nw][nw]
    cplx<double>* sD = sV + 3 * nw2;          // [3][nw][nw]
    double* sc = (double*)(sD + 3 * nw2);     // [nw][27]
    double* sE = sc + nw * 27;                // [D3_NWMAX]
    int* sGb = (int*)(sE + D3_NWMAX);         // [D3_NWMAX + 1]
    int* sLab = sGb + D3_NWMAX + 1;           // [D3_NWMAX]
    for (int i = tid; i < nw; i += D3_THREADS) sE[i] = Eall[k * nw + i];
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
        for (int n = 0; n < nw; ++n) sLab[n] = -1;  // past the last Kramers border: in no group
        for (int g = 0; g + 1 < nb; ++g)
            for (int n = sGb[g]; n < sGb[g + 1]; ++n) sLab[n] = g;
    }
    const cplx<T>* Xk = Xall + k * nch * (int64_t)nw2;
    for (int t = tid; t < 3 * nw2; t += D3_THREADS) {
        const int c = t / nw2, x = (t % nw2) / nw, y = t % nw;
        const cplx<T> v = Xk[(int64_t)ch.V[c] * nw2 + x * nw + y];
        const cplx<double> vd = mk((double)v.re, (double)v.im);
        const double de = sE[y] - sE[x];
        const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        sV[t] = vd;
        sD[t] = mk(m * vd.re, m * vd.im);
    }
    __syncthreads();
    auto V = [&](int a, int x, int y) { return sV[(a * nw + x) * nw + y]; };
    auto D = [&](int a, int x, int y) { return sD[(a * nw + x) * nw + y]; };
    auto W = [&](int a, int b, int x, int y) {
        const cplx<T> v = Xk[(int64_t)ch.W[a * 3 + b] * nw2 + x * nw + y];
        return mk((double)v.re, (double)v.im);
    };
    auto einv = [&](int x, int y) {
        const double de = sE[x] - sE[y];
        return fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
    };
    for (int t = tid; t < nw * 27; t += D3_THREADS) {
        const int m = t / 27, r = t % 27, a = r / 9, b = (r / 3) % 3, c = r % 3;
        const int gm = sLab[m];
        double val = 0.0;
        if (gm >= 0) {
            const cplx<T> w3 = Xk[(int64_t)ch.W3[r] * nw2 + m * nw + m];
            cplx<double> s = mk((double)w3.re, (double)w3.im);
            for (int l = 0; l < nw; ++l) {
                if (sLab[l] == gm) continue;
                s += W(a, b, m, l) * D(c, l, m) - D(c, m, l) * W(a, b, l, m);
                cplx<double> dvnl = W(a, c, m, l), dvln = W(a, c, l, m);
                cplx<double> ddln = W(b, c, l, m), ddnl = W(b, c, m, l);
                for (int q = 0; q < nw; ++q) {
                    if (sLab[q] == gm) {  // q in the group
                        dvnl += V(a, m, q) * D(c, q, l);
                        dvln -= D(c, l, q) * V(a, q, m);
                        ddln -= D(b, l, q) * V(c, q, m) + D(c, l, q) * V(b, q, m);
                        ddnl += V(b, m, q) * D(c, q, l) + V(c, m, q) * D(b, q, l);
                    } else {  // q out
                        dvnl -= D(c, m, q) * V(a, q, l);
                        dvln += V(a, l, q) * D(c, q, m);
                        ddln += V(b, l, q) * D(c, q, m) + V(c, l, q) * D(b, q, m);
                        ddnl -= D(b, m, q) * V(c, q, l) + D(c, m, q) * V(b, q, l);
                    }
                }
                ddln = (-einv(l, m)) * ddln;
                ddnl = (-einv(m, l)) * ddnl;
                s += dvnl * D(b, l, m) + V(a, m, l) * ddln - ddnl * V(a, l, m) - D(b, m, l) * dvln;
            }
            val = s.re;
        }
        sc[t] = val;
    }
    __syncthreads();
    for (int t = tid; t < nw * 27; t += D3_THREADS) {
        const int n = t / 27, r = t % 27, g = sLab[n];
        double s = 0.0;
        if (mode == 1) {
            s = sc[t];
        } else if (g >= 0) {
            for (int m = sGb[g]; m < sGb[g + 1]; ++m) s += sc[m * 27 + r];
            s /= sGb[g + 1] - sGb[g];
        }
        out[(k * nw + n) * 27 + r] = s;
    }
}

extern "C" int wbmi_der3e_bands(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                int prec, const int32_t* chV3_host, const int32_t* chW9_host,
                                const int32_t* chW27_host, double degen_thresh, int degen_Kramers, int mode,
                                double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !chV3_host || !chW9_host || !chW27_host ||
        !out_dev || (mode != 0 && mode != 1))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_der3e_bands: bad arguments");
    if (nw > D3_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_der3e_bands: nw <= %d", D3_NWMAX);
    wbmi_der3e_ch ch;
    for (int c = 0; c < 3; ++c) ch.V[c] = chV3_host[c];
    for (int c = 0; c < 9; ++c) ch.W[c] = chW9_host[c];
    for (int c = 0; c < 27; ++c) ch.W3[c] = chW27_host[c];
    auto bad = [&](int s) { return s < 0 || s >= nch; };
    for (int c = 0; c < 27; ++c)
        if ((c < 3 && bad(ch.V[c])) || (c < 9 && bad(ch.W[c])) || bad(ch.W3[c]))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_der3e_bands: channel out of range");
    const size_t lds = sizeof(cplx<double>) * 6 * nw * nw + sizeof(double) * (27 * nw + D3_NWMAX) +
                       sizeof(int) * (2 * D3_NWMAX + 1);
    if (prec == WBMI_F64) {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_der3e<double>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_der3e<double>, dim3(nk), dim3(D3_THREADS), lds, ctx->stream, nw, nch, ch, degen_thresh,
                           degen_Kramers, mode, E_dev, (const cplx<double>*)Xbar_dev, out_dev);
    } else if (prec == WBMI_F32) {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_der3e<float>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_der3e<float>, dim3(nk), dim3(D3_THREADS), lds, ctx->stream, nw, nch, ch, degen_thresh,
                           degen_Kramers, mode, E_dev, (const cplx<float>*)Xbar_dev, out_dev);
    } else {
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_der3e_bands: bad precision");
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
