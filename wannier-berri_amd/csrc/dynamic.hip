// Frequency-dependent calculators of libwbmi (calculators/dynamic.py:26-241):
// the Kubo interband optical conductivity (Formula_OptCond 170-181), the joint
// density of states (JDOS 145-167) and the spin Hall conductivity (SHC 204-241).
//
// Per k: the generalized Berry connection A^a = i D_H^a (+ Abar^a with external
// terms), D_H = -V o dEinv (data_K.py:290-334); for every ordered pair (m, n)
// of band groups (degen_thresh) whose occupations can differ inside the Fermi
// range, M^{ab} = sum_{i in m, j in n} i A^a_ij A^b_ji, and
//   out[EF][w][a][b] += (f(E_n) - f(E_m)) (E_n - E_m) / (E_n - E_m - w - i eta) M^{ab}
// (Gaussian smearing: the imaginary part is pi G(E_n - E_m - w)).
//
// Workgroup = (a tile of 256 (EF, w) outputs, one per thread with nine complex
// accumulators) x (a chunk of k-points).  For each k the workgroup forms A in
// LDS and walks the group pairs 256 at a time: every thread forms one pair's
// M into an LDS slot, then every thread adds all slots into its outputs.  The
// tile is added to `out` once, with FP64 atomics, at the end.
#include <cmath>

#include "wbmi_common.h"

#define OC_THREADS 256
#define OC_KCHUNK 16
#define OC_NWMAX 128  // band groups of one k (LDS arrays of this size)
#define OC_LDS_NW 32  // the per-k matrices (A, Y) live in LDS up to this nw, else in the workspace

struct OcArgs {
    wbmi_optcond_params p;
    int nw, nch;
};

// band groups of one k (grid/tetrahedron.py:132-136) into gb[0..ng], their mean energies
// into gm; returns ng.  One thread.
__device__ inline int band_groups(const double* E, int nw, double degen_thresh, int degen_Kramers, int* gb,
                                  double* gm) {
    int nb = 0;
    gb[nb++] = 0;
    for (int i = 1; i < nw; ++i)
        if (E[i] - E[i - 1] > degen_thresh) gb[nb++] = i;
    gb[nb++] = nw;
    if (degen_Kramers) {
        int m = 0;
        for (int i = 0; i < nb; ++i)
            if (gb[i] % 2 == 0) gb[m++] = gb[i];
        nb = m;
    }
    for (int g = 0; g < nb - 1; ++g) {
        double s = 0.0;
        for (int i = gb[g]; i < gb[g + 1]; ++i) s += E[i];
        gm[g] = s / (gb[g + 1] - gb[g]);
    }
    return nb - 1;
}

__device__ inline double fermi_dirac(double E, double mu, double kBT) {
    if (kBT == 0.0) return E <= mu ? 1.0 : 0.0;
    if (mu > E + 30.0 * kBT) return 1.0;
    if (mu < E - 30.0 * kBT) return 0.0;
    return 1.0 / (exp((E - mu) / kBT) + 1.0);
}

// A^a_ij = i D_H^a_ij (+ Abar^a_ij), D_H = -V dEinv (0 where |E_i - E_j| < 1e-7), of every k into
// Y[k][3][nw][nw] (KIND 0: -i A, the spin Hall Y), for nw > OC_LDS_NW where it does not fit in LDS.
template <typename T, int KIND>
__global__ void __launch_bounds__(256) k_build_y(int nw, int nch, wbmi_optcond_params p,
                                                 const double* __restrict__ Eall, const cplx<T>* __restrict__ Xall,
                                                 cplx<double>* __restrict__ Y) {
    const int64_t k = blockIdx.x;
    const int nw2 = nw * nw;
    const double* E = Eall + k * nw;
    const cplx<T>* Xk = Xall + k * nch * (int64_t)nw2;
    for (int t = threadIdx.x; t < 3 * nw2; t += blockDim.x) {
        const int c = t / nw2, i = (t % nw2) / nw, j = t % nw;
        const double de = E[i] - E[j];
        const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        const cplx<T> v = Xk[(int64_t)p.ch[WBMI_ROLE_V + c] * nw2 + i * nw + j];
        cplx<double> av = mk((double)v.im * m, -(double)v.re * m);  // i * (-m V)
        if (p.external_terms) {
            const cplx<T> x = Xk[(int64_t)p.ch[WBMI_ROLE_A + c] * nw2 + i * nw + j];
            av.re += (double)x.re;
            av.im += (double)x.im;
        }
        Y[k * 3 * nw2 + t] = KIND == 0 ? mk(av.im, -av.re) : av;
    }
}

// GA: A read from the workspace (nw > OC_LDS_NW, filled by k_build_y) instead of LDS.
template <typename T, bool GA>
__global__ void __launch_bounds__(OC_THREADS) k_optcond(int nk, OcArgs a, const double* __restrict__ Eall,
                                                        const cplx<T>* __restrict__ Xall,
                                                        const cplx<double>* __restrict__ Aws, double* __restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const wbmi_optcond_params& p = a.p;
    const int nw = a.nw, nw2 = nw * nw, tid = threadIdx.x;
    cplx<double>* sA = (cplx<double>*)smem;       // [3][nw][nw] (LDS form only)
    cplx<double>* sM = sA + (GA ? 0 : 3 * nw2);  // [OC_THREADS][9]
    double* sEm = (double*)(sM + OC_THREADS * 9);
    double* sEn = sEm + OC_THREADS;
    double* sE = sEn + OC_THREADS;
    double* sGm = sE + OC_NWMAX;
    int* sOk = (int*)(sGm + OC_NWMAX);
    int* sGb = sOk + OC_THREADS;
    int& sNg = sGb[OC_NWMAX + 1];
    const int o = blockIdx.x * OC_THREADS + tid;  // output index EF * nomega + w
    const bool has_out = o < p.nEF * p.nomega;
    const double mu = has_out ? p.Efermi[o / p.nomega] : 0.0;
    const double w = has_out ? p.omega[o % p.nomega] : 0.0;
    const double occ_lo = p.EFmin - 30.0 * p.kBT, occ_hi = p.EFmax + 30.0 * p.kBT;
    cplx<double> acc[9];
#pragma unroll
    for (int ab = 0; ab < 9; ++ab) acc[ab] = mk(0.0, 0.0);
    const int k1 = min(nk, (int)(blockIdx.y + 1) * OC_KCHUNK);
    for (int k = blockIdx.y * OC_KCHUNK; k < k1; ++k) {
        for (int i = tid; i < nw; i += OC_THREADS) sE[i] = Eall[(int64_t)k * nw + i];
        __syncthreads();
        if (tid == 0) sNg = band_groups(sE, nw, p.degen_thresh, p.degen_Kramers, sGb, sGm);
        // A^a_ij = i D_H^a_ij (+ Abar^a_ij), D_H = -V dEinv (0 where |E_i - E_j| < 1e-7)
        const cplx<double>* Ak = GA ? Aws + (int64_t)k * 3 * nw2 : sA;
        if (!GA) {
            const cplx<T>* Xk = Xall + (int64_t)k * a.nch * nw2;
            for (int t = tid; t < 3 * nw2; t += OC_THREADS) {
                const int c = t / nw2, i = (t % nw2) / nw, j = t % nw;
                const double de = sE[i] - sE[j];
                const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
                const cplx<T> v = Xk[(int64_t)p.ch[WBMI_ROLE_V + c] * nw2 + i * nw + j];
                cplx<double> av = mk((double)v.im * m, -(double)v.re * m);  // i * (-m V)
                if (p.external_terms) {
                    const cplx<T> x = Xk[(int64_t)p.ch[WBMI_ROLE_A + c] * nw2 + i * nw + j];
                    av.re += (double)x.re;
                    av.im += (double)x.im;
                }
                sA[t] = av;
            }
        }
        __syncthreads();
        const int ng = sNg, ncand = ng * ng;
        for (int c0 = 0; c0 < ncand; c0 += OC_THREADS) {
            const int cnd = c0 + tid;
            int ok = 0;
            if (cnd < ncand) {
                const int gm = cnd / ng, gn = cnd % ng;
                const double Em = sGm[gm], En = sGm[gn];
                ok = !((Em < occ_lo && En < occ_lo) || (Em > occ_hi && En > occ_hi));
                if (ok) {
                    sEm[tid] = Em;
                    sEn[tid] = En;
                    for (int ab = 0; ab < 9; ++ab) {
                        const int ca = ab / 3, cb = ab % 3;
                        cplx<double> s = mk(0.0, 0.0);
                        for (int i = sGb[gm]; i < sGb[gm + 1]; ++i)
                            for (int j = sGb[gn]; j < sGb[gn + 1]; ++j)
                                s += Ak[(ca * nw + i) * nw + j] * Ak[(cb * nw + j) * nw + i];
                        sM[tid * 9 + ab] = times_i(s);
                    }
                }
            }
            sOk[tid] = ok;
            __syncthreads();
            const int ns = ncand - c0 < OC_THREADS ? ncand - c0 : OC_THREADS;
            if (has_out) {
                for (int s = 0; s < ns; ++s) {
                    if (!sOk[s]) continue;
                    const double Em = sEm[s], En = sEn[s];
                    const double fE = fermi_dirac(En, mu, p.kBT) - fermi_dirac(Em, mu, p.kBT);
                    if (fE == 0.0) continue;
                    // (E_n - E_m) / (x - i eta), x = E_n - E_m - w
                    const double x = En - Em - w;
                    const double den = 1.0 / (x * x + p.eta * p.eta);
                    double fr = x * den, fi = p.eta * den;
                    if (p.smr_gaussian) {
                        const double y = x / p.eta;
                        fi = fabs(x) < p.eta * sqrt(200.0) ? M_PI / (sqrt(M_PI) * p.eta) * exp(-y * y) : 0.0;
                    }
                    const double sc = (En - Em) * fE;
                    const cplx<double> f = mk(sc * fr, sc * fi);
#pragma unroll
                    for (int ab = 0; ab < 9; ++ab) acc[ab] += f * sM[s * 9 + ab];
                }
            }
            __syncthreads();
        }
    }
    if (has_out) {
#pragma unroll
        for (int ab = 0; ab < 9; ++ab) {
            unsafeAtomicAdd(out + ((int64_t)o * 9 + ab) * 2, acc[ab].re);
            unsafeAtomicAdd(out + ((int64_t)o * 9 + ab) * 2 + 1, acc[ab].im);
        }
    }
}

extern "C" size_t wbmi_optical_conductivity_work_bytes(int nk, int nw) {
    return nw > OC_LDS_NW ? sizeof(cplx<double>) * 3 * (size_t)nk * nw * nw : 0;
}

template <typename T>
static int launch_optcond(wbmi_ctx* ctx, int nk, const OcArgs& a, const double* E, const void* X, void* work,
                          double* out) {
    const int nw = a.nw;
    const bool ga = nw > OC_LDS_NW;
    const int64_t nout = (int64_t)a.p.nEF * a.p.nomega;
    dim3 g((unsigned)ceil_div(nout, OC_THREADS), (unsigned)ceil_div(nk, OC_KCHUNK));
    const size_t lds = sizeof(cplx<double>) * ((ga ? 0 : 3 * (size_t)nw * nw) + OC_THREADS * 9) +
                       sizeof(double) * (2 * OC_THREADS + 2 * OC_NWMAX) + sizeof(int) * (OC_THREADS + OC_NWMAX + 2);
    if (ga) {
        hipLaunchKernelGGL((k_build_y<T, 1>), dim3(nk), dim3(256), 0, ctx->stream, nw, a.nch, a.p, E,
                           (const cplx<T>*)X, (cplx<double>*)work);
        WBMI_LAUNCH_CHECK(ctx);
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_optcond<T, true>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_optcond<T, true>), g, dim3(OC_THREADS), lds, ctx->stream, nk, a, E, (const cplx<T>*)X,
                           (const cplx<double>*)work, out);
    } else {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_optcond<T, false>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_optcond<T, false>), g, dim3(OC_THREADS), lds, ctx->stream, nk, a, E,
                           (const cplx<T>*)X, (const cplx<double>*)nullptr, out);
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

static int optcond(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev, int prec,
                   const wbmi_optcond_params* params, void* work_dev, size_t work_bytes, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !params || !out_dev || !params->Efermi ||
        !params->omega || params->nEF <= 0 || params->nomega <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_optical_conductivity: bad arguments");
    if (nw > OC_NWMAX)
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_optical_conductivity: nw <= %d", OC_NWMAX);
    const size_t need = wbmi_optical_conductivity_work_bytes(nk, nw);
    if (need > 0 && (!work_dev || work_bytes < need))
        return wbmi_fail(ctx, WBMI_ERR_ARG,
                         "wbmi_optical_conductivity: nw = %d > %d needs a workspace of %zu bytes "
                         "(wbmi_optical_conductivity_ws)", nw, OC_LDS_NW, need);
    if (!(params->eta > 0) || params->kBT < 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_optical_conductivity: smearing width must be positive");
    for (int c = 0; c < 3; ++c) {
        const int v = params->ch[WBMI_ROLE_V + c], x = params->ch[WBMI_ROLE_A + c];
        if (v < 0 || v >= nch || (params->external_terms && (x < 0 || x >= nch)))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_optical_conductivity: missing V or A channel");
    }
    OcArgs a;
    a.p = *params;
    a.nw = nw;
    a.nch = nch;
    if (prec == WBMI_F64) return launch_optcond<double>(ctx, nk, a, E_dev, Xbar_dev, work_dev, out_dev);
    if (prec == WBMI_F32) return launch_optcond<float>(ctx, nk, a, E_dev, Xbar_dev, work_dev, out_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_optical_conductivity: bad precision");
}

extern "C" int wbmi_optical_conductivity(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                         const void* Xbar_dev, int prec, const wbmi_optcond_params* params,
                                         double* out_dev) {
    return optcond(ctx, nk, nw, nch, E_dev, Xbar_dev, prec, params, nullptr, 0, out_dev);
}

extern "C" int wbmi_optical_conductivity_ws(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                            const void* Xbar_dev, int prec, const wbmi_optcond_params* params,
                                            void* work_dev, size_t work_bytes, double* out_dev) {
    return optcond(ctx, nk, nw, nch, E_dev, Xbar_dev, prec, params, work_dev, work_bytes, out_dev);
}

// Joint density of states (dynamic.py:145-167): per k, every ordered pair (m, n) of band
// groups with E_m < max(Efermi), E_n > min(Efermi) and
// min(omega) - 5 eta < E_m - E_n < max(omega) + 5 eta adds
//   out[EF][w] += |m| |n| (f(E_n) - f(E_m)) smear(E_m - E_n - w).
// Workgroup = (256 (EF, w) outputs, one per thread) x (a chunk of k-points); the pair
// list of a k is shared, each thread accumulates its own output.
#define JD_THREADS 256
#define JD_KCHUNK 32
#define JD_NWMAX 256

__global__ void __launch_bounds__(JD_THREADS) k_jdos(int nk, wbmi_optcond_params p, int nw,
                                                     const double* __restrict__ Eall, double* __restrict__ out) {
    __shared__ double sE[JD_NWMAX], sGm[JD_NWMAX];
    __shared__ int sGb[JD_NWMAX + 1];
    __shared__ int sNg;
    const int tid = threadIdx.x;
    const int o = blockIdx.x * JD_THREADS + tid;  // output index EF * nomega + w
    const bool has_out = o < p.nEF * p.nomega;
    const double mu = has_out ? p.Efermi[o / p.nomega] : 0.0;
    const double w = has_out ? p.omega[o % p.nomega] : 0.0;
    const double wlo = p.omega_min - 5.0 * p.eta, whi = p.omega_max + 5.0 * p.eta;
    double acc = 0.0;
    const int k1 = min(nk, (int)(blockIdx.y + 1) * JD_KCHUNK);
    for (int k = blockIdx.y * JD_KCHUNK; k < k1; ++k) {
        for (int i = tid; i < nw; i += JD_THREADS) sE[i] = Eall[(int64_t)k * nw + i];
        __syncthreads();
        if (tid == 0) sNg = band_groups(sE, nw, p.degen_thresh, p.degen_Kramers, sGb, sGm);
        __syncthreads();
        const int ng = sNg;
        if (has_out)
            for (int gm = 0; gm < ng; ++gm) {
                const double Em = sGm[gm];
                if (!(Em < p.EFmax)) continue;
                for (int gn = 0; gn < ng; ++gn) {
                    const double En = sGm[gn], d = Em - En;
                    if (!(En > p.EFmin) || !(d > wlo) || !(d < whi)) continue;
                    const double fE = fermi_dirac(En, mu, p.kBT) - fermi_dirac(Em, mu, p.kBT);
                    if (fE == 0.0) continue;
                    const double x = d - w;
                    double sm;
                    if (p.smr_gaussian) {
                        const double y = x / p.eta;
                        sm = fabs(x) < p.eta * sqrt(200.0) ? exp(-y * y) / (sqrt(M_PI) * p.eta) : 0.0;
                    } else {
                        sm = 1.0 / (M_PI * p.eta) * (p.eta * p.eta) / (x * x + p.eta * p.eta);
                    }
                    acc += (double)((sGb[gm + 1] - sGb[gm]) * (sGb[gn + 1] - sGb[gn])) * fE * sm;
                }
            }
        __syncthreads();
    }
    if (has_out) unsafeAtomicAdd(out + o, acc);
}

extern "C" int wbmi_jdos(wbmi_ctx* ctx, int nk, int nw, const double* E_dev, const wbmi_optcond_params* params,
                         double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || !E_dev || !params || !out_dev || !params->Efermi || !params->omega ||
        params->nEF <= 0 || params->nomega <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_jdos: bad arguments");
    if (nw > JD_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_jdos: nw <= %d", JD_NWMAX);
    if (!(params->eta > 0) || params->kBT < 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_jdos: smearing width must be positive");
    const int64_t nout = (int64_t)params->nEF * params->nomega;
    dim3 g((unsigned)ceil_div(nout, JD_THREADS), (unsigned)ceil_div(nk, JD_KCHUNK));
    hipLaunchKernelGGL(k_jdos, g, dim3(JD_THREADS), 0, ctx->stream, nk, *params, nw, E_dev, out_dev);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// ---------------------------------------------------------------------------
// Kubo spin Hall conductivity (dynamic.py:204-241) with the Ryoo spin current
// (formula/covariant.py:689-756).
// ---------------------------------------------------------------------------

// J^{as} = (J' + J'^+) / 2, one workgroup per (k, component a*3+s), one thread per (i, j);
// both J'_ij and J'_ji are formed by the same thread, so no barrier is needed.
//   simple (covariant.py:721-727): J'^{as}_ij = sum_l SS^s_il v^a_lj, the Velocity
//     v^a_lj = V^a_lj (+ i Abar^a_lj (E_l - E_j) with external terms, covariant.py:321-328)
//   Ryoo (covariant.py:746-756): J'^{as}_ij = -i (E_j SA^{as}_ij - SHA^{as}_ij) + sum_l SS^s_il V^a_lj
//   Qiao (covariant.py:729-744): J'^{as}_ij = v^a_j SS^s_ij + E_j (-i SR^{as}_ij + sum_l SS^s_il D^a_lj)
//                                             - (-i SHR^{as}_ij + sum_l SH^s_il D^a_lj),
//     v^a_j = Re V^a_jj, D^a_lj = -V^a_lj / (E_l - E_j) (0 below 1e-7).
// J: complex128 [nk, 9, nw, nw].
template <typename T>
__global__ void __launch_bounds__(256) k_spin_current(int nk, int nw, int nch, wbmi_shc_params p,
                                                      const double* __restrict__ Eall,
                                                      const cplx<T>* __restrict__ Xall, cplx<double>* __restrict__ J) {
    const int k = blockIdx.x, comp = blockIdx.y, a = comp / 3, sp = comp % 3;
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + (int64_t)k * nch * nw2;
    const bool qiao = p.kind == WBMI_SHC_QIAO, simple = p.kind == WBMI_SHC_SIMPLE;
    const bool ext = p.base.external_terms != 0;
    // Ryoo: SA, SHA; Qiao: SR, SHR (the same roles in the formula), and SH; simple: none
    const cplx<T>* SA = simple ? nullptr : Xk + (qiao ? p.ch_SR[comp] : p.ch_SA[comp]) * nw2;
    const cplx<T>* SHA = simple ? nullptr : Xk + (qiao ? p.ch_SHR[comp] : p.ch_SHA[comp]) * nw2;
    const cplx<T>* Ab = simple && ext ? Xk + p.base.ch[WBMI_ROLE_A + a] * nw2 : nullptr;
    const cplx<T>* S = Xk + p.ch_S[sp] * nw2;
    const cplx<T>* SH = Xk + (qiao ? p.ch_SH[sp] : p.ch_S[sp]) * nw2;
    const cplx<T>* V = Xk + p.base.ch[WBMI_ROLE_V + a] * nw2;
    const double* E = Eall + (int64_t)k * nw;
    auto ld = [](const cplx<T>* X, int64_t t) { return mk((double)X[t].re, (double)X[t].im); };
    auto dh = [&](int l, int j) {  // D^a_lj
        const double de = E[l] - E[j];
        const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        const cplx<double> v = ld(V, l * nw + j);
        return mk(-m * v.re, -m * v.im);
    };
    auto jprime = [&](int i, int j) {
        if (simple) {
            cplx<double> v = mk(0.0, 0.0);
            for (int l = 0; l < nw; ++l) {
                cplx<double> vel = ld(V, l * nw + j);
                if (ext) {
                    const cplx<double> x = ld(Ab, l * nw + j);
                    const double de = E[l] - E[j];
                    vel += mk(-x.im * de, x.re * de);  // i Abar (E_l - E_j)
                }
                v += ld(S, i * nw + l) * vel;
            }
            return v;
        }
        const cplx<double> sa = ld(SA, i * nw + j), sha = ld(SHA, i * nw + j);
        // -i (E_j SA - SHA)  (Qiao: -i (E_j SR - SHR))
        cplx<double> v = mk(E[j] * sa.im - sha.im, -(E[j] * sa.re - sha.re));
        if (!qiao) {
            for (int l = 0; l < nw; ++l) v += ld(S, i * nw + l) * ld(V, l * nw + j);
        } else {
            cplx<double> ks = mk(0.0, 0.0), ls = mk(0.0, 0.0);
            for (int l = 0; l < nw; ++l) {
                const cplx<double> d = dh(l, j);
                ks += ld(S, i * nw + l) * d;
                ls += ld(SH, i * nw + l) * d;
            }
            v += E[j] * ks - ls;
            v += ld(V, j * nw + j).re * ld(S, i * nw + j);
        }
        return v;
    };
    for (int t = threadIdx.x; t < nw2; t += blockDim.x) {
        const int i = t / nw, j = t % nw;
        const cplx<double> x = jprime(i, j), y = jprime(j, i);
        J[((int64_t)k * 9 + comp) * nw2 + t] = mk(0.5 * (x.re + y.re), 0.5 * (x.im - y.im));
    }
}

// Shift current: generalised derivative of the Berry connection (dynamic.py:246-288),
// Ag[n][m][p][q] for one (p, q) per workgroup, one thread per (n, m):
//   Ag = i (W^{pq} + HD^{pq} + DV^{pq}) / (E_m - E_n)   (0 below 1e-7)
//        [+ Abar'^{pq} + AD^{pq} - i AA^{pq} + AD'^{pq}  with external terms]
//   HD^{ca}_nm = sum_l V^c_nl P^a_lm - V^c_nn P^a_nm - sum_l P^a_nl V^c_lm + P^a_nm V^c_mm,
//   DV^{ca}_nm = D^c_nm (V^a_nn - V^a_mm) + D^a_nm (V^c_nn - V^c_mm),
//   AD^{ac}_nm = D^a_nm (A^c_nn - A^c_mm) + D^c_nm (A^a_nn - A^a_mm),
//   AA^{ab}_nm = A^a_nm (A^b_nn - A^b_mm),  AD'^{ca} = HD^{ca} with A in place of V,
//   P^a_nm = -V^a_nm (E_n - E_m) / ((E_n - E_m)^2 + sc_eta^2),  D = D_H.
// Out: complex128 [nk, 9, nw, nw] (component p*3 + q).
template <typename T>
__global__ void __launch_bounds__(256) k_shift_ag(int nk, int nw, int nch, wbmi_optcond_params p, double sc_eta,
                                                  const double* __restrict__ Eall, const cplx<T>* __restrict__ Xall,
                                                  cplx<double>* __restrict__ Ag) {
    const int k = blockIdx.x, comp = blockIdx.y, cp = comp / 3, cq = comp % 3;
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + (int64_t)k * nch * nw2;
    const double* E = Eall + (int64_t)k * nw;
    const bool ext = p.external_terms != 0;
    auto ld = [&](int role, int i, int j) {
        const cplx<T> x = Xk[(int64_t)p.ch[role] * nw2 + (int64_t)i * nw + j];
        return mk((double)x.re, (double)x.im);
    };
    auto dinv = [&](int i, int j) {  // 1 / (E_i - E_j), 0 below 1e-7
        const double de = E[i] - E[j];
        return fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
    };
    auto Dh = [&](int a, int i, int j) { return -dinv(i, j) * ld(WBMI_ROLE_V + a, i, j); };
    auto Pv = [&](int a, int i, int j) {
        const double de = E[i] - E[j];
        return -(de / (de * de + sc_eta * sc_eta)) * ld(WBMI_ROLE_V + a, i, j);
    };
    // HD-type term with X (V or A) in the c slot and P in the a slot, positions (c, a) = (cp, cq)
    auto xp = [&](int role0, int n, int m) {
        const int c = cp, a = cq;
        cplx<double> s = mk(0.0, 0.0);
        for (int l = 0; l < nw; ++l) s += ld(role0 + c, n, l) * Pv(a, l, m) - Pv(a, n, l) * ld(role0 + c, l, m);
        s -= ld(role0 + c, n, n) * Pv(a, n, m);
        s += Pv(a, n, m) * ld(role0 + c, m, m);
        return s;
    };
    for (int t = threadIdx.x; t < nw2; t += blockDim.x) {
        const int n = t / nw, m = t % nw;
        const int c = cp, a = cq;
        cplx<double> g = ld(WBMI_ROLE_W + c * 3 + a, n, m) + xp(WBMI_ROLE_V, n, m);
        g += Dh(c, n, m) * (ld(WBMI_ROLE_V + a, n, n) - ld(WBMI_ROLE_V + a, m, m));
        g += Dh(a, n, m) * (ld(WBMI_ROLE_V + c, n, n) - ld(WBMI_ROLE_V + c, m, m));
        const double di = dinv(m, n);
        cplx<double> v = mk(-g.im * di, g.re * di);  // i g / (E_m - E_n)
        if (ext) {
            // positions (p, q) = (cp, cq): Abar'[.., a = p, d = q], AD[.., a = p, c = q],
            // AA[.., a = p, b = q], AD' (HD with A) at (c = p, a = q)
            const int pp = cp, qq = cq;
            v += ld(WBMI_ROLE_AD + pp * 3 + qq, n, m);
            v += Dh(pp, n, m) * (ld(WBMI_ROLE_A + qq, n, n) - ld(WBMI_ROLE_A + qq, m, m));
            v += Dh(qq, n, m) * (ld(WBMI_ROLE_A + pp, n, n) - ld(WBMI_ROLE_A + pp, m, m));
            const cplx<double> aa = ld(WBMI_ROLE_A + pp, n, m) * (ld(WBMI_ROLE_A + qq, n, n) - ld(WBMI_ROLE_A + qq, m, m));
            v += mk(aa.im, -aa.re);  // -i AA
            v += xp(WBMI_ROLE_A, n, m);
        }
        Ag[((int64_t)k * 9 + comp) * nw2 + t] = v;
    }
}

// Injection current (dynamic.py:327-361): X^{ab}_nm = (v^a_n - v^a_m) A^b_nm, v^a_n = Re V^a_nn,
// A = i D_H (+ Abar).  Out: complex128 [nk, 9, nw, nw] (component a*3 + b).
template <typename T>
__global__ void __launch_bounds__(256) k_inject_x(int nk, int nw, int nch, wbmi_optcond_params p,
                                                  const double* __restrict__ Eall, const cplx<T>* __restrict__ Xall,
                                                  cplx<double>* __restrict__ X) {
    const int k = blockIdx.x, comp = blockIdx.y, a = comp / 3, b = comp % 3;
    const int64_t nw2 = (int64_t)nw * nw;
    const cplx<T>* Xk = Xall + (int64_t)k * nch * nw2;
    const double* E = Eall + (int64_t)k * nw;
    const cplx<T>* Va = Xk + (int64_t)p.ch[WBMI_ROLE_V + a] * nw2;
    const cplx<T>* Vb = Xk + (int64_t)p.ch[WBMI_ROLE_V + b] * nw2;
    const cplx<T>* Ab = Xk + (int64_t)p.ch[WBMI_ROLE_A + b] * nw2;
    for (int t = threadIdx.x; t < nw2; t += blockDim.x) {
        const int n = t / nw, m = t % nw;
        const double de = E[n] - E[m];
        const double mi = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        // A = i (-mi V) (+ Abar)
        cplx<double> A = mk(mi * (double)Vb[t].im, -mi * (double)Vb[t].re);
        if (p.external_terms) A += mk((double)Ab[t].re, (double)Ab[t].im);
        const double dv = (double)Va[(int64_t)n * nw + n].re - (double)Va[(int64_t)m * nw + m].re;
        X[((int64_t)k * 9 + comp) * nw2 + t] = dv * A;
    }
}

#define P27_PAIRS 128  // pair slots per round (27 complex sums each)

// Pair walk shared by the spin Hall conductivity (KIND 0), the shift current (1) and the
// injection current (2).  Per k: Y^b in LDS (KIND 0: -i A, else A; A = i D_H (+ Abar));
// for every pair (m, n) of band groups that can change occupation (128 at a time),
//   P^{x,b} = sum_{i in m, j in n} X^x_ij Y^b_ji   (x = 0..8, from the per-k workspace)
// into an LDS slot; then every thread adds all slots into its (EF, w) output:
//   KIND 0: out[abs] += (f_n - f_m) Im P^{as,b} / (2 (E_m - E_n - w - i eta))   (Gaussian: pi G / 2)
//   KIND 1: out[abc] += (f_n - f_m) (-Im P^{ca,b} - Im P^{ba,c}) (g(E_m - E_n - w) + g(E_n - E_m - w))
//   KIND 2: out[abc] += (f_n - f_m) P^{ab,c} g(E_m - E_n - w)
// (g = the smearing function).  One FP64 atomic add per output per workgroup.
// GY: Y read from the workspace (nw > OC_LDS_NW, filled by k_build_y) instead of LDS.
template <typename T, int KIND, bool GY>
__global__ void __launch_bounds__(OC_THREADS) k_pair27(int nk, wbmi_optcond_params p, int nw, int nch,
                                                       const double* __restrict__ Eall,
                                                       const cplx<T>* __restrict__ Xall,
                                                       const cplx<double>* __restrict__ W,
                                                       const cplx<double>* __restrict__ Yws, double* __restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int nw2 = nw * nw, tid = threadIdx.x;
    cplx<double>* sYl = (cplx<double>*)smem;             // [3][nw][nw] (LDS form only)
    cplx<double>* sP = sYl + (GY ? 0 : 3 * nw2);        // [P27_PAIRS][27]
    double* sEm = (double*)(sP + P27_PAIRS * 27);
    double* sEn = sEm + P27_PAIRS;
    double* sE = sEn + P27_PAIRS;
    double* sGm = sE + OC_NWMAX;
    int* sOk = (int*)(sGm + OC_NWMAX);
    int* sGb = sOk + P27_PAIRS;
    int& sNg = sGb[OC_NWMAX + 1];
    const int o = blockIdx.x * OC_THREADS + tid;
    const bool has_out = o < p.nEF * p.nomega;
    const double mu = has_out ? p.Efermi[o / p.nomega] : 0.0;
    const double w = has_out ? p.omega[o % p.nomega] : 0.0;
    const double occ_lo = p.EFmin - 30.0 * p.kBT, occ_hi = p.EFmax + 30.0 * p.kBT;
    auto smear = [&](double x) {
        if (p.smr_gaussian) {
            const double y = x / p.eta;
            return fabs(x) < p.eta * sqrt(200.0) ? exp(-y * y) / (sqrt(M_PI) * p.eta) : 0.0;
        }
        return 1.0 / (M_PI * p.eta) * (p.eta * p.eta) / (x * x + p.eta * p.eta);
    };
    cplx<double> acc[27];
#pragma unroll
    for (int c = 0; c < 27; ++c) acc[c] = mk(0.0, 0.0);
    const int k1 = min(nk, (int)(blockIdx.y + 1) * OC_KCHUNK);
    for (int k = blockIdx.y * OC_KCHUNK; k < k1; ++k) {
        for (int i = tid; i < nw; i += OC_THREADS) sE[i] = Eall[(int64_t)k * nw + i];
        __syncthreads();
        if (tid == 0) sNg = band_groups(sE, nw, p.degen_thresh, p.degen_Kramers, sGb, sGm);
        const cplx<double>* sY = GY ? Yws + (int64_t)k * 3 * nw2 : sYl;
        if (!GY) {
            const cplx<T>* Xk = Xall + (int64_t)k * nch * nw2;
            for (int t = tid; t < 3 * nw2; t += OC_THREADS) {
                const int c = t / nw2, i = (t % nw2) / nw, j = t % nw;
                const double de = sE[i] - sE[j];
                const double m = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
                const cplx<T> v = Xk[(int64_t)p.ch[WBMI_ROLE_V + c] * nw2 + i * nw + j];
                cplx<double> a = mk(m * (double)v.im, -m * (double)v.re);  // i D_H = i (-m V)
                if (p.external_terms) {
                    const cplx<T> x = Xk[(int64_t)p.ch[WBMI_ROLE_A + c] * nw2 + i * nw + j];
                    a += mk((double)x.re, (double)x.im);
                }
                sYl[t] = KIND == 0 ? mk(a.im, -a.re) : a;  // KIND 0: -i A
            }
        }
        __syncthreads();
        const cplx<double>* Wk = W + (int64_t)k * 9 * nw2;
        const int ng = sNg, ncand = ng * ng;
        for (int c0 = 0; c0 < ncand; c0 += P27_PAIRS) {
            const int cnd = c0 + tid;
            if (tid < P27_PAIRS) {
                int ok = 0;
                if (cnd < ncand) {
                    const int gm = cnd / ng, gn = cnd % ng;
                    const double Em = sGm[gm], En = sGm[gn];
                    ok = !((Em < occ_lo && En < occ_lo) || (Em > occ_hi && En > occ_hi));
                    if (ok) {
                        sEm[tid] = Em;
                        sEn[tid] = En;
                        for (int x = 0; x < 9; ++x) {
                            const cplx<double>* Wx = Wk + x * nw2;
                            cplx<double> q0 = mk(0.0, 0.0), q1 = mk(0.0, 0.0), q2 = mk(0.0, 0.0);
                            for (int i = sGb[gm]; i < sGb[gm + 1]; ++i)
                                for (int j = sGb[gn]; j < sGb[gn + 1]; ++j) {
                                    const cplx<double> xv = Wx[i * nw + j];
                                    q0 += xv * sY[j * nw + i];
                                    q1 += xv * sY[nw2 + j * nw + i];
                                    q2 += xv * sY[2 * nw2 + j * nw + i];
                                }
                            sP[tid * 27 + x * 3 + 0] = q0;
                            sP[tid * 27 + x * 3 + 1] = q1;
                            sP[tid * 27 + x * 3 + 2] = q2;
                        }
                    }
                }
                sOk[tid] = ok;
            }
            __syncthreads();
            const int ns = ncand - c0 < P27_PAIRS ? ncand - c0 : P27_PAIRS;
            if (has_out) {
                for (int s = 0; s < ns; ++s) {
                    if (!sOk[s]) continue;
                    const double Em = sEm[s], En = sEn[s];
                    const double fE = fermi_dirac(En, mu, p.kBT) - fermi_dirac(Em, mu, p.kBT);
                    if (fE == 0.0) continue;
                    const cplx<double>* P = sP + s * 27;
                    if (KIND == 0) {
                        const double x = Em - En - w;
                        const double den = 1.0 / (x * x + p.eta * p.eta);
                        const double fr = x * den;
                        const double fi = p.smr_gaussian ? M_PI * smear(x) : p.eta * den;
                        const cplx<double> f = mk(0.5 * fE * fr, 0.5 * fE * fi);
#pragma unroll
                        for (int a = 0; a < 3; ++a)
#pragma unroll
                            for (int b = 0; b < 3; ++b)
#pragma unroll
                                for (int sp = 0; sp < 3; ++sp) {
                                    const double mv = P[(a * 3 + sp) * 3 + b].im;
                                    acc[(a * 3 + b) * 3 + sp] += mv * f;
                                }
                    } else if (KIND == 1) {
                        const double f = fE * (smear(Em - En - w) + smear(En - Em - w));
#pragma unroll
                        for (int a = 0; a < 3; ++a)
#pragma unroll
                            for (int b = 0; b < 3; ++b)
#pragma unroll
                                for (int c = 0; c < 3; ++c) {
                                    const double mv = -P[(c * 3 + a) * 3 + b].im - P[(b * 3 + a) * 3 + c].im;
                                    acc[(a * 3 + b) * 3 + c].re += f * mv;
                                }
                    } else {
                        const double f = fE * smear(Em - En - w);
#pragma unroll
                        for (int q = 0; q < 27; ++q) acc[q] += f * P[q];
                    }
                }
            }
            __syncthreads();
        }
    }
    if (has_out) {
#pragma unroll
        for (int c = 0; c < 27; ++c) {
            unsafeAtomicAdd(out + ((int64_t)o * 27 + c) * 2, acc[c].re);
            unsafeAtomicAdd(out + ((int64_t)o * 27 + c) * 2 + 1, acc[c].im);
        }
    }
}

// Wk: the per-k workspace (X^x, 9 nw^2 per k); for nw > OC_LDS_NW, Y follows it (3 nw^2 per k).
template <typename T, int KIND>
static int launch_pair27(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E, const void* X,
                         const wbmi_optcond_params* params, const cplx<double>* Wk, double* out) {
    const bool gy = nw > OC_LDS_NW;
    const int64_t nout = (int64_t)params->nEF * params->nomega;
    dim3 g((unsigned)ceil_div(nout, OC_THREADS), (unsigned)ceil_div(nk, OC_KCHUNK));
    const size_t lds = sizeof(cplx<double>) * ((gy ? 0 : 3 * (size_t)nw * nw) + P27_PAIRS * 27) +
                       sizeof(double) * (2 * P27_PAIRS + 2 * OC_NWMAX) + sizeof(int) * (P27_PAIRS + OC_NWMAX + 2);
    if (gy) {
        cplx<double>* Y = const_cast<cplx<double>*>(Wk) + 9 * (size_t)nk * nw * nw;
        hipLaunchKernelGGL((k_build_y<T, KIND == 0 ? 0 : 1>), dim3(nk), dim3(256), 0, ctx->stream, nw, nch, *params,
                           E, (const cplx<T>*)X, Y);
        WBMI_LAUNCH_CHECK(ctx);
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_pair27<T, KIND, true>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_pair27<T, KIND, true>), g, dim3(OC_THREADS), lds, ctx->stream, nk, *params, nw, nch, E,
                           (const cplx<T>*)X, Wk, (const cplx<double>*)Y, out);
    } else {
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_pair27<T, KIND, false>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_pair27<T, KIND, false>), g, dim3(OC_THREADS), lds, ctx->stream, nk, *params, nw, nch,
                           E, (const cplx<T>*)X, Wk, (const cplx<double>*)nullptr, out);
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// arguments shared by the three pair-walk calculators
static int check_pair27(wbmi_ctx* ctx, const char* what, int nk, int nw, int nch, const double* E, const void* X,
                        const wbmi_optcond_params* p, void* work, size_t work_bytes, double* out, bool need_w) {
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E || !X || !p || !out || !p->Efermi || !p->omega || p->nEF <= 0 ||
        p->nomega <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "%s: bad arguments", what);
    if (nw > OC_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "%s: nw <= %d", what, OC_NWMAX);
    if (!(p->eta > 0) || p->kBT < 0) return wbmi_fail(ctx, WBMI_ERR_ARG, "%s: smearing width must be positive", what);
    if (!work || work_bytes < wbmi_spin_hall_work_bytes(nk, nw))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "%s: workspace of %zu bytes needed", what, wbmi_spin_hall_work_bytes(nk, nw));
    auto bad = [&](int c) { return c < 0 || c >= nch; };
    for (int c = 0; c < 3; ++c)
        if (bad(p->ch[WBMI_ROLE_V + c]) || (p->external_terms && bad(p->ch[WBMI_ROLE_A + c])))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "%s: missing V or A channel", what);
    if (need_w)
        for (int c = 0; c < 9; ++c)
            if (bad(p->ch[WBMI_ROLE_W + c]) || (p->external_terms && bad(p->ch[WBMI_ROLE_AD + c])))
                return wbmi_fail(ctx, WBMI_ERR_ARG, "%s: missing Ham'' or AA' channel", what);
    return WBMI_OK;
}

extern "C" int wbmi_shift_current(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                  int prec, const wbmi_optcond_params* params, double sc_eta, void* work_dev,
                                  size_t work_bytes, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    int rc = check_pair27(ctx, "wbmi_shift_current", nk, nw, nch, E_dev, Xbar_dev, params, work_dev, work_bytes,
                          out_dev, true);
    if (rc) return rc;
    if (!(sc_eta > 0)) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_shift_current: sc_eta must be positive");
    cplx<double>* Wk = (cplx<double>*)work_dev;
    if (prec == WBMI_F64) {
        hipLaunchKernelGGL(k_shift_ag<double>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, sc_eta,
                           E_dev, (const cplx<double>*)Xbar_dev, Wk);
        WBMI_LAUNCH_CHECK(ctx);
        return launch_pair27<double, 1>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, Wk, out_dev);
    }
    if (prec == WBMI_F32) {
        hipLaunchKernelGGL(k_shift_ag<float>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, sc_eta,
                           E_dev, (const cplx<float>*)Xbar_dev, Wk);
        WBMI_LAUNCH_CHECK(ctx);
        return launch_pair27<float, 1>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, Wk, out_dev);
    }
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_shift_current: bad precision");
}

extern "C" int wbmi_injection_current(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                      const void* Xbar_dev, int prec, const wbmi_optcond_params* params,
                                      void* work_dev, size_t work_bytes, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    int rc = check_pair27(ctx, "wbmi_injection_current", nk, nw, nch, E_dev, Xbar_dev, params, work_dev, work_bytes,
                          out_dev, false);
    if (rc) return rc;
    cplx<double>* Wk = (cplx<double>*)work_dev;
    if (prec == WBMI_F64) {
        hipLaunchKernelGGL(k_inject_x<double>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, E_dev,
                           (const cplx<double>*)Xbar_dev, Wk);
        WBMI_LAUNCH_CHECK(ctx);
        return launch_pair27<double, 2>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, Wk, out_dev);
    }
    if (prec == WBMI_F32) {
        hipLaunchKernelGGL(k_inject_x<float>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, E_dev,
                           (const cplx<float>*)Xbar_dev, Wk);
        WBMI_LAUNCH_CHECK(ctx);
        return launch_pair27<float, 2>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, Wk, out_dev);
    }
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_injection_current: bad precision");
}

extern "C" size_t wbmi_spin_hall_work_bytes(int nk, int nw) {
    // X^x per k (9 nw^2), then Y per k (3 nw^2) when it does not fit in LDS
    return sizeof(cplx<double>) * (nw > OC_LDS_NW ? 12 : 9) * (size_t)nk * nw * nw;
}

template <typename T>
static int launch_shc(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E, const void* X,
                      const wbmi_shc_params* params, cplx<double>* J, double* out) {
    hipLaunchKernelGGL(k_spin_current<T>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, E,
                       (const cplx<T>*)X, J);
    WBMI_LAUNCH_CHECK(ctx);
    return launch_pair27<T, 0>(ctx, nk, nw, nch, E, X, &params->base, J, out);
}

extern "C" int wbmi_spin_hall(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                              int prec, const wbmi_shc_params* params, void* work_dev, size_t work_bytes,
                              double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !params || !out_dev || !params->base.Efermi ||
        !params->base.omega || params->base.nEF <= 0 || params->base.nomega <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: bad arguments");
    if (nw > OC_NWMAX) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_spin_hall: nw <= %d", OC_NWMAX);
    if (!(params->base.eta > 0) || params->base.kBT < 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: smearing width must be positive");
    if (!work_dev || work_bytes < wbmi_spin_hall_work_bytes(nk, nw))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: workspace of %zu bytes needed",
                         wbmi_spin_hall_work_bytes(nk, nw));
    const int kind = params->kind;
    if (kind != WBMI_SHC_RYOO && kind != WBMI_SHC_QIAO && kind != WBMI_SHC_SIMPLE)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: bad spin current kind %d", kind);
    if (kind != WBMI_SHC_SIMPLE && !params->base.external_terms)
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_spin_hall: the Ryoo and Qiao spin currents need external terms");
    auto bad = [&](int c) { return c < 0 || c >= nch; };
    for (int c = 0; c < 3; ++c) {
        if (bad(params->base.ch[WBMI_ROLE_V + c]) || bad(params->ch_S[c]) ||
            (params->base.external_terms && bad(params->base.ch[WBMI_ROLE_A + c])))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: missing V, A or SS channel");
    }
    for (int c = 0; c < 9; ++c) {
        if (kind == WBMI_SHC_RYOO && (bad(params->ch_SA[c]) || bad(params->ch_SHA[c])))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: missing SA or SHA channel");
        if (kind == WBMI_SHC_QIAO && (bad(params->ch_SR[c]) || bad(params->ch_SHR[c]) || (c < 3 && bad(params->ch_SH[c]))))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: missing SH, SR or SHR channel");
    }
    if (prec == WBMI_F64)
        return launch_shc<double>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, (cplx<double>*)work_dev, out_dev);
    if (prec == WBMI_F32)
        return launch_shc<float>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, (cplx<double>*)work_dev, out_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_spin_hall: bad precision");
}

// Spin Berry curvature per band (tabulate.py SpinBerry, formula/covariant.py:759-789):
//   c_m^{abs} = -2 Im sum_{l outside m's group} J^{as}_ml / (E_m - E_l) (D^b_lm - i Abar^b_lm),
// D^b_lm = -V^b_lm / (E_l - E_m) (both inverses 0 below 1e-7; Abar only with external terms),
// out[k][n][a*9+b*3+s] = the mean of c over n's group (per_band: c_n itself, the band's share
// of its group's trace).  One workgroup per k: W = D - i Abar in LDS, one thread per (m, abs)
// for c, then the group means.  J from k_spin_current.
template <typename T>
__global__ void __launch_bounds__(OC_THREADS) k_tab_spin_berry(int nk, int nw, int nch, wbmi_optcond_params p,
                                                               const double* __restrict__ Eall,
                                                               const cplx<T>* __restrict__ Xall,
                                                               const cplx<double>* __restrict__ J,
                                                               double* __restrict__ out, int per_band) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int k = blockIdx.x, tid = threadIdx.x, nw2 = nw * nw;
    cplx<double>* sW = (cplx<double>*)smem;  // [3][nw][nw]
    double* sC = (double*)(sW + 3 * nw2);    // [nw][27]
    double* sE = sC + nw * 27;               // [OC_NWMAX]
    double* sGm = sE + OC_NWMAX;             // [OC_NWMAX]
    int* sGb = (int*)(sGm + OC_NWMAX);       // [OC_NWMAX + 1]
    int* sG = sGb + OC_NWMAX + 1;            // band -> group [OC_NWMAX]
    for (int i = tid; i < nw; i += OC_THREADS) sE[i] = Eall[(int64_t)k * nw + i];
    __syncthreads();
    if (tid == 0) {
        const int ng = band_groups(sE, nw, p.degen_thresh, p.degen_Kramers, sGb, sGm);
        for (int i = 0; i < nw; ++i) sG[i] = -1;  // bands past the last Kramers border: no group
        for (int g = 0; g < ng; ++g)
            for (int i = sGb[g]; i < sGb[g + 1]; ++i) sG[i] = g;
    }
    const cplx<T>* Xk = Xall + (int64_t)k * nch * nw2;
    for (int t = tid; t < 3 * nw2; t += OC_THREADS) {
        const int c = t / nw2, l = (t % nw2) / nw, m = t % nw;
        const double de = sE[l] - sE[m];
        const double inv = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
        const cplx<T> v = Xk[(int64_t)p.ch[WBMI_ROLE_V + c] * nw2 + l * nw + m];
        cplx<double> w = mk(-inv * (double)v.re, -inv * (double)v.im);
        if (p.external_terms) {
            const cplx<T> x = Xk[(int64_t)p.ch[WBMI_ROLE_A + c] * nw2 + l * nw + m];
            w.re += (double)x.im;  // - i Abar
            w.im -= (double)x.re;
        }
        sW[t] = w;
    }
    __syncthreads();
    const cplx<double>* Jk = J + (int64_t)k * 9 * nw2;
    for (int t = tid; t < nw * 27; t += OC_THREADS) {
        const int m = t / 27, r = t % 27, a = r / 9, b = (r / 3) % 3, s = r % 3;
        const cplx<double>* Jm = Jk + (int64_t)(a * 3 + s) * nw2 + m * nw;
        const cplx<double>* Wb = sW + b * nw2;
        const int gm = sG[m];
        double sum = 0.0;
        for (int l = 0; l < nw && gm >= 0; ++l) {
            if (sG[l] == gm) continue;
            const double de = sE[m] - sE[l];
            const double inv = fabs(de) < 1e-7 ? 0.0 : 1.0 / de;
            const cplx<double> j = Jm[l], w = Wb[l * nw + m];
            sum += inv * (j.re * w.im + j.im * w.re);
        }
        sC[t] = -2.0 * sum;
    }
    __syncthreads();
    for (int t = tid; t < nw * 27; t += OC_THREADS) {
        const int n = t / 27, r = t % 27, g = sG[n];
        double s = 0.0;
        if (per_band) {
            s = sC[t];
        } else if (g >= 0) {
            for (int m = sGb[g]; m < sGb[g + 1]; ++m) s += sC[m * 27 + r];
            s /= sGb[g + 1] - sGb[g];
        }
        out[((int64_t)k * nw + n) * 27 + r] = s;
    }
}

template <typename T>
static int launch_spin_berry(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E, const void* X,
                             const wbmi_shc_params* params, cplx<double>* J, double* out, int per_band) {
    hipLaunchKernelGGL(k_spin_current<T>, dim3(nk, 9), dim3(256), 0, ctx->stream, nk, nw, nch, *params, E,
                       (const cplx<T>*)X, J);
    WBMI_LAUNCH_CHECK(ctx);
    const size_t lds = sizeof(cplx<double>) * 3 * nw * nw + sizeof(double) * (27 * nw + 2 * OC_NWMAX) +
                       sizeof(int) * (2 * OC_NWMAX + 1);
    hipLaunchKernelGGL(k_tab_spin_berry<T>, dim3(nk), dim3(OC_THREADS), lds, ctx->stream, nk, nw, nch, params->base,
                       E, (const cplx<T>*)X, (const cplx<double>*)J, out, per_band);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

static int spin_berry(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev, int prec,
                      const wbmi_shc_params* params, void* work_dev, size_t work_bytes, double* out_dev, int per_band) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nch <= 0 || !E_dev || !Xbar_dev || !params || !out_dev)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: bad arguments");
    if (nw > OC_LDS_NW) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_tabulate_spin_berry: nw <= %d", OC_LDS_NW);
    if (!work_dev || work_bytes < wbmi_spin_hall_work_bytes(nk, nw))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: workspace of %zu bytes needed",
                         wbmi_spin_hall_work_bytes(nk, nw));
    const int kind = params->kind;
    if (kind != WBMI_SHC_RYOO && kind != WBMI_SHC_QIAO && kind != WBMI_SHC_SIMPLE)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: bad spin current kind %d", kind);
    if (kind != WBMI_SHC_SIMPLE && !params->base.external_terms)
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED,
                         "wbmi_tabulate_spin_berry: the Ryoo and Qiao spin currents need external terms");
    auto bad = [&](int c) { return c < 0 || c >= nch; };
    for (int c = 0; c < 3; ++c) {
        if (bad(params->base.ch[WBMI_ROLE_V + c]) || bad(params->ch_S[c]) ||
            (params->base.external_terms && bad(params->base.ch[WBMI_ROLE_A + c])))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: missing V, A or SS channel");
    }
    for (int c = 0; c < 9; ++c) {
        if (kind == WBMI_SHC_RYOO && (bad(params->ch_SA[c]) || bad(params->ch_SHA[c])))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: missing SA or SHA channel");
        if (kind == WBMI_SHC_QIAO && (bad(params->ch_SR[c]) || bad(params->ch_SHR[c]) || (c < 3 && bad(params->ch_SH[c]))))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: missing SH, SR or SHR channel");
    }
    if (prec == WBMI_F64)
        return launch_spin_berry<double>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, (cplx<double>*)work_dev, out_dev,
                                         per_band);
    if (prec == WBMI_F32)
        return launch_spin_berry<float>(ctx, nk, nw, nch, E_dev, Xbar_dev, params, (cplx<double>*)work_dev, out_dev,
                                        per_band);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate_spin_berry: bad precision");
}

extern "C" int wbmi_tabulate_spin_berry(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                        const void* Xbar_dev, int prec, const wbmi_shc_params* params,
                                        void* work_dev, size_t work_bytes, double* out_dev) {
    return spin_berry(ctx, nk, nw, nch, E_dev, Xbar_dev, prec, params, work_dev, work_bytes, out_dev, 0);
}

extern "C" int wbmi_spin_berry_bands(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                                     const void* Xbar_dev, int prec, const wbmi_shc_params* params, void* work_dev,
                                     size_t work_bytes, double* out_dev) {
    return spin_berry(ctx, nk, nw, nch, E_dev, Xbar_dev, prec, params, work_dev, work_bytes, out_dev, 1);
}
