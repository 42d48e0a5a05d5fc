// Entry points of the covariant formula traces + Fermi-level scan of libwbmi
// (calculators/static.py:60-169) and the histogram finalisation.  The kernels
// are in scan_kernels.h; scan_{f64,f32}.hip, scan_tetra_{f64,f32}.hip and
// scan_tab_{f64,f32}.hip instantiate them.
#include "scan_kernels.h"

static int check_roles(wbmi_ctx* ctx, const wbmi_scan_params* p, int nch) {
    auto need = [&](int base, int n) -> int {
        for (int i = 0; i < n; ++i)
            if (p->ch[base + i] < 0 || p->ch[base + i] >= nch)
                return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: channel role %d missing or out of range",
                                 base + i);
        return WBMI_OK;
    };
    int rc = WBMI_OK;
    bool ex = p->external_terms;
    switch (p->formula) {
        case WBMI_FORMULA_IDENTITY: break;
        case WBMI_FORMULA_OMEGA:
        case WBMI_FORMULA_VELOMEGA:
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_O, 3)))) return rc;
            break;
        case WBMI_FORMULA_DEROMEGA:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_W, 9))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_AD, 9)) || (rc = need(WBMI_ROLE_O, 3)) ||
                       (rc = need(WBMI_ROLE_OD, 9))))
                return rc;
            break;
        case WBMI_FORMULA_DERMORB:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_W, 9))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_AD, 9)) || (rc = need(WBMI_ROLE_O, 3)) ||
                       (rc = need(WBMI_ROLE_OD, 9)) || (rc = need(WBMI_ROLE_B, 3)) || (rc = need(WBMI_ROLE_BD, 9)) ||
                       (rc = need(WBMI_ROLE_C, 3)) || (rc = need(WBMI_ROLE_CD, 9))))
                return rc;
            break;
        case WBMI_FORMULA_DERSPIN:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_S, 3)) || (rc = need(WBMI_ROLE_SD, 9))) return rc;
            break;
        case WBMI_FORMULA_VELSPIN:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_S, 3))) return rc;
            break;
        case WBMI_FORMULA_SPIN:
            if ((rc = need(WBMI_ROLE_S, 3))) return rc;
            break;
        case WBMI_FORMULA_INVMASS:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_W, 9))) return rc;
            break;
        case WBMI_FORMULA_VELVEL:
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            break;
        case WBMI_FORMULA_VELVELVEL:
        case WBMI_FORMULA_MASSVEL:
            if (p->comp_slice < 0 || p->comp_slice > 2)
                return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: comp_slice must be 0, 1 or 2");
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            if (p->formula == WBMI_FORMULA_MASSVEL && (rc = need(WBMI_ROLE_W, 9))) return rc;
            break;
        case WBMI_FORMULA_HALL_SURF:
        case WBMI_FORMULA_HALL_SEA:
            if ((rc = need(WBMI_ROLE_V, 3)) || (rc = need(WBMI_ROLE_W, 9))) return rc;
            break;
        case WBMI_FORMULA_BANDVALUES:
            if (!((nch == 9 && p->comp_slice == 0) || (nch == 27 && p->comp_slice >= 0 && p->comp_slice <= 2)))
                return wbmi_fail(ctx, WBMI_ERR_ARG,
                                 "wbmi_static_scan: band values need nch 9 (comp_slice 0) or 27 (comp_slice 0..2)");
            break;
        case WBMI_FORMULA_OMEGASPIN:
        case WBMI_FORMULA_OMEGAOMEGA:
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_O, 3)))) return rc;
            if (p->formula == WBMI_FORMULA_OMEGASPIN && (rc = need(WBMI_ROLE_S, 3))) return rc;
            break;
        case WBMI_FORMULA_OMEGAHPLUS:
        case WBMI_FORMULA_MORB_HPM:
        case WBMI_FORMULA_VELHPLUS:
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_O, 3)) || (rc = need(WBMI_ROLE_B, 3)) ||
                       (rc = need(WBMI_ROLE_C, 3))))
                return rc;
            break;
        default:
            return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_static_scan: formula %d is not available", p->formula);
    }
    return WBMI_OK;
}

extern "C" int wbmi_static_scan(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                int prec, const wbmi_scan_params* params, double weight, double* hist_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > SC_NWMAX || nch <= 0 || !E_dev || !params || !hist_dev ||
        (params->formula != WBMI_FORMULA_IDENTITY && !Xbar_dev))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: bad arguments (nw <= %d)", SC_NWMAX);
    if (params->nEF_extra <= 0 || !(params->dEF > 0) || params->fder < 0 || params->fder > 3)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: bad Fermi grid");
    if (params->band_weight && params->fder == 0)
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_static_scan: selection of bands for the Fermi sea is not implemented");
    int rc = check_roles(ctx, params, nch);
    if (rc) return rc;
    ScanArgs a;
    a.p = *params;
    a.nw = nw;
    a.nch = nch;
    a.stage_va = nw <= SC_LDS_NW;
    a.weight = weight;
    a.tetra = 0;
    a.nEF = 0;
    a.Ec = nullptr;
    a.ec_stride = 0;
    a.Ef = nullptr;
    a.hist_stride = 0;
    if (prec == WBMI_F64) return scan_dispatch_f64(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    if (prec == WBMI_F32) return scan_dispatch_f32(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: bad precision");
}

static int scan_tetra(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                      const double* Ecorners_dev, const void* Xbar_dev, int prec, const wbmi_scan_params* params,
                      const double* Efermi_dev, int nEF, double weight, double* hist_dev, bool kres);

extern "C" int wbmi_static_scan_tetra(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                                      const double* Ecorners_dev, const void* Xbar_dev, int prec,
                                      const wbmi_scan_params* params, const double* Efermi_dev, int nEF,
                                      double weight, double* hist_dev) {
    return scan_tetra(ctx, nk, nw, nch, E_dev, ncorner, Ecorners_dev, Xbar_dev, prec, params, Efermi_dev, nEF, weight,
                      hist_dev, false);
}

extern "C" int wbmi_static_scan_tetra_k(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                                        const double* Ecorners_dev, const void* Xbar_dev, int prec,
                                        const wbmi_scan_params* params, const double* Efermi_dev, int nEF,
                                        double weight, double* hist_dev) {
    return scan_tetra(ctx, nk, nw, nch, E_dev, ncorner, Ecorners_dev, Xbar_dev, prec, params, Efermi_dev, nEF, weight,
                      hist_dev, true);
}

static int ncomp_of_formula(int f);

static int scan_tetra(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                      const double* Ecorners_dev, const void* Xbar_dev, int prec, const wbmi_scan_params* params,
                      const double* Efermi_dev, int nEF, double weight, double* hist_dev, bool kres) {
    if (ncorner != 8 && ncorner != 4)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_tetra: ncorner must be 8 or 4");
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > SC_NWMAX || nch <= 0 || !E_dev || !Ecorners_dev || !params || !hist_dev ||
        !Efermi_dev || nEF <= 0 || (params->formula != WBMI_FORMULA_IDENTITY && !Xbar_dev))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_tetra: bad arguments (nw <= %d)", SC_NWMAX);
    if (params->fder < -1 || params->fder > 3 || !(params->EFmax >= params->EFmin))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_tetra: bad fder or Fermi range");
    if (params->band_weight && (params->fder == 0 || params->fder == -1))
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED,
                         "wbmi_static_scan_tetra: selection of bands for the Fermi sea (or the hole-like group) is not implemented");
    int rc = check_roles(ctx, params, nch);
    if (rc) return rc;
    ScanArgs a;
    a.p = *params;
    a.nw = nw;
    a.nch = nch;
    a.stage_va = nw <= SC_LDS_NW;
    a.weight = weight;
    a.tetra = ncorner;
    a.nEF = nEF;
    a.Ec = Ecorners_dev;
    a.ec_stride = (int64_t)nk * nw;
    a.Ef = Efermi_dev;
    a.hist_stride = kres ? 2 * (int64_t)nEF * ncomp_of_formula(params->formula) : 0;
    if (prec == WBMI_F64) return scan_dispatch_tetra_f64(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    if (prec == WBMI_F32) return scan_dispatch_tetra_f32(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_tetra: bad precision");
}

static int ncomp_of_formula(int f) {
    switch (f) {
        case WBMI_FORMULA_IDENTITY: return 1;
        case WBMI_FORMULA_OMEGA:
        case WBMI_FORMULA_MORB_HPM:
        case WBMI_FORMULA_SPIN: return 3;
        default: return 9;
    }
}

extern "C" int wbmi_static_scan_k(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                  int prec, const wbmi_scan_params* params, double weight, double* hist_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > SC_NWMAX || nch <= 0 || !E_dev || !params || !hist_dev ||
        (params->formula != WBMI_FORMULA_IDENTITY && !Xbar_dev))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_k: bad arguments (nw <= %d)", SC_NWMAX);
    if (params->nEF_extra <= 0 || !(params->dEF > 0) || params->fder < 0 || params->fder > 3)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_k: bad Fermi grid");
    if (params->band_weight && params->fder == 0)
        return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_static_scan_k: selection of bands for the Fermi sea is not implemented");
    int rc = check_roles(ctx, params, nch);
    if (rc) return rc;
    ScanArgs a;
    a.p = *params;
    a.nw = nw;
    a.nch = nch;
    a.stage_va = nw <= SC_LDS_NW;
    a.weight = weight;
    a.tetra = 0;
    a.nEF = 0;
    a.Ec = nullptr;
    a.ec_stride = 0;
    a.Ef = nullptr;
    a.hist_stride = (int64_t)params->nEF_extra * ncomp_of_formula(params->formula);
    if (prec == WBMI_F64) return scan_dispatch_f64(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    if (prec == WBMI_F32) return scan_dispatch_f32(ctx, nk, a, E_dev, Xbar_dev, hist_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan_k: bad precision");
}

extern "C" int wbmi_tabulate(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                             int prec, const wbmi_scan_params* params, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > SC_NWMAX || nch <= 0 || !E_dev || !params || !out_dev ||
        (params->formula != WBMI_FORMULA_IDENTITY && !Xbar_dev))
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate: bad arguments (nw <= %d)", SC_NWMAX);
    if (params->band_weight) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate: band_weight must be NULL");
    int rc = check_roles(ctx, params, nch);
    if (rc) return rc;
    ScanArgs a;
    a.p = *params;
    a.p.fder = 1;  // no Fermi-sea group
    a.nw = nw;
    a.nch = nch;
    a.stage_va = nw <= SC_LDS_NW;
    a.weight = 1.0;
    a.tetra = 0;
    a.nEF = 0;
    a.Ec = nullptr;
    a.ec_stride = 0;
    a.Ef = nullptr;
    a.hist_stride = 0;
    if (prec == WBMI_F64) return scan_dispatch_tab_f64(ctx, nk, a, E_dev, Xbar_dev, out_dev);
    if (prec == WBMI_F32) return scan_dispatch_tab_f32(ctx, nk, a, E_dev, Xbar_dev, out_dev);
    return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_tabulate: bad precision");
}

// ---------------------------------------------------------------------------
// histogram -> result (static.py:137-155)
// ---------------------------------------------------------------------------

#define FIN_THREADS 256
// One workgroup: per component, an in-place block prefix sum over the Fermi
// levels (each thread owns a contiguous run), then the fder stencil.
__global__ void __launch_bounds__(FIN_THREADS) k_finalize(double* __restrict__ hist, int nEF, int ncomp, int fder,
                                                         double dEF, double scale, double* __restrict__ out,
                                                         int64_t hstride, int64_t ostride) {
    __shared__ double part[FIN_THREADS];
    // one histogram (k-resolved: one per k-point, blockIdx.x) and one component (blockIdx.y)
    // per workgroup
    hist += (int64_t)blockIdx.x * hstride;
    out += (int64_t)blockIdx.x * ostride;
    const int tid = threadIdx.x;
    const int L = (nEF + FIN_THREADS - 1) / FIN_THREADS;
    const int i0 = tid * L < nEF ? tid * L : nEF, i1 = i0 + L < nEF ? i0 + L : nEF;
    {
        const int c = blockIdx.y;
        double sum = 0.0;
        for (int i = i0; i < i1; ++i) sum += hist[(int64_t)i * ncomp + c];
        part[tid] = sum;
        __syncthreads();
        // inclusive scan of the per-thread sums (Hillis-Steele)
        for (int off = 1; off < FIN_THREADS; off <<= 1) {
            const double v = tid >= off ? part[tid - off] : 0.0;
            __syncthreads();
            part[tid] += v;
            __syncthreads();
        }
        double run = tid > 0 ? part[tid - 1] : 0.0;
        for (int i = i0; i < i1; ++i) {
            run += hist[(int64_t)i * ncomp + c];
            hist[(int64_t)i * ncomp + c] = run;
        }
        __syncthreads();
    }
    const int extra = fder == 0 ? 0 : (fder == 3 ? 2 : 1);
    const int nout = nEF - 2 * extra;
    const int c = blockIdx.y;
    for (int i = tid; i < nout; i += FIN_THREADS) {
        const int t = i * ncomp + c;
        const double* r = hist + c;
        double v;
        if (fder == 0) v = r[(int64_t)i * ncomp];
        else if (fder == 1) v = (r[(int64_t)(i + 2) * ncomp] - r[(int64_t)i * ncomp]) / (2 * dEF);
        else if (fder == 2)
            v = (r[(int64_t)(i + 2) * ncomp] + r[(int64_t)i * ncomp] - 2 * r[(int64_t)(i + 1) * ncomp]) / (dEF * dEF);
        else
            v = (r[(int64_t)(i + 4) * ncomp] - r[(int64_t)i * ncomp] -
                 2 * (r[(int64_t)(i + 3) * ncomp] - r[(int64_t)(i + 1) * ncomp])) /
                (2 * dEF * dEF * dEF);
        out[t] = v * scale;
    }
}

// out += scale * add (the directly accumulated rows of the tetrahedron histogram); n elements
// in batches of `per`, batch b of `add` at b * hstride
__global__ void k_add_scaled(double* __restrict__ out, const double* __restrict__ add, double scale, int64_t n,
                             int64_t per, int64_t hstride) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += scale * add[(i / per) * hstride + i % per];
}

extern "C" int wbmi_finalize_tetra(wbmi_ctx* ctx, double* hist_dev, int nEF, int ncomp, double scale,
                                   double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (!hist_dev || !out_dev || ncomp <= 0 || nEF <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize_tetra: bad arguments");
    hipLaunchKernelGGL(k_finalize, dim3(1, ncomp), dim3(FIN_THREADS), 0, ctx->stream, hist_dev, nEF, ncomp, 0, 1.0, scale,
                       out_dev, (int64_t)0, (int64_t)0);
    WBMI_LAUNCH_CHECK(ctx);
    const int64_t n = (int64_t)nEF * ncomp;
    hipLaunchKernelGGL(k_add_scaled, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, out_dev,
                       hist_dev + n, scale, n, n, 2 * n);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" int wbmi_finalize_tetra_k(wbmi_ctx* ctx, double* hist_dev, int nk, int nEF, int ncomp, double scale,
                                     double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (!hist_dev || !out_dev || nk <= 0 || ncomp <= 0 || nEF <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize_tetra_k: bad arguments");
    const int64_t per = (int64_t)nEF * ncomp;
    hipLaunchKernelGGL(k_finalize, dim3(nk, ncomp), dim3(FIN_THREADS), 0, ctx->stream, hist_dev, nEF, ncomp, 0, 1.0, scale,
                       out_dev, 2 * per, per);
    WBMI_LAUNCH_CHECK(ctx);
    const int64_t n = per * nk;
    hipLaunchKernelGGL(k_add_scaled, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, out_dev,
                       hist_dev + per, scale, n, per, 2 * per);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" int wbmi_finalize(wbmi_ctx* ctx, double* hist_dev, int nEF_extra, int ncomp, int fder, double dEF,
                             double scale, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    int extra = fder == 0 ? 0 : (fder == 3 ? 2 : 1);
    if (!hist_dev || !out_dev || ncomp <= 0 || fder < 0 || fder > 3 || nEF_extra - 2 * extra <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize: bad arguments");
    hipLaunchKernelGGL(k_finalize, dim3(1, ncomp), dim3(FIN_THREADS), 0, ctx->stream, hist_dev, nEF_extra, ncomp, fder, dEF,
                       scale, out_dev, (int64_t)0, (int64_t)0);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" int wbmi_finalize_k(wbmi_ctx* ctx, double* hist_dev, int nk, int nEF_extra, int ncomp, int fder, double dEF,
                               double scale, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    int extra = fder == 0 ? 0 : (fder == 3 ? 2 : 1);
    if (!hist_dev || !out_dev || nk <= 0 || ncomp <= 0 || fder < 0 || fder > 3 || nEF_extra - 2 * extra <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize_k: bad arguments");
    const int64_t hs = (int64_t)nEF_extra * ncomp, os = (int64_t)(nEF_extra - 2 * extra) * ncomp;
    hipLaunchKernelGGL(k_finalize, dim3(nk, ncomp), dim3(FIN_THREADS), 0, ctx->stream, hist_dev, nEF_extra, ncomp, fder, dEF,
                       scale, out_dev, hs, os);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
