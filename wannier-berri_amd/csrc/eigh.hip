// Batched Hermitian eigensolver of libwbmi: cyclic two-sided Jacobi.
//
// One 64-lane workgroup per k-point; the matrix A and the accumulated
// rotations V stay in LDS for the whole solve.  Each Jacobi step rotates
// nw/2 disjoint (p, q) pairs of the round-robin (circle) ordering at once:
// all column updates (A G and V G), a barrier, all row updates (G^+ A).
// The complex rotation is G = diag(1, e^{-i phi}) R diag(1, e^{i phi}) with
// a_pq = |a_pq| e^{i phi} and R the real symmetric Jacobi rotation
// (Golub & Van Loan, Matrix Computations, sec. 8.5.2).  Sweeps stop when a
// whole sweep applies no rotation.  Eigenvalues are returned ascending with
// eigenvectors as columns, as np.linalg.eigh does (data_K/data_K.py:213).
#include <cfloat>

#include "wbmi_common.h"

template <typename T>
struct Eps;
template <>
struct Eps<double> {
    static constexpr double v = 2.220446049250313e-16;
};
template <>
struct Eps<float> {
    static constexpr float v = 1.1920929e-07f;
};

// pair i of round r in the circle ordering of n (even) players
__device__ inline void rr_pair(int n, int r, int i, int& p, int& q) {
    int m = n - 1;
    p = (i == 0) ? 0 : ((i + r - 1) % m) + 1;
    q = ((n - 1 - i + r - 1) % m) + 1;
}

template <typename T>
__global__ void __launch_bounds__(64) k_eigh_jacobi(int nw, const cplx<T>* __restrict__ H, int64_t Hstride,
                                                    double* __restrict__ Eout, cplx<T>* __restrict__ Uout,
                                                    int max_sweeps) {
    extern __shared__ char smem[];
    const int n = nw + (nw & 1);  // padded to even; the dummy index is decoupled
    const int ld = n + 1;         // LDS row stride (one element of padding against bank conflicts)
    cplx<T>* A = (cplx<T>*)smem;
    cplx<T>* V = A + n * ld;
    T* prm = (T*)(V + n * ld);  // per pair: c, s*cos(phi), s*sin(phi)
    int* flag = (int*)(prm + 3 * (n / 2));
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    const cplx<T>* Hk = H + k * Hstride;

    for (int t = tid; t < n * n; t += 64) {
        int i = t / n, j = t % n;
        cplx<T> a = mk((T)0, (T)0);
        if (i < nw && j < nw) {
            cplx<T> x = Hk[i * nw + j], y = Hk[j * nw + i];
            a = mk((T)0.5 * (x.re + y.re), (T)0.5 * (x.im - y.im));
        }
        A[i * ld + j] = a;
        V[i * ld + j] = mk((T)(i == j ? 1 : 0), (T)0);
    }
    __syncthreads();

    const T eps = Eps<T>::v;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (tid == 0) *flag = 0;
        __syncthreads();
        for (int r = 0; r < n - 1; ++r) {
            if (tid < n / 2) {
                int p, q;
                rr_pair(n, r, tid, p, q);
                cplx<T> apq = A[p * ld + q];
                T g = sqrt(apq.re * apq.re + apq.im * apq.im);
                T app = A[p * ld + p].re, aqq = A[q * ld + q].re;
                T c = 1, sc = 0, ss = 0;
                T scale = fabs(app) > fabs(aqq) ? fabs(app) : fabs(aqq);
                if (g > (T)0.5 * eps * scale && g > (T)(sizeof(T) == 8 ? 1e-300 : 1e-36)) {
                    T tau = (aqq - app) / (2 * g);
                    T t = (tau >= 0 ? (T)1 : (T)-1) / (fabs(tau) + sqrt((T)1 + tau * tau));
                    c = (T)1 / sqrt((T)1 + t * t);
                    T s = t * c;
                    sc = s * apq.re / g;
                    ss = s * apq.im / g;
                    *flag = 1;
                }
                prm[3 * tid] = c;
                prm[3 * tid + 1] = sc;
                prm[3 * tid + 2] = ss;
            }
            __syncthreads();
            // columns of A and V:  col_p <- c col_p - s e^{-i phi} col_q ; col_q <- s e^{i phi} col_p + c col_q
            for (int t = tid; t < n * (n / 2); t += 64) {
                int pr = t / n, i = t % n;
                int p, q;
                rr_pair(n, r, pr, p, q);
                T c = prm[3 * pr];
                cplx<T> se = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                if (se.re == 0 && se.im == 0) continue;
                cplx<T> ap = A[i * ld + p], aq = A[i * ld + q];
                A[i * ld + p] = c * ap - mulc(aq, se);
                A[i * ld + q] = se * ap + c * aq;
                cplx<T> vp = V[i * ld + p], vq = V[i * ld + q];
                V[i * ld + p] = c * vp - mulc(vq, se);
                V[i * ld + q] = se * vp + c * vq;
            }
            __syncthreads();
            // rows of A:  row_p <- c row_p - s e^{i phi} row_q ; row_q <- s e^{-i phi} row_p + c row_q
            for (int t = tid; t < n * (n / 2); t += 64) {
                int pr = t / n, j = t % n;
                int p, q;
                rr_pair(n, r, pr, p, q);
                T c = prm[3 * pr];
                cplx<T> se = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                if (se.re == 0 && se.im == 0) continue;
                cplx<T> ap = A[p * ld + j], aq = A[q * ld + j];
                A[p * ld + j] = c * ap - se * aq;
                A[q * ld + j] = mulc(ap, se) + c * aq;
            }
            __syncthreads();
        }
        if (*flag == 0) break;
        __syncthreads();
    }
    // ascending order: rank of each eigenvalue (ties broken by index)
    for (int i = tid; i < nw; i += 64) {
        T ei = A[i * ld + i].re;
        int rank = 0;
        for (int j = 0; j < nw; ++j) {
            T ej = A[j * ld + j].re;
            rank += (ej < ei) || (ej == ei && j < i);
        }
        Eout[k * nw + rank] = (double)ei;
        for (int m = 0; m < nw; ++m) Uout[(k * nw + m) * nw + rank] = V[m * ld + i];
    }
}

extern "C" int wbmi_eigh(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride, double* E_dev,
                         void* U_dev, int prec, int refine, const void* H64_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > 64 || !H_dev || !E_dev || !U_dev || Hstride < (int64_t)nw * nw)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: bad arguments (nw <= 64, Hstride >= nw*nw)");
    if (refine) return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_eigh: refinement is not available in this build");
    (void)H64_dev;
    int n = nw + (nw & 1);
    const int sweeps = 40;
    if (prec == WBMI_F64) {
        size_t lds = 2 * sizeof(cplx<double>) * n * (n + 1) + 3 * sizeof(double) * (n / 2) + 16;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi<double>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_eigh_jacobi<double>, dim3(nk), dim3(64), lds, ctx->stream, nw,
                           (const cplx<double>*)H_dev, Hstride, E_dev, (cplx<double>*)U_dev, sweeps);
    } else if (prec == WBMI_F32) {
        size_t lds = 2 * sizeof(cplx<float>) * n * (n + 1) + 3 * sizeof(float) * (n / 2) + 16;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi<float>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_eigh_jacobi<float>, dim3(nk), dim3(64), lds, ctx->stream, nw,
                           (const cplx<float>*)H_dev, Hstride, E_dev, (cplx<float>*)U_dev, sweeps);
    } else {
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: bad precision");
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
