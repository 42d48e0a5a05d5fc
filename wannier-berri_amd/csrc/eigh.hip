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

#include "mfma_c64.h"
#include "wbmi_common.h"

// update items per lane per batch (loads of a batch are issued together)
#define JB 8

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

// pair i of round r (0 <= i < n/2, 0 <= r < n-1) in the circle ordering of n
// (even) players: i + r - 1 < 2(n-1) and n - 2 - i + r < 2(n-1), so one
// conditional subtraction replaces the modulo.
__device__ inline void rr_pair(int n, int r, int i, int& p, int& q) {
    const int m = n - 1;
    int a = i + r - 1;
    a = a >= m ? a - m : a;
    p = (i == 0) ? 0 : a + 1;
    int b = n - 2 - i + r;
    b = b >= m ? b - m : b;
    q = b + 1;
}

// Items (pair, index) of a round, pair-major, dealt to the 64 lanes:
// lane t starts at item t and steps by 64 items.
struct ItemWalk {
    int pr, i, dp, di;
    __device__ ItemWalk(int n, int tid) : pr(tid / n), i(tid % n), dp(64 / n), di(64 % n) {}
    __device__ inline void next(int n) {
        pr += dp;
        i += di;
        if (i >= n) {
            i -= n;
            ++pr;
        }
    }
};

// T: arithmetic and LDS type; Tin: type of H; TO: type of the stored eigenvectors
template <typename T, typename Tin, typename TO = T>
__global__ void __launch_bounds__(64) k_eigh_jacobi(int nw, const cplx<Tin>* __restrict__ H, int64_t Hstride,
                                                    double* __restrict__ Eout, cplx<TO>* __restrict__ Uout,
                                                    int max_sweeps) {
    extern __shared__ char smem[];
    const int n = nw + (nw & 1);  // padded to even; the dummy index is decoupled
    const int ld = n + 1;         // LDS row stride (one element of padding against bank conflicts)
    const int np = n / 2;
    cplx<T>* A = (cplx<T>*)smem;
    cplx<T>* V = A + n * ld;
    T* prm = (T*)(V + n * ld);  // per pair: c, s*cos(phi), s*sin(phi)
    int* flag = (int*)(prm + 3 * np);
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    const cplx<Tin>* Hk = H + k * Hstride;

    for (int i = tid; i < n; i += 64)
        for (int j = 0; j < n; ++j) {
            cplx<T> a = mk((T)0, (T)0);
            if (i < nw && j < nw) {
                cplx<Tin> x = Hk[i * nw + j], y = Hk[j * nw + i];
                a = mk((T)(0.5 * (x.re + y.re)), (T)(0.5 * (x.im - y.im)));
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
            bool rot = false;
            if (tid < np) {
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
                    rot = true;
                    *flag = 1;
                }
                prm[3 * tid] = c;
                prm[3 * tid + 1] = sc;
                prm[3 * tid + 2] = ss;
            }
            // the workgroup is one wavefront: __any is uniform over it
            if (!__any(rot)) continue;
            __syncthreads();
            // columns of A and V:  col_p <- c col_p - s e^{-i phi} col_q ; col_q <- s e^{i phi} col_p + c col_q
            // Items of a phase touch disjoint elements: each lane loads JB items, then rotates and stores them.
            for (ItemWalk w(n, tid); w.pr < np;) {
                int ii[JB], pp[JB], qq[JB];
                bool ok[JB];
                T cc[JB];
                cplx<T> se[JB], ap[JB], aq[JB], vp[JB], vq[JB];
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    ok[u] = w.pr < np;
                    const int pr = ok[u] ? w.pr : 0;
                    rr_pair(n, r, pr, pp[u], qq[u]);
                    ii[u] = w.i;
                    w.next(n);
                    cc[u] = prm[3 * pr];
                    se[u] = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                    ap[u] = A[ii[u] * ld + pp[u]];
                    aq[u] = A[ii[u] * ld + qq[u]];
                    vp[u] = V[ii[u] * ld + pp[u]];
                    vq[u] = V[ii[u] * ld + qq[u]];
                }
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    if (!ok[u]) continue;
                    A[ii[u] * ld + pp[u]] = cc[u] * ap[u] - mulc(aq[u], se[u]);
                    A[ii[u] * ld + qq[u]] = se[u] * ap[u] + cc[u] * aq[u];
                    V[ii[u] * ld + pp[u]] = cc[u] * vp[u] - mulc(vq[u], se[u]);
                    V[ii[u] * ld + qq[u]] = se[u] * vp[u] + cc[u] * vq[u];
                }
            }
            __syncthreads();
            // rows of A:  row_p <- c row_p - s e^{i phi} row_q ; row_q <- s e^{-i phi} row_p + c row_q
            for (ItemWalk w(n, tid); w.pr < np;) {
                int jj[JB], pp[JB], qq[JB];
                bool ok[JB];
                T cc[JB];
                cplx<T> se[JB], ap[JB], aq[JB];
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    ok[u] = w.pr < np;
                    const int pr = ok[u] ? w.pr : 0;
                    rr_pair(n, r, pr, pp[u], qq[u]);
                    jj[u] = w.i;
                    w.next(n);
                    cc[u] = prm[3 * pr];
                    se[u] = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                    ap[u] = A[pp[u] * ld + jj[u]];
                    aq[u] = A[qq[u] * ld + jj[u]];
                }
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    if (!ok[u]) continue;
                    A[pp[u] * ld + jj[u]] = cc[u] * ap[u] - se[u] * aq[u];
                    A[qq[u] * ld + jj[u]] = mulc(ap[u], se[u]) + cc[u] * aq[u];
                }
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
        for (int m = 0; m < nw; ++m) {
            const cplx<T> v = V[m * ld + i];
            Uout[(k * nw + m) * nw + rank] = mk((TO)v.re, (TO)v.im);
        }
    }
}

// ---------------------------------------------------------------------------
// FP32 Jacobi for padded sizes n = NT in {16, 32, 64}: the rotations of
// k_eigh_jacobi, in the same order, with the size known at compile time.  A
// round's parameters are stored once per pair as {c, s cos phi, s sin phi,
// p | q << 16}, so an update item is one 16-byte parameter read plus its
// element loads and stores (four of each in the column phase, two in the row
// phase), with no per-item index arithmetic beyond two adds.  Lane t owns row
// (column phase) or column (row phase) t % NT of the pairs t / NT,
// t / NT + 64 / NT, ...
// ---------------------------------------------------------------------------
template <int NT>
__host__ __device__ constexpr size_t jacobi_c_lds() {
    return 2 * sizeof(cplx<float>) * NT * (NT + 1) + sizeof(float4) * (NT / 2) + 16;
}

// U0 (optional, NT <= 32): approximate eigenvectors (columns), e.g. those of
// the same k-point index in the neighbouring K-block.  The sweeps then run on
// A = U0^+ H U0 (both products on the FP64 matrix cores, stored in FP32) with
// the rotations accumulated onto V = U0: a nearly diagonal start needs about
// three active sweeps instead of six or seven.  Any unitary U0 gives the same
// eigen-decomposition; a poor one only costs sweeps.
// (no minimum waves per SIMD: with 4 the kernel keeps 116 VGPRs and no AGPRs but runs 15 %
// slower than at 2 waves per SIMD with 64 AGPRs, its LDS traffic being the limit)
template <int NT, typename Tin, typename TO>
__global__ void __launch_bounds__(64) k_eigh_jacobi_c(int nw, const cplx<Tin>* __restrict__ H, int64_t Hstride,
                                                      double* __restrict__ Eout, cplx<TO>* Uout,
                                                      int max_sweeps, const cplx<TO>* U0) {
    typedef float T;
    constexpr int n = NT, ld = NT + 1, np = NT / 2;
    constexpr int LPR = 64 / NT;          // pair stride of a lane
    constexpr int IPL = np / LPR;         // update items per lane and phase
    constexpr int B = IPL < 8 ? IPL : 8;  // items per load batch
    static_assert(NT % 2 == 0 && 64 % NT == 0 && IPL % B == 0, "NT must be 16, 32 or 64");
    extern __shared__ char smem[];
    cplx<T>* A = (cplx<T>*)smem;
    cplx<T>* V = A + n * ld;
    float4* prm = (float4*)(V + n * ld);  // 16-byte aligned: 16 n (n + 1) bytes precede it
    int* flag = (int*)(prm + np);
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    const cplx<Tin>* Hk = H + k * Hstride;

    bool warm = false;
    if constexpr (NT <= 32) warm = U0 != nullptr;
    if (!warm) {
        for (int i = tid; i < n; i += 64)
            for (int j = 0; j < n; ++j) {
                cplx<T> a = mk((T)0, (T)0);
                if (i < nw && j < nw) {
                    cplx<Tin> x = Hk[i * nw + j], y = Hk[j * nw + i];
                    a = mk((T)(0.5 * (x.re + y.re)), (T)(0.5 * (x.im - y.im)));
                }
                A[i * ld + j] = a;
                V[i * ld + j] = mk((T)(i == j ? 1 : 0), (T)0);
            }
        __syncthreads();
    } else if constexpr (NT <= 32) {
        // V = U0 (identity on the padding index of odd nw)
        const cplx<TO>* U0k = U0 + k * nw * nw;
        for (int t = tid; t < n * n; t += 64) {
            const int i = t / n, j = t % n;
            cplx<T> v = mk((T)(i == j ? 1 : 0), (T)0);
            if (i < nw && j < nw) {
                const cplx<TO> u = U0k[i * nw + j];
                v = mk((T)u.re, (T)u.im);
            }
            V[i * ld + j] = v;
        }
        __syncthreads();
        // T = H U0 into A, then A = U0^+ T (one 32 x 32 block per product)
        MfmaBlock C;
        C.zero();
        mfma_block(
            C, n, n, n, 0, 0,
            [&](int r, int m) {
                if (r >= nw || m >= nw) return mk(0.0, 0.0);
                const cplx<Tin> x = Hk[r * nw + m], y = Hk[m * nw + r];
                return mk(0.5 * ((double)x.re + (double)y.re), 0.5 * ((double)x.im - (double)y.im));
            },
            [&](int m, int c) { return mk((double)V[m * ld + c].re, (double)V[m * ld + c].im); });
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mfma_row(bi, v), c = mfma_col(bj);
                    if (r < n && c < n) A[r * ld + c] = mk((T)C.re[bi][bj][v], (T)C.im[bi][bj][v]);
                }
        __syncthreads();
        C.zero();
        mfma_block(
            C, n, n, n, 0, 0, [&](int r, int m) { return mk((double)V[m * ld + r].re, -(double)V[m * ld + r].im); },
            [&](int m, int c) { return mk((double)A[m * ld + c].re, (double)A[m * ld + c].im); });
        __syncthreads();  // every read of T is done
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mfma_row(bi, v), c = mfma_col(bj);
                    if (r < n && c < n) A[r * ld + c] = mk((T)C.re[bi][bj][v], (T)C.im[bi][bj][v]);
                }
        __syncthreads();
    }

    const int lr = tid % NT;
    const int pr0 = tid / NT;
    const T eps = Eps<T>::v;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (tid == 0) *flag = 0;
        __syncthreads();
        for (int r = 0; r < n - 1; ++r) {
            bool rot = false;
            if (tid < np) {
                int p, q;
                rr_pair(n, r, tid, p, q);
                cplx<T> apq = A[p * ld + q];
                T g = sqrt(apq.re * apq.re + apq.im * apq.im);
                T app = A[p * ld + p].re, aqq = A[q * ld + q].re;
                T c = 1, sc = 0, ss = 0;
                T scale = fabs(app) > fabs(aqq) ? fabs(app) : fabs(aqq);
                if (g > (T)0.5 * eps * scale && g > (T)1e-36) {
                    // hardware reciprocal and reciprocal square root (1 ulp): c^2 + s^2 = 1 to
                    // rounding as with correctly rounded division, in far fewer instructions
                    const T rg = __builtin_amdgcn_rcpf(g);
                    const T tau = (aqq - app) * (T)0.5 * rg;
                    const T ta = __builtin_amdgcn_rcpf(fabs(tau) + sqrt((T)1 + tau * tau));
                    const T t = tau >= 0 ? ta : -ta;
                    c = __builtin_amdgcn_rsqf((T)1 + t * t);
                    const T s = t * c;
                    sc = s * apq.re * rg;
                    ss = s * apq.im * rg;
                    rot = true;
                    *flag = 1;
                }
                prm[tid] = make_float4(c, sc, ss, __int_as_float(p | (q << 16)));
            }
            // the workgroup is one wavefront: __any is uniform over it
            if (!__any(rot)) continue;
            __syncthreads();
            // columns of A and V:  col_p <- c col_p - s e^{-i phi} col_q ; col_q <- s e^{i phi} col_p + c col_q
#pragma unroll
            for (int b0 = 0; b0 < IPL; b0 += B) {
                float4 P[B];
                int op[B], oq[B];
                cplx<T> ap[B], aq[B], vp[B], vq[B];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    P[u] = prm[pr0 + LPR * (b0 + u)];
                    const int pq = __float_as_int(P[u].w);
                    op[u] = lr * ld + (pq & 0xffff);
                    oq[u] = lr * ld + (pq >> 16);
                    ap[u] = A[op[u]];
                    aq[u] = A[oq[u]];
                    vp[u] = V[op[u]];
                    vq[u] = V[oq[u]];
                }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const T c = P[u].x;
                    const cplx<T> se = mk(P[u].y, P[u].z);
                    A[op[u]] = c * ap[u] - mulc(aq[u], se);
                    A[oq[u]] = se * ap[u] + c * aq[u];
                    V[op[u]] = c * vp[u] - mulc(vq[u], se);
                    V[oq[u]] = se * vp[u] + c * vq[u];
                }
            }
            __syncthreads();
            // rows of A:  row_p <- c row_p - s e^{i phi} row_q ; row_q <- s e^{-i phi} row_p + c row_q
#pragma unroll
            for (int b0 = 0; b0 < IPL; b0 += B) {
                float4 P[B];
                int op[B], oq[B];
                cplx<T> ap[B], aq[B];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    P[u] = prm[pr0 + LPR * (b0 + u)];
                    const int pq = __float_as_int(P[u].w);
                    op[u] = (pq & 0xffff) * ld + lr;
                    oq[u] = (pq >> 16) * ld + lr;
                    ap[u] = A[op[u]];
                    aq[u] = A[oq[u]];
                }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const T c = P[u].x;
                    const cplx<T> se = mk(P[u].y, P[u].z);
                    A[op[u]] = c * ap[u] - se * aq[u];
                    A[oq[u]] = mulc(ap[u], se) + c * aq[u];
                }
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
        for (int m = 0; m < nw; ++m) {
            const cplx<T> v = V[m * ld + i];
            Uout[(k * nw + m) * nw + rank] = mk((TO)v.re, (TO)v.im);
        }
    }
}

// ---------------------------------------------------------------------------
// 64 < n <= 128: FP32 Jacobi with A in LDS (132 KB at n = 128) and the
// eigenvectors in row chunks.  The rotations act on V from the right, so the
// rows of V are independent; every workgroup of the grid row blockIdx.y runs
// the same (deterministic) Jacobi on A and accumulates rows [v0, v0 + vr) of
// V.  Uniform control flow: the round flag is written by wave 0 and read by
// all waves after a barrier, double-buffered by round parity.
// ---------------------------------------------------------------------------
#define JROWS_THREADS 256

struct ItemWalkN {
    int pr, i, dp, di;
    __device__ ItemWalkN(int rows, int tid, int nth) : pr(tid / rows), i(tid % rows), dp(nth / rows), di(nth % rows) {}
    __device__ inline void next(int rows) {
        pr += dp;
        i += di;
        if (i >= rows) {
            i -= rows;
            ++pr;
        }
    }
};

__host__ __device__ inline size_t jrows_lds(int n, int vr) {
    const size_t ld = n + 1;
    return sizeof(cplx<float>) * (n + vr) * ld + sizeof(float) * 3 * (n / 2) + sizeof(int) * (n + 4) + 16;
}

template <typename Tin, typename TO>
__global__ void __launch_bounds__(JROWS_THREADS) k_eigh_jacobi_rows(int nw, const cplx<Tin>* __restrict__ H,
                                                                   int64_t Hstride, double* __restrict__ Eout,
                                                                   cplx<TO>* __restrict__ Uout, int max_sweeps,
                                                                   int vr) {
    typedef float T;
    extern __shared__ char smem[];
    const int n = nw + (nw & 1), ld = n + 1, np = n / 2;
    cplx<T>* A = (cplx<T>*)smem;
    cplx<T>* V = A + n * ld;  // rows v0 .. v0 + nv - 1 of V
    T* prm = (T*)(V + vr * ld);
    int* rank = (int*)(prm + 3 * np);
    int* flags = rank + n;  // [0] sweep, [1 + r % 2] round
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    const int v0 = blockIdx.y * vr;
    const int nv = n - v0 < vr ? n - v0 : vr;
    const cplx<Tin>* Hk = H + k * Hstride;
    for (int t = tid; t < n * n; t += JROWS_THREADS) {
        const int i = t / n, j = t % n;
        cplx<T> a = mk((T)0, (T)0);
        if (i < nw && j < nw) {
            const cplx<Tin> x = Hk[i * nw + j], y = Hk[j * nw + i];
            a = mk((T)(0.5 * (x.re + y.re)), (T)(0.5 * (x.im - y.im)));
        }
        A[i * ld + j] = a;
    }
    for (int t = tid; t < nv * n; t += JROWS_THREADS) {
        const int r = t / n, j = t % n;
        V[r * ld + j] = mk((T)(v0 + r == j ? 1 : 0), (T)0);
    }
    __syncthreads();
    const T eps = Eps<T>::v;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (tid == 0) flags[0] = 0;
        __syncthreads();
        for (int r = 0; r < n - 1; ++r) {
            bool rot = false;
            if (tid < np) {
                int p, q;
                rr_pair(n, r, tid, p, q);
                const cplx<T> apq = A[p * ld + q];
                const T g = sqrt(apq.re * apq.re + apq.im * apq.im);
                const T app = A[p * ld + p].re, aqq = A[q * ld + q].re;
                T c = 1, sc = 0, ss = 0;
                const T scale = fabs(app) > fabs(aqq) ? fabs(app) : fabs(aqq);
                if (g > (T)0.5 * eps * scale && g > (T)1e-36) {
                    const T tau = (aqq - app) / (2 * g);
                    const T t = (tau >= 0 ? (T)1 : (T)-1) / (fabs(tau) + sqrt((T)1 + tau * tau));
                    c = (T)1 / sqrt((T)1 + t * t);
                    const T sn = t * c;
                    sc = sn * apq.re / g;
                    ss = sn * apq.im / g;
                    rot = true;
                }
                prm[3 * tid] = c;
                prm[3 * tid + 1] = sc;
                prm[3 * tid + 2] = ss;
            }
            if (tid < 64) {
                const bool any = __any(rot);
                if (tid == 0) {
                    flags[1 + (r & 1)] = any;
                    if (any) flags[0] = 1;
                }
            }
            __syncthreads();
            if (!flags[1 + (r & 1)]) continue;
            // columns of A (rows i < n) and of the V chunk (rows < nv)
            for (int part = 0; part < 2; ++part) {
                cplx<T>* M = part == 0 ? A : V;
                const int rows = part == 0 ? n : nv;
                for (ItemWalkN w(rows, tid, JROWS_THREADS); w.pr < np;) {
                    int ii[JB], pp[JB], qq[JB];
                    bool ok[JB];
                    T cc[JB];
                    cplx<T> se[JB], ap[JB], aq[JB];
#pragma unroll
                    for (int u = 0; u < JB; ++u) {
                        ok[u] = w.pr < np;
                        const int pr = ok[u] ? w.pr : 0;
                        rr_pair(n, r, pr, pp[u], qq[u]);
                        ii[u] = w.i;
                        w.next(rows);
                        cc[u] = prm[3 * pr];
                        se[u] = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                        ap[u] = M[ii[u] * ld + pp[u]];
                        aq[u] = M[ii[u] * ld + qq[u]];
                    }
#pragma unroll
                    for (int u = 0; u < JB; ++u) {
                        if (!ok[u]) continue;
                        M[ii[u] * ld + pp[u]] = cc[u] * ap[u] - mulc(aq[u], se[u]);
                        M[ii[u] * ld + qq[u]] = se[u] * ap[u] + cc[u] * aq[u];
                    }
                }
            }
            __syncthreads();
            // rows of A
            for (ItemWalkN w(n, tid, JROWS_THREADS); w.pr < np;) {
                int jj[JB], pp[JB], qq[JB];
                bool ok[JB];
                T cc[JB];
                cplx<T> se[JB], ap[JB], aq[JB];
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    ok[u] = w.pr < np;
                    const int pr = ok[u] ? w.pr : 0;
                    rr_pair(n, r, pr, pp[u], qq[u]);
                    jj[u] = w.i;
                    w.next(n);
                    cc[u] = prm[3 * pr];
                    se[u] = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                    ap[u] = A[pp[u] * ld + jj[u]];
                    aq[u] = A[qq[u] * ld + jj[u]];
                }
#pragma unroll
                for (int u = 0; u < JB; ++u) {
                    if (!ok[u]) continue;
                    A[pp[u] * ld + jj[u]] = cc[u] * ap[u] - se[u] * aq[u];
                    A[qq[u] * ld + jj[u]] = mulc(ap[u], se[u]) + cc[u] * aq[u];
                }
            }
            __syncthreads();
        }
        if (flags[0] == 0) break;
        __syncthreads();
    }
    // ascending order (ties by index); every chunk computes the same ranks
    for (int i = tid; i < nw; i += JROWS_THREADS) {
        const T ei = A[i * ld + i].re;
        int rk = 0;
        for (int j = 0; j < nw; ++j) {
            const T ej = A[j * ld + j].re;
            rk += (ej < ei) || (ej == ei && j < i);
        }
        rank[i] = rk;
        if (blockIdx.y == 0) Eout[k * nw + rk] = (double)ei;
    }
    __syncthreads();
    for (int t = tid; t < nv * nw; t += JROWS_THREADS) {
        const int r = t / nw, i = t % nw;
        if (v0 + r >= nw) continue;
        const cplx<T> v = V[r * ld + i];
        Uout[(k * nw + v0 + r) * nw + rank[i]] = mk((TO)v.re, (TO)v.im);
    }
}

// ---------------------------------------------------------------------------
// FP64 refinement steps of approximate eigenvectors (Ogita & Aishima,
// "Iterative refinement for symmetric eigenvalue decomposition", Japan J.
// Indust. Appl. Math. 35 (2018), Algorithm 1, written for Hermitian matrices):
//   R = I - X^+ X,  S = X^+ A X,  lam_i = S_ii / (1 - R_ii),
//   delta = 2 (||S - diag(lam)|| + max|lam| ||R||)   (Frobenius bounds),
//   E_ij = (S_ij + lam_j R_ij) / (lam_j - lam_i) if |lam_j - lam_i| > delta, else R_ij / 2,
//   X <- X + X E.
// Each step squares the error of the vectors of separated eigenvalues;
// clusters below delta are re-orthogonalised only.
// ---------------------------------------------------------------------------

#define OA_THREADS 256
// stop refining once ||E||_F is below this: complex64 vectors are limited by
// their storage; complex128 vectors after a step with ||E|| < 1e-10 are
// accurate to ~||E||^2 ||H|| / gap, i.e. to FP64 rounding
#define OA_TOL_C64 1e-8
#define OA_TOL_C128 1e-10
// complex64 vectors: a correction with ||E||_F below this is the last one.  The
// step converges quadratically, so the vectors it leaves are accurate to about
// ||E||^2 = 1e-8 (well below complex64 rounding) where the gaps are not small,
// and the cluster pass that follows rotates every pair still mixed by more than
// OA_TOL_C64 (close pairs converge more slowly).
#define OA_LAST_C64 1e-4

// C = X Y (nw x nw, complex FP64) on the matrix cores: 16 x 16 output blocks
// dealt to the waves of the workgroup, four independent accumulators per block
// (re re, re im, im re, im im) so that consecutive MFMAs never wait on each
// other; out(r, c, value) for r, c < nw.  Lane layout as in mfma_c64.h.
template <int NTHREADS, class FX, class FY, class FO>
__device__ __attribute__((always_inline)) void oa_gemm(int nw, FX X, FY Y, FO out) {
    const int wave = threadIdx.x >> 6, nwave = NTHREADS / 64;
    const int li = threadIdx.x & 15, lk = (threadIdx.x & 63) >> 4;
    const int nb = (nw + 15) / 16;
    for (int t = wave; t < nb * nb; t += nwave) {
        const int r0 = (t / nb) * 16, c0 = (t % nb) * 16;
        const bool rok = r0 + li < nw, cok = c0 + li < nw;
        wbmi_v4d rr = {0.0, 0.0, 0.0, 0.0}, ri = rr, ir = rr, ii = rr;
#pragma unroll 2
        for (int k0 = 0; k0 < nw; k0 += 4) {
            const int k = k0 + lk;
            const bool kok = k < nw;
            const cplx<double> a = (kok && rok) ? X(r0 + li, k) : mk(0.0, 0.0);
            const cplx<double> b = (kok && cok) ? Y(k, c0 + li) : mk(0.0, 0.0);
            rr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, rr, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, ri, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, ir, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, ii, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = r0 + lk + 4 * v, c = c0 + li;
            if (r < nw && c < nw) out(r, c, mk(rr[v] - ii[v], ri[v] + ir[v]));
        }
    }
}

// A = (H + H^+)/2 of k-point Hk into LDS
__device__ inline void oa_load_a(int nw, int ld, const cplx<double>* __restrict__ Hk, cplx<double>* sA) {
    for (int t = threadIdx.x; t < nw * nw; t += OA_THREADS) {
        const int i = t / nw, j = t % nw;
        const cplx<double> x = Hk[i * nw + j], y = Hk[j * nw + i];
        sA[i * ld + j] = mk(0.5 * (x.re + y.re), 0.5 * (x.im - y.im));
    }
}

// Scratch of cluster_jacobi in LDS: per pair {c, s cos phi, s sin phi} and
// (p, q), plus a round flag pair.
__host__ __device__ inline size_t cj_lds(int n) {
    const int np = (n + 1) / 2;
    return sizeof(double) * 3 * np + sizeof(int) * (2 * np + 4);
}

// Finishes the refinement where it stalls: a pair closer than the
// Ogita-Aishima threshold delta is only re-orthogonalised by the correction
// step, so its two vectors can stay any unitary mix of the pair's
// eigenvectors.  With S = X^+ A X of the refined X (Hermitian, diagonal to
// the refinement's accuracy for separated eigenvalues), cyclic Jacobi sweeps
// (round-robin order, rotations of Golub & Van Loan, sec. 8.5.2, as in
// k_eigh_jacobi) rotate only the pairs with
//   |S_pq| > floor_abs  and  |S_pq| > theta |S_qq - S_pp|,
// i.e. whose mixing angle exceeds theta: none after a converged refinement of
// a separated spectrum, the unresolved pairs (and larger clusters) otherwise.
// The rotations update S (columns and rows) and the columns of X.
// Uniform control flow: the round flag is set by the rotating threads and
// read by every thread after a barrier (double-buffered by round parity).
template <int NTH, class FS, class FX>
__device__ void cluster_jacobi(int n, FS S, FX X, char* scratch, double theta, double floor_abs, int max_sweeps) {
    const int ne = n + (n & 1), np = ne / 2;
    double* prm = (double*)scratch;
    int* pq = (int*)(prm + 3 * np);
    int* flags = pq + 2 * np;  // [0] sweep, [1 + (r & 1)] round
    const int tid = threadIdx.x;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (tid == 0) flags[0] = flags[1] = flags[2] = 0;
        __syncthreads();
        for (int r = 0; r < ne - 1; ++r) {
            // the next round's flag: its last readers (round r - 1) passed a barrier since, or read 0
            if (tid == 0) flags[1 + ((r + 1) & 1)] = 0;
            if (tid < np) {
                int p, q;
                rr_pair(ne, r, tid, p, q);
                double c = 1.0, sc = 0.0, ss = 0.0;
                if (p < n && q < n) {
                    const cplx<double> apq = S(p, q);
                    const double g = sqrt(apq.re * apq.re + apq.im * apq.im);
                    const double app = S(p, p).re, aqq = S(q, q).re;
                    if (g > floor_abs && g > theta * fabs(aqq - app)) {
                        const double tau = (aqq - app) / (2.0 * g);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        const double sn = t * c;
                        sc = sn * apq.re / g;
                        ss = sn * apq.im / g;
                        flags[1 + (r & 1)] = 1;
                        flags[0] = 1;
                    }
                }
                prm[3 * tid] = c;
                prm[3 * tid + 1] = sc;
                prm[3 * tid + 2] = ss;
                pq[2 * tid] = p;
                pq[2 * tid + 1] = q;
            }
            __syncthreads();
            if (flags[1 + (r & 1)] == 0) continue;
            // columns of S and X: col_p <- c col_p - s e^{-i phi} col_q ; col_q <- s e^{i phi} col_p + c col_q
            for (int t = tid; t < np * n; t += NTH) {
                const int pr = t / n, i = t % n;
                const int p = pq[2 * pr], q = pq[2 * pr + 1];
                if (p >= n || q >= n) continue;
                const double c = prm[3 * pr];
                const cplx<double> se = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                if (c == 1.0 && se.re == 0.0 && se.im == 0.0) continue;
                const cplx<double> ap = S(i, p), aq = S(i, q), xp = X(i, p), xq = X(i, q);
                S(i, p) = c * ap - mulc(aq, se);
                S(i, q) = se * ap + c * aq;
                X(i, p) = c * xp - mulc(xq, se);
                X(i, q) = se * xp + c * xq;
            }
            __syncthreads();
            // rows of S: row_p <- c row_p - s e^{i phi} row_q ; row_q <- s e^{-i phi} row_p + c row_q
            for (int t = tid; t < np * n; t += NTH) {
                const int pr = t / n, j = t % n;
                const int p = pq[2 * pr], q = pq[2 * pr + 1];
                if (p >= n || q >= n) continue;
                const double c = prm[3 * pr];
                const cplx<double> se = mk(prm[3 * pr + 1], prm[3 * pr + 2]);
                if (c == 1.0 && se.re == 0.0 && se.im == 0.0) continue;
                const cplx<double> ap = S(p, j), aq = S(q, j);
                S(p, j) = c * ap - se * aq;
                S(q, j) = mulc(ap, se) + c * aq;
            }
            __syncthreads();
        }
        const bool more = flags[0] != 0;
        __syncthreads();
        if (!more) break;
    }
}

// Refinement steps until the correction satisfies ||E||_F < tol (at most
// `iters`), then eigenvalues as Rayleigh quotients of the final vectors.
// Well-separated spectra stop after two steps (||E|| ~ eps32, then ~eps32^2).
// A pair closer than delta of the first step (FP32 vectors mix it) is only
// re-orthogonalised there; the following steps separate it, converging
// quadratically once the mixing angle is small (six to eight steps for a
// relative gap of 4e-6 at nw = 18..40).
// (at least 5 waves per SIMD: 72 VGPRs and no AGPRs, 7 waves per SIMD)
template <typename TU>
__global__ void __launch_bounds__(OA_THREADS, 5) k_eigh_refine(int nw, const cplx<double>* __restrict__ H64,
                                                            int64_t Hstride, double* __restrict__ Eout,
                                                            cplx<TU>* __restrict__ U, int iters, double tol) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int ld = nw + 1;
    cplx<double>* sX = (cplx<double>*)smem;   // X (approximate eigenvectors)
    cplx<double>* sA = sX + nw * ld;          // A (for the whole solve)
    cplx<double>* sT = sA + nw * ld;          // A X, then R, then X + X E
    cplx<double>* sS = sT + nw * ld;          // S, then E
    double* lam = (double*)(sS + nw * ld);
    double* red = lam + nw;                   // [0] ||S-D||^2, [1] ||R||^2, [2] max|lam|, [3] ||E||^2
    char* cjs = (char*)(red + 4);             // cluster_jacobi scratch
    int* rank = (int*)(cjs + cj_lds(nw));
    const int tid = threadIdx.x;
    const int64_t k = blockIdx.x;
    const cplx<double>* Hk = H64 + k * Hstride;
    cplx<TU>* Uk = U + k * nw * nw;
    for (int t = tid; t < nw * nw; t += OA_THREADS) {
        const cplx<TU> u = Uk[t];
        sX[(t / nw) * ld + t % nw] = mk((double)u.re, (double)u.im);
    }
    oa_load_a(nw, ld, Hk, sA);
    for (int it = 0; it < iters; ++it) {
        if (tid < 4) red[tid] = 0.0;
        __syncthreads();
        // T = A X
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return sA[r * ld + m]; }, [&](int m, int c) { return sX[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sT[r * ld + c] = v; });
        __syncthreads();
        // S = X^+ T ; then R = I - X^+ X into sT
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return conj(sX[m * ld + r]); },
                    [&](int m, int c) { return sT[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sS[r * ld + c] = v; });
        __syncthreads();
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return conj(sX[m * ld + r]); },
                    [&](int m, int c) { return sX[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sT[r * ld + c] = mk((r == c ? 1.0 : 0.0) - v.re, -v.im); });
        __syncthreads();
        for (int i = tid; i < nw; i += OA_THREADS) lam[i] = sS[i * ld + i].re / (1.0 - sT[i * ld + i].re);
        __syncthreads();
        // Frobenius norms and max|lam|
        double s2 = 0.0, r2 = 0.0, lm = 0.0;
        for (int t = tid; t < nw * nw; t += OA_THREADS) {
            const int i = t / nw, j = t % nw;
            cplx<double> sv = sS[i * ld + j];
            if (i == j) sv.re -= lam[i];
            s2 += sv.re * sv.re + sv.im * sv.im;
            const cplx<double> rv = sT[i * ld + j];
            r2 += rv.re * rv.re + rv.im * rv.im;
        }
        for (int i = tid; i < nw; i += OA_THREADS) lm = fmax(lm, fabs(lam[i]));
        for (int off = 32; off > 0; off >>= 1) {
            s2 += __shfl_down(s2, off);
            r2 += __shfl_down(r2, off);
            lm = fmax(lm, __shfl_down(lm, off));
        }
        if ((tid & 63) == 0) {
            atomicAdd(&red[0], s2);
            atomicAdd(&red[1], r2);
            atomicMax((unsigned long long*)&red[2], (unsigned long long)__double_as_longlong(lm));
        }
        __syncthreads();
        const double delta = 2.0 * (sqrt(red[0]) + red[2] * sqrt(red[1]));
        // E into sS, ||E||^2 into red[3]
        double e2 = 0.0;
        for (int t = tid; t < nw * nw; t += OA_THREADS) {
            const int i = t / nw, j = t % nw;
            const double dl = lam[j] - lam[i];
            const cplx<double> r = sT[i * ld + j];
            cplx<double> e;
            if (fabs(dl) > delta) {
                const cplx<double> sv = sS[i * ld + j];
                e = (1.0 / dl) * mk(sv.re + lam[j] * r.re, sv.im + lam[j] * r.im);
            } else {
                e = 0.5 * r;
            }
            sS[i * ld + j] = e;
            e2 += e.re * e.re + e.im * e.im;
        }
        for (int off = 32; off > 0; off >>= 1) e2 += __shfl_down(e2, off);
        if ((tid & 63) == 0) atomicAdd(&red[3], e2);
        __syncthreads();
        // every thread reads red[3] here; it is cleared only after the next barrier
        const bool converged = red[3] < tol * tol || (sizeof(TU) == 4 && red[3] < OA_LAST_C64 * OA_LAST_C64);
        // X + X E into sT (R is no longer read), which becomes X
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return sX[r * ld + m]; }, [&](int m, int c) { return sS[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sT[r * ld + c] = sX[r * ld + c] + v; });
        __syncthreads();
        cplx<double>* t = sX;
        sX = sT;
        sT = t;
        if (converged) break;
    }
    if (sizeof(TU) == 8) {
        // complex128 output: one Newton-Schulz step X <- X (I + R/2), R = I - X^+ X, so that the
        // basis is orthonormal to rounding also where a close pair leaves the vectors themselves
        // accurate only to ~eps |H| / gap (the correction steps do not preserve orthogonality)
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return conj(sX[m * ld + r]); },
                    [&](int m, int c) { return sX[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sS[r * ld + c] = mk(0.5 * ((r == c ? 1.0 : 0.0) - v.re), -0.5 * v.im); });
        __syncthreads();
        oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return sX[r * ld + m]; }, [&](int m, int c) { return sS[m * ld + c]; },
                    [&](int r, int c, cplx<double> v) { sT[r * ld + c] = sX[r * ld + c] + v; });
        __syncthreads();
        cplx<double>* t = sX;
        sX = sT;
        sT = t;
    }
    // S = X^+ A X of the refined columns; pairs left mixed are rotated apart
    const double scale = red[2];  // max |lam| of the last step
    oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return sA[r * ld + m]; }, [&](int m, int c) { return sX[m * ld + c]; },
                [&](int r, int c, cplx<double> v) { sT[r * ld + c] = v; });
    __syncthreads();
    oa_gemm<OA_THREADS>(nw, [&](int r, int m) { return conj(sX[m * ld + r]); },
                [&](int m, int c) { return sT[m * ld + c]; },
                [&](int r, int c, cplx<double> v) { sS[r * ld + c] = v; });
    __syncthreads();
    cluster_jacobi<OA_THREADS>(
        nw, [&](int i, int j) -> cplx<double>& { return sS[i * ld + j]; },
        [&](int i, int j) -> cplx<double>& { return sX[i * ld + j]; }, cjs, tol, 64.0 * DBL_EPSILON * scale, 10);
    // eigenvalues: Rayleigh quotients x_c^+ A x_c / x_c^+ x_c = S_cc / x_c^+ x_c
    for (int c = tid; c < nw; c += OA_THREADS) {
        double den = 0.0;
        for (int m = 0; m < nw; ++m) {
            const cplx<double> x = sX[m * ld + c];
            den += x.re * x.re + x.im * x.im;
        }
        lam[c] = sS[c * ld + c].re / den;
    }
    __syncthreads();
    // ascending order (ties by index)
    for (int i = tid; i < nw; i += OA_THREADS) {
        int rk = 0;
        for (int j = 0; j < nw; ++j) rk += (lam[j] < lam[i]) || (lam[j] == lam[i] && j < i);
        rank[i] = rk;
    }
    __syncthreads();
    for (int t = tid; t < nw * nw; t += OA_THREADS) {
        const int r = t / nw, c = t % nw;
        const cplx<double> x = sX[r * ld + c];
        Uk[r * nw + rank[c]] = mk((TU)x.re, (TU)x.im);
    }
    for (int i = tid; i < nw; i += OA_THREADS) Eout[k * nw + rank[i]] = lam[i];
}

// ---------------------------------------------------------------------------
// Refinement for 44 < nw <= 128: the same steps with the four nw x nw matrices
// (X, T, S/E, R) in a per-k workspace of global memory and every product as
// 32 x 32 wave tasks on the FP64 matrix cores (operands read through L2).
// ---------------------------------------------------------------------------
#define OAG_THREADS 512
#define OAG_NWMAX 128

// C = P Q (n x n) as wave tasks; out(r, c, value) for r, c < n
template <class FP, class FQ, class FO>
__device__ __attribute__((always_inline)) void wg_gemm(int n, FP P, FQ Q, FO out) {
    const int wave = threadIdx.x >> 6, nwave = OAG_THREADS / 64;
    const int nb = (n + 31) / 32;
    for (int t = wave; t < nb * nb; t += nwave) {
        const int rb = t / nb, cb = t % nb;
        MfmaBlock C;
        C.zero();
        mfma_block(C, n, n, n, rb * 32, cb * 32, P, Q);
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = rb * 32 + mfma_row(bi, v), c = cb * 32 + mfma_col(bj);
                    if (r < n && c < n) out(r, c, C.get(bi, bj, v));
                }
    }
}

__device__ inline double block_sum_to(double v, double* slot) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(slot, v);
    return v;
}

template <typename TU>
__global__ void __launch_bounds__(OAG_THREADS) k_eigh_refine_g(int n, const cplx<double>* __restrict__ H64,
                                                              int64_t Hstride, double* __restrict__ Eout,
                                                              cplx<TU>* __restrict__ U, int iters, double tol,
                                                              cplx<double>* __restrict__ work) {
    __shared__ double lam[OAG_NWMAX];
    __shared__ double red[4];  // [0] ||S-D||^2, [1] ||R||^2, [2] max|lam|, [3] ||E||^2
    __shared__ int rank[OAG_NWMAX];
    __shared__ __attribute__((aligned(16))) char cjs[sizeof(double) * 3 * (OAG_NWMAX / 2) +
                                                     sizeof(int) * (OAG_NWMAX + 4)];
    const int tid = threadIdx.x, n2 = n * n;
    const int64_t k = blockIdx.x;
    const cplx<double>* Hk = H64 + k * Hstride;
    cplx<TU>* Uk = U + k * n2;
    cplx<double>* X = work + k * 4 * (int64_t)n2;
    cplx<double>* T = X + n2;
    cplx<double>* S = T + n2;
    cplx<double>* R = S + n2;
    auto Aop = [&](int r, int m) {
        const cplx<double> x = Hk[r * n + m], y = Hk[m * n + r];
        return mk(0.5 * (x.re + y.re), 0.5 * (x.im - y.im));
    };
    for (int t = tid; t < n2; t += OAG_THREADS) {
        const cplx<TU> u = Uk[t];
        X[t] = mk((double)u.re, (double)u.im);
    }
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        if (tid < 4) red[tid] = 0.0;
        wg_gemm(n, Aop, [&](int m, int c) { return X[m * n + c]; }, [&](int r, int c, cplx<double> v) { T[r * n + c] = v; });
        __syncthreads();
        wg_gemm(n, [&](int r, int m) { return conj(X[m * n + r]); }, [&](int m, int c) { return T[m * n + c]; },
                [&](int r, int c, cplx<double> v) { S[r * n + c] = v; });
        wg_gemm(n, [&](int r, int m) { return conj(X[m * n + r]); }, [&](int m, int c) { return X[m * n + c]; },
                [&](int r, int c, cplx<double> v) { R[r * n + c] = mk((r == c ? 1.0 : 0.0) - v.re, -v.im); });
        __syncthreads();
        for (int i = tid; i < n; i += OAG_THREADS) lam[i] = S[i * n + i].re / (1.0 - R[i * n + i].re);
        __syncthreads();
        double s2 = 0.0, r2 = 0.0, lm = 0.0;
        for (int t = tid; t < n2; t += OAG_THREADS) {
            const int i = t / n, j = t % n;
            cplx<double> sv = S[t];
            if (i == j) sv.re -= lam[i];
            s2 += sv.re * sv.re + sv.im * sv.im;
            r2 += R[t].re * R[t].re + R[t].im * R[t].im;
        }
        for (int i = tid; i < n; i += OAG_THREADS) lm = fmax(lm, fabs(lam[i]));
        for (int off = 32; off > 0; off >>= 1) lm = fmax(lm, __shfl_down(lm, off));
        block_sum_to(s2, &red[0]);
        block_sum_to(r2, &red[1]);
        if ((tid & 63) == 0) atomicMax((unsigned long long*)&red[2], (unsigned long long)__double_as_longlong(lm));
        __syncthreads();
        const double delta = 2.0 * (sqrt(red[0]) + red[2] * sqrt(red[1]));
        double e2 = 0.0;
        for (int t = tid; t < n2; t += OAG_THREADS) {
            const int i = t / n, j = t % n;
            const double dl = lam[j] - lam[i];
            const cplx<double> r = R[t];
            cplx<double> e;
            if (fabs(dl) > delta) {
                const cplx<double> sv = S[t];
                e = (1.0 / dl) * mk(sv.re + lam[j] * r.re, sv.im + lam[j] * r.im);
            } else {
                e = 0.5 * r;
            }
            S[t] = e;
            e2 += e.re * e.re + e.im * e.im;
        }
        block_sum_to(e2, &red[3]);
        __syncthreads();
        // every thread reads red[3] here; it is cleared only after the next barrier
        const bool converged = red[3] < tol * tol || (sizeof(TU) == 4 && red[3] < OA_LAST_C64 * OA_LAST_C64);
        wg_gemm(n, [&](int r, int m) { return X[r * n + m]; }, [&](int m, int c) { return S[m * n + c]; },
                [&](int r, int c, cplx<double> v) { T[r * n + c] = X[r * n + c] + v; });
        __syncthreads();
        cplx<double>* t = X;
        X = T;
        T = t;
        if (converged) break;
    }
    if (sizeof(TU) == 8) {
        // Newton-Schulz orthonormalisation X <- X (I + R/2), R = I - X^+ X
        wg_gemm(n, [&](int r, int m) { return conj(X[m * n + r]); }, [&](int m, int c) { return X[m * n + c]; },
                [&](int r, int c, cplx<double> v) { R[r * n + c] = mk(0.5 * ((r == c ? 1.0 : 0.0) - v.re), -0.5 * v.im); });
        __syncthreads();
        wg_gemm(n, [&](int r, int m) { return X[r * n + m]; }, [&](int m, int c) { return R[m * n + c]; },
                [&](int r, int c, cplx<double> v) { T[r * n + c] = X[r * n + c] + v; });
        __syncthreads();
        cplx<double>* t = X;
        X = T;
        T = t;
    }
    // S = X^+ A X of the refined columns; pairs left mixed are rotated apart
    const double scale = red[2];  // max |lam| of the last step
    wg_gemm(n, Aop, [&](int m, int c) { return X[m * n + c]; }, [&](int r, int c, cplx<double> v) { T[r * n + c] = v; });
    __syncthreads();
    wg_gemm(n, [&](int r, int m) { return conj(X[m * n + r]); }, [&](int m, int c) { return T[m * n + c]; },
            [&](int r, int c, cplx<double> v) { S[r * n + c] = v; });
    __syncthreads();
    cluster_jacobi<OAG_THREADS>(
        n, [&](int i, int j) -> cplx<double>& { return S[i * n + j]; },
        [&](int i, int j) -> cplx<double>& { return X[i * n + j]; }, cjs, tol, 64.0 * DBL_EPSILON * scale, 10);
    // Rayleigh quotients x_c^+ A x_c / x_c^+ x_c = S_cc / x_c^+ x_c
    for (int c = tid; c < n; c += OAG_THREADS) {
        double den = 0.0;
        for (int m = 0; m < n; ++m) {
            const cplx<double> x = X[m * n + c];
            den += x.re * x.re + x.im * x.im;
        }
        lam[c] = S[c * n + c].re / den;
    }
    __syncthreads();
    for (int i = tid; i < n; i += OAG_THREADS) {
        int rk = 0;
        for (int j = 0; j < n; ++j) rk += (lam[j] < lam[i]) || (lam[j] == lam[i] && j < i);
        rank[i] = rk;
    }
    __syncthreads();
    for (int t = tid; t < n2; t += OAG_THREADS) {
        const int r = t / n, c = t % n;
        const cplx<double> x = X[t];
        Uk[r * n + rank[c]] = mk((TU)x.re, (TU)x.im);
    }
    for (int i = tid; i < n; i += OAG_THREADS) Eout[k * n + rank[i]] = lam[i];
}

// Scratch of wbmi_eigh_noalloc: the per-k refinement workspace (X, T, S/E, R for
// 44 < nw <= 128) in a file-scope device array, so that callers that cannot allocate
// (graph capture, serving loops) get the refined solver without a workspace.  Each GPU
// holds its own copy for the life of the process; calls on one GPU share it (their
// kernels write every element before reading it, so nothing needs clearing, and
// overlapping calls on one GPU race).  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
#define WBMI_EIGH_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) char scratch[WBMI_EIGH_SCRATCH_BYTES];

int wbmi_eigh_scratch_init(void** ptr) {
    *ptr = nullptr;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)) == hipSuccess && *ptr ? 0 : 1;
}

// bytes of global workspace wbmi_eigh needs (refinement for nw > 44)
static size_t eigh_work_bytes(int nk, int nw, int refine) {
    return (refine && nw > 44) ? sizeof(cplx<double>) * 4 * (size_t)nw * nw * nk : 0;
}

extern "C" size_t wbmi_eigh_work_bytes(int nk, int nw, int prec, int refine) {
    (void)prec;
    return nk > 0 && nw > 0 ? eigh_work_bytes(nk, nw, refine) : 0;
}

// FP32 Jacobi of the FP64 Hamiltonian (eigenvectors stored as TU)
template <typename TU>
static int jacobi_f32_of_f64(wbmi_ctx* ctx, int nk, int nw, const cplx<double>* H64, int64_t Hstride, double* E,
                             cplx<TU>* U, const cplx<TU>* U0 = nullptr) {
    const int n = nw + (nw & 1);
    const int sweeps = 40;
    if (n == 16 || n == 32 || n == 64) {
        const void* fn = n == 16 ? (const void*)k_eigh_jacobi_c<16, double, TU>
                       : n == 32 ? (const void*)k_eigh_jacobi_c<32, double, TU>
                                 : (const void*)k_eigh_jacobi_c<64, double, TU>;
        const size_t lds = n == 16 ? jacobi_c_lds<16>() : n == 32 ? jacobi_c_lds<32>() : jacobi_c_lds<64>();
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (n == 16)
            hipLaunchKernelGGL((k_eigh_jacobi_c<16, double, TU>), dim3(nk), dim3(64), lds, ctx->stream, nw, H64,
                               Hstride, E, U, sweeps, U0);
        else if (n == 32)
            hipLaunchKernelGGL((k_eigh_jacobi_c<32, double, TU>), dim3(nk), dim3(64), lds, ctx->stream, nw, H64,
                               Hstride, E, U, sweeps, U0);
        else
            hipLaunchKernelGGL((k_eigh_jacobi_c<64, double, TU>), dim3(nk), dim3(64), lds, ctx->stream, nw, H64,
                               Hstride, E, U, sweeps, (const cplx<TU>*)nullptr);
    } else if (n <= 64) {
        const size_t lds = 2 * sizeof(cplx<float>) * n * (n + 1) + 3 * sizeof(float) * (n / 2) + 16;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi<float, double, TU>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_eigh_jacobi<float, double, TU>), dim3(nk), dim3(64), lds, ctx->stream, nw, H64, Hstride,
                           E, U, sweeps);
    } else {
        // rows of V per workgroup: what the LDS leaves after A
        int vr = 1;
        while (vr < n && jrows_lds(n, vr + 1) <= 160 * 1024) ++vr;
        const int nchunk = (n + vr - 1) / vr;
        const size_t lds = jrows_lds(n, vr);
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi_rows<double, TU>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_eigh_jacobi_rows<double, TU>), dim3(nk, nchunk), dim3(JROWS_THREADS), lds, ctx->stream,
                           nw, H64, Hstride, E, U, sweeps, vr);
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// FP32 Jacobi of the FP64 Hamiltonian, then FP64 Ogita-Aishima steps; eigenvectors stored as TU
template <typename TU>
static int eigh_refined(wbmi_ctx* ctx, int nk, int nw, const cplx<double>* H64, int64_t Hstride, double* E,
                        cplx<TU>* U, int steps, void* work, size_t work_bytes, const cplx<TU>* U0 = nullptr) {
    const double tol = sizeof(TU) == 4 ? OA_TOL_C64 : OA_TOL_C128;
    int rc = jacobi_f32_of_f64<TU>(ctx, nk, nw, H64, Hstride, E, U, U0);
    if (rc) return rc;
    if (nw > 44) {
        if (!work || work_bytes < eigh_work_bytes(nk, nw, steps))
            return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: refinement at nw = %d needs %zu bytes of workspace", nw,
                             eigh_work_bytes(nk, nw, steps));
        hipLaunchKernelGGL((k_eigh_refine_g<TU>), dim3(nk), dim3(OAG_THREADS), 0, ctx->stream, nw, H64, Hstride, E, U,
                           steps, tol, (cplx<double>*)work);
        WBMI_LAUNCH_CHECK(ctx);
        return WBMI_OK;
    }
    const size_t lds2 =
        4 * sizeof(cplx<double>) * nw * (nw + 1) + sizeof(double) * (nw + 4) + cj_lds(nw) + sizeof(int) * nw + 16;
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_refine<TU>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
    hipLaunchKernelGGL((k_eigh_refine<TU>), dim3(nk), dim3(OA_THREADS), lds2, ctx->stream, nw, H64, Hstride, E, U,
                       steps, tol);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" int wbmi_eigh(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride, double* E_dev,
                         void* U_dev, int prec, int refine, const void* H64_dev, void* work_dev, size_t work_bytes) {
    if (!ctx) return WBMI_ERR_ARG;
    if (nk <= 0 || nw <= 0 || nw > 128 || !E_dev || !U_dev || Hstride < (int64_t)nw * nw)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: bad arguments (nw <= 128, Hstride >= nw*nw)");
    if (refine < 0 || refine > 16) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: refine must be 0..16");
    if (prec != WBMI_F64 && prec != WBMI_F32) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: bad precision");
    if (refine) {
        // the FP64 Hamiltonian: H_dev (complex128) for prec = WBMI_F64, H64_dev for prec = WBMI_F32
        const void* H64 = prec == WBMI_F64 ? H_dev : H64_dev;
        if (!H64) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: refinement needs the FP64 Hamiltonian");
        if (prec == WBMI_F64)
            return eigh_refined<double>(ctx, nk, nw, (const cplx<double>*)H64, Hstride, E_dev, (cplx<double>*)U_dev,
                                        refine, work_dev, work_bytes);
        return eigh_refined<float>(ctx, nk, nw, (const cplx<double>*)H64, Hstride, E_dev, (cplx<float>*)U_dev, refine,
                                   work_dev, work_bytes);
    }
    if (!H_dev) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: H_dev is null");
    if (nw > 64) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh: nw > 64 needs refine >= 1");
    const int n = nw + (nw & 1);
    const int sweeps = 40;
    if (prec == WBMI_F64) {
        size_t lds = 2 * sizeof(cplx<double>) * n * (n + 1) + 3 * sizeof(double) * (n / 2) + 16;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi<double, double>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_eigh_jacobi<double, double>), dim3(nk), dim3(64), lds, ctx->stream, nw,
                           (const cplx<double>*)H_dev, Hstride, E_dev, (cplx<double>*)U_dev, sweeps);
    } else {
        size_t lds = 2 * sizeof(cplx<float>) * n * (n + 1) + 3 * sizeof(float) * (n / 2) + 16;
        WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_eigh_jacobi<float, float>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_eigh_jacobi<float, float>), dim3(nk), dim3(64), lds, ctx->stream, nw,
                           (const cplx<float>*)H_dev, Hstride, E_dev, (cplx<float>*)U_dev, sweeps);
    }
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" size_t wbmi_eigh_noalloc_max_bytes(void) { return WBMI_EIGH_SCRATCH_BYTES; }

extern "C" int wbmi_eigh_noalloc(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride, double* E_dev,
                                 void* U_dev, int prec, int refine, const void* H64_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    const size_t need = nk > 0 && nw > 0 ? eigh_work_bytes(nk, nw, refine) : 0;
    if (need > WBMI_EIGH_SCRATCH_BYTES)
        return wbmi_fail(ctx, WBMI_ERR_ARG,
                         "wbmi_eigh_noalloc: %d k-points at nw = %d need %zu bytes of scratch, more than its %zu "
                         "(at most %zu k-points per call at this nw)",
                         nk, nw, need, WBMI_EIGH_SCRATCH_BYTES,
                         WBMI_EIGH_SCRATCH_BYTES / eigh_work_bytes(1, nw, refine));
    if (need > 0 && !ctx->eigh_scratch) return wbmi_fail(ctx, WBMI_ERR_HIP, "wbmi_eigh_noalloc: no scratch address");
    return wbmi_eigh(ctx, nk, nw, H_dev, Hstride, E_dev, U_dev, prec, refine, H64_dev, ctx->eigh_scratch,
                     WBMI_EIGH_SCRATCH_BYTES);
}

extern "C" int wbmi_eigh_guess(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride, double* E_dev,
                               void* U_dev, int prec, int refine, const void* H64_dev, const void* U0_dev,
                               void* work_dev, size_t work_bytes) {
    if (!ctx) return WBMI_ERR_ARG;
    if (!U0_dev)
        return wbmi_eigh(ctx, nk, nw, H_dev, Hstride, E_dev, U_dev, prec, refine, H64_dev, work_dev, work_bytes);
    if (nk <= 0 || nw <= 0 || nw > 128 || !E_dev || !U_dev || Hstride < (int64_t)nw * nw)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh_guess: bad arguments (nw <= 128, Hstride >= nw*nw)");
    if (refine < 1 || refine > 16) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh_guess: refine must be 1..16");
    if (prec != WBMI_F64 && prec != WBMI_F32) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh_guess: bad precision");
    const void* H64 = prec == WBMI_F64 ? H_dev : H64_dev;
    if (!H64) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_eigh_guess: refinement needs the FP64 Hamiltonian");
    if (prec == WBMI_F64)
        return eigh_refined<double>(ctx, nk, nw, (const cplx<double>*)H64, Hstride, E_dev, (cplx<double>*)U_dev,
                                    refine, work_dev, work_bytes, (const cplx<double>*)U0_dev);
    return eigh_refined<float>(ctx, nk, nw, (const cplx<double>*)H64, Hstride, E_dev, (cplx<float>*)U_dev, refine,
                               work_dev, work_bytes, (const cplx<float>*)U0_dev);
}
