// Covariant formula traces + Fermi-level scan of libwbmi.
//
// Replaces the per-(k, band group) loop of StaticCalculator.__call__
// (calculators/static.py:60-136).  One 256-thread workgroup per k-point.
//
// Band partition.  The reference evaluates, for every band group G, the
// formula with inn = G and out = all other bands.  Every sum in the formulas
// is then either over l in out (l not in G) or over n' in G, so with a label
// per band (lab[x] = group of x) the traces of ALL groups come from one set
// of per-band values v_n, n in G, using the masks
//     Off(x, y) = lab[x] != lab[y]      (x, y in different groups)
//     On(x, y)  = lab[x] == lab[y]
// and the masked matrices
//     D'^a = -V^a o dEinv o Off,  V'' = V o On,  A'' = A o On.
// The Fermi-sea group (all bands below EFmin, data_K.py:178-185) and the
// cumulative cuts [0, c) of non-additive formulas (static.py:109-117) are
// two-set partitions evaluated the same way.
//
// All arithmetic is FP64 (energies, dEinv, traces, histogram); channels may
// be stored in FP32 or FP64.
#include <cmath>

#include "wbmi_common.h"

#define SC_THREADS 256
#define SC_NWMAX 64

__constant__ int c_alpha[3] = {1, 2, 0};  // utility.py:45-46
__constant__ int c_beta[3] = {2, 0, 1};

struct ScanArgs {
    wbmi_scan_params p;
    int nw, nch;
    double weight;
};

template <typename T>
struct KCtx {
    const cplx<T>* X;  // channels of this k: [nch][nw][nw]
    int nw;
    const int* ch;
    const double* E;
    const double* M;  // dEinv o Off
    const int* lab;
    const int* gs;
    const int* ge;
    __device__ inline cplx<double> get(int role, int x, int y) const {
        cplx<T> v = X[((int64_t)ch[role] * nw + x) * nw + y];
        return mk((double)v.re, (double)v.im);
    }
    __device__ inline cplx<double> V(int a, int x, int y) const { return get(WBMI_ROLE_V + a, x, y); }
    __device__ inline cplx<double> W(int a, int b, int x, int y) const { return get(WBMI_ROLE_W + 3 * a + b, x, y); }
    __device__ inline cplx<double> A(int a, int x, int y) const { return get(WBMI_ROLE_A + a, x, y); }
    __device__ inline cplx<double> Ad(int a, int d, int x, int y) const { return get(WBMI_ROLE_AD + 3 * a + d, x, y); }
    __device__ inline cplx<double> O(int c, int x, int y) const { return get(WBMI_ROLE_O + c, x, y); }
    __device__ inline cplx<double> Od(int c, int d, int x, int y) const { return get(WBMI_ROLE_OD + 3 * c + d, x, y); }
    __device__ inline cplx<double> B(int a, int x, int y) const { return get(WBMI_ROLE_B + a, x, y); }
    __device__ inline cplx<double> C(int a, int x, int y) const { return get(WBMI_ROLE_C + a, x, y); }
    // D'^a_xy = -V^a_xy dEinv_xy Off(x, y)
    __device__ inline cplx<double> Dp(int a, int x, int y) const {
        double m = M[x * nw + y];
        if (m == 0.0) return mk(0.0, 0.0);
        return (-m) * V(a, x, y);
    }
};

// pre-Hermitian Omega summand s^c_xy for x, y in the same group
// (formula/covariant.py:175-200): Omega.nn = s + s^+
template <typename T>
__device__ cplx<double> omega_s(const KCtx<T>& K, int c, int x, int y, bool internal, bool external) {
    const int al = c_alpha[c], be = c_beta[c];
    cplx<double> s = mk(0.0, 0.0);
    for (int l = 0; l < K.nw; ++l) {
        if (K.lab[l] == K.lab[x]) continue;
        cplx<double> dxa = K.Dp(al, x, l);
        if (internal) s += times_i(-(dxa * K.Dp(be, l, y)));
        if (external) s += K.Dp(be, x, l) * K.A(al, l, y) - dxa * K.A(be, l, y);
    }
    if (external) {
        for (int m = K.gs[x]; m < K.ge[x]; ++m) s += times_i(-(K.A(al, x, m) * K.A(be, m, y)));
        s += 0.5 * K.O(c, x, y);
    }
    return s;
}

// pre-Hermitian Morb_H summand h^c_xy for x, y in the same group
// (formula/covariant.py:390-414)
template <typename T>
__device__ cplx<double> morbh_s(const KCtx<T>& K, int c, int x, int y, bool internal, bool external) {
    const int al = c_alpha[c], be = c_beta[c];
    cplx<double> s = mk(0.0, 0.0);
    for (int l = 0; l < K.nw; ++l) {
        if (K.lab[l] == K.lab[x]) continue;
        cplx<double> dxa = K.Dp(al, x, l);
        if (internal) s += times_i(-(K.E[l] * (dxa * K.Dp(be, l, y))));
        if (external) s += K.Dp(be, x, l) * K.B(al, l, y) - dxa * K.B(be, l, y);
    }
    if (external) {
        for (int m = K.gs[x]; m < K.ge[x]; ++m) s += times_i(-(K.E[m] * (K.A(al, x, m) * K.A(be, m, y))));
        s += 0.5 * K.C(c, x, y);
    }
    return s;
}

// Shared-memory layout (dynamic)
struct Smem {
    double* E;
    double* M;
    cplx<double>* P;
    cplx<double>* acc;  // [nw][9]
    double* cut;        // [nw+1][9]
    double* gw;         // group weights (mean energy)
    int* lab;
    int* gs;
    int* ge;
    int* act;
    int* gb;   // group borders [nw+1]
    int* gin;  // group in range
    int* misc; // [0] = ngroups, [1] = bandmax
    int* ch;   // channel slot per role [WBMI_NROLES]
};

__device__ inline void* align16(void* p) { return (void*)(((uintptr_t)p + 15) & ~(uintptr_t)15); }

__device__ Smem carve(char* base, int nw) {
    Smem s;
    char* p = base;
    s.P = (cplx<double>*)align16(p);
    p = (char*)(s.P + nw * nw);
    s.acc = (cplx<double>*)align16(p);
    p = (char*)(s.acc + nw * 9);
    s.M = (double*)align16(p);
    p = (char*)(s.M + nw * nw);
    s.E = (double*)p;
    p = (char*)(s.E + nw);
    s.cut = (double*)p;
    p = (char*)(s.cut + (nw + 1) * 9);
    s.gw = (double*)p;
    p = (char*)(s.gw + nw);
    s.lab = (int*)p;
    s.gs = s.lab + nw;
    s.ge = s.gs + nw;
    s.act = s.ge + nw;
    s.gb = s.act + nw;
    s.gin = s.gb + nw + 1;
    s.misc = s.gin + nw;
    s.ch = s.misc + 4;
    return s;
}

static size_t smem_bytes(int nw) {
    size_t b = 0;
    b += 16 + sizeof(cplx<double>) * nw * nw;
    b += 16 + sizeof(cplx<double>) * nw * 9;
    b += 16 + sizeof(double) * nw * nw;
    b += sizeof(double) * nw;
    b += sizeof(double) * (nw + 1) * 9;
    b += sizeof(double) * nw;
    b += sizeof(int) * (5 * nw + 1 + nw + 4 + WBMI_NROLES);
    return b + 64;
}

__device__ inline int ncomp_of(int f) {
    switch (f) {
        case WBMI_FORMULA_IDENTITY: return 1;
        case WBMI_FORMULA_OMEGA: case WBMI_FORMULA_MORB_HPM: return 3;
        default: return 9;
    }
}

// build M = dEinv o Off for the current labels (data_K.py:290-298)
__device__ void build_M(const Smem& S, int nw) {
    for (int t = threadIdx.x; t < nw * nw; t += SC_THREADS) {
        int x = t / nw, y = t % nw;
        double m = 0.0;
        if (S.lab[x] != S.lab[y]) {
            double de = S.E[x] - S.E[y];
            if (fabs(de) >= 1e-7) m = 1.0 / de;
        }
        S.M[t] = m;
    }
}

// Per-band values for the current partition into S.acc[n][comp] (real part
// kept in .re).  Only bands with act[n] are evaluated.
template <typename T>
__device__ void eval_partition(const Smem& S, const KCtx<T>& K, const wbmi_scan_params& p) {
    const int nw = K.nw;
    const int tid = threadIdx.x;
    const bool in = p.internal_terms, ex = p.external_terms;
    build_M(S, nw);
    for (int t = tid; t < nw * 9; t += SC_THREADS) S.acc[t] = mk(0.0, 0.0);
    __syncthreads();

    switch (p.formula) {
        case WBMI_FORMULA_IDENTITY: {
            for (int n = tid; n < nw; n += SC_THREADS)
                if (S.act[n]) S.acc[n * 9] = mk(1.0, 0.0);
            break;
        }
        case WBMI_FORMULA_OMEGA: {
            for (int t = tid; t < nw * 3; t += SC_THREADS) {
                int n = t / 3, c = t % 3;
                if (!S.act[n]) continue;
                S.acc[n * 9 + c] = mk(2.0 * omega_s(K, c, n, n, in, ex).re, 0.0);
            }
            break;
        }
        case WBMI_FORMULA_MORB_HPM: {
            for (int t = tid; t < nw * 3; t += SC_THREADS) {
                int n = t / 3, c = t % 3;
                if (!S.act[n]) continue;
                double v = 2.0 * morbh_s(K, c, n, n, in, ex).re;
                if (p.sign != 0) v += p.sign * S.E[n] * 2.0 * omega_s(K, c, n, n, in, ex).re;
                S.acc[n * 9 + c] = mk(v, 0.0);
            }
            break;
        }
        case WBMI_FORMULA_VELOMEGA: {
            // Re sum_{n' in G} V^a_nn' Omega^b_n'n, Omega = s + s^+ (formula/formula.py:141-148)
            for (int t = tid; t < nw * 9; t += SC_THREADS) {
                int n = t / 9, a = (t % 9) / 3, b = t % 3;
                if (!S.act[n]) continue;
                cplx<double> v = mk(0.0, 0.0);
                for (int m = S.gs[n]; m < S.ge[n]; ++m) {
                    cplx<double> om = omega_s(K, b, m, n, in, ex) + conj(omega_s(K, b, n, m, in, ex));
                    v += K.V(a, n, m) * om;
                }
                S.acc[t] = mk(v.re, 0.0);
            }
            break;
        }
        case WBMI_FORMULA_DEROMEGA: {
            // formula/covariant.py:227-256.  acc holds the complex pre-Hermitian
            // diagonal sum_terms S^{cd}_nn; the trace of S + S^+ is 2 Re.
            if (ex) {
                for (int t = tid; t < nw * 9; t += SC_THREADS) {
                    int n = t / 9, c = (t % 9) / 3, d = t % 3;
                    if (!S.act[n]) continue;
                    // 0.5 dO.nn: O'^{cd}_nn - (D'^d O^c)_nn + (O^c D'^d)_nn
                    cplx<double> t5 = K.Od(c, d, n, n);
                    for (int l = 0; l < nw; ++l) t5 += K.O(c, n, l) * K.Dp(d, l, n) - K.Dp(d, n, l) * K.O(c, l, n);
                    cplx<double> s = 0.5 * t5;
                    // -i sigma sum_{n' in G} A^a_nn' dA.nn^{bd}_n'n
                    for (int sg = 0; sg < 2; ++sg) {
                        int a = sg == 0 ? c_alpha[c] : c_beta[c];
                        int b = sg == 0 ? c_beta[c] : c_alpha[c];
                        double sig = sg == 0 ? 1.0 : -1.0;
                        for (int m = S.gs[n]; m < S.ge[n]; ++m) {
                            cplx<double> y2 = K.Ad(b, d, m, n);
                            for (int l = 0; l < nw; ++l)
                                y2 += K.A(b, m, l) * K.Dp(d, l, n) - K.Dp(d, m, l) * K.A(b, l, n);
                            s += times_i(-(sig * (K.A(a, n, m) * y2)));
                        }
                    }
                    S.acc[t] = s;
                }
            }
            __syncthreads();
            for (int d = 0; d < 3; ++d) {
                for (int b = 0; b < 3; ++b) {
                    // the two components c that pair D^a with index b:
                    //   sigma=+1: beta_c == b, a = alpha_c;  sigma=-1: alpha_c == b, a = beta_c
                    int c1 = -1, c2 = -1;
                    for (int c = 0; c < 3; ++c) {
                        if (c_beta[c] == b) c1 = c;
                        if (c_alpha[c] == b) c2 = c;
                    }
                    if (in) {
                        // X1^{bd}_ln = dD.ln = -dEinv_ln [W + V^b D'^d + V^d D'^b - D'^b V''^d - D'^d V''^b]_ln
                        for (int t = tid; t < nw * nw; t += SC_THREADS) {
                            int l = t / nw, n = t % nw;
                            cplx<double> z = mk(0.0, 0.0);
                            double m = S.M[l * nw + n];
                            if (S.act[n] && m != 0.0) {
                                z = K.W(b, d, l, n);
                                for (int q = 0; q < nw; ++q) z += K.V(b, l, q) * K.Dp(d, q, n) + K.V(d, l, q) * K.Dp(b, q, n);
                                for (int q = S.gs[n]; q < S.ge[n]; ++q)
                                    z -= K.Dp(b, l, q) * K.V(d, q, n) + K.Dp(d, l, q) * K.V(b, q, n);
                                z = (-m) * z;
                            }
                            S.P[t] = z;
                        }
                        __syncthreads();
                        // -i sigma sum_l D'^a_nl X1_ln
                        for (int t = tid; t < nw * 2; t += SC_THREADS) {
                            int n = t / 2, w = t % 2;
                            int c = w == 0 ? c1 : c2;
                            if (!S.act[n]) continue;
                            int a = w == 0 ? c_alpha[c] : c_beta[c];
                            double sig = w == 0 ? 1.0 : -1.0;
                            cplx<double> acc = mk(0.0, 0.0);
                            for (int l = 0; l < nw; ++l) acc += K.Dp(a, n, l) * S.P[l * nw + n];
                            S.acc[n * 9 + c * 3 + d] += times_i(-(sig * acc));
                        }
                        __syncthreads();
                    }
                    if (ex) {
                        // Y1^{bd}_ln = dA.ln = A'^{bd} - D'^d A''^b + A^b D'^d
                        for (int t = tid; t < nw * nw; t += SC_THREADS) {
                            int l = t / nw, n = t % nw;
                            cplx<double> z = mk(0.0, 0.0);
                            if (S.act[n] && S.lab[l] != S.lab[n]) {
                                z = K.Ad(b, d, l, n);
                                for (int q = S.gs[n]; q < S.ge[n]; ++q) z -= K.Dp(d, l, q) * K.A(b, q, n);
                                for (int q = 0; q < nw; ++q) z += K.A(b, l, q) * K.Dp(d, q, n);
                            }
                            S.P[t] = z;
                        }
                        __syncthreads();
                        // -sigma sum_l D'^a_nl Y1_ln
                        for (int t = tid; t < nw * 2; t += SC_THREADS) {
                            int n = t / 2, w = t % 2;
                            int c = w == 0 ? c1 : c2;
                            if (!S.act[n]) continue;
                            int a = w == 0 ? c_alpha[c] : c_beta[c];
                            double sig = w == 0 ? 1.0 : -1.0;
                            cplx<double> acc = mk(0.0, 0.0);
                            for (int l = 0; l < nw; ++l) acc += K.Dp(a, n, l) * S.P[l * nw + n];
                            S.acc[n * 9 + c * 3 + d] -= sig * acc;
                        }
                        __syncthreads();
                    }
                }
                if (ex) {
                    for (int a = 0; a < 3; ++a) {
                        int c1 = -1, c2 = -1;  // sigma=+1: alpha_c == a ; sigma=-1: beta_c == a
                        for (int c = 0; c < 3; ++c) {
                            if (c_alpha[c] == a) c1 = c;
                            if (c_beta[c] == a) c2 = c;
                        }
                        // X2^{ad}_nl = dD.nl = -dEinv_nl [W + V''^a D'^d + V''^d D'^a - D'^a V^d - D'^d V^a]_nl
                        for (int t = tid; t < nw * nw; t += SC_THREADS) {
                            int n = t / nw, l = t % nw;
                            cplx<double> z = mk(0.0, 0.0);
                            double m = S.M[n * nw + l];
                            if (S.act[n] && m != 0.0) {
                                z = K.W(a, d, n, l);
                                for (int q = S.gs[n]; q < S.ge[n]; ++q)
                                    z += K.V(a, n, q) * K.Dp(d, q, l) + K.V(d, n, q) * K.Dp(a, q, l);
                                for (int q = 0; q < nw; ++q) z -= K.Dp(a, n, q) * K.V(d, q, l) + K.Dp(d, n, q) * K.V(a, q, l);
                                z = (-m) * z;
                            }
                            S.P[t] = z;
                        }
                        __syncthreads();
                        // -sigma sum_l X2_nl A^b_ln
                        for (int t = tid; t < nw * 2; t += SC_THREADS) {
                            int n = t / 2, w = t % 2;
                            int c = w == 0 ? c1 : c2;
                            if (!S.act[n]) continue;
                            int b = w == 0 ? c_beta[c] : c_alpha[c];
                            double sig = w == 0 ? 1.0 : -1.0;
                            cplx<double> acc = mk(0.0, 0.0);
                            for (int l = 0; l < nw; ++l) acc += S.P[n * nw + l] * K.A(b, l, n);
                            S.acc[n * 9 + c * 3 + d] -= sig * acc;
                        }
                        __syncthreads();
                    }
                }
            }
            for (int t = tid; t < nw * 9; t += SC_THREADS) S.acc[t] = mk(2.0 * S.acc[t].re, 0.0);
            break;
        }
        default:
            break;
    }
    __syncthreads();
}

__device__ inline void scatter(double* hist, const wbmi_scan_params& p, double E, const double* v, int ncomp,
                               double weight) {
    int bin;
    if (E < p.EFmin) {
        bin = 0;
    } else if (E <= p.EFmax) {
        bin = (int)ceil((E - p.EFmin) / p.dEF);
        if (bin >= p.nEF_extra) return;
    } else {
        return;
    }
    for (int c = 0; c < ncomp; ++c) unsafeAtomicAdd(hist + (int64_t)bin * ncomp + c, weight * v[c]);
}

template <typename T>
__global__ void __launch_bounds__(SC_THREADS) k_static_scan(ScanArgs args, const double* __restrict__ Eall,
                                                            const cplx<T>* __restrict__ Xall,
                                                            double* __restrict__ hist) {
    extern __shared__ char smem_raw[];
    const wbmi_scan_params& p = args.p;
    const int nw = args.nw;
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    Smem S = carve(smem_raw, nw);
    for (int n = tid; n < nw; n += SC_THREADS) S.E[n] = Eall[k * nw + n];
    for (int r = tid; r < WBMI_NROLES; r += SC_THREADS) S.ch[r] = p.ch[r];
    __syncthreads();
    const int ncomp = ncomp_of(p.formula);
    const bool additive = p.formula == WBMI_FORMULA_IDENTITY || p.formula == WBMI_FORMULA_OMEGA ||
                          p.formula == WBMI_FORMULA_DEROMEGA || p.formula == WBMI_FORMULA_VELOMEGA;

    // band groups (grid/tetrahedron.py:132-162, data_K.py:172-186)
    if (tid == 0) {
        int ng = 0;
        S.gb[0] = 0;
        int nb = 1;
        for (int i = 1; i < nw; ++i)
            if (S.E[i] - S.E[i - 1] > p.degen_thresh) S.gb[nb++] = i;
        S.gb[nb++] = nw;
        if (p.degen_Kramers) {
            int m = 0;
            for (int i = 0; i < nb; ++i)
                if (S.gb[i] % 2 == 0) S.gb[m++] = S.gb[i];
            nb = m;
        }
        ng = nb - 1;
        int first_in = -1;
        for (int g = 0; g < ng; ++g) {
            int a = S.gb[g], b = S.gb[g + 1];
            int inr = (S.E[b - 1] >= p.EFmin) && (S.E[a] <= p.EFmax);
            S.gin[g] = inr;
            if (inr && first_in < 0) first_in = a;
            double sum = 0.0;
            for (int i = a; i < b; ++i) sum += S.E[i];
            S.gw[g] = sum / (b - a);
        }
        int bandmax = 0;
        if (p.fder == 0) {
            for (int i = 0; i < nw; ++i)
                if (S.E[i] < p.EFmin) bandmax = i + 1;
            if (first_in >= 0 && first_in < bandmax) bandmax = first_in;
        }
        S.misc[0] = ng;
        S.misc[1] = bandmax;
    }
    __syncthreads();
    const int ng = S.misc[0];
    const int bandmax = S.misc[1];

    KCtx<T> K;
    K.X = Xall + k * (int64_t)args.nch * nw * nw;
    K.nw = nw;
    K.ch = S.ch;
    K.E = S.E;
    K.M = S.M;
    K.lab = S.lab;
    K.gs = S.gs;
    K.ge = S.ge;

    double v[9];
    if (additive) {
        // main partition: every group at once
        for (int g = 0; g < ng; ++g)
            for (int n = S.gb[g] + tid; n < S.gb[g + 1]; n += SC_THREADS) {
                S.lab[n] = g;
                S.gs[n] = S.gb[g];
                S.ge[n] = S.gb[g + 1];
                S.act[n] = S.gin[g];
            }
        __syncthreads();
        eval_partition(S, K, p);
        if (tid < ng && S.gin[tid]) {
            int g = tid;
            for (int c = 0; c < ncomp; ++c) {
                double s = 0.0;
                for (int n = S.gb[g]; n < S.gb[g + 1]; ++n) s += S.acc[n * 9 + c].re;
                v[c] = s;
            }
            scatter(hist, p, S.gw[g], v, ncomp, args.weight);
        }
        __syncthreads();
        if (bandmax > 0) {
            // Fermi-sea group (0, bandmax) with weight -inf
            for (int n = tid; n < nw; n += SC_THREADS) {
                int in = n < bandmax;
                S.lab[n] = in ? 0 : 1;
                S.gs[n] = in ? 0 : bandmax;
                S.ge[n] = in ? bandmax : nw;
                S.act[n] = in;
            }
            __syncthreads();
            eval_partition(S, K, p);
            if (tid == 0) {
                for (int c = 0; c < ncomp; ++c) {
                    double s = 0.0;
                    for (int n = 0; n < bandmax; ++n) s += S.acc[n * 9 + c].re;
                    v[c] = s;
                }
                scatter(hist, p, -INFINITY, v, ncomp, args.weight);
            }
            __syncthreads();
        }
    } else {
        // non-additive: traces of the cumulative cuts [0, c) (static.py:109-117)
        for (int cidx = 1; cidx <= nw; ++cidx) {
            bool need = (bandmax == cidx);
            for (int g = 0; g < ng; ++g)
                if (S.gin[g] && (S.gb[g] == cidx || S.gb[g + 1] == cidx)) need = true;
            if (!need) continue;
            for (int n = tid; n < nw; n += SC_THREADS) {
                int in = n < cidx;
                S.lab[n] = in ? 0 : 1;
                S.gs[n] = in ? 0 : cidx;
                S.ge[n] = in ? cidx : nw;
                S.act[n] = in;
            }
            __syncthreads();
            eval_partition(S, K, p);
            if (tid == 0)
                for (int c = 0; c < ncomp; ++c) {
                    double s = 0.0;
                    for (int n = 0; n < cidx; ++n) s += S.acc[n * 9 + c].re;
                    S.cut[cidx * 9 + c] = s;
                }
            __syncthreads();
        }
        if (tid == 0) {
            for (int c = 0; c < ncomp; ++c) S.cut[c] = 0.0;
            for (int g = 0; g < ng; ++g) {
                if (!S.gin[g]) continue;
                for (int c = 0; c < ncomp; ++c) v[c] = S.cut[S.gb[g + 1] * 9 + c] - S.cut[S.gb[g] * 9 + c];
                scatter(hist, p, S.gw[g], v, ncomp, args.weight);
            }
            if (bandmax > 0) {
                for (int c = 0; c < ncomp; ++c) v[c] = S.cut[bandmax * 9 + c];
                scatter(hist, p, -INFINITY, v, ncomp, args.weight);
            }
        }
    }
}

static int check_roles(wbmi_ctx* ctx, const wbmi_scan_params* p, int nch) {
    auto need = [&](int base, int n) -> int {
        for (int i = 0; i < n; ++i)
            if (p->ch[base + i] < 0 || p->ch[base + i] >= nch)
                return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: channel role %d missing or out of range",
                                 base + i);
        return WBMI_OK;
    };
    int rc = WBMI_OK;
    bool in = p->internal_terms, ex = p->external_terms;
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
        case WBMI_FORMULA_MORB_HPM:
            if ((rc = need(WBMI_ROLE_V, 3))) return rc;
            if (ex && ((rc = need(WBMI_ROLE_A, 3)) || (rc = need(WBMI_ROLE_O, 3)) || (rc = need(WBMI_ROLE_B, 3)) ||
                       (rc = need(WBMI_ROLE_C, 3))))
                return rc;
            break;
        default:
            return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_static_scan: formula %d is not available", p->formula);
    }
    (void)in;
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
    int rc = check_roles(ctx, params, nch);
    if (rc) return rc;
    ScanArgs a;
    a.p = *params;
    a.nw = nw;
    a.nch = nch;
    a.weight = weight;
    size_t lds = smem_bytes(nw);
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_static_scan<double>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_static_scan<float>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (prec == WBMI_F64)
        hipLaunchKernelGGL(k_static_scan<double>, dim3(nk), dim3(SC_THREADS), lds, ctx->stream, a, E_dev,
                           (const cplx<double>*)Xbar_dev, hist_dev);
    else if (prec == WBMI_F32)
        hipLaunchKernelGGL(k_static_scan<float>, dim3(nk), dim3(SC_THREADS), lds, ctx->stream, a, E_dev,
                           (const cplx<float>*)Xbar_dev, hist_dev);
    else
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: bad precision");
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// ---------------------------------------------------------------------------
// histogram -> result (static.py:137-155)
// ---------------------------------------------------------------------------

__global__ void k_finalize(const double* __restrict__ hist, int nEF, int ncomp, int fder, double dEF, double scale,
                           double* __restrict__ out) {
    extern __shared__ double cum[];  // [nEF][ncomp] prefix sums
    for (int c = threadIdx.x; c < ncomp; c += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < nEF; ++i) {
            s += hist[i * ncomp + c];
            cum[i * ncomp + c] = s;
        }
    }
    __syncthreads();
    int extra = fder == 0 ? 0 : (fder == 3 ? 2 : 1);
    int nout = nEF - 2 * extra;
    for (int t = threadIdx.x; t < nout * ncomp; t += blockDim.x) {
        int i = t / ncomp, c = t % ncomp;
        const double* r = cum + c;
        double v;
        if (fder == 0) v = r[i * ncomp];
        else if (fder == 1) v = (r[(i + 2) * ncomp] - r[i * ncomp]) / (2 * dEF);
        else if (fder == 2) v = (r[(i + 2) * ncomp] + r[i * ncomp] - 2 * r[(i + 1) * ncomp]) / (dEF * dEF);
        else v = (r[(i + 4) * ncomp] - r[i * ncomp] - 2 * (r[(i + 3) * ncomp] - r[(i + 1) * ncomp])) / (2 * dEF * dEF * dEF);
        out[t] = v * scale;
    }
}

extern "C" int wbmi_finalize(wbmi_ctx* ctx, const double* hist_dev, int nEF_extra, int ncomp, int fder, double dEF,
                             double scale, double* out_dev) {
    if (!ctx) return WBMI_ERR_ARG;
    int extra = fder == 0 ? 0 : (fder == 3 ? 2 : 1);
    if (!hist_dev || !out_dev || ncomp <= 0 || fder < 0 || fder > 3 || nEF_extra - 2 * extra <= 0)
        return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize: bad arguments");
    size_t lds = sizeof(double) * (size_t)nEF_extra * ncomp;
    if (lds > 160 * 1024) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_finalize: too many Fermi levels for one pass");
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_finalize, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)lds));
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), lds, ctx->stream, hist_dev, nEF_extra, ncomp, fder, dEF, scale,
                       out_dev);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}
