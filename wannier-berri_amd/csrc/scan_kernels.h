// Covariant formula traces + Fermi-level scan of libwbmi.
//
// Replaces the per-(k, band group) loop of StaticCalculator.__call__
// (calculators/static.py:60-136).  One workgroup per k-point (512 threads;
// DerOmega / DerMorb: scan_threads); the formula is a template parameter, so
// every calculator gets its own kernel.
//
// Band partition.  The reference evaluates, for every band group G, the
// formula with inn = G and out = all other bands.  Every sum in the formulas
// is then either over l in out (l not in G) or over n' in G, so with a label
// per band (lab[x] = group of x) the traces of ALL groups come from one set
// of per-band values v_n, n in G, using the masks
//     Off(x, y) = lab[x] != lab[y]      (x, y in different groups)
//     On(x, y)  = lab[x] == lab[y]
// and the masked matrices
//     D'^a = -V^a o M,  M = dEinv o Off,  V'' = V o On,  A'' = A o On.
// The Fermi-sea group (all bands below EFmin, data_K.py:178-185) and the
// cumulative cuts [0, c) of non-additive formulas (static.py:109-117) are
// two-set partitions evaluated the same way.
//
// DerOmega.  With G^{bd} = V^b D'^d, H^{ad} = D'^a V^d and K^{bd} = A^b D'^d,
//   dD.ln = X1^{bd} = -M o [W^{bd} + G^{bd} + G^{db} - D'^b V''^d - D'^d V''^b]
//   dD.nl = X2^{ad} = -M o [W^{ad} + V''^a D'^d + V''^d D'^a - H^{ad} - H^{da}]
//   dA.ln = Y1^{bd} = A'^{bd} - D'^d A''^b + K^{bd}
// (formula/elementary.py:63-70, formula/formula.py:108-118).  X1 and X2 are
// symmetric in (b, d), so the elementwise terms are folded half-and-half into
// each product: Q1 = G + (W - S)/2, Q2 = -H + (W + R)/2, Y1 = K + A' - D'A''.
// The traces use only diagonal elements of D'^a X1, X2 A^b and D'^a Y1, so
// every product element is consumed where it is computed: a wave task forms
// one 16 x 32 block of one of the products on the FP64 matrix cores, adds the
// elementwise terms, multiplies by the consumer weights, reduces over the
// summed index with cross-lane shuffles and adds the result to the per-band
// rows with LDS atomics.  The waves take tasks from an LDS counter.  No
// product is stored and no barrier separates them.
//
// For DerOmega the operands V and A stay in LDS (rows padded by one element)
// when nw <= 32.
// All arithmetic is FP64 (energies, dEinv, traces, histogram); channels may
// be stored in FP32 or FP64.
#pragma once

#include <cmath>
#include <type_traits>

#include "mfma_c64.h"
#include "wbmi_common.h"

#define SC_THREADS 512            // workgroup size (and launch bound) of the scan kernels
#ifndef SC_TASK_THREADS
#define SC_TASK_THREADS 768       // DerOmega / DerMorb: 12 waves (3 per SIMD), one workgroup per CU
#endif
#define SC_NWMAX 128
#ifndef SC_STAGE_A
#define SC_STAGE_A 1  // A staged in LDS next to V (1) or read from the channels (0)
#endif
#define SC_M_NW 64    // dEinv o Off lives in LDS up to this nw, else it is formed on the fly
#define SC_LDS_NW 32  // V and A live in LDS up to this nw

static __constant__ int c_alpha[3] = {1, 2, 0};  // utility.py:45-46
static __constant__ int c_beta[3] = {2, 0, 1};

struct ScanArgs {
    wbmi_scan_params p;
    int nw, nch, stage_va;
    double weight;
    // tetrahedron method: `tetra` = number of corners, 8 (parallelepiped: corner ix*4 + iy*2 + iz
    // at k + (2i-1) dK/2, 12 tetrahedra around the centre) or 4 (tetrahedral K-block: the vertices,
    // one tetrahedron); corner energies [tetra][nk][nw], ascending Fermi levels Ef[nEF]; hist holds
    // nEF step rows (prefix-summed later) followed by nEF rows added as they are
    int tetra, nEF;
    const double* Ec;
    int64_t ec_stride;
    const double* Ef;
    // k-resolved scans: k-point k accumulates into hist + k * hist_stride (0: one histogram)
    int64_t hist_stride;
};

// dynamic shared-memory layout
struct Smem {
    char* V;            // [3][nw][ld] staged V (element type scan_stage)
    char* A;            // [3][nw][ld] staged A (external terms)
    cplx<double>* acc;  // [nw][9]
    double* M;          // [nw][ld] dEinv o Off
    double* E;
    double* cut;        // [nw+1][9]
    double* gw;         // group weights (mean energy)
    int* lab;
    int* gs;
    int* ge;
    int* act;
    int* gb;    // group borders [nw+1]
    int* gin;   // group in range
    int* misc;  // [0] = ngroups, [1], [2] = sea group [lo, hi) (empty: lo == hi)
    int* ch;    // channel slot per role [WBMI_NROLES]
    // tetrahedron method only (empty regions otherwise)
    double* tC;    // [nw][8] corner energies
    double* tEmn;  // [nw] min over centre and corners
    double* tEmx;  // [nw] max over centre and corners
    double* tv;    // [nw+1][9] group values
    int* tg;       // [nw] group of each band
    int* tlo;      // [nw] first Fermi level of the band's partial range
    int* tcnt;     // [nw] length of that range
    int* tpre;     // [nw+1] prefix sums of tcnt
};

__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }

// byte offsets of the LDS regions and their total size
struct SmemOffsets {
    size_t V, A, acc, M, E, cut, gw, ints, tet, tints, total;
};

// vsz: bytes of one staged V / A element
__host__ __device__ inline SmemOffsets smem_offsets(int nw, bool stage_v, bool stage_a, bool tetra, size_t vsz) {
    const size_t ld = nw + 1;
    SmemOffsets o;
    size_t off = 0;
    o.V = off;
    off = al16(off + (stage_v ? vsz * 3 * nw * ld : 0));
    o.A = off;
    off = al16(off + (stage_a ? vsz * 3 * nw * ld : 0));
    o.acc = off;
    off = al16(off + sizeof(cplx<double>) * nw * 9);
    o.M = off;
    off = al16(off + (nw <= SC_M_NW ? sizeof(double) * nw * ld : 0));
    o.E = off;
    off = al16(off + sizeof(double) * nw);
    o.cut = off;
    off = al16(off + sizeof(double) * (nw + 1) * 9);
    o.gw = off;
    off = al16(off + sizeof(double) * nw);
    o.ints = off;
    off = al16(off + sizeof(int) * (6 * nw + 1 + 4 + WBMI_NROLES));
    o.tet = off;
    off = al16(off + (tetra ? sizeof(double) * (8 * nw + 2 * nw + (nw + 1) * 9) : 0));
    o.tints = off;
    off = al16(off + (tetra ? sizeof(int) * (4 * nw + 1) : 0));
    o.total = off;
    return o;
}

__device__ __attribute__((always_inline)) inline Smem make_smem(char* base, int nw, bool stage_v, bool stage_a,
                                                                bool tetra, size_t vsz) {
    const SmemOffsets o = smem_offsets(nw, stage_v, stage_a, tetra, vsz);
    Smem s;
    // LDS addresses even where a region is empty (never read then): no null tests on LDS pointers
    s.V = base + o.V;
    s.A = base + o.A;
    s.acc = (cplx<double>*)(base + o.acc);
    s.M = (double*)(base + o.M);  // empty region (never read) when nw > SC_M_NW
    s.E = (double*)(base + o.E);
    s.cut = (double*)(base + o.cut);
    s.gw = (double*)(base + o.gw);
    int* q = (int*)(base + o.ints);
    s.lab = q;
    s.gs = q + nw;
    s.ge = q + 2 * nw;
    s.act = q + 3 * nw;
    s.gb = q + 4 * nw;
    s.gin = q + 5 * nw + 1;
    s.misc = q + 6 * nw + 1;
    s.ch = q + 6 * nw + 5;
    double* t = (double*)(base + o.tet);
    s.tC = t;
    s.tEmn = t + 8 * nw;
    s.tEmx = t + 9 * nw;
    s.tv = t + 10 * nw;
    int* ti = (int*)(base + o.tints);
    s.tg = ti;
    s.tlo = ti + nw;
    s.tcnt = ti + 2 * nw;
    s.tpre = ti + 3 * nw;
    return s;
}

// Element access to the channels of one k-point.
// LM: dEinv o Off read from LDS (nw <= SC_M_NW) or formed from E and the labels.
// ST: V read from its LDS copy (DerOmega, DerMorb and Omega at nw <= SC_LDS_NW) or from the
// channels; SA: the same for A (staged together with V, and only with external terms, the only
// case that reads it).
// Both are compile-time so that every access is a plain LDS or global load.
template <typename T, bool LM, bool ST, bool SA, int NW = 0, typename TS = T>
struct KCtx {
    const cplx<T>* X;  // channels of this k: [nch][nw][nw]
    int nw, ld;  // run-time size (NW = 0); NW > 0 fixes both at compile time
    __device__ inline int n_() const {
        if constexpr (NW > 0) return NW;
        else return nw;
    }
    __device__ inline int ld_() const {
        if constexpr (NW > 0) return NW + 1;
        else return ld;
    }
    const int* ch;
    const double* E;
    const double* M;  // dEinv o Off, row stride ld
    const int* lab;
    const int* gs;
    const int* ge;
    const cplx<TS>* sV;  // LDS copies of V and A (ST), element type TS
    const cplx<TS>* sA;
    const double* bv;    // WBMI_FORMULA_BANDVALUES: this k's caller values [nw][nbv]
    int nbv;
    __device__ inline cplx<double> get(int role, int x, int y) const {
        cplx<T> v = X[((int64_t)ch[role] * n_() + x) * n_() + y];
        return mk((double)v.re, (double)v.im);
    }
    __device__ inline cplx<double> V(int a, int x, int y) const {
        if constexpr (ST) {
            const cplx<TS> v = sV[(a * n_() + x) * ld_() + y];
            return mk((double)v.re, (double)v.im);
        } else {
            return get(WBMI_ROLE_V + a, x, y);
        }
    }
    __device__ inline cplx<double> A(int a, int x, int y) const {
        if constexpr (SA) {
            const cplx<TS> v = sA[(a * n_() + x) * ld_() + y];
            return mk((double)v.re, (double)v.im);
        } else {
            return get(WBMI_ROLE_A + a, x, y);
        }
    }
    __device__ inline cplx<double> W(int a, int b, int x, int y) const { return get(WBMI_ROLE_W + 3 * a + b, x, y); }
    __device__ inline cplx<double> Ad(int a, int d, int x, int y) const { return get(WBMI_ROLE_AD + 3 * a + d, x, y); }
    __device__ inline cplx<double> O(int c, int x, int y) const { return get(WBMI_ROLE_O + c, x, y); }
    __device__ inline cplx<double> Od(int c, int d, int x, int y) const { return get(WBMI_ROLE_OD + 3 * c + d, x, y); }
    __device__ inline cplx<double> B(int a, int x, int y) const { return get(WBMI_ROLE_B + a, x, y); }
    __device__ inline cplx<double> C(int a, int x, int y) const { return get(WBMI_ROLE_C + a, x, y); }
    __device__ inline cplx<double> Bd(int a, int d, int x, int y) const { return get(WBMI_ROLE_BD + 3 * a + d, x, y); }
    __device__ inline cplx<double> Cd(int c, int d, int x, int y) const { return get(WBMI_ROLE_CD + 3 * c + d, x, y); }
    __device__ inline cplx<double> S(int b, int x, int y) const { return get(WBMI_ROLE_S + b, x, y); }
    __device__ inline cplx<double> Sd(int b, int d, int x, int y) const { return get(WBMI_ROLE_SD + 3 * b + d, x, y); }
    // X o On: elements between bands of the same group
    __device__ inline cplx<double> Ain(int a, int x, int y) const { return lab[x] == lab[y] ? A(a, x, y) : mk(0.0, 0.0); }
    __device__ inline cplx<double> Vin(int a, int x, int y) const { return lab[x] == lab[y] ? V(a, x, y) : mk(0.0, 0.0); }
    // dEinv_xy Off(x, y) (data_K.py:290-298): from LDS, or formed from E and the labels
    __device__ inline double m(int x, int y) const {
        if constexpr (LM) {
            return M[x * ld_() + y];
        } else {
            const double de = E[x] - E[y];
            return (lab[x] != lab[y] && fabs(de) >= 1e-7) ? 1.0 / de : 0.0;
        }
    }
    // D'^a_xy = -V^a_xy dEinv_xy Off(x, y)  (branch-free: M = 0 gives an exact zero for finite V)
    __device__ inline cplx<double> Dp(int a, int x, int y) const { return (-m(x, y)) * V(a, x, y); }
};

// pre-Hermitian Omega summand s^c_xy for x, y in the same group
// (formula/covariant.py:175-200): Omega.nn = s + s^+.  The sum over the
// off-group bands l runs over [l0, l1); `tail` adds the within-group and O terms.
template <typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) cplx<double> omega_s(const KCtx<T, LM, ST, SA, NW, TS>& K, int c, int x, int y, bool internal,
                                                              bool external, int l0, int l1, bool tail) {
    const int al = c_alpha[c], be = c_beta[c];
    cplx<double> s = mk(0.0, 0.0);
    for (int l = l0; l < l1; ++l) {
        if (K.lab[l] == K.lab[x]) continue;
        cplx<double> dxa = K.Dp(al, x, l);
        if (internal) s += times_i(-(dxa * K.Dp(be, l, y)));
        if (external) s += K.Dp(be, x, l) * K.A(al, l, y) - dxa * K.A(be, l, y);
    }
    if (external && tail) {
        for (int m = K.gs[x]; m < K.ge[x]; ++m) s += times_i(-(K.A(al, x, m) * K.A(be, m, y)));
        s += 0.5 * K.O(c, x, y);
    }
    return s;
}
template <typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) cplx<double> omega_s(const KCtx<T, LM, ST, SA, NW, TS>& K, int c, int x, int y, bool internal,
                                                              bool external) {
    return omega_s(K, c, x, y, internal, external, 0, K.n_(), true);
}

// pre-Hermitian Morb_H summand h^c_xy for x, y in the same group
// (formula/covariant.py:390-414)
template <typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) cplx<double> morbh_s(const KCtx<T, LM, ST, SA, NW, TS>& K, int c, int x, int y, bool internal, bool external) {
    const int al = c_alpha[c], be = c_beta[c];
    cplx<double> s = mk(0.0, 0.0);
    for (int l = 0; l < K.n_(); ++l) {
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

// Group block of the inverse effective mass (formula/elementary.py:28-33 through
// formula/formula.py:95-107 with A = V, dA = W): for x, y in one band group,
//   M^{bd}_xy = W^{bd}_xy - sum_{l off} D^d_xl V^b_ly + sum_{l off} V^b_xl D^d_ly
// (D' carries the off-group mask, so the sums run over every l).
template <typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) cplx<double> mass_xy(const KCtx<T, LM, ST, SA, NW, TS>& K, int b, int d, int x,
                                                               int y) {
    cplx<double> v = K.W(b, d, x, y);
    for (int l = 0; l < K.n_(); ++l) {
        v -= K.Dp(d, x, l) * K.V(b, l, y);
        v += K.V(b, x, l) * K.Dp(d, l, y);
    }
    return v;
}

// Contraction of a rank-4 product X[a][b][c][d] to the Hall tensor [j][i]
// (calculators/static.py:419-423): 0.5 sum over (a, b) = (al_j, be_j) (+), (be_j, al_j) (-)
// and (c, d) = (be_i, al_i) (+), (al_i, be_i) (-).  `term(a, b, c, d)` gives X.
template <typename Fn>
__device__ __attribute__((always_inline)) cplx<double> hall_contract(int j, int i, Fn term) {
    const int aj = c_alpha[j], bj = c_beta[j], ai = c_alpha[i], bi = c_beta[i];
    cplx<double> v = term(aj, bj, bi, ai);
    v -= term(aj, bj, ai, bi);
    v -= term(bj, aj, bi, ai);
    v += term(bj, aj, ai, bi);
    return mk(0.5 * v.re, 0.5 * v.im);
}

template <int F>
__host__ __device__ constexpr int ncomp_of() {
    return F == WBMI_FORMULA_IDENTITY ? 1
           : (F == WBMI_FORMULA_OMEGA || F == WBMI_FORMULA_MORB_HPM || F == WBMI_FORMULA_SPIN) ? 3
                                                                                                : 9;
}

template <int F>
__host__ __device__ constexpr bool uses_A() {
    return F == WBMI_FORMULA_OMEGA || F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_VELOMEGA ||
           F == WBMI_FORMULA_MORB_HPM || F == WBMI_FORMULA_VELHPLUS || F == WBMI_FORMULA_DERMORB ||
           F == WBMI_FORMULA_OMEGASPIN || F == WBMI_FORMULA_OMEGAOMEGA || F == WBMI_FORMULA_OMEGAHPLUS;
}

// The two curl components c whose index pair (alpha_c, beta_c) contains x.
struct Pair {
    int c, partner;
    double sig;
};
// x in the second slot for sigma=+1: (sigma=+1: beta_c == x, a = alpha_c; sigma=-1: alpha_c == x, a = beta_c)
__device__ inline Pair pair_second(int x, int which) {
    for (int c = 0; c < 3; ++c) {
        if (which == 0 && c_beta[c] == x) return Pair{c, c_alpha[c], 1.0};
        if (which == 1 && c_alpha[c] == x) return Pair{c, c_beta[c], -1.0};
    }
    return Pair{0, 0, 0.0};
}
// x in the first slot for sigma=+1: (sigma=+1: alpha_c == x, b = beta_c; sigma=-1: beta_c == x, b = alpha_c)
__device__ inline Pair pair_first(int x, int which) {
    for (int c = 0; c < 3; ++c) {
        if (which == 0 && c_alpha[c] == x) return Pair{c, c_beta[c], 1.0};
        if (which == 1 && c_beta[c] == x) return Pair{c, c_alpha[c], -1.0};
    }
    return Pair{0, 0, 0.0};
}

// build M = dEinv o Off for the current labels (data_K.py:290-298)
__device__ __attribute__((always_inline)) void build_M(const Smem& S, int nw) {
    const int ld = nw + 1;
    for (int t = threadIdx.x; t < nw * nw; t += (int)blockDim.x) {
        int x = t / nw, y = t % nw;
        double m = 0.0;
        if (S.lab[x] != S.lab[y]) {
            double de = S.E[x] - S.E[y];
            if (fabs(de) >= 1e-7) m = 1.0 / de;
        }
        S.M[x * ld + y] = m;
    }
}

// sum of v over the lanes of a wave that differ only in the lane bits of `mask`
__device__ inline double wave_sum(double v, int mask) {
    for (int b = 1; b < 64; b <<= 1)
        if (mask & b) v += __shfl_xor(v, b);
    return v;
}

// Wave tasks.  One wave forms the 16-row x 32-column block (rh, cb) of a dense product
// P = X Y on the matrix cores (FP64 MFMA, mfma_c64.h), adds the elementwise terms of
// each element (x, y) and accumulates the consumers' weighted elements into
// the workgroup's per-band rows acc[band][comp] (LDS FP64 atomic adds, .re) after a
// shuffle reduction over the summed index.  The band of an element is its column y (kind 1 forms its
// product transposed for this).  Kinds (task indices b, d; for kinds 5, 6 b = c):
//   0  Q1^{bd} = V^b D'^d + (W - D'^b V''^d - D'^d V''^b)/2, y in the partition, x off-side:
//      DerOmega T1; DerMorb s2 (x 2 E_x), s7 (conj B^a_xy), s1 (raw V^b D'^d)
//   1  Q2^{bd} = -D'^b V^d + (W + V''^b D'^d + V''^d D'^b)/2, x in, y off-side: DerOmega T3
//   2  Y1^{bd} = A^b D'^d + A' - D'^d A''^b, x off-side: DerOmega T2;
//      x same-side (DerMorb): s5 terms of A' + A^b D'^d
//   3  YB^{bd} = B^b D'^d + B' - D'^d B''^b, x off-side: DerMorb s6
//   4  DA^{bd} = D'^d A^b, same side: DerMorb s5 (-D'^d A^b) and the Omega block (-D'^al A^be + D'^be A^al)
//   5  DD^c = D'^al D'^be, same side: Omega block, internal term
//   6  AA^c = A''^al A''^be, same side: Omega block, external term
//   7  VA^{d be} = V''^d A''^be, same side: DerMorb s4
// DerOmega contributions carry the band weight dw(n) = 2 (DerOmega) or
// 2 sign E_n (DerMorb's Eav o dOmega term; only cut partitions are summed, so a
// same-side element may be credited to either of its bands).
struct TaskCtx {
    bool dm, in, ex;
    double sign;
};

template <int KIND>
__host__ __device__ constexpr int task_nt() {
    return KIND == 0 ? 5 : KIND == 1 ? 4 : KIND == 2 ? 2 : KIND == 3 ? 2 : KIND == 4 ? 5 : KIND == 7 ? 1 : 3;
}

// M1 (kind 0 only): the task also carries the kind-1 task (d, b).  With V
// Hermitian and D' anti-Hermitian, (D'^d V^b)_yx = -conj((V^b D'^d)_xy), so
// kind 1's transposed product is -conj of kind 0's at the same element, and
// its elementwise channel W^{db}_yx = conj(W^{bd}_xy): one product and one
// load serve both consumers.
template <int KIND, bool M1, typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) void formula_task(const Smem& S, const KCtx<T, LM, ST, SA, NW, TS>& K, const TaskCtx& tc_,
                                                           cplx<double>* acc, int b, int d, int rh, int cb) {
    const int nw = K.n_(), lane = threadIdx.x & 63;
    const int r0 = rh * 16, c0 = cb * 32;  // a 16 x 32 half block
    const bool dm = tc_.dm, in = tc_.in, ex = tc_.ex;
    const double sign = tc_.sign;
    const int al = c_alpha[b], be = c_beta[b];  // kinds 5, 6: b = c
    MfmaHalf P;
    P.zero();
    if constexpr (KIND == 0)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.V(b, r, m); }, [&](int m, int c) { return K.Dp(d, m, c); });
    else if constexpr (KIND == 1)  // transposed: P_xy = (D'^b V^d)_yx, so the band n = y is a column as for the others
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.V(d, m, r); }, [&](int m, int c) { return K.Dp(b, c, m); });
    else if constexpr (KIND == 2)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.A(b, r, m); }, [&](int m, int c) { return K.Dp(d, m, c); });
    else if constexpr (KIND == 3)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.B(b, r, m); }, [&](int m, int c) { return K.Dp(d, m, c); });
    else if constexpr (KIND == 4)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.Dp(d, r, m); }, [&](int m, int c) { return K.A(b, m, c); });
    else if constexpr (KIND == 5)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.Dp(al, r, m); }, [&](int m, int c) { return K.Dp(be, m, c); });
    else if constexpr (KIND == 6)
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.Ain(al, r, m); }, [&](int m, int c) { return K.Ain(be, m, c); });
    else
        mfma_half(P, nw, nw, nw, r0, c0, [&](int r, int m) { return K.Vin(b, r, m); }, [&](int m, int c) { return K.Ain(d, m, c); });

    // targets: partner component tp, accumulator column tcol, factor tw
    constexpr int NT = task_nt<KIND>();
    int tp[NT], tcol[NT];
    double tw[NT];
#pragma unroll
    for (int w = 0; w < NT; ++w) {
        tp[w] = 0;
        tcol[w] = 0;
        tw[w] = 0.0;
    }
    if constexpr (KIND == 0 || KIND == 1) {
        // the two index pairs of X1 / X2: targets coincide for b == d (one lane adds twice)
        const int nt = b == d ? 2 : 4;
        const double f = b == d ? 2.0 : 1.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const Pair p = KIND == 1 ? pair_first(w < 2 ? b : d, w % 2) : pair_second(w < 2 ? b : d, w % 2);
            tp[w] = p.partner;
            tcol[w] = p.c * 3 + (w < 2 ? d : b);
            tw[w] = w < nt ? f * p.sig : 0.0;
        }
        if constexpr (KIND == 0) {
            // s1: the c with beta_c = d, partner alpha_c, column (c, b)
            for (int c = 0; c < 3; ++c)
                if (c_beta[c] == d) {
                    tp[4] = c_alpha[c];
                    tcol[4] = c * 3 + b;
                    tw[4] = 2.0;
                }
        }
    } else if constexpr (KIND == 2 || KIND == 3) {
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const Pair p = pair_second(b, w);
            tp[w] = p.partner;
            tcol[w] = p.c * 3 + d;
            tw[w] = p.sig;
        }
    } else if constexpr (KIND == 4) {
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const Pair p = pair_second(b, w);
            tp[w] = p.partner;
            tcol[w] = p.c * 3 + d;
            tw[w] = p.sig;
        }
        // Omega block: D'^d A^b enters s^c with -1 for (al_c, be_c) = (d, b), +1 for (b, d)
        for (int c = 0; c < 3; ++c) {
            const double coef = (c_alpha[c] == d && c_beta[c] == b) ? -1.0 : (c_alpha[c] == b && c_beta[c] == d) ? 1.0 : 0.0;
            if (coef != 0.0)
                for (int e = 0; e < 3; ++e) {
                    tp[2 + e] = e;
                    tcol[2 + e] = c * 3 + e;
                    tw[2 + e] = 2.0 * sign * coef;
                }
        }
    } else if constexpr (KIND == 5 || KIND == 6) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            tp[e] = e;
            tcol[e] = b * 3 + e;
            tw[e] = 2.0 * sign;
        }
    } else {
        for (int c = 0; c < 3; ++c)
            if (c_beta[c] == d) {
                tp[0] = c_alpha[c];
                tcol[0] = c * 3 + b;
                tw[0] = 2.0;
            }
    }
    // targets of the carried kind-1 task (b1, d1) = (d, b)
    constexpr int NT1 = M1 ? 4 : 1;
    const int b1 = d, d1 = b;
    int tp1[NT1], tcol1[NT1];
    double tw1[NT1];
#pragma unroll
    for (int w = 0; w < NT1; ++w) {
        tp1[w] = 0;
        tcol1[w] = 0;
        tw1[w] = 0.0;
    }
    if constexpr (M1) {
        const int nt = b1 == d1 ? 2 : 4;
        const double f = b1 == d1 ? 2.0 : 1.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const Pair p = pair_first(w < 2 ? b1 : d1, w % 2);
            tp1[w] = p.partner;
            tcol1[w] = p.c * 3 + (w < 2 ? d1 : b1);
            tw1[w] = w < nt ? f * p.sig : 0.0;
        }
    }

    // part[u][w]: u = the lane's band slot, its column bj (the band n is the column y of every element)
    constexpr int NU = 2;
    double part[NU][NT], part1[NU][NT1];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
#pragma unroll
        for (int w = 0; w < NT; ++w) part[u][w] = 0.0;
#pragma unroll
        for (int w = 0; w < NT1; ++w) part1[u][w] = 0.0;
    }
    {
        // the elementwise channel (W, A', B') of this row half: all loads issued before use
        cplx<double> gz[2][4];
        if constexpr (KIND <= 3) {
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int x = r0 + mfma_hrow(v), y = c0 + mfma_col(bj);
                    const bool ok = x < nw && y < nw;
                    const int xx = ok ? x : 0, yy = ok ? y : 0;
                    if constexpr (KIND == 0)
                        gz[bj][v] = K.W(b, d, xx, yy);  // (l, n)
                    else if constexpr (KIND == 1)
                        gz[bj][v] = conj(K.W(b, d, xx, yy));  // (n, l) of the Hermitian W, read along rows
                    else if constexpr (KIND == 2)
                        gz[bj][v] = ex ? K.Ad(b, d, xx, yy) : mk(0.0, 0.0);
                    else
                        gz[bj][v] = K.Bd(b, d, xx, yy);
                }
        }
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int x = r0 + mfma_hrow(v), y = c0 + mfma_col(bj);
            if (x >= nw || y >= nw) continue;
            const int i = bj, j = bj;  // part slot
            const cplx<double> pv = P.get(bj, v);
            const int n = y, l = x;
            if (!S.act[n]) continue;
            const bool same = K.lab[l] == K.lab[n];
            const bool single = S.ge[n] - S.gs[n] == 1;  // group {n}: the within-group sums have one term
            const double dw = dm ? 2.0 * sign * K.E[n] : 2.0;
            if constexpr (KIND == 0) {
                if (same) continue;
                // (W^{bd} - D'^b V''^d - D'^d V''^b)_ln / 2, with D'_xy = -m_xy V_xy and m_nl = -m_ln
                const double mln = K.m(l, n);
                cplx<double> z = gz[bj][v];
                if (single)
                    z += mln * (K.V(b, l, n) * K.V(d, n, n) + K.V(d, l, n) * K.V(b, n, n));
                else
                    for (int q = S.gs[n]; q < S.ge[n]; ++q)
                        z -= K.Dp(b, l, q) * K.V(d, q, n) + K.Dp(d, l, q) * K.V(b, q, n);
                const cplx<double> Pm = (-mln) * (pv + 0.5 * z);
                const double wD = (in ? dw : 0.0) + (dm && in ? 2.0 * K.E[l] : 0.0);
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const cplx<double> Dnl = mln * K.V(tp[w], n, l);
                    double t = wD * (Dnl * Pm).im;  // Re(-i D' X1) (T1, s2)
                    if (dm && ex) t -= 2.0 * (conj(K.B(tp[w], l, n)) * Pm).re;  // s7
                    part[j][w] += t;
                }
                if (dm && in) part[j][4] += ((mln * K.V(tp[4], n, l)) * pv).im;  // s1: Re(-i D'^al V^d D'^be)
                if constexpr (M1) {
                    // kind 1 (b1, d1): (W^{b1 d1} + V''^b1 D'^d1 + V''^d1 D'^b1)_nl / 2, with P1 = -conj(pv)
                    const double mnl = -mln;
                    cplx<double> z1 = conj(gz[bj][v]);
                    if (single)
                        z1 -= mnl * (K.V(b1, n, n) * K.V(d1, n, l) + K.V(d1, n, n) * K.V(b1, n, l));
                    else
                        for (int q = S.gs[n]; q < S.ge[n]; ++q)
                            z1 += K.V(b1, n, q) * K.Dp(d1, q, l) + K.V(d1, n, q) * K.Dp(b1, q, l);
                    const cplx<double> Pm1 = (-mnl) * (0.5 * z1 + conj(pv));
#pragma unroll
                    for (int w = 0; w < 4; ++w) part1[i][w] -= dw * (Pm1 * K.A(tp1[w], l, n)).re;
                }
            } else if constexpr (KIND == 1) {
                if (same) continue;
                // (W^{bd} + V''^b D'^d + V''^d D'^b)_nl / 2
                const double mnl = K.m(n, l);
                cplx<double> z = gz[bj][v];
                if (single)
                    z -= mnl * (K.V(b, n, n) * K.V(d, n, l) + K.V(d, n, n) * K.V(b, n, l));
                else
                    for (int q = S.gs[n]; q < S.ge[n]; ++q)
                        z += K.V(b, n, q) * K.Dp(d, q, l) + K.V(d, n, q) * K.Dp(b, q, l);
                const cplx<double> Pm = (-mnl) * (0.5 * z - pv);
#pragma unroll
                for (int w = 0; w < 4; ++w) part[i][w] -= dw * (Pm * K.A(tp[w], l, n)).re;
            } else if constexpr (KIND == 2) {
                if (!same) {
                    // A'^{bd} - D'^d A''^b
                    const double mln = K.m(l, n);
                    cplx<double> z = gz[bj][v];
                    if (single)
                        z += mln * (K.V(d, l, n) * K.A(b, n, n));
                    else
                        for (int q = S.gs[n]; q < S.ge[n]; ++q) z -= K.Dp(d, l, q) * K.A(b, q, n);
                    const cplx<double> Y1 = pv + z;
#pragma unroll
                    for (int w = 0; w < 2; ++w) part[j][w] -= dw * ((mln * K.V(tp[w], n, l)) * Y1).re;
                } else if (dm) {
                    // s5: Re(-2i sig A^a_nl E_l (A' + A^b D'^d)_ln)
                    const cplx<double> dA = pv + K.Ad(b, d, l, n);
#pragma unroll
                    for (int w = 0; w < 2; ++w) part[j][w] += 2.0 * K.E[l] * (K.A(tp[w], n, l) * dA).im;
                }
            } else if constexpr (KIND == 3) {
                if (same) continue;
                // B'^{bd} - D'^d B''^b
                cplx<double> z = gz[bj][v];
                for (int q = S.gs[n]; q < S.ge[n]; ++q) z -= K.Dp(d, l, q) * K.B(b, q, n);
                const cplx<double> YB = pv + z;
#pragma unroll
                for (int w = 0; w < 2; ++w) part[j][w] -= 2.0 * (K.Dp(tp[w], n, l) * YB).re;  // s6
            } else if constexpr (KIND == 4) {
                if (!same) continue;
                // s5: -Re(-2i sig A^a_nl E_l (D'^d A^b)_ln) ; Omega block: 2 sign coef Re(DA_ln V^e_nl)
#pragma unroll
                for (int w = 0; w < 2; ++w) part[j][w] -= 2.0 * K.E[l] * (K.A(tp[w], n, l) * pv).im;
#pragma unroll
                for (int w = 2; w < 5; ++w) part[j][w] += (pv * K.V(tp[w], n, l)).re;
            } else if constexpr (KIND == 5 || KIND == 6) {
                if (!same) continue;
                // Omega block term -i P_ln: 2 sign Re(-i P_ln V^e_nl)
#pragma unroll
                for (int w = 0; w < 3; ++w) part[j][w] += (pv * K.V(tp[w], n, l)).im;
            } else {
                if (!same) continue;
                // s4: Re(-2i A^al_np (V''^d A''^be)_pn)
                part[j][0] += (K.A(tp[0], n, l) * pv).im;
            }
        }
    }
    // reduce over the summed index l: the rows, lane bits 4-5
    const int mask = 0x30;
    const bool writer = (lane >> 4) == 0;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int n = c0 + mfma_col(u);
#pragma unroll
        for (int w = 0; w < NT; ++w) {
            const double sum = wave_sum(part[u][w], mask);
            if (writer && tw[w] != 0.0 && n < nw) atomicAdd(&acc[n * 9 + tcol[w]].re, tw[w] * sum);
        }
        if constexpr (M1) {
#pragma unroll
            for (int w = 0; w < NT1; ++w) {
                const double sum = wave_sum(part1[u][w], mask);
                if (writer && tw1[w] != 0.0 && n < nw) atomicAdd(&acc[n * 9 + tcol1[w]].re, tw1[w] * sum);
            }
        }
    }
}

// DerOmega (formula/covariant.py:212-259) or, dm = true, DerMorb
// (formula/covariant.py:463-552) per-band values into S.acc[n][c*3+d].
// DerMorb = DerMorb_H + sign [Eav o dOmega + (Omega V + V Omega)/2] on cut
// partitions; its trace terms (s1 ... s7, the Omega block) are listed at
// formula_task.
// DM is compile-time so that the DerOmega kernel carries only the kinds it runs (0 - 2): its
// register allocation and code size are not those of the DerMorb kinds.
template <bool DM, typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) void eval_deromega(const Smem& S, const KCtx<T, LM, ST, SA, NW, TS>& K, bool in, bool ex,
                                                            int sign) {
    constexpr bool dm = DM;
    const int nw = K.n_(), tid = threadIdx.x;
    // per-band terms, lanes running over (n, c, d)
    for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
        const int n = t % nw, c = (t / nw) / 3, d = (t / nw) % 3;
        double v = 0.0;
        if (S.act[n] && ex) {
            cplx<double> s = mk(0.0, 0.0);
            const int g0 = S.gs[n], g1 = S.ge[n];
            for (int sg = 0; sg < 2; ++sg) {
                const int a = sg == 0 ? c_alpha[c] : c_beta[c];
                const int b = sg == 0 ? c_beta[c] : c_alpha[c];
                const double sig = sg == 0 ? 1.0 : -1.0;
                // T4: -i sigma sum_{n' in G} A^a_nn' dA.nn^{bd}_n'n
                for (int m = g0; m < g1; ++m) {
                    cplx<double> y2 = K.Ad(b, d, m, n);
#pragma unroll 4
                    for (int l = 0; l < nw; ++l) y2 += K.A(b, m, l) * K.Dp(d, l, n) - K.Dp(d, m, l) * K.A(b, l, n);
                    s += times_i(-(sig * (K.A(a, n, m) * y2)));
                }
            }
            // T5: 0.5 dO.nn = 0.5 [O'^{cd}_nn - (D'^d O^c)_nn + (O^c D'^d)_nn].  O is Hermitian
            // and D' anti-Hermitian, so (D'^d O^c)_nn = -conj((O^c D'^d)_nn) and only the real part,
            // Re O'_nn + 2 Re sum_l conj(O_ln) D'_ln, is used: column n of O, which consecutive
            // threads (consecutive n) read from consecutive addresses
            double t5 = K.Od(c, d, n, n).re;
#pragma unroll 8
            for (int l = 0; l < nw; ++l) t5 += 2.0 * (conj(K.O(c, l, n)) * K.Dp(d, l, n)).re;
            s.re += 0.5 * t5;
            v = (dm ? 2.0 * sign * S.E[n] : 2.0) * s.re;
            if (dm) {
                // s3: dH.nn = C'^{cd}_nn - (D'^d C^c)_nn + (C^c D'^d)_nn
                cplx<double> s3 = K.Cd(c, d, n, n);
                // Omega block, elementwise part: 2 sign Re sum_{l in G} (O^c_nl / 2) V^d_ln
                cplx<double> ov = mk(0.0, 0.0);
                for (int l = 0; l < nw; ++l) {
                    s3 += K.C(c, n, l) * K.Dp(d, l, n) - K.Dp(d, n, l) * K.C(c, l, n);
                    if (K.lab[l] == K.lab[n]) ov += K.O(c, n, l) * K.V(d, l, n);
                }
                v += s3.re + sign * ov.re;
            }
        }
        S.acc[n * 9 + c * 3 + d] = mk(v, 0.0);
    }
    if (tid == 0) S.misc[3] = 0;  // task counter
    __syncthreads();
    // wave tasks (kind, b, d, row block, column block), round-robin over the waves
    const TaskCtx tcx{dm, in, ex, (double)sign};
    int cnt[8];
    cnt[0] = (in || (dm && ex)) ? 9 : 0;
    cnt[1] = ex ? 9 : 0;
    cnt[2] = ex ? 9 : 0;
    cnt[3] = dm && ex ? 9 : 0;
    cnt[4] = dm && ex ? 9 : 0;
    cnt[5] = dm && in && sign != 0 ? 3 : 0;
    cnt[6] = dm && ex && sign != 0 ? 3 : 0;
    cnt[7] = dm && ex ? 9 : 0;
    if (dm && sign == 0) cnt[1] = 0;  // T3 feeds only the dOmega term
    // kind-1 tasks ride on the kind-0 tasks of the transposed index pair when both run
    const bool merge01 = cnt[0] > 0 && cnt[1] > 0;
    if (merge01) cnt[1] = 0;
    // a task is one 16-row x 32-column half block of one product; the waves take the next
    // task from an LDS counter, kinds in the order above (the heaviest, kind 0 with its
    // carried kind-1 task, first), so the waves finish together whatever the mix of kinds
    const int nblk = (nw + 31) / 32, nrh = (nw + 15) / 16, nb2 = nrh * nblk;
    int ntask = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) ntask += cnt[q] * nb2;
    const int lane = tid & 63;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(&S.misc[3], 1);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntask) break;
        int kind = -1, rem = 0, base = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int nq = cnt[q] * nb2;
            if (kind < 0 && t - base < nq) {
                kind = q;
                rem = t - base;
            }
            base += nq;
        }
        const int idx = rem / nb2, rh = (rem % nb2) / nblk, cb = rem % nblk;
        const int b = kind == 5 || kind == 6 ? idx : idx / 3, d = kind == 5 || kind == 6 ? 0 : idx % 3;
        switch (kind) {
            case 0:
                if (merge01)
                    formula_task<0, true>(S, K, tcx, S.acc, b, d, rh, cb);
                else
                    formula_task<0, false>(S, K, tcx, S.acc, b, d, rh, cb);
                break;
            case 1: formula_task<1, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
            case 2: formula_task<2, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
            default:
                if constexpr (DM) {
                    switch (kind) {
                        case 3: formula_task<3, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
                        case 4: formula_task<4, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
                        case 5: formula_task<5, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
                        case 6: formula_task<6, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
                        default: formula_task<7, false>(S, K, tcx, S.acc, b, d, rh, cb); break;
                    }
                }
                break;
        }
    }
}

// Per-band values for the current partition into S.acc[n][comp] (.re).
template <int F, typename T, bool LM, bool ST, bool SA, int NW, typename TS>
__device__ __attribute__((always_inline)) void eval_partition(const Smem& S, const KCtx<T, LM, ST, SA, NW, TS>& K, const wbmi_scan_params& p) {
    const int nw = K.n_();
    const int tid = threadIdx.x;
    const bool in = p.internal_terms, ex = p.external_terms;
    if constexpr (LM) build_M(S, nw);
    for (int t = tid; t < nw * 9; t += (int)blockDim.x) S.acc[t] = mk(0.0, 0.0);
    __syncthreads();
    if (F == WBMI_FORMULA_IDENTITY) {
        for (int n = tid; n < nw; n += (int)blockDim.x)
            if (S.act[n]) S.acc[n * 9] = mk(1.0, 0.0);
    } else if (F == WBMI_FORMULA_OMEGA) {
        if constexpr (ST) {
            // V and A in LDS: four adjacent lanes share (n, c) and split the sum over l,
            // reduced by two shuffles (a uniform loop, so every lane reaches them)
            const int L4 = (nw + 3) / 4;
            for (int t0 = 0; t0 < nw * 12; t0 += (int)blockDim.x) {
                const int t = t0 + tid;
                const bool ok = t < nw * 12;
                const int n = ok ? t / 12 : 0, c = (t / 4) % 3, qq = t % 4;
                const bool act = ok && S.act[n];
                double v = 0.0;
                if (act) {
                    const int l0 = qq * L4, l1 = l0 + L4 < nw ? l0 + L4 : nw;
                    v = omega_s(K, c, n, n, in, ex, l0, l1, qq == 0).re;
                }
                v += __shfl_xor(v, 1);
                v += __shfl_xor(v, 2);
                if (act && qq == 0) S.acc[n * 9 + c] = mk(2.0 * v, 0.0);
            }
        } else {
            for (int t = tid; t < nw * 3; t += (int)blockDim.x) {
                int n = t / 3, c = t % 3;
                if (!S.act[n]) continue;
                S.acc[n * 9 + c] = mk(2.0 * omega_s(K, c, n, n, in, ex).re, 0.0);
            }
        }
    } else if (F == WBMI_FORMULA_MORB_HPM) {
        for (int t = tid; t < nw * 3; t += (int)blockDim.x) {
            int n = t / 3, c = t % 3;
            if (!S.act[n]) continue;
            double v = 2.0 * morbh_s(K, c, n, n, in, ex).re;
            if (p.sign != 0) v += p.sign * S.E[n] * 2.0 * omega_s(K, c, n, n, in, ex).re;
            S.acc[n * 9 + c] = mk(v, 0.0);
        }
    } else if (F == WBMI_FORMULA_VELOMEGA) {
        // Re sum_{n' in G} V^a_nn' Omega^b_n'n, Omega = s + s^+ (formula/formula.py:141-148)
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            int n = t / 9, a = (t % 9) / 3, b = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) {
                cplx<double> om = omega_s(K, b, m, n, in, ex) + conj(omega_s(K, b, n, m, in, ex));
                v += K.V(a, n, m) * om;
            }
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_VELHPLUS) {
        // Re sum_{n' in G} V^a_nn' Hpm^b_n'n with the group block of Morb_Hpm,
        // Hpm = Morb_H.nn + sign Eav o Omega.nn (formula/covariant.py:424-449,
        // formula/formula.py:141-148; FormulaProduct is additive)
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            int n = t / 9, a = (t % 9) / 3, b = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) {
                cplx<double> h = morbh_s(K, b, m, n, in, ex) + conj(morbh_s(K, b, n, m, in, ex));
                if (p.sign != 0)
                    h += (p.sign * 0.5 * (S.E[m] + S.E[n])) *
                         (omega_s(K, b, m, n, in, ex) + conj(omega_s(K, b, n, m, in, ex)));
                v += K.V(a, n, m) * h;
            }
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_DERSPIN) {
        // generalised derivative of the spin (formula/formula.py:95-107 with A = Xbar('SS'),
        // dA = Xbar('SS', 1)): Re [S'^{bd}_nn - sum_{l off} D^d_nl S^b_ln + sum_{l off} S^b_nl D^d_ln]
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, b = (t % 9) / 3, d = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = K.Sd(b, d, n, n);
            for (int l = 0; l < nw; ++l) {
                v -= K.Dp(d, n, l) * K.S(b, l, n);
                v += K.S(b, n, l) * K.Dp(d, l, n);
            }
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_SPIN) {
        // Re S^s_nn: the group trace of Xbar('SS') (formula/covariant.py:331-335)
        for (int t = tid; t < nw * 3; t += (int)blockDim.x) {
            const int n = t / 3, sp = t % 3;
            if (S.act[n]) S.acc[n * 9 + sp] = mk(K.S(sp, n, n).re, 0.0);
        }
    } else if (F == WBMI_FORMULA_VELSPIN) {
        // Re sum_{n' in G} V^a_nn' S^b_n'n (formula/formula.py:121-148, covariant.py:811-814)
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, a = (t % 9) / 3, b = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) v += K.V(a, n, m) * K.S(b, m, n);
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_INVMASS) {
        // inverse effective mass, the generalised derivative of the velocity (formula/elementary.py:28-33,
        // formula/formula.py:95-107 with A = V, dA = W): Re [W^{bd}_nn - sum_{l off} D^d_nl V^b_ln
        // + sum_{l off} V^b_nl D^d_ln]
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, b = (t % 9) / 3, d = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = K.W(b, d, n, n);
            for (int l = 0; l < nw; ++l) {
                v -= K.Dp(d, n, l) * K.V(b, l, n);
                v += K.V(b, n, l) * K.Dp(d, l, n);
            }
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_VELVEL) {
        // Re sum_{n' in G} V^a_nn' V^b_n'n (FormulaProduct of two velocities, formula/covariant.py:817-820)
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, a = (t % 9) / 3, b = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) v += K.V(a, n, m) * K.V(b, m, n);
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_VELVELVEL) {
        // [a][b][c] = Re sum_{m, l in G} V^a_nm V^b_ml V^c_ln (formula/formula.py:121-148), a = comp_slice
        const int a = p.comp_slice;
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, b = (t % 9) / 3, c = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m)
                for (int l = S.gs[n]; l < S.ge[n]; ++l) v += K.V(a, n, m) * K.V(b, m, l) * K.V(c, l, n);
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_MASSVEL) {
        // [a][b][c] = Re sum_{m in G} M^{ab}_nm V^c_mn, a = comp_slice
        const int a = p.comp_slice;
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, b = (t % 9) / 3, c = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) v += mass_xy(K, a, b, n, m) * K.V(c, m, n);
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_HALL_SURF) {
        // VelMassVel X[a][b][c][d] = Re sum_{m, l in G} V^a_nm M^{bc}_ml V^d_ln, contracted
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, j = (t % 9) / 3, i = t % 3;
            if (!S.act[n]) continue;
            const cplx<double> v = hall_contract(j, i, [&](int a, int b, int c, int d) {
                cplx<double> x = mk(0.0, 0.0);
                for (int m = S.gs[n]; m < S.ge[n]; ++m)
                    for (int l = S.gs[n]; l < S.ge[n]; ++l) x += K.V(a, n, m) * mass_xy(K, b, c, m, l) * K.V(d, l, n);
                return x;
            });
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_HALL_SEA) {
        // MassMass X[a][b][c][d] = Re sum_{m in G} M^{ab}_nm M^{cd}_mn, its axes moved as
        // transpose(0, 4, 1, 2, 3) (Y[p][q][r][s] = X[q][r][s][p]), contracted
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, j = (t % 9) / 3, i = t % 3;
            if (!S.act[n]) continue;
            const cplx<double> v = hall_contract(j, i, [&](int p0, int q, int r, int s4) {
                cplx<double> x = mk(0.0, 0.0);
                for (int m = S.gs[n]; m < S.ge[n]; ++m) x += mass_xy(K, q, r, n, m) * mass_xy(K, s4, p0, m, n);
                return x;
            });
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_OMEGASPIN || F == WBMI_FORMULA_OMEGAOMEGA || F == WBMI_FORMULA_OMEGAHPLUS) {
        // Re sum_{n' in G} Omega^a_nn' Y^b_n'n, Omega = s + s^+, Y = S (OmegaS), Omega (OmegaOmega)
        // or Morb_Hpm with sign +1 (OmegaHplus) (formula/formula.py:121-148, covariant.py:849-866)
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, a = (t % 9) / 3, b = t % 3;
            if (!S.act[n]) continue;
            cplx<double> v = mk(0.0, 0.0);
            for (int m = S.gs[n]; m < S.ge[n]; ++m) {
                const cplx<double> om = omega_s(K, a, n, m, in, ex) + conj(omega_s(K, a, m, n, in, ex));
                cplx<double> y;
                if (F == WBMI_FORMULA_OMEGASPIN) {
                    y = K.S(b, m, n);
                } else {
                    const cplx<double> ob = omega_s(K, b, m, n, in, ex) + conj(omega_s(K, b, n, m, in, ex));
                    if (F == WBMI_FORMULA_OMEGAOMEGA)
                        y = ob;
                    else
                        y = morbh_s(K, b, m, n, in, ex) + conj(morbh_s(K, b, n, m, in, ex)) +
                            (0.5 * (S.E[m] + S.E[n])) * ob;
                }
                v += om * y;
            }
            S.acc[t] = mk(v.re, 0.0);
        }
    } else if (F == WBMI_FORMULA_BANDVALUES) {
        // the caller's per-band shares of the group traces, 9 of them from row comp_slice * 9
        for (int t = tid; t < nw * 9; t += (int)blockDim.x) {
            const int n = t / 9, c = t % 9;
            if (S.act[n]) S.acc[t] = mk(K.bv[n * K.nbv + p.comp_slice * 9 + c], 0.0);
        }
    } else if (F == WBMI_FORMULA_DEROMEGA) {
        eval_deromega<false>(S, K, in, ex, 0);
    } else if (F == WBMI_FORMULA_DERMORB) {
        eval_deromega<true>(S, K, in, ex, p.sign);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// tetrahedron method (grid/tetrahedron.py:15-128, 269-281)
// ---------------------------------------------------------------------------

#define TETRA_SPLIT 1e-12  // minimum separation of sorted corner energies (tetrahedron.py:17-21)
#define TETRA_PAD 1e-11    // above max + TETRA_PAD every tetrahedron is fully occupied

// Occupied fraction (der 0) of a tetrahedron with linear energy between the corner values,
// or its der-th derivative in the Fermi level.  der 0 evaluates the three cubic pieces as
// products of ratios; der >= 1 differentiates their power-series form, with the
// coefficients grouped as the reference groups them (its rounding where corners nearly
// coincide).  No contraction into FMAs, so that the rounding is that of separate operations.
__device__ inline double tetra_occ(double ef, double a, double b, double c, double d, int der) {
#pragma clang fp contract(off)
    double t;
    if (a > b) { t = a; a = b; b = t; }
    if (c > d) { t = c; c = d; d = t; }
    if (a > c) { t = a; a = c; c = t; }
    if (b > d) { t = b; b = d; d = t; }
    if (b > c) { t = b; b = c; c = t; }
    const double e1 = a;
    const double e2 = b - e1 < TETRA_SPLIT ? e1 + TETRA_SPLIT : b;
    const double e3 = c - e2 < TETRA_SPLIT ? e2 + TETRA_SPLIT : c;
    const double e4 = d - e3 < TETRA_SPLIT ? e3 + TETRA_SPLIT : d;
    if (der == 0) {
        if (ef >= e4) return 1.0;
        if (ef < e1) return 0.0;
        if (ef >= e3) return 1.0 - ((ef - e4) / (e1 - e4)) * ((ef - e4) / (e2 - e4)) * ((ef - e4) / (e3 - e4));
        if (ef >= e2) {
            const double a13 = (ef - e1) / (e3 - e1), a14 = (ef - e1) / (e4 - e1);
            const double a23 = (ef - e2) / (e3 - e2), a24 = (ef - e2) / (e4 - e2);
            return a23 * a24 + a13 * (a14 * (1.0 - a24) + a24 * (1.0 - a23));
        }
        return ((ef - e1) / (e2 - e1)) * ((ef - e1) / (e3 - e1)) * ((ef - e1) / (e4 - e1));
    }
    if (ef >= e4 || ef < e1) return 0.0;
    double c1, c2, c3;
    if (ef >= e3) {
        const double r3 = 1.0 / ((e4 - e1) * (e4 - e2) * (e4 - e3));
        c1 = 3.0 * (e4 * e4) * r3;
        c2 = -3.0 * e4 * r3;
        c3 = r3;
    } else if (ef >= e2) {
        const double r2 = 1.0 / ((e3 - e1) * (e4 - e1) * (e3 - e2) * (e4 - e2));
        c1 = (-2.0 * e1 * ((e3 - e2) * (e4 - e2)) + (2.0 * e2 * e4 + e2 * e2) * (e1 - e3) +
              (e1 * e2 + e2 * e3 + e1 * e3) * (e2 - e4)) * r2;
        c2 = (((e3 - e2) * (e4 - e2)) - (e1 - e3) * (2.0 * e2 + e4) - (e3 + e1 + e2) * (e2 - e4)) * r2;
        c3 = r2 * (e1 + e2 - e3 - e4);
    } else {
        const double r1 = 1.0 / ((e2 - e1) * (e3 - e1) * (e4 - e1));
        c1 = 3.0 * (e1 * e1) * r1;
        c2 = -3.0 * e1 * r1;
        c3 = r1;
    }
    if (der == 1) return c1 + ef * (2.0 * c2 + 3.0 * c3 * ef);
    if (der == 2) return 2.0 * c2 + 6.0 * c3 * ef;
    return 6.0 * c3;
}

// Weight of one band at one k: the mean over the 12 tetrahedra joining the centre to the
// two triangles (F00, F01, F11), (F00, F10, F11) of each face of the parallelepiped, faces
// taken as x = side, y = side, z = side for side 0 then 1.  C[ix*4 + iy*2 + iz].
__device__ inline double tetra_band(double ef, double ec, const double* C, int der, int ncorner) {
    // tetrahedral K-blocks: one tetrahedron, the 4 corners themselves (tetrahedron.py:204-209)
    if (ncorner == 4) return tetra_occ(ef, C[0], C[1], C[2], C[3], der);
    double occ = 0.0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        // face x = side: F[a][b] = C[side*4 + a*2 + b]
        occ += tetra_occ(ef, ec, C[side * 4], C[side * 4 + 1], C[side * 4 + 3], der);
        occ += tetra_occ(ef, ec, C[side * 4], C[side * 4 + 2], C[side * 4 + 3], der);
        // face y = side: F[a][b] = C[a*4 + side*2 + b]
        occ += tetra_occ(ef, ec, C[side * 2], C[side * 2 + 1], C[4 + side * 2 + 1], der);
        occ += tetra_occ(ef, ec, C[side * 2], C[4 + side * 2], C[4 + side * 2 + 1], der);
        // face z = side: F[a][b] = C[a*4 + b*2 + side]
        occ += tetra_occ(ef, ec, C[side], C[2 + side], C[6 + side], der);
        occ += tetra_occ(ef, ec, C[side], C[4 + side], C[6 + side], der);
    }
    return occ / 12.0;
}

// first index e with Ef[e] >= x (lower) or Ef[e] > x (upper) in the ascending Ef[0..n)
__device__ inline int ef_bound(const double* __restrict__ Ef, int n, double x, bool upper) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (upper ? Ef[m] <= x : Ef[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Adds the in-range groups' values S.tv with their tetrahedron weights (static.py:119-125;
// TetraWeights.weights_all_band_groups, tetrahedron.py:212-255): band b of group g adds
// tv[g] / |g| * w_b(Ef).  w_b is 0 below the band's lowest energy over centre and corners
// and (der 0) 1 above its highest: that part is one step row of the histogram; the levels
// in between are evaluated one by one, spread over all threads.
static __device__ void tetra_accumulate(const Smem& S, const ScanArgs& args, int ncomp, double* __restrict__ hist) {
    const int nw = args.nw, nEF = args.nEF, der = args.p.fder, tid = threadIdx.x;
    for (int b = tid; b < nw; b += (int)blockDim.x) {
        const int g = S.tg[b];
        int lo = 0, hi = 0;
        if (g >= 0 && S.gin[g]) {  // g < 0: a band in no group (the odd last band under degen_Kramers)
            lo = ef_bound(args.Ef, nEF, S.tEmn[b], false);
            hi = ef_bound(args.Ef, nEF, S.tEmx[b] + TETRA_PAD, true);
            const double f = args.weight / (S.gb[g + 1] - S.gb[g]);
            if (der == 0 && hi < nEF)
                for (int c = 0; c < ncomp; ++c) unsafeAtomicAdd(hist + (int64_t)hi * ncomp + c, f * S.tv[g * 9 + c]);
            if (der == -1) {
                // 1 - occupation: 1 below the band's span (rows 0 .. hi-1), minus the occupation inside
                for (int c = 0; c < ncomp; ++c) {
                    unsafeAtomicAdd(hist + c, f * S.tv[g * 9 + c]);
                    if (hi < nEF) unsafeAtomicAdd(hist + (int64_t)hi * ncomp + c, -f * S.tv[g * 9 + c]);
                }
            }
        }
        S.tlo[b] = lo;
        S.tcnt[b] = hi - lo;
    }
    __syncthreads();
    if (tid == 0) {
        S.tpre[0] = 0;
        for (int b = 0; b < nw; ++b) S.tpre[b + 1] = S.tpre[b] + S.tcnt[b];
    }
    __syncthreads();
    const int total = S.tpre[nw];
    for (int t = tid; t < total; t += (int)blockDim.x) {
        // band b with tpre[b] <= t < tpre[b + 1]
        int lo = 0, hi = nw;
        while (hi - lo > 1) {
            const int m = (lo + hi) >> 1;
            if (S.tpre[m] <= t) lo = m;
            else hi = m;
        }
        const int b = lo, e = S.tlo[b] + (t - S.tpre[b]), g = S.tg[b];
        const double w = der == -1 ? -tetra_band(args.Ef[e], S.E[b], S.tC + b * 8, 0, args.tetra)
                                   : tetra_band(args.Ef[e], S.E[b], S.tC + b * 8, der, args.tetra);
        const double f = args.weight * w / (S.gb[g + 1] - S.gb[g]);
        for (int c = 0; c < ncomp; ++c)
            unsafeAtomicAdd(hist + (int64_t)(nEF + e) * ncomp + c, f * S.tv[g * 9 + c]);
    }
}

// select_bands share of the group [a, b) (utility.py:398-403): 1 without a selection
__device__ inline double group_selected(const wbmi_scan_params& p, int a, int b) {
    if (!p.band_weight) return 1.0;
    double s = 0.0;
    for (int n = a; n < b; ++n) s += p.band_weight[n];
    return s / (double)(b - a);
}

// Fermi-level bin of a group with mean energy E (static.py:131-136): 0 for
// E < EFmin (added to every level), ceil((E - EFmin)/dEF) up to EFmax, -1 = none.
__device__ inline int fermi_bin(const wbmi_scan_params& p, double E) {
    if (E < p.EFmin) return 0;
    if (E <= p.EFmax) {
        const int bin = (int)ceil((E - p.EFmin) / p.dEF);
        return bin < p.nEF_extra ? bin : -1;
    }
    return -1;
}

// MODE: 0 = Fermi-level bins, 1 = tetrahedron weights, 2 = per-band table (hist is then
// out[nk][nw][ncomp]: each band gets its group's trace / group size, or, for the identity
// formula, the group's mean energy)
// element type of the staged V and A: complex128 for DerOmega / DerMorb (one workgroup per CU
// holds them anyway; converted once at staging instead of at every access), the channels'
// storage type for Omega (two complex64 workgroups per CU)
template <typename T, int F>
struct scan_stage {
    typedef typename std::conditional<F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_DERMORB, double, T>::type type;
};

// workgroup size of k_static_scan<T, F, .., MODE> (its launch bound).  The tetrahedron
// variants (MODE 1) of DerOmega / DerMorb keep 8-wave workgroups: with their weight code
// inlined, the 168-register bound of 12-wave workgroups spills several hundred registers
template <typename T, int F, int MODE>
__host__ __device__ constexpr int scan_threads() {
    constexpr bool tasks = F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_DERMORB;
    if (!tasks || MODE == 1) return SC_THREADS;
    return SC_TASK_THREADS;
}

#ifndef SC_TASK_MINW
#define SC_TASK_MINW 1  // minimum waves per SIMD asked of the DerOmega / DerMorb Fermi-bin kernels
#endif
template <typename T, int F, int MODE>
__host__ __device__ constexpr int scan_min_waves() {
    constexpr bool tasks = F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_DERMORB;
    return tasks && MODE != 1 ? SC_TASK_MINW : 1;
}

// NW > 0: the kernel for nw = NW only, every size and row stride a compile-time constant
template <typename T, int F, bool LM, bool ST, int MODE, int NW>
__global__ void __launch_bounds__((scan_threads<T, F, MODE>()), (scan_min_waves<T, F, MODE>())) k_static_scan(ScanArgs args, const double* __restrict__ Eall,
                                                            const cplx<T>* __restrict__ Xall,
                                                            double* __restrict__ hist) {
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    const wbmi_scan_params& p = args.p;
    const int nw = NW > 0 ? NW : args.nw, ld = nw + 1;
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x;
    if (MODE != 2) hist += k * args.hist_stride;
    constexpr bool tasks = F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_DERMORB;
    static_assert(tasks || F == WBMI_FORMULA_OMEGA || !ST, "V and A are staged for DerOmega, DerMorb and Omega");
    using TS = typename scan_stage<T, F>::type;
    constexpr bool SA = ST && SC_STAGE_A;
    constexpr bool stage_v = ST;
    const bool stage_a = SA && p.external_terms;
    constexpr bool tetra = MODE == 1;  // compile-time: the Fermi-bin kernels carry no tetrahedron code
    constexpr bool tab = MODE == 2;
    const Smem S = make_smem(smem_raw, nw, stage_v, stage_a, tetra, sizeof(cplx<TS>));
    for (int n = tid; n < nw; n += (int)blockDim.x) S.E[n] = Eall[k * nw + n];
    for (int r = tid; r < WBMI_NROLES; r += (int)blockDim.x) S.ch[r] = p.ch[r];
    if (tetra)
        for (int t = tid; t < args.tetra * nw; t += (int)blockDim.x) {
            const int c = t / nw, b = t % nw;
            S.tC[b * 8 + c] = args.Ec[c * args.ec_stride + k * nw + b];
        }
    __syncthreads();
    if (tetra) {
        // band energy range over the centre and the corners (tetrahedron.py:189-193)
        for (int b = tid; b < nw; b += (int)blockDim.x) {
            double mn = S.E[b], mx = S.E[b];
            for (int c = 0; c < args.tetra; ++c) {
                mn = fmin(mn, S.tC[b * 8 + c]);
                mx = fmax(mx, S.tC[b * 8 + c]);
            }
            S.tEmn[b] = mn;
            S.tEmx[b] = mx;
        }
        __syncthreads();
    }
    const cplx<T>* Xk = Xall + k * (int64_t)args.nch * nw * nw;
    if (stage_v) {
        // staged only for nw <= SC_LDS_NW: 3 nw^2 <= 3072 elements per array, NS per thread of
        // the workgroup.  Each copy loop is branch-free (indices clamped into range; a thread past
        // the end rewrites the last element with its own value), so it is one basic block and the
        // scheduler issues all of a thread's loads before its first LDS store
        constexpr int NTH = scan_threads<T, F, MODE>();
        constexpr int NS = (3 * SC_LDS_NW * SC_LDS_NW + NTH - 1) / NTH;
        const int nw2 = nw * nw, n3 = 3 * nw2;
        int so[NS], lo[NS];  // source element (channel-relative) and LDS element of each slot
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int t = min(tid + NTH * u, n3 - 1);
            const int a = t >= 2 * nw2 ? 2 : t >= nw2 ? 1 : 0, xy = t - a * nw2;
            const int x = xy / nw, y = xy - x * nw;
            so[u] = a * 4096 + xy;  // a in the bits above xy (nw^2 <= 1024)
            lo[u] = (a * nw + x) * ld + y;
        }
        const cplx<T>* srcv[3] = {Xk + (int64_t)S.ch[WBMI_ROLE_V] * nw2, Xk + (int64_t)S.ch[WBMI_ROLE_V + 1] * nw2,
                                  Xk + (int64_t)S.ch[WBMI_ROLE_V + 2] * nw2};
        cplx<TS>* dv = (cplx<TS>*)S.V;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int a = so[u] >> 12, xy = so[u] & 4095;
            const cplx<T> v = (a == 0 ? srcv[0] : a == 1 ? srcv[1] : srcv[2])[xy];
            dv[lo[u]] = mk((TS)v.re, (TS)v.im);
        }
        if (stage_a) {
            const cplx<T>* srca[3] = {Xk + (int64_t)S.ch[WBMI_ROLE_A] * nw2, Xk + (int64_t)S.ch[WBMI_ROLE_A + 1] * nw2,
                                      Xk + (int64_t)S.ch[WBMI_ROLE_A + 2] * nw2};
            cplx<TS>* da = (cplx<TS>*)S.A;
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                const int a = so[u] >> 12, xy = so[u] & 4095;
                const cplx<T> v = (a == 0 ? srca[0] : a == 1 ? srca[1] : srca[2])[xy];
                da[lo[u]] = mk((TS)v.re, (TS)v.im);
            }
        }
    }
    constexpr int ncomp = ncomp_of<F>();
    constexpr bool additive = F == WBMI_FORMULA_IDENTITY || F == WBMI_FORMULA_OMEGA ||
                              F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_VELOMEGA ||
                              F == WBMI_FORMULA_VELHPLUS || F == WBMI_FORMULA_DERSPIN ||
                              F == WBMI_FORMULA_VELSPIN || F == WBMI_FORMULA_SPIN || F == WBMI_FORMULA_INVMASS ||
                              F == WBMI_FORMULA_VELVEL || F == WBMI_FORMULA_VELVELVEL || F == WBMI_FORMULA_MASSVEL ||
                              F == WBMI_FORMULA_HALL_SURF || F == WBMI_FORMULA_HALL_SEA ||
                              F == WBMI_FORMULA_BANDVALUES || F == WBMI_FORMULA_OMEGASPIN ||
                              F == WBMI_FORMULA_OMEGAOMEGA || F == WBMI_FORMULA_OMEGAHPLUS;

    // band groups (grid/tetrahedron.py:132-162, data_K.py:172-186), formed by wave 0: each lane
    // owns bands lane and lane + 64 (nw <= 128); borders from two ballots and popcounts, one
    // lane per group for its range test and mean energy, wave reductions for the sea group
    if (tid < 64) {
        const int lane = tid;
        const double thr = p.degen_thresh;
        auto is_border = [&](int i) {  // i in [0, nw]
            bool b = i == 0 || i == nw || (i < nw && S.E[i] - S.E[i - 1] > thr);
            if (p.degen_Kramers && (i & 1)) b = false;
            return b;
        };
        const bool b0 = lane <= nw && is_border(lane);
        const bool b1 = lane + 64 <= nw && is_border(lane + 64);
        const bool b2 = lane == 0 && nw == 128 && is_border(128);
        const uint64_t m0 = __ballot(b0), m1 = __ballot(b1);
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int n0 = __popcll(m0), n1 = __popcll(m1);
        if (b0) S.gb[__popcll(m0 & below)] = lane;
        if (b1) S.gb[n0 + __popcll(m1 & below)] = lane + 64;
        if (b2) S.gb[n0 + n1] = 128;
        const int nb = n0 + n1 + (nw == 128 && is_border(128) ? 1 : 0);
        const int ng = nb - 1;
        // degen_Kramers keeps only even borders, so with odd nw the last band is in no group
        // (grid/tetrahedron.py:132-136): it is an "out" band of every group and adds nothing itself
        for (int i = lane; i < nw; i += 64) S.tg[i] = -1;
        wave_lds_sync();
        int first_in = nw, last_in = 0;
        for (int g = lane; g < ng; g += 64) {
            const int a = S.gb[g], b = S.gb[g + 1];
            int inr;
            if (tab) {
                inr = 1;  // every group (Tabulator, tabulate.py:25-31)
            } else if (tetra) {
                // range over centre and corners against [Ef[0], Ef[-1]] (tetrahedron.py:140-152)
                double mn = S.tEmn[a], mx = S.tEmx[a];
                for (int i = a; i < b; ++i) {
                    mn = fmin(mn, S.tEmn[i]);
                    mx = fmax(mx, S.tEmx[i]);
                    S.tg[i] = g;
                }
                inr = (mx >= p.EFmin) && (mn <= p.EFmax);
            } else {
                inr = (S.E[b - 1] >= p.EFmin) && (S.E[a] <= p.EFmax);
            }
            S.gin[g] = inr;
            if (inr) {
                first_in = min(first_in, a);
                last_in = max(last_in, b);
            }
            double sum = 0.0;
            for (int i = a; i < b; ++i) sum += S.E[i];
            S.gw[g] = sum / (b - a);
        }
        for (int o = 32; o >= 1; o >>= 1) {
            first_in = min(first_in, __shfl_xor(first_in, o));
            last_in = max(last_in, __shfl_xor(last_in, o));
        }
        int sea_lo = 0, sea_hi = 0;
        if (p.fder == 0 && !tab) {
            // bands below the range, up to the first group in it (data_K.py:178-185;
            // tetra: highest energy over centre and corners below Ef[0], tetrahedron.py:155-162)
            int hi = 0, lo = 0;
            for (int i = lane; i < nw; i += 64) {
                if ((tetra ? S.tEmx[i] : S.E[i]) < p.EFmin) hi = i + 1;
                // tetra with Emin: bands entirely below Emin stay out (tetrahedron.py:246-251)
                if (tetra && p.erange && S.tEmx[i] < p.Emin) lo = i + 1;
            }
            for (int o = 32; o >= 1; o >>= 1) {
                hi = max(hi, __shfl_xor(hi, o));
                lo = max(lo, __shfl_xor(lo, o));
            }
            sea_hi = first_in < hi ? first_in : hi;
            sea_lo = lo < sea_hi ? lo : sea_hi;
        } else if (tetra && p.fder == -1) {
            // hole-like: bands whose lowest energy lies above Ef[-1], from the end of the last
            // group in range on (tetrahedron.py:165-172,246-253)
            int bandmin = nw, bandmax = nw;
            for (int i = lane; i < nw; i += 64) {
                if (S.tEmn[i] > p.EFmax) bandmin = min(bandmin, i);
                // with Emax: bands entirely above Emax stay out (tetrahedron.py:252-258)
                if (p.erange && S.tEmn[i] > p.Emax) bandmax = min(bandmax, i);
            }
            for (int o = 32; o >= 1; o >>= 1) {
                bandmin = min(bandmin, __shfl_xor(bandmin, o));
                bandmax = min(bandmax, __shfl_xor(bandmax, o));
            }
            if (bandmin < last_in) bandmin = last_in;
            if (bandmin < bandmax) {
                sea_lo = bandmin;
                sea_hi = bandmax;
            }
        }
        if (lane == 0) {
            S.misc[0] = ng;
            S.misc[1] = sea_lo;
            S.misc[2] = sea_hi;
        }
    }
    __syncthreads();
    const int ng = S.misc[0];
    const int sea_lo = S.misc[1], sea_hi = S.misc[2];

    KCtx<T, LM, ST, SA, NW, TS> K;
    K.X = Xk;
    K.nw = nw;
    K.ld = ld;
    K.ch = S.ch;
    K.E = S.E;
    K.M = S.M;
    K.lab = S.lab;
    K.gs = S.gs;
    K.ge = S.ge;
    K.sV = (const cplx<TS>*)S.V;
    K.sA = (const cplx<TS>*)S.A;
    K.bv = (const double*)(const void*)Xall + k * (int64_t)nw * args.nch;
    K.nbv = args.nch;

    if (additive || tab) {
        // main partition: every group at once (each group's trace with inn = the group,
        // out = the rest: the additive scan and the per-band table of any formula); a band in
        // no group (label -1) is an out band of every group
        for (int n = tid; n < nw; n += (int)blockDim.x) {
            S.lab[n] = -1;
            S.gs[n] = n;
            S.ge[n] = n;
            S.act[n] = 0;
        }
        __syncthreads();
        for (int g = 0; g < ng; ++g)
            for (int n = S.gb[g] + tid; n < S.gb[g + 1]; n += (int)blockDim.x) {
                S.lab[n] = g;
                S.gs[n] = S.gb[g];
                S.ge[n] = S.gb[g + 1];
                S.act[n] = S.gin[g];
            }
        __syncthreads();
        eval_partition<F>(S, K, p);
        for (int t = tid; t < ng * ncomp; t += (int)blockDim.x) {
            const int g = t / ncomp, c = t % ncomp;
            if (tab) {
                double v;
                if (F == WBMI_FORMULA_IDENTITY) {
                    v = S.gw[g];  // Eavln: mean energy of the group
                } else {
                    v = 0.0;
                    for (int n = S.gb[g]; n < S.gb[g + 1]; ++n) v += S.acc[n * 9 + c].re;
                    v /= (S.gb[g + 1] - S.gb[g]);
                }
                for (int n = S.gb[g]; n < S.gb[g + 1]; ++n) hist[(k * nw + n) * ncomp + c] = v;
                continue;
            }
            if (tetra) {
                if (!S.gin[g]) continue;
                double s = 0.0;
                for (int n = S.gb[g]; n < S.gb[g + 1]; ++n) s += S.acc[n * 9 + c].re;
                S.tv[g * 9 + c] = s * group_selected(p, S.gb[g], S.gb[g + 1]);
                continue;
            }
            const int bin = S.gin[g] ? fermi_bin(p, S.gw[g]) : -1;
            if (bin < 0) continue;
            double s = 0.0;
            for (int n = S.gb[g]; n < S.gb[g + 1]; ++n) s += S.acc[n * 9 + c].re;
            unsafeAtomicAdd(hist + (int64_t)bin * ncomp + c, args.weight * s * group_selected(p, S.gb[g], S.gb[g + 1]));
        }
        __syncthreads();
        if (sea_hi > sea_lo) {
            // Fermi-sea group [sea_lo, sea_hi) (weight -inf; hole-like tetra: the bands above),
            // added at every level; every other band is its "out" (with Emin / Emax on both
            // sides of it: each side its own block)
            for (int n = tid; n < nw; n += (int)blockDim.x) {
                int inn = n >= sea_lo && n < sea_hi;
                S.lab[n] = inn ? 0 : 1;
                S.gs[n] = inn ? sea_lo : (n < sea_lo ? 0 : sea_hi);
                S.ge[n] = inn ? sea_hi : (n < sea_lo ? sea_lo : nw);
                S.act[n] = inn;
            }
            __syncthreads();
            eval_partition<F>(S, K, p);
            for (int c = tid; c < ncomp; c += (int)blockDim.x) {
                double s = 0.0;
                for (int n = sea_lo; n < sea_hi; ++n) s += S.acc[n * 9 + c].re;
                unsafeAtomicAdd(hist + c, args.weight * s);  // bin 0: every level
            }
            __syncthreads();
        }
    } else {
        // non-additive: traces of the cumulative cuts [0, c) (static.py:109-117)
        for (int cidx = 1; cidx <= nw; ++cidx) {
            bool need = sea_hi > sea_lo && (cidx == sea_hi || cidx == sea_lo);
            for (int g = 0; g < ng; ++g)
                if (S.gin[g] && (S.gb[g] == cidx || S.gb[g + 1] == cidx)) need = true;
            if (!need) continue;
            for (int n = tid; n < nw; n += (int)blockDim.x) {
                int inn = n < cidx;
                S.lab[n] = inn ? 0 : 1;
                S.gs[n] = inn ? 0 : cidx;
                S.ge[n] = inn ? cidx : nw;
                S.act[n] = inn;
            }
            __syncthreads();
            eval_partition<F>(S, K, p);
            if (tid == 0)
                for (int c = 0; c < ncomp; ++c) {
                    double s = 0.0;
                    for (int n = 0; n < cidx; ++n) s += S.acc[n * 9 + c].re;
                    S.cut[cidx * 9 + c] = s;
                }
            __syncthreads();
        }
        for (int c = tid; c < ncomp; c += (int)blockDim.x) S.cut[c] = 0.0;
        __syncthreads();
        for (int t = tid; t < ng * ncomp; t += (int)blockDim.x) {
            const int g = t / ncomp, c = t % ncomp;
            if (tetra) {
                if (S.gin[g])
                    S.tv[g * 9 + c] = (S.cut[S.gb[g + 1] * 9 + c] - S.cut[S.gb[g] * 9 + c]) *
                                      group_selected(p, S.gb[g], S.gb[g + 1]);
                continue;
            }
            const int bin = S.gin[g] ? fermi_bin(p, S.gw[g]) : -1;
            if (bin < 0) continue;
            const double s = S.cut[S.gb[g + 1] * 9 + c] - S.cut[S.gb[g] * 9 + c];
            unsafeAtomicAdd(hist + (int64_t)bin * ncomp + c, args.weight * s * group_selected(p, S.gb[g], S.gb[g + 1]));
        }
        if (sea_hi > sea_lo)
            for (int c = tid; c < ncomp; c += (int)blockDim.x)
                unsafeAtomicAdd(hist + c, args.weight * (S.cut[sea_hi * 9 + c] - S.cut[sea_lo * 9 + c]));
    }
    if (tetra) {
        __syncthreads();
        tetra_accumulate(S, args, ncomp, hist);
    }
}

template <typename T, int F, bool LM, bool ST, int MODE, int NW = 0>
static int launch_scan_t(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist) {
    const bool stage_a = ST && SC_STAGE_A && a.p.external_terms;
    // DerOmega / DerMorb: 12 waves (3 per SIMD at <= 168 VGPRs), one workgroup per CU (at nw = 32
    // LDS holds V and A, 118 KB as complex128)
    const int threads = scan_threads<T, F, MODE>();
    size_t lds = smem_offsets(a.nw, ST, stage_a, MODE == 1, sizeof(cplx<typename scan_stage<T, F>::type>)).total;
    if (lds > 160 * 1024) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_static_scan: %zu bytes of LDS needed", lds);
    WBMI_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_static_scan<T, F, LM, ST, MODE, NW>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_static_scan<T, F, LM, ST, MODE, NW>), dim3(nk), dim3(threads), lds,
                       ctx->stream, a, E,
                       (const cplx<T>*)X, hist);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

// dEinv o Off in LDS up to SC_M_NW (the layout of smem_offsets), formed on the fly above;
// V and A in LDS up to SC_LDS_NW for the formulas with wave tasks
template <typename T, int F, int MODE>
static int launch_scan(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist) {
    constexpr bool staged = F == WBMI_FORMULA_DEROMEGA || F == WBMI_FORMULA_DERMORB || F == WBMI_FORMULA_OMEGA;
    if constexpr (staged) {
        // nw = 32 (a single 32-column block, the spin blocks of the spinor models) has its own
        // Fermi-bin kernels with the size fixed at compile time
        if constexpr (MODE == 0)
            if (a.stage_va && a.nw == 32) return launch_scan_t<T, F, true, true, MODE, 32>(ctx, nk, a, E, X, hist);
        if (a.stage_va) return launch_scan_t<T, F, true, true, MODE>(ctx, nk, a, E, X, hist);
    }
    if (a.nw <= SC_M_NW) return launch_scan_t<T, F, true, false, MODE>(ctx, nk, a, E, X, hist);
    return launch_scan_t<T, F, false, false, MODE>(ctx, nk, a, E, X, hist);
}

template <typename T, int MODE>
static int dispatch_scan(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist) {
    switch (a.p.formula) {
        case WBMI_FORMULA_IDENTITY: return launch_scan<T, WBMI_FORMULA_IDENTITY, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_OMEGA: return launch_scan<T, WBMI_FORMULA_OMEGA, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_DEROMEGA: return launch_scan<T, WBMI_FORMULA_DEROMEGA, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_VELOMEGA: return launch_scan<T, WBMI_FORMULA_VELOMEGA, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_MORB_HPM: return launch_scan<T, WBMI_FORMULA_MORB_HPM, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_VELHPLUS: return launch_scan<T, WBMI_FORMULA_VELHPLUS, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_DERMORB: return launch_scan<T, WBMI_FORMULA_DERMORB, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_DERSPIN: return launch_scan<T, WBMI_FORMULA_DERSPIN, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_VELSPIN: return launch_scan<T, WBMI_FORMULA_VELSPIN, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_SPIN: return launch_scan<T, WBMI_FORMULA_SPIN, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_INVMASS: return launch_scan<T, WBMI_FORMULA_INVMASS, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_VELVEL: return launch_scan<T, WBMI_FORMULA_VELVEL, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_VELVELVEL: return launch_scan<T, WBMI_FORMULA_VELVELVEL, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_MASSVEL: return launch_scan<T, WBMI_FORMULA_MASSVEL, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_HALL_SURF: return launch_scan<T, WBMI_FORMULA_HALL_SURF, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_HALL_SEA: return launch_scan<T, WBMI_FORMULA_HALL_SEA, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_BANDVALUES: return launch_scan<T, WBMI_FORMULA_BANDVALUES, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_OMEGASPIN: return launch_scan<T, WBMI_FORMULA_OMEGASPIN, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_OMEGAOMEGA: return launch_scan<T, WBMI_FORMULA_OMEGAOMEGA, MODE>(ctx, nk, a, E, X, hist);
        case WBMI_FORMULA_OMEGAHPLUS: return launch_scan<T, WBMI_FORMULA_OMEGAHPLUS, MODE>(ctx, nk, a, E, X, hist);
        default: return wbmi_fail(ctx, WBMI_ERR_UNSUPPORTED, "wbmi_static_scan: formula %d", a.p.formula);
    }
}

// one translation unit per (precision, Fermi-bin | tetrahedron) instantiates the kernels
int scan_dispatch_f64(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
int scan_dispatch_f32(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
int scan_dispatch_tetra_f64(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
int scan_dispatch_tetra_f32(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
int scan_dispatch_tab_f64(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
int scan_dispatch_tab_f32(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist);
