/*
 * wbmi.h -- C ABI of the MI355X Berry-quantity library (libwbmi.so, gfx950).
 *
 * The library evaluates one K-block of the wannier-berri Berry path:
 *
 *     R-space matrices --(R->k)--> k-space matrices --(eigh)--> E, U
 *     --(U^+ X U)--> band-basis matrices --(covariant formula + Fermi scan)-->
 *     FP64 Fermi-level histogram (accumulated: this is the k-sum)
 *
 * Conventions
 *   - Every array argument named *_dev is a caller-owned DEVICE pointer
 *     (e.g. a torch tensor's data_ptr()), contiguous, row-major.  Arrays named
 *     *_host are host pointers read during the call.
 *   - Complex numbers are interleaved (re, im).  `prec` selects the element
 *     type of complex k-space arrays: WBMI_F64 = complex128, WBMI_F32 = complex64.
 *     Energies, histograms and all inputs of the Fermi scan are always FP64.
 *   - Work is enqueued on the context's stream; calls return after enqueueing
 *     (wbmi_sync waits).  Return 0 on success, a WBMI_ERR_* code otherwise;
 *     wbmi_last_error() gives the message.  Argument errors mirror the
 *     reference's ValueError / AssertionError cases and are detected on the
 *     host before any launch.
 *   - One context per host thread and stream; contexts share nothing.
 */
#ifndef WBMI_H
#define WBMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WBMI_OK 0
#define WBMI_ERR_ARG 1
#define WBMI_ERR_HIP 2
#define WBMI_ERR_UNSUPPORTED 3

#define WBMI_F64 0
#define WBMI_F32 1
/* wbmi_rotate only: complex64 storage, products on the fp16 matrix cores with operands
 * split in two halves (about FP32 accuracy) for nw <= 32; WBMI_F32 above */
#define WBMI_F32_SPLIT 2

/* formula ids for wbmi_static_scan */
#define WBMI_FORMULA_IDENTITY 0   /* formula/covariant.py:10-22            */
#define WBMI_FORMULA_OMEGA 1      /* formula/covariant.py:161-203 (AHC)    */
#define WBMI_FORMULA_DEROMEGA 2   /* formula/covariant.py:212-259 (BCD)    */
#define WBMI_FORMULA_VELOMEGA 3   /* formula/covariant.py:798-801          */
#define WBMI_FORMULA_MORB_HPM 4   /* formula/covariant.py:424-449          */
#define WBMI_FORMULA_DERMORB 5    /* formula/covariant.py:522-552          */
#define WBMI_FORMULA_VELHPLUS 6   /* formula/covariant.py:804-808          */
#define WBMI_FORMULA_DERSPIN 7    /* formula/covariant.py:338-342 (GME spin, sea) */
#define WBMI_FORMULA_VELSPIN 8    /* formula/covariant.py:811-814 (GME spin, surface) */
#define WBMI_FORMULA_SPIN 9       /* formula/covariant.py:331-335 (Spin: Re S^s_nn)   */
#define WBMI_FORMULA_INVMASS 10   /* formula/elementary.py:28-33 (Ohmic, Fermi sea)  */
#define WBMI_FORMULA_VELVEL 11    /* formula/covariant.py:817-820 (Ohmic, Fermi surface) */
/* rank-3 formulas: 27 components, computed 9 at a time; wbmi_scan_params.comp_slice = a
 * selects the components [a][b][c] of one call (hist / out rows of 9, index b*3+c) */
#define WBMI_FORMULA_VELVELVEL 12 /* formula/covariant.py:823-827 (NLDrude, f'')     */
#define WBMI_FORMULA_MASSVEL 13   /* formula/covariant.py:830-833 (NLDrude, Fermi surface) */
/* classic Hall conductivity, the rank-4 products contracted to [j][i] as
 * calculators/static.py:407-446 contracts them (9 components) */
#define WBMI_FORMULA_HALL_SURF 14 /* VelMassVel, covariant.py:842-846, static.py:407-424 */
#define WBMI_FORMULA_HALL_SEA 15  /* MassMass, covariant.py:836-839, static.py:427-446   */
/* per-band values computed by the caller: Xbar_dev is float64 [nk, nw, nch] (nch = 9 or 27,
 * `prec` is not read), and a band's value is row comp_slice * 9 .. + 8 of its nch; the
 * trace of a group is the sum over its bands (additive), as for the other formulas */
#define WBMI_FORMULA_BANDVALUES 16
/* products of the group blocks of Omega with the spin, Omega and Morb_Hpm (sign +1)
 * (formula/covariant.py:849-866), Fermi-surface Zeeman terms of the AHC (static.py:605-645) */
#define WBMI_FORMULA_OMEGASPIN 17
#define WBMI_FORMULA_OMEGAOMEGA 18
#define WBMI_FORMULA_OMEGAHPLUS 19

/* channel roles: index into wbmi_scan_params.ch (value = channel slot in the
 * per-k stack, -1 = absent).  Two-index roles are row-major over (x, y). */
#define WBMI_ROLE_V 0     /* Xbar('Ham',1)[..,a]            3 */
#define WBMI_ROLE_W 3     /* Xbar('Ham',2)[..,a,b]          9 */
#define WBMI_ROLE_A 12    /* Xbar('AA',0)[..,a]             3 */
#define WBMI_ROLE_AD 15   /* Xbar('AA',1)[..,a,d]           9 */
#define WBMI_ROLE_O 24    /* Xbar('rotAA',0)[..,c]          3 */
#define WBMI_ROLE_OD 27   /* Xbar('rotAA',1)[..,c,d]        9 */
#define WBMI_ROLE_B 36    /* Xbar('BB',0)[..,a]             3 */
#define WBMI_ROLE_BD 39   /* Xbar('BB',1)[..,a,d]           9 */
#define WBMI_ROLE_C 48    /* Xbar('CC',0)[..,a]             3 */
#define WBMI_ROLE_CD 51   /* Xbar('CC',1)[..,a,d]           9 */
#define WBMI_ROLE_S 60    /* Xbar('SS',0)[..,b]             3 */
#define WBMI_ROLE_SD 63   /* Xbar('SS',1)[..,b,d]           9 */
#define WBMI_NROLES 72

typedef struct wbmi_ctx wbmi_ctx;
typedef struct wbmi_r2k_plan wbmi_r2k_plan;

/* Parameters of one Fermi-level scan (calculators/static.py:26-58). */
typedef struct wbmi_scan_params {
    int formula;          /* WBMI_FORMULA_*                                   */
    int sign;             /* Morb_Hpm / DerMorb sign (+1, -1, 0)              */
    int internal_terms;   /* formula/formula.py:11-25                         */
    int external_terms;   /* already 0 when force_internal_terms_only         */
    int fder;             /* 0 = Fermi sea (adds the below-range group)       */
    int nEF_extra;        /* number of extended Fermi levels                  */
    double EFmin, EFmax, dEF;
    double degen_thresh;  /* calculator.py:20                                 */
    int degen_Kramers;
    int ch[WBMI_NROLES];
    int comp_slice;       /* rank-3 formulas: first tensor index of this call  */
    /* select_bands (calculators/static.py:39,133,136; utility.py:398-403): device
     * [nw] float64, the number of times each band is listed; a band group adds its
     * value times the sum over its bands / its size (groups with none listed add
     * nothing).  NULL = every band.  fder >= 1 only, in Fermi-level bins and with
     * tetrahedra (tetrahedron.py:238-255; the reference refuses it for the Fermi
     * sea and the hole-like group, data_K.py:179-180); the table entry points
     * refuse a non-NULL value.                                                   */
    const double* band_weight;
    /* Emin / Emax of the tetrahedron method (grid/tetrahedron.py:246-258), read when
     * erange != 0: the Fermi-sea group (fder 0) leaves out the bands whose energies over
     * centre and corners all lie below Emin; the hole-like group (fder -1) those above
     * Emax.  Fermi-level bins ignore them, as the reference does (data_K.py:172-186). */
    int erange;
    double Emin, Emax;
} wbmi_scan_params;

/* Parameters of the Kubo optical conductivity (calculators/dynamic.py:26-196). */
typedef struct wbmi_optcond_params {
    int nEF, nomega;
    const double* Efermi; /* device [nEF]                                     */
    const double* omega;  /* device [nomega]                                  */
    double EFmin, EFmax;  /* min / max of Efermi (occupation window)          */
    double omega_min, omega_max; /* min / max of omega (JDOS pair window)      */
    double kBT;           /* 0: step occupations (utility.py:172-182)         */
    double eta;           /* smr_fixed_width                                  */
    int smr_gaussian;     /* 1: Gaussian, 0: Lorentzian smearing              */
    double degen_thresh;
    int degen_Kramers;
    int external_terms;   /* A = i D_H + Abar (data_K.py:329-334)             */
    int ch[WBMI_NROLES];  /* channel slots: WBMI_ROLE_V (and WBMI_ROLE_A)     */
} wbmi_optcond_params;

/* Parameters of the spin Hall conductivity (calculators/dynamic.py:204-241):
 * the optical-conductivity parameters (V and A roles; with external_terms 0 A
 * is not read), the spin current `kind` (formula/covariant.py:703-756) and the
 * channel slots of Xbar('SS')[s] and, for Ryoo, Xbar('SA')[a, s],
 * Xbar('SHA')[a, s], for Qiao Xbar('SH')[s], Xbar('SR')[a, s],
 * Xbar('SHR')[a, s] (row-major (a, s)).  Ryoo and Qiao need external terms. */
#define WBMI_SHC_RYOO 0
#define WBMI_SHC_QIAO 1
#define WBMI_SHC_SIMPLE 2   /* {S, v} / 2 with the Velocity (covariant.py:721-727) */
typedef struct wbmi_shc_params {
    wbmi_optcond_params base;
    int kind;
    int ch_S[3];
    int ch_SA[9];
    int ch_SHA[9];
    int ch_SH[3];
    int ch_SR[9];
    int ch_SHR[9];
} wbmi_shc_params;

/* ---- context ------------------------------------------------------------ */
int wbmi_ctx_create(int device, void* stream, wbmi_ctx** out);
int wbmi_ctx_destroy(wbmi_ctx* ctx);
int wbmi_ctx_set_stream(wbmi_ctx* ctx, void* stream);
int wbmi_sync(wbmi_ctx* ctx);
const char* wbmi_last_error(const wbmi_ctx* ctx);
int wbmi_version(void);

/* ---- R-space channel stack ---------------------------------------------
 * Builds out[R, ch, m, n] = sum_t sign_t * base[R, m, n, a_t] * prod_(d in t) i r_d
 * with r = R_cart + t_n - t_m (the Wannier-centre-shifted R vector).
 * Replaces Rvectors.derivative (fourier/rvectors.py:487-494), the rotAA curl
 * (data_K/data_K_R.py:48-51) and the per-quantity copies of get_R_mat.
 *   cRshift_dev : float64 [nR, nw, nw, 3]
 *   base_dev    : complex128 [nR, nw, nw, nbase]
 *   spec_host   : int32 [nch, 2, 4] = (sign, a, d1, d2) per term; sign 0 =
 *                 unused term, d = -1 = no derivative.
 *   out_dev     : complex128 [nR, nch, nw, nw]                                  */
int wbmi_build_channels(wbmi_ctx* ctx, int nR, int nw, const double* cRshift_dev,
                        const void* base_dev, int nbase, const int32_t* spec_host,
                        int nch, void* out_dev);

/* ---- R -> k on an FFT sub-grid -------------------------------------------
 * Replaces FFT_R_to_k grid mode (fourier/fft.py:58-192) together with
 * Rvectors.apply_expdK (fourier/rvectors.py:468,477-485):
 *   out[k, e] = sum_R X[R, e] exp(2 pi i (j/NKFFT + dK).R),  k = (j0*N1+j1)*N2+j2
 * as three separable DFT passes over the R bounding box.  `nelem` is the
 * number of complex elements per R (nch*nw*nw).                             */
int wbmi_r2k_plan_create(wbmi_ctx* ctx, int nR, const int32_t* iRvec_host,
                         const int32_t NKFFT[3], int64_t nelem, wbmi_r2k_plan** out);
int wbmi_r2k_plan_destroy(wbmi_r2k_plan* plan);
/* bytes of box storage (complex128) and of the per-call workspace (the pass phase tables and
 * the intermediate pass outputs): calls that do not share a workspace (and, for an aliased
 * plan, a box) may run concurrently on different streams                     */
int wbmi_r2k_plan_sizes(const wbmi_r2k_plan* plan, size_t* box_bytes, size_t* work_bytes);
/* scatter X[R, e] (complex128) into the box; needed once per system when no
 * axis aliases, and is repeated by wbmi_r2k_exec otherwise                  */
int wbmi_r2k_box(wbmi_ctx* ctx, const wbmi_r2k_plan* plan, const void* XR_dev, void* box_dev);
/* out_dev: [nk, nelem] complex of `prec`; XR_dev is read only when an axis
 * aliases (R extent > NKFFT), in which case box_dev is rebuilt per call.    */
int wbmi_r2k_exec(wbmi_ctx* ctx, const wbmi_r2k_plan* plan, const void* XR_dev,
                  void* box_dev, const double dK[3], void* work_dev, void* out_dev, int prec);
/* wbmi_r2k_exec in two steps, for plans whose box does not depend on dK (no axis aliases):
 * the z pass box -> w1 ([L0*L1][N2][nelem] complex of `prec`, wbmi_r2k_w1_bytes) depends only
 * on dK[2], so a caller that evaluates many K-blocks of one grid can keep w1 resident per
 * distinct dK[2] and run only the y and x passes per K-block (wbmi_r2k_yx).  The reference
 * recomputes the whole FFT per K-block (fourier/fft.py:58-192, grid/Kpoint.py K-block loop);
 * the split is this library's, the result equals wbmi_r2k_exec's.  Both steps take the
 * workspace of wbmi_r2k_plan_sizes.                                                      */
int wbmi_r2k_w1_bytes(const wbmi_r2k_plan* plan, int prec, size_t* bytes);
int wbmi_r2k_zpass(wbmi_ctx* ctx, const wbmi_r2k_plan* plan, const void* box_dev, const double dK[3],
                   void* work_dev, void* w1_dev, int prec);
int wbmi_r2k_yx(wbmi_ctx* ctx, const wbmi_r2k_plan* plan, const void* w1_dev, const double dK[3],
                void* work_dev, void* out_dev, int prec);

/* ---- R -> k on an arbitrary k-list ---------------------------------------
 * Phase-matrix transform (fourier/fft.py:125-130,170-171), no expdK phase.
 *   klist_dev : float64 [nk, 3];  iRvec_dev : int32 [nR, 3]
 *   XR_dev    : complex128 [nR, nelem];  out_dev : [nk, nelem] (prec)       */
int wbmi_r2k_klist(wbmi_ctx* ctx, int nR, const int32_t* iRvec_dev, int64_t nelem,
                   const void* XR_dev, int nk, const double* klist_dev, void* out_dev, int prec);
/* The same transform with an explicit workgroup tile (0 .. WBMI_KLIST_NTILES-1; 0 is what
 * wbmi_r2k_klist uses): no counterpart in the reference, for the per-nw autotuner
 * (wannierberri_amd/autotune.py) that times the tiles on the caller's shapes.
 * wbmi_r2k_klist_tile_shape gives a tile's k-points x elements per workgroup.         */
#define WBMI_KLIST_NTILES 5
int wbmi_r2k_klist_tile(wbmi_ctx* ctx, int nR, const int32_t* iRvec_dev, int64_t nelem,
                        const void* XR_dev, int nk, const double* klist_dev, void* out_dev, int prec,
                        int tile);
int wbmi_r2k_klist_tile_shape(int tile, int* k_rows, int* elements);

/* ---- batched Hermitian eigensolver ----------------------------------------
 * Replaces np.linalg.eigh in Data_K.E_K (data_K/data_K.py:211-218).
 * H is read as 0.5 (H + H^+) (data_K_R.py:26 hermitises HH_K).
 *   H_dev : [nk, Hstride] complex (prec), matrix k at H_dev + k*Hstride
 *   E_dev : float64 [nk, nw] ascending;  U_dev : [nk, nw, nw] complex (prec),
 *   column j = eigenvector of E[j].  nw <= 128 (nw <= 64 without refinement).
 *   refine = n in 1..16: FP32 Jacobi of the FP64 Hamiltonian (H_dev for
 *   prec = WBMI_F64; H64_dev, complex128 with the same stride, for
 *   prec = WBMI_F32), then FP64 Ogita-Aishima steps (Japan J. Indust. Appl.
 *   Math. 35 (2018), Alg. 1) until the correction has ||E||_F below 1e-8
 *   (complex64 vectors) or 1e-10 (complex128 vectors), at most n of them;
 *   complex128 vectors get a final Newton-Schulz orthonormalisation.
 *   E are the FP64 Rayleigh quotients of the refined vectors.  Separated
 *   spectra take 2-3 steps; a pair closer than the first step's threshold
 *   (~1e-4 |H| at nw = 32) about ten.  For nw > 64 the FP32 Jacobi runs one
 *   workgroup per chunk of eigenvector rows.
 *   refine = 0: plain Jacobi in the arithmetic of `prec` (nw <= 64).
 *   work_dev / work_bytes: device workspace of at least
 *   wbmi_eigh_work_bytes(...) bytes (nonzero only for refinement at
 *   nw > 44); may be NULL / 0 otherwise.                                   */
int wbmi_eigh(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride,
              double* E_dev, void* U_dev, int prec, int refine, const void* H64_dev,
              void* work_dev, size_t work_bytes);
size_t wbmi_eigh_work_bytes(int nk, int nw, int prec, int refine);

/* ---- the eigensolver without allocation ------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  wbmi_eigh with its
 * workspace taken from a fixed scratch array in the library, one per GPU, of
 * wbmi_eigh_noalloc_max_bytes() = 64 MiB: sizes whose wbmi_eigh_work_bytes
 * exceed it are refused (at nw = 128 with refinement: 64 k-points per call;
 * nw <= 44 needs none).  Allocates nothing; calls on one GPU share the scratch,
 * so calls that overlap on one GPU race.                                    */
int wbmi_eigh_noalloc(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride,
                      double* E_dev, void* U_dev, int prec, int refine, const void* H64_dev);
size_t wbmi_eigh_noalloc_max_bytes(void);

/* ---- batched eigensolver with a starting basis ---------------------------
 * No counterpart in the reference (its eigh starts from nothing,
 * data_K.py:211-218): the same result as wbmi_eigh with refine >= 1, started
 * from U0_dev, [nk, nw, nw] complex of `prec` (columns = approximate
 * eigenvectors, e.g. the previous K-block's U at the same k index, which is
 * one grid step away).  For nw + (nw & 1) in {16, 32} the FP32 Jacobi runs on
 * U0^+ H U0 with the rotations accumulated onto U0 (about half the sweeps);
 * other sizes start from the identity.  Any unitary U0 gives the same
 * eigen-decomposition.  U0_dev may equal U_dev (each k-point reads its guess
 * before it writes); U0_dev = NULL is wbmi_eigh.                           */
int wbmi_eigh_guess(wbmi_ctx* ctx, int nk, int nw, const void* H_dev, int64_t Hstride, double* E_dev,
                    void* U_dev, int prec, int refine, const void* H64_dev, const void* U0_dev, void* work_dev,
                    size_t work_bytes);

/* ---- Kubo optical conductivity ---------------------------------------------
 * Replaces OpticalConductivity.__call__ (calculators/dynamic.py:26-196 with
 * Formula_OptCond 170-181) on one K-block: for every pair (m, n) of band
 * groups, M^{ab} = sum_{i in m, j in n} i A^a_ij A^b_ji, A = i D_H (+ Abar),
 * added as (f(E_n) - f(E_m)) (E_n - E_m) / (E_n - E_m - w - i eta) M^{ab}
 * (Gaussian: imaginary part pi G) into out_dev [nEF, nomega, 3, 3] complex128
 * (interleaved doubles), accumulated across calls; the caller scales by
 * factor_opt / (nk V).  nw <= 32 (larger: wbmi_optical_conductivity_ws). */
/* Joint density of states (calculators/dynamic.py:145-167): per k, band-group
 * pairs (m, n) with E_m < EFmax, E_n > EFmin and omega_min - 5 eta <
 * E_m - E_n < omega_max + 5 eta add |m| |n| (f(E_n) - f(E_m)) smear(E_m - E_n - w).
 *   E_dev   : float64 [nk, nw]
 *   params  : the optical-conductivity parameters (channels not read)
 *   out_dev : float64 [nEF, nomega], accumulated (zero it first); the caller
 *             scales by 1 / (nk V).                                           */
int wbmi_jdos(wbmi_ctx* ctx, int nk, int nw, const double* E_dev, const wbmi_optcond_params* params,
              double* out_dev);

/* Kubo spin Hall conductivity with the Ryoo or Qiao spin current
 * (dynamic.py:204-241, formula/covariant.py:689-756; external terms):
 *   J^{as} = (J' + J'^+)/2,
 *   Ryoo: J'^{as}_mn = -i (E_n SA^{as}_mn - SHA^{as}_mn) + sum_l SS^s_ml V^a_ln,
 *   Qiao: J'^{as}_mn = v^a_n SS^s_mn + E_n (-i SR^{as}_mn + sum_l SS^s_ml D^a_ln)
 *                      - (-i SHR^{as}_mn + sum_l SH^s_ml D^a_ln),
 *         v^a_n = Re V^a_nn, D = D_H,
 *   B^b = -i (i D_H + Abar)^b,
 *   out[EF][w][a][b][s] += (f(E_n) - f(E_m)) sum_{i in m, j in n}
 *                          Im(J^{as}_ij B^b_ji) / (2 (E_m - E_n - w - i eta))
 * over band-group pairs (m, n) (Gaussian: imaginary part pi G / 2).
 *   out_dev  : float64 [nEF, nomega, 27, 2] (complex), accumulated; the caller
 *              scales by factor_shc / (nk V).
 *   work_dev : wbmi_spin_hall_work_bytes(nk, nw) bytes (J per k).            */
size_t wbmi_spin_hall_work_bytes(int nk, int nw);
int wbmi_spin_hall(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                   int prec, const wbmi_shc_params* params, void* work_dev, size_t work_bytes,
                   double* out_dev);

/* Shift current (dynamic.py:244-324) and injection current (dynamic.py:327-361)
 * with the pair walk of wbmi_spin_hall:
 *   shift:     out[EF][w][a][b][c] += (f(E_n) - f(E_m)) sum_{i in m, j in n}
 *              (-Im A'^{ca}_ij A^b_ji - Im A'^{ba}_ij A^c_ji) (g(E_m - E_n - w) + g(E_n - E_m - w)),
 *              A' the generalised derivative of A regularised by sc_eta (roles V, W, and
 *              with external terms A, AD);
 *   injection: out[EF][w][a][b][c] += (f(E_n) - f(E_m)) sum (v^a_i - v^a_j) A^b_ij A^c_ji g(E_m - E_n - w),
 * A = i D_H (+ Abar), g = the smearing function.
 *   out_dev : float64 [nEF, nomega, 27, 2] (complex), accumulated; the caller scales.
 *   work_dev: wbmi_spin_hall_work_bytes(nk, nw) bytes.                        */
int wbmi_shift_current(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                       int prec, const wbmi_optcond_params* params, double sc_eta, void* work_dev,
                       size_t work_bytes, double* out_dev);
int wbmi_injection_current(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                           int prec, const wbmi_optcond_params* params, void* work_dev, size_t work_bytes,
                           double* out_dev);

int wbmi_optical_conductivity(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                              int prec, const wbmi_optcond_params* params, double* out_dev);
/* The same for nw > 32 (up to 128), where A does not fit in LDS: work_dev holds A for every k,
 * wbmi_optical_conductivity_work_bytes(nk, nw) bytes (0 for nw <= 32, where
 * wbmi_optical_conductivity needs none).  The pair-walk calculators (wbmi_spin_hall,
 * wbmi_shift_current, wbmi_injection_current) take nw <= 128 the same way, in the
 * larger workspace that wbmi_spin_hall_work_bytes reports for nw > 32.            */
size_t wbmi_optical_conductivity_work_bytes(int nk, int nw);
int wbmi_optical_conductivity_ws(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                                 int prec, const wbmi_optcond_params* params, void* work_dev, size_t work_bytes,
                                 double* out_dev);

/* ---- gauge rotation U^+ X U -----------------------------------------------
 * Replaces Data_K._rotate (data_K/data_K.py:130-132) for the channels of a
 * per-k stack X[k, nch, nw, nw] (in place).  herm_host[c] = 1 forces the
 * channel Hermitian first (0.5 (X + X^+), fourier/fft.py:183-186,
 * data_K_R.py:87).  Channels with skip_host[c] = 1 are left untouched.
 * Complex64 channels (prec = WBMI_F32) are rotated with FP64 arithmetic;
 * only the stored result is rounded.  Flag arrays may be NULL (all 0).
 * nw <= 128; work_dev / work_bytes: device workspace of at least
 * wbmi_rotate_work_bytes(...) bytes (nonzero only for nw > 64).            */
int wbmi_rotate(wbmi_ctx* ctx, int nk, int nw, int nch, const void* U_dev,
                void* X_dev, const int32_t* herm_host, const int32_t* skip_host, int prec,
                void* work_dev, size_t work_bytes);
size_t wbmi_rotate_work_bytes(int nk, int nw, int prec);

/* ---- covariant formula traces + Fermi scan --------------------------------
 * Replaces the per-(k, band-group) loop and Fermi scan of
 * StaticCalculator.__call__ (calculators/static.py:60-136), band grouping
 * Data_K.get_bands_in_range_groups (data_K/data_K.py:172-194,
 * grid/tetrahedron.py:132-162) and the formula traces.
 *   E_dev    : float64 [nk, nw]
 *   Xbar_dev : [nk, nch, nw, nw] rotated channels (prec)
 *   hist_dev : float64 [nEF_extra, 3^ndim], accumulated with weight * value
 *              at bin ceil((E-EFmin)/dEF) (bin 0 for the below-range group).
 * `weight` carries the k-sum factor (K-block factor / nk).                   */
int wbmi_static_scan(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                     const void* Xbar_dev, int prec, const wbmi_scan_params* params,
                     double weight, double* hist_dev);

/* k-resolved scan (static.py:60-101,156-164, k_resolved=True): k-point k
 * accumulates into its own histogram, hist_dev float64 [nk, nEF_extra, 3^ndim]. */
int wbmi_static_scan_k(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev,
                       const void* Xbar_dev, int prec, const wbmi_scan_params* params,
                       double weight, double* hist_dev);

/* ---- the same with tetrahedron weights -----------------------------------
 * Replaces the tetra=True branch of StaticCalculator.__call__
 * (calculators/static.py:80-85,119-125) with TetraWeightsParal
 * (grid/tetrahedron.py:185-281): each group's trace is weighted over the
 * Fermi levels by the mean of its bands' 12-tetrahedron occupations (fder 0)
 * or their fder-th derivative (1..3).  Groups are kept, and the fder-0 group
 * of bands below the range is formed, from each band's energy range over the
 * centre and the 8 corners.
 *   ncorner      : 8 (parallelepiped K-blocks, 12 tetrahedra around each k) or
 *                  4 (tetrahedral K-blocks of GridTetra, one tetrahedron)
 *   Ecorners_dev : float64 [ncorner, nk, nw] band energies at the corners:
 *                  k + ((2i-1), (2j-1), (2l-1)) * dK_fullBZ / 2, corner i*4 + j*2 + l
 *                  (data_K/data_K_R.py:126-141), or k + vertex_i (99-119)
 *   Efermi_dev   : float64 [nEF], ascending
 *   params       : fder = 0..3, or -1 for hole_like (weights 1 - occupation and
 *                  the group of bands above the range, tetrahedron.py:165-172);
 *                  EFmin = Efermi[0], EFmax = Efermi[nEF-1];
 *                  nEF_extra and dEF are not read
 *   hist_dev     : float64 [2 * nEF, 3^ndim]: rows [0, nEF) are steps
 *                  (prefix-summed by wbmi_finalize_tetra), rows [nEF, 2 nEF)
 *                  are added as they are.                                    */
int wbmi_static_scan_tetra(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                           const double* Ecorners_dev, const void* Xbar_dev, int prec,
                           const wbmi_scan_params* params, const double* Efermi_dev, int nEF,
                           double weight, double* hist_dev);

/* ---- per-band table (k-resolved tabulation) ------------------------------
 * Replaces Tabulator.__call__ (calculators/tabulate.py:20-48): for every k and
 * band, the trace of its band group (degen_thresh, degen_Kramers; inn = the
 * group, out = the rest) divided by the group size; for WBMI_FORMULA_IDENTITY
 * the group's mean energy (formula Eavln).  params: formula, sign,
 * internal/external terms, degen_thresh, degen_Kramers, ch (the Fermi-level
 * fields are not read).
 *   out_dev : float64 [nk, nw, 3^ndim] (ndim 0 for the identity formula)     */
int wbmi_tabulate(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                  int prec, const wbmi_scan_params* params, double* out_dev);

/* Per-band table of a channel diagonal (tabulate.py Velocity and Spin:
 * formula/covariant.py:321-336, whose group traces are sums of diagonal
 * elements): out[k][n][c] = mean over the band group of n of Re X^{ch3[c]}_mm.
 *   ch3_host : int32 [3] channel slots (host);  out_dev : float64 [nk, nw, 3] */
int wbmi_tabulate_diag(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                       int prec, const int32_t* ch3_host, double degen_thresh, int degen_Kramers,
                       double* out_dev);

/* k-resolved: hist_dev [nk, 2 nEF, 3^ndim], one histogram per k-point */
int wbmi_static_scan_tetra_k(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, int ncorner,
                             const double* Ecorners_dev, const void* Xbar_dev, int prec,
                             const wbmi_scan_params* params, const double* Efermi_dev, int nEF,
                             double weight, double* hist_dev);
/* wbmi_finalize_tetra of nk histograms [nk, 2 nEF, ncomp] -> out [nk, nEF, ncomp] */
int wbmi_finalize_tetra_k(wbmi_ctx* ctx, double* hist_dev, int nk, int nEF, int ncomp, double scale,
                          double* out_dev);

/* Per-band table of the inverse effective mass (tabulate.py InvMass,
 * formula/elementary.py:28-33): out[k][n][b*3+d] = group mean of
 * Re [W^{bd}_nn + sum_{l outside the group} (V^b_nl D^d_ln - D^d_nl V^b_ln)].
 *   chV3_host, chW9_host : int32 channel slots of V and Xbar('Ham', 2) (host)
 *   out_dev : float64 [nk, nw, 9]                                               */
int wbmi_tabulate_invmass(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                          int prec, const int32_t* chV3_host, const int32_t* chW9_host,
                          double degen_thresh, int degen_Kramers, double* out_dev);

/* Per-band spin Berry curvature (tabulate.py:218-221 SpinBerry, formula/covariant.py:759-789
 * SpinOmega): out[k][n][a*9+b*3+s] = group mean over m of
 *   -2 Im sum_{l outside the group} J^{as}_ml / (E_m - E_l) (D^b_lm - i Abar^b_lm),
 * J the spin current of params->kind (as in wbmi_spin_hall), D = D_H, Abar only with
 * params->base.external_terms.  Reads params->base.ch (V, A), degen_thresh and
 * degen_Kramers; Efermi / omega are not used.
 *   work_dev: wbmi_spin_hall_work_bytes(nk, nw) bytes (J per k).
 *   out_dev : float64 [nk, nw, 27]                                                 */
int wbmi_tabulate_spin_berry(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                             int prec, const wbmi_shc_params* params, void* work_dev, size_t work_bytes,
                             double* out_dev);
/* Third k-derivative of the band energies (tabulate.py Der3E, formula/covariant.py:126-150 with
 * InvMass, DerDcov, DerWln of formula/elementary.py:28-70), nw <= 32:
 *   mode 0: out[k][n][a*9+b*3+c] = group mean of Re Der3E.nn[m, m, a, b, c] (the tabulator),
 *   mode 1: the band's own Re Der3E.nn[n, n, a, b, c], its share of the group's trace (input of
 *           NLDrude_FermiSea, calculators/static.py:519-529, binned with WBMI_FORMULA_BANDVALUES).
 *   chV3_host, chW9_host, chW27_host : int32 channel slots of Xbar('Ham', 1), ('Ham', 2), ('Ham', 3)
 *   out_dev : float64 [nk, nw, 27]                                                           */
int wbmi_der3e_bands(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev, int prec,
                     const int32_t* chV3_host, const int32_t* chW9_host, const int32_t* chW27_host,
                     double degen_thresh, int degen_Kramers, int mode, double* out_dev);

/* The same per band, before the group mean: out[k][n][abs] = c_n^{abs}, the band's share of its
 * group's trace (the trace of a group is the sum over its bands).  Input of the static spin Hall
 * conductivity (calculators/static.py:647-660), which bins these with WBMI_FORMULA_BANDVALUES. */
int wbmi_spin_berry_bands(wbmi_ctx* ctx, int nk, int nw, int nch, const double* E_dev, const void* Xbar_dev,
                          int prec, const wbmi_shc_params* params, void* work_dev, size_t work_bytes,
                          double* out_dev);

/* out_dev [nEF, ncomp] = scale * (prefix sums of rows [0, nEF) + rows [nEF, 2 nEF))
 * of a wbmi_static_scan_tetra histogram (static.py:149-155, no fder stencil).
 * The step rows of hist_dev are overwritten with their prefix sums.         */
int wbmi_finalize_tetra(wbmi_ctx* ctx, double* hist_dev, int nEF, int ncomp, double scale,
                        double* out_dev);

/* ---- histogram -> EnergyResult data ---------------------------------------
 * Prefix sum over the Fermi levels, fder finite-difference stencil and scaling
 * (calculators/static.py:126-155).  out_dev: float64 [nEF_extra-2*extra, ncomp].
 * hist_dev is overwritten with its prefix sums (any nEF_extra).             */
/* wbmi_finalize of nk histograms [nk, nEF_extra, ncomp] -> out [nk, nEF, ncomp]. */
int wbmi_finalize_k(wbmi_ctx* ctx, double* hist_dev, int nk, int nEF_extra, int ncomp, int fder,
                    double dEF, double scale, double* out_dev);
int wbmi_finalize(wbmi_ctx* ctx, double* hist_dev, int nEF_extra, int ncomp,
                  int fder, double dEF, double scale, double* out_dev);

#ifdef __cplusplus
}
#endif
#endif /* WBMI_H */
