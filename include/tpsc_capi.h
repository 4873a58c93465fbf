/*
 * tpsc_capi.h -- C ABI of the MI355X-native TPSC hot path (libtpsc_hip.so).
 *
 * The reference (dbergeron1/TPSC_conductivity) has no plugin API: its boundary is the
 * public method surface of the class chain cond_opt > green > chi > chi0 > green0 >
 * hamiltonien.  Each entry point below replaces one of those methods, keeps its order
 * of use and its semantics, and is what an FFI (ctypes, cgo, JNI) would bind.  The
 * reference's own callers are tpsc_main.cpp:61-114 and the lazy chain at
 * cond_opt.cpp:2550-2560.
 *
 * Conventions: every call returns 0 on success or a negative TPSC_E_* code, with the
 * message in tpsc_last_error(ctx).  A context owns its device buffers and one HIP
 * stream; calls are synchronous for the scalars they return; a context is used from one
 * host thread at a time.  Arrays copied out by tpsc_get keep the reference layouts:
 * packed wedge idx(l,m) = m + l(l+1)/2 (m <= l), frequency/tau-major a[idx + j*nk8].
 */
#ifndef TPSC_CAPI_H
#define TPSC_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPSC_OK 0
#define TPSC_E_INVALID -1     /* bad argument or unsupported size          */
#define TPSC_E_DEVICE -2      /* HIP runtime error                         */
#define TPSC_E_NOMEM -3       /* device allocation failed                  */
#define TPSC_E_ORDER -4       /* stage called before its inputs exist      */
#define TPSC_E_NOROOT -5      /* a root finder did not converge            */

/* Model and grid, as the constructors take them:
 * cond_opt(int nk0, int Nw0, double n, double T, double eta, double params[4]={t',t'',t3,U})
 * (cond_opt.h:25, cond_opt.cpp:10; green.h:26; chi.h:25; hamiltonien.cpp:14-34). */
typedef struct tpsc_params {
    int nk0;         /* N, lattice N x N; 2^k (4..256) or 3*2^k (12..192)       */
    int Nw0;         /* total number of fermionic +-iwn; power of two, 16..4096 */
    double n;        /* density                                                */
    double T;        /* temperature                                            */
    double eta;      /* broadening, kept for interface parity                  */
    double tp, tpp, t3, U;
    int coarse_pass; /* 1: find_Usp_Uch() runs the (N/2, Nw/2) pass first, as the
                        reference does (chi.cpp:285-310); 0: full grid only      */
    int precision;   /* TPSC_PREC_FP64 or TPSC_PREC_MIXED (no reference counterpart:
                        the reference computes in FP64 only)                     */
    double t;        /* nearest-neighbour hopping; 0 means 1, the reference's fixed value
                        (hamiltonien.cpp:26,121).  A free t makes d/dt defined.  */
} tpsc_params;

/* precision values.  MIXED: the slab transforms (2D DCT/DST per tau or iqn slice) and their
 * pointwise products run in FP32 and the tau arrays (chi0(q,tau), chi0(r,tau), V, Sigma(k,tau),
 * h, f0) are stored in FP32; the tau = 0 / beta boundary slices, spline moments, tail moments,
 * pencil Fourier sums, chi0(q,iqn), the Usp/Uch/mu root reductions and the Lambda_jj Matsubara
 * sums stay in FP64, and Sigma(k,iwn) is stored assembled in FP32 complex.  tpsc_get widens
 * every array to FP64. */
#define TPSC_PREC_FP64 0
#define TPSC_PREC_MIXED 1
/* In the mixed path the chi0 stage (slabs and pencils) runs in FP64 for a point whose spin
 * enhancement 1/(1 - Usp chi0max/2) exceeds 8: chi_sp multiplies the relative error of chi0 by
 * that factor.  The point's first Usp search runs on the FP32 chi0; if it finds the enhancement
 * above 8, chi0 and the roots are redone in FP64, and later solves of the point (the same
 * parameters) start in FP64 (TPSC_S_CHI0_FP64). */

typedef struct tpsc_ctx tpsc_ctx;

/* Fields for tpsc_get (array copy-out) and tpsc_get_scalar. */
enum tpsc_field {
    TPSC_F_EK = 1,        /* eps(k) on the nk x nk quadrant            (hamiltonien.cpp:49-68) */
    TPSC_F_CHI0 = 2,      /* chi0(q,iqn), nwb*nk8                       (chi0.cpp:669-690)      */
    TPSC_F_CHIRT = 3,     /* chi0(r,tau), nwb*nk8                       (chi0.cpp:504)          */
    TPSC_F_CHI0_INF = 4,  /* 1/qn^2 tail coefficient, nk8                (chi0.cpp:673)          */
    TPSC_F_CHI0_INF2 = 5, /* 1/qn^4 tail coefficient, nk8                (chi0.cpp:682)          */
    TPSC_F_DTAU_CHIRT0 = 6,
    TPSC_F_V = 7,         /* V(r,tau), nwb*nk8                          (chi.cpp:166)           */
    TPSC_F_SIGMA_TAU = 8, /* Sigma(k,tau), (Nw+1)*nk8                   (green.cpp:1872-1885)   */
    TPSC_F_SELF = 9,      /* R(k,n) complex, (Nw/2)*nk8 interleaved re/im (green.cpp:2006)      */
    TPSC_F_SIGMA_INF2 = 10,
    TPSC_F_SIGMA_INF3 = 11,
    TPSC_F_SIGMA_INF4 = 12,
    TPSC_F_CHIJJ1 = 13,   /* bubble Lambda_jj1(iqn), nwb                (cond_opt.cpp:3784-3817) */
    TPSC_F_CHIJJ2 = 14,   /* Maki-Thompson Lambda_jj2(iqn), nwb         (cond_opt.cpp:3790-3807) */
    TPSC_F_DTAU_SIGMA = 15, /* 2*nk8                                    (green.cpp:1743-1817)   */
    TPSC_F_CHI0_TAU = 16, /* chi0(q,tau) before the tau->iqn transform (kept only after tpsc_chi0) */
    TPSC_F_SIGMA_KW = 17, /* Sigma(k,iwn), n = 0..Nw/2-1, (Nw/2)*nk8 complex interleaved re/im:
                             Sigma_inf/z + C2/z^2 + C3/z^3 + R/z^4, z = iwn (green.cpp:3039);
                             the only Sigma(k,iwn) field of TPSC_PREC_MIXED                  */
    TPSC_F_CHIJJ3 = 18,   /* Aslamazov-Larkin Lambda_jj3 on the sparse boson set, al_N values,
                             entry 0 = 0                                (cond_opt.cpp:6396-6398) */
    TPSC_F_CHIJJ_ALL = 19, /* Lambda_jj1 + Lambda_jj2 + Lambda_jj3 on that set, al_N values
                                                                        (cond_opt.cpp:6460-6467) */
    TPSC_F_AL_FREQ = 20,  /* the sparse boson indices n (as doubles), al_N values
                                                                        (cond_opt.cpp:5123-5156) */
    TPSC_F_CHIJJ1_K = 21, /* Lambda_jj1 by angle of k, nwb * N_theta_k values [b + n*N_theta_k]
                                                        (cond_opt.cpp:3140-3174, 3792-3847) */
    TPSC_F_CHIJJ2_K = 22, /* Lambda_jj2 by angle of k, the same layout  (cond_opt.cpp:3799-3809) */
    TPSC_F_CHIJJ3_K = 23, /* Lambda_jj3 by angle of k on the sparse set, al_N * N_theta_k values
                             [b + j*N_theta_k], entry j = 0 zero; only when N_theta_k > 1
                                                                        (cond_opt.cpp:6261-6398) */
    /* tangents of the last tpsc_loss_backward (no reference counterpart), parameters p = t, t',
     * mu0, U: */
    TPSC_F_GRAD_TERMS = 24,   /* 16 values [p*4 + k]: d/dp of loss term k (Re Sigma, Im Sigma,
                                 chi_sp(pi,pi), Lambda_jj; 0 for terms not in the loss) */
    TPSC_F_GRAD_DMU = 25,     /* 4 values: dmu/dp of the interacting mu (Lambda_jj term only) */
    TPSC_F_GRAD_DLAMBDA = 26, /* 4*nwb values [p*nwb + n]: d/dp [Lambda_jj1 + Lambda_jj2](iq_n)
                                 (Lambda_jj term only) */
    TPSC_F_SIGMA_BND = 27     /* Sigma(k, tau = 0) and Sigma(k, tau = beta), 2*nk8: the end samples
                                 the self-energy spline uses (FP64 in both paths; green.cpp:1822-1892) */
};
enum tpsc_scalar {
    TPSC_S_MU0 = 1, TPSC_S_CHI0MAX, TPSC_S_QMAX_L, TPSC_S_QMAX_M, TPSC_S_USP, TPSC_S_UCH,
    TPSC_S_DBLOCC, TPSC_S_DIFREL_SP, TPSC_S_DIFREL_CH, TPSC_S_SIGMA_INF, TPSC_S_TRSG,
    TPSC_S_TRSG_CORR, TPSC_S_MU, TPSC_S_TRSG2, TPSC_S_DBLOCC2, TPSC_S_KX0, TPSC_S_CHIJJ1_INF2,
    TPSC_S_CHIJJ2_INF2, TPSC_S_CHIJJ2_INF4, TPSC_S_NEVAL_SP, TPSC_S_NEVAL_CH, TPSC_S_NEVAL_MU,
    TPSC_S_NK, TPSC_S_NK8, TPSC_S_NWB, TPSC_S_NQ, TPSC_S_AL_N, TPSC_S_AL_MMIN,
    TPSC_S_STAGES, /* bitmask of the completed stages, TPSC_STAGE_* (no reference counterpart:
                      the reference keeps this state in its members) */
    TPSC_S_N_THETA, /* the angle bins N_theta_k of the last tpsc_chijj call */
    TPSC_S_CHI0_FP64, /* 1 if the chi0 stage of the full grid runs in FP64: always in the FP64 path;
                        in the mixed path once the point's spin enhancement 1/(1 - Usp chi0max/2)
                        exceeded 8 (no reference counterpart) */
    TPSC_S_ROOTS_ON_DEVICE /* bitmask of the root searches of the last solve (1: Usp, 2: Uch, 4: mu)
                              that settled in the device Halley iteration rather than in the host
                              bracketed fallback (no reference counterpart) */
};
enum tpsc_stage {
    TPSC_STAGE_CHI0 = 1, TPSC_STAGE_USP_UCH = 2, TPSC_STAGE_VERTEX = 4, TPSC_STAGE_SIGMA = 8,
    TPSC_STAGE_TRACE_SG = 16, TPSC_STAGE_MU = 32, TPSC_STAGE_TRACE_SG2 = 64, TPSC_STAGE_CHIJJ = 128,
    TPSC_STAGE_AL = 256
};

/* constructor / destructor (cond_opt.cpp:10-62 and the base-class constructors). Builds
 * eps(k) and mu0 (green0.cpp:18-23). */
int tpsc_create(const tpsc_params* p, tpsc_ctx** out);
void tpsc_destroy(tpsc_ctx* ctx);
const char* tpsc_last_error(const tpsc_ctx* ctx); /* ctx NULL: the last failed context-free call
                                                      of the calling thread */

/* green0::find_mu, green0.cpp:240-262 (finite-lattice mu0, root of ddensity_mu_PBC) */
int tpsc_find_mu0(tpsc_ctx* ctx, double* mu0);

/* chi0::calc_chiqw_FFT_spline(int nq0, int nw0), chi0.cpp:279-718; also set_grid
 * (green0.h:44) to (nq0, nw0) */
int tpsc_chi0(tpsc_ctx* ctx, int nq0, int nw0);

/* chi::sum_rule_sp / sum_rule_ch at one candidate vertex, chi.cpp:833-843
 * (channel 0: spin, 1: charge; the charge rule uses the current Usp) */
int tpsc_sum_rule(tpsc_ctx* ctx, int channel, double Ux, double* value);

/* The value of tpsc_sum_rule without a context and without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Inputs are device arrays in
 * the layouts tpsc_get copies out: chi0_qn = chi0(q,iqn) (TPSC_F_CHI0, (Nw/2+1)*nk8 doubles),
 * chi0_inf and chi0_inf2 (nk8 doubles each), nk8 = (N/2+1)(N/2+2)/2; Usp is read by the charge
 * channel only.  The value of chi::sum_rule_sp (channel 0) or sum_rule_ch (channel 1) at Ux
 * (chi.cpp:833-843) goes to value[0], a device address, asynchronously on `stream` (a
 * hipStream_t; NULL = the default stream).  The weights and partial sums live in a fixed
 * 128 KiB array of the library on each GPU (held until the process ends), which limits N to
 * 256; larger N gives TPSC_E_INVALID.  Calls that overlap on one GPU share that array and race.
 * Errors: TPSC_E_INVALID, TPSC_E_DEVICE, with the message in tpsc_last_error(NULL) on the
 * calling thread. */
int tpsc_sum_rule_noalloc(const double* chi0_qn, const double* chi0_inf, const double* chi0_inf2, int N, int Nw,
                          double T, double n, double U, double Usp, int channel, double Ux, double* value,
                          void* stream);

/* chi::find_Usp_Uch(int nq0, int nw0, double Usc[2]), chi.cpp:846-876 */
int tpsc_find_Usp_Uch_grid(tpsc_ctx* ctx, int nq0, int nw0, double* Usp, double* Uch);

/* chi::find_Usp_Uch(), chi.cpp:210-417 (coarse pass when params.coarse_pass, then full) */
int tpsc_find_Usp_Uch(tpsc_ctx* ctx, double* Usp, double* Uch);

/* chi::calc_chirt, chi.cpp:37-206 (V(r,tau) for tau in [0, beta/2]) */
int tpsc_vertex_rt(tpsc_ctx* ctx);

/* green::calc_Self_FFT_spline, green.cpp:1673-2031, without its leading find_Usp_Uch()
 * and calc_chirt() calls (use tpsc_calc_self_energy_full for the whole method) */
int tpsc_self_energy(tpsc_ctx* ctx);
int tpsc_calc_self_energy_full(tpsc_ctx* ctx);

/* green::traceSelfG, green.cpp:3001-3104: Tr(Sigma G0) and its Sigma_inf-corrected value */
int tpsc_trace_self_G(tpsc_ctx* ctx, double* trace, double* trace_corr);

/* green::traceDiffG(mu), green.cpp:3108-3182 */
int tpsc_trace_diff_G(tpsc_ctx* ctx, double mu, double* value);

/* green::find_mu, green.cpp:3185-3282 (then traceSelfG2, green.cpp:2929-2997) */
int tpsc_find_mu(tpsc_ctx* ctx, double* mu);

/* cond_opt::calc_chijj_vertex_corr_all_optim(chi3, mmin, N_theta_k, k_chijj, nk_chijj),
 * cond_opt.cpp:2537-6467.  Bubble + Maki-Thompson + kx0; with chi3 != 0 also the
 * Aslamazov-Larkin term on the sparse boson set of parameter mmin >= 0 (TPSC_F_CHIJJ3,
 * TPSC_F_CHIJJ_ALL, TPSC_F_AL_FREQ; Nw <= 2048).  N_theta_k in 1..64: the contributions are also
 * summed per angle bin b = floor(atan((N/2 - ky)/(N/2 - kx)) / (pi/(2 N_theta_k))) of k
 * (TPSC_F_CHIJJ1_K, _CHIJJ2_K, _CHIJJ3_K; the reference driver uses 1, tpsc_main.cpp:243).
 * Runs the missing upstream stages first, as cond_opt.cpp:2550-2560 does. */
int tpsc_chijj(tpsc_ctx* ctx, int chi3, int mmin, int N_theta_k);

/* one full solve: constructor state -> calc_Self_FFT_spline -> traceSelfG -> find_mu ->
 * calc_chijj_vertex_corr_all_optim (tpsc_main.cpp:61-75 and :243) */
int tpsc_solve(tpsc_ctx* ctx);

/* green::Gwn_local (green.cpp:209-266): the local Green function
 * G_loc(iw_n) = (1/N^2) sum_k 1/(iw_n - e_k + mu - Sigma(k,iw_n)), n = 0..Nw/2-1, written to out
 * as Nw/2 complex values interleaved re/im.  Needs the self-energy and mu (TPSC_E_ORDER). */
int tpsc_gwn_local(tpsc_ctx* ctx, double* out);

/* chi::sp_corr_length (chi.cpp:451-640, run by green::calc_Self_FFT_spline, green.cpp:1680):
 * out = {vF, xi_y, xi_x, xi_as_y, xi_as_x}: the Fermi velocity at the hot spot of the
 * non-interacting Fermi surface (chi.cpp:419-447, green0.cpp:626-675), the inverse half width
 * at half maximum of chi_sp(q, 0) along y and x through qmax, and the asymptotic
 * xi0 sqrt(Usp/(2/chi0max - Usp)).  Needs Usp and chi0 on the full grid (TPSC_E_ORDER). */
int tpsc_sp_corr_length(tpsc_ctx* ctx, double* out);

/* mu0 as an input parameter: every later stage uses this mu0 instead of solving the density
 * equation (green0.cpp:240-262); NaN returns to the solved mu0.  Resets the computed stages.
 * No reference counterpart: it makes d/dmu0 a partial derivative at fixed n. */
int tpsc_fix_mu0(tpsc_ctx* ctx, double mu0);

/* Gradient of a loss L(t, t', mu0, U) built from the solve's outputs (FP64 path only), as a
 * forward / backward pair.  No reference counterpart.  L is the sum of the terms in `terms`
 * (bitwise or of TPSC_LOSS_*); w: nk8 weights over the packed wedge (the Sigma terms).
 * forward: runs the chain at the context's parameters (mu0 fixed if tpsc_fix_mu0 was called):
 *   calc_Self_FFT_spline (chi0 -> Usp, Uch -> V -> Sigma, green.cpp:1673-2031), and for the
 *   Lambda_jj term traceSelfG, find_mu and calc_chijj_vertex_corr_all_optim (tpsc_main.cpp:61-75,
 *   243); returns L.
 * backward: grad[4] = dloss * dL/d(t, t', mu0, U) by tangent-linear propagation through the
 *   same chain: one pass per parameter, the tau slabs recomputed instead of stored, the Usp / Uch
 *   roots differentiated implicitly through their sum rules (chi.cpp:833-843) and mu through
 *   Tr(G - G0)(mu) = 0 (green.cpp:3185-3282).  The Lambda_jj term needs Nw <= 1024.
 * The per-term derivatives, dmu/dp and d Lambda_jj(iq_n)/dp are in TPSC_F_GRAD_*. */
#define TPSC_LOSS_RE_SIGMA 1 /* sum_k w_k Re Sigma(k, i pi T)                                   */
#define TPSC_LOSS_IM_SIGMA 2 /* sum_k w_k Im Sigma(k, i pi T)                                   */
#define TPSC_LOSS_CHISP_Q 4  /* chi_sp(q = (pi, pi), iq_0) = chi0/(1 - Usp chi0/2)              */
#define TPSC_LOSS_LAMBDA 8   /* sum_{n=0}^{8} [Lambda_jj1 + Lambda_jj2](iq_n), bubble + MT      */
int tpsc_loss_forward(tpsc_ctx* ctx, const double* w, int terms, double* loss);
int tpsc_loss_backward(tpsc_ctx* ctx, double dloss, double* grad);
/* the same pair for L = sum_k w_k Re Sigma(k, i pi T) (terms = TPSC_LOSS_RE_SIGMA) */
int tpsc_sigma_loss_forward(tpsc_ctx* ctx, const double* w, double* loss);
int tpsc_sigma_loss_backward(tpsc_ctx* ctx, double dloss, double* grad);

/* n weights u_k ~ U(0,1) from std::mt19937_64(seed) and std::uniform_real_distribution
 * (the loss weights of the gradient tests); needs no context */
int tpsc_random_weights(unsigned long long seed, size_t n, double* out);

/* ---- batches of independent parameter points (no reference counterpart) ---- */
typedef struct tpsc_batch_result {
    int status;                                  /* TPSC_OK or the failing call's code      */
    int queued;                                  /* 1: solved on the queued path (workers =
                                                    1), 0: by tpsc_solve                    */
    double mu0, Usp, Uch, mu, dblOcc, kx0;
    double seconds;                              /* wall time of the point's solve (queued
                                                    path: the queue's time per point)       */
} tpsc_batch_result;

/* Solves npts independent points, each with the results of tpsc_create + tpsc_solve, largest
 * N^2 Nw first.  workers = 1: one host thread queues every point on one context per grid
 * without a host wait (root searches settled on the device in at most 8 passes each), drains
 * the streams once and solves again with tpsc_solve the points the queue could not finish
 * (a search unsettled after 8 passes, a failed mu0 search, a mixed-precision point that needs
 * the FP64 chi0 stage).  workers >= 2: host threads, each driving its own contexts and
 * streams with tpsc_set_params + tpsc_solve, so that the points' kernels interleave on the
 * GPU.  workers = 0: chosen by tpsc_batch_tune.  The queued path keeps its contexts (one per
 * grid and precision) and its pinned result slots from call to call; tpsc_batch_release frees
 * them.  Calls on the queued path from several host threads run one after another.
 * chijj (optional, npts*ldc doubles): Lambda_jj1 + Lambda_jj2 at n = 0..min(Nw0/2, ldc-1). */
int tpsc_solve_batch(const tpsc_params* pts, int npts, int workers, tpsc_batch_result* out, double* chijj,
                     int ldc);
/* times the solve of the 8 largest points at 1, 2, 4 and 8 workers (points/s in rates[4]
 * if not NULL) and returns the fastest worker count */
int tpsc_batch_tune(const tpsc_params* pts, int npts, int* workers, double* rates);
/* frees the contexts and result slots the queued batch path keeps between calls */
int tpsc_batch_release(void);
/* npts random points drawn from std::mt19937_64(seed) in the order N in {32,48,64,96,128},
 * T ~ U[0.05,0.5], n ~ U[0.8,1], U ~ U[1,6], t' ~ U[-0.3,0]; Nw the smallest power of two
 * with (Nw-1) pi T >= 20; FP64, no coarse pass. */
int tpsc_batch_points(unsigned long long seed, int npts, tpsc_params* out);

/* re-targets a context to another parameter point on the same grid and precision (n, T,
 * hoppings, U, coarse_pass may change): the constructors' eps(k) and mu0 are redone and every
 * computed stage is dropped; the device arrays are kept.  No reference counterpart. */
int tpsc_set_params(tpsc_ctx* ctx, const tpsc_params* p);

/* the parameters the context was created with (or last set) */
int tpsc_get_params(tpsc_ctx* ctx, tpsc_params* out);

/* Writes, into directory dir (NULL or "" = the current directory), the text files the
 * reference writes along the same call sequence, for the stages this context has completed,
 * with the reference's file names and stream formats:
 *   e_k_*                                    hamiltonien.cpp:72-82
 *   chi0_qmax_wn_*, chi0_wn0_q_*             chi0.cpp:720-737
 *   TPSC_infos_*, TPSCparams_*               chi.cpp:375-415, :476-560 (sp_corr_length),
 *                                            green.cpp:3088-3103 (traceSelfG),
 *                                            green.cpp:3259-3266 (find_mu), :2978-2994 (traceSelfG2)
 *   mu0_mu_D_*                               green.cpp:3268-3276
 *   kx0_*, kx0_bin_*, chijj1_qn_*, chijj2_qn_*, chijj_with_vc1_qn_*   cond_opt.cpp:4195-4279
 *   chijj3_qn_*, chijj_with_all_vc_qn_*      cond_opt.cpp:6447-6481
 *   Gwn_loc_*                                green.cpp:209-266 (after find_mu)
 *   FS_*, FS_discret_AkF_*, FS_discret_*     green0.cpp:51-105, 626-676 (constructor)
 *   FS_inter_AkF_*, FS_inter_*, FS_inter_AkF_Pade_*, FS_inter_Pade_*   green.cpp:2210-2341 (after find_mu)
 *   Self_Pade_k_*, Akw_Pade_k_*              green.cpp:942-1065 (MDC of tpsc_main.cpp:101)
 *   chijj1_k_qn_*, chijj1_k_tot_qn_*, chijj2_k_qn_*, chijj2_k_tot_qn_*   cond_opt.cpp:3889-3946
 *   chijj3_k_qn_*, chijj3_k_tot_qn_*         cond_opt.cpp:6483-6520 (N_theta_k > 1)
 * The files the reference builds in several appends are written whole.  *nfiles (optional)
 * receives the number of files written. */
int tpsc_write_files(tpsc_ctx* ctx, const char* dir, int* nfiles);

/* ---- post-processing of a solved point (host FP64 over the device results, as the reference's
 * main runs it after find_mu, tpsc_main.cpp:77-101) ---------------------------------------------
 * green0::find_mu_inf_syst (green0.cpp:221-236, 266-287): mu0 of the infinite lattice, the root of
 * (2/pi^2) int_[0,pi]^2 f(eps(k) - mu) dk = n by nested adaptive Gauss-Kronrod quadrature; 0 at
 * half filling without t', t''.  Needs no computed stage. */
int tpsc_mu0_infinite(tpsc_ctx* ctx, double* mu0);
/* green0::find_FS(nk) (green0.cpp:626-676): for kx0 = j pi/(nk-1), the root kx in [kx0, pi] of
 * eps(kx, kx - kx0) = mu0 (finite-lattice mu0); kxy[2j], kxy[2j+1] = (kx, kx - kx0), -1 from the
 * first j without a root on; *count (optional) = the number found. */
int tpsc_fermi_surface(tpsc_ctx* ctx, int nk, double* kxy, int* count);
/* the lattice Fermi surface along the diagonals x = l + m, y = m: per l < N/2+1, the m of largest
 * A(k, omega = 0) and that A.  kind 0: non-interacting, eta = 1e-5 (green0::find_FS_PBC,
 * green0.cpp:51-105); 1: interacting, Sigma(k, i eta) by the 3-point polynomial of Sigma(k, iw_n),
 * n < 3; 2: the 4-point Pade; eta = 1e-6 (green::find_FS_inter, green.cpp:2210-2341; needs mu). */
int tpsc_fs_lattice(tpsc_ctx* ctx, int kind, int* m_idx, double* Ak);
/* green::MDC(n, Nn, w, eta, Neta, use_Pade) (green.cpp:814-1068): Sigma(k, w + i eta_e) and
 * A(k, w + i eta_e) at every wedge point k, continued from Sigma(k, i w_{n_j}), j < Nn (the stored
 * frequencies j = 0..Nn-1 at the points n[j], as the reference indexes them) by the Nn-point Pade
 * (use_pade) or the 4- or 3-point polynomial.  self_w: nk8*Neta complex [k*Neta + e] interleaved
 * re/im; akw: nk8*Neta [k*Neta + e]; either may be NULL.  Needs Sigma and mu. */
int tpsc_mdc(tpsc_ctx* ctx, const int* n, int Nn, double w, const double* eta, int Neta, int use_pade,
             double* self_w, double* akw);
/* green::EDC(kx, ky, Nk, w, nw0, m_freq, NP_max, eta, N_eta) with calc_spectral_weight
 * (green.cpp:589-700, 1073-1137): A(k_p, w_i + i eta_e) from the Pade through the first
 * NP = min(N_freq, NP_max) points of the non-uniform Matsubara set of parameter m_freq
 * (create_non_uniform_ind_freq, green.cpp:556-583).  spectr: Nk*nw0*Neta [(p*nw0 + i)*Neta + e].
 * k indices fold into [0, N/2] as the reference's. */
int tpsc_edc(tpsc_ctx* ctx, const int* kx, const int* ky, int Nk, const double* w, int nw0, int m_freq, int NP_max,
             const double* eta, int Neta, double* spectr);
/* green::calc_Self_Re_w(k, n, Nn, w, Nw, eta, Neta) (green.cpp:408-551): at k = (k[0], k[1]),
 * 0 <= k[i] <= N/2, Sigma(w_i + i eta_e) and A from the Nn-point Pade (Nn >= 4), and from the
 * continued-fraction form Sigma = Sinf/(z + b1 + S(z)), S the Pade of the transformed remainder on
 * max(Nn/8, 4) points ("Jfrac").  self_w, self_jfrac: Nw*Neta complex [i*Neta + e]; akw, akw_jfrac:
 * Nw*Neta; any may be NULL. */
int tpsc_self_real_freq(tpsc_ctx* ctx, const int* k, const int* n, int Nn, const double* w, int Nw, const double* eta,
                        int Neta, double* self_w, double* akw, double* self_jfrac, double* akw_jfrac);
/* green::save_Self_bin (green.cpp:1496-1514), FP64 path: self_bin_* in dir (NULL = cwd), the
 * (Nw/2+1)*nk8 complex values of the reference's Self_kw_array after calc_Self_FFT_spline: the
 * remainder R(k, n), n < Nw/2, then (Sigma(k, tau_{Nw/2}), Sigma(k, beta)). */
int tpsc_save_self_bin(tpsc_ctx* ctx, const char* dir);
/* chi0::save_chi (chi0.cpp:61-96): chi0_* in dir, int nbq = N/2+1, int nwmax = Nw/2, then
 * chi0(q, iq_n), n = 0..nwmax, nk8 doubles each. */
int tpsc_save_chi0_bin(tpsc_ctx* ctx, const char* dir);
/* green::load_Self_bin (green.cpp:1516-1540): the (Nw/2+1)*nk8 complex values of a self_bin_*
 * file of an N x N, Nw grid into out (interleaved).  Needs no context. */
int tpsc_read_self_bin(const char* file, int N, int Nw, double* out);
/* chi0::loadChi0 (chi0.cpp:99-156): the header of a chi0_* file and, when out is not NULL and cap
 * (doubles) suffices, its (nwmax+1) * nbq(nbq+1)/2 values.  Needs no context. */
int tpsc_read_chi0_bin(const char* file, int* nbq, int* nwmax, double* out, size_t cap);

/* copy-out in the reference layouts; bytes must be >= the field size */
int tpsc_get(tpsc_ctx* ctx, int field, void* dst, size_t bytes);
size_t tpsc_field_size(tpsc_ctx* ctx, int field);
int tpsc_get_scalar(tpsc_ctx* ctx, int which, double* value);

/* Per-kernel timing of the last calls when enabled: names are the kernel families
 * ("chi0_slab", "sigma_pencil", ...).  tpsc_prof_get fills up to cap entries. */
int tpsc_prof_enable(tpsc_ctx* ctx, int on);
int tpsc_prof_reset(tpsc_ctx* ctx);
int tpsc_prof_count(tpsc_ctx* ctx);
int tpsc_prof_get(tpsc_ctx* ctx, int i, const char** name, double* total_ms, int* calls);

/* keep a device copy of chi0(q,tau) at the next tpsc_chi0 call (test hook; allocates
 * nwb*nk8 doubles once) */
int tpsc_keep_chi_tau(tpsc_ctx* ctx, int on);

/* device properties useful to callers (HBM bytes of the context's buffers) */
int tpsc_device_bytes(tpsc_ctx* ctx, double* bytes);

/* generique::find_zero (generique.cpp:604-913; with limits, :917-1234): a root of f inside
 * init = {start, lo, hi}; lims NULL or {min, max}, which a widened bracket may not leave.
 * The |f| tolerance is 1e-16 without limits and 1e-15 with them, as in the reference; the
 * method is a bracketed hybrid of inverse quadratic interpolation and bisection (the
 * library's host fallback of the sum-rule and mu searches and its Fermi-surface traces).
 * Host only; needs no device.  Returns 0 with *root set, or TPSC_E_NOROOT with *root the best
 * point seen; *nevals (optional) receives the number of evaluations of f. */
int tpsc_find_zero(double (*f)(double x, void* user), void* user, const double init[3], const double* lims,
                   double* root, int* nevals);

/* ---- analytic continuation (host, FP64; needs no device) ---------------------------------
 * generique::pade_cont_frac_coef_rec (generique.h:112, generique.cpp:1650-1663): coefficients
 * of the Thiele continued-fraction Pade approximant through (z[i], f[i]), i < n.  z, f, coef:
 * n complex values, interleaved re/im. */
int tpsc_pade_coef(const double* z, const double* f, int n, double* coef);
/* generique::pade (generique.cpp:1630-1648): the approximant at x[k], k < nx, into out
 * (complex, interleaved). */
int tpsc_pade_eval(const double* z, const double* coef, int n, const double* x, int nx, double* out);
/* Optical conductivity sigma(w_k) = Im Lambda(w_k + i eta) / w_k, Lambda continued by the
 * Pade approximant through Lambda(iq_n) = lam[n], q_n = 2 pi n T, n < npade (the bosonic
 * continuation of chi.cpp:1055-1093 applied to Lambda_jj; the reference itself stops at
 * Lambda_jj(iq_n)).  w_k = 0 is evaluated at w = 1e-6.  lam: Lambda_jj1 + Lambda_jj2 from
 * tpsc_get. */
int tpsc_pade_conductivity(const double* lam, int nlam, double T, int npade, const double* omega, int nw,
                           double eta, double* sigma);

#ifdef __cplusplus
}
#endif
#endif
