/*
 * tpsc_oracle.h -- FP64 CPU restatement of the TPSC hot path (test infrastructure).
 *
 * This is the ORACLE: the correctness reference for the HIP product library and the
 * CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as a checker.  The product path never links it.
 *
 * It restates, in plain C99 + OpenMP, the algorithm of dbergeron1/TPSC_conductivity for
 * the chain  eps(k) -> mu0 -> chi0(q,iqn) -> (Usp,Uch) -> V(r,tau) -> Sigma(k,iwn) -> mu
 * -> Lambda_jj(iqn) (bubble + Maki-Thompson + diamagnetic kx0), and the Aslamazov-Larkin
 * term Lambda_jj3 on its sparse boson set.  Every function cites the
 * reference file:line it follows.  The FFTW r2r/dft calls of the reference are replaced
 * by direct-definition transforms (REDFT00 / RODFT00 by cosine/sine matrices, complex DFT
 * by a radix-2 FFT), LAPACK dgbsv_ by a restated banded LU with partial pivoting.
 *
 * Layouts are those of the reference:
 *   quadrant grid   ek[m + l*nk]                 nk = N/2+1, kx = l*pi/(nk-1)
 *   packed wedge    idx(l,m) = m + l(l+1)/2       m <= l, nk8 = nk(nk+1)/2
 *   tau/freq-major  a[idx + j*nk8]
 *   Lambda grid     a[m + l*nk + j*nq]           kx index l+1 in 1..nk-2, nq = nk(nk-2)
 */
#ifndef TPSC_ORACLE_H
#define TPSC_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nk0;          /* N: lattice is N x N                                  */
    int Nw0;          /* total number of fermionic +-iwn (tau intervals)      */
    double n;         /* density                                              */
    double T;         /* temperature                                          */
    double eta;       /* broadening (kept for interface parity; unused here)  */
    double tp, tpp, t3, U;
    int coarse_pass;  /* 1: run the coarse (N/2, Nw/2) Usp/Uch pass first    */
    double t;         /* nearest-neighbour hopping; 0 means 1 (the reference's fixed
                         value, hamiltonien.cpp:26,121)                        */
} orc_params;

typedef struct orc {
    orc_params p;
    /* current grid (set_grid changes it, green0.h:44) */
    int nbk, nbw, nwmax;
    double *ek, *er;
    double mu0;
    /* chi0 stage */
    double *chiqw;          /* nwb*nk8: chi0(q,tau) then chi0(q,iqn)            */
    double *chiqt;          /* copy of chi0(q,tau) kept for tests (nwb*nk8)     */
    double *chirt;          /* nwb*nk8: chi0(r,tau)                             */
    double *dtau_chirt0, *dtau_Grt0p, *dtau_chiqt0;  /* nk8 each                */
    double *chi0_inf, *chi0_inf2;                   /* nk8 each                */
    double chi0max; int qmax[2];
    /* sum rules */
    double nsr, Usp, Uch, dblOcc, difRelsp, difRelch, Usp_coarse, Uch_coarse;
    int nevals_sp, nevals_ch;
    /* vertex */
    double *V;              /* nwb*nk8 */
    /* self-energy */
    double *Sigma_tau;      /* (Nw+1)*nk8: Sigma(k,tau) kept for tests          */
    double *dtau_Sigma;     /* 2*nk8 */
    double complex *Self;   /* nwb*nk8: R(k,n), n < Nw/2                        */
    double Sigma_inf, *Sigma_inf2, *Sigma_inf3, *Sigma_inf4;
    double trSG, trSG_corr, trSG2, dblOcc2;
    double mu; int mu_ok; int nevals_mu;
    /* Lambda_jj */
    double *chijj1, *chijj2;        /* nwb each (n = 0..Nw/2) */
    double kx0;
    double chijj1_inf2, chijj2_inf2, chijj2_inf4;
    int N_theta;
    double *chijj1_k, *chijj2_k;    /* (nwb)*N_theta angle-binned contributions */
    /* Aslamazov-Larkin term (orc_chijj_al) */
    int al_N, al_N_theta;
    long *al_freq;                  /* al_N sparse boson indices                   */
    double *chijj3, *chijj_all;     /* al_N each: Lambda_jj3, Lambda_jj1+2+3       */
    double *chijj3_k;               /* al_N*al_N_theta angle-binned Lambda_jj3     */
    double *al_hw, *al_hg2;         /* hG(q,iq_p) p < Nw/2 ([q + p*nq]), 1/qn^2 coefficient */
    int mu0_fixed;                  /* 1: mu0 is an input (orc_fix_mu0), not solved from n */
} orc;

/* ---- lifecycle / model (hamiltonien.cpp:14-34,49-68; green0.cpp:13-37) ---- */
orc *orc_create(const orc_params *p);
void orc_destroy(orc *o);
void orc_set_grid(orc *o, int nk0, int Nw0);
double orc_dispk(const orc *o, double kx, double ky);
double orc_grad_ek_x(const orc *o, double kx, double ky);
double orc_hess_ek_xx(const orc *o, double kx, double ky);
double orc_ddensity_mu_PBC(orc *o, double mu);
int orc_find_mu0(orc *o);

/* ---- stages ---- */
int orc_chi0(orc *o, int nq0, int nw0);                 /* chi0.cpp:279-718 */
double orc_sum_rule(orc *o, int channel, double Ux);    /* chi.cpp:738-843  */
int orc_find_Usp_Uch_grid(orc *o, int nq0, int nw0, double Usc[2]);  /* chi.cpp:846-876 */
int orc_find_Usp_Uch(orc *o);                           /* chi.cpp:210-417  */
int orc_vertex(orc *o);                                 /* chi.cpp:37-206   */
int orc_self_energy(orc *o);                            /* green.cpp:1673-2031 */
int orc_trace_self_G(orc *o);                           /* green.cpp:3001-3104 */
double orc_traceDiffG(orc *o, double mu);               /* green.cpp:3108-3182 */
int orc_find_mu(orc *o);                                /* green.cpp:3185-3282 */
int orc_trace_self_G2(orc *o);                          /* green.cpp:2929-2997 */
int orc_chijj(orc *o, int N_theta);                     /* cond_opt.cpp:2537-4290 */
int orc_solve(orc *o);                                  /* tpsc_main.cpp:61-75 + E4 */
int orc_solve_timed(orc *o, double t[6]);               /* the same without the coarse pass, stage times */
int orc_solve_from_roots(orc *o, double Usp, double Uch); /* the chain after the roots (Uch NaN: searched) */
int orc_omp_max_threads(void);
int orc_gwn_local(orc *o, double complex *out);          /* green.cpp:209-266 */
void orc_fix_mu0(orc *o, double mu0);                    /* mu0 as an input parameter   */
double orc_hot_spot_vF(orc *o);                          /* chi.cpp:419-447, green0.cpp:626-675 */
void orc_sp_corr_from(const double *c0, int nbq, int qx, int qy, double chi0max, double Usp,
                      double out[4]);                    /* chi.cpp:451-640 */
int orc_sp_corr_length(orc *o, double out[5]);
double orc_sigma_loss(orc *o, const double *w);          /* sum_k w_k Re Sigma(k,iw_0)  */
double orc_loss(orc *o, const double *w, int terms); /* the loss terms of tpsc_loss_forward */
int orc_al_freqs(int nbw, int mmin, long *out);         /* cond_opt.cpp:5123-5156 */
int orc_chijj_al(orc *o, int mmin, int N_theta);        /* cond_opt.cpp:4292-6467 (chi3) */
double complex orc_al_tail(double Usp, double Uch, double chi0inf, double T, double tau,
                           int kind);                   /* cond_opt.cpp:5556-5618 */
void orc_al_tail_out(double Usp, double Uch, double chi0inf, double T, double tau, int kind,
                     double out[2]);

/* ---- self-checks of the reference, as oracle functions ---- */
double orc_chiqw_discret(orc *o, int qx, int qy, int n);   /* chi0.cpp:1588-1651 */
double orc_trace_chi0(orc *o, double *tr_corr);           /* chi0.cpp:1498-1584 */
double complex orc_self_kw_brute(orc *o, int kx, int ky, int n); /* green.cpp:1401-1494 */

/* ---- numerics (generique.cpp) ---- */
typedef double (*orc_fn)(void *ctx, double x);
int orc_find_zero(orc_fn f, void *ctx, const double init[3], double *root,
                  const double *lims, int *nevals);        /* generique.cpp:604-1234 */
void orc_spline_coeffs_rel(const double *x0, const double *V, int N0,
                           double *coeffs);               /* generique.cpp:214-322 */
void orc_roots_2pol(const double complex coef[3], double complex roots[2]); /* :1666-1726 */
void orc_pade_coef(const double complex *func, const double complex *x, int N,
                   double complex *coef);                 /* generique.h:112, .cpp:1652 */
double complex orc_pade_eval(double complex x, int N, const double complex *x0,
                             const double complex *coef); /* generique.cpp:1630-1649 */
/* out-pointer forms for FFI callers that cannot pass C complex by value */
void orc_self_kw_brute_out(orc *o, int kx, int ky, int n, double out[2]);
void orc_pade_eval_out(const double z[2], int N, const double complex *x0,
                       const double complex *coef, double out[2]);
/* transforms with FFTW's REDFT00 / RODFT00 / dft definitions */
void orc_r2r_2d(double *a, int n0, int n1, int kind0, int kind1); /* kind 0: REDFT00, 1: RODFT00 */
void orc_redft00_1d(double *a, int n);
void orc_dft(double complex *a, int n, int sign);      /* sign +1: FFTW_BACKWARD */

#ifdef __cplusplus
}
#endif
#endif
