// tpsc_launch.hpp -- host-side launchers of the TPSC kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "tpsc_device.hpp"
#include "tpsc_roots.hpp"

namespace tpsc {

// Device-resident arrays of one parameter point, sized for the full grid.
// Layouts follow the reference: wedge idx(l,m) = m + l(l+1)/2, tau/frequency-major
// a[idx + j*nk8]; Lambda grid a[m + l*nk + j*nq] with kx index l+1.
// The tau arrays chi, chirt, V, Stau, h, f0 hold FP32 values in the mixed path (the
// pointer type is nominal there); Self holds FP32 complex.  Boundary arrays (chirt0, V0, Sb,
// hVb, f0Vb) are FP64 in both paths.
struct DevArrays {
    double* ek = nullptr;        // nk*nk quadrant
    double* ekw = nullptr;       // nk8  wedge-packed eps(k)
    double* wq = nullptr;        // nk8  wedge weight 8/2^(edges+diag)
    double* er = nullptr;        // nk8
    double* chi = nullptr;       // nwb*nk8 : chi0(q,tau)
    double* chiw = nullptr;      // nwb*nk8 : chi0(q,iqn), FP64 (== chi in the FP64 path)
    double* chirt = nullptr;     // nwb*nk8 : chi0(r,tau)
    double* dchirt0 = nullptr;   // nk8
    double* dGrt0p = nullptr;    // nk8
    double* dchiqt0 = nullptr;   // nk8
    double* chi0inf = nullptr;   // nk8
    double* chi0inf2 = nullptr;  // nk8
    double* srm = nullptr;       // (Nw+1)*kSRM suffix moments of the sum-rule terms (k_sr_moments)
    double* srmax = nullptr;     // Nw+1        suffix maxima of |chi0| per row
    double* V = nullptr;         // nwb*nk8
    double* Stau = nullptr;      // (Nw+1)*nk8
    double* dSig = nullptr;      // 2*nk8  dSigma/dtau at tau = 0, beta
    double* Sb = nullptr;        // 2*nk8  Sigma(k, tau = 0), Sigma(k, tau = beta)
    double* chirt0 = nullptr;    // nk8    chi0(r, tau = 0)
    double* V0 = nullptr;        // nk8    V(r, tau = 0)
    cd* Self = nullptr;          // nk8*(Nw/2), point-major: Sigma(k,iw_n) at [k*(Nw/2) + n]
    double* Sinf2 = nullptr;     // nk8
    double* Sinf3 = nullptr;
    double* Sinf4 = nullptr;
    double* Sinf = nullptr;      // 1
    double* dek = nullptr;       // nq   (v_x on the Lambda grid)
    double* d2ek = nullptr;      // nk*nk
    double* h = nullptr;         // (Nw+1)*nq
    double* f0 = nullptr;        // (Nw+1)*nq
    double* dhV = nullptr;       // 2*nq
    double* df0V = nullptr;      // 2*nq
    double* hVb = nullptr;       // 2*nq   hV(k, tau = 0), hV(k, tau = beta)
    double* f0Vb = nullptr;      // 2*nq   f0V(k, tau = 0), f0V(k, tau = beta)
    double* bscr = nullptr;      // 4*nq   staging of the Lambda boundary kernel
    double* hT = nullptr;        // (Nw+1)*nq  hV in point-major layout [q][j]
    double* f0T = nullptr;       // (Nw+1)*nq  f0V in point-major layout
    double* poles = nullptr;     // nq * 64    two-pole data per Lambda point (k_jj_poles)
    double* jjpart = nullptr;    // nblk * 2*nwb  per-workgroup Lambda partials
    double* jjscal = nullptr;    // nblk * 4      per-workgroup k0, inf sums
    double* kx0part = nullptr;   // 2*nk
    double* jjred = nullptr;     // 2*nwb + 4 reduced Lambda rows and scalars
    double* tmp = nullptr;       // scratch, >= 4*nk8
    double* part = nullptr;      // reduction partials
    double* red = nullptr;       // reduction result(s)
    HalleyState* root = nullptr;  // state of the device-driven root search
    cd* twN = nullptr;           // Nw   exp(-2 pi i t/Nw)
    cd* twX = nullptr;           // Nw   phase tables of the pencil transforms (k_fill_tables)
    cd* twS = nullptr;
    cd* twM = nullptr;
    cd* ph = nullptr;            // Nw+1
    int* imax = nullptr;         // argmax result
    double* mu0v = nullptr;      // 1  mu0 of the current grid (k_mu0_root), read by the chi0 and h/f0 kernels
    double* psc = nullptr;       // kPscDoubles scalars of a queued solve (layout below)
    unsigned* ticket = nullptr;  // 1  arrival count of a root-search pass (zero between passes)
};

// d.psc of a queued solve: the mirrors of the Usp, Uch and mu searches (8 doubles each: status,
// root, x, evaluations, f, f', lo, hi; the root read by later stages at [kPsc* + 1]), chi0max,
// and the mu0 search {found, mu0, evaluations}
constexpr int kPscUsp = 0, kPscUch = 8, kPscMu = 16, kPscChi0max = 24, kPscMu0 = 25, kPscDoubles = 32;

// Aslamazov-Larkin work arrays (FP64 in both precisions; allocated at the first chi3 call).
// A plane is (Nw+1)*nq doubles; complex tau arrays are two planes (re, im).
struct ALArrays {
    double* Grt = nullptr;      // (Nw+1)*nk8  G0(r,-tau_j) on the wedge
    double* dGrt = nullptr;     // 2*nk8       its tau-slopes at 0, beta
    double* dhG = nullptr;      // 2*nq        slopes of hG(q,tau) at 0, beta
    double* hw = nullptr;       // nq*(Nw/2)   hG(q,iq_p), point-major [q][p]
    double* hg2 = nullptr;      // nq          1/qn^2 coefficient of hG
    double* chiwP = nullptr;    // nk8*nwb     chi0(q,iq_p), point-major [w][p]
    double* P1p = nullptr;      // 2 planes    point-major [c][q][t]: HG(q,tau), then P1(k,tau)
    double* P2p = nullptr;      // 2 planes    point-major P2(k,tau)
    double* HGs = nullptr;      // 2 planes    slab-major HG(q,tau) [c][t][q]
    double* P1s = nullptr;      // 2 planes    slab-major P1 (and hG(q,tau) at set-up)
    double* P2s = nullptr;      // 2 planes    slab-major P2
    double* dHq = nullptr;      // 4*nq        HG slopes [c][side][q]
    double* dP = nullptr;       // 8*nq        P1/P2 slopes [which][side][c][k]
    double* contrib = nullptr;  // nwb*nq      per-frequency, per-k contributions [n][k]
    double* red = nullptr;      // nwb         Lambda_jj3 on the sparse set
    int N = 0, Nw = 0;          // grid the arrays were sized for
};

// Tangent arrays of the chain eps, mu0 -> chi0 -> Usp, Uch -> V -> Sigma (FP64; allocated at
// the first gradient call).  Layouts are those of the forward arrays of the same name.
struct TanArrays {
    double* dek = nullptr;     // nk*nk   d eps(k) on the quadrant
    double* dekw = nullptr;    // nk8     d eps(k) on the wedge
    double* der = nullptr;     // nk8     d eps(r)
    double* chi = nullptr;     // nwb*nk8 dchi0(q,tau), then dchi0(q,iqn) in place
    double* chirt = nullptr;   // nwb*nk8 dchi0(r,tau)
    double* chirt0 = nullptr;  // nk8
    double* dchirt0 = nullptr; // nk8
    double* dGrt0p = nullptr;  // nk8
    double* dchiqt0 = nullptr; // nk8
    double* chi0inf = nullptr; // nk8
    double* chi0inf2 = nullptr;// nk8
    double* V = nullptr;       // nwb*nk8
    double* V0 = nullptr;      // nk8
    double* dSig = nullptr;    // 2*nk8
    double* Sb = nullptr;      // 2*nk8
    double* Stau = nullptr;    // (Nw+1)*nk8
    cd* Self = nullptr;        // nk8*(Nw/2)
    double* Sinf2 = nullptr;   // nk8
    double* Sinf3 = nullptr;
    double* Sinf4 = nullptr;
    double* Sinf = nullptr;    // 1
    double* scr = nullptr;     // 6*nk8
    double* part = nullptr;    // 4 * sum-rule grid
    double* red = nullptr;     // 8
    double* w = nullptr;       // nk8 loss weights
    // the Lambda_jj stage (allocated at the first gradient with the Lambda_jj term)
    double* vel = nullptr;     // nq          d v_x(k) on the Lambda grid
    double* h = nullptr;       // (Nw+1)*nq   d hV(k,tau), tau-major
    double* f0 = nullptr;      // (Nw+1)*nq   d f0V(k,tau)
    double* hT = nullptr;      // (Nw+1)*nq   the same, point-major
    double* f0T = nullptr;
    double* dhV = nullptr;     // 2*nq        d of the tau-slopes at 0, beta
    double* df0V = nullptr;
    double* hVb = nullptr;     // 2*nq        d of the values at 0, beta
    double* f0Vb = nullptr;
    double* bscr = nullptr;    // 12*nq       staging of k_hf0_bnd_tan
    double* jjpart = nullptr;  // jj_nblk*2*nwb  per-workgroup tangent rows
    double* jjred = nullptr;   // 2*nwb + 2   their sums, then (dF/dp, dF/dmu) of the mu root
    double* jjtr = nullptr;    // 2*ceil(nk8/4) partials of k_trace_diffG_tan
    int jj_nblk = 0;
};

struct Grid {
    int N, nk, Nw, nwb;
    long nk8, nq;
    double T;
    int prec;  // 0: FP64; 1: mixed (FP32 slab arrays, FP32-complex Sigma(k,iwn))
};

constexpr int kRedBlocks = 1024;
constexpr int kRowGroups = 32;       // frequency-row groups of the reduction grids
constexpr int kPartCap = 8192;       // capacity of DevArrays::part
// Root-search passes over at most this many workgroups take their Halley step in the workgroup
// that finishes last (one launch per pass: small grids are launch-bound); larger passes launch
// the step kernel after them, since each workgroup's release fence (a write-back of its XCD's
// L2) costs more there than the launch it saves (256^2: sum-rule passes of 1056 workgroups,
// trace passes of 2097).
constexpr int kFusedStepMax = 640;
// series terms of the sum-rule passes' small-chi rows and the moments they take, M_2 .. M_{K+3}
constexpr int kSRK = 18;
constexpr int kSRM = kSRK + 2;
// ... used on grids of at least this many chi0(q,iqn) values (16 MB in FP64): on smaller ones a
// pass is launch-bound, and the two moment kernels would cost more than the rows they save
constexpr long kSRMinElems = 2L << 20;
constexpr int kJJK = 16;  // k-points per Lambda workgroup for Nw > 256
// Lambda pencil shape: up to Nw = 256 one wavefront per workgroup (its barriers are then free)
// and 4 points per workgroup, so that small grids still spread over the CUs; beyond, 256
// threads and kJJK points
constexpr int jj_points(int Nw) { return Nw <= 256 ? 4 : kJJK; }
constexpr int jj_threads(int Nw) { return Nw <= 256 ? 64 : 256; }

// ---- launchers; each returns hipGetLastError() of its launch
hipError_t launch_fill_tw(const DevArrays& d, int L, hipStream_t s);
hipError_t launch_chi0_tau0(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_chi0_slab(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_chi0_pencil(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_chi0max(const Grid& g, const DevArrays& d, hipStream_t s, double* out2 = nullptr);
hipError_t launch_sum_rule(const Grid& g, const DevArrays& d, double sU, double* out, hipStream_t s);
// the partial sums of the sum-rule trace at sU over explicit arrays (chi0(q,iqn) nwb*nk8,
// its two tail coefficients and the wedge weights, nk8 each): sum_rule_parts(N) doubles to part,
// their count to *nparts
int sum_rule_parts(int N);
hipError_t launch_sum_rule_part(int N, int Nw, double T, const double* chiw, const double* inf1, const double* inf2,
                                const double* wq, double sU, double* part, int* nparts, hipStream_t s);
// Device-driven Halley search (tpsc_roots.hpp) in d.root: initialise it, then one pass of
// the sum rule (channel sign sgn, Ux = the current iterate) or of traceDiffG (pmu = the current
// iterate) with its first two derivatives, followed by the step.  T and mu0 are read from the
// state, so a pass depends on the point only through d.root.  mirror (mapped host memory)
// holds {status, root, x, passes, f, f', lo, hi}: "running" from the initialisation, the
// outcome once the search has settled.
// mu0p (device, may be null): the state's mu0; a0src (device Usp, may be null): the charge sum
// rule's constant a0 = a0A - (nsr + Usp nsr^2/(2U) - nsr^2), formed on the device
hipError_t launch_halley_init(HalleyState* st, const HalleyState& s0, const double* chi0max, double* mirror,
                              hipStream_t s, const double* mu0p = nullptr, const double* a0src = nullptr,
                              double a0A = 0.0, double nsr = 0.0, double U = 0.0);
// psc[slot], psc[3 + slot], psc[6 + slot] = root, status, evaluations of the search in st
hipError_t launch_sum_rule3_pass(const Grid& g, const DevArrays& d, double sgn, double* mirror, hipStream_t s);
hipError_t launch_sr_moments(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_trace_diffG3_pass(const Grid& g, const DevArrays& d, double* mirror, hipStream_t s);
// green0::find_mu on the device (one workgroup, the whole bracketing search): out = {found, mu0, evaluations}
hipError_t launch_mu0_root(const Grid& g, const DevArrays& d, double n, double preset, double* out, hipStream_t s);
hipError_t launch_model(const Grid& g, const DevArrays& d, double t, double tp, double tpp, double t3,
                        hipStream_t s);
hipError_t launch_reduce_cols(const double* part, int nrows, int ncols, double* out, hipStream_t s);
hipError_t launch_reduce_cols2(const double* p1, int ncols1, double* o1, const double* p2, int ncols2, double* o2,
                               int nrows, hipStream_t s, const double* csrc = nullptr, int ncopy = 0,
                               double* cdst = nullptr);
// Optional device sources of the point scalars (the queued solve of tpsc_capi.cpp; null: the values
// passed): dsc = {Usp, Uch, mu} (device), from which the kernels form U(3Usp+Uch)/8 and /4 with the
// host's rounding; mu0p = the device mu0.
hipError_t launch_vertex_slab(const Grid& g, const DevArrays& d, double U, double Usp, double Uch, hipStream_t s,
                              const double* dsc = nullptr);
hipError_t launch_vertex_pencil(const Grid& g, const DevArrays& d, double U, double Usp, double Uch, hipStream_t s,
                                const double* dsc = nullptr);
hipError_t launch_sigma_bnd(const Grid& g, const DevArrays& d, double mu0, double fact, hipStream_t s,
                            const double* mu0p = nullptr, double U = 0.0, const double* dsc = nullptr);
hipError_t launch_sigma_slab(const Grid& g, const DevArrays& d, double mu0, hipStream_t s,
                             const double* mu0p = nullptr);
hipError_t launch_sigma_pencil(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_trace_diffG(const Grid& g, const DevArrays& d, double pmu, double mu0, double* out,
                              hipStream_t s);
hipError_t launch_trace_selfG(const Grid& g, const DevArrays& d, double mu0, double mu, int which, double* out,
                              hipStream_t s);
hipError_t launch_hf0_slab(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_hf0_bnd(const Grid& g, const DevArrays& d, double mu0, double fact4, hipStream_t s,
                          const double* mu0p = nullptr, double U = 0.0, const double* dsc = nullptr);
hipError_t launch_hVf0V_slab(const Grid& g, const DevArrays& d, hipStream_t s);
hipError_t launch_jj_poles(const Grid& g, const DevArrays& d, double mu, hipStream_t s, const double* dsc = nullptr);
hipError_t launch_jj_transpose(const Grid& g, const DevArrays& d, hipStream_t s);
// partial rows to part (nblk x ntheta x 2*nwb); bins: the angle bin of each Lambda point, or null
hipError_t launch_jj_pencil(const Grid& g, const DevArrays& d, double mu, double* part, const int* bins, int ntheta,
                            hipStream_t s, const double* dsc = nullptr);
// out[i] = (float)in[i], i < n
hipError_t launch_narrow(const double* in, float* out, long n, hipStream_t s);
// out[r*ntheta + b]: the sum of row r of part (nrows x n) over the columns k with bins[k] == b
hipError_t launch_sum_rows_binned(const double* part, int nrows, int n, const int* bins, int ntheta, double* out,
                                  hipStream_t s);
hipError_t launch_kx0_rows(const Grid& g, const DevArrays& d, double mu, hipStream_t s, const double* dsc = nullptr,
                           double* out = nullptr);
hipError_t launch_finalize(const double* part, int n, double* out, hipStream_t s);

constexpr int kGlocChunks = 64;  // k chunks of k_gloc (partials: kGlocChunks x Nw doubles)
hipError_t launch_gloc(const Grid& g, const DevArrays& d, double mu, double* part, double* out, hipStream_t s);
hipError_t launch_chi0_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                           hipStream_t s);
hipError_t launch_sum_rule_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double Usp, double Uch,
                               double* out4, hipStream_t s);
hipError_t launch_vertex_slab_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double U, double dU,
                                  double Usp, double dUsp, double Uch, double dUch, hipStream_t s);
hipError_t launch_axpy(long n, double a, const double* x, double* y, hipStream_t s);
hipError_t launch_sigma_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                            double fact, double dfact, hipStream_t s);
// which 0: sum_k w_k Re Sigma(k, i pi T); 1: sum_k w_k Im Sigma(k, i pi T)
hipError_t launch_sigma_loss(const Grid& g, const double* w, const double* Sinf, const double* Sinf2,
                             const double* Sinf3, const cd* Self, int which, double* out, hipStream_t s);
// tangent kernels of the Lambda_jj stage (tpsc_grad_jj.hip)
bool grad_jj_supported(const Grid& g);
hipError_t launch_trace_diffG_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                                  double mu, double* out2, hipStream_t s);
hipError_t launch_hVf0V_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                            hipStream_t s);
hipError_t launch_hf0_bnd_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                              double fact4, double dfact4, hipStream_t s);
hipError_t launch_jj_pencil_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu, double dmu,
                                int nblk, hipStream_t s);
hipError_t launch_transpose_d(const double* in, double* out, long ncols, int nrows, hipStream_t s);
hipError_t launch_sum_rows(const double* part, int nrows, int n, double* out, hipStream_t s);
hipError_t launch_al_setup(const Grid& g, const DevArrays& d, const ALArrays& a, double mu0, hipStream_t s);
hipError_t launch_al_freq(const Grid& g, const DevArrays& d, const ALArrays& a, double Usp, double Uch, double mu,
                          double U, long nn, double* contrib, hipStream_t s);

bool supported_N(int N);
bool supported_Nw(int Nw);

}  // namespace tpsc
