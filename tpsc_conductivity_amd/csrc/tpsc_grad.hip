// tpsc_grad.hip -- tangent-linear (forward-mode) kernels of the chain
//   eps(k), mu0 -> chi0 -> (Usp, Uch) -> V -> Sigma(k,iwn)
// for the gradient of L = sum_k w_k Re Sigma(k, i w_0) with respect to (t, t', mu0, U).
//
// One tangent pass per parameter: every slab kernel here recomputes the primal G0 slabs it
// needs (tau-slab recomputation: nothing of the forward's slab intermediates is stored) and
// applies the product rule to the reference's pointwise formulas; the transforms are linear
// and are the same DCT-I / spline-FFT kernels as the forward.  The spline + DFT pencils
// (chi0 tau -> iqn, V tau transform, Sigma tau -> iwn) are linear in their inputs and run
// unchanged on the tangent arrays (launchers with a DevArrays view of those arrays).
// dx = d(eps_k - mu0) is the tangent of the dispersion argument; FP64 only.
#include "tpsc_launch.hpp"

namespace tpsc {

#define GRAD_THREADS 1024

// Fermi function f(x) = 1/(e^{x/T} + 1) and 1 - f, without overflow
__device__ __forceinline__ void fermi2(double x, double T, double& f, double& g)
{
    if (x >= 0.0) {
        const double e = exp(-x / T);
        f = e / (1.0 + e);
        g = 1.0 / (1.0 + e);
    } else {
        const double e = exp(x / T);
        f = 1.0 / (1.0 + e);
        g = e / (1.0 + e);
    }
}

// tau = 0 data (chi0.cpp:323-447) and their tangents; one block.  Primal and tangent of
//   dG(r,0+) = DCT[x (1-f)/N^2]                 tangent DCT[((1-f) + x f(1-f)/T) dx/N^2]
//   dG(r,0-) = (r = 0 ? -mu0 : eps_r) - dG(r,0+)
//   G(r,0)   = DCT[-(1-f)/N^2]                  tangent DCT[-f(1-f)/T dx/N^2]
//   chi0(r,0) = -2G(G+1) (r = 0), -2G^2;  d/dtau chi0(r,0) from dG(r,0+-) and G (chi0.cpp:394-447)
// Tangent outputs: tchi[0..nk8) = dchi0(q,0), tchirt[0..nk8) = dchi0(r,0), tchirt0,
// tdchirt0, tdGrt0p, tdchiqt0.  scr: 5*nk8 doubles.
template <int N>
__global__ void __launch_bounds__(GRAD_THREADS)
    k_chi0_tau0_tan(const double* __restrict__ ekw, const double* __restrict__ dekw, const double* __restrict__ er,
                    const double* __restrict__ der, double mu0, double dmu0, double T, double* tchi, double* tchirt,
                    double* tchirt0, double* tdchirt0, double* tdGrt0p, double* tdchiqt0, double* scr)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int tid = threadIdx.x, nt = blockDim.x;
    double *PdGp = scr, *TdGp = scr + nk8, *Pg = scr + 2 * nk8, *Tg = scr + 3 * nk8, *stg = scr + 4 * nk8;
    for (int pass = 0; pass < 4; pass++) {
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            const double x = ekw[i] - mu0, dx = dekw[i] - dmu0;
            double f, g;
            fermi2(x, T, f, g);
            double v;
            if (pass == 0) v = x * g / nk02;
            else if (pass == 1) v = (g + x * f * g / T) * dx / nk02;
            else if (pass == 2) v = -g / nk02;
            else v = -f * g / T * dx / nk02;
            put_sym(slab, nk, l, m, v);
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        double* o = (pass == 0) ? PdGp : (pass == 1) ? TdGp : (pass == 2) ? Pg : Tg;
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            o[i] = slab[m + l * nk];
        }
        __syncthreads();
    }
    // dchi0(r,0) and the tangent of d/dtau chi0(r,0)
    for (long i = tid; i < nk8; i += nt) {
        const double g = Pg[i], dg = Tg[i];
        const double c = (i == 0) ? -2.0 * (2.0 * g + 1.0) * dg : -4.0 * g * dg;
        tchirt[i] = c;
        tchirt0[i] = c;
        tdGrt0p[i] = TdGp[i];
        const double Pm = ((i == 0) ? -mu0 : er[i]) - PdGp[i];
        const double Tm = ((i == 0) ? -dmu0 : der[i]) - TdGp[i];
        const double v = (i == 0) ? -2.0 * (TdGp[i] * (1.0 + g) + PdGp[i] * dg + Tm * g + Pm * dg)
                                  : -2.0 * (TdGp[i] * g + PdGp[i] * dg + Tm * g + Pm * dg);
        tdchirt0[i] = v;
        stg[i] = v;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; pass++) {
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym(slab, nk, l, m, (pass == 0) ? tchirt[i] : stg[i]);
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        double* o = (pass == 0) ? tchi : tdchiqt0;
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            o[i] = slab[m + l * nk];
        }
        __syncthreads();
    }
}

// G0(k,+tau)/N^2 (chi0.cpp:468-471, form 1 at tau = beta/2 :535-538) and G0(k,-tau)/N^2
// (:486-489) with their x-derivatives g (f/T - t) and g (t - (1-f)/T)
__device__ __forceinline__ double g0p_val(double x, double t, double T, double nk02, int form)
{
    if (x >= 0.0) return exp(-t * x) * (-1.0 / (exp(-x / T) + 1.0)) / nk02;
    if (form == 0) return -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02;
    return -exp((1.0 / T - t) * x) / ((exp(x / T) + 1.0) * nk02);
}
__device__ __forceinline__ double g0m_val(double x, double t, double T, double nk02)
{
    if (x >= 0.0) return -exp((t - 1.0 / T) * x) * (-1.0 / (exp(-x / T) + 1.0)) / nk02;
    return exp(t * x) / ((exp(x / T) + 1.0) * nk02);
}

// slices j = 1 .. nwb-1 of chi0(r,tau) and chi0(q,tau) (chi0.cpp:454-559), tangents:
//   dchi0(r,tau) = -2 (dG(r,-tau) G(r,tau) + G(r,-tau) dG(r,tau)), tau = beta/2: 4 G dG
// The block stages G(r,tau) in tchirt[j] and dG(r,tau) in tchi[j] before their final values.
template <int N>
__global__ void __launch_bounds__(GRAD_THREADS)
    k_chi0_slab_tan(const double* __restrict__ ekw, const double* __restrict__ dekw, double mu0, double dmu0,
                    double T, int Nw, double* tchi, double* tchirt)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int nwb = Nw / 2 + 1;
    const int j = blockIdx.x + 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* cr = tchirt + (long)j * nk8;
    double* cq = tchi + (long)j * nk8;
    const bool mid = (j == nwb - 1);
    const double t = mid ? 1.0 / (2.0 * T) : j / (Nw * T);
    const int npass = mid ? 2 : 4;
    for (int pass = 0; pass < npass; pass++) {
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            const double x = ekw[i] - mu0, dx = dekw[i] - dmu0;
            double f, g;
            fermi2(x, T, f, g);
            double v;
            if (pass < 2) {
                const double p = g0p_val(x, t, T, nk02, mid ? 1 : 0);
                v = (pass == 0) ? p : p * (f / T - t) * dx;
            } else {
                const double p = g0m_val(x, t, T, nk02);
                v = (pass == 2) ? p : p * (t - g / T) * dx;
            }
            put_sym(slab, nk, l, m, v);
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            const double v = slab[m + l * nk];
            if (pass == 0) cr[i] = v;                        // G(r,tau)
            else if (pass == 1) cq[i] = mid ? 4.0 * cr[i] * v : v;  // dG(r,tau), or dchi0(r,beta/2)
            else if (pass == 2) cq[i] = -2.0 * v * cq[i];    // -2 G(r,-tau) dG(r,tau)
            else cr[i] = cq[i] - 2.0 * v * cr[i];            // - 2 dG(r,-tau) G(r,tau)
        }
        __syncthreads();
    }
    if (mid)
        for (long i = tid; i < nk8; i += nt) cr[i] = cq[i];
    __syncthreads();
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        put_sym(slab, nk, l, m, cr[i]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        cq[i] = slab[m + l * nk];
    }
}

// Implicit derivatives of the sum rules (chi.cpp:738-843): per (q, n) partials of
//   A_sp = sum w [r_sp^2 - 1] dchi0,  B_sp = sum w chi0^2 r_sp^2 / 2,   r_sp = 1/(1 - Usp chi0/2)
//   A_ch = sum w [r_ch^2 - 1] dchi0,  B_ch = -sum w chi0^2 r_ch^2 / 2,  r_ch = 1/(1 + Uch chi0/2)
// with the chi0 tail (chi0_inf/qn^2 + chi0_inf2/qn^4) and its tangent beyond Nw/2; w = the
// sum-rule weights x T/N^2.  part[k*nblk + b], k < 4.
__global__ void __launch_bounds__(256)
    k_sum_rule_tan(const double* __restrict__ chi, const double* __restrict__ inf1, const double* __restrict__ inf2,
                   const double* __restrict__ tchi, const double* __restrict__ tinf1,
                   const double* __restrict__ tinf2, const double* __restrict__ wq, int nk8, int nwmax, double T,
                   double Nq, double Usp, double Uch, double* part)
{
    __shared__ double sc[4 * 4];
    const int q = blockIdx.x * 256 + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (q < nk8) {
        const double w16 = 2.0 * wq[q];
        for (int j = blockIdx.y; j <= 2 * nwmax; j += gridDim.y) {
            const double w = ((j == 0) ? 0.5 * w16 : w16) * T / Nq;
            double c, dc;
            if (j <= nwmax) {
                c = chi[(long)j * nk8 + q];
                dc = tchi[(long)j * nk8 + q];
            } else {
                const double wn = 2 * j * kPi * T;
                const double i2 = 1.0 / (wn * wn);
                c = (inf1[q] + inf2[q] * i2) * i2;
                dc = (tinf1[q] + tinf2[q] * i2) * i2;
            }
            const double rs = 1.0 / (1.0 - Usp * c / 2.0), rc = 1.0 / (1.0 + Uch * c / 2.0);
            acc[0] += w * (rs * rs - 1.0) * dc;
            acc[1] += w * 0.5 * c * c * rs * rs;
            acc[2] += w * (rc * rc - 1.0) * dc;
            acc[3] -= w * 0.5 * c * c * rc * rc;
        }
    }
    block_sum_k<4>(acc, sc);
    const int nblk = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[k * nblk + b] = acc[k];
}

// tangent of the q-space vertex integrand (chi.cpp:85-170), slice j, then DCT-I -> r:
//   normFact U/8 [3 Usp (chi_sp - chi0) + Uch (chi_ch - chi0)]
//   d chi_sp = (dchi0 + dUsp chi0^2/2) r_sp^2,  d chi_ch = (dchi0 - dUch chi0^2/2) r_ch^2
template <int N>
__global__ void __launch_bounds__(GRAD_THREADS)
    k_vertex_slab_tan(const double* __restrict__ chi, const double* __restrict__ tchi, double* tV, double normFact,
                      double U, double dU, double Usp, double dUsp, double Uch, double dUch)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const long j = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double c = chi[i + j * nk8], dc = tchi[i + j * nk8];
        const double rs = 1.0 / (1.0 - Usp * c / 2.0), rc = 1.0 / (1.0 + Uch * c / 2.0);
        const double csp = c * rs, cch = c * rc;
        const double dcsp = (dc + dUsp * c * c / 2.0) * rs * rs, dcch = (dc - dUch * c * c / 2.0) * rc * rc;
        const double P = 3.0 * Usp * (csp - c) + Uch * (cch - c);
        const double dP = 3.0 * dUsp * (csp - c) + 3.0 * Usp * (dcsp - dc) + dUch * (cch - c) + Uch * (dcch - dc);
        put_sym(slab, nk, l, m, normFact * (dU * P + U * dP) / 8.0);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        tV[i + j * nk8] = slab[m + l * nk];
    }
}

// y[i] += a x[i]
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* y)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// tangent of the Sigma boundary data (green.cpp:1743-1817; forward k_sigma_bnd): block side
// 0 (tau = 0+) / 1 (tau = beta-), with g = G(r,0+) = DCT[-(1-f)/N^2] or G(r,beta-) = DCT[-f/N^2]
//   dSigma/dtau integrand: side 0  fact dchi0'(r,0) g + V0 dG(r,0+)
//                          side 1 -fact dchi0'(r,0) g + V0 (-dG(r,0+) + xi_r)
//   Sigma(r, 0 / beta) integrand: g V0
// Outputs tdSig[side*nk8 + k], tSb[side*nk8 + k].  scr: 6*nk8 doubles.
template <int N>
__global__ void __launch_bounds__(GRAD_THREADS)
    k_sigma_bnd_tan(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ er,
                    const double* __restrict__ der, const double* __restrict__ V0, const double* __restrict__ tV0,
                    const double* __restrict__ dchirt0, const double* __restrict__ tdchirt0,
                    const double* __restrict__ dGrt0p, const double* __restrict__ tdGrt0p, double mu0, double dmu0,
                    double T, double fact, double dfact, double* tdSig, double* tSb, double* scr)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int side = blockIdx.x;
    double* Pg = scr + 3 * side * nk8;
    double* stg = Pg + nk8;  // 2*nk8
    for (int pass = 0; pass < 2; pass++) {
        for (int s = tid; s < nk * nk; s += nt) {
            const double x = ek[s] - mu0, dx = dek[s] - dmu0;
            double f, g;
            fermi2(x, T, f, g);
            double v;
            if (pass == 0) v = (side == 0) ? -g / nk02 : -f / nk02;
            else v = (side == 0) ? -f * g / T * dx / nk02 : f * g / T * dx / nk02;
            slab[s] = v;
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        if (pass == 0) {
            for (long i = tid; i < nk8; i += nt) {
                int l, m;
                wedge_lm(i, l, m);
                Pg[i] = slab[m + l * nk];
            }
        } else {
            for (long i = tid; i < nk8; i += nt) {
                int l, m;
                wedge_lm(i, l, m);
                const double g = Pg[i], dg = slab[m + l * nk];
                const double v = V0[i], dv = tV0[i];
                const double c = dchirt0[i], dc = tdchirt0[i];
                const double dfc = dfact * c * g + fact * dc * g + fact * c * dg;
                if (side == 0) {
                    stg[i] = dfc + dv * dGrt0p[i] + v * tdGrt0p[i];
                } else {
                    const double xi = (i == 0) ? -mu0 : er[i], dxi = (i == 0) ? -dmu0 : der[i];
                    stg[i] = -dfc + dv * (-dGrt0p[i] + xi) + v * (-tdGrt0p[i] + dxi);
                }
                stg[nk8 + i] = dg * v + g * dv;
            }
        }
        __syncthreads();
    }
    for (int pass = 0; pass < 2; pass++) {
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym(slab, nk, l, m, stg[pass * nk8 + i]);
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        double* o = (pass == 0) ? tdSig : tSb;
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            o[i + side * nk8] = slab[m + l * nk];
        }
        __syncthreads();
    }
}

// tangent of Sigma(k,tau_j) = DCT[G0(r,tau_j) V(r,|tau_j|)] (green.cpp:1830-1892; forward
// k_sigma_slab): DCT[dG0 V + G0 dV], G0 and dG0 recomputed; G0(r,tau) staged in tS[j].
template <int N>
__global__ void __launch_bounds__(GRAD_THREADS)
    k_sigma_slab_tan(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ V,
                     const double* __restrict__ tV, double mu0, double dmu0, double T, int Nw, double* tS)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int j = blockIdx.x;
    const int nwb = Nw / 2 + 1;
    const int jv = (j < nwb) ? j : Nw - j;
    const double t = j / (Nw * T);
    const int tid = threadIdx.x, nt = blockDim.x;
    double* o = tS + (long)j * nk8;
    const double* Vj = V + (long)jv * nk8;
    const double* tVj = tV + (long)jv * nk8;
    for (int pass = 0; pass < 2; pass++) {
        for (int s = tid; s < nk * nk; s += nt) {
            const double x = ek[s] - mu0, dx = dek[s] - dmu0;
            double f, g;
            fermi2(x, T, f, g);
            const double p = (x > 0) ? -exp(-t * x) / (exp(-x / T) + 1.0) / nk02
                                     : -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02;
            slab[s] = (pass == 0) ? p : p * (f / T - t) * dx;
        }
        __syncthreads();
        slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
        if (pass == 0)
            for (long i = tid; i < nk8; i += nt) {
                int l, m;
                wedge_lm(i, l, m);
                o[i] = slab[m + l * nk];
            }
        __syncthreads();
    }
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double v = slab[m + l * nk] * Vj[i] + o[i] * tVj[i];
        o[i] = v;
    }
    __syncthreads();
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        put_sym(slab, nk, l, m, o[i]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        o[i] = slab[m + l * nk];
    }
}

// Sigma(k, i w0) = Sinf/z + C2/z^2 + C3/z^3 + R/z^4 at z = i w0 (green.cpp:3039):
// which 0: sum_k w_k Re Sigma = sum_k w_k (-C2_k / w0^2 + Re R_k(0) / w0^4);
// which 1: sum_k w_k Im Sigma = sum_k w_k (-Sinf / w0 + C3_k / w0^3 + Im R_k(0) / w0^4).
// Self: FP64 remainder R, point-major [k*(Nw/2) + n].
__global__ void k_sigma_loss(const double* __restrict__ w, const double* __restrict__ Sinfp,
                             const double* __restrict__ Sinf2, const double* __restrict__ Sinf3,
                             const cd* __restrict__ Self, long nk8, int H, double w0, int which, double* out)
{
    __shared__ double sc[16];
    double acc = 0.0;
    const double i1 = 1.0 / w0, i2 = i1 * i1;
    const double Si = Sinfp[0];
    for (long k = threadIdx.x; k < nk8; k += blockDim.x)
        acc += (which == 0) ? w[k] * (-Sinf2[k] * i2 + Self[k * H].x * i2 * i2)
                            : w[k] * (-Si * i1 + Sinf3[k] * i2 * i1 + Self[k * H].y * i2 * i2);
    acc = block_sum(acc, sc);
    if (threadIdx.x == 0) out[0] = acc;
}

// ================================================================== launchers
static size_t grad_lds(int N)
{
    const int nk = N / 2 + 1;
    return (((size_t)nk * nk * sizeof(double) + 15) & ~(size_t)15) + 2 * sizeof(double) * (size_t)N;
}

#define GRAD_DISPATCH_N(N, ...)                                     \
    switch (N) {                                                    \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;       \
        case 8: { constexpr int NN = 8; __VA_ARGS__; } break;       \
        case 12: { constexpr int NN = 12; __VA_ARGS__; } break;     \
        case 24: { constexpr int NN = 24; __VA_ARGS__; } break;     \
        case 48: { constexpr int NN = 48; __VA_ARGS__; } break;     \
        case 96: { constexpr int NN = 96; __VA_ARGS__; } break;     \
        case 192: { constexpr int NN = 192; __VA_ARGS__; } break;   \
        case 16: { constexpr int NN = 16; __VA_ARGS__; } break;     \
        case 32: { constexpr int NN = 32; __VA_ARGS__; } break;     \
        case 64: { constexpr int NN = 64; __VA_ARGS__; } break;     \
        case 128: { constexpr int NN = 128; __VA_ARGS__; } break;   \
        case 256: { constexpr int NN = 256; __VA_ARGS__; } break;   \
        default: return hipErrorInvalidValue;                       \
    }

template <typename K>
static void grad_set_lds(K kern, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

hipError_t launch_chi0_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                           hipStream_t s)
{
    const size_t lds = grad_lds(g.N);
    GRAD_DISPATCH_N(g.N, {
        grad_set_lds(k_chi0_tau0_tan<NN>, lds);
        hipLaunchKernelGGL(k_chi0_tau0_tan<NN>, dim3(1), dim3(GRAD_THREADS), lds, s, d.ekw, t.dekw, d.er, t.der, mu0,
                           dmu0, g.T, t.chi, t.chirt, t.chirt0, t.dchirt0, t.dGrt0p, t.dchiqt0, t.scr);
        grad_set_lds(k_chi0_slab_tan<NN>, lds);
        hipLaunchKernelGGL(k_chi0_slab_tan<NN>, dim3(g.nwb - 1), dim3(GRAD_THREADS), lds, s, d.ekw, t.dekw, mu0, dmu0,
                           g.T, g.Nw, t.chi, t.chirt);
    });
    return hipGetLastError();
}

hipError_t launch_sum_rule_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double Usp, double Uch,
                               double* out4, hipStream_t s)
{
    const double Nq = 4.0 * (g.nk - 1) * (g.nk - 1);
    const dim3 gr((unsigned)((g.nk8 + 255) / 256), kRowGroups);
    hipLaunchKernelGGL(k_sum_rule_tan, gr, dim3(256), 0, s, d.chiw, d.chi0inf, d.chi0inf2, t.chi, t.chi0inf,
                       t.chi0inf2, d.wq, (int)g.nk8, g.Nw / 2, g.T, Nq, Usp, Uch, t.part);
    return launch_sum_rows(t.part, 4, (int)(gr.x * gr.y), out4, s);
}

hipError_t launch_vertex_slab_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double U, double dU,
                                  double Usp, double dUsp, double Uch, double dUch, hipStream_t s)
{
    const double normFact = g.T / ((double)g.N * g.N);
    const size_t lds = grad_lds(g.N);
    GRAD_DISPATCH_N(g.N, {
        grad_set_lds(k_vertex_slab_tan<NN>, lds);
        hipLaunchKernelGGL(k_vertex_slab_tan<NN>, dim3(g.nwb), dim3(GRAD_THREADS), lds, s, d.chiw, t.chi, t.V,
                           normFact, U, dU, Usp, dUsp, Uch, dUch);
    });
    return hipGetLastError();
}

hipError_t launch_axpy(long n, double a, const double* x, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, y);
    return hipGetLastError();
}

hipError_t launch_sigma_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                            double fact, double dfact, hipStream_t s)
{
    const size_t lds = grad_lds(g.N);
    GRAD_DISPATCH_N(g.N, {
        grad_set_lds(k_sigma_bnd_tan<NN>, lds);
        hipLaunchKernelGGL(k_sigma_bnd_tan<NN>, dim3(2), dim3(GRAD_THREADS), lds, s, d.ek, t.dek, d.er, t.der, d.V0,
                           t.V0, d.dchirt0, t.dchirt0, d.dGrt0p, t.dGrt0p, mu0, dmu0, g.T, fact, dfact, t.dSig, t.Sb,
                           t.scr);
        grad_set_lds(k_sigma_slab_tan<NN>, lds);
        hipLaunchKernelGGL(k_sigma_slab_tan<NN>, dim3(g.Nw + 1), dim3(GRAD_THREADS), lds, s, d.ek, t.dek, d.V, t.V,
                           mu0, dmu0, g.T, g.Nw, t.Stau);
    });
    return hipGetLastError();
}

hipError_t launch_sigma_loss(const Grid& g, const double* w, const double* Sinf, const double* Sinf2,
                             const double* Sinf3, const cd* Self, int which, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sigma_loss, dim3(1), dim3(1024), 0, s, w, Sinf, Sinf2, Sinf3, Self, g.nk8, g.Nw / 2,
                       kPi * g.T, which, out);
    return hipGetLastError();
}

}  // namespace tpsc
