// tpsc_grad_jj.hip -- tangent-linear (forward-mode) kernels of the Lambda_jj stage
//   (eps, v_x, mu0, mu, V, Sigma) -> h, f0 -> hV, f0V -> Lambda_jj(iq_n) (bubble + Maki-Thompson)
// for gradients through calc_chijj_vertex_corr_all_optim (cond_opt.cpp:2537-3871); FP64 only.
//
// One tangent pass per parameter, on top of the tangents of eps(k), V and Sigma(k, iwn) that
// tpsc_grad.hip provides:
//  * the interacting mu moves with its root: Tr(G - G0)(mu) = 0 (green.cpp:3108-3282) gives
//    dmu = -F_p / F_mu (k_trace_diffG_tan);
//  * the h / f0 integrands (cond_opt.cpp:2697-2827) are recomputed slab by slab with their
//    tangents (dual numbers), transformed, multiplied by V and dV (k_hVf0V_tan), and the same
//    for their tau = 0, beta values and slopes (k_hf0_bnd_tan);
//  * the pencil (k_jj_pencil_tan) runs the reference's per-point steps on value and tangent:
//    splines and transforms are linear and run unchanged on the tangent samples; the Dyson G,
//    the two-pole asymptote (roots of z^2 + (mu - e) z - Sigma_inf, differentiated implicitly),
//    the products G(tau) J(-tau), G(tau) G(-tau) and the end slopes take the product rule.
// Straightforward code: this path serves gradient checks and fitting, not the timed solve.
#include "tpsc_launch.hpp"

namespace tpsc {

#define GRAD_JJ_THREADS 1024

// ------------------------------------------------------------------ dual numbers
struct dual {
    double v, d;  // value, tangent
};
__device__ __forceinline__ dual operator+(dual a, dual b) { return dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator-(dual a) { return dual{-a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator/(dual a, dual b)
{
    const double q = a.v / b.v;
    return dual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ dual operator+(dual a, double s) { return dual{a.v + s, a.d}; }
__device__ __forceinline__ dual operator+(double s, dual a) { return dual{s + a.v, a.d}; }
__device__ __forceinline__ dual operator-(dual a, double s) { return dual{a.v - s, a.d}; }
__device__ __forceinline__ dual operator-(double s, dual a) { return dual{s - a.v, -a.d}; }
__device__ __forceinline__ dual operator*(double s, dual a) { return dual{s * a.v, s * a.d}; }
__device__ __forceinline__ dual operator*(dual a, double s) { return dual{a.v * s, a.d * s}; }
__device__ __forceinline__ dual operator/(dual a, double s) { return dual{a.v / s, a.d / s}; }
__device__ __forceinline__ dual operator/(double s, dual b)
{
    const double q = s / b.v;
    return dual{q, -q * b.d / b.v};
}
__device__ __forceinline__ dual dexp(dual a)
{
    const double e = exp(a.v);
    return dual{e, e * a.d};
}

// complex value and complex tangent
struct cdual {
    cd v, d;
};
__device__ __forceinline__ cdual operator+(cdual a, cdual b) { return cdual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ cdual operator-(cdual a, cdual b) { return cdual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ cdual operator-(cdual a) { return cdual{-a.v, -a.d}; }
__device__ __forceinline__ cdual operator*(cdual a, cdual b) { return cdual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ cdual operator/(cdual a, cdual b)
{
    const cd q = a.v / b.v;
    return cdual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ cdual operator*(double s, cdual a) { return cdual{s * a.v, s * a.d}; }
__device__ __forceinline__ cdual operator*(cdual a, double s) { return cdual{a.v * s, a.d * s}; }
__device__ __forceinline__ cdual operator+(cdual a, double s) { return cdual{a.v + cd{s, 0.0}, a.d}; }
__device__ __forceinline__ cdual cexpdual(cdual z)
{
    const cd e = cexpd(z.v);
    return cdual{e, e * z.d};
}
__device__ __forceinline__ cdual cdl(dual re, dual im) { return cdual{cd{re.v, im.v}, cd{re.d, im.d}}; }
__device__ __forceinline__ dual rpart(cdual z) { return dual{z.v.x, z.d.x}; }
__device__ __forceinline__ dual ipart(cdual z) { return dual{z.v.y, z.d.y}; }

// ------------------------------------------------------------------ pointwise integrands
// h and f0 integrands at tau = t (cond_opt.cpp:2697-2735, 2795-2827), x = e - mu0, v = v_x(k)
__device__ __forceinline__ void hf0_dual(dual x, dual v, double T, double t, double nf, dual& hv, dual& fv)
{
    if (x.v > 0) {
        const dual E = dexp(-(x * t)), ex = dexp(-(x / T));
        const dual den = ex + 1.0;
        hv = nf * v * E / den;
        fv = nf * v * E * (t / den - ex / (T * den * den));
    } else {
        const dual E = dexp(x * (1.0 / T - t)), ex = dexp(x / T);
        const dual den = ex + 1.0;
        hv = nf * v * E / den;
        fv = nf * v * E * (t / den - 1.0 / (T * den * den));
    }
}

// tau-slopes of the h (which 0) and f0 (which 1) integrands at tau = 0+ (side 0) or beta-
// (side 1) (cond_opt.cpp:2745-2827)
__device__ __forceinline__ dual slope_dual(int which, int side, dual x, dual v, double T, double nf)
{
    if (which == 0) {
        if (x.v > 0) {
            const dual ex = dexp(-(x / T));
            return (side == 0) ? -(nf * v * x / (ex + 1.0)) : -(nf * v * x * ex / (ex + 1.0));
        }
        const dual ex = dexp(x / T);
        return (side == 0) ? -(nf * v * x * ex / (ex + 1.0)) : -(nf * v * x / (ex + 1.0));
    }
    if (x.v > 0) {
        const dual ex = dexp(-(x / T)), e1 = ex + 1.0;
        return (side == 0) ? nf * v * (x * ex / (T * e1 * e1) + 1.0 / e1)
                           : nf * v * ex * ((-x) * (1.0 / (T * e1) - ex / (T * e1 * e1)) + 1.0 / e1);
    }
    const dual ex = dexp(x / T), e1 = ex + 1.0;
    return (side == 0) ? nf * v * ex * (x / (T * e1 * e1) + 1.0 / e1)
                       : nf * v * ((-x) * (1.0 / (T * e1) - 1.0 / (T * e1 * e1)) + 1.0 / e1);
}

// ------------------------------------------------------------------ mu
// F(mu) = sum_k 2 w_k sum_{j=0}^{Nw} [Re G(k, iw_j) - Re G0(k, iw_j)] (green.cpp:3108-3182, the
// terms of k_trace_diffG3).  With a = e - mu + Re Sigma, b = w_j - Im Sigma, c = e - mu0:
// part[blk] = dF/dp at fixed mu = sum 2w [(a^2-b^2) r^2 da + 2ab r^2 db - (c^2-w^2) r0^2 dc],
// part[nblk + blk] = dF/dmu = sum 2w (b^2-a^2) r^2, r = 1/(a^2+b^2), r0 = 1/(c^2+w^2).
__global__ void __launch_bounds__(256)
    k_trace_diffG_tan(const double* __restrict__ ekw, const double* __restrict__ dekw, const double* __restrict__ wq,
                      const cd* __restrict__ Self, const cd* __restrict__ tSelf, const double* __restrict__ Sinf2,
                      const double* __restrict__ tSinf2, const double* __restrict__ Sinf3,
                      const double* __restrict__ tSinf3, const double* __restrict__ Sinf4,
                      const double* __restrict__ tSinf4, const double* __restrict__ Sinfp,
                      const double* __restrict__ tSinfp, int nk8, int Nw, double T, double mu0, double dmu0,
                      double mu, double* part)
{
    __shared__ double sc[2 * 4];
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int H = Nw / 2;
    double acc[2] = {0.0, 0.0};
    if (q < nk8) {
        const double Si = Sinfp[0], dSi = tSinfp[0];
        const double e = ekw[q], de = dekw[q], w = 2.0 * wq[q];
        const double s2 = Sinf2[q], s3 = Sinf3[q], s4 = Sinf4[q];
        const double d2 = tSinf2[q], d3 = tSinf3[q], d4 = tSinf4[q];
        const double c = e - mu0, dc = de - dmu0;
        for (int j = lane; j <= Nw; j += 64) {
            const double wn = (2 * j + 1) * kPi * T;
            double rS, iS, drS, diS;
            if (j < H) {
                const cd s = sigma_mom(Si, s2, s3, Self[(long)q * H + j], wn);
                const cd ds = sigma_mom(dSi, d2, d3, tSelf[(long)q * H + j], wn);
                rS = s.x, iS = s.y, drS = ds.x, diS = ds.y;
            } else {
                const double iw = 1.0 / wn, iw2 = iw * iw;
                rS = (-s2 + s4 * iw2) * iw2;
                iS = (-Si + s3 * iw2) * iw;
                drS = (-d2 + d4 * iw2) * iw2;
                diS = (-dSi + d3 * iw2) * iw;
            }
            const double a = e - mu + rS, b = wn - iS;
            const double r = 1.0 / (a * a + b * b), r0 = 1.0 / (c * c + wn * wn);
            const double da = de + drS, db = -diS;
            acc[0] += w * ((a * a - b * b) * r * r * da + 2.0 * a * b * r * r * db - (c * c - wn * wn) * r0 * r0 * dc);
            acc[1] += w * (b * b - a * a) * r * r;
        }
    }
    block_sum_k<2>(acc, sc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = acc[0];
        part[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ------------------------------------------------------------------ hV, f0V tau slabs
// Block (j, which): the tangent of hV(k, tau_j) = DST x DCT[2 h(r,tau_j) V(r,|tau_j|)] (which 0)
// or of f0V (which 1), cond_opt.cpp:2697-2735, 3029-3114.  The primal h (or f0) slab is
// recomputed and parked in the block's own output plane (each thread reads back what it
// wrote) while the tangent integrand is transformed; then 2 (dh V + h dV) is transformed.
template <int N>
__global__ void __launch_bounds__(GRAD_JJ_THREADS)
    k_hVf0V_tan(const double* __restrict__ ek, const double* __restrict__ tek, const double* __restrict__ vel,
                const double* __restrict__ tvel, const double* __restrict__ V, const double* __restrict__ tV,
                double mu0, double dmu0, double T, int Nw, double* th, double* tf0)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int j = blockIdx.x, which = blockIdx.y;
    const double t = (double)j / (T * Nw);
    const int jv = (j > Nw / 2) ? Nw - j : j;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* const o = ((which == 0) ? th : tf0) + (long)j * nq;
    for (int pass = 0; pass < 2; pass++) {
        for (long s = tid; s < nq; s += nt) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            const long i = m + (long)(l + 1) * nk;
            dual hv, fv;
            hf0_dual(dual{ek[i] - mu0, tek[i] - dmu0}, dual{vel[s], tvel[s]}, T, t, nf, hv, fv);
            const dual a = (which == 0) ? hv : fv;
            slab[s] = (pass == 0) ? a.v : a.d;
        }
        __syncthreads();
        slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
        if (pass == 0)
            for (long s = tid; s < nq; s += nt) o[s] = slab[s];
        __syncthreads();
    }
    const double* Vj = V + (long)jv * nk8;
    const double* tVj = tV + (long)jv * nk8;
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        const long r = fold_idx(l + 1, m, N);
        slab[s] = 2.0 * (slab[s] * Vj[r] + o[s] * tVj[r]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    for (long s = tid; s < nq; s += nt) o[s] = slab[s];
}

// ------------------------------------------------------------------ boundary data
// Tangents of the tau = 0, beta values and slopes of hV, f0V (k_hf0_bnd; cond_opt.cpp:
// 2697-2735, 2745-3114).  Block b: which = b/2 (0: h, 1: f0), side = b%2.  fact4 = U(3Usp+Uch)/4.
// scr: 3*nq per block (primal h(r,tau_b), its tangent, the primal slope transform).
template <int N>
__global__ void __launch_bounds__(GRAD_JJ_THREADS)
    k_hf0_bnd_tan(const double* __restrict__ ek, const double* __restrict__ tek, const double* __restrict__ vel,
                  const double* __restrict__ tvel, const double* __restrict__ V0, const double* __restrict__ tV0,
                  const double* __restrict__ dchirt0, const double* __restrict__ tdchirt0, double mu0, double dmu0,
                  double T, int Nw, double fact4, double dfact4, double* tdhV, double* tdf0V, double* thVb,
                  double* tf0Vb, double* scr)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int which = blockIdx.x >> 1, side = blockIdx.x & 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* const S0 = scr + (long)blockIdx.x * 3 * nq;  // h(r, tau_b)
    double* const S1 = S0 + nq;                           // its tangent
    double* const S2 = S1 + nq;                           // transformed slope integrand
    const double tb = (side == 0) ? 0.0 : (double)Nw / (T * Nw);
    // passes 0, 1: the integrand at tau_b and its tangent; 2, 3: the slope integrand and its tangent
    for (int pass = 0; pass < 4; pass++) {
        for (long s = tid; s < nq; s += nt) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            const long i = m + (long)(l + 1) * nk;
            const dual x{ek[i] - mu0, tek[i] - dmu0}, v{vel[s], tvel[s]};
            dual a;
            if (pass < 2) {
                dual hv, fv;
                hf0_dual(x, v, T, tb, nf, hv, fv);
                a = (which == 0) ? hv : fv;
            } else {
                a = slope_dual(which, side, x, v, T, nf);
            }
            slab[s] = (pass & 1) ? a.d : a.v;
        }
        __syncthreads();
        slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
        if (pass < 3) {
            double* o = (pass == 0) ? S0 : (pass == 1) ? S1 : S2;
            for (long s = tid; s < nq; s += nt) o[s] = slab[s];
            __syncthreads();
        }
    }
    // slope: DST x DCT[2 d V0 +- h fact4 dchi0(r,0)] (the sign of the tau = beta side is minus)
    const double sg = (side == 0) ? 1.0 : -1.0;
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        const long r = fold_idx(l + 1, m, N);
        const double v0 = V0[r], dv0 = tV0[r], c0 = dchirt0[r], dc0 = tdchirt0[r];
        slab[s] = 2.0 * (slab[s] * v0 + S2[s] * dv0) +
                  sg * (S1[s] * fact4 * c0 + S0[s] * dfact4 * c0 + S0[s] * fact4 * dc0);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    double* o = ((which == 0) ? tdhV : tdf0V) + side * nq;
    for (long s = tid; s < nq; s += nt) o[s] = slab[s];
    __syncthreads();
    // value: DST x DCT[2 h V0]
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        const long r = fold_idx(l + 1, m, N);
        slab[s] = 2.0 * (S1[s] * V0[r] + S0[s] * tV0[r]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    double* ob = ((which == 0) ? thVb : tf0Vb) + side * nq;
    for (long s = tid; s < nq; s += nt) ob[s] = slab[s];
}

// ------------------------------------------------------------------ the Lambda pencil
// two-pole data of a point (cond_opt.cpp:3409-3467) with tangents: roots r1, r2 of
// z^2 + (mu - e) z - Sigma_inf (dr = [(de - dmu) r + dSigma_inf] / (2r + mu - e)), residues
// a = r1/(r1-r2), r2/(r2-r1), b = +-CJ1/(r1-r2), den = 1 + exp(-|Re r|/T ...)
struct PolesT {
    cdual r[2], a[2], b[2], den[2];
};
__device__ __forceinline__ PolesT poles_tan(dual e, dual mu, dual Sinf, dual CJ1, double T)
{
    PolesT P;
    const cd B{mu.v - e.v, 0.0}, Cc{-Sinf.v, 0.0};
    const cd Rr = csqrtd(B * B - 4.0 * Cc);
    cd z[2] = {(-B - Rr) / 2.0, (-B + Rr) / 2.0};
    for (int k = 0; k < 2; k++) {
        // one Newton step on z^2 + B z + C, as the reference polishes its roots
        const cd p = 2.0 * z[k] + B;
        const cd y = z[k] * z[k] + B * z[k] + Cc;
        if (cabs2(p) != 0.0 && cabs2(y) > 1.0e-16) {
            const cd zn = z[k] - y / p;
            const cd yn = zn * zn + B * zn + Cc;
            if (cabs2(yn) < cabs2(y)) z[k] = zn;
        }
        const cd dB{mu.d - e.d, 0.0}, dC{-Sinf.d, 0.0};
        const cd dz = -((dB * z[k] + dC) / (2.0 * z[k] + B));
        P.r[k] = cdual{z[k], dz};
    }
    const cdual diff = P.r[0] - P.r[1];
    P.a[0] = P.r[0] / diff;
    P.a[1] = -(P.r[1] / diff);
    const cdual cj{cd{CJ1.v, 0.0}, cd{CJ1.d, 0.0}};
    P.b[0] = cj / diff;
    P.b[1] = -P.b[0];
    for (int k = 0; k < 2; k++) {
        const cdual rt = P.r[k] * (1.0 / T);
        P.den[k] = ((P.r[k].v.x > 0) ? cexpdual(-rt) : cexpdual(rt)) + 1.0;
    }
    return P;
}

// Per-k tangent of the Lambda_jj contributions (the steps of k_jj_pencil, cond_opt.cpp:
// 3179-3822) of every Lambda-grid point q = blk, blk + nblk, ...: workgroup blk adds them into
// its own tangent rows part[blk*2*nwb + n] (bubble) and part[blk*2*nwb + nwb + n] (MT).
// Primal arrays as the forward left them (hT, f0T point-major, FP64 Sigma remainder R);
// tangent arrays t* alike.  LDS: samples and transforms of value and tangent.
template <int L, int NT>
__global__ void __launch_bounds__(NT)
    k_jj_pencil_tan(const double* __restrict__ ek, const double* __restrict__ tek, const double* __restrict__ vel,
                    const double* __restrict__ tvel, const double* __restrict__ hT, const double* __restrict__ thT,
                    const double* __restrict__ f0T, const double* __restrict__ tf0T, const double* __restrict__ dhV,
                    const double* __restrict__ tdhV, const double* __restrict__ df0V,
                    const double* __restrict__ tdf0V, const double* __restrict__ hVb, const double* __restrict__ thVb,
                    const double* __restrict__ f0Vb, const double* __restrict__ tf0Vb, const cd* __restrict__ Self,
                    const cd* __restrict__ tSelf, const double* __restrict__ Sinf2, const double* __restrict__ tSinf2,
                    const double* __restrict__ Sinf3, const double* __restrict__ tSinf3,
                    const double* __restrict__ Sinfp, const double* __restrict__ tSinfp, int N, double T, double mu,
                    double dmu, const cd* __restrict__ tw, const cd* __restrict__ twX, const cd* __restrict__ twS,
                    const cd* __restrict__ twM, const cd* __restrict__ phT, double* part)
{
    constexpr int H = L / 2;
    constexpr int NWB = H + 1;
    constexpr int nw = NT / 64;
    __shared__ double Y[2 * (L + 1)], dY[2 * (L + 1)];
    __shared__ cd X[L], dX[L];
    __shared__ double ES[2 * (2 * NT + 8)];
    __shared__ double sc[16 * nw];
    __shared__ double Jb[2];
    double* Y1 = Y;
    double* Y2 = Y + (L + 1);
    double* dY1 = dY;
    double* dY2 = dY + (L + 1);
    cd* W = (cd*)Y;  // slot values while Y1, Y2 are free
    cd* dW = (cd*)dY;
    const int nk = N / 2 + 1;
    const long nq = (long)nk * (nk - 2);
    const double nf = 1.0 / ((double)N * N);
    const dual Sinf{Sinfp[0], tSinfp[0]};
    const double hstep = 1.0 / (L * T), iT = 1.0 / T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long blk = blockIdx.x, nblk = gridDim.x;
    double* const c1 = part + blk * 2 * NWB;
    double* const c2 = c1 + NWB;
    for (int n = tid; n < 2 * NWB; n += NT) c1[n] = 0.0;
    auto slot_jx = [](int s) { return (s >= H) ? s - H : s + H; };
    auto dyson_idx = [](int s) { return (s < H) ? H - s - 1 : s - H; };
    for (long q = blk; q < nq; q += nblk) {
        const int l = (int)(q / nk), m = (int)(q % nk);
        const long ki = fold_idx(l + 1, m, N);
        const long ie = m + (long)(l + 1) * nk;
        const dual e{ek[ie], tek[ie]}, v{vel[q], tvel[q]};
        const double wt = (m == 0 || m == nk - 1) ? 2.0 : 4.0;
        const dual mud{mu, dmu};
        // (1) f0V, hV pencils with their ends from the boundary data; splines; phased FFT
        for (int j = tid; j <= L; j += NT) {
            const long o = q * (L + 1) + j;
            Y1[j] = (j == 0) ? f0Vb[q] : (j == L) ? f0Vb[nq + q] : f0T[o];
            Y2[j] = (j == 0) ? hVb[q] : (j == L) ? hVb[nq + q] : hT[o];
            dY1[j] = (j == 0) ? tf0Vb[q] : (j == L) ? tf0Vb[nq + q] : tf0T[o];
            dY2[j] = (j == 0) ? thVb[q] : (j == L) ? thVb[nq + q] : thT[o];
        }
        const dual CJ1{-hVb[q] - hVb[nq + q], -thVb[q] - thVb[nq + q]};
        const dual cj01{-f0Vb[q] - f0Vb[nq + q], -tf0Vb[q] - tf0Vb[nq + q]};
        const dual cj02{df0V[q] + df0V[q + nq], tdf0V[q] + tdf0V[q + nq]};
        const dual cj2{dhV[q] + dhV[q + nq], tdhV[q] + tdhV[q + nq]};
        __syncthreads();
        {
            double* const ys[2] = {Y1, Y2};
            const double d0s[2] = {df0V[q], dhV[q]}, dNs[2] = {df0V[q + nq], dhV[q + nq]};
            block_spline_fixed<L, NT, 2>(ys, hstep, d0s, dNs, ES, nullptr);
            double* const ts[2] = {dY1, dY2};
            const double t0s[2] = {tdf0V[q], tdhV[q]}, tNs[2] = {tdf0V[q + nq], tdhV[q + nq]};
            block_spline_fixed<L, NT, 2>(ts, hstep, t0s, tNs, ES, nullptr);
        }
        for (int j = tid; j < L; j += NT) {
            X[j] = twX[j] * cd{(Y1[j + 1] - Y1[j]) / hstep, (Y2[j + 1] - Y2[j]) / hstep};
            dX[j] = twX[j] * cd{(dY1[j + 1] - dY1[j]) / hstep, (dY2[j + 1] - dY2[j]) / hstep};
        }
        const dual cj03{-(Y1[0] + Y1[L]), -(dY1[0] + dY1[L])}, cj3{-(Y2[0] + Y2[L]), -(dY2[0] + dY2[L])};
        __syncthreads();
        block_fft<L, NT>(X, +1, tw);
        block_fft<L, NT>(dX, +1, tw);
        // (2) Matsubara slots: Dyson G, J0 (f0V), J (hV); sums r[0..5] and their tangents
        double r6[12];
        for (int k = 0; k < 12; k++) r6[k] = 0.0;
        const dual C2{Sinf2[ki], tSinf2[ki]}, C3{Sinf3[ki], tSinf3[ki]};
        for (int s = tid; s < L; s += NT) {
            const int jx = slot_jx(s);
            const cd Fj = X[jx], Fc = conjg(X[L - 1 - jx]);
            const cd tFj = dX[jx], tFc = conjg(dX[L - 1 - jx]);
            const cd Af = cd{0.5 * (Fj.x + Fc.x), 0.5 * (Fj.y + Fc.y)};
            const cd Bh = cd{0.5 * (Fj.y - Fc.y), -0.5 * (Fj.x - Fc.x)};
            const cd tAf = cd{0.5 * (tFj.x + tFc.x), 0.5 * (tFj.y + tFc.y)};
            const cd tBh = cd{0.5 * (tFj.y - tFc.y), -0.5 * (tFj.x - tFc.x)};
            const double knj = (jx < H) ? (2 * jx + 1) * kPi * T : (2 * (jx - L) + 1) * kPi * T;
            const double ik = 1.0 / knj, ik2 = ik * ik;
            const cd ts = twS[jx];
            const cd af = ts * Af, ah = ts * Bh, taf = ts * tAf, tah = ts * tBh;
            const cdual J0 = cdl((-cj02 + dual{af.x, taf.x} * ik2) * ik2,
                                 (-cj01 + (cj03 + dual{af.y, taf.y} * ik) * ik2) * ik);
            const cdual Jv = cdl((-cj2 + dual{ah.x, tah.x} * ik2) * ik2,
                                 (-CJ1 + (cj3 + dual{ah.y, tah.y} * ik) * ik2) * ik);
            const double kn = (2.0 * (s - H) + 1) * kPi * T;
            const long o = ki * H + dyson_idx(s);
            cd sg = sigma_mom(Sinf.v, C2.v, C3.v, Self[o], fabs(kn));
            cd tsg = sigma_mom(Sinf.d, C2.d, C3.d, tSelf[o], fabs(kn));
            if (s < H) {
                sg = conjg(sg);
                tsg = conjg(tsg);
            }
            const cdual den{cd{-e.v + mu - sg.x, kn - sg.y}, cd{-e.d + dmu - tsg.x, -tsg.y}};
            const cdual Tc{cd{T, 0.0}, cd{0.0, 0.0}};
            const cdual Gv = Tc / den;
            const cdual t = (Gv * Gv * J0) * iT;
            const double ikn = 1.0 / kn;
            const cdual G2 = Tc / cdl(mud - e, dual{kn, 0.0} + Sinf * ikn);
            const cdual Jn = Gv * Jv - G2 * cdl(dual{0.0, 0.0}, -(CJ1 * ikn));
            const cdual Gn = Gv - G2;
            r6[0] += t.v.x;
            r6[1] += t.v.y;
            r6[2] += Gn.v.x;
            r6[3] += Gn.v.y;
            r6[4] += kn * Gn.v.y;
            r6[5] -= kn * Jn.v.y;
            r6[6] += t.d.x;
            r6[7] += t.d.y;
            r6[8] += Gn.d.x;
            r6[9] += Gn.d.y;
            r6[10] += kn * Gn.d.y;
            r6[11] -= kn * Jn.d.y;
            W[s] = cd{Gn.v.x + Jn.v.y, Gn.v.y + Jn.v.x};
            dW[s] = cd{Gn.d.x + Jn.d.y, Gn.d.y + Jn.d.x};
        }
        // block sums of the twelve slot sums (fixed order: waves, then lanes by wave_sum)
        for (int k = 0; k < 12; k++) r6[k] = wave_sum(r6[k]);
        if (lane == 0)
            for (int k = 0; k < 12; k++) sc[k * nw + wave] = r6[k];
        __syncthreads();
        double S6[12];
        for (int k = 0; k < 12; k++) {
            double a = 0.0;
            for (int i = 0; i < nw; i++) a += sc[k * nw + i];
            S6[k] = a;
        }
        // (3) two-pole asymptote
        const PolesT P = poles_tan(e, mud, Sinf, CJ1, T);
        cdual ad[2], bd[2];
        for (int k = 0; k < 2; k++) {
            ad[k] = P.a[k] / P.den[k];
            bd[k] = P.b[k] / P.den[k];
        }
        // (4) to imaginary time
        block_fft<L, NT>(W, -1, tw);
        block_fft<L, NT>(dW, -1, tw);
        if (tid == 0) {
            cdual Jinf{cd{0.0, 0.0}, cd{0.0, 0.0}};
            for (int k = 0; k < 2; k++) {
                const cdual ej = (P.r[k].v.x > 0) ? cdual{cd{1.0, 0.0}, cd{0.0, 0.0}} : cexpdual(P.r[k] * iT);
                Jinf = Jinf + bd[k] * ej;
            }
            Jb[0] = (cd{W[0].y, 0.0} / phT[L]).x + Jinf.v.x;
            Jb[1] = (cd{dW[0].y, 0.0} / phT[L]).x + Jinf.d.x;
        }
        for (int j = tid; j < L; j += NT) {
            const double t = j / (L * T);
            cdual G0t{cd{0.0, 0.0}, cd{0.0, 0.0}}, Jt{cd{0.0, 0.0}, cd{0.0, 0.0}};
            for (int k = 0; k < 2; k++) {
                const cdual r = P.r[k];
                const cdual ep = (r.v.x > 0) ? cexpdual(r * (-t)) : cexpdual(r * (1.0 / T - t));
                const cdual ej = (r.v.x > 0) ? cexpdual(r * (t - 1.0 / T)) : cexpdual(r * t);
                G0t = G0t - ad[k] * ep;
                Jt = Jt + bd[k] * ej;
            }
            const cd ph = phT[j];
            const cd R = ph * W[j], dR = ph * dW[j];
            X[j] = cd{R.x + G0t.v.x, R.y + Jt.v.x};
            dX[j] = cd{dR.x + G0t.d.x, dR.y + Jt.d.x};
        }
        __syncthreads();
        const dual G0{X[0].x, dX[0].x}, J0t{X[0].y, dX[0].y}, Jbeta{Jb[0], Jb[1]};
        // pole sums at tau = 0+, beta- (k_jj_poles)
        cdual fp[2], fm[2], fj[2], fjb[2];
        for (int k = 0; k < 2; k++) {
            const cdual r = P.r[k], a = P.a[k], b = P.b[k], dn = P.den[k];
            const bool pos = r.v.x > 0;
            const cdual ex = pos ? cexpdual(-(r * iT)) : cexpdual(r * iT);
            fp[k] = pos ? a / dn : a * ex / dn;
            fm[k] = pos ? a * ex / dn : a / dn;
            fj[k] = pos ? r * b * ex / dn : r * b / dn;
            fjb[k] = pos ? r * b / dn : r * b * ex / dn;
        }
        const dual s4{S6[4], S6[10]}, s5{S6[5], S6[11]};
        const dual dGp = s4 + rpart(P.r[0] * fp[0] + P.r[1] * fp[1]);
        const dual dGm = -s4 + rpart(P.r[0] * fm[0] + P.r[1] * fm[1]);
        const dual dJ0 = s5 + rpart(fj[0] + fj[1]);
        const dual dJb = -s5 + rpart(fjb[0] + fjb[1]);
        // (5) Maki-Thompson G(tau) J(-tau) -> Y1, bubble -G(tau) G(-tau) -> Y2 (and tangents)
        for (int j = tid; j < L; j += NT) {
            const cd g = X[j], tg = dX[j];
            Y1[j] = g.x * g.y;
            dY1[j] = tg.x * g.y + g.x * tg.y;
            if (j == 0) {
                Y2[0] = G0.v * (G0.v + 1.0);
                dY2[0] = G0.d * (2.0 * G0.v + 1.0);
            } else {
                Y2[j] = -(g.x * X[L - j].x);
                dY2[j] = -(tg.x * X[L - j].x + g.x * dX[L - j].x);
            }
        }
        if (tid == 0) {
            const dual y1L = (-G0 - 1.0) * Jbeta, y2L = G0 * (G0 + 1.0);
            Y1[L] = y1L.v;
            dY1[L] = y1L.d;
            Y2[L] = y2L.v;
            dY2[L] = y2L.d;
        }
        __syncthreads();
        const dual gd0 = dGp * J0t + G0 * dJ0;
        const dual gdN = dGm * Jbeta - (G0 + 1.0) * dJb;
        const dual MJ1{Y1[0] - Y1[L], dY1[0] - dY1[L]};
        const dual dchi = dGp * (G0 + 1.0) + G0 * dGm;
        double integ[2] = {0.0, 0.0}, tinteg[2] = {0.0, 0.0};
        {
            double* const ys[2] = {Y1, Y2};
            const double d0s[2] = {gd0.v, dchi.v}, dNs[2] = {gdN.v, -dchi.v};
            block_spline_fixed<L, NT, 2>(ys, hstep, d0s, dNs, ES, integ);
            double* const ts[2] = {dY1, dY2};
            const double t0s[2] = {gd0.d, dchi.d}, tNs[2] = {gdN.d, -dchi.d};
            block_spline_fixed<L, NT, 2>(ts, hstep, t0s, tNs, ES, tinteg);
        }
        {
            const double a = wave_sum(integ[1]), b = wave_sum(tinteg[1]);
            if (lane == 0) {
                sc[12 * nw + wave] = a;
                sc[13 * nw + wave] = b;
            }
        }
        const dual MJ3{Y1[0] - Y1[L], dY1[0] - dY1[L]};
        // (6) packed transform of the two moment sequences
        for (int j = tid; j < L; j += NT) {
            const double a = (Y1[j + 1] - Y1[j]) / hstep;
            const double b = (j < H) ? (Y2[j + 1] - Y2[j]) / hstep : -(Y2[L - j] - Y2[L - 1 - j]) / hstep;
            const double ta = (dY1[j + 1] - dY1[j]) / hstep;
            const double tb = (j < H) ? (dY2[j + 1] - dY2[j]) / hstep : -(dY2[L - j] - dY2[L - 1 - j]) / hstep;
            X[j] = cd{a, b};
            dX[j] = cd{ta, tb};
        }
        __syncthreads();
        block_fft<L, NT>(X, -1, tw);
        block_fft<L, NT>(dX, -1, tw);
        const dual wv = (nf * wt) * v;
        for (int n = tid; n <= H; n += NT) {
            dual bub;
            if (n == 0) {
                double a = 0.0, b = 0.0;
                for (int i = 0; i < nw; i++) {
                    a += sc[12 * nw + i];
                    b += sc[13 * nw + i];
                }
                bub = dual{a, b};
                const dual m0 = wv * dual{S6[0], S6[6]};
                c2[0] -= m0.d;
            } else {
                const cd Fn = X[n], Fc = conjg(X[(L - n) & (L - 1)]);
                const cd tFn = dX[n], tFc = conjg(dX[(L - n) & (L - 1)]);
                const double iw = 1.0 / (2 * n * kPi * T), iw2 = iw * iw;
                const cd tm = twM[n];
                const cd a = tm * cd{0.5 * (Fn.x + Fc.x), 0.5 * (Fn.y + Fc.y)};
                const cd ta = tm * cd{0.5 * (tFn.x + tFc.x), 0.5 * (tFn.y + tFc.y)};
                const dual imJ = (-2.0 * MJ1 + (2.0 * MJ3 + 2.0 * dual{a.y, ta.y} * iw) * iw2) * iw;
                c2[n] -= (wv * imJ * iw).d;
                const cd b = tm * cd{0.5 * (Fn.y - Fc.y), -0.5 * (Fn.x - Fc.x)};
                const cd tb = tm * cd{0.5 * (tFn.y - tFc.y), -0.5 * (tFn.x - tFc.x)};
                bub = (-2.0 * dchi + dual{b.x, tb.x} * iw2) * iw2;
            }
            c1[n] -= (2.0 * wv * v * bub).d;
        }
        __syncthreads();
    }
}

// ================================================================== launchers
static size_t grad_jj_lds(int N)
{
    const int nk = N / 2 + 1;
    const size_t nq = (size_t)nk * (nk - 2);
    return ((nq * sizeof(double) + 15) & ~(size_t)15) + 2 * sizeof(double) * (size_t)N;
}

#define GRAD_JJ_DISPATCH_N(N, ...)                                  \
    switch (N) {                                                    \
        case 8: { constexpr int NN = 8; __VA_ARGS__; } break;       \
        case 12: { constexpr int NN = 12; __VA_ARGS__; } break;     \
        case 16: { constexpr int NN = 16; __VA_ARGS__; } break;     \
        case 24: { constexpr int NN = 24; __VA_ARGS__; } break;     \
        case 32: { constexpr int NN = 32; __VA_ARGS__; } break;     \
        case 48: { constexpr int NN = 48; __VA_ARGS__; } break;     \
        case 64: { constexpr int NN = 64; __VA_ARGS__; } break;     \
        case 96: { constexpr int NN = 96; __VA_ARGS__; } break;     \
        case 128: { constexpr int NN = 128; __VA_ARGS__; } break;   \
        case 192: { constexpr int NN = 192; __VA_ARGS__; } break;   \
        case 256: { constexpr int NN = 256; __VA_ARGS__; } break;   \
        default: return hipErrorInvalidValue;                       \
    }

template <typename K>
static void grad_jj_set_lds(K kern, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

bool grad_jj_supported(const Grid& g)
{
    if (g.N < 8) return false;
    switch (g.Nw) {
        case 16: case 32: case 64: case 128: case 256: case 512: case 1024: return true;
        default: return false;
    }
}

hipError_t launch_trace_diffG_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                                  double mu, double* out2, hipStream_t s)
{
    const int nb = (int)((g.nk8 + 3) / 4);
    hipLaunchKernelGGL(k_trace_diffG_tan, dim3(nb), dim3(256), 0, s, d.ekw, t.dekw, d.wq, d.Self, t.Self, d.Sinf2,
                       t.Sinf2, d.Sinf3, t.Sinf3, d.Sinf4, t.Sinf4, d.Sinf, t.Sinf, (int)g.nk8, g.Nw, g.T, mu0, dmu0,
                       mu, t.jjtr);
    return launch_sum_rows(t.jjtr, 2, nb, out2, s);
}

hipError_t launch_hVf0V_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                            hipStream_t s)
{
    const size_t lds = grad_jj_lds(g.N);
    GRAD_JJ_DISPATCH_N(g.N, {
        grad_jj_set_lds(k_hVf0V_tan<NN>, lds);
        hipLaunchKernelGGL(k_hVf0V_tan<NN>, dim3(g.Nw + 1, 2), dim3(GRAD_JJ_THREADS), lds, s, d.ek, t.dek, d.dek,
                           t.vel, d.V, t.V, mu0, dmu0, g.T, g.Nw, t.h, t.f0);
    });
    return hipGetLastError();
}

hipError_t launch_hf0_bnd_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu0, double dmu0,
                              double fact4, double dfact4, hipStream_t s)
{
    const size_t lds = grad_jj_lds(g.N);
    GRAD_JJ_DISPATCH_N(g.N, {
        grad_jj_set_lds(k_hf0_bnd_tan<NN>, lds);
        hipLaunchKernelGGL(k_hf0_bnd_tan<NN>, dim3(4), dim3(GRAD_JJ_THREADS), lds, s, d.ek, t.dek, d.dek, t.vel, d.V0,
                           t.V0, d.dchirt0, t.dchirt0, mu0, dmu0, g.T, g.Nw, fact4, dfact4, t.dhV, t.df0V, t.hVb,
                           t.f0Vb, t.bscr);
    });
    return hipGetLastError();
}

hipError_t launch_jj_pencil_tan(const Grid& g, const DevArrays& d, const TanArrays& t, double mu, double dmu,
                                int nblk, hipStream_t s)
{
    constexpr int NT = 256;
    switch (g.Nw) {
#define JJT_CASE(LV)                                                                                                \
    case LV:                                                                                                        \
        hipLaunchKernelGGL((k_jj_pencil_tan<LV, NT>), dim3(nblk), dim3(NT), 0, s, d.ek, t.dek, d.dek, t.vel, d.hT,    \
                           t.hT, d.f0T, t.f0T, d.dhV, t.dhV, d.df0V, t.df0V, d.hVb, t.hVb, d.f0Vb, t.f0Vb, d.Self,     \
                           t.Self, d.Sinf2, t.Sinf2, d.Sinf3, t.Sinf3, d.Sinf, t.Sinf, g.N, g.T, mu, dmu, d.twN,     \
                           d.twX, d.twS, d.twM, d.ph, t.jjpart);                                                     \
        break;
        JJT_CASE(16)
        JJT_CASE(32)
        JJT_CASE(64)
        JJT_CASE(128)
        JJT_CASE(256)
        JJT_CASE(512)
        JJT_CASE(1024)
#undef JJT_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tpsc
