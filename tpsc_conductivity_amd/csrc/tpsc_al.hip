// tpsc_al.hip -- Aslamazov-Larkin term Lambda_jj3(iq_n) of the current-current correlation
// (cond_opt.cpp:4292-6467, chi3 = true) on gfx950, FP64 throughout.
//
// Set-up (once per call):
//   k_al_grt_slab    G0(r,-tau) on the wedge, tau_j = j beta/Nw, and its slopes (:4374-4445)
//   k_al_hg_slab     hG(q,tau) = DST x DCT [h(r,tau) G0(r,-tau)], slopes (:5159-5261)
//   k_al_hw_pencil   hG(q,iq_p), p < Nw/2, and its 1/qn^2 coefficient hg2 (:5266-5330)
// Per boson frequency q_n of the sparse set (:5350-6400):
//   k_al_HG_pencil   HG(q,iq_p) = Gamma(q,iq_p) [hG(p) - hG(p+n)] / [chi0(p) - chi0(p+n)] on the
//                    window p in [-Nw/2 - n/2, Nw/2 - n/2), to tau with the two-pole tail
//   k_al_prod_slab   HG(r,tau) G0(r,-tau) and HG(r,beta-tau) G0(r,-tau) back to k, slopes
//   k_al_k_pencil    splines + shifted fermionic transforms of both products, then the sum
//                    over k and w_m of v_k G(k,iw_m) G(k,iw_m+iq_n) [HG - HG2]
// Layout: slab-major arrays [tau][q] feed the slab kernels, point-major rows [q][tau] feed the
// pencils (k_transpose in between); complex tau arrays are two real planes (re, im).
// Lambda grid q = m + l*nk: kx = (l+1) pi/(nk-1), ky = m pi/(nk-1), nq = nk (nk-2).
#include "tpsc_launch.hpp"

namespace tpsc {

#define AL_SLAB_THREADS 1024
#define AL_PEN_THREADS 256
#define AL_K_THREADS 512

// G0(k,-tau)/N^2 (cond_opt.cpp:4396-4400) and its tau = 0 slope (:4417-4421)
__device__ __forceinline__ double al_g0m(double x, double t, double T, double nf)
{
    return (x < 0) ? nf * exp(x * t) / (exp(x / T) + 1.0) : nf * exp(x * (t - 1.0 / T)) / (exp(-x / T) + 1.0);
}
__device__ __forceinline__ double al_dg0m(double x, double T, double nf)
{
    return (x < 0) ? nf * x / (exp(x / T) + 1.0) : nf * x * exp(-x / T) / (exp(-x / T) + 1.0);
}

// Block j <= Nw: Grt[j*nk8 + w] = G0(r_w, -tau_j).  Block Nw+1: dGrt[w] = d/dtau G0(r,-tau) at
// tau = 0 and dGrt[nk8 + w] = eps(r) - mu0 delta_r0 - dGrt[w] at tau = beta.
template <int N>
__global__ void __launch_bounds__(AL_SLAB_THREADS)
    k_al_grt_slab(const double* __restrict__ ekw, const double* __restrict__ er, double mu0, double T, int Nw,
                  double* Grt, double* dGrt)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int j = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    const double t = (double)j / (T * Nw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double x = ekw[i] - mu0;
        put_sym(slab, nk, l, m, (j <= Nw) ? al_g0m(x, t, T, nf) : al_dg0m(x, T, nf));
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double v = slab[m + l * nk];
        if (j <= Nw) {
            Grt[(long)j * nk8 + i] = v;
        } else {
            dGrt[i] = v;
            dGrt[nk8 + i] = ((i == 0) ? -mu0 : er[i]) - v;
        }
    }
}

// Block j <= Nw: hGs[j*nq + q] = DST x DCT [h(r,tau_j) G0(r,-tau_j)] with h(r,tau) the
// DST x DCT of the k-space integrand (cond_opt.cpp:4302-4330, 5231-5261).
// Blocks Nw+1+side: slope at tau = 0 / beta of the same product (:5159-5229),
// dh(r) G0(r,-tau_b) + h(r,tau_b) dG0(r,-tau_b), back to q -> dhG[side*nq + q].
template <int N>
__global__ void __launch_bounds__(AL_SLAB_THREADS)
    k_al_hg_slab(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ Grt,
                 const double* __restrict__ dGrt, double mu0, double T, int Nw, double* hGs, double* dhG)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    constexpr int PER = (int)((nq + AL_SLAB_THREADS - 1) / AL_SLAB_THREADS);
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int jt = (b <= Nw) ? b : ((b == Nw + 1) ? 0 : Nw);
    const double t = (double)jt / (T * Nw);
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        slab[s] = hf0_val(0, ek[m + (l + 1) * nk], dek[s], mu0, T, t, nf);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    if (b <= Nw) {
        const double* g = Grt + (long)b * nk8;
        for (long s = tid; s < nq; s += AL_SLAB_THREADS) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            slab[s] *= g[fold_idx(l + 1, m, N)];
        }
        __syncthreads();
        slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
        double* o = hGs + (long)b * nq;
        for (long s = tid; s < nq; s += AL_SLAB_THREADS) o[s] = slab[s];
        return;
    }
    const int side = b - Nw - 1;
    double st[PER];
#pragma unroll
    for (int c = 0; c < PER; c++) {
        const long s = tid + (long)c * AL_SLAB_THREADS;
        st[c] = 0.0;
        if (s < nq) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            st[c] = slab[s] * dGrt[side * nk8 + fold_idx(l + 1, m, N)];
        }
    }
    __syncthreads();
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        slab[s] = h_slope_val(side, ek[m + (l + 1) * nk], dek[s], mu0, T, nf);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    const double* g = Grt + (long)jt * nk8;
#pragma unroll
    for (int c = 0; c < PER; c++) {
        const long s = tid + (long)c * AL_SLAB_THREADS;
        if (s < nq) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            slab[s] = slab[s] * g[fold_idx(l + 1, m, N)] + st[c];
        }
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) dhG[side * nq + s] = slab[s];
}

// Per Lambda point q: clamped spline of hG(q,tau_j), j = 0..L, slopes dhG; hw[q*H] = twice its
// integral, hw[q*H + p] = 2 Re[(1 - e^{-2 pi i p/L}) F_p], F = FORWARD DFT of 6a_j, 0 < p < H;
// hg2[q] = 2 (dhG(0) - dhG(beta)) (cond_opt.cpp:5266-5330).  hGp: point-major rows [q][j].
template <int L>
__global__ void __launch_bounds__(AL_PEN_THREADS)
    k_al_hw_pencil(const double* __restrict__ hGp, const double* __restrict__ dhG, long nq, double T,
                   const cd* __restrict__ tw, const cd* __restrict__ twM, double* hw, double* hg2)
{
    constexpr int NT = AL_PEN_THREADS;
    constexpr int H = L / 2;
    __shared__ double Y[L + 1];
    __shared__ cd X[L];
    __shared__ double ES[2 * NT + 8];
    __shared__ double sc[NT / 64];
    const long q = blockIdx.x;
    const int tid = threadIdx.x;
    const double h = 1.0 / (L * T);
    const double* row = hGp + q * (L + 1);
    for (int j = tid; j <= L; j += NT) Y[j] = row[j];
    const double d0 = dhG[q], dN = dhG[nq + q];
    __syncthreads();
    double integ[1] = {0.0};
    {
        double* const ys[1] = {Y};
        block_spline_fixed<L, NT, 1>(ys, h, &d0, &dN, ES, integ);
    }
    const double tot = block_sum(integ[0], sc);
    for (int j = tid; j < L; j += NT) X[j] = cd{(Y[j + 1] - Y[j]) / h, 0.0};
    __syncthreads();
    block_fft<L, NT>(X, -1, tw);
    double* o = hw + q * H;
    if (tid == 0) {
        o[0] = 2 * tot;
        hg2[q] = 2 * (d0 - dN);
    }
    for (int p = 1 + tid; p < H; p += NT) {
        const cd f = twM[p] * X[p];
        o[p] = 2 * f.x;
    }
}

// Per Lambda point q and boson index nn (cond_opt.cpp:5350-5620):
//   HG_j = Gamma(p1) D_j, D_j = [hG(p1) - hG(p2)]/[chi0(p1) - chi0(p2)], p1 = j (j < H - nn/2) or
//   j - L, p2 = p1 + nn; at p1 = -nn/2 (nn even, chi0(p1) = chi0(p2)) the Richardson value
//   (4 D_{j-1} - D_{j-2})/3; Gamma = 3 Usp chi_sp + Uch chi_ch.
//   HG(q,tau_t) = FORWARD DFT_t [T (HG_j - Hinf_j)] + hg2 tail(tau_t), Hinf the two-pole
//   asymptote, tail its imaginary-time form; slopes at 0+ and beta-.
// chiwP: chi0(q,iq_p) point-major [w][p], p <= H.  Writes HGp planes [c][q][t], t <= L, and
// dHq[(c*2 + side)*nq + q].
template <int L>
__global__ void __launch_bounds__(AL_PEN_THREADS)
    k_al_HG_pencil(const double* __restrict__ chiwP, const double* __restrict__ chi0inf,
                   const double* __restrict__ chi0inf2, const double* __restrict__ hw,
                   const double* __restrict__ hg2, int N, double T, double Usp, double Uch, long nn,
                   const cd* __restrict__ tw, double* HGp, double* dHq)
{
    constexpr int NT = AL_PEN_THREADS;
    constexpr int H = L / 2;
    __shared__ double cw[H + 1], hv[H];
    __shared__ cd X[L];
    __shared__ double sc[NT / 64];
    const int nk = N / 2 + 1;
    const long nq = (long)nk * (nk - 2);
    const long q = blockIdx.x;
    const int tid = threadIdx.x;
    const int l = (int)(q / nk), m = (int)(q % nk);
    const long r = fold_idx(l + 1, m, N);
    for (int p = tid; p <= H; p += NT) cw[p] = chiwP[r * (H + 1) + p];
    for (int p = tid; p < H; p += NT) hv[p] = hw[q * H + p];
    const double ci = chi0inf[r], ci2 = chi0inf2[r], g2 = hg2[q];
    const int nc = (int)(nn / 2);
    const double tpT = 2.0 * kPi * T;
    __syncthreads();
    auto chi0 = [&](long p) -> double {
        if (p < 0) p = -p;
        if (p <= H) return cw[p];
        const double iw = 1.0 / (tpT * p), iw2 = iw * iw;
        return (ci + ci2 * iw2) * iw2;
    };
    auto hG = [&](long p) -> double {
        if (p < 0) p = -p;
        if (p == 0) return hv[0];
        const double iw = 1.0 / (tpT * p), iw2 = iw * iw;
        return (-g2 + hv[(p < H) ? p : H - 1] * iw2) * iw2;
    };
    auto Dq = [&](int j) -> double {
        const long p1 = (j < H - nc) ? j : j - L, p2 = p1 + nn;
        return (hG(p1) - hG(p2)) / (chi0(p1) - chi0(p2));
    };
    const double asp = 0.5 * Usp * ci, ach = 0.5 * Uch * ci;
    double dr = 0.0;
    for (int j = tid; j < L; j += NT) {
        const long p1 = (j < H - nc) ? j : j - L;
        const double c1 = chi0(p1);
        const double G = 3 * Usp * c1 / (1.0 - Usp * c1 / 2) + Uch * c1 / (1.0 + Uch * c1 / 2);
        const double D = (nn % 2 == 0 && p1 == -(nn / 2)) ? (4 * Dq(j - 1) - Dq(j - 2)) / 3.0 : Dq(j);
        const double wn = tpT * p1, w2 = wn * wn;
        const double Hinf = g2 * (3.0 * Usp / (-w2 + asp) + Uch / (-w2 - ach));
        const double f = T * (G * D - Hinf);
        X[j] = cd{f, 0.0};
        dr -= wn * f;
    }
    const double dtauH0 = block_sum(dr, sc);
    block_fft<L, NT>(X, -1, tw);
    // two-pole tail: spin poles +-i a, charge poles +-c (cond_opt.cpp:5556-5618)
    const double beta = 1.0 / T;
    const double a = sqrt(asp), c = sqrt(ach);
    const cd zs0 = cd{0.0, a}, zs1 = cd{0.0, -a};
    const cd es0 = cexpd(cd{0.0, -a * beta}), es1 = cexpd(cd{0.0, a * beta});  // e^{-zs0 beta}, e^{-zs1 beta}
    const cd A0 = cd{1.0, 0.0} / ((es0 - cd{1.0, 0.0}) * (zs0 - zs1));
    const cd A1 = cd{1.0, 0.0} / ((es1 - cd{1.0, 0.0}) * (zs1 - zs0));
    const double ec = exp(-c * beta);
    const double B0 = 1.0 / ((ec - 1.0) * (2.0 * c)), B1 = 1.0 / ((1.0 - ec) * (-2.0 * c));
    double* Pr = HGp + q * (L + 1);
    double* Pi = Pr + nq * (L + 1);
    for (int t = tid; t < L; t += NT) {
        const double tau = t / (L * T);
        const cd tsp = cis(-a * tau) * A0 + cis(a * tau) * A1;
        const double tch = exp(-c * tau) * B0 + exp(-c * (beta - tau)) * B1;
        const cd v = X[t] + g2 * (3.0 * Usp * tsp + cd{Uch * tch, 0.0});
        Pr[t] = v.x;
        Pi[t] = v.y;
    }
    if (tid == 0) {
        // tau = beta
        const cd tspb = cd{1.0, 0.0} / ((cd{1.0, 0.0} - es1) * (zs0 - zs1)) +
                        cd{1.0, 0.0} / ((cd{1.0, 0.0} - es0) * (zs1 - zs0));
        const double tchb = ec / ((ec - 1.0) * (2.0 * c)) + 1.0 / ((1.0 - ec) * (-2.0 * c));
        const cd vb = X[0] + g2 * (3.0 * Usp * tspb + cd{Uch * tchb, 0.0});
        Pr[L] = vb.x;
        Pi[L] = vb.y;
        // slopes at 0+ and beta-
        const cd s0 = -(zs0 * A0) - zs1 * A1;
        const double c0 = -c * B0 + c * ec * B1;
        const cd sb = -(zs0 * es0 * A0) - zs1 * es1 * A1;
        const double cb = -c * ec * B0 + c * B1;
        const cd d0 = cd{0.0, dtauH0} + g2 * (3.0 * Usp * s0 + cd{Uch * c0, 0.0});
        const cd dN = cd{0.0, dtauH0} + g2 * (3.0 * Usp * sb + cd{Uch * cb, 0.0});
        dHq[0 * nq + q] = d0.x;
        dHq[1 * nq + q] = dN.x;
        dHq[2 * nq + q] = d0.y;
        dHq[3 * nq + q] = dN.y;
    }
}

// Products of HG(r,tau) with G0(r,-tau), back to k (cond_opt.cpp:5622-5905, 6000-6170).
// Blocks b < 2(L+1): t = b/2, part c = b%2 (re, im):
//   S = DST x DCT [HG_c(q,tau_t)/N^2] (q -> r), then
//   P1_c(k,tau_t)   = DST x DCT [S G0(r,-tau_t)]
//   P2_c(k,tau_L-t) = DST x DCT [S G0(r,-tau_{L-t})]       (HG2(r,tau) = HG(r,beta-tau))
// Blocks 2(L+1) + 4 which + 2 side + c: the slopes of P1 (which 0) and P2 (which 1) at
// tau = 0 / beta from HG(r,tau_b') dG0(r,-tau_b) + dHG(r) G0(r,-tau_b), with
// dHG2(r,0) = -dHG(r,beta), dHG2(r,beta) = -dHG(r,0).
// Planes: HGs, P1s, P2s [c][t][q]; dHq [c][side][q]; dP [which][side][c][q].
template <int N>
__global__ void __launch_bounds__(AL_SLAB_THREADS)
    k_al_prod_slab(const double* __restrict__ HGs, const double* __restrict__ dHq, const double* __restrict__ Grt,
                   const double* __restrict__ dGrt, int L, double* P1s, double* P2s, double* dP)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    constexpr int PER = (int)((nq + AL_SLAB_THREADS - 1) / AL_SLAB_THREADS);
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int tid = threadIdx.x;
    const long plane = (long)(L + 1) * nq;
    const int b = blockIdx.x;
    int tA, c, which = 0, side = 0;
    const bool main_blk = b < 2 * (L + 1);
    if (main_blk) {
        tA = b >> 1;
        c = b & 1;
    } else {
        const int bb = b - 2 * (L + 1);
        which = bb >> 2;
        side = (bb >> 1) & 1;
        c = bb & 1;
        tA = which ? (side ? 0 : L) : (side ? L : 0);
    }
    const double* src = HGs + c * plane + (long)tA * nq;
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) slab[s] = nf * src[s];
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    double st[PER];
    int fo[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const long s = tid + (long)i * AL_SLAB_THREADS;
        st[i] = 0.0;
        fo[i] = 0;
        if (s < nq) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            fo[i] = (int)fold_idx(l + 1, m, N);
            st[i] = slab[s];
        }
    }
    __syncthreads();
    if (main_blk) {
#pragma unroll 1
        for (int pass = 0; pass < 2; pass++) {
            const int tg = pass ? L - tA : tA;
            const double* g = Grt + (long)tg * nk8;
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const long s = tid + (long)i * AL_SLAB_THREADS;
                if (s < nq) slab[s] = st[i] * g[fo[i]];
            }
            __syncthreads();
            slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
            double* o = (pass ? P2s : P1s) + c * plane + (long)tg * nq;
            for (long s = tid; s < nq; s += AL_SLAB_THREADS) o[s] = slab[s];
            __syncthreads();
        }
        return;
    }
    const double* dg = dGrt + side * nk8;
#pragma unroll
    for (int i = 0; i < PER; i++) st[i] *= dg[fo[i]];
    const int ds = which ? 1 - side : side;
    const double sgn = which ? -1.0 : 1.0;
    const double* dsrc = dHq + (long)(c * 2 + ds) * nq;
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) slab[s] = nf * dsrc[s];
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    const double* g = Grt + (long)(side ? L : 0) * nk8;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const long s = tid + (long)i * AL_SLAB_THREADS;
        if (s < nq) slab[s] = sgn * slab[s] * g[fo[i]] + st[i];
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    double* o = dP + (long)((which * 2 + side) * 2 + c) * nq;
    for (long s = tid; s < nq; s += AL_SLAB_THREADS) o[s] = slab[s];
}

// Dyson G(k, i w_f) = 1/(i w_f + mu - e - Sigma(k,iw_f)), Sigma from the stored form for
// |f + 1/2| < H, from its moments beyond, conjugated for f < 0 (cond_opt.cpp:6314-6377)
template <typename SelfT>
__device__ __forceinline__ cd al_G(const SelfT* __restrict__ Self, long ki, int H, long f, double C2, double C3,
                                   double Sinf, double e, double mu, double T)
{
    const double kn = (2.0 * f + 1) * kPi * T;
    const long g = (f >= 0) ? f : -f - 1;
    cd sg = (g < H) ? sigma_kw(Self, ki * H + g, Sinf, C2, C3, fabs(kn)) : sigma_mom(Sinf, C2, C3, cd{0.0, 0.0}, fabs(kn));
    if (f < 0) sg = conjg(sg);
    return cd{1.0, 0.0} / (cd{mu - e, kn} - sg);
}

// Per Lambda point k (cond_opt.cpp:5907-6400): clamped splines of the four real pencils
// (P1 re/im, P2 re/im) with the slopes dP; HG(k, i w_(f+nn)) from P1 and HG2(k, i w_f) from P2
// by  -i (z_0 + z_L)/w - (z'_0 + z'_L)/w^2 + i (z''_0 + z''_L)/w^3 + (1 - e^{-i w beta/L}) F/w^4,
// F = FORWARD DFT of e^{-i pi t/L} 6a_t (re + i im); f = j (j < H - nn/2) or j - L; then
// contrib[k] = Re[-i U T wt v_k / (N^2 2 qn) sum_j G(f) G(f+nn) (HG_j - HG2_j)].
template <int L, int NT, typename SelfT>
__global__ void __launch_bounds__(NT)
    k_al_k_pencil(const double* __restrict__ P1p, const double* __restrict__ P2p, const double* __restrict__ dP,
                  const double* __restrict__ ek, const double* __restrict__ dek, const SelfT* __restrict__ Self,
                  const double* __restrict__ Sinf2, const double* __restrict__ Sinf3, const double* __restrict__ Sinfp,
                  int N, double T, double mu, double U, long nn, const cd* __restrict__ tw,
                  const cd* __restrict__ twX, const cd* __restrict__ twS, double* contrib)
{
    constexpr int H = L / 2;
    __shared__ __align__(16) double Yb[4 * (L + 1)];
    __shared__ cd X[L];
    __shared__ double ES[4 * (2 * NT + 8)];
    __shared__ double sc[2 * (NT / 64)];
    const int nk = N / 2 + 1;
    const long nq = (long)nk * (nk - 2);
    const long plane = nq * (L + 1);
    const long k = blockIdx.x;
    const int tid = threadIdx.x;
    const int l = (int)(k / nk), m = (int)(k % nk);
    const long ki = fold_idx(l + 1, m, N);
    const double hstep = 1.0 / (L * T);
    const int nc = (int)(nn / 2);
    double* Y[4] = {Yb, Yb + (L + 1), Yb + 2 * (L + 1), Yb + 3 * (L + 1)};
    const double* rows[4] = {P1p + k * (L + 1), P1p + plane + k * (L + 1), P2p + k * (L + 1),
                             P2p + plane + k * (L + 1)};
    for (int j = tid; j <= L; j += NT) {
        Y[0][j] = rows[0][j];
        Y[1][j] = rows[1][j];
        Y[2][j] = rows[2][j];
        Y[3][j] = rows[3][j];
    }
    // dP [which][side][c][q]
    const double d0[4] = {dP[0 * nq + k], dP[1 * nq + k], dP[4 * nq + k], dP[5 * nq + k]};
    const double dN[4] = {dP[2 * nq + k], dP[3 * nq + k], dP[6 * nq + k], dP[7 * nq + k]};
    __syncthreads();
    const cd cw1 = cd{Y[0][0] + Y[0][L], Y[1][0] + Y[1][L]};
    const cd cw2 = cd{Y[2][0] + Y[2][L], Y[3][0] + Y[3][L]};
    block_spline_fixed<L, NT, 4>(Y, hstep, d0, dN, ES, nullptr);
    const cd dd1 = cd{Y[0][0] + Y[0][L], Y[1][0] + Y[1][L]};  // z''(0) + z''(beta)
    const cd dd2 = cd{Y[2][0] + Y[2][L], Y[3][0] + Y[3][L]};
    const cd ds1 = cd{d0[0] + dN[0], d0[1] + dN[1]};
    const cd ds2 = cd{d0[2] + dN[2], d0[3] + dN[3]};
    const double ih = 1.0 / hstep;
    // P1 -> HG(k, i w_(f+nn)), kept in the P1 half of Yb
    for (int t = tid; t < L; t += NT)
        X[t] = conjg(twX[t]) * cd{(Y[0][t + 1] - Y[0][t]) * ih, (Y[1][t + 1] - Y[1][t]) * ih};
    __syncthreads();
    block_fft<L, NT>(X, -1, tw);
    cd* HGK = (cd*)Yb;
    for (int j = tid; j < L; j += NT) {
        const long f = (j < H - nc) ? j : j - L;
        const long s = f + nn;
        const double w = (2 * s + 1) * kPi * T;
        const double iw = 1.0 / w, iw2 = iw * iw;
        // -i cw/w - ds/w^2 + i dd/w^3
        cd v = cd{(cw1.y - ds1.x * iw - dd1.y * iw2) * iw, (-cw1.x - ds1.y * iw + dd1.x * iw2) * iw};
        if (s < L) {
            const long fi = (s < 0) ? s + L : s;
            v = v + (conjg(twS[fi]) * X[fi]) * (iw2 * iw2);
        }
        HGK[j] = v;
    }
    __syncthreads();
    for (int t = tid; t < L; t += NT)
        X[t] = conjg(twX[t]) * cd{(Y[2][t + 1] - Y[2][t]) * ih, (Y[3][t + 1] - Y[3][t]) * ih};
    __syncthreads();
    block_fft<L, NT>(X, -1, tw);
    const double C2 = Sinf2[ki], C3 = Sinf3[ki], Sinf = Sinfp[0];
    const double e = ek[m + (l + 1) * nk];
    double acc[2] = {0.0, 0.0};
    for (int j = tid; j < L; j += NT) {
        const long f = (j < H - nc) ? j : j - L;
        const double w = (2 * f + 1) * kPi * T;
        const double iw = 1.0 / w, iw2 = iw * iw;
        const long fi = (f < 0) ? f + L : f;
        const cd v2 = cd{(cw2.y - ds2.x * iw - dd2.y * iw2) * iw, (-cw2.x - ds2.y * iw + dd2.x * iw2) * iw} +
                      (conjg(twS[fi]) * X[fi]) * (iw2 * iw2);
        const cd G1 = al_G(Self, ki, H, f, C2, C3, Sinf, e, mu, T);
        const cd G2 = al_G(Self, ki, H, f + nn, C2, C3, Sinf, e, mu, T);
        const cd t = (G1 * G2) * (HGK[j] - v2);
        acc[0] += t.x;
        acc[1] += t.y;
    }
    block_sum_k<2>(acc, sc);
    if (tid == 0) {
        const double wt = (m == 0 || m == nk - 1) ? 2.0 : 4.0;
        const double qn = 2.0 * nn * kPi * T;
        const double fac = U * T * wt * dek[k] / ((double)N * N * 2.0 * qn);
        contrib[k] = fac * acc[1];  // Re[-i fac acc]
    }
}

// ================================================================== launchers
static size_t al_slab_lds(int N, bool lambda)
{
    const int nk = N / 2 + 1;
    const size_t n = lambda ? (size_t)nk * (nk - 2) : (size_t)nk * nk;
    return ((n * sizeof(double) + 15) & ~(size_t)15) + 2 * sizeof(double) * (size_t)N;
}

#define AL_DISPATCH_N(N, ...)                                       \
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
#define AL_DISPATCH_L(Lv, ...)                                      \
    switch (Lv) {                                                   \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;     \
        case 32: { constexpr int LL = 32; __VA_ARGS__; } break;     \
        case 64: { constexpr int LL = 64; __VA_ARGS__; } break;     \
        case 128: { constexpr int LL = 128; __VA_ARGS__; } break;   \
        case 256: { constexpr int LL = 256; __VA_ARGS__; } break;   \
        case 512: { constexpr int LL = 512; __VA_ARGS__; } break;   \
        case 1024: { constexpr int LL = 1024; __VA_ARGS__; } break; \
        case 2048: { constexpr int LL = 2048; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;                       \
    }

// ST: stored form of Sigma(k,iwn) (FP64 remainder, or the assembled FP32 value in the mixed path)
#define DISPATCH_AL_SELF(prec, ...) \
    if (prec) {                     \
        using ST = cf;              \
        __VA_ARGS__;                \
    } else {                        \
        using ST = cd;              \
        __VA_ARGS__;                \
    }

template <typename K>
static void al_set_lds(K kern, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

hipError_t launch_al_setup(const Grid& g, const DevArrays& d, const ALArrays& a, double mu0, hipStream_t s)
{
    AL_DISPATCH_N(g.N, {
        const size_t lw = al_slab_lds(g.N, false), lq = al_slab_lds(g.N, true);
        al_set_lds(k_al_grt_slab<NN>, lw);
        hipLaunchKernelGGL(k_al_grt_slab<NN>, dim3(g.Nw + 2), dim3(AL_SLAB_THREADS), lw, s, d.ekw, d.er, mu0, g.T,
                           g.Nw, a.Grt, a.dGrt);
        al_set_lds(k_al_hg_slab<NN>, lq);
        hipLaunchKernelGGL(k_al_hg_slab<NN>, dim3(g.Nw + 3), dim3(AL_SLAB_THREADS), lq, s, d.ek, d.dek, a.Grt,
                           a.dGrt, mu0, g.T, g.Nw, a.P1s, a.dhG);
    });
    hipError_t e = launch_transpose_d(a.P1s, a.P1p, g.nq, g.Nw + 1, s);
    if (e != hipSuccess) return e;
    AL_DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL(k_al_hw_pencil<LL>, dim3((unsigned)g.nq), dim3(AL_PEN_THREADS), 0, s, a.P1p, a.dhG, g.nq,
                           g.T, d.twN, d.twM, a.hw, a.hg2);
    });
    // chi0(q,iq_p) point-major for the per-frequency pencils
    return launch_transpose_d(d.chiw, a.chiwP, g.nk8, g.nwb, s);
}

hipError_t launch_al_freq(const Grid& g, const DevArrays& d, const ALArrays& a, double Usp, double Uch, double mu,
                          double U, long nn, double* contrib, hipStream_t s)
{
    const long plane = (long)(g.Nw + 1) * g.nq;
    AL_DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL(k_al_HG_pencil<LL>, dim3((unsigned)g.nq), dim3(AL_PEN_THREADS), 0, s, a.chiwP, d.chi0inf,
                           d.chi0inf2, a.hw, a.hg2, g.N, g.T, Usp, Uch, nn, d.twN, a.P1p, a.dHq);
    });
    hipError_t e;
    for (int c = 0; c < 2; c++)
        if ((e = launch_transpose_d(a.P1p + c * plane, a.HGs + c * plane, g.Nw + 1, (int)g.nq, s)) != hipSuccess)
            return e;
    AL_DISPATCH_N(g.N, {
        const size_t lq = al_slab_lds(g.N, true);
        al_set_lds(k_al_prod_slab<NN>, lq);
        hipLaunchKernelGGL(k_al_prod_slab<NN>, dim3(2 * (g.Nw + 1) + 8), dim3(AL_SLAB_THREADS), lq, s, a.HGs, a.dHq,
                           a.Grt, a.dGrt, g.Nw, a.P1s, a.P2s, a.dP);
    });
    for (int c = 0; c < 2; c++) {
        if ((e = launch_transpose_d(a.P1s + c * plane, a.P1p + c * plane, g.nq, g.Nw + 1, s)) != hipSuccess) return e;
        if ((e = launch_transpose_d(a.P2s + c * plane, a.P2p + c * plane, g.nq, g.Nw + 1, s)) != hipSuccess) return e;
    }
    DISPATCH_AL_SELF(g.prec, AL_DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL((k_al_k_pencil<LL, AL_K_THREADS, ST>), dim3((unsigned)g.nq), dim3(AL_K_THREADS), 0, s,
                           a.P1p, a.P2p, a.dP, d.ek, d.dek, (const ST*)d.Self, d.Sinf2, d.Sinf3, d.Sinf, g.N, g.T, mu,
                           U, nn, d.twN, d.twX, d.twS, contrib);
    }));
    return hipGetLastError();
}

}  // namespace tpsc
