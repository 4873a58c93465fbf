// tpsc_kernels.hip -- HIP kernels of the TPSC hot path for gfx950.
//
// Stage map (reference = dbergeron1/TPSC_conductivity):
//   k_chi0_tau0, k_chi0_slab     chi0.cpp:310-559   G0(k,+-tau) -> G(r,+-tau) -> chi0(r,tau) -> chi0(q,tau)
//   k_chi0_pencil                chi0.cpp:566-698   clamped spline in tau + DFT -> chi0(q,iqn), tails
//   k_sum_rule                   chi.cpp:738-830    sum-rule traces (one candidate U per launch)
//   k_mu0_root                   green0.cpp:182-203, 240-262 finite-lattice density and the mu0 root
//   k_vertex_slab/_pencil        chi.cpp:85-170     V(r,tau)
//   k_sigma_bnd/_slab/_pencil    green.cpp:1743-2018 Sigma(k,iwn) moments + remainder
//   k_trace_diffG, k_trace_selfG green.cpp:2929-3182 traces for mu and consistency checks
//   k_hf0_slab, k_hf0_bnd,
//   k_hVf0V_slab, k_jj_pencil,
//   k_kx0_rows                   cond_opt.cpp:2584-4173 Lambda_jj bubble + Maki-Thompson + kx0
// Slab kernels hold one (nk x nk) slab in LDS (133 KB at N=256 in FP64) and transform it with
// wavefront FFTs; pencil kernels hold one tau-pencil per workgroup.
// Precision: S is the storage and slab-arithmetic type of the tau arrays (chi0(q,tau),
// chi0(r,tau), V, Sigma(k,tau), h, f0); chi0(q,iqn) is stored in FP64 in both paths.  S = double is the FP64 path.  S = float is the mixed path: the
// slab transforms and pointwise products run in FP32, while the tau = 0 / beta boundary data,
// the spline moments, the tail moments, the pencil Fourier sums, the root-finder reductions
// and the Lambda_jj Matsubara sums stay in FP64.  Self holds the remainder R(k,n) as FP64
// complex (S = double) or the assembled Sigma(k,iwn) as FP32 complex (S = float).
#include <algorithm>
#include <unordered_map>

#include "tpsc_launch.hpp"

namespace tpsc {

#define SLAB_THREADS 1024
// waves per SIMD a slab kernel is compiled for: two FP32 workgroups (68 KB LDS each at N=256)
// fit one CU when the kernel stays within 64 VGPRs; an FP64 slab (133 KB) runs alone
template <typename S>
struct SlabOcc {
    static constexpr int waves = 4;
};
template <>
struct SlabOcc<float> {
    static constexpr int waves = 8;
};
#define PEN_THREADS 256

// ================================================================== small helpers
__global__ void k_fill_tw(cd* tw, int L)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < L) {
        double s, c;
        sincospi(2.0 * (double)t / (double)L, &s, &c);
        tw[t] = cd{c, -s};
    }
}

// sum n partials in a fixed order (single block)
__global__ void k_finalize(const double* __restrict__ part, int n, double* out)
{
    __shared__ double sc[16];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
    a = block_sum(a, sc);
    if (threadIdx.x == 0) out[0] = a;
}

// out[b] = sum of the n partials part[b*n .. b*n+n) in a fixed order; one block per value
__global__ void k_finalize_k(const double* __restrict__ part, int n, double* out)
{
    __shared__ double sc[16];
    const double* p = part + (long)blockIdx.x * n;
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += p[i];
    a = block_sum(a, sc);
    if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// ================================================================== device-driven root search
// mirror (mapped host memory) = {status, root, x, passes, f, f', lo, hi} of the search: written
// here with status "running", and by the step that settles it (one system-scope write per
// search instead of one per pass)
// chi0max != nullptr (the Usp search): the bracket is (0, 2/chi0max - 2 eps) from the device
// maximum of chi0, with the start s0.x kept when it lies inside and the midpoint otherwise, so
// that the host need not wait for chi0max before the first pass
// Point scalars formed on the device by the queued solve (tpsc_capi.cpp) where the host passes
// them by value otherwise, with the host's rounding (no contraction into fused multiply-adds):
// U (3 Usp + Uch) / div, and the charge sum rule's constant part
// A0 - (nsr + Usp nsr^2 / (2U) - nsr^2), A0 = n - n^2/2 (chi.cpp:830-843).
__device__ __forceinline__ double vfact(double U, double Usp, double Uch, double div)
{
#pragma clang fp contract(off)
    return U * (3.0 * Usp + Uch) / div;
}
__device__ __forceinline__ double uch_a0(double A0, double nsr, double U, double Usp)
{
#pragma clang fp contract(off)
    return A0 - (nsr + Usp * nsr * nsr / (2.0 * U) - nsr * nsr);
}

// a0src (device Usp, may be null): a0 = uch_a0(a0A, nsr, U, *a0src); mu0p (may be null): the
// state's mu0 from the device
__global__ void k_halley_init(HalleyState* st, HalleyState s0, const double* chi0max, double* mirror,
                              const double* mu0p, const double* a0src, double a0A, double nsr, double U)
{
    HalleyState s = s0;
    if (chi0max) {
        const double lo = 0.0, hi = 2.0 / chi0max[0] - 2 * DBL_EPSILON;
        const double x0 = (s0.x > lo && s0.x < hi) ? s0.x : 0.5 * (lo + hi);
        s = halley_start(x0, lo, hi, s0.itmax, s0.a0, s0.a1);
        s.T = s0.T;
        s.mu0 = s0.mu0;
    }
    if (mu0p) s.mu0 = *mu0p;
    if (a0src) s.a0 = uch_a0(a0A, nsr, U, *a0src);
    *st = s;
    mirror[0] = s.status;
    mirror[1] = s.root;
    mirror[2] = s.x;
    mirror[3] = s.ne;
}

// One Halley step after a pass, by one workgroup of 256 threads: the three partial sums of
// part[k * n + i] (k = 0, 1, 2), each in the order of k_finalize_k, the update of st, and the
// outcome in mirror once the search has settled.  A settled search is left as it is.
__device__ __forceinline__ void halley_step_block(const double* part, int n, HalleyState* st, double* mirror,
                                                  double* sc)
{
    const HalleyState s0 = *st;
    if (s0.status != kRootRunning) return;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        v[0] += part[i];
        v[1] += part[(long)n + i];
        v[2] += part[2L * n + i];
    }
    block_sum_k<3>(v, sc);
    if (threadIdx.x == 0) {
        HalleyState s = s0;
        halley_update(s, v);
        *st = s;
        if (s.status != kRootRunning) {
            mirror[0] = s.status;
            mirror[1] = s.root;
            mirror[2] = s.x;
            mirror[3] = s.ne;
            mirror[4] = s.f;
            mirror[5] = s.f1;
            mirror[6] = s.lo;
            mirror[7] = s.hi;
        }
    }
}

// True in the workgroup of a pass that finishes last, once every workgroup's partial sums
// (stored by its thread 0) are visible to it; that workgroup resets the ticket for the next pass.
// No workgroup waits for another: the others return at once.  Only thread 0 releases (its
// part[] stores, written back from its XCD's L2 before the ticket) and only the last workgroup
// acquires: a full fence in every workgroup would also invalidate each XCD's L2 while the
// other workgroups still stream the pass's input from it.
__device__ __forceinline__ bool last_workgroup(unsigned* ticket, unsigned nblk)
{
    __shared__ unsigned last;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nblk - 1;
    }
    __syncthreads();
    if (!last) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// the Halley step as a kernel of its own, launched after a pass of more than kFusedStepMax
// workgroups
__global__ void __launch_bounds__(256) k_halley_step(const double* part, int n, HalleyState* st, double* mirror)
{
    __shared__ double sc[3 * 4];
    halley_step_block(part, n, st, mirror, sc);
}

// ================================================================== chi0 slabs
// G(k,t)/N^2, chi0.cpp:468-471 (form==0) and :535-538 (form==1, tau = beta/2)
__device__ __forceinline__ double g0_plus(double x, double t, double T, double nk02, int form)
{
    if (x >= 0.0) return exp(-t * x) * (-1.0 / (exp(-x / T) + 1.0)) / nk02;
    if (form == 0) return -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02;
    return -exp((1.0 / T - t) * x) / ((exp(x / T) + 1.0) * nk02);
}
// G(k,-t)/N^2, chi0.cpp:486-489
__device__ __forceinline__ double g0_minus(double x, double t, double T, double nk02)
{
    if (x >= 0.0) return -exp((t - 1.0 / T) * x) * (-1.0 / (exp(-x / T) + 1.0)) / nk02;
    return exp(t * x) / ((exp(x / T) + 1.0) * nk02);
}

// tau = 0 slice plus the tau-derivative data (chi0.cpp:323-447); one block.
// tmp[0:nk8] <- G(r,0), tmp[nk8:2nk8] <- dG(r,0-), tmp[2nk8:3nk8] <- staging
// The boundary data are tail moments: this kernel runs in FP64 for both precisions and
// only rounds the slice-0 values it stores into the S arrays.
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS)
    k_chi0_tau0(const double* __restrict__ ekw, const double* __restrict__ er, const double* __restrict__ mu0p,
                double T, S* chi, S* chirt, double* dchirt0, double* dGrt0p, double* dchiqt0, double* chirt0,
                double* tmp)
{
    const double mu0 = *mu0p;
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    double* G0r = tmp;
    double* dGm = tmp + nk8;
    double* stg = tmp + 2 * nk8;
    const int tid = threadIdx.x, nt = blockDim.x;
    // d/dtau G(r,0+) = DCT[-ekmu Gkt / N^2]
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double x = ekw[i] - mu0;
        const double Gkt = -1.0 / (exp(-x / T) + 1.0);
        put_sym(slab, nk, l, m, -x * Gkt / nk02);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double v = slab[m + l * nk];
        dGrt0p[i] = v;
        dGm[i] = ((i == 0) ? -mu0 : er[i]) - v;
    }
    __syncthreads();
    // G(r,0)
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double x = ekw[i] - mu0;
        put_sym(slab, nk, l, m, (-1.0 / (exp(-x / T) + 1.0)) / nk02);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double g = slab[m + l * nk];
        G0r[i] = g;
        const double c = (i == 0) ? -2.0 * g * (g + 1.0) : -2.0 * g * g;
        chirt[i] = (S)c;
        chirt0[i] = c;
        stg[i] = c;
    }
    __syncthreads();
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        put_sym(slab, nk, l, m, stg[i]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        chi[i] = (S)slab[m + l * nk];
    }
    __syncthreads();
    // d/dtau chi0(r,0) and its transform
    for (long i = tid; i < nk8; i += nt) {
        const double g = G0r[i];
        const double v = (i == 0) ? -2.0 * (dGrt0p[i] * (1.0 + g) + dGm[i] * g)
                                  : -2.0 * (dGrt0p[i] * g + dGm[i] * g);
        dchirt0[i] = v;
        stg[i] = v;
    }
    __syncthreads();
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        put_sym(slab, nk, l, m, stg[i]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        dchiqt0[i] = slab[m + l * nk];
    }
}

// slices j = 1 .. nwb-1 (chi0.cpp:454-559); block b handles j = b + 1
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS, SlabOcc<S>::waves)
    k_chi0_slab(const double* __restrict__ ekw, const double* __restrict__ mu0p, double T, int Nw, S* chi, S* chirt)
{
    const double mu0 = *mu0p;
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    S* slab = (S*)smem;
    auto* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int nwb = Nw / 2 + 1;
    const int j = blockIdx.x + 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    S* cr = chirt + (long)j * nk8;
    S* cq = chi + (long)j * nk8;
    if (j < nwb - 1) {
        const double t = j / (Nw * T);
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym_t(slab, nk, l, m, (S)g0_plus(ekw[i] - mu0, t, T, nk02, 0));
        }
        __syncthreads();
        slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
        for (long i = tid; i < nk8; i += nt) {  // stage G(r,t) in chirt[j]
            int l, m;
            wedge_lm(i, l, m);
            cr[i] = slab[m + l * nk];
        }
        __syncthreads();
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym_t(slab, nk, l, m, (S)g0_minus(ekw[i] - mu0, t, T, nk02));
        }
        __syncthreads();
        slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
        for (long i = tid; i < nk8; i += nt) {  // chi0(r,t) = -2 G(r,-t) G(r,t)
            int l, m;
            wedge_lm(i, l, m);
            cr[i] = (S)(-2) * slab[m + l * nk] * cr[i];
        }
        __syncthreads();
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym_t(slab, nk, l, m, cr[i]);
        }
        __syncthreads();
    } else {
        const double t = 1.0 / (2.0 * T);
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            put_sym_t(slab, nk, l, m, (S)g0_plus(ekw[i] - mu0, t, T, nk02, 1));
        }
        __syncthreads();
        slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
        for (int s = tid; s < nk * nk; s += nt) slab[s] = (S)2 * slab[s] * slab[s];
        __syncthreads();
        for (long i = tid; i < nk8; i += nt) {
            int l, m;
            wedge_lm(i, l, m);
            cr[i] = slab[l + m * nk];
        }
        __syncthreads();
    }
    slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        cq[i] = slab[m + l * nk];
    }
}

// ================================================================== chi0 pencils
// chi0.cpp:586-698: per wedge q, clamped spline of chi0(q,tau_j), j=0..Nw/2, slopes
// (dchiqt0[q], 0); F = BACKWARD DFT_Nw of the antisymmetrised 6a_j; tails.
// Reads chi0(q,tau) from chit, writes chi0(q,iqn) to chiw in FP64 (chiw may alias chit in
// the FP64 path: each block reads its column completely before writing it).
template <int L, typename S>
__global__ void __launch_bounds__(PEN_THREADS)
    k_chi0_pencil(const S* chit, double* chiw, const double* __restrict__ dchiqt0, double* chi0inf,
                  double* chi0inf2, long nk8, double T, const cd* __restrict__ tw)
{
    constexpr int NS = L / 2;  // intervals
    constexpr int PJ = (NS + PEN_THREADS - 1) / PEN_THREADS;  // intervals per thread
    // The spline samples, then in place their moments (NS + 1 doubles), live in the FFT buffer
    // until the transform input is formed: 36 KB of LDS per workgroup, four per CU.
    static_assert(NS + 1 <= 2 * L, "the spline array fits in the transform buffer");
    __shared__ cd X[L];
    __shared__ double ES[2 * PEN_THREADS + 8];
    __shared__ double sc[16];
    const long q = xcd_remap(blockIdx.x, gridDim.x);
    if (q >= nk8) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* y = (double*)X;
    for (int j = tid; j <= NS; j += nt) y[j] = (double)chit[q + (long)j * nk8];
    const double d0 = dchiqt0[q], dN = 0.0;
    const double h = 1.0 / (L * T);
    __syncthreads();
    // y -> moments M in place; integ: this thread's part of the spline's integral
    double integ = 0.0, a6v[PJ];
    {
        double* const ys[1] = {y};
        block_spline_fixed<NS, PEN_THREADS, 1>(ys, h, &d0, &dN, ES, &integ);
    }
    const double* M = y;
#pragma unroll
    for (int c = 0; c < PJ; c++) {
        const int j = tid + c * nt;
        a6v[c] = (j < NS) ? (M[j + 1] - M[j]) / h : 0.0;
    }
    __syncthreads();  // every read of y and M is done: X may be written
#pragma unroll
    for (int c = 0; c < PJ; c++) {
        const int j = tid + c * nt;
        if (j < NS) {
            X[j] = cd{a6v[c], 0.0};
            X[L - 1 - j] = cd{-a6v[c], 0.0};
        }
    }
    integ = block_sum(integ, sc);
    block_fft<L, PEN_THREADS>(X, +1, tw);
    if (tid == 0) chiw[q] = 2 * integ;
    if (q != 0) {
        const double inf = -2.0 * d0;
        for (int j = 1 + tid; j <= NS; j += nt) {
            const double wn = 2 * j * kPi * T;
            // cos, sin of 2 pi j / L from the twiddle table tw[j] = exp(-2 pi i j / L)
            const cd e = tw[j];
            chiw[q + (long)j * nk8] =
                inf / (wn * wn) + ((1.0 - e.x) * X[j].x - e.y * X[j].y) / (wn * wn * wn * wn);
        }
        if (tid == 0) {
            chi0inf[q] = inf;
            chi0inf2[q] = 2.0 * X[NS].x;
        }
    } else {
        for (int j = 1 + tid; j <= NS; j += nt) chiw[(long)j * nk8] = 0.0;
        if (tid == 0) {
            chi0inf[0] = 0.0;
            chi0inf2[0] = 0.0;
        }
    }
}

// chi0max and its wedge position (chi0.cpp:703-717): largest value, ties to the
// larger packed index (the reference scans l, m downwards with a strict >).
// out2 (may be null): a second copy of the maximum
__global__ void k_chi0max(const double* __restrict__ chi, long nk8, double* out, int* imax, double* out2)
{
    __shared__ double sv[1024];
    __shared__ long si[1024];
    double best = 0.0;
    long bi = -1;
    for (long i = threadIdx.x; i < nk8; i += blockDim.x) {
        const double v = chi[i];
        if (v > best || (v == best && bi >= 0 && i > bi) ) {
            if (v > 0.0 || bi >= 0) {
                best = v;
                bi = i;
            }
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double v2 = sv[threadIdx.x + s];
            const long i2 = si[threadIdx.x + s];
            if (i2 >= 0 && (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 > si[threadIdx.x]) ||
                            si[threadIdx.x] < 0)) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = (si[0] >= 0) ? sv[0] : 0.0;
        imax[0] = (int)((si[0] >= 0) ? si[0] : 0);
        if (out2) out2[0] = out[0];
    }
}

// ================================================================== reductions
// sum-rule trace sum_{n=0}^{Nw} sum_q (w T/Nq) [chi/(1 + sU chi/2) - chi]  (chi.cpp:738-830)
// sU = -Usp for the spin channel, +Uch for the charge channel.  Grid: x over wedge points
// (256 per block), y over RG row groups; rows j = y, y+RG, ...  One partial per block.
__global__ void __launch_bounds__(256)
    k_sum_rule(const double* __restrict__ chi, const double* __restrict__ inf1, const double* __restrict__ inf2,
               const double* __restrict__ wq, int nk8, int nwmax, double T, double Nq, double sU, double* part)
{
    __shared__ double sc[16];
    const int q = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (q < nk8) {
        const double w16 = 2.0 * wq[q];
        const double a1 = inf1[q], a2 = inf2[q];
        for (int j = blockIdx.y; j <= 2 * nwmax; j += gridDim.y) {
            const double w = (j == 0) ? 0.5 * w16 : w16;
            double c;
            if (j <= nwmax)
                c = chi[(long)j * nk8 + q];
            else {
                const double wn = 2 * j * kPi * T;
                c = a1 / (wn * wn) + a2 / (wn * wn * wn * wn);
            }
            acc += w * T / Nq * (c / (1 + sU * c / 2.0) - c);
        }
    }
    acc = block_sum(acc, sc);
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// The same trace and its first two derivatives with respect to Ux, where sU = sgn * Ux:
// d/dUx [c/(1 + sU c/2)] = -sgn c^2/2 / (1 + sU c/2)^2, d2/dUx2 = c^3/2 / (1 + sU c/2)^3.
// Partials: part[k * nblk + blk], k = 0, 1, 2.  A thread sums its rows j = y, y + RG, ... in
// increasing j: the stored rows j <= nwmax in batches of kSR loads issued before any use (one
// dependent load per row would leave HBM idle), then the asymptotic rows, whose 1/wn^2 is
// common to the block.
constexpr int kSR = 8;
__device__ __forceinline__ void sum_rule3_term(double c, double w, double sU, double sgn, double acc[3])
{
    const double r = rcp_nr(1 + sU * c / 2.0);
    acc[0] += w * (c * r - c);
    acc[1] -= w * sgn * 0.5 * c * c * r * r;
    acc[2] += w * 0.5 * c * c * c * r * r * r;
}
// Rows of small chi: with y = -sU/2 and |y c| <= rho for every element of rows j >= n0,
//   c r - c = sum_{k>=1} y^k c^(k+1),  c^2 r^2 = sum_{k>=0} (k+1) y^k c^(k+2),
//   c^3 r^3 = sum_{k>=0} (k+1)(k+2)/2 y^k c^(k+3),   r = 1/(1 - y c),
// so their part of the three sums is a polynomial in y over the moments
// M_m(n0) = sum_{j>=n0} sum_q w c^m, m = 2..kSRK+3, stored once per chi0 (k_sr_moments,
// k_sr_suffix).  rho = 1/16 and kSRK = 18: the first omitted term of the c^3 r^3 series is
// below 210 * 16^-19 ~ 3e-21 of the moment M_3 it is measured against.  Every chi0(q,iqn) and
// its asymptotic continuation are positive, so the moments carry no cancellation.
constexpr double kSRRho = 1.0 / 16.0;
// Ux (and T) come from the root search st, so that a captured pass graph serves every point;
// a settled search leaves nothing to do.  ticket != null: the workgroup that finishes last takes
// the Halley step.  srm != null: rows j >= n0(Ux) (all rows whose elements satisfy
// |Ux c/2| <= rho, from the suffix maxima srmax) are not read; workgroup (0, 0) adds their
// part from the suffix moments srm[n0].
__global__ void __launch_bounds__(256)
    k_sum_rule3(const double* __restrict__ chi, const double* __restrict__ inf1, const double* __restrict__ inf2,
                const double* __restrict__ wq, int nk8, int nwmax, double Nq, double sgn, HalleyState* st,
                double* part, unsigned* ticket, double* mirror, const double* __restrict__ srm,
                const double* __restrict__ srmax)
{
    __shared__ double sc[3 * 4];
    if (st->status != kRootRunning) return;
    const double Ux = st->x, T = st->T;
    const int q = blockIdx.x * 256 + threadIdx.x;
    const int RG = gridDim.y;
    const double sU = sgn * Ux;
    const int nrows = 2 * nwmax + 1;
    int n0 = nrows;
    if (srm) {
        // srmax is non-increasing in j: n0 = the number of rows above the threshold
        const double thr = kSRRho / (0.5 * fabs(Ux));
        double cnt = 0.0;
        for (int j = threadIdx.x; j < nrows; j += 256) cnt += (srmax[j] > thr) ? 1.0 : 0.0;
        n0 = (int)block_sum(cnt, sc);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    if (q < nk8) {
        const double wb = 2.0 * wq[q] * T / Nq;
        const double a1 = inf1[q], a2 = inf2[q];
        const int jst = min(nwmax, n0 - 1), jas = min(2 * nwmax, n0 - 1);
        int j = blockIdx.y;
        for (; j + (kSR - 1) * RG <= jst; j += kSR * RG) {
            double c[kSR];
#pragma unroll
            for (int u = 0; u < kSR; u++) c[u] = chi[(long)(j + u * RG) * nk8 + q];
#pragma unroll
            for (int u = 0; u < kSR; u++) sum_rule3_term(c[u], (j + u * RG == 0) ? 0.5 * wb : wb, sU, sgn, acc);
        }
        for (; j <= jst; j += RG) sum_rule3_term(chi[(long)j * nk8 + q], (j == 0) ? 0.5 * wb : wb, sU, sgn, acc);
        for (; j <= jas; j += RG) {
            const double wn = 2 * j * kPi * T;
            const double i2 = rcp_nr(wn * wn);
            sum_rule3_term((a1 + a2 * i2) * i2, wb, sU, sgn, acc);
        }
    }
    if (srm && n0 < nrows && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        const double* M = srm + (long)n0 * kSRM;  // M[m - 2] = M_m(n0)
        const double y = -0.5 * sU;
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        for (int k = kSRK; k >= 0; k--) {
            if (k >= 1) f0 = f0 * y + M[k - 1];
            f1 = f1 * y + (k + 1) * M[k];
            f2 = f2 * y + 0.5 * (k + 1) * (k + 2) * M[k + 1];
        }
        acc[0] += y * f0;
        acc[1] -= sgn * 0.5 * f1;
        acc[2] += 0.5 * f2;
    }
    block_sum_k<3>(acc, sc);
    const int nblk = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[k * nblk + b] = acc[k];
    if (ticket && last_workgroup(ticket, nblk)) halley_step_block(part, nblk, st, mirror, sc);
}

// Row moments of the sum-rule terms (one workgroup per row j <= 2 nwmax, stored rows from chi,
// the rest from the asymptote a1/wn^2 + a2/wn^4 as k_sum_rule3 forms them): R[j][m - 2] =
// sum_q w_j wb_q c^m, m = 2..kSRK+3 (w_0 = 1/2), and rmax[j] = max_q |c|.
__global__ void __launch_bounds__(256)
    k_sr_moments(const double* __restrict__ chi, const double* __restrict__ inf1, const double* __restrict__ inf2,
                 const double* __restrict__ wq, int nk8, int nwmax, double Nq, double T, double* R, double* rmax)
{
    __shared__ double sc[kSRM * 4];
    const int j = blockIdx.x;
    double acc[kSRM];
#pragma unroll
    for (int m = 0; m < kSRM; m++) acc[m] = 0.0;
    double mx = 0.0;
    const double wn = 2 * j * kPi * T;
    const double i2 = (j > nwmax) ? rcp_nr(wn * wn) : 0.0;
    for (int q = threadIdx.x; q < nk8; q += 256) {
        const double wb = 2.0 * wq[q] * T / Nq * ((j == 0) ? 0.5 : 1.0);
        const double c = (j <= nwmax) ? chi[(long)j * nk8 + q] : (inf1[q] + inf2[q] * i2) * i2;
        // c^i, i < kSRM, as a product tree (c^i = c^h c^(i-h), h the largest power of two below
        // i): five multiplies deep instead of a chain of kSRM
        double cp[kSRM];
        cp[0] = 1.0;
        cp[1] = c;
#pragma unroll
        for (int i = 2; i < kSRM; i++) {
            const int h = 1 << (31 - __builtin_clz(i - 1));
            cp[i] = cp[h] * cp[i - h];
        }
        const double w2 = wb * c * c;
#pragma unroll
        for (int m = 0; m < kSRM; m++) acc[m] = fma(w2, cp[m], acc[m]);
        mx = fmax(mx, fabs(c));
    }
    block_sum_k<kSRM>(acc, sc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < kSRM; m++) R[(long)j * kSRM + m] = acc[m];
    // block maximum through the same scratch
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) sc[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) rmax[j] = fmax(fmax(sc[0], sc[1]), fmax(sc[2], sc[3]));
}

// Suffix sums (columns 0..kSRM-1 of R, in place) and suffix maxima (rmax, in place) over the
// rows 0..nrows-1: column b of one workgroup of 1024 threads, thread t owning rows
// [t ch, t ch + ch); the thread totals are combined by a fixed-order scan, so the result does
// not depend on timing.
__global__ void __launch_bounds__(1024) k_sr_suffix(double* R, double* rmax, int nrows)
{
    __shared__ double tot[1024];
    const int b = blockIdx.x;  // b < kSRM: a moment column; b == kSRM: the maxima
    const bool ismax = b == kSRM;
    const int t = threadIdx.x;
    const int ch = (nrows + 1023) / 1024;
    constexpr int kMaxCh = 8;  // nrows <= 8192
    double v[kMaxCh];
    double s = 0.0;
    for (int i = ch - 1; i >= 0; i--) {
        const int j = t * ch + i;
        const double x = (j < nrows) ? (ismax ? rmax[j] : R[(long)j * kSRM + b]) : 0.0;
        s = ismax ? fmax(s, x) : s + x;
        v[i] = s;
    }
    tot[t] = s;
    __syncthreads();
    // inclusive suffix scan of the thread totals (Hillis-Steele, fixed order)
    for (int o = 1; o < 1024; o <<= 1) {
        const double a = tot[t], c = (t + o < 1024) ? tot[t + o] : 0.0;
        __syncthreads();
        tot[t] = ismax ? fmax(a, c) : a + c;
        __syncthreads();
    }
    const double after = (t + 1 < 1024) ? tot[t + 1] : 0.0;  // rows of the threads after t
    for (int i = 0; i < ch; i++) {
        const int j = t * ch + i;
        if (j < nrows) {
            if (ismax)
                rmax[j] = fmax(v[i], after);
            else
                R[(long)j * kSRM + b] = v[i] + after;
        }
    }
}

// wedge points per thread in k_mu0_root: 16 waves hold the 8385 points of N = 256
constexpr int kMu0K = 9;

// green0::find_mu (green0.cpp:240-262) in one workgroup, on block-wide sums of the
// finite-lattice density n(mu) = (2/N^2) sum_wedge w f(eps - mu) (green0.cpp:182-203): every
// thread sees the same sums, so the control flow stays uniform.  One evaluation gives n - dens
// and its first two derivatives from the same exponential (f' = f(1-f)/T,
// f'' = f(1-f)(1-2f)/T^2), and the Halley iteration of tpsc_roots.hpp runs from the reference's
// start 0 in its bracket [-8, 8] (green0.cpp:248-257).  n(mu) is increasing, so it fails only
// when the root lies outside the bracket or the derivative underflows (no level within ~700 T
// of the iterate); the derivative-free bracketed search (bracket_root, which widens the
// interval) then takes over.  out = {found, mu0, evaluations}.
// Each thread keeps its kMu0K strided wedge points (eps, weight) in registers for all the
// evaluations; a padded slot has weight 0 and adds +0.  The launcher sizes the workgroup so
// that blockDim.x * kMu0K >= nk8.
// The root goes to out (mapped host memory) and to the device slot mu0v, which the chi0 and
// h/f0 kernels read, so that the host need not wait for it; a failed search leaves mu0v as it
// was (the previous mu0).  A preset value (not NaN) is stored without a search.
__global__ void __launch_bounds__(1024) k_mu0_root(const double* __restrict__ ekw, const double* __restrict__ wq,
                                                   long nk8, double Nk2, double T, double dens, double preset,
                                                   double* out, double* mu0v)
{
    if (!isnan(preset)) {
        if (threadIdx.x == 0) {
            out[0] = 1.0;
            out[1] = preset;
            out[2] = 0.0;
            *mu0v = preset;
        }
        return;
    }
    __shared__ double sc[3 * 16];
    double e[kMu0K], w[kMu0K];
#pragma unroll
    for (int k = 0; k < kMu0K; k++) {
        const long i = threadIdx.x + (long)k * blockDim.x;
        e[k] = (i < nk8) ? ekw[i] : 0.0;
        w[k] = (i < nk8) ? 2.0 * wq[i] / Nk2 : 0.0;
    }
    const double iT = 1.0 / T;
    // v = {n(mu) - dens, n'(mu), n''(mu)}
    auto density3 = [&](double mup, double* v) {
        v[0] = v[1] = v[2] = 0.0;
#pragma unroll
        for (int k = 0; k < kMu0K; k++) {
            const double f = 1.0 / (exp((e[k] - mup) * iT) + 1.0), g = f * (1.0 - f) * iT;
            v[0] += w[k] * f;
            v[1] += w[k] * g;
            v[2] += w[k] * g * (1.0 - 2.0 * f) * iT;
        }
        block_sum_k<3>(v, sc);
        v[0] -= dens;
    };
    HalleyState hs = halley_start(0.0, -8.0, 8.0, 60, 0.0, 0.0);
    while (hs.status == kRootRunning) {
        double v[3];
        density3(hs.x, v);
        halley_update(hs, v);
    }
    bool ok = hs.status == kRootConverged;
    double root = hs.root;
    int ne = hs.ne;
    if (!ok) {
        int nb = 0;
        auto density = [&](double mup) {
            double v[3];
            density3(mup, v);
            return v[0];
        };
        ok = bracket_root(density, 0.0, -8.0, 8.0, nullptr, 1.0e-16, &root, &nb);
        ne += nb;
    }
    if (threadIdx.x == 0) {
        out[0] = ok ? 1.0 : 0.0;
        out[1] = root;
        out[2] = ne;
        if (ok) *mu0v = root;
    }
}

// ================================================================== vertex V(r,tau)
// chi.cpp:85-119, block j = iqn index
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS, SlabOcc<S>::waves)
    k_vertex_slab(const double* __restrict__ chi, S* V, double normFact, double U, double Usp, double Uch,
                  const double* __restrict__ dsc)
{
    if (dsc) {
        Usp = dsc[kPscUsp + 1];
        Uch = dsc[kPscUch + 1];
    }
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    extern __shared__ double smem[];
    S* slab = (S*)smem;
    auto* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const long j = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double c = chi[i + j * nk8];
        const double csp = c / (1.0 - Usp * c / 2.0), cch = c / (1.0 + Uch * c / 2.0);
        put_sym_t(slab, nk, l, m, (S)(normFact * U * (3.0 * Usp * (csp - c) + Uch * (cch - c)) / 8.0));
    }
    __syncthreads();
    slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        V[i + j * nk8] = slab[m + l * nk];
    }
}

// chi.cpp:123-170: REDFT00 over n (length Nw/2+1) for two wedge points per block, as one
// complex FFT of length Nw of the even extensions; then V += fact * chi0(r,tau)
template <int L, typename S>
__global__ void __launch_bounds__(PEN_THREADS)
    k_vertex_pencil(S* V, const S* __restrict__ chirt, const double* __restrict__ chirt0, double* V0, long nk8,
                    double fact, const cd* __restrict__ tw, double U, const double* __restrict__ dsc)
{
    if (dsc) fact = vfact(U, dsc[kPscUsp + 1], dsc[kPscUch + 1], 8.0);
    constexpr int H = L / 2;
    __shared__ cd X[L];
    const long q = 2 * (long)xcd_remap(blockIdx.x, gridDim.x);
    if (q >= nk8) return;
    const bool two = (q + 1) < nk8;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = tid; j <= H; j += nt) {
        const double a = (double)V[q + (long)j * nk8];
        const double b = two ? (double)V[q + 1 + (long)j * nk8] : 0.0;
        X[j] = cd{a, b};
        if (j > 0 && j < H) X[L - j] = cd{a, b};
    }
    __syncthreads();
    block_fft<L, PEN_THREADS>(X, -1, tw);
    for (int j = tid; j <= H; j += nt) {
        const long o = q + (long)j * nk8;
        V[o] = (S)(X[j].x + fact * (double)chirt[o]);
        if (two) V[o + 1] = (S)(X[j].y + fact * (double)chirt[o + 1]);
    }
    if (tid == 0) {  // tau = 0 in FP64 for the boundary kernels
        V0[q] = X[0].x + fact * chirt0[q];
        if (two) V0[q + 1] = X[0].y + fact * chirt0[q + 1];
    }
}

// ================================================================== self-energy
__device__ __forceinline__ void store_self(cd* Self, long o, cd R, double, double, double, double) { Self[o] = R; }
__device__ __forceinline__ void store_self(cf* Self, long o, cd R, double Sinf, double s2, double s3, double wn)
{
    const cd v = sigma_mom(Sinf, s2, s3, R, wn);
    Self[o] = cf{(float)v.x, (float)v.y};
}

// Sigma_inf = -(Sigma(k=0,0) + Sigma(k=0,beta)) (green.cpp:1996); one thread
// Boundary data of Sigma(k,tau) in FP64, block b: side = b % 2 (0: tau = 0, 1: tau = beta),
// pass = b / 2: the slope dSigma/dtau (pass 0, green.cpp:1743-1817) or the value
// Sigma(k,tau_b) = DCT[G0(r,tau_b) V(r,0)] (pass 1, green.cpp:1822-1892 at j = 0, Nw;
// V(r,beta) = V(r,0)).  Each block transforms G0(k,tau_b) to r itself (the two passes of a
// side run side by side: two transforms in a row per block instead of three).
// G0(k,0+) = -1/(e^{-x/T}+1) and G0(k,beta-) = -1/(e^{x/T}+1), both over N^2.
// tmp: staging, nk8 per block.
template <int N>
__global__ void __launch_bounds__(SLAB_THREADS)
    k_sigma_bnd(const double* __restrict__ ek, const double* __restrict__ er, const double* __restrict__ V0,
                const double* __restrict__ dchirt0, const double* __restrict__ dGrt0p, double mu0, double T,
                double fact, double* dSig, double* Sb, double* tmp, const double* __restrict__ mu0p, double U,
                const double* __restrict__ dsc)
{
    if (mu0p) mu0 = *mu0p;
    if (dsc) fact = vfact(U, dsc[kPscUsp + 1], dsc[kPscUch + 1], 8.0);
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int side = blockIdx.x & 1, pass = blockIdx.x >> 1;
    double* stg = tmp + blockIdx.x * nk8;
    for (int s = tid; s < nk * nk; s += nt) {
        const double x = ek[s] - mu0;
        slab[s] = (side == 0) ? -1.0 / (exp(-x / T) + 1.0) / nk02 : -1.0 / (exp(x / T) + 1.0) / nk02;
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const double g = slab[m + l * nk];
        const double v = V0[i];
        if (pass == 1)
            stg[i] = g * v;
        else if (side == 0)
            stg[i] = fact * dchirt0[i] * g + v * dGrt0p[i];
        else {
            const double xi = (i == 0) ? -mu0 : er[i];
            stg[i] = -fact * dchirt0[i] * g + v * (-dGrt0p[i] + xi);
        }
    }
    __syncthreads();
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        put_sym(slab, nk, l, m, stg[i]);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 0, 0>(slab, nk, nk, nk, tw);
    double* o = (pass == 0) ? dSig : Sb;
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        o[i + side * nk8] = slab[m + l * nk];
    }
}

// Sigma(k,tau_j), j = 0..Nw (green.cpp:1822-1892): G0(k,tau) -> r, x V(r,|tau|), -> k
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS, SlabOcc<S>::waves)
    k_sigma_slab(const double* __restrict__ ek, const S* __restrict__ V, double mu0, double T, int Nw, S* Stau,
                 const double* __restrict__ mu0p)
{
    if (mu0p) mu0 = *mu0p;
    constexpr int nk = N / 2 + 1;
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = (double)N * N;
    extern __shared__ double smem[];
    S* slab = (S*)smem;
    auto* tw = slab_tw(slab, (long)nk * nk);
    fill_twiddles(tw, N);
    const int j = blockIdx.x;
    const int nwb = Nw / 2 + 1;
    const int jv = (j < nwb) ? j : Nw - j;
    const double t = j / (Nw * T);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int s = tid; s < nk * nk; s += nt) {
        const double x = ek[s] - mu0;
        slab[s] = (S)((x > 0) ? -exp(-t * x) / (exp(-x / T) + 1.0) / nk02
                              : -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02);
    }
    __syncthreads();
    slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
    const S* Vj = V + (long)jv * nk8;
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        const S v = slab[m + l * nk] * Vj[i];
        put_sym_t(slab, nk, l, m, v);
    }
    __syncthreads();
    slab_r2r_2d<S, N, 0, 0>(slab, nk, nk, nk, tw);
    S* o = Stau + (long)j * nk8;
    for (long i = tid; i < nk8; i += nt) {
        int l, m;
        wedge_lm(i, l, m);
        o[i] = slab[m + l * nk];
    }
}

// green.cpp:1925-2018: per wedge k, spline of Sigma(k,tau_j) (j=0..Nw) with slopes dSig,
// F = BACKWARD DFT[exp(i pi j/Nw) 6a_j], R(k,n) = (1 - exp(i pi (2n+1)/Nw)) F_n
// threads of the self-energy pencil: 512 for the long transforms of the mixed path, which
// stores the assembled FP32 Sigma(k,iwn) (at Nw = 2048: 0.59 against 0.63 ms at 256); 256 in
// FP64, which stores the remainder (0.48 against 0.57 ms at 512)
template <typename SelfT>
constexpr int sigma_pen_threads(int L) { return (L >= 1024 && sizeof(SelfT) == sizeof(cf)) ? 512 : PEN_THREADS; }
template <int L, typename S, typename SelfT, int NT = sigma_pen_threads<SelfT>(L)>
__global__ void __launch_bounds__(NT)
    k_sigma_pencil(const S* __restrict__ Stau, const double* __restrict__ Sb, const double* __restrict__ dSig,
                   long nk8, double T, SelfT* Self,
                   double* Sinf2, double* Sinf3, double* Sinf4, double* Sinfp,
                   const cd* __restrict__ tw, const cd* __restrict__ twX, const cd* __restrict__ twS)
{
    __shared__ double y[L + 1];  // the samples, then in place their spline moments
    __shared__ cd X[L];
    __shared__ double ES[2 * NT + 8];
    const long k = xcd_remap(blockIdx.x, gridDim.x);
    if (k >= nk8) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = tid; j <= L; j += nt)
        y[j] = (j == 0) ? Sb[k] : (j == L) ? Sb[nk8 + k] : (double)Stau[k + (long)j * nk8];
    const double d0 = dSig[k], dN = dSig[k + nk8];
    const double h = 1.0 / (L * T);
    __syncthreads();
    {
        double* const ys[1] = {y};
        block_spline_fixed<L, NT, 1>(ys, h, &d0, &dN, ES, nullptr);
    }
    const double* M = y;
    for (int j = tid; j < L; j += nt) {
        const double a6 = (M[j + 1] - M[j]) / h;
        X[j] = twX[j] * a6;
    }
    __syncthreads();
    block_fft<L, NT>(X, +1, tw);
    // Sigma_inf = -(Sigma(k=0, 0+) + Sigma(k=0, beta-)) from the boundary values; stored by k = 0
    const double s2 = d0 + dN, s3 = -(M[0] + M[L]), Si = -(Sb[0] + Sb[nk8]);
    if (k == 0 && tid == 0) Sinfp[0] = Si;
    for (int n = tid; n < L / 2; n += nt) {
        const cd R = twS[n] * X[n];
        store_self(Self, k * (L / 2) + n, R, Si, s2, s3, (2 * n + 1) * kPi * T);
        if (n == L / 2 - 1) Sinf4[k] = R.x;
    }
    if (tid == 0) {
        Sinf2[k] = s2;
        Sinf3[k] = s3;
    }
}

// phase tables of the pencil transforms, evaluated from the same double arguments as the
// reference: twX[j] = exp(i j pi/L) (green.cpp:1957, cond_opt.cpp:3290); twS[j] = 1 -
// exp(i a_j pi/L) with a_j = 2j+1 (j < L/2) or 2(j-L)+1 (green.cpp:2006,
// cond_opt.cpp:3305-3310); twM[n] = 1 - exp(-2 n i pi/L) (cond_opt.cpp:3731,3805);
// ph[j] = exp(i (j pi (L-1))/L) (cond_opt.cpp:3635), ph[L] = exp(i pi (L-1)) (:3693)
__global__ void k_fill_tables(int L, cd* twX, cd* twS, cd* twM, cd* ph)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < L) {
        twX[j] = cis(j * kPi / L);
        const double a = (j < L / 2) ? (2 * j + 1) * kPi / L : (2 * (j - L) + 1) * kPi / L;
        const cd e = cis(a);
        twS[j] = cd{1.0 - e.x, -e.y};
        const cd em = cis(-2 * j * kPi / L);
        twM[j] = cd{1.0 - em.x, -em.y};
        ph[j] = cis((j * kPi * (L - 1)) / L);
    }
    if (j == L) ph[L] = cis(kPi * (L - 1));
}

// ================================================================== traces

// Trace reductions over (k, n).  Sigma(k,iwn) is stored point-major, Self[k*(Nw/2) + n], so
// one wavefront takes one wedge point and its lanes walk the frequencies (coalesced); a block
// of 256 threads covers 4 wedge points and writes one partial per value, summed afterwards
// in a fixed order (k_finalize, k_finalize_k).
constexpr int kTraceK = 4;  // wedge points per trace block

// traceDiffG(pmu) partials (green.cpp:3108-3182), n = 0..Nw
template <typename SelfT>
__global__ void __launch_bounds__(256)
    k_trace_diffG(const double* __restrict__ ekw, const double* __restrict__ wq, const SelfT* __restrict__ Self,
                  const double* __restrict__ Sinf2, const double* __restrict__ Sinf3,
                  const double* __restrict__ Sinf4, const double* __restrict__ Sinf, int nk8, int Nw, double T,
                  double mu0, double pmu, double* part)
{
    __shared__ double sc[16];
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * kTraceK + (threadIdx.x >> 6);
    const int H = Nw / 2;
    double acc = 0.0;
    if (q < nk8) {
        const double Si = Sinf[0];
        const double e = ekw[q], w = 2.0 * wq[q];
        const double s2 = Sinf2[q], s3 = Sinf3[q], s4 = Sinf4[q];
        const SelfT* Sq = Self + (long)q * H;
        for (int j = lane; j <= Nw; j += 64) {
            const double wn = (2 * j + 1) * kPi * T;
            double rS, iS;
            if (j >= H) {
                rS = -s2 / (wn * wn) + s4 / (wn * wn * wn * wn);
                iS = -Si / wn + s3 / (wn * wn * wn);
            } else {
                const cd s = sigma_kw(Sq, j, Si, s2, s3, wn);
                rS = s.x;
                iS = s.y;
            }
            const double tmp = w / (((wn - iS) * (wn - iS) + (e - pmu + rS) * (e - pmu + rS)) *
                                    (wn * wn + (e - mu0) * (e - mu0)));
            acc += tmp * ((mu0 - pmu + rS) * (e - pmu + rS) * (e - mu0) + iS * (iS - 2 * wn) * (e - mu0) -
                          wn * wn * (mu0 - pmu + rS));
        }
    }
    acc = block_sum(acc, sc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// traceDiffG and its first two derivatives with respect to pmu.  With a = e - pmu + Re Sigma,
// b = wn - Im Sigma: Re G = -a/(a^2+b^2), d/dpmu Re G = (b^2-a^2)/(a^2+b^2)^2,
// d2/dpmu2 Re G = (6ab^2 - 2a^3)/(a^2+b^2)^3.  Partials part[k * nblk + blk], k = 0, 1, 2.
// A lane sums its frequencies j = lane, lane + 64, ... in increasing j: the stored ones
// (j < Nw/2) in batches of kTD loads issued before any use, then the asymptotic ones.
constexpr int kTD = 4;
__device__ __forceinline__ void trace_diffG3_term(double wn, double rS, double iS, double e, double c, double w,
                                                  double mu0, double pmu, double acc[3])
{
    const double a = e - pmu + rS, b = wn - iS;
    const double a2 = a * a, b2 = b * b, r = rcp_nr(a2 + b2);
    const double r0 = rcp_nr(wn * wn + c * c);
    // Re G - Re G0 = [c (a^2+b^2) - a (c^2+wn^2)] r r0, written without cancellation
    acc[0] += w * r * r0 * ((mu0 - pmu + rS) * a * c + iS * (iS - 2 * wn) * c - wn * wn * (mu0 - pmu + rS));
    acc[1] += w * (b2 - a2) * r * r;
    acc[2] += w * (6.0 * a * b2 - 2.0 * a * a2) * r * r * r;
}
// pmu (and T, mu0) come from the root search st, so that a captured pass graph serves every
// point; a settled search leaves nothing to do.  ticket != null: the workgroup that finishes
// last takes the Halley step.
template <typename SelfT>
__global__ void __launch_bounds__(256)
    k_trace_diffG3(const double* __restrict__ ekw, const double* __restrict__ wq, const SelfT* __restrict__ Self,
                   const double* __restrict__ Sinf2, const double* __restrict__ Sinf3,
                   const double* __restrict__ Sinf4, const double* __restrict__ Sinf, int nk8, int Nw,
                   HalleyState* st, double* part, unsigned* ticket, double* mirror)
{
    __shared__ double sc[3 * 4];
    if (st->status != kRootRunning) return;
    const double pmu = st->x, T = st->T, mu0 = st->mu0;
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * kTraceK + (threadIdx.x >> 6);
    const int H = Nw / 2;
    double acc[3] = {0.0, 0.0, 0.0};
    if (q < nk8) {
        const double Si = Sinf[0];
        const double e = ekw[q], w = 2.0 * wq[q];
        const double s2 = Sinf2[q], s3 = Sinf3[q], s4 = Sinf4[q];
        const SelfT* Sq = Self + (long)q * H;
        const double c = e - mu0;
        int j = lane;
        for (; j + (kTD - 1) * 64 < H; j += kTD * 64) {
            SelfT v[kTD];
#pragma unroll
            for (int u = 0; u < kTD; u++) v[u] = Sq[j + u * 64];
#pragma unroll
            for (int u = 0; u < kTD; u++) {
                const double wn = (2 * (j + u * 64) + 1) * kPi * T;
                const cd s = sigma_kw(v + u, 0, Si, s2, s3, wn);
                trace_diffG3_term(wn, s.x, s.y, e, c, w, mu0, pmu, acc);
            }
        }
        for (; j < H; j += 64) {
            const double wn = (2 * j + 1) * kPi * T;
            const cd s = sigma_kw(Sq, j, Si, s2, s3, wn);
            trace_diffG3_term(wn, s.x, s.y, e, c, w, mu0, pmu, acc);
        }
        for (; j <= Nw; j += 64) {
            const double wn = (2 * j + 1) * kPi * T;
            const double iw = 1.0 / wn, iw2 = iw * iw;
            trace_diffG3_term(wn, (-s2 + s4 * iw2) * iw2, (-Si + s3 * iw2) * iw, e, c, w, mu0, pmu, acc);
        }
    }
    block_sum_k<3>(acc, sc);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[k * gridDim.x + blockIdx.x] = acc[k];
    if (ticket && last_workgroup(ticket, gridDim.x)) halley_step_block(part, gridDim.x, st, mirror, sc);
}

// which = 0: Tr Sigma G0 main sum (green.cpp:3018-3051); 1: its Sigma_inf correction
// (:3053-3084, the n < Nw/2 terms plus one per k); 2: Tr Sigma G (green.cpp:2944-2976)
template <typename SelfT>
__global__ void __launch_bounds__(256)
    k_trace_selfG(const double* __restrict__ ekw, const double* __restrict__ wq, const SelfT* __restrict__ Self,
                  const double* __restrict__ Sinf2, const double* __restrict__ Sinf3, const double* __restrict__ Sinf,
                  int nk8, int Nw, double T, double mu0, double mu, double nk02, int which, double* part)
{
    __shared__ double sc[16];
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * kTraceK + (threadIdx.x >> 6);
    const int H = Nw / 2;
    double acc = 0.0;
    if (q < nk8) {
        const double Si = Sinf[0];
        const double e = ekw[q], w = wq[q];
        const SelfT* Sq = Self + (long)q * H;
        for (int j = lane; j < H; j += 64) {
            const double wn = (2 * j + 1) * kPi * T;
            if (which == 1) {
                acc += w * Si * T * 2 / (nk02 * (wn * wn + (e - mu0) * (e - mu0)));
                if (j == 0) {
                    if (fabs(e - mu0) > 1.0e-12)
                        acc += w * Si * (1.0 / (exp((e - mu0) / T) + 1.0) - 0.5) / (e - mu0) / nk02;
                    else {
                        const double ex = (e - mu0) < 0 ? exp((e - mu0) / T) : exp(-(e - mu0) / T);
                        acc -= w * Si * ex / ((ex + 1.0) * (ex + 1.0) * nk02 * T);
                    }
                }
                continue;
            }
            const cd s = sigma_kw(Sq, j, Si, Sinf2[q], Sinf3[q], wn);
            if (which == 0) {
                const double tmp = w * T * 2 / (nk02 * (wn * wn + (e - mu0) * (e - mu0)));
                acc += tmp * (s.x * (mu0 - e) + s.y * wn);
            } else {
                const double tmp = w * T * 2 / (nk02 * ((wn - s.y) * (wn - s.y) + (e - mu + s.x) * (e - mu + s.x)));
                acc += tmp * (s.y * wn - s.y * s.y - (e - mu) * s.x - s.x * s.x);
            }
        }
    }
    acc = block_sum(acc, sc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// column sums of a row-major (nrows x ncols) partial table in a fixed order: a block of
// 64 columns x 16 row groups, row group partials added in order
__device__ __forceinline__ void reduce_cols_block(const double* __restrict__ part, int nrows, int ncols, double* out,
                                                  int blk)
{
    __shared__ double t[16][65];
    const int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
    const int c = blk * 64 + cx;
    double a = 0.0;
    if (c < ncols)
        for (int r = ry; r < nrows; r += 16) a += part[(long)r * ncols + c];
    t[ry][cx] = a;
    __syncthreads();
    if (ry == 0 && c < ncols) {
        double sum = 0.0;
        for (int i = 0; i < 16; i++) sum += t[i][cx];
        out[c] = sum;
    }
}
__global__ void __launch_bounds__(1024) k_reduce_cols(const double* __restrict__ part, int nrows, int ncols,
                                                      double* out)
{
    reduce_cols_block(part, nrows, ncols, out, blockIdx.x);
}
// two tables of the same row count in one launch (blocks of the first, then of the second);
// block 0 also copies ncopy doubles from csrc to cdst (null: nothing)
__global__ void __launch_bounds__(1024) k_reduce_cols2(const double* __restrict__ p1, int ncols1, double* o1,
                                                       const double* __restrict__ p2, int ncols2, double* o2,
                                                       int nrows, const double* __restrict__ csrc, int ncopy,
                                                       double* cdst)
{
    const int b1 = (ncols1 + 63) / 64;
    if ((int)blockIdx.x < b1)
        reduce_cols_block(p1, nrows, ncols1, o1, blockIdx.x);
    else
        reduce_cols_block(p2, nrows, ncols2, o2, blockIdx.x - b1);
    if (csrc && blockIdx.x == 0)
        for (int i = threadIdx.x; i < ncopy; i += blockDim.x) cdst[i] = csrc[i];
}

// dispersion tables on the device (hamiltonien.cpp:49-68,125-155; green0 wedge weights)
__global__ void k_model(int N, double t, double tp, double tpp, double t3, double* ek, double* ekw, double* wq,
                        double* er, double* d2ek, double* dek)
{
    const int nk = N / 2 + 1;
    const long nk8 = (long)nk * (nk + 1) / 2, nq = (long)nk * (nk - 2);
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const double dk = kPi / (nk - 1);
    if (i < (long)nk * nk) {
        const int l = (int)(i / nk), m = (int)(i % nk);
        const double kx = l * dk, ky = m * dk;
        const double cx = cos(kx), cy = cos(ky), c2x = cos(2 * kx), c2y = cos(2 * ky);
        ek[i] = -2.0 * t * (cx + cy) - 4.0 * tp * cx * cy - 2.0 * tpp * (c2x + c2y) - 4 * t3 * (c2x * cy + c2y * cx);
        d2ek[i] = 2.0 * t * cx + 4.0 * tp * cx * cy + 8.0 * tpp * c2x + 4 * t3 * (4.0 * c2x * cy + c2y * cx);
        if (m <= l) {
            const long w = widx(l, m);
            ekw[w] = ek[i];
            wq[w] = wedge_weight(l, m, nk);
            er[w] = (w == 1) ? -t : (w == 2) ? -tp : (w == 3) ? -tpp : (w == 4) ? -t3 : 0.0;
        }
    }
    if (i < nq) {
        const int l = (int)(i / nk), m = (int)(i % nk);
        const double kx = (l + 1) * dk, ky = m * dk;
        dek[i] = 2.0 * t * sin(kx) + 4.0 * tp * sin(kx) * cos(ky) + 4.0 * tpp * sin(2 * kx) +
                 4 * t3 * (2.0 * sin(2 * kx) * cos(ky) + cos(2 * ky) * sin(kx));
    }
    (void)nk8;
}

// ================================================================== Lambda_jj slabs
// h(r,tau_j) and f0(r,tau_j), j = 0..Nw (cond_opt.cpp:2697-2735, 2795-2827):
// RODFT00 along kx (index l+1), REDFT00 along ky.
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS, SlabOcc<S>::waves)
    k_hf0_slab(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ mu0p,
               double T, int Nw, S* h, S* f0)
{
    const double mu0 = *mu0p;
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    S* slab = (S*)smem;
    auto* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int j = blockIdx.x;
    const double t = (double)j / (T * Nw);
    const int tid = threadIdx.x, nt = blockDim.x;
    S* const oh = h + (long)j * nq;
    S* const of = f0 + (long)j * nq;
    // h and f0 together (hf0_pair); f0 waits in its output row, read back by the thread that
    // wrote it, while h is transformed
    for (int s = tid; s < (int)nq; s += nt) {
        const int l = s / nk, m = s % nk;
        double hv, fv;
        hf0_pair(ek[m + (l + 1) * nk], dek[s], mu0, T, t, nf, hv, fv);
        slab[s] = (S)hv;
        of[s] = (S)fv;
    }
    __syncthreads();
    slab_r2r_2d<S, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    for (int s = tid; s < (int)nq; s += nt) oh[s] = slab[s];
    __syncthreads();
    for (int s = tid; s < (int)nq; s += nt) slab[s] = of[s];
    __syncthreads();
    slab_r2r_2d<S, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    for (int s = tid; s < (int)nq; s += nt) of[s] = slab[s];
}

// Boundary data of hV and f0V in FP64 (cond_opt.cpp:2697-2735, 2745-3114 at tau = 0, beta).
// Block b: which = (b/2)%2 (0: h, 1: f0), side = b%2 (0: tau = 0, 1: tau = beta), part = b/4:
//   (1) h(r,tau_b) (or f0) by DST x DCT of the k-space integrand at t = 0 or beta
//   part 1: (2) the tau-derivative integrand -> r, combined with V(r,0), dchi0(r,0) -> k:
//       slopes dhV / df0V (cond_opt.cpp:2745-3014)
//   part 0: (3) 2 h(r,tau_b) V(r,0) -> k: hVb / f0Vb, the tau = 0, beta values of hV / f0V
// Both parts of a (which, side) run (1) themselves and then side by side: three transforms in
// a row per block instead of four.  Independent of the tau slabs.  bscr: staging, nq per
// (which, side).
template <int N>
__global__ void __launch_bounds__(SLAB_THREADS)
    k_hf0_bnd(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ V0,
              const double* __restrict__ dchirt0, double mu0, double T, int Nw, double fact4, double* dhV,
              double* df0V, double* hVb, double* f0Vb, double* bscr, const double* __restrict__ mu0p, double U,
              const double* __restrict__ dsc)
{
    if (mu0p) mu0 = *mu0p;
    if (dsc) fact4 = vfact(U, dsc[kPscUsp + 1], dsc[kPscUch + 1], 4.0);
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    const double nf = 1.0 / ((double)N * N);
    extern __shared__ double smem[];
    double* slab = smem;
    cd* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int which = (blockIdx.x >> 1) & 1, side = blockIdx.x & 1, part = blockIdx.x >> 2;
    const int tid = threadIdx.x, nt = blockDim.x;
    double* hr = bscr + (blockIdx.x & 3) * nq;
    const double tb = (side == 0) ? 0.0 : (double)Nw / (T * Nw);  // tau_j of k_hf0_slab at j = 0, Nw
    // (1)
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        slab[s] = hf0_val(which, ek[m + (l + 1) * nk], dek[s], mu0, T, tb, nf);
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    if (part == 0) {
        // (3), from h(r,tau_b) in place
        for (long s = tid; s < nq; s += nt) {
            const int l = (int)(s / nk), m = (int)(s % nk);
            slab[s] = 2.0 * slab[s] * V0[fold_idx(l + 1, m, N)];
        }
        __syncthreads();
        slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
        double* ob = ((which == 0) ? hVb : f0Vb) + side * nq;
        for (long s = tid; s < nq; s += nt) ob[s] = slab[s];
        return;
    }
    for (long s = tid; s < nq; s += nt) hr[s] = slab[s];
    __syncthreads();
    // (2)
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        const double e = ek[m + (l + 1) * nk], v = dek[s];
        double val;
        if (which == 0) {
            val = h_slope_val(side, e, v, mu0, T, nf);
        } else {
            if (e > mu0) {
                const double ex = exp((mu0 - e) / T);
                val = (side == 0)
                          ? nf * v * ((e - mu0) * ex / (T * (ex + 1.0) * (ex + 1.0)) + 1.0 / (ex + 1.0))
                          : nf * v * ex * ((mu0 - e) * (1.0 / (T * (ex + 1.0)) - ex / (T * (ex + 1.0) * (ex + 1.0))) + 1.0 / (ex + 1.0));
            } else {
                const double ex = exp((e - mu0) / T);
                val = (side == 0)
                          ? nf * v * ex * ((e - mu0) / (T * (ex + 1.0) * (ex + 1.0)) + 1.0 / (ex + 1.0))
                          : nf * v * ((mu0 - e) * (1.0 / (T * (ex + 1.0)) - 1.0 / (T * (ex + 1.0) * (ex + 1.0))) + 1.0 / (ex + 1.0));
            }
        }
        slab[s] = val;
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    for (long s = tid; s < nq; s += nt) {
        const int l = (int)(s / nk), m = (int)(s % nk);
        const long r = fold_idx(l + 1, m, N);
        const double d = slab[s], v = V0[r], hs = hr[s];
        slab[s] = (side == 0) ? 2.0 * d * v + hs * fact4 * dchirt0[r] : 2.0 * d * v - hs * fact4 * dchirt0[r];
    }
    __syncthreads();
    slab_r2r_2d<double, N, 1, 0>(slab, nk - 2, nk, nk, tw);
    double* o = ((which == 0) ? dhV : df0V) + side * nq;
    for (long s = tid; s < nq; s += nt) o[s] = slab[s];
}

// hV(k,tau) = DST x DCT [2 h(r,tau) V(r,|tau|)], same for f0, in place (cond_opt.cpp:3029-3114)
template <int N, typename S>
__global__ void __launch_bounds__(SLAB_THREADS, SlabOcc<S>::waves)
    k_hVf0V_slab(const S* __restrict__ V, int Nw, S* h, S* f0)
{
    constexpr int nk = N / 2 + 1;
    constexpr long nq = (long)nk * (nk - 2);
    constexpr long nk8 = (long)nk * (nk + 1) / 2;
    extern __shared__ double smem[];
    S* slab = (S*)smem;
    auto* tw = slab_tw(slab, nq);
    fill_twiddles(tw, N);
    const int tid = threadIdx.x, nt = blockDim.x;
    // slabs j = blockIdx.x, + gridDim.x, ...: the launch may hold fewer workgroups than slabs
    for (int j = blockIdx.x; j <= Nw; j += gridDim.x) {
        const int jv = (j > Nw / 2) ? Nw - j : j;
        const S* Vj = V + (long)jv * nk8;
        for (int which = 0; which < 2; which++) {
            S* a = ((which == 0) ? h : f0) + (long)j * nq;
            for (long s = tid; s < nq; s += nt) {
                const int l = (int)(s / nk), m = (int)(s % nk);
                slab[s] = (S)2 * a[s] * Vj[fold_idx(l + 1, m, N)];
            }
            __syncthreads();
            slab_r2r_2d<S, N, 1, 0>(slab, nk - 2, nk, nk, tw);
            for (long s = tid; s < nq; s += nt) a[s] = slab[s];
            __syncthreads();
        }
    }
}

constexpr int kHVBlocks = 256;  // workgroups of the FP32 hV/f0V slab kernel

// ================================================================== Lambda_jj pencils
struct Poles {
    cd r1, r2, a1, a2, b1, b2, den1, den2;
};

// roots of z^2 + (mu - e) z - Sigma_inf with Newton polish (generique.cpp:1666-1726)
__device__ __forceinline__ void roots2(cd A, cd B, cd Cc, cd* roots)
{
    if (!(cabs2(A) > 1.0e-15)) return;
    const cd Rr = csqrtd(B * B - 4.0 * A * Cc);
    roots[0] = (-B - Rr) / (2.0 * A);
    roots[1] = (-B + Rr) / (2.0 * A);
    for (int j = 0; j < 2; j++) {
        cd x = roots[j];
        cd yv = A * x * x + B * x + Cc;
        cd p = 2.0 * A * x + B;
        const cd y0 = yv;
        int it = 0;
        while (cabs2(p) != 0.0 && it < 10 && cabs2(yv) > 1.0e-16) {
            const cd x2 = x, y2 = yv;
            const cd cc = yv - p * x;
            x = (-cc) / p;
            yv = A * x * x + B * x + Cc;
            p = 2.0 * A * x + B;
            if (cabs2(y2) < cabs2(yv)) {
                x = 0.5 * (x + x2);
                yv = A * x * x + B * x + Cc;
                p = 2.0 * A * x + B;
            }
            it++;
        }
        if (cabs2(yv) < cabs2(y0)) roots[j] = x;
    }
}

__device__ __forceinline__ Poles make_poles(double e, double mu, double Sinf, double CJ1, double T)
{
    Poles P;
    cd roots[2] = {cd{0, 0}, cd{0, 0}};
    roots2(cd{1.0, 0.0}, cd{-e + mu, 0.0}, cd{-Sinf, 0.0}, roots);
    P.r1 = roots[0];
    P.r2 = roots[1];
    P.a1 = P.r1 / (P.r1 - P.r2);
    P.a2 = P.r2 / (P.r2 - P.r1);
    P.b1 = CJ1 / (P.r1 - P.r2);
    P.b2 = -P.b1;
    P.den1 = (P.r1.x > 0 ? cexpd(-(P.r1 / T)) : cexpd(P.r1 / T)) + cd{1.0, 0.0};
    P.den2 = (P.r2.x > 0 ? cexpd(-(P.r2 / T)) : cexpd(P.r2 / T)) + cd{1.0, 0.0};
    return P;
}

// Dyson G at Matsubara slot s (n = s - Nw/2), with the conjugate Sigma for n < 0
// (cond_opt.cpp:3354-3407).  dyson_n: the stored frequency index of slot s.
__device__ __forceinline__ int dyson_n(int s, int H) { return (s < H) ? H - s - 1 : s - H; }
// from the stored value raw = Self[ki * H + dyson_n(s, H)]
template <typename SelfT>
__device__ __forceinline__ cd dyson_G_raw(SelfT raw, int s, int H, double C2, double C3, double Sinf, double e,
                                          double mu, double T, double& kn)
{
    kn = (2.0 * (s - H) + 1) * kPi * T;
    cd sg = sigma_kw(&raw, 0, Sinf, C2, C3, fabs(kn));
    if (s < H) sg = conjg(sg);
    const cd den = cd{0.0 - e, kn} + cd{mu, 0.0} - sg;
    return sdiv_rcp(T, den);
}
template <typename SelfT>
__device__ __forceinline__ cd dyson_G(const SelfT* __restrict__ Self, long ki, long nk8, int s, int H, double C2,
                                      double C3, double Sinf, double e, double mu, double T, double& kn)
{
    return dyson_G_raw(Self[ki * H + dyson_n(s, H)], s, H, C2, C3, Sinf, e, mu, T, kn);
}

// Two-pole asymptote data of every Lambda-grid point (cond_opt.cpp:3409-3467): the roots of
// z^2 + (mu - e) z - Sigma_inf, the residues, and the tau = 0+ / beta- pole sums used for the
// boundary slopes.  One thread per point.
struct JJPoles {
    Poles P;
    cd f1p, f1m, fj1, fj1b, f2p, f2m, fj2, fj2b;
    // e1(t) ej1(t) = C1 for all t (C1 = e^{-beta r1} if Re r1 > 0, else e^{beta r1}); used when
    // beta |Re r1| < 600, so that 1/e1 stays finite.  Same for the second pole.
    cd C1, C2;
    int rec1, rec2;
    cd a1d, a2d, b1d, b2d;  // a1/den1, a2/den2, b1/den1, b2/den2
    // the pencil's per-step factors exp(-dt r), exp(dt r) of each pole (dt = NT/(L T), its
    // threads' stride in tau), and the two-pole part of J at tau = beta
    cd s1, sj1, s2, sj2;
    double Jinf_beta;
};
static_assert(sizeof(JJPoles) <= 64 * sizeof(double), "DevArrays::poles holds 64 doubles per point");

// The two factors of pole r in the tau = t asymptote (cond_opt.cpp:3409-3467), both of modulus
// <= 1 on [0, beta]: e(t) = exp(-t r) (Re r > 0) or exp((beta - t) r), and its partner
// ej(t) = exp(r (t - beta)) or exp(r t), taken as C / e(t) when rec is set (e ej = C).
__device__ __forceinline__ cd pole_e(cd r, double t, double T)
{
    return (r.x > 0) ? cexpd(-t * r) : cexpd((1.0 / T - t) * r);
}
__device__ __forceinline__ cd pole_ej(cd r, double t, double T, int rec, cd C, cd e)
{
    if (rec) return C / e;
    return (r.x > 0) ? cexpd(r * (t - 1.0 / T)) : cexpd(r * t);
}
__device__ __forceinline__ bool pole_off(cd v)
{
    const double a = fabs(v.x) + fabs(v.y);
    return !(a >= 1.0e-280 && a <= 2.0);
}

__global__ void k_jj_poles(const double* __restrict__ ek, const double* __restrict__ hVb,
                           const double* __restrict__ Sinfp, int N, double T, double mu, double dt, JJPoles* out,
                           const double* __restrict__ dsc)
{
    if (dsc) mu = dsc[kPscMu + 1];
    const int nk = N / 2 + 1;
    const long nq = (long)nk * (nk - 2);
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int l = (int)(q / nk), m = (int)(q % nk);
    const double e = ek[m + (l + 1) * nk];
    const double CJ1 = -hVb[q] - hVb[nq + q];
    JJPoles o;
    const Poles P = make_poles(e, mu, Sinfp[0], CJ1, T);
    o.P = P;
    const double beta = 1.0 / T;
    o.C1 = (P.r1.x > 0) ? cexpd(-(beta * P.r1)) : cexpd(beta * P.r1);
    o.C2 = (P.r2.x > 0) ? cexpd(-(beta * P.r2)) : cexpd(beta * P.r2);
    o.rec1 = fabs(P.r1.x) * beta < 600.0;
    o.rec2 = fabs(P.r2.x) * beta < 600.0;
    o.a1d = P.a1 / P.den1;
    o.a2d = P.a2 / P.den2;
    o.b1d = P.b1 / P.den1;
    o.b2d = P.b2 / P.den2;
    o.s1 = cexpd(-dt * P.r1);
    o.sj1 = cexpd(dt * P.r1);
    o.s2 = cexpd(-dt * P.r2);
    o.sj2 = cexpd(dt * P.r2);
    {
        const cd ej1 = (P.r1.x > 0) ? cd{1.0, 0.0} : cexpd(P.r1 * beta);
        const cd ej2 = (P.r2.x > 0) ? cd{1.0, 0.0} : cexpd(P.r2 * beta);
        o.Jinf_beta = (P.b1 * ej1 / P.den1 + P.b2 * ej2 / P.den2).x;
    }
    if (P.r1.x > 0) {
        const cd ex = cexpd(-(P.r1 / T));
        o.f1p = P.a1 / P.den1;
        o.f1m = P.a1 * ex / P.den1;
        o.fj1 = P.r1 * P.b1 * ex / P.den1;
        o.fj1b = P.r1 * P.b1 / P.den1;
    } else {
        const cd ex = cexpd(P.r1 / T);
        o.f1p = P.a1 * ex / P.den1;
        o.f1m = P.a1 / P.den1;
        o.fj1 = P.r1 * P.b1 / P.den1;
        o.fj1b = P.r1 * P.b1 * ex / P.den1;
    }
    if (P.r2.x > 0) {
        const cd ex = cexpd(-(P.r2 / T));
        o.f2p = P.a2 / P.den2;
        o.f2m = P.a2 * ex / P.den2;
        o.fj2 = P.r2 * P.b2 * ex / P.den2;
        o.fj2b = P.r2 * P.b2 / P.den2;
    } else {
        const cd ex = cexpd(P.r2 / T);
        o.f2p = P.a2 * ex / P.den2;
        o.f2m = P.a2 / P.den2;
        o.fj2 = P.r2 * P.b2 / P.den2;
        o.fj2b = P.r2 * P.b2 * ex / P.den2;
    }
    out[q] = o;
}

// [j][q] (j < nj, q < nq) -> [q][j]: kTT x kTT tiles through LDS, both sides coalesced (a
// wave moves 64 consecutive elements of one row per access; 16 accesses per thread in flight).
// Rows of odd length (nq = nk (nk - 2), nj = Nw + 1) rule out vector accesses.  The Lambda
// pencils then read each tau-pencil as one contiguous row.  Two arrays per launch (grid z).
constexpr int kTT = 64;
template <typename S>
__global__ void __launch_bounds__(256) k_transpose(const S* __restrict__ in0, S* __restrict__ out0,
                                                   const S* __restrict__ in1, S* __restrict__ out1, long nq, int nj)
{
    const S* __restrict__ in = blockIdx.z ? in1 : in0;  // z: which of the two arrays
    S* __restrict__ out = blockIdx.z ? out1 : out0;
    __shared__ S t[kTT][kTT + 1];
    const long q0 = (long)blockIdx.x * kTT;
    const int j0 = blockIdx.y * kTT;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    {
        // all 16 loads issued before the first LDS write
        const long q = q0 + tx;
        S v[kTT / 4];
#pragma unroll
        for (int r = 0; r < kTT / 4; r++) {
            const int j = j0 + ty + 4 * r;
            v[r] = (j < nj && q < nq) ? in[(long)j * nq + q] : S(0);
        }
#pragma unroll
        for (int r = 0; r < kTT / 4; r++) t[ty + 4 * r][tx] = v[r];
    }
    __syncthreads();
    const int j = j0 + tx;
#pragma unroll
    for (int r = 0; r < kTT; r += 4) {
        const long q = q0 + ty + r;
        if (j < nj && q < nq) out[q * nj + j] = t[tx][ty + r];
    }
}

// Per-k Lambda_jj contributions (cond_opt.cpp:3179-3822).  One workgroup of NT threads per
// Lambda-grid point at a time; workgroup blk handles the points q = blk + kk*nblk,
// kk < jj_points(L), and accumulates into its own partial rows, one pair per angle bin b of the
// point (bins[q], cond_opt.cpp:3792-3793; bin 0 when bins is null):
// part[(blk*ntheta + b)*2*nwb + n] (bubble) and part[(blk*ntheta + b)*2*nwb + nwb + n]
// (Maki-Thompson).
//
// The six length-L transforms of the reference run as three complex FFTs of pairs of
// sequences whose transforms have a known symmetry:
//   (1) the f0V and hV spline moments 6a_j (real) with the phase exp(i pi j/L): for real a,
//       A(n) = sum_j a_j e^{2 pi i j (n + 1/2)/L} obeys A(L-1-n) = conj A(n), so the FFT F of
//       (a + i b) e^{i pi j/L} splits into A(n) = (F_n + conj F_{L-1-n})/2 and
//       B(n) = (F_n - conj F_{L-1-n})/(2i);
//   (2) G(ikn) - G2 and conj(J(ikn)) (both Hermitian in n -> -n-1): ph_j * FFT_-(.)_j is real
//       for each, so ph_j * FFT_-(Gn + i conj Jn)_j = G(tau_j) + i J(tau_j) before the two-pole
//       terms (FFT_+(Jn)_j / ph_j = conj(ph_j FFT_-(conj Jn)_j));
//   (3) the Maki-Thompson and bubble moments (real): A_n = (F_n + conj F_{L-n})/2,
//       B_n = (F_n - conj F_{L-n})/(2i).
// Imaginary parts of G(tau), J(tau) are zero in exact arithmetic and are not formed.
// LDS: the spline samples / moments (Y1, Y2, in place), which also hold the packed slot
// values W of step (2), and one transform buffer X.  hT, f0T are the hV, f0V pencils in
// point-major layout [q][j] (k_transpose); the pole data come from k_jj_poles.

#ifndef JJ_WAVES_PER_SIMD
#define JJ_WAVES_PER_SIMD 2  // two 256-thread workgroups per CU (74 KB LDS each), no register cap
#endif
// L = 4096 takes 139 KB of LDS: one workgroup per CU
template <int L, int NT, typename S, typename SelfT>
__global__ void __launch_bounds__(NT, (L > 2048 ? 1 : JJ_WAVES_PER_SIMD))
    k_jj_pencil(const double* __restrict__ ek, const double* __restrict__ dek, const double* __restrict__ d2ek,
                const S* __restrict__ hT, const S* __restrict__ f0T, const double* __restrict__ dhV,
                const double* __restrict__ df0V, const double* __restrict__ hVb, const double* __restrict__ f0Vb,
                const JJPoles* __restrict__ poles, const SelfT* __restrict__ Self, const double* __restrict__ Sinf2,
                const double* __restrict__ Sinf3, const double* __restrict__ Sinfp, int N, double T, double mu,
                const cd* __restrict__ tw,
                const cd* __restrict__ twX, const cd* __restrict__ twS, const cd* __restrict__ twM,
                const cd* __restrict__ phT, double* part, double* scal, const int* __restrict__ bins, int ntheta,
                const double* __restrict__ dsc)
{
    if (dsc) mu = dsc[kPscMu + 1];
    constexpr int H = L / 2;
    constexpr int NWB = H + 1;
    constexpr int CS = (L + 1 + NT - 1) / NT;  // spline chunk
    __shared__ __align__(16) double Yb[2 * (L + 1)];
    __shared__ cd X[L];
    __shared__ double sc[8 * (NT / 64)];
    __shared__ double ES[2 * (2 * NT + 8)];
    __shared__ double Jbeta_s;
    double* Y1 = Yb;
    double* Y2 = Yb + (L + 1);
    cd* W = (cd*)Yb;  // L complex values, used while Y1, Y2 are free
    const int nk = N / 2 + 1;
    const long nq = (long)nk * (nk - 2);
    const double nf = 1.0 / ((double)N * N);
    const double Sinf = Sinfp[0];
    const double hstep = 1.0 / (L * T), ihstep = L * T, iT = 1.0 / T;
    const int tid = threadIdx.x;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const long nblk = gridDim.x;
    double* const rows = part + (long)blk * 2 * NWB * ntheta;
    for (int n = tid; n < 2 * NWB * ntheta; n += NT) rows[n] = 0.0;  // the barrier below orders it
    double k0acc = 0.0, inf1 = 0.0, inf2 = 0.0, inf4 = 0.0;
    // The interior samples j = 1..L-1 of the f0V and hV pencils of the next point go from HBM
    // straight into LDS (global_load_lds, no registers) while the current point's last
    // accumulation runs: they are issued once Yb is free (after the last transform, which
    // uses it as its second buffer).  FP64 rows land in Y1 + 1 and Y2 + 1 themselves;
    // FP32 rows land in the upper half of Yb (stage0, stage1) and are widened at the top of
    // the next point.  Each wave copies 256-byte chunks c = wave, wave + NT/64, ... of a row,
    // one dword per lane.
    constexpr bool kWide = sizeof(S) == sizeof(double);
    char* const stage0 = kWide ? (char*)(Y1 + 1) : (char*)Y2;
    char* const stage1 = kWide ? (char*)(Y2 + 1) : (char*)Y2 + 4 * L;
    auto dma_rows = [&](long qq) {
        constexpr int NDW = (L - 1) * (int)sizeof(S) / 4;  // dwords per row
        // the LDS destination (M0) must be wave-uniform: the wave index is made provably so
        const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        const char* src0 = (const char*)(f0T + qq * (L + 1) + 1);
        const char* src1 = (const char*)(hT + qq * (L + 1) + 1);
        for (int c = wv; c * 64 < NDW; c += NT / 64) {
            const int i = c * 64 + lane;
            if (i < NDW) {
                __builtin_amdgcn_global_load_lds((const void*)(src0 + 4 * i), (void*)(stage0 + 256 * c), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(src1 + 4 * i), (void*)(stage1 + 256 * c), 4, 0, 0);
            }
        }
    };
    if (blk < nq) dma_rows(blk);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS DMA has landed (the barrier below covers the others)
    __syncthreads();
    // point kk of workgroup blk is q = blk + kk * nblk: the workgroups resident on one XCD
    // (consecutive blk) work on consecutive q at the same time and share the cache lines of
    // the tau-major hV / f0V rows and of Sigma
    constexpr int KP = jj_points(L);
    for (int kk = 0; kk < KP; kk++) {
        const long q = (long)blk + kk * nblk;
        if (q >= nq) break;
        const int l = (int)(q / nk), m = (int)(q % nk);
        const long ki = fold_idx(l + 1, m, N);
        const double e = ek[m + (l + 1) * nk];
        const double v = dek[q];
        const double wt = (m == 0 || m == nk - 1) ? 2.0 : 4.0;
        double* const c1 = rows + (long)(bins ? bins[q] : 0) * 2 * NWB;
        double* const c2 = c1 + NWB;
        // (1) f0V and hV pencils (interior samples already in LDS), their splines, one packed FFT
        if constexpr (kWide) {
            if (tid == 0) {
                Y1[0] = f0Vb[q];
                Y1[L] = f0Vb[nq + q];
                Y2[0] = hVb[q];
                Y2[L] = hVb[nq + q];
            }
        } else {
            const float* sf = (const float*)stage0;
            const float* sh = (const float*)stage1;
            float fa[CS], ha[CS];
#pragma unroll
            for (int c = 0; c < CS; c++) {
                const int j = tid + c * NT;
                fa[c] = (j > 0 && j < L) ? sf[j - 1] : 0.0f;
                ha[c] = (j > 0 && j < L) ? sh[j - 1] : 0.0f;
            }
            __syncthreads();  // the staged rows overlap Y2
#pragma unroll
            for (int c = 0; c < CS; c++) {
                const int j = tid + c * NT;
                if (j <= L) {
                    Y1[j] = (j == 0) ? f0Vb[q] : (j == L) ? f0Vb[nq + q] : (double)fa[c];
                    Y2[j] = (j == 0) ? hVb[q] : (j == L) ? hVb[nq + q] : (double)ha[c];
                }
            }
        }
        const double CJ1 = -hVb[q] - hVb[nq + q];
        __syncthreads();
        {
            double* const ys[2] = {Y1, Y2};
            const double d0s[2] = {df0V[q], dhV[q]}, dNs[2] = {df0V[q + nq], dhV[q + nq]};
            block_spline_fixed<L, NT, 2>(ys, hstep, d0s, dNs, ES, nullptr);
        }
        for (int j = tid; j < L; j += NT) {
            const double a = (Y1[j + 1] - Y1[j]) * ihstep, b = (Y2[j + 1] - Y2[j]) * ihstep;
            X[j] = twX[j] * cd{a, b};
        }
        const double cj03 = -(Y1[0] + Y1[L]), cj3 = -(Y2[0] + Y2[L]);
        __syncthreads();
        block_fft_pp<L, NT>(X, W, +1, tw);  // Y1, Y2 (W) are free: their last reads are above
        // (2) per Matsubara slot s (n = s - L/2): Dyson G, J0 (f0V), J (hV) -> W
        double r6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // qn0 (re, im), sum Gn (re, im), dGp, dJ0
        {
            const double C2 = Sinf2[ki], C3 = Sinf3[ki];
            const double cj01 = -f0Vb[q] - f0Vb[nq + q], cj02 = df0V[q] + df0V[q + nq];
            const double cj2 = dhV[q] + dhV[q + nq];
            // the global reads of slot s + NT (Sigma, twS) are issued before the work of slot s
            const SelfT* Sk = Self + ki * H;
            auto slot_jx = [&](int s) { return (s >= H) ? s - H : s + H; };  // transform index of slot s
            SelfT sv_next = Sk[dyson_n(tid, H)];
            cd ts_next = twS[slot_jx(tid)];
#pragma unroll 1
            for (int s = tid; s < L; s += NT) {
                const SelfT sv = sv_next;
                const cd ts = ts_next;
                if (s + NT < L) {
                    sv_next = Sk[dyson_n(s + NT, H)];
                    ts_next = twS[slot_jx(s + NT)];
                }
                const int jx = slot_jx(s);
                const cd Fj = X[jx], Fc = conjg(X[L - 1 - jx]);
                const cd Af = cd{0.5 * (Fj.x + Fc.x), 0.5 * (Fj.y + Fc.y)};
                const cd Bh = cd{0.5 * (Fj.y - Fc.y), -0.5 * (Fj.x - Fc.x)};
                const double knj = (jx < H) ? (2 * jx + 1) * kPi * T : (2 * (jx - L) + 1) * kPi * T;
                // -i c1/k - c2/k^2 + i c3/k^3 + a/k^4 with one reciprocal of k (cond_opt.cpp:3302-3311)
                const double ik = rcp_nr(knj), ik2 = ik * ik;
                const cd af = ts * Af, ah = ts * Bh;
                const cd J0 = cd{(-cj02 + af.x * ik2) * ik2, (-cj01 + (cj03 + af.y * ik) * ik2) * ik};
                const cd Jv = cd{(-cj2 + ah.x * ik2) * ik2, (-CJ1 + (cj3 + ah.y * ik) * ik2) * ik};
                double kn;
                const cd Gv = dyson_G_raw(sv, s, H, C2, C3, Sinf, e, mu, T, kn);
                const cd t = (Gv * Gv * J0) * iT;
                r6[0] += t.x;
                r6[1] += t.y;
                const double ikn = rcp_nr(kn);
                const cd G2 = sdiv_rcp(T, cd{-e + mu, kn + Sinf * ikn});
                const cd Jn = Gv * Jv - G2 * cd{0.0, -CJ1 * ikn};
                const cd Gn = Gv - G2;
                r6[2] += Gn.x;
                r6[3] += Gn.y;
                r6[4] += kn * Gn.y;
                r6[5] -= kn * Jn.y;
                W[s] = cd{Gn.x + Jn.y, Gn.y + Jn.x};  // Gn + i conj(Jn)
            }
        }
        // the six sums: wave sums parked in sc[k * nw + wave], added in the order of block_sum_k
        // where they are used (after the transform: fewer registers live across it)
        constexpr int nw = NT / 64;
        {
            const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
            for (int k = 0; k < 6; k++) r6[k] = wave_sum(r6[k]);
            if (lane == 0)
#pragma unroll
                for (int k = 0; k < 6; k++) sc[k * nw + wave] = r6[k];
        }
        __syncthreads();  // the W writes and the parked sums; X's last reads (step 2) are above
        auto sum6 = [&](int k) {
            double t = 0.0;
            for (int i = 0; i < nw; i++) t += sc[k * nw + i];
            return t;
        };
        // (3) two-pole asymptote: pole data from k_jj_poles
        const JJPoles& pp = poles[q];
        const Poles& P = pp.P;
        // (4) to imaginary time: X_j <- (G(tau_j), J(tau_j))
        block_fft_pp<L, NT>(W, X, -1, tw);
        if (tid == 0) Jbeta_s = (cd{W[0].y, 0.0} / phT[L]).x + pp.Jinf_beta;  // W[0].y = FFT_+(Jn)_0 (ph_0 = 1)
        {
            // the pole factors are geometric in j: from tau_j to tau_{j+NT} e_p gains
            // exp(-dt r_p) and ej_p exp(+dt r_p), dt = NT/(L T) (the step factors come with the
            // pole data); a factor that leaves [1e-280, 2] on the way (underflow of a factor
            // that grows again, overflow of a step) is evaluated directly again
            double t = tid / (L * T);
            cd e1 = pole_e(P.r1, t, T), e2 = pole_e(P.r2, t, T);
            cd ej1 = pole_ej(P.r1, t, T, pp.rec1, pp.C1, e1), ej2 = pole_ej(P.r2, t, T, pp.rec2, pp.C2, e2);
            const cd s1 = pp.s1, sj1 = pp.sj1, s2 = pp.s2, sj2 = pp.sj2;
            cd ph_next = phT[tid];
#pragma unroll 1
            for (int j = tid; j < L; j += NT) {
                const cd ph = ph_next;
                if (j + NT < L) ph_next = phT[j + NT];
                const double G0t = (-(pp.a1d * e1) - pp.a2d * e2).x;
                const double Jinf = (pp.b1d * ej1 + pp.b2d * ej2).x;
                const cd R = ph * W[j];
                X[j] = cd{R.x + G0t, R.y + Jinf};
                if (j + NT < L) {
                    t = (j + NT) / (L * T);
                    e1 = e1 * s1;
                    e2 = e2 * s2;
                    ej1 = ej1 * sj1;
                    ej2 = ej2 * sj2;
                    if (pole_off(e1)) e1 = pole_e(P.r1, t, T);
                    if (pole_off(e2)) e2 = pole_e(P.r2, t, T);
                    if (pole_off(ej1)) ej1 = pole_ej(P.r1, t, T, pp.rec1, pp.C1, e1);
                    if (pole_off(ej2)) ej2 = pole_ej(P.r2, t, T, pp.rec2, pp.C2, e2);
                }
            }
        }
        __syncthreads();
        const double G0 = X[0].x, J0t = X[0].y, Jbeta = Jbeta_s;
        // sums of this point with the pole parts
        const double k0r = sum6(2) + (pp.f1m + pp.f2m).x;
        const double s4 = sum6(4), s5 = sum6(5);
        const double dGp = s4 + (P.r1 * pp.f1p + P.r2 * pp.f2p).x;
        const double dGm = -s4 + (P.r1 * pp.f1m + P.r2 * pp.f2m).x;
        const double dJ0 = s5 + (pp.fj1 + pp.fj2).x;
        const double dJb = -s5 + (pp.fj1b + pp.fj2b).x;
        // (5) Maki-Thompson samples G(tau) J(-tau) -> Y1; bubble samples -G(tau) G(-tau) -> Y2 over
        // the whole of [0, beta].  They are symmetric about beta/2 (the product is the same in
        // both orders), so the clamped spline over [0, beta] with end slopes (dchi, -dchi) is, on
        // [0, beta/2], the reference's spline with slopes (dchi, 0), and its integral is twice
        // that of the half: the two splines run in one call.
        for (int j = tid; j < L; j += NT) {
            const cd gj = X[j];
            Y1[j] = gj.x * gj.y;
            Y2[j] = (j == 0) ? G0 * (G0 + 1.0) : -(gj.x * X[L - j].x);
        }
        if (tid == 0) {
            Y1[L] = (-G0 - 1.0) * Jbeta;
            Y2[L] = G0 * (G0 + 1.0);
        }
        __syncthreads();
        const double gd0 = dGp * J0t + G0 * dJ0;
        const double gdN = dGm * Jbeta - (G0 + 1.0) * dJb;
        const double MJ1 = Y1[0] - Y1[L];
        const double dchi = dGp * (G0 + 1.0) + G0 * dGm;
        double integ[2] = {0.0, 0.0};
        {
            double* const ys[2] = {Y1, Y2};
            const double d0s[2] = {gd0, dchi}, dNs[2] = {gdN, -dchi};
            block_spline_fixed<L, NT, 2>(ys, hstep, d0s, dNs, ES, integ);
        }
        // the bubble over [0, beta] is used by thread 0 only (n = 0): each wave parks its sum in
        // sc[6 nw + wave], which thread 0 adds in the accumulation below (after two barriers)
        {
            const double ws = wave_sum(integ[1]);
            if ((tid & 63) == 0) sc[6 * (NT / 64) + (tid >> 6)] = ws;
        }
        const double MJ3 = Y1[0] - Y1[L];
        // (6) packed transform of the two moment sequences
        for (int j = tid; j < L; j += NT) {
            const double a = (Y1[j + 1] - Y1[j]) * ihstep;
            const double b = (j < H) ? (Y2[j + 1] - Y2[j]) * ihstep : -(Y2[L - j] - Y2[L - 1 - j]) * ihstep;
            X[j] = cd{a, b};
        }
        __syncthreads();  // Y1, Y2 read for the last time; the bubble's wave sums are in sc
        block_fft_pp<L, NT>(X, W, -1, tw);
        if (kk + 1 < KP && q + nblk < nq) dma_rows(q + nblk);  // W is free again
        const double wv = nf * wt * v;
#pragma unroll 1
        for (int n = tid; n <= H; n += NT) {
            const cd Fn = X[n], Fc = conjg(X[(L - n) & (L - 1)]);
            double bub;
            if (n == 0) {
                bub = sum6(6);  // the bubble's wave sums
                c2[0] -= wv * sum6(0);
            } else {
                const double iw = 1.0 / (2 * n * kPi * T), iw2 = iw * iw;
                const cd tm = twM[n];
                const cd a = tm * cd{0.5 * (Fn.x + Fc.x), 0.5 * (Fn.y + Fc.y)};
                // -2 MJ1/wn + 2 MJ3/wn^3 + 2 a/wn^4, then / wn
                const double imJ = (-2.0 * MJ1 + (2.0 * MJ3 + 2.0 * a.y * iw) * iw2) * iw;
                c2[n] -= wv * imJ * iw;
                const cd b = tm * cd{0.5 * (Fn.y - Fc.y), -0.5 * (Fn.x - Fc.x)};
                bub = (-2.0 * dchi + b.x * iw2) * iw2;
            }
            c1[n] -= 2.0 * wv * v * bub;
        }
        k0acc -= 2.0 * wt * nf * d2ek[m + (l + 1) * nk] * k0r;
        inf1 += 4.0 * dchi * wv * v;
        inf2 += wv * 2.0 * MJ1;
        inf4 -= wv * 2.0 * MJ3;
        __builtin_amdgcn_s_waitcnt(0);  // the next point's rows have landed in LDS
        __syncthreads();
    }
    if (tid == 0) {
        scal[blk * 4 + 0] = k0acc;
        scal[blk * 4 + 1] = inf1;
        scal[blk * 4 + 2] = inf2;
        scal[blk * 4 + 3] = inf4;
    }
}

// kx = 0 and kx = pi rows of the diamagnetic term (cond_opt.cpp:3993-4173); block b:
// row l = (b < nk ? 0 : nk-1), m = b % nk
template <typename SelfT>
__global__ void k_kx0_rows(const double* __restrict__ ek, const double* __restrict__ d2ek, const SelfT* __restrict__ Self,
                           const double* __restrict__ Sinf2, const double* __restrict__ Sinf3,
                           const double* __restrict__ Sinfp, int N, int L, double T, double mu, double* out,
                           const double* __restrict__ dsc)
{
    if (dsc) mu = dsc[kPscMu + 1];
    __shared__ double sc[16];
    const int nk = N / 2 + 1;
    const long nk8 = (long)nk * (nk + 1) / 2;
    const int l = (blockIdx.x < nk) ? 0 : nk - 1, m = blockIdx.x % nk;
    const long ki = fold_idx(l, m, N);
    const double C2 = Sinf2[ki], C3 = Sinf3[ki], Sinf = Sinfp[0];
    const double e = ek[m + l * nk];
    const int H = L / 2;
    cd k0 = cd{0, 0};
    for (int s = threadIdx.x; s < L; s += blockDim.x) {
        double kn;
        const cd Gv = dyson_G(Self, ki, nk8, s, H, C2, C3, Sinf, e, mu, T, kn);
        const cd G0v = T / cd{-e + mu, kn + Sinf / kn};
        k0 = k0 + (Gv - G0v);
    }
    double kv[2] = {k0.x, k0.y};
    block_sum_k<2>(kv, sc);
    k0 = cd{kv[0], kv[1]};
    if (threadIdx.x == 0) {
        const Poles P = make_poles(e, mu, Sinf, 0.0, T);
        const cd f1m = (P.r1.x > 0) ? P.a1 * cexpd(-(P.r1 / T)) / P.den1 : P.a1 / P.den1;
        const cd f2m = (P.r2.x > 0) ? P.a2 * cexpd(-(P.r2 / T)) / P.den2 : P.a2 / P.den2;
        const double k0r = k0.x + (f1m + f2m).x;
        double wt = 2;
        if (m == 0 || m == nk - 1) wt = wt / 2;
        out[blockIdx.x] = -2.0 * wt / ((double)N * N) * d2ek[m + l * nk] * k0r;
    }
}

// ================================================================== local Green function
// G_loc(iw_j) = sum over the wedge of (w_k/N^2) / (iw_j - e_k + mu - Sigma(k,iw_j)), j < Nw/2
// (green.cpp:209-266).  Thread = frequency j (Self is point-major, so a warp reads one k's
// row contiguously); blockIdx.y = k chunk: k = y, y + KCH, ...; partials
// part[y*2H + 2j + (0 re, 1 im)], y < kGlocChunks, summed over y in a fixed order by
// k_reduce_cols.
template <typename SelfT>
__global__ void __launch_bounds__(256)
    k_gloc(const double* __restrict__ ekw, const double* __restrict__ wq, const SelfT* __restrict__ Self,
           const double* __restrict__ Sinf2, const double* __restrict__ Sinf3, const double* __restrict__ Sinfp,
           long nk8, int H, double T, double mu, double nk02, double* part)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= H) return;
    const double wn = (2 * j + 1) * kPi * T;
    const double Sinf = Sinfp[0];
    double re = 0.0, im = 0.0;
    for (long k = blockIdx.y; k < nk8; k += gridDim.y) {
        const cd sg = sigma_kw(Self, k * H + j, Sinf, Sinf2[k], Sinf3[k], wn);
        const cd g = (wq[k] / nk02) / cd{mu - ekw[k] - sg.x, wn - sg.y};
        re += g.x;
        im += g.y;
    }
    double* o = part + (long)blockIdx.y * 2 * H;
    o[2 * j] = re;
    o[2 * j + 1] = im;
}

// ================================================================== launchers
// LDS bytes of a slab kernel: the slab (nk x nk, or (nk-2) x nk for Lambda_jj) of elem-byte
// values, 16-byte aligned, then the N-point twiddle table of the same precision
static size_t slab_lds(int N, bool lambda, size_t elem)
{
    const int nk = N / 2 + 1;
    const size_t n = lambda ? (size_t)nk * (nk - 2) : (size_t)nk * nk;
    return ((n * elem + 15) & ~(size_t)15) + 2 * elem * (size_t)N;
}

// lattice sizes with compiled slab transforms: 2^k (4..256) and 3*2^k (12..192)
bool supported_N(int N)
{
    switch (N) {
        case 4: case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
            return true;
        default:
            return false;
    }
}
bool supported_Nw(int Nw)
{
    return Nw == 16 || Nw == 32 || Nw == 64 || Nw == 128 || Nw == 256 || Nw == 512 || Nw == 1024 || Nw == 2048 ||
           Nw == 4096;
}

#define DISPATCH_N(N, ...)                          \
    switch (N) {                                    \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;     \
        case 8: { constexpr int NN = 8; __VA_ARGS__; } break;     \
        case 12: { constexpr int NN = 12; __VA_ARGS__; } break;   \
        case 24: { constexpr int NN = 24; __VA_ARGS__; } break;   \
        case 48: { constexpr int NN = 48; __VA_ARGS__; } break;   \
        case 96: { constexpr int NN = 96; __VA_ARGS__; } break;   \
        case 192: { constexpr int NN = 192; __VA_ARGS__; } break; \
        case 16: { constexpr int NN = 16; __VA_ARGS__; } break;   \
        case 32: { constexpr int NN = 32; __VA_ARGS__; } break;   \
        case 64: { constexpr int NN = 64; __VA_ARGS__; } break;   \
        case 128: { constexpr int NN = 128; __VA_ARGS__; } break; \
        case 256: { constexpr int NN = 256; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;       \
    }
#define DISPATCH_L(Lv, ...)                           \
    switch (Lv) {                                     \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;     \
        case 32: { constexpr int LL = 32; __VA_ARGS__; } break;     \
        case 64: { constexpr int LL = 64; __VA_ARGS__; } break;     \
        case 128: { constexpr int LL = 128; __VA_ARGS__; } break;   \
        case 256: { constexpr int LL = 256; __VA_ARGS__; } break;   \
        case 512: { constexpr int LL = 512; __VA_ARGS__; } break;   \
        case 1024: { constexpr int LL = 1024; __VA_ARGS__; } break; \
        case 2048: { constexpr int LL = 2048; __VA_ARGS__; } break; \
        case 4096: { constexpr int LL = 4096; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;         \
    }

#define DISPATCH_LJJ(Lv, ...)                         \
    switch (Lv) {                                     \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;     \
        case 32: { constexpr int LL = 32; __VA_ARGS__; } break;     \
        case 64: { constexpr int LL = 64; __VA_ARGS__; } break;     \
        case 128: { constexpr int LL = 128; __VA_ARGS__; } break;   \
        case 256: { constexpr int LL = 256; __VA_ARGS__; } break;   \
        case 512: { constexpr int LL = 512; __VA_ARGS__; } break;   \
        case 1024: { constexpr int LL = 1024; __VA_ARGS__; } break; \
        case 2048: { constexpr int LL = 2048; __VA_ARGS__; } break; \
        case 4096: { constexpr int LL = 4096; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;         \
    }

// SS: storage / slab type of the tau and iqn arrays; ST: stored form of Sigma(k,iwn)
#define DISPATCH_S(prec, ...)     \
    if (prec) {                   \
        using SS = float;         \
        using ST = cf;            \
        __VA_ARGS__;              \
    } else {                      \
        using SS = double;        \
        using ST = cd;            \
        __VA_ARGS__;              \
    }

// the dynamic-LDS limit of a kernel, raised once to the largest size asked for (the attribute
// holds for the process; each host thread keeps its own record of what it has set)
template <typename K>
static void set_lds(K kern, size_t bytes)
{
    thread_local std::unordered_map<const void*, size_t> done;
    const void* f = reinterpret_cast<const void*>(kern);
    size_t& have = done[f];
    if (have >= bytes) return;
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    have = bytes;
}

hipError_t launch_fill_tw(const DevArrays& d, int L, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_tw, dim3((L + 255) / 256), dim3(256), 0, s, d.twN, L);
    hipLaunchKernelGGL(k_fill_tables, dim3((L + 256) / 256), dim3(256), 0, s, L, d.twX, d.twS, d.twM, d.ph);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* part, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, part, n, out);
    return hipGetLastError();
}

hipError_t launch_chi0_tau0(const Grid& g, const DevArrays& d, hipStream_t s)
{
    const size_t lds = slab_lds(g.N, false, sizeof(double));
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        set_lds(k_chi0_tau0<NN, SS>, lds);
        hipLaunchKernelGGL((k_chi0_tau0<NN, SS>), dim3(1), dim3(SLAB_THREADS), lds, s, d.ekw, d.er, d.mu0v, g.T,
                           (SS*)d.chi, (SS*)d.chirt, d.dchirt0, d.dGrt0p, d.dchiqt0, d.chirt0, d.tmp);
    }));
    return hipGetLastError();
}

hipError_t launch_chi0_slab(const Grid& g, const DevArrays& d, hipStream_t s)
{
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        const size_t lds = slab_lds(g.N, false, sizeof(SS));
        set_lds(k_chi0_slab<NN, SS>, lds);
        hipLaunchKernelGGL((k_chi0_slab<NN, SS>), dim3(g.nwb - 1), dim3(SLAB_THREADS), lds, s, d.ekw, d.mu0v, g.T, g.Nw,
                           (SS*)d.chi, (SS*)d.chirt);
    }));
    return hipGetLastError();
}

hipError_t launch_chi0_pencil(const Grid& g, const DevArrays& d, hipStream_t s)
{
    DISPATCH_S(g.prec, DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL((k_chi0_pencil<LL, SS>), dim3((unsigned)g.nk8), dim3(PEN_THREADS), 0, s,
                           (const SS*)d.chi, d.chiw, d.dchiqt0, d.chi0inf, d.chi0inf2, g.nk8, g.T, d.twN);
    }));
    return hipGetLastError();
}

hipError_t launch_chi0max(const Grid& g, const DevArrays& d, hipStream_t s, double* out2)
{
    hipLaunchKernelGGL(k_chi0max, dim3(1), dim3(1024), 0, s, d.chiw, g.nk8, d.red, d.imax, out2);
    return hipGetLastError();
}

static dim3 red_grid(const Grid& g) { return dim3((unsigned)((g.nk8 + 255) / 256), kRowGroups); }
static dim3 trace_grid(const Grid& g) { return dim3((unsigned)((g.nk8 + kTraceK - 1) / kTraceK)); }

hipError_t launch_sum_rule(const Grid& g, const DevArrays& d, double sU, double* out, hipStream_t s)
{
    int np = 0;
    const hipError_t e = launch_sum_rule_part(g.N, g.Nw, g.T, d.chiw, d.chi0inf, d.chi0inf2, d.wq, sU, d.part, &np, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, d.part, np, out);
    return hipGetLastError();
}

int sum_rule_parts(int N) { return (int)((((long)(N / 2 + 1) * (N / 2 + 2) / 2) + 255) / 256) * kRowGroups; }

hipError_t launch_sum_rule_part(int N, int Nw, double T, const double* chiw, const double* inf1, const double* inf2,
                                const double* wq, double sU, double* part, int* nparts, hipStream_t s)
{
    const int nk = N / 2 + 1;
    const long nk8 = (long)nk * (nk + 1) / 2;
    const double Nq = 4.0 * (nk - 1) * (nk - 1);
    const dim3 gr((unsigned)((nk8 + 255) / 256), kRowGroups);
    hipLaunchKernelGGL(k_sum_rule, gr, dim3(256), 0, s, chiw, inf1, inf2, wq, (int)nk8, Nw / 2, T, Nq, sU, part);
    *nparts = (int)(gr.x * gr.y);
    return hipGetLastError();
}

hipError_t launch_halley_init(HalleyState* st, const HalleyState& s0, const double* chi0max, double* mirror,
                              hipStream_t s, const double* mu0p, const double* a0src, double a0A, double nsr,
                              double U)
{
    hipLaunchKernelGGL(k_halley_init, dim3(1), dim3(1), 0, s, st, s0, chi0max, mirror, mu0p, a0src, a0A, nsr, U);
    return hipGetLastError();
}

hipError_t launch_sum_rule3_pass(const Grid& g, const DevArrays& d, double sgn, double* mirror, hipStream_t s)
{
    const double Nq = 4.0 * (g.nk - 1) * (g.nk - 1);
    const dim3 gr = red_grid(g);
    const int nblk = (int)(gr.x * gr.y);
    const bool fused = nblk <= kFusedStepMax;
    const bool series = d.srm && g.nk8 * g.nwb >= kSRMinElems;  // the moments exist (launch_sr_moments)
    hipLaunchKernelGGL(k_sum_rule3, gr, dim3(256), 0, s, d.chiw, d.chi0inf, d.chi0inf2, d.wq, (int)g.nk8, g.Nw / 2,
                       Nq, sgn, d.root, d.part, fused ? d.ticket : nullptr, mirror, series ? d.srm : nullptr,
                       series ? d.srmax : nullptr);
    if (!fused) hipLaunchKernelGGL(k_halley_step, dim3(1), dim3(256), 0, s, d.part, nblk, d.root, mirror);
    return hipGetLastError();
}

hipError_t launch_sr_moments(const Grid& g, const DevArrays& d, hipStream_t s)
{
    if (!d.srm || g.nk8 * g.nwb < kSRMinElems) return hipSuccess;
    const double Nq = 4.0 * (g.nk - 1) * (g.nk - 1);
    const int nrows = g.Nw + 1;  // 2 nwmax + 1
    if (nrows > 8192) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sr_moments, dim3(nrows), dim3(256), 0, s, d.chiw, d.chi0inf, d.chi0inf2, d.wq, (int)g.nk8,
                       g.Nw / 2, Nq, g.T, d.srm, d.srmax);
    hipLaunchKernelGGL(k_sr_suffix, dim3(kSRM + 1), dim3(1024), 0, s, d.srm, d.srmax, nrows);
    return hipGetLastError();
}

hipError_t launch_model(const Grid& g, const DevArrays& d, double t, double tp, double tpp, double t3,
                        hipStream_t s)
{
    const long n = std::max<long>((long)g.nk * g.nk, g.nq);
    hipLaunchKernelGGL(k_model, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.N, t, tp, tpp, t3, d.ek, d.ekw,
                       d.wq, d.er, d.d2ek, d.dek);
    return hipGetLastError();
}

hipError_t launch_reduce_cols(const double* part, int nrows, int ncols, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_cols, dim3((ncols + 63) / 64), dim3(1024), 0, s, part, nrows, ncols, out);
    return hipGetLastError();
}

hipError_t launch_reduce_cols2(const double* p1, int ncols1, double* o1, const double* p2, int ncols2, double* o2,
                               int nrows, hipStream_t s, const double* csrc, int ncopy, double* cdst)
{
    const unsigned nb = (unsigned)((ncols1 + 63) / 64 + (ncols2 + 63) / 64);
    hipLaunchKernelGGL(k_reduce_cols2, dim3(nb), dim3(1024), 0, s, p1, ncols1, o1, p2, ncols2, o2, nrows, csrc, ncopy,
                       cdst);
    return hipGetLastError();
}

hipError_t launch_mu0_root(const Grid& g, const DevArrays& d, double n, double preset, double* out, hipStream_t s)
{
    const double Nk2 = 4.0 * (g.nk - 1) * (g.nk - 1);
    // fewest waves that hold the wedge at kMu0K points per thread (barriers get cheaper)
    const long waves = (g.nk8 + 64L * kMu0K - 1) / (64L * kMu0K);
    if (waves > 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mu0_root, dim3(1), dim3((unsigned)(64 * waves)), 0, s, d.ekw, d.wq, g.nk8, Nk2, g.T, n,
                       preset, out, d.mu0v);
    return hipGetLastError();
}

hipError_t launch_vertex_slab(const Grid& g, const DevArrays& d, double U, double Usp, double Uch, hipStream_t s,
                              const double* dsc)
{
    const double normFact = g.T / ((double)g.N * g.N);
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        const size_t lds = slab_lds(g.N, false, sizeof(SS));
        set_lds(k_vertex_slab<NN, SS>, lds);
        hipLaunchKernelGGL((k_vertex_slab<NN, SS>), dim3(g.nwb), dim3(SLAB_THREADS), lds, s, d.chiw, (SS*)d.V,
                           normFact, U, Usp, Uch, dsc);
    }));
    return hipGetLastError();
}

hipError_t launch_vertex_pencil(const Grid& g, const DevArrays& d, double U, double Usp, double Uch, hipStream_t s,
                                const double* dsc)
{
    const double fact = U * (3.0 * Usp + Uch) / 8.0;
    DISPATCH_S(g.prec, DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL((k_vertex_pencil<LL, SS>), dim3((unsigned)((g.nk8 + 1) / 2)), dim3(PEN_THREADS), 0, s,
                           (SS*)d.V, (const SS*)d.chirt, d.chirt0, d.V0, g.nk8, fact, d.twN, U, dsc);
    }));
    return hipGetLastError();
}

hipError_t launch_sigma_bnd(const Grid& g, const DevArrays& d, double mu0, double fact, hipStream_t s,
                            const double* mu0p, double U, const double* dsc)
{
    const size_t lds = slab_lds(g.N, false, sizeof(double));
    DISPATCH_N(g.N, {
        set_lds(k_sigma_bnd<NN>, lds);
        hipLaunchKernelGGL(k_sigma_bnd<NN>, dim3(4), dim3(SLAB_THREADS), lds, s, d.ek, d.er, d.V0, d.dchirt0, d.dGrt0p,
                           mu0, g.T, fact, d.dSig, d.Sb, d.tmp, mu0p, U, dsc);
    });
    return hipGetLastError();
}

hipError_t launch_sigma_slab(const Grid& g, const DevArrays& d, double mu0, hipStream_t s, const double* mu0p)
{
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        const size_t lds = slab_lds(g.N, false, sizeof(SS));
        set_lds(k_sigma_slab<NN, SS>, lds);
        hipLaunchKernelGGL((k_sigma_slab<NN, SS>), dim3(g.Nw + 1), dim3(SLAB_THREADS), lds, s, d.ek, (const SS*)d.V,
                           mu0, g.T, g.Nw, (SS*)d.Stau, mu0p);
    }));
    return hipGetLastError();
}

hipError_t launch_sigma_pencil(const Grid& g, const DevArrays& d, hipStream_t s)
{
    DISPATCH_S(g.prec, DISPATCH_L(g.Nw, {
        hipLaunchKernelGGL((k_sigma_pencil<LL, SS, ST>), dim3((unsigned)g.nk8), dim3(sigma_pen_threads<ST>(LL)), 0, s,
                           (const SS*)d.Stau, d.Sb, d.dSig, g.nk8, g.T, (ST*)d.Self, d.Sinf2, d.Sinf3, d.Sinf4, d.Sinf,
                           d.twN, d.twX, d.twS);
    }));
    return hipGetLastError();
}

hipError_t launch_trace_diffG(const Grid& g, const DevArrays& d, double pmu, double mu0, double* out,
                              hipStream_t s)
{
    const dim3 gr = trace_grid(g);
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_trace_diffG<ST>, gr, dim3(256), 0, s, d.ekw, d.wq, (const ST*)d.Self, d.Sinf2, d.Sinf3,
                           d.Sinf4, d.Sinf, (int)g.nk8, g.Nw, g.T, mu0, pmu, d.part);
    });
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, d.part, (int)(gr.x * gr.y), out);
    return hipGetLastError();
}

hipError_t launch_trace_diffG3_pass(const Grid& g, const DevArrays& d, double* mirror, hipStream_t s)
{
    const dim3 gr = trace_grid(g);
    const int nblk = (int)gr.x;
    const bool fused = nblk <= kFusedStepMax;
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_trace_diffG3<ST>, gr, dim3(256), 0, s, d.ekw, d.wq, (const ST*)d.Self, d.Sinf2, d.Sinf3,
                           d.Sinf4, d.Sinf, (int)g.nk8, g.Nw, d.root, d.part, fused ? d.ticket : nullptr, mirror);
    });
    if (!fused) hipLaunchKernelGGL(k_halley_step, dim3(1), dim3(256), 0, s, d.part, nblk, d.root, mirror);
    return hipGetLastError();
}

hipError_t launch_trace_selfG(const Grid& g, const DevArrays& d, double mu0, double mu, int which, double* out,
                              hipStream_t s)
{
    const double nk02 = 4.0 * (g.nk - 1) * (g.nk - 1);
    const dim3 gr = trace_grid(g);
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_trace_selfG<ST>, gr, dim3(256), 0, s, d.ekw, d.wq, (const ST*)d.Self, d.Sinf2, d.Sinf3,
                           d.Sinf, (int)g.nk8, g.Nw, g.T, mu0, mu, nk02, which, d.part);
    });
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, d.part, (int)(gr.x * gr.y), out);
    return hipGetLastError();
}

hipError_t launch_hf0_slab(const Grid& g, const DevArrays& d, hipStream_t s)
{
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        const size_t lds = slab_lds(g.N, true, sizeof(SS));
        set_lds(k_hf0_slab<NN, SS>, lds);
        hipLaunchKernelGGL((k_hf0_slab<NN, SS>), dim3(g.Nw + 1), dim3(SLAB_THREADS), lds, s, d.ek, d.dek, d.mu0v, g.T,
                           g.Nw, (SS*)d.h, (SS*)d.f0);
    }));
    return hipGetLastError();
}

hipError_t launch_hf0_bnd(const Grid& g, const DevArrays& d, double mu0, double fact4, hipStream_t s,
                          const double* mu0p, double U, const double* dsc)
{
    const size_t lds = slab_lds(g.N, true, sizeof(double));
    DISPATCH_N(g.N, {
        set_lds(k_hf0_bnd<NN>, lds);
        hipLaunchKernelGGL(k_hf0_bnd<NN>, dim3(8), dim3(SLAB_THREADS), lds, s, d.ek, d.dek, d.V0, d.dchirt0, mu0, g.T,
                           g.Nw, fact4, d.dhV, d.df0V, d.hVb, d.f0Vb, d.bscr, mu0p, U, dsc);
    });
    return hipGetLastError();
}

hipError_t launch_hVf0V_slab(const Grid& g, const DevArrays& d, hipStream_t s)
{
    DISPATCH_S(g.prec, DISPATCH_N(g.N, {
        const size_t lds = slab_lds(g.N, true, sizeof(SS));
        set_lds(k_hVf0V_slab<NN, SS>, lds);
        // FP32: at most kHVBlocks workgroups, each looping over slabs.  This kernel runs beside
        // the latency-bound self-energy pencil and mu search; two of its workgroups fill a CU's
        // waves, one leaves room for theirs.  An FP64 slab (133 KB of LDS) holds a CU alone
        // either way, and fewer of them only lengthen the kernel (measured at 256^2 x 2048:
        // -0.05 ms per solve in FP32, +0.2 ms in FP64).
        const unsigned nb = (unsigned)((sizeof(SS) == sizeof(float)) ? std::min(g.Nw + 1, kHVBlocks) : g.Nw + 1);
        hipLaunchKernelGGL((k_hVf0V_slab<NN, SS>), dim3(nb), dim3(SLAB_THREADS), lds, s, (const SS*)d.V, g.Nw,
                           (SS*)d.h, (SS*)d.f0);
    }));
    return hipGetLastError();
}

hipError_t launch_jj_poles(const Grid& g, const DevArrays& d, double mu, hipStream_t s, const double* dsc)
{
    const double dt = jj_threads(g.Nw) / (g.Nw * g.T);  // = NT / (L T) of k_jj_pencil
    hipLaunchKernelGGL(k_jj_poles, dim3((unsigned)((g.nq + 127) / 128)), dim3(128), 0, s, d.ek, d.hVb, d.Sinf, g.N,
                       g.T, mu, dt, (JJPoles*)d.poles, dsc);
    return hipGetLastError();
}

hipError_t launch_jj_transpose(const Grid& g, const DevArrays& d, hipStream_t s)
{
    const dim3 gr((unsigned)((g.nq + kTT - 1) / kTT), (unsigned)((g.Nw + 1 + kTT - 1) / kTT), 2);
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_transpose<SS>, gr, dim3(256), 0, s, (const SS*)d.h, (SS*)d.hT, (const SS*)d.f0,
                           (SS*)d.f0T, g.nq, g.Nw + 1);
    });
    return hipGetLastError();
}

hipError_t launch_jj_pencil(const Grid& g, const DevArrays& d, double mu, double* part, const int* bins, int ntheta,
                            hipStream_t s, const double* dsc)
{
    const unsigned nblk = (unsigned)((g.nq + jj_points(g.Nw) - 1) / jj_points(g.Nw));
    DISPATCH_S(g.prec, DISPATCH_LJJ(g.Nw, {
        constexpr int NTJ = jj_threads(LL);
        hipLaunchKernelGGL((k_jj_pencil<LL, NTJ, SS, ST>), dim3(nblk), dim3(NTJ), 0, s, d.ek, d.dek,
                           d.d2ek, (const SS*)d.hT, (const SS*)d.f0T, d.dhV, d.df0V, d.hVb, d.f0Vb,
                           (const JJPoles*)d.poles, (const ST*)d.Self, d.Sinf2, d.Sinf3, d.Sinf, g.N, g.T, mu, d.twN,
                           d.twX, d.twS, d.twM, d.ph, part, d.jjscal, bins, ntheta, dsc);
    }));
    return hipGetLastError();
}

// out[i] = (float)in[i]
__global__ void k_narrow(const double* __restrict__ in, float* __restrict__ out, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

hipError_t launch_narrow(const double* in, float* out, long n, hipStream_t s)
{
    const long nb = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)std::max<long>(nb, 1)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

// out[r*ntheta + b] = sum over k < n with bins[k] == b of part[r*n + k], in increasing k
__global__ void __launch_bounds__(256) k_sum_rows_binned(const double* __restrict__ part, int n,
                                                         const int* __restrict__ bins, int ntheta, double* out)
{
    __shared__ double sc[16];
    const int r = blockIdx.x, b = blockIdx.y;
    const double* row = part + (long)r * n;
    double a = 0.0;
    for (int k = threadIdx.x; k < n; k += 256)
        if (bins[k] == b) a += row[k];
    a = block_sum(a, sc);
    if (threadIdx.x == 0) out[(long)r * ntheta + b] = a;
}

hipError_t launch_sum_rows_binned(const double* part, int nrows, int n, const int* bins, int ntheta, double* out,
                                  hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_rows_binned, dim3(nrows, ntheta), dim3(256), 0, s, part, n, bins, ntheta, out);
    return hipGetLastError();
}

hipError_t launch_gloc(const Grid& g, const DevArrays& d, double mu, double* part, double* out, hipStream_t s)
{
    const int H = g.Nw / 2;
    const double nk02 = 4.0 * (g.nk - 1) * (g.nk - 1);
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_gloc<ST>, dim3((H + 255) / 256, kGlocChunks), dim3(256), 0, s, d.ekw, d.wq,
                           (const ST*)d.Self, d.Sinf2, d.Sinf3, d.Sinf, g.nk8, H, g.T, mu, nk02, part);
    });
    return launch_reduce_cols(part, kGlocChunks, 2 * H, out, s);
}

// [r][c] (nrows x ncols) -> [c][r]
hipError_t launch_transpose_d(const double* in, double* out, long ncols, int nrows, hipStream_t s)
{
    const dim3 gr((unsigned)((ncols + kTT - 1) / kTT), (unsigned)((nrows + kTT - 1) / kTT));
    hipLaunchKernelGGL(k_transpose<double>, gr, dim3(256), 0, s, in, out, in, out, ncols, nrows);  // grid z = 1
    return hipGetLastError();
}

// out[r] = sum of part[r*n .. r*n+n) in a fixed order, r < nrows
hipError_t launch_sum_rows(const double* part, int nrows, int n, double* out, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize_k, dim3(nrows), dim3(256), 0, s, part, n, out);
    return hipGetLastError();
}

hipError_t launch_kx0_rows(const Grid& g, const DevArrays& d, double mu, hipStream_t s, const double* dsc,
                           double* out)
{
    DISPATCH_S(g.prec, {
        hipLaunchKernelGGL(k_kx0_rows<ST>, dim3(2 * g.nk), dim3(256), 0, s, d.ek, d.d2ek, (const ST*)d.Self, d.Sinf2,
                           d.Sinf3, d.Sinf, g.N, g.Nw, g.T, mu, out ? out : d.kx0part, dsc);
    });
    return hipGetLastError();
}

}  // namespace tpsc
