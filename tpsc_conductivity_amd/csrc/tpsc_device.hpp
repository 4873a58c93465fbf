// tpsc_device.hpp -- device building blocks of the TPSC hot path on gfx950 (CDNA4).
//
//  * cd            : complex double; division by the plain formula with one reciprocal of
//                    (c^2+d^2)
//  * wave_r2r_*    : FFTW REDFT00 / RODFT00 of LDS-resident rows, two real rows packed
//                    into one complex sequence, one length-N FFT per 64-lane wavefront
//                    (R values per lane, cross-lane radix-2 stages with xor shuffles)
//  * block_fft     : workgroup Stockham radix-4/2 complex FFT of a tau-pencil in LDS
//  * block_spline  : clamped cubic spline moments on a uniform grid, solved as a
//                    causal + anti-causal first-order recursion (pole sqrt(3)-2) over
//                    chunks with a 32-sample warm-up, plus the two boundary modes
//  * block reductions in FP64
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpsc {

constexpr double kPi = 3.14159265358979323846;

struct cd {
    double x, y;
};
__host__ __device__ inline cd mk(double a, double b) { return cd{a, b}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.x, -a.y}; }
__host__ __device__ inline cd operator*(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.x, s * a.y}; }
__host__ __device__ inline cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline cd operator/(cd a, double s) { return cd{a.x / s, a.y / s}; }
__host__ __device__ inline cd operator/(cd a, cd b)
{
    const double r = 1.0 / (b.x * b.x + b.y * b.y);
    return cd{(a.x * b.x + a.y * b.y) * r, (a.y * b.x - a.x * b.y) * r};
}
__host__ __device__ inline cd operator/(double s, cd b) { return cd{s, 0.0} / b; }
__host__ __device__ inline cd conjg(cd a) { return cd{a.x, -a.y}; }
__host__ __device__ inline double cabs2(cd a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ inline cd cexpd(cd z)
{
    double s, c;
    sincos(z.y, &s, &c);
    const double e = exp(z.x);
    return cd{e * c, e * s};
}
__device__ inline cd cis(double ang)
{
    double s, c;
    sincos(ang, &s, &c);
    return cd{c, s};
}
__device__ inline cd csqrtd(cd z)
{
    // principal branch, cut along the negative real axis
    const double r = hypot(z.x, z.y);
    if (r == 0.0) return cd{0.0, z.y};
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (r + z.x));
        return cd{t, z.y / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (r - z.x));
    return cd{fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

__device__ inline long widx(long l, long m) { return m + (l * (l + 1)) / 2; }
// weight of wedge point (l, m), m <= l, in sums over the full zone: 8 images, halved on each
// of the edges l, m in {0, nk - 1} and on the diagonal m = l (green0.cpp:192-195)
__device__ inline double wedge_weight(int l, int m, int nk)
{
    double ww = 8;
    if (l == 0 || l == nk - 1) ww /= 2;
    if (m == 0 || m == nk - 1) ww /= 2;
    if (m == l) ww /= 2;
    return ww;
}
__device__ inline long fold_idx(long x, long y, long N)
{
    if (x > N / 2) x = N - x;
    if (y > N / 2) y = N - y;
    if (y > x) {
        const long t = x;
        x = y;
        y = t;
    }
    return widx(x, y);
}

// blockIdx -> logical block so that consecutive logical blocks run on the same XCD
// (the hardware deals workgroups round-robin over 8 XCDs); neighbouring pencils then
// share L2 lines of the tau-major arrays.
__device__ inline int xcd_remap(int b, int nb)
{
    const int nx = 8;
    const int per = (nb + nx - 1) / nx;
    const int x = b % nx, i = b / nx;
    const int full = nb - (per - 1) * nx;  // XCDs that get `per` blocks
    int base = (x < full) ? x * per : full * per + (x - full) * (per - 1);
    int lb = base + i;
    return lb;
}

// single-precision complex for the FP32 slab transforms of the mixed-precision path
struct cf {
    float x, y;
};
__host__ __device__ inline cf operator+(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cf operator-(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cf operator*(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

template <typename Real>
struct cplx_of;
template <>
struct cplx_of<double> {
    using type = cd;
};
template <>
struct cplx_of<float> {
    using type = cf;
};

// v from lane (lane ^ M), M a power of two < 64.  M = 1, 2: DPP quad permutation; M = 4:
// quad permutation 3-2-1-0 then the 8-lane mirror (i -> 7 - i = i ^ 7); M = 8: 8-lane mirror
// then the 16-lane mirror (i ^ 7 ^ 15); these are VALU moves.  M = 16: ds_swizzle in
// bitmask mode (xor within 32 lanes); M = 32: ds_bpermute.
template <int M>
__device__ __forceinline__ int xor_lane_i(int x)
{
    if constexpr (M == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    else if constexpr (M == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    else if constexpr (M == 4) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x1B, 0xF, 0xF, false);
        return __builtin_amdgcn_update_dpp(0, t, 0x141, 0xF, 0xF, false);
    } else if constexpr (M == 8) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);
        return __builtin_amdgcn_update_dpp(0, t, 0x140, 0xF, 0xF, false);
    } else if constexpr (M == 16)
        return __builtin_amdgcn_ds_swizzle(x, 0x401F);
    else
        return __shfl_xor(x, M, 64);
}
template <int M>
__device__ __forceinline__ float xor_lane(float v)
{
    return __int_as_float(xor_lane_i<M>(__float_as_int(v)));
}
template <int M>
__device__ __forceinline__ double xor_lane(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane_i<M>((int)(b & 0xffffffffLL));
    const int hi = xor_lane_i<M>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int M>
__device__ __forceinline__ cd shfl_xor_c(cd v)
{
    return cd{xor_lane<M>(v.x), xor_lane<M>(v.y)};
}
template <int M>
__device__ __forceinline__ cf shfl_xor_c(cf v)
{
    return cf{xor_lane<M>(v.x), xor_lane<M>(v.y)};
}

// ---------------------------------------------------------------------------------
// in-register DFT of length R (R = 1, 2, 3, 4, 8), forward sign, natural order output
// ---------------------------------------------------------------------------------
template <typename C>
__device__ inline C mul_mi(C a) { return C{a.y, -a.x}; }  // a * (-i)
template <typename C>
__device__ inline C mul_pi(C a) { return C{-a.y, a.x}; }  // a * (+i)

template <int R, typename C>
__device__ __forceinline__ void reg_dft(C* v)
{
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 3) {
        // w = exp(-2 pi i/3) = -1/2 - i sqrt(3)/2: X1,2 = a - (b + c)/2 -/+ i (sqrt(3)/2) (b - c)
        using F = decltype(C{}.x);
        const F h = (F)0.86602540378443864676;
        const C a = v[0], s = v[1] + v[2], d = v[1] - v[2];
        const C m = C{a.x - (F)0.5 * s.x, a.y - (F)0.5 * s.y};
        v[0] = a + s;
        v[1] = m + mul_mi(C{h * d.x, h * d.y});
        v[2] = m - mul_mi(C{h * d.x, h * d.y});
    } else if constexpr (R == 4) {
        const C a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        const C s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = a1 - a3;
        v[0] = s02 + s13;
        v[2] = s02 - s13;
        v[1] = d02 + mul_mi(d13);
        v[3] = d02 - mul_mi(d13);
    } else if constexpr (R == 8) {
        using F = decltype(C{}.x);
        const F h = (F)0.70710678118654752440;
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        reg_dft<4>(e);
        reg_dft<4>(o);
        const C w1 = C{h, -h}, w3 = C{-h, -h};
        o[1] = o[1] * w1;
        o[2] = mul_mi(o[2]);
        o[3] = o[3] * w3;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = e[k] + o[k];
            v[k + 4] = e[k] - o[k];
        }
    }
}

// ---------------------------------------------------------------------------------
// REDFT00 / RODFT00 of LDS rows by one length-L complex FFT per row pair.
//   kind 0 (REDFT00): rows of n = L/2+1 values, even extension z_j = x_{min(j,L-j)}
//   kind 1 (RODFT00): rows of n = L/2-1 values, odd extension z_0 = z_{L/2} = 0,
//                     z_j = x_{j-1}, z_{L-j} = -x_{j-1}
// Row r element e lives at base[r*rs + e*es].  The DFT of a real even (odd) sequence is
// real (imaginary), so packing rows r and r+1 as real and imaginary parts yields both
// transforms from one FFT.  L = R*G with G (lanes per FFT) the largest power of two that
// divides L, at most 64, and R in {1, 2, 3, 4, 8} (L = 2^k or 3*2^k).  The arithmetic runs
// in the precision of Real (FP64 or FP32).
// Must be called by every thread of the block (uniform control flow); caller syncs.
// tw: LDS table tw[t] = exp(-2 pi i t / L), t < L, in the precision of Real.
// ---------------------------------------------------------------------------------
// cross-lane DIF stages of span HH, HH/2, ..., 1 (lane p of a G-lane group).  The lower
// lane of a pair keeps v + o, the upper one (o - v) w; both are fma(s, v, o) w' with
// (s, w') = (1, 1) or (-1, w), so no lane evaluates both branches.  Exact: fma(+-1, v, o) rounds
// as o +- v, and the product with 1 + 0i returns its operand.
template <int HH, int R, int L, typename C>
__device__ __forceinline__ void dif_stages(C* v, int p, const C* tw)
{
    if constexpr (HH >= 1) {
        using F = decltype(C{}.x);
        const bool upper = (p & HH) != 0;
        const C wt = tw[(p & (HH - 1)) * (L / (2 * HH))];
        const C w = upper ? wt : C{(F)1, (F)0};
        const F sg = upper ? (F)-1 : (F)1;
#pragma unroll
        for (int k1 = 0; k1 < R; k1++) {
            const C o = shfl_xor_c<HH>(v[k1]);
            v[k1] = C{fma(sg, v[k1].x, o.x), fma(sg, v[k1].y, o.y)} * w;
        }
        dif_stages<HH / 2, R, L>(v, p, tw);
    }
}

template <typename Real, int R, int G, int KIND>
__device__ __forceinline__ void wave_r2r_rows(Real* base, int nrows, int rs, int es,
                                              const typename cplx_of<Real>::type* tw)
{
    using C = typename cplx_of<Real>::type;
    constexpr int L = R * G;
    constexpr int P = 64 / G;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int sub = lane / G, p = lane % G;
    const int npairs = (nrows + 1) >> 1;
    constexpr int half = L / 2;
    const int rounds = (npairs + nwaves * P - 1) / (nwaves * P);
    int br = 0;  // bit reversal of p over log2(G) bits
#pragma unroll
    for (int b = 1, q = G >> 1; b < G; b <<= 1, q >>= 1)
        if (p & b) br |= q;
    for (int rd = 0; rd < rounds; rd++) {
        const int task = (rd * nwaves + wave) * P + sub;
        const bool valid = task < npairs;
        const int r1 = 2 * task, r2 = 2 * task + 1;
        const bool has2 = valid && (r2 < nrows);
        const long o1 = (long)r1 * rs, o2 = (long)r2 * rs;
        C v[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int j = p + G * k;
            Real a = 0, b = 0;
            if (KIND == 0) {
                const int e = (j <= half) ? j : L - j;
                if (valid) a = base[o1 + (long)e * es];
                if (has2) b = base[o2 + (long)e * es];
            } else if (j != 0 && j != half) {
                const int e = (j < half) ? j - 1 : L - j - 1;
                if (valid) a = base[o1 + (long)e * es];
                if (has2) b = base[o2 + (long)e * es];
                if (j > half) {
                    a = -a;
                    b = -b;
                }
            }
            v[k] = C{a, b};
        }
        reg_dft<R>(v);
        // twiddle by exp(-2 pi i p k1 / L)
#pragma unroll
        for (int k1 = 1; k1 < R; k1++) v[k1] = v[k1] * tw[(p * k1) % L];
        // G-point DIF across lanes; lane p ends with output index bitrev(p)
        dif_stages<G / 2, R, L>(v, p, tw);
        if (valid) {
#pragma unroll
            for (int k1 = 0; k1 < R; k1++) {
                const int k = k1 + R * br;
                if (KIND == 0) {
                    if (k <= half) {
                        base[o1 + (long)k * es] = v[k1].x;
                        if (has2) base[o2 + (long)k * es] = v[k1].y;
                    }
                } else if (k >= 1 && k <= half - 1) {
                    base[o1 + (long)(k - 1) * es] = -v[k1].y;
                    if (has2) base[o2 + (long)(k - 1) * es] = v[k1].x;
                }
            }
        }
    }
}

// fill tw[t] = exp(-2 pi i t / L) for t < L (block-cooperative; caller syncs); the angle
// is evaluated in FP64 and rounded to the table precision
template <typename C>
__device__ __forceinline__ void fill_twiddles(C* tw, int L)
{
    using F = decltype(C{}.x);
    for (int t = threadIdx.x; t < L; t += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)t / (double)L, &s, &c);
        tw[t] = C{(F)c, (F)(-s)};
    }
}

// 2D transform of an n0 x n1 LDS slab (row-major, row stride ld): axis 1 then axis 0.
// KIND0 along axis 0 (slow index), KIND1 along axis 1 (0: REDFT00, 1: RODFT00).  Both use
// FFT length N.
template <typename Real, int N, int KIND0, int KIND1>
__device__ __forceinline__ void slab_r2r_2d(Real* slab, int n0, int n1, int ld, const typename cplx_of<Real>::type* tw)
{
    constexpr int G = ((N & -N) < 64) ? (N & -N) : 64;  // lanes: the power-of-two part of N, at most 64
    constexpr int R = N / G;
    static_assert(R == 1 || R == 2 || R == 3 || R == 4 || R == 8, "transform length 2^k or 3*2^k");
    wave_r2r_rows<Real, R, G, KIND1>(slab, n0, ld, 1, tw);  // rows (axis 1)
    __syncthreads();
    wave_r2r_rows<Real, R, G, KIND0>(slab, n1, 1, ld, tw);  // columns (axis 0)
    __syncthreads();
}

// ---------------------------------------------------------------------------------
// Workgroup Stockham FFT in place in LDS: radix-8 passes, then one radix-4 or radix-2 pass
// for the remaining factor of L.  sign = -1: exp(-2 pi i jk/L) (FFTW_FORWARD); sign = +1:
// FFTW_BACKWARD.  twg: global table exp(-2 pi i t / L), t < L.
// Pass with butterfly span Ns: butterfly b (k = b mod Ns) reads x[b + r L/RAD], multiplies
// by w^(r k L/(Ns RAD)), takes an RAD-point DFT and writes y[(b/Ns) Ns RAD + k + m Ns].
// The twiddles w^r come from three table reads (w, w^2, w^4) and products, issued before
// the LDS reads and the barrier so their latency overlaps them.
// ---------------------------------------------------------------------------------
// PP (ping-pong): the pass reads src and writes a different buffer dst, so the barrier
// between its reads and its writes is not needed; in place (src == dst) it is.
template <int RAD, int L, int NT, bool PP = false>
__device__ __forceinline__ void fft_pass(const cd* src, cd* dst, int sign, const cd* __restrict__ twg, int Ns)
{
    constexpr int NB = L / RAD;
    constexpr int PER = (NB + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int stride = L / (Ns * RAD);
    cd w1[PER], w2[PER], w4[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        const int k = b % Ns;
        cd a = cd{1.0, 0.0}, c2 = cd{1.0, 0.0}, c4 = cd{1.0, 0.0};
        if (Ns > 1 && b < NB) {
            a = twg[(k * stride) & (L - 1)];
            if (RAD >= 4) c2 = twg[(2 * k * stride) & (L - 1)];
            if (RAD >= 8) c4 = twg[(4 * k * stride) & (L - 1)];
        }
        if (sign > 0) {
            a.y = -a.y;
            c2.y = -c2.y;
            c4.y = -c4.y;
        }
        w1[i] = a;
        w2[i] = c2;
        w4[i] = c4;
    }
    cd v[PER][RAD];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        if (b < NB) {
#pragma unroll
            for (int r = 0; r < RAD; r++) v[i][r] = src[b + r * NB];
        }
    }
    if constexpr (!PP) __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        if (b < NB) {
            const int k = b % Ns;
            if (Ns > 1) {
                v[i][1] = v[i][1] * w1[i];
                if (RAD >= 4) {
                    v[i][2] = v[i][2] * w2[i];
                    v[i][3] = v[i][3] * (w1[i] * w2[i]);
                }
                if (RAD >= 8) {
                    const cd w3 = w1[i] * w2[i];
                    v[i][4] = v[i][4] * w4[i];
                    v[i][5] = v[i][5] * (w4[i] * w1[i]);
                    v[i][6] = v[i][6] * (w4[i] * w2[i]);
                    v[i][7] = v[i][7] * (w4[i] * w3);
                }
            }
            if (sign > 0) {
#pragma unroll
                for (int r = 0; r < RAD; r++) v[i][r].y = -v[i][r].y;
            }
            reg_dft<RAD>(v[i]);
            if (sign > 0) {
#pragma unroll
                for (int r = 0; r < RAD; r++) v[i][r].y = -v[i][r].y;
            }
            const int o = (b / Ns) * Ns * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; r++) dst[o + r * Ns] = v[i][r];
        }
    }
    __syncthreads();
}

template <int L, int NT>
__device__ __forceinline__ void block_fft(cd* buf, int sign, const cd* __restrict__ twg)
{
    constexpr int LOG2 = __builtin_ctz(L);
    constexpr int P8 = LOG2 / 3;
    constexpr int REM = LOG2 - 3 * P8;
    int Ns = 1;
#pragma unroll
    for (int pass = 0; pass < P8; pass++) {
        fft_pass<8, L, NT>(buf, buf, sign, twg, Ns);
        Ns *= 8;
    }
    if constexpr (REM == 2) fft_pass<4, L, NT>(buf, buf, sign, twg, Ns);
    if constexpr (REM == 1) fft_pass<2, L, NT>(buf, buf, sign, twg, Ns);
}

// The same transform with a second LDS buffer tmp of L values: the passes alternate between
// buf and tmp and take one barrier each instead of two.  buf must be complete on entry (the
// caller's barrier); the result is in buf (after an odd number of passes it is copied back).
template <int L, int NT>
__device__ __forceinline__ void block_fft_pp(cd* buf, cd* tmp, int sign, const cd* __restrict__ twg)
{
    constexpr int LOG2 = __builtin_ctz(L);
    constexpr int P8 = LOG2 / 3;
    constexpr int REM = LOG2 - 3 * P8;
    constexpr int NP = P8 + (REM > 0 ? 1 : 0);
    cd* a = buf;
    cd* b = tmp;
    int Ns = 1;
#pragma unroll
    for (int pass = 0; pass < P8; pass++) {
        fft_pass<8, L, NT, true>(a, b, sign, twg, Ns);
        cd* t = a;
        a = b;
        b = t;
        Ns *= 8;
    }
    if constexpr (REM == 2) fft_pass<4, L, NT, true>(a, b, sign, twg, Ns);
    if constexpr (REM == 1) fft_pass<2, L, NT, true>(a, b, sign, twg, Ns);
    if constexpr (NP % 2 == 1) {
        for (int j = threadIdx.x; j < L; j += NT) buf[j] = tmp[j];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// Clamped cubic spline on a uniform grid x_j = j h, j = 0..N (N intervals).
// Moments M_j = S''(x_j) solve
//   2 M_0 + M_1 = r_0,  M_{j-1} + 4 M_j + M_{j+1} = r_j,  M_{N-1} + 2 M_N = r_N
// with r_0 = 6/h ((y_1-y_0)/h - d0), r_j = 6/h^2 (y_{j+1} - 2 y_j + y_{j-1}),
// r_N = 6/h (dN - (y_N-y_{N-1})/h).  The interior operator has Green's function
// A z^{|j|}, z = sqrt(3)-2, A = 1/(2 sqrt 3); P_j = A sum_{i=1}^{N-1} z^{|j-i|} r_i
// = A (u_j + w_j - r_j) with u (w) the causal (anti-causal) recursion s <- r + z s.
// M_j = P_j + alpha z^j + beta z^{N-j}; alpha, beta from the two boundary rows.
// The boundary modes reach 2 KW samples into the sequence (|z|^66 < 1e-37, below FP64
// rounding of the moments).
// ---------------------------------------------------------------------------------
constexpr int KW = 33;

// z^e of the spline pole z = sqrt(3) - 2 by repeated multiplication, z^(N-1) by repeated
// squaring, and the number of neighbouring chunks whose carries count (|z^(C (d-1))| > 1e-19,
// at most maxd)
constexpr double kSplineZ = 1.7320508075688772935 - 2.0;
constexpr double spline_zpow(int e)
{
    double p = 1.0;
    for (int i = 0; i < e; i++) p *= kSplineZ;
    return p;
}
constexpr double spline_zpow_sq(int e)
{
    double r = 1.0, zz = kSplineZ;
    while (e > 0) {
        if (e & 1) r *= zz;
        zz *= zz;
        e >>= 1;
    }
    return r;
}
// z^(S0 + C i), i = 0..I, as compile-time constants: v[i] is the product of S0 + C i factors z
// taken in turn from 1, as a run-time loop p *= z forms it (the same roundings)
template <int I>
struct SplineZTable {
    double v[I + 1];
};
template <int S0, int C, int I>
constexpr SplineZTable<I> spline_ztable()
{
    SplineZTable<I> t{};
    for (int i = 0; i <= I; i++) t.v[i] = spline_zpow(S0 + C * i);
    return t;
}
constexpr int kZSelMax = 16;  // longest table picked by selects instead of a multiply loop
constexpr int spline_carry_terms(double zc, int maxd)
{
    int d = 0;
    double zd = 1.0;
    while (d < maxd && (zd > 1.0e-19 || zd < -1.0e-19)) {
        d++;
        zd *= zc;
    }
    return d;
}

// The clamped spline above, in place, for K sequences of length N (known at compile time) at
// once, their latencies overlapping: yM[k] holds y_0..y_N on entry and the
// moments M_0..M_N on exit.  Each thread keeps its chunk of the right-hand side in registers, so
// no scratch array is needed.  The two first-order recursions run once from a zero state to get
// the chunk-end states, and again after the exchange, seeded with the carries of the
// neighbouring chunks, S_{t-1} = sum_{d>=1} z^{C(d-1)} E_{t-d}, truncated once z^{C(d-1)} <
// 1e-19.  The chunk layout is fixed at compile time, with almost no divergent branches: the
// sequences are padded with zeros past N (the anti-causal recursion then reaches N in the zero
// state it starts from, so the moments do not change), the carries take their compile-time
// number of terms with out-of-range neighbours read as zero, full chunks store without
// per-element tests, and only the chunks within 2 KW of an end take the boundary modes.
// trap (optional): trap[k] receives this thread's part of sum_j h (y_j + y_{j+1})/2
// - h^3 (M_j + M_{j+1})/24, the integral of spline k over [0, N h]; the caller adds the parts.
// ES: LDS scratch of K*(2*NT + 8) doubles.
template <int N, int NT, int K>
__device__ __forceinline__ void block_spline_fixed(double* const* yM, double h, const double* d0, const double* dN,
                                                   double* ES, double* trap)
{
    constexpr int C = (N + 1 + NT - 1) / NT;
    constexpr double z = kSplineZ;
    constexpr double A = 0.28867513459481288225;
    constexpr double zc = spline_zpow(C);
    constexpr int D = spline_carry_terms(zc, NT - 1);
    constexpr int tN = N / C, cN = N % C;                  // thread and slot of j = N
    constexpr int t1 = 1 / C, c1 = 1 % C;                  // of j = 1
    constexpr int tM = (N - 1) / C, cM = (N - 1) % C;      // of j = N - 1
    constexpr double zN1 = spline_zpow_sq(N - 1);
    constexpr double dg = 2.0 + z, off = zN1 * (1.0 + 2.0 * z);
    constexpr double idet = 1.0 / (dg * dg - off * off);
    constexpr int TL = 2 * KW / C;                                     // last thread with j0 <= 2 KW
    constexpr int MH = (2 * KW >= cN + 1) ? (2 * KW - cN - 1) / C + 1 : 0;  // last m with N - jl <= 2 KW
    const int tid = threadIdx.x;
    const int j0 = tid * C;
    const bool full = j0 + C - 1 <= N;
    const double ih = 1.0 / h;
    double* bnd = ES + 2 * K * NT;  // per sequence: r_0, r_N, M_0, M_1, M_{N-1}, M_N (8 slots)
    double ri[K][C], u[K][C];
    double tr[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double* y = yM[k];
        double yv[C + 2];  // y[j0 - 1 .. j0 + C], indices clamped to [0, N]
#pragma unroll
        for (int i = 0; i < C + 2; i++) yv[i] = y[min(max(j0 - 1 + i, 0), N)];
        tr[k] = 0.0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int j = j0 + c;
            const double r = 6.0 * ih * ih * (yv[c + 2] - 2.0 * yv[c + 1] + yv[c]);
            ri[k][c] = (j >= 1 && j <= N - 1) ? r : 0.0;
            if (trap && j < N) tr[k] += h * (yv[c + 1] + yv[c + 2]) * 0.5;
        }
        if (tid == 0) bnd[8 * k + 0] = 6.0 * ih * ((yv[2] - yv[1]) * ih - d0[k]);
        if (tid == tN) bnd[8 * k + 1] = 6.0 * ih * (dN[k] - (yv[cN + 1] - yv[cN]) * ih);
    }
    // chunk-end states of the two recursions from a zero start
#pragma unroll
    for (int k = 0; k < K; k++) {
        double sa = 0.0;
#pragma unroll
        for (int c = 0; c < C; c++) sa = ri[k][c] + z * sa;
        ES[2 * k * NT + tid] = sa;
        sa = 0.0;
#pragma unroll
        for (int c = C - 1; c >= 0; c--) sa = ri[k][c] + z * sa;
        ES[2 * k * NT + NT + tid] = sa;
    }
    __syncthreads();  // all reads of y are done
#pragma unroll
    for (int k = 0; k < K; k++) {
        double Su = 0.0, Sw = 0.0, zd = 1.0;
#pragma unroll
        for (int d = 1; d <= D; d++) {
            const double el = ES[2 * k * NT + max(tid - d, 0)];
            const double er = ES[2 * k * NT + NT + min(tid + d, NT - 1)];
            Su += zd * (tid - d >= 0 ? el : 0.0);
            Sw += zd * (tid + d < NT ? er : 0.0);
            zd *= zc;
        }
        // the recursions again, seeded with the carries from the neighbouring chunks
        double sa = Su;
#pragma unroll
        for (int c = 0; c < C; c++) {
            sa = ri[k][c] + z * sa;
            u[k][c] = sa;
        }
        sa = Sw;
#pragma unroll
        for (int c = C - 1; c >= 0; c--) {
            sa = ri[k][c] + z * sa;
            u[k][c] = A * (u[k][c] + sa - ri[k][c]);
        }
        double* y = yM[k];
        if (full) {
#pragma unroll
            for (int c = 0; c < C; c++) y[j0 + c] = u[k][c];
        } else if (j0 <= N) {
#pragma unroll
            for (int c = 0; c < C; c++)
                if (j0 + c <= N) y[j0 + c] = u[k][c];
        }
        // end moments for the boundary modes (read after the barrier, never rewritten)
        if (tid == 0) bnd[8 * k + 2] = u[k][0];
        if (tid == t1) bnd[8 * k + 3] = u[k][c1];
        if (tid == tM) bnd[8 * k + 4] = u[k][cM];
        if (tid == tN) bnd[8 * k + 5] = u[k][cN];
    }
    __syncthreads();
    const bool edge = j0 <= N && (j0 <= 2 * KW || j0 + C - 1 >= N - 2 * KW);
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (edge) {
            const double R0 = bnd[8 * k + 0] - 2.0 * bnd[8 * k + 2] - bnd[8 * k + 3];
            const double RN = bnd[8 * k + 1] - bnd[8 * k + 4] - 2.0 * bnd[8 * k + 5];
            const double alpha = (dg * R0 - off * RN) * idet;
            const double beta = (dg * RN - off * R0) * idet;
            // z^j and z^(N-j) of this chunk by repeated multiplication from the chunk's first
            // power: z^j0 (thread tid <= TL) and z^(N-jl) (thread tN - m, m <= MH: exponent
            // cN + 1 + (m - 1) C) picked from compile-time tables, no serial loop
            double zlo[C], zhi[C];
            // the thread index through an empty asm, so the picks stay here instead of being
            // hoisted out of a caller's loop into registers held across it
            int ts = tid;
            asm volatile("" : "+v"(ts));
            double zp = 1.0;
            if constexpr (TL <= kZSelMax) {
                constexpr auto tlo = spline_ztable<0, C, TL>();
#pragma unroll
                for (int t = 1; t <= TL; t++) zp = (ts == t) ? tlo.v[t] : zp;
            } else if (j0 <= 2 * KW) {
                for (int e = 0; e < j0; e++) zp *= z;
            }
#pragma unroll
            for (int c = 0; c < C; c++) {
                zlo[c] = zp;
                zp *= z;
            }
            const int jl = min(j0 + C - 1, N);
            zp = 1.0;
            if constexpr (MH <= kZSelMax) {
                if constexpr (MH >= 1) {
                    constexpr auto thi = spline_ztable<cN + 1, C, MH - 1>();
#pragma unroll
                    for (int m = 1; m <= MH; m++) zp = (ts == tN - m) ? thi.v[m - 1] : zp;
                }
            } else if (N - jl <= 2 * KW) {
                for (int e = 0; e < N - jl; e++) zp *= z;
            }
#pragma unroll
            for (int c = C - 1; c >= 0; c--) {
                zhi[c] = zp;
                if (j0 + c <= jl) zp *= z;
            }
            double* y = yM[k];
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int j = j0 + c;
                double add = 0.0;
                if (j <= 2 * KW) add += alpha * zlo[c];
                if (N - j <= 2 * KW) add += beta * zhi[c];
                if (add != 0.0 && j <= N) {
                    const double m = u[k][c] + add;
                    y[j] = m;
                    u[k][c] = m;
                }
            }
        }
        if (trap) {
            double tm = 0.0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int j = j0 + c;
                tm += (j > N) ? 0.0 : (j == 0 || j == N) ? u[k][c] : 2.0 * u[k][c];
            }
            trap[k] = tr[k] - h * h * h * tm * (1.0 / 24.0);
        }
    }
    __syncthreads();
}


// ---------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// 1/x for finite, nonzero x of normal range: the hardware reciprocal (v_rcp_f64, relative
// error below 2^-22) refined by two Newton steps, which leave it within an ulp of the quotient;
// five instructions against about ten of an IEEE division.  x = 0 gives NaN, not inf.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// s / b (real s, complex b != 0) with one rcp_nr of |b|^2
__device__ __forceinline__ cd sdiv_rcp(double s, cd b)
{
    const double r = s * rcp_nr(b.x * b.x + b.y * b.y);
    return cd{b.x * r, -b.y * r};
}

// block-wide sums of K values at once; every thread gets the results.
// scratch: LDS double[>= K * nwaves]
template <int K>
__device__ __forceinline__ void block_sum_k(double* v, double* scratch)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) scratch[k * nw + wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int i = 0; i < nw; i++) s += scratch[k * nw + i];
        v[k] = s;
    }
    __syncthreads();
}

// block-wide sum; every thread gets the result. scratch: LDS double[>= nwaves]
__device__ inline double block_sum(double v, double* scratch)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += scratch[i];
    __syncthreads();
    return s;
}


// ---------------------------------------------------------------------------------
// slab layout helpers, Sigma(k,iwn) storage forms and the Lambda-grid integrand h / f0
// (shared by tpsc_kernels.hip and tpsc_al.hip)
// ---------------------------------------------------------------------------------
// wedge (l,m) of a packed index, m <= l
// The first guess for l comes from the hardware FP32 square root (idx < 2^24 is exact in FP32,
// the guess is within one of the row); the two loops make it exact.
__device__ __forceinline__ void wedge_lm(long idx, int& l, int& m)
{
    const int id = (int)idx;
    int L = (int)((__builtin_amdgcn_sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((L + 1) * (L + 2) / 2 <= id) L++;
    while (L * (L + 1) / 2 > id) L--;
    l = L;
    m = id - L * (L + 1) / 2;
}

// write a wedge value to both symmetric positions of an nk x nk LDS slab
__device__ __forceinline__ void put_sym(double* slab, int nk, int l, int m, double v)
{
    slab[m + l * nk] = v;
    slab[l + m * nk] = v;
}

// twiddle table placed after an n-element slab in LDS, 16-byte aligned
template <typename Real>
__device__ __forceinline__ typename cplx_of<Real>::type* slab_tw(Real* slab, long n)
{
    const long bytes = (n * (long)sizeof(Real) + 15) & ~15L;
    return (typename cplx_of<Real>::type*)((char*)slab + bytes);
}

template <typename Real>
__device__ __forceinline__ void put_sym_t(Real* slab, int nk, int l, int m, Real v)
{
    slab[m + l * nk] = v;
    slab[l + m * nk] = v;
}

__device__ __forceinline__ cd sigma_mom(double Sinf, double s2, double s3, cd R, double wn)
{
    // Sinf/z + s2/z^2 + s3/z^3 + R/z^4, z = i wn (green.cpp:3039), with one reciprocal of wn
    const double iw = 1.0 / wn, iw2 = iw * iw;
    return cd{(-s2 + R.x * iw2) * iw2, (-Sinf + (s3 + R.y * iw) * iw2) * iw};
}

// Sigma(k, i w_n), w_n > 0, at offset o of the stored form: the remainder R (FP64 path,
// assembled with the moments) or the assembled value itself (mixed path)
__device__ __forceinline__ cd sigma_kw(const cd* __restrict__ Self, long o, double Sinf, double s2, double s3,
                                       double wn)
{
    return sigma_mom(Sinf, s2, s3, Self[o], wn);
}
__device__ __forceinline__ cd sigma_kw(const cf* __restrict__ Self, long o, double, double, double, double)
{
    const cf v = Self[o];
    return cd{(double)v.x, (double)v.y};
}
// k-space integrand of h (which = 0) or f0 (which = 1) at tau = t (cond_opt.cpp:2697-2735,
// 2795-2827); nf = 1/N^2, v = v_x(k)
__device__ __forceinline__ double hf0_val(int which, double e, double v, double mu0, double T, double t, double nf)
{
    if (which == 0) {
        const double x = e - mu0;
        return (x > 0) ? nf * v * exp(-x * t) / (exp(-x / T) + 1.0)
                       : nf * v * exp(x * (1.0 / T - t)) / (exp(x / T) + 1.0);
    }
    if (e > mu0) {
        const double ex = exp((mu0 - e) / T);
        return nf * v * exp((mu0 - e) * t) * (t / (ex + 1.0) - ex / (T * (ex + 1.0) * (ex + 1.0)));
    }
    const double ex = exp((e - mu0) / T);
    return nf * v * exp((e - mu0) * (1.0 / T - t)) * (t / (ex + 1.0) - 1.0 / (T * (ex + 1.0) * (ex + 1.0)));
}

// h and f0 of the same point at once: they share both exponentials, exp(-x t) (x = e - mu0 > 0)
// or exp(x (1/T - t)), and exp(-|x|/T), whose arguments are the same numbers (mu0 - e = -x
// exactly), so hv, fv equal hf0_val(0, ...), hf0_val(1, ...) bit for bit.
__device__ __forceinline__ void hf0_pair(double e, double v, double mu0, double T, double t, double nf, double& hv,
                                         double& fv)
{
    const double x = e - mu0;
    if (x > 0) {
        const double E = exp(-x * t), ex = exp(-x / T);
        hv = nf * v * E / (ex + 1.0);
        fv = nf * v * E * (t / (ex + 1.0) - ex / (T * (ex + 1.0) * (ex + 1.0)));
    } else {
        const double E = exp(x * (1.0 / T - t)), ex = exp(x / T);
        hv = nf * v * E / (ex + 1.0);
        fv = nf * v * E * (t / (ex + 1.0) - 1.0 / (T * (ex + 1.0) * (ex + 1.0)));
    }
}


// tau-derivative of the h integrand at tau = 0+ (side 0) or beta- (side 1)
// (cond_opt.cpp:2745-2781)
__device__ __forceinline__ double h_slope_val(int side, double e, double v, double mu0, double T, double nf)
{
    if (e > mu0) {
        const double ex = exp((mu0 - e) / T);
        return (side == 0) ? -nf * v * (e - mu0) / (ex + 1.0) : -nf * v * (e - mu0) * ex / (ex + 1.0);
    }
    const double ex = exp((e - mu0) / T);
    return (side == 0) ? -nf * v * (e - mu0) * ex / (ex + 1.0) : -nf * v * (e - mu0) / (ex + 1.0);
}

}  // namespace tpsc
