// tpsc_roots.hpp -- root searches shared by host and device code: the Halley iteration (host
// loop or device-driven passes) and the derivative-free bracketed search bracket_root (host, or
// every thread of one workgroup in step).
//
// Root of a monotone f inside (lo, hi) from x by Halley steps, with the bracket narrowed by the
// sign of every evaluation and a bisection whenever a step leaves it.  One evaluation gives
// v = {f, f', f''} (one pass over a device array).  Converged when a step is below
// 1e-13 max(|x|, 1): the cubic order leaves the last iterate at the rounding level of f.
// Fails (the caller then falls back to the bracketing search) if f' changes sign, the root
// lands on an end of the interval, or itmax evaluations are spent.  A step that leaves the
// bracket is replaced by the Newton step, and that by a bisection.
//
// f may carry a part known on the host, a0 + a1 x (so f' gains a1), added to each evaluation
// before the step; the device-driven loop keeps the state in device memory and advances it
// in a one-block kernel after each pass, so that passes run back to back without a host
// round trip.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace tpsc {

enum : int { kRootRunning = 0, kRootConverged = 1, kRootFailed = 2 };

struct HalleyState {
    double x, lo, hi, lo0, hi0, root;
    double a0, a1;  // f += a0 + a1 x, f' += a1
    double f, f1;   // last evaluation (diagnostics)
    double T, mu0;  // inputs of the pass kernels that change from point to point
    int dir, ne, status, itmax;
};

__host__ __device__ inline HalleyState halley_start(double x, double lo, double hi, int itmax, double a0, double a1)
{
    HalleyState s{};
    s.x = x;
    s.lo = s.lo0 = lo;
    s.hi = s.hi0 = hi;
    s.root = x;
    s.a0 = a0;
    s.a1 = a1;
    s.itmax = itmax;
    s.status = kRootRunning;
    return s;
}

// advance s by one evaluation v = {f, f', f''} at s.x (without the a0 + a1 x part)
__host__ __device__ inline void halley_update(HalleyState& s, const double v[3])
{
    s.ne++;
    const double f = v[0] + (s.a0 + s.a1 * s.x), f1 = v[1] + s.a1, f2 = v[2];
    s.f = f;
    s.f1 = f1;
    if (!std::isfinite(f) || !std::isfinite(f1) || f1 == 0.0) {
        s.status = kRootFailed;
        return;
    }
    if (s.dir == 0) s.dir = (f1 > 0) ? 1 : -1;
    if ((f1 > 0 ? 1 : -1) != s.dir) {
        s.status = kRootFailed;
        return;
    }
    if (f == 0.0) {
        s.root = s.x;
        s.status = kRootConverged;
        return;
    }
    if ((f > 0) == (s.dir > 0))
        s.hi = s.x;
    else
        s.lo = s.x;
    const double den = 2.0 * f1 * f1 - f * f2;
    const double dn = -f / f1;  // Newton step
    const double dh = (den != 0.0 && std::isfinite(den)) ? -2.0 * f * f1 / den : dn;
    if (fabs(dh) <= 1.0e-13 * fmax(fabs(s.x), 1.0)) {  // converged: the step is below the tolerance
        double xn = s.x + dh;
        if (!(xn >= s.lo && xn <= s.hi)) xn = s.x;
        if (xn <= s.lo0 || xn >= s.hi0) {
            s.status = kRootFailed;
            return;
        }
        s.root = xn;
        s.status = kRootConverged;
        return;
    }
    double xn = s.x + dh;
    if (!(xn > s.lo && xn < s.hi)) xn = s.x + dn;
    if (!(xn > s.lo && xn < s.hi)) xn = 0.5 * (s.lo + s.hi);
    s.x = xn;
    if (s.ne >= s.itmax) s.status = kRootFailed;
}

// Derivative-free bracketed search, for the roots whose function has no cheap derivative (the
// host fallbacks of the sum-rule and mu searches, the Fermi-surface traces) and for the mu0
// search when its Halley iteration fails.  Replaces the reference's find_zero
// (generique.cpp:604-913, and :917-1234 with limits): same inputs {start, lo, hi}, the same
// optional limits and |f| tolerance, a different method.
//
// 1. Bracket.  f is taken at lo and hi, and at the start when it lies strictly inside; the
//    part of [lo, hi] whose ends differ in sign and that has the start as an end is kept
//    ([start, hi] when both parts do, as in the reference).  Without a
//    sign change the interval grows: the end with the smaller |f| moves outward by the current
//    width (the width doubles each time), clamped to the limits when given; kRootWiden moves at
//    most.  No sign change then, or a limit reached without one, is a failure.
// 2. Shrink.  Chandrupatla's hybrid of inverse quadratic interpolation and bisection
//    (T. R. Chandrupatla, Advances in Engineering Software 28 (1997) 145-149): the next point
//    is a + t (b - a) on the bracket [a, b] (a the newest point, c the end just dropped).  The
//    quadratic through (a, b, c) is used when its inverse is monotone on the bracket, i.e.
//    phi^2 < xi and (1 - phi)^2 < 1 - xi with xi = (a - b)/(c - b), phi = (fa - fb)/(fc - fb);
//    otherwise t = 1/2.  t is kept at least tol/|b - a| away from both ends so that every
//    step moves by more than the tolerance.
// Ends when |f| <= ftol at the best point, f is exactly zero, or the bracket is narrower than
// 2 tol, tol = 2 eps |x| + xabs; the point of smaller |f| is returned.  maxev bounds the number
// of evaluations (a bisection halves the bracket, so 200 is never reached in double).
constexpr int kRootWiden = 64;

template <typename Fn>
__host__ __device__ bool bracket_root(const Fn& f, double x0, double lo, double hi, const double* lims, double ftol,
                                      double* root, int* nevals, int maxev = 200, double xabs = 1.0e-15)
{
    int ne = 0;
    auto F = [&](double x) {
        ne++;
        return f(x);
    };
    auto finish = [&](double x, bool ok) {
        *root = x;
        if (nevals) *nevals = ne;
        return ok;
    };
    auto opposite = [](double u, double v) { return (u < 0.0 && v > 0.0) || (u > 0.0 && v < 0.0); };
    double a = lo, b = hi;
    double fa = F(a), fb = F(b);
    if (!(fabs(fa) > ftol)) return finish(a, std::isfinite(fa));
    if (!(fabs(fb) > ftol)) return finish(b, std::isfinite(fb));
    if (x0 > lo && x0 < hi) {
        const double f0 = F(x0);
        if (!(fabs(f0) > ftol)) return finish(x0, std::isfinite(f0));
        if (opposite(f0, fb))  // [x0, hi] first: the side the reference searches when both halves bracket
            a = x0, fa = f0;
        else if (opposite(f0, fa))
            b = x0, fb = f0;
    }
    for (int k = 0; !opposite(fa, fb); k++) {
        if (k == kRootWiden || !std::isfinite(fa) || !std::isfinite(fb)) return finish(fabs(fa) < fabs(fb) ? a : b, false);
        const double w = b - a;
        if (fabs(fa) <= fabs(fb)) {
            double na = a - w;
            if (lims && na < lims[0]) na = lims[0];
            if (!(na < a)) return finish(a, false);
            b = a, fb = fa;
            a = na, fa = F(a);
            if (!(fabs(fa) > ftol)) return finish(a, std::isfinite(fa));
        } else {
            double nb = b + w;
            if (lims && nb > lims[1]) nb = lims[1];
            if (!(nb > b)) return finish(b, false);
            a = b, fa = fb;
            b = nb, fb = F(b);
            if (!(fabs(fb) > ftol)) return finish(b, std::isfinite(fb));
        }
    }
    double c = b, fc = fb;  // no third point yet: the first step bisects
    double t = 0.5;
    for (;;) {
        const double xt = a + t * (b - a), ft = F(xt);
        if (opposite(ft, fa)) {  // root in [xt, a]: a becomes the far end, b is dropped
            c = b, fc = fb;
            b = a, fb = fa;
        } else {  // root in [xt, b]: a is dropped
            c = a, fc = fa;
        }
        a = xt, fa = ft;
        const bool aBest = fabs(fa) < fabs(fb);
        const double xm = aBest ? a : b, fm = aBest ? fa : fb;
        if (!std::isfinite(fm)) return finish(xm, false);
        if (!(fabs(fm) > ftol)) return finish(xm, true);
        const double tol = 2.0 * DBL_EPSILON * fabs(xm) + xabs, width = fabs(b - a);
        if (width <= 2.0 * tol) return finish(xm, true);
        if (ne >= maxev) return finish(xm, false);
        const double tl = tol / width;
        const double xi = (a - b) / (c - b), phi = (fa - fb) / (fc - fb);
        if (phi * phi < xi && (1.0 - phi) * (1.0 - phi) < 1.0 - xi)
            t = fa / (fb - fa) * fc / (fb - fc) + (c - a) / (b - a) * fa / (fc - fa) * fb / (fc - fb);
        else
            t = 0.5;
        t = fmin(1.0 - tl, fmax(tl, t));
    }
}

// The reference's calling form of its root search: init = {start, lo, hi}, optional limits,
// |f| tolerance 1e-16 without limits and 1e-15 with them (generique.cpp:604-607,917-920).
template <typename Fn>
__host__ __device__ bool find_root(const Fn& f, const double init[3], double* root, const double* lims, int* nevals)
{
    return bracket_root(f, init[0], init[1], init[2], lims, lims ? 1.0e-15 : 1.0e-16, root, nevals);
}

}  // namespace tpsc
