// tpsc_roots.hpp -- root searches shared by host and device code: the Halley iteration (host
// loop or device-driven passes) and the bracketing search find_zero_t (host, or every thread of
// one workgroup in step).
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

// generique.cpp:604-913 (lims == nullptr) and :917-1234 (lims given): bracketed inverse
// quadratic steps with secant fallback and bisection on stalls; secant search first when
// the initial interval has no sign change.  Returns the abscissa of the smallest |f| kept.
// The reference leaves the secant start unset when f(init[0]) is the smallest; it starts
// from init[0] here.
template <typename Fn>
__host__ __device__ bool find_zero_t(const Fn& f, const double init[3], double* root, const double* lims, int* nevals)
{
    const double tol = lims ? 1.0e-15 : 1.0e-16;
    const double eps = (lims ? 2.0 : 10.0) * DBL_EPSILON;
    const int itmax = lims ? 12 : 20;
    const double dytol = 0.01;
    int it = 0, ne = 0;
    double a = 0.0, b = 0.0, c = 0.0;
    const double xmin = init[1], xmax = init[2];
    double x[3] = {init[0], xmin, xmax}, y[3] = {0.0, 0.0, 0.0};
    auto F = [&](double v) {
        ne++;
        return f(v);
    };
    auto done = [&](double v, bool ok) {
        *root = v;
        if (nevals) *nevals = ne;
        return ok;
    };
    y[1] = F(xmin);
    y[2] = F(xmax);
    if (fabs(y[1]) < tol) return done(xmin, true);
    if (fabs(y[2]) < tol) return done(xmax, true);
    if (y[1] * y[2] > 0) {
        y[0] = F(x[0]);
        if (fabs(y[0]) < tol) return done(x[0], true);
        if (y[0] * y[2] < 0) {
            const double xt = x[0];
            x[0] = (x[0] + x[2]) / 2;
            x[1] = xt;
            y[1] = y[0];
        } else {
            double yt = y[0], xt = x[0];
            if (fabs(y[1]) < fabs(yt)) yt = y[1], xt = x[1];
            if (fabs(y[2]) < fabs(yt)) yt = y[2], xt = x[2];
            double xt1 = xt + 10 * eps, yt1 = F(xt1);
            xt1 = xt - yt * ((xt1 - xt) / (yt1 - yt));
            yt1 = F(xt1);
            it++;
            while (fabs(yt1) > tol && it < itmax && yt1 * yt > 0 && std::isfinite(xt1)) {
                xt = xt1;
                yt = yt1;
                xt1 = xt + 10 * eps;
                yt1 = F(xt1);
                xt1 = xt - yt * ((xt1 - xt) / (yt1 - yt));
                yt1 = F(xt1);
                it++;
            }
            if (fabs(yt1) < tol) return done(xt1, true);
            if (yt1 * yt < 0) {
                x[0] = (xt + xt1) / 2;
                x[1] = xt;
                y[1] = yt;
                x[2] = xt1;
                y[2] = yt1;
            }
            if (it >= itmax || !std::isfinite(xt1)) return done(xt1, false);
        }
    }
    if (lims && (x[1] < lims[0] || x[1] > lims[1] || x[2] < lims[0] || x[2] > lims[1])) return done(x[0], false);
    if (fabs(xmax - xmin) < eps) return done(xmin, true);
    y[0] = F(x[0]);
    if (fabs(y[0]) < tol) return done(x[0], true);

    int i3, imid, ifar, im, ip;
    auto sides = [&]() {
        int i1, i2;
        if (y[0] * y[1] > 0)
            i1 = 0, i2 = 1;
        else if (y[0] * y[2] > 0)
            i1 = 0, i2 = 2;
        else
            i1 = 1, i2 = 2;
        i3 = 3 - i1 - i2;
        if (x[i3] > x[i1]) {
            if (x[i1] > x[i2])
                ifar = i2, imid = i1;
            else
                ifar = i1, imid = i2;
            im = imid, ip = i3;
        } else {
            if (x[i1] < x[i2])
                ifar = i2, imid = i1;
            else
                ifar = i1, imid = i2;
            im = i3, ip = imid;
        }
    };
    sides();
    x[ifar] = -y[imid] * (x[i3] - x[imid]) / (y[i3] - y[imid]) + x[imid];
    y[ifar] = F(x[ifar]);
    double dy = y[ifar];
    while (fabs(dy) > tol && it < itmax) {
        sides();
        const double den = -x[0] * x[0] * x[2] + x[0] * x[0] * x[1] - x[0] * x[1] * x[1] + x[0] * x[2] * x[2] -
                           x[2] * x[2] * x[1] + x[2] * x[1] * x[1];
        if (fabs(den) > 0.0) {
            a = (x[0] * y[2] - x[0] * y[1] + x[2] * y[1] - y[2] * x[1] - y[0] * x[2] + y[0] * x[1]) / den;
            b = -(-y[2] * x[1] * x[1] + x[1] * x[1] * y[0] - x[2] * x[2] * y[0] - y[1] * x[0] * x[0] +
                  y[2] * x[0] * x[0] + y[1] * x[2] * x[2]) / den;
            c = (y[2] * x[0] * x[0] * x[1] - y[1] * x[0] * x[0] * x[2] - x[1] * x[1] * x[0] * y[2] +
                 x[2] * x[2] * x[0] * y[1] + x[1] * x[1] * y[0] * x[2] - x[2] * x[2] * y[0] * x[1]) / den;
        }
        auto secant_or_stop = [&](bool& stop) {
            if (fabs(y[i3] - y[imid]) > tol)
                x[ifar] = -y[imid] * (x[i3] - x[imid]) / (y[i3] - y[imid]) + x[imid];
            else
                stop = true;
        };
        bool stop = false;
        if ((b * b - 4.0 * a * c) > 0.0 && fabs(den) > 0.0) {
            const double xp = (-b + sqrt(b * b - 4.0 * a * c)) / (2 * a);
            const double xm = (-b - sqrt(b * b - 4.0 * a * c)) / (2 * a);
            if (xp > x[im] && xp < x[ip])
                x[ifar] = xp;
            else if (xm > x[im] && xm < x[ip])
                x[ifar] = xm;
            else
                secant_or_stop(stop);
        } else
            secant_or_stop(stop);
        if (stop) return done(x[imid], true);
        y[ifar] = F(x[ifar]);
        double dyprec = fabs(dy);
        dy = y[ifar];
        if (fabs(dyprec - fabs(dy)) < dytol * dyprec || dyprec < fabs(dy)) {
            x[ifar] = (x[i3] + x[imid]) / 2.0;
            y[ifar] = F(x[ifar]);
            dy = y[ifar];
        }
        it++;
    }
    int best = 0;
    if (fabs(y[1]) < fabs(y[best])) best = 1;
    if (fabs(y[2]) < fabs(y[best])) best = 2;
    return done(x[best], true);
}

}  // namespace tpsc
