// Device special functions for the Furman-Pivi emitter, written from published methods:
//   P(a, x)        series / Lentz continued fraction for the incomplete gamma function
//                  (Abramowitz & Stegun 6.5.29 and 6.5.31; Lentz 1976, Thompson & Barnett 1986)
//   I_x(a, b)      Lentz continued fraction of the incomplete beta function (A&S 26.5.8)
//   inverses       safeguarded Halley iteration inside a shrinking bracket, started from the
//                  Wilson-Hilferty approximation (gamma) and the A&S 26.5.22 approximation
//                  (beta), each with the A&S 26.2.23 normal quantile.
// All are FP64.  The inverses stop after a Halley (or Newton) step of at most kStepTol |x|:
// the error left is then of order kStepTol^3 (kStepTol^2 for Newton), far below rounding.
// The test comes before the bracket safeguard: at convergence the step rounds to zero and the
// new iterate equals the bracket end just set, which the safeguard would replace by a bisection
// point far from the root (the reference stops at 1e-9 / 1e-12).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace sf {

constexpr double kTiny = 1e-300;
constexpr double kEps = 1.1102230246251565e-16;
constexpr double kStepTol = 1e-8;

// 1/x from the hardware estimate and two Newton steps r <- r + r (1 - x r) (each at least squares
// the relative error), for the continued fractions' inner loops, where a correctly rounded
// division costs about twice as many dependent instructions.  Finite nonzero x only.  (A host
// build of this header, as the host test makes, divides.)
#ifndef OECL_SF_FAST_RCP
#define OECL_SF_FAST_RCP 1
#endif
__device__ __forceinline__ double rcp(double x)
{
#if OECL_SF_FAST_RCP && defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.), r, r);
    return fma(fma(-x, r, 1.), r, r);
#else
    return 1. / x;
#endif
}

// Normal quantile, A&S 26.2.23 (|error| < 4.5e-4): only used as a starting point.
__device__ __forceinline__ double norm_quantile_rough(double p)
{
    double q = p < 0.5 ? p : 1. - p;
    if (q <= 0.) q = 1e-300;
    double t = sqrt(-2. * log(q));
    double z = t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                       (1. + t * (1.432788 + t * (0.189269 + t * 0.001308)));
    return p < 0.5 ? -z : z;
}

// Regularized lower incomplete gamma P(a, x), a > 0, given lga = lgamma(a).
__device__ inline double gammainc_lg(double a, double x, double lga)
{
    if (!(x > 0.)) return 0.;
    double lpref = -x + a * log(x) - lga;
    if (x < a + 1.) {
        // P = e^{-x} x^a / Gamma(a+1) * sum_n x^n / ((a+1)...(a+n))
        double term = 1. / a, sum = term, den = a;
        for (int n = 1; n < 2000; ++n) {
            den += 1.;
            term *= x * rcp(den);
            sum += term;
            if (fabs(term) < fabs(sum) * kEps) break;
        }
        return sum * exp(lpref);
    }
    // Q by the continued fraction 1/(x+1-a- 1(1-a)/(x+3-a- 2(2-a)/(x+5-a- ...))), Lentz.
    double b = x + 1. - a;
    double c = 1. / kTiny;
    double d = 1. / b;
    double h = d;
    for (int i = 1; i < 2000; ++i) {
        double an = -i * (i - a);
        b += 2.;
        d = an * d + b;
        if (fabs(d) < kTiny) d = kTiny;
        c = b + an * rcp(c);
        if (fabs(c) < kTiny) c = kTiny;
        d = rcp(d);
        double del = d * c;
        h *= del;
        if (fabs(del - 1.) < kEps) break;
    }
    return 1. - exp(lpref) * h;
}

// Regularized lower incomplete gamma P(a, x), a > 0.
__device__ inline double gammainc(double a, double x) { return gammainc_lg(a, x, lgamma(a)); }

// Inverse of P(a, .) at probability p.
__device__ inline double gammaincinv(double a, double p)
{
    if (p <= 0.) return 0.;
    if (p >= 1.) return fmax(100., a + 100. * sqrt(a));
    double lga = lgamma(a);
    double x;
    if (a > 1.) {
        // Far lower tail (x << a): P(a, x) = x^a e^-x / Gamma(a + 1) (1 + x / (a + 1) + ...), so
        // log x = (log p + lgamma(a + 1) + x - log(1 + x / (a + 1))) / a, two fixed-point passes
        // from x = (p Gamma(a + 1))^(1/a).  Elsewhere Wilson-Hilferty.
        const double lp = log(p) + lgamma(a + 1.);
        x = exp(lp / a);
        if (x < 0.2 * a) {
            x = exp((lp + x - log1p(x / (a + 1.))) / a);
            x = exp((lp + x - log1p(x / (a + 1.))) / a);
        } else {
            double z = norm_quantile_rough(p);
            double s = 1. - 1. / (9. * a) + z / (3. * sqrt(a));
            x = a * s * s * s;
            if (x < 1e-3) x = 1e-3;
        }
    } else {
        // P(a, x) ~ x^a / Gamma(a + 1) for small x.
        x = exp((log(p) + lgamma(a + 1.)) / a);
        if (!(x > 0.)) x = 1e-300;
    }
    double lo = 0., hi = INFINITY;
    for (int it = 0; it < 200; ++it) {
        double f = gammainc_lg(a, x, lga) - p;
        if (f == 0.) return x;
        if (f > 0.) hi = x;
        else lo = x;
        double dens = exp((a - 1.) * log(x) - x - lga);  // dP/dx
        double xn;
        if (dens > 0. && isfinite(dens)) {
            double u = f / dens;
            double corr = 1. - 0.5 * u * ((a - 1.) / x - 1.);
            double stepv = (corr > 0.5 && corr < 2.) ? u / corr : u;
            xn = x - stepv;
            // converged; tested before the bracket, which a step rounding to zero lands on
            if (fabs(stepv) <= kStepTol * fabs(x)) return xn;
        } else {
            xn = NAN;
        }
        if (!(xn > lo && xn < hi)) xn = isinf(hi) ? 2. * x + 1e-300 : 0.5 * (lo + hi);
        if (fabs(xn - x) <= 4. * kEps * fabs(xn)) return xn;
        if (!isinf(hi) && (hi - lo) <= 4. * kEps * hi) return 0.5 * (lo + hi);
        x = xn;
    }
    return x;
}

// Lentz continued fraction for I_x(a, b) (used where x < (a + 1)/(a + b + 2)).
__device__ inline double betacf(double a, double b, double x)
{
    double qab = a + b, qap = a + 1., qam = a - 1.;
    double c = 1., d = 1. - qab * x / qap;
    if (fabs(d) < kTiny) d = kTiny;
    d = 1. / d;
    double h = d;
    for (int m = 1; m < 5000; ++m) {
        int m2 = 2 * m;
        double aa = m * (b - m) * x * rcp((qam + m2) * (a + m2));
        d = 1. + aa * d;
        if (fabs(d) < kTiny) d = kTiny;
        c = 1. + aa * rcp(c);
        if (fabs(c) < kTiny) c = kTiny;
        d = rcp(d);
        h *= d * c;
        aa = -(a + m) * (qab + m) * x * rcp((a + m2) * (qap + m2));
        d = 1. + aa * d;
        if (fabs(d) < kTiny) d = kTiny;
        c = 1. + aa * rcp(c);
        if (fabs(c) < kTiny) c = kTiny;
        d = rcp(d);
        double del = d * c;
        h *= del;
        if (fabs(del - 1.) < kEps) break;
    }
    return h;
}

// Regularized incomplete beta I_x(a, b), given lab = lgamma(a + b) - lgamma(a) - lgamma(b).
__device__ inline double betainc_lg(double a, double b, double x, double lab)
{
    if (x <= 0.) return 0.;
    if (x >= 1.) return 1.;
    double lbt = lab + a * log(x) + b * log1p(-x);
    if (x < (a + 1.) / (a + b + 2.)) return exp(lbt) * betacf(a, b, x) / a;
    return 1. - exp(lbt) * betacf(b, a, 1. - x) / b;
}

// Regularized incomplete beta I_x(a, b).
__device__ inline double betainc(double a, double b, double x)
{
    return betainc_lg(a, b, x, lgamma(a + b) - lgamma(a) - lgamma(b));
}

// Inverse of I_.(a, b) at probability p.
__device__ inline double betaincinv(double a, double b, double p)
{
    if (p <= 0.) return 0.;
    if (p >= 1.) return 1.;
    double x;
    if (a >= 1. && b >= 1.) {
        // A&S 26.5.22
        double y = norm_quantile_rough(1. - p);
        double lam = (y * y - 3.) / 6.;
        double h = 2. / (1. / (2. * a - 1.) + 1. / (2. * b - 1.));
        double w = y * sqrt(h + lam) / h -
                   (1. / (2. * b - 1.) - 1. / (2. * a - 1.)) * (lam + 5. / 6. - 2. / (3. * h));
        x = a / (a + b * exp(2. * w));
    } else {
        double lna = log(a / (a + b)), lnb = log(b / (a + b));
        double t = exp(a * lna) / a, u = exp(b * lnb) / b, w = t + u;
        if (p < t / w) x = pow(a * w * p, 1. / a);
        else x = 1. - pow(b * w * (1. - p), 1. / b);
    }
    if (!(x > 0. && x < 1.)) x = 0.5;
    double lbeta = lgamma(a) + lgamma(b) - lgamma(a + b);
    const double lab = lgamma(a + b) - lgamma(a) - lgamma(b);
    double lo = 0., hi = 1.;
    for (int it = 0; it < 300; ++it) {
        double f = betainc_lg(a, b, x, lab) - p;
        if (f == 0.) return x;
        if (f > 0.) hi = x;
        else lo = x;
        double dens = exp((a - 1.) * log(x) + (b - 1.) * log1p(-x) - lbeta);
        double xn;
        if (dens > 0. && isfinite(dens)) {
            double u = f / dens;
            double corr = 1. - 0.5 * u * ((a - 1.) / x - (b - 1.) / (1. - x));
            double stepv = (corr > 0.5 && corr < 2.) ? u / corr : u;
            xn = x - stepv;
            // converged; tested before the bracket, which a step rounding to zero lands on
            if (fabs(stepv) <= kStepTol * fmax(fabs(x), 1e-300)) return xn;
        } else {
            xn = NAN;
        }
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 4. * kEps * fmax(fabs(xn), 1e-300)) return xn;
        if ((hi - lo) <= 4. * kEps * hi) return 0.5 * (lo + hi);
        x = xn;
    }
    return x;
}

// Binomial probabilities C(10, k) p^k (1-p)^(10-k) (gsl_ran_binomial_pdf's value for the
// emitter's m = 10), with p^k and (1-p)^j formed by repeated multiplication in a fixed order:
// pk[k] = (...((1 * p) * p)...) * p, k factors.  Evaluated as (C * pk[k]) * qk[10 - k].
struct Binom10 {
    double pk[11], qk[11];
    __device__ explicit Binom10(double p)
    {
        const double q = 1. - p;
        pk[0] = 1.;
        qk[0] = 1.;
#pragma unroll
        for (int i = 1; i <= 10; ++i) {
            pk[i] = pk[i - 1] * p;
            qk[i] = qk[i - 1] * q;
        }
    }
    __device__ double pmf(int k) const
    {
        const double c10[11] = {1., 10., 45., 120., 210., 252., 210., 120., 45., 10., 1.};
        return c10[k] * pk[k] * qk[10 - k];
    }
};

}  // namespace sf
