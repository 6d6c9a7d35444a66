// setup.hpp -- host-side set-up tables of the propagation step: grids,
// stencil weights, FFT radix plans and twiddles, form factors and the
// impurity interpolation table.  All computed once per context.
#pragma once

#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace hetd {
namespace setup {

constexpr long double kPiL = 3.141592653589793238462643383279502884L;

// Centred npd-point weights of the id-th derivative at unit spacing, c[j]
// for offset j - npd/2.  They equal the centred rows of the reference's
// table (newder.f:99-128 over mates.f:172-14043): the Fornberg closed forms
// d_k = (-1)^{k+1} (m!)^2/(k (m-k)! (m+k)!), c_k = 2 d_k/k, c_0 = -2 sum 1/k^2.
inline void fd_weights(int npd, int id, double* c) {
    const int m = npd / 2;
    for (int j = 0; j < npd; ++j) c[j] = 0.0;
    long double s2 = 0.0L;
    for (int k = 1; k <= m; ++k) {
        long double ratio = 1.0L;
        for (int j = 1; j <= k; ++j) ratio *= (long double)(m - k + j) / (long double)(m + j);
        const long double sg = (k & 1) ? 1.0L : -1.0L;
        if (id == 1) {
            const long double d = sg * ratio / (long double)k;
            c[m + k] = (double)d;
            c[m - k] = (double)(-d);
        } else {
            const long double d = 2.0L * sg * ratio / ((long double)k * (long double)k);
            c[m + k] = (double)d;
            c[m - k] = (double)d;
            s2 += 1.0L / ((long double)k * (long double)k);
        }
    }
    if (id == 2) c[m] = (double)(-2.0L * s2);
}

// Radix plan: 4s first, a spare 2 merged with a 4 into an 8, then 2, 3, 5.
// Returns false for a length with a prime factor above 5.
inline bool radix_plan(int n, int* rad, int* nst, int maxst) {
    std::vector<int> f;
    int m = n, n4 = 0, n2 = 0;
    while (m % 4 == 0) { ++n4; m /= 4; }
    while (m % 2 == 0) { ++n2; m /= 2; }
    if (n2 && n4) { f.push_back(8); --n4; --n2; }
    for (int i = 0; i < n4; ++i) f.push_back(4);
    for (int i = 0; i < n2; ++i) f.push_back(2);
    while (m % 3 == 0) { f.push_back(3); m /= 3; }
    while (m % 5 == 0) { f.push_back(5); m /= 5; }
    if (m != 1 || (int)f.size() > maxst) return false;
    *nst = (int)f.size();
    for (size_t i = 0; i < f.size(); ++i) rad[i] = f[i];
    return true;
}

inline void twiddles(int n, std::vector<double>& tw64, std::vector<float>& tw32) {
    tw64.resize(2 * (size_t)n);
    tw32.resize(2 * (size_t)n);
    for (int m = 0; m < n; ++m) {
        const long double a = -2.0L * kPiL * (long double)m / (long double)n;
        tw64[2 * m] = (double)cosl(a);
        tw64[2 * m + 1] = (double)sinl(a);
        tw32[2 * m] = (float)cosl(a);
        tw32[2 * m + 1] = (float)sinl(a);
    }
}

// pi/2 - Si(x), x > 0 (the quantity fforma's sinint returns, fforma.f90:143-196).
// Power series of Si for x <= 4; for x > 4 the continued fraction
// E1(z) = e^{-z}/(z + 1 - 1^2/(z + 3 - 2^2/(z + 5 - ...))) at z = ix by the
// modified Lentz method, with pi/2 - Si(x) = -Im E1(ix).
inline long double sinint(long double x) {
    if (x <= 4.0L) {
        long double term = x, sum = x, x2 = x * x;
        for (int k = 1; k < 60; ++k) {
            term *= -x2 / ((long double)(2 * k) * (long double)(2 * k + 1));
            const long double add = term / (long double)(2 * k + 1);
            sum += add;
            if (fabsl(add) < 1e-22L * fabsl(sum)) break;
        }
        return kPiL / 2.0L - sum;
    }
    long double br = 1.0L, bi = x, cr = 1e300L, ci = 0.0L;
    long double den = br * br + bi * bi;
    long double dr = br / den, di = -bi / den, hr = dr, hi = di;
    for (int i = 2; i < 100000; ++i) {
        const long double a = -(long double)(i - 1) * (long double)(i - 1);
        br += 2.0L;
        const long double tr = a * dr + br, ti = a * di + bi;
        den = tr * tr + ti * ti;
        dr = tr / den;
        di = -ti / den;
        den = cr * cr + ci * ci;
        cr = br + a * cr / den;
        ci = bi - a * ci / den;
        const long double er = cr * dr - ci * di, ei = cr * di + ci * dr;
        const long double nhr = hr * er - hi * ei, nhi = hr * ei + hi * er;
        hr = nhr;
        hi = nhi;
        if (fabsl(er - 1.0L) + fabsl(ei) < 1e-21L) break;
    }
    return -(cosl(x) * hi - sinl(x) * hr);
}

// Fourier transform of the Orsay-Trento-cored Lennard-Jones potential at
// |p| (cycles/A): 8 eps sigma^6 (sigma^6 I11 - I5)/p with
// I_n = int_h^inf sin(2 pi p r)/r^n dr in closed form (fforma.f90:40-137),
// and the p = 0 limit bforce (fforma.f90:256-280).  Long double.
inline double vlj_hat(double p, double eps, double sigma, double hc) {
    const long double s6 = powl((long double)sigma, 6);
    if (std::fabs(p) < 1e-10) {
        const long double pi4 = 4.0L * kPiL, eps4 = 4.0L * eps;
        return (double)(pi4 * eps4 * s6 * s6 / (9.0L * powl((long double)hc, 9)) -
                        pi4 * eps4 * s6 / (3.0L * powl((long double)hc, 3)));
    }
    const long double a = 2.0L * kPiL * (long double)p, x = hc, y = a * x, y2 = y * y;
    const long double si = sinint(y), cy = cosl(y), sy = sinl(y);
    const long double bd = 40320.0L - 720.0L * y2 + 24.0L * y2 * y2 - 2.0L * y2 * y2 * y2 + y2 * y2 * y2 * y2;
    const long double cd = 362880.0L - 5040.0L * y2 + 120.0L * y2 * y2 - 6.0L * y2 * y2 * y2 + y2 * y2 * y2 * y2;
    const long double t1 = (a * bd * cy + cd * sy / x) / (3628800.0L * powl(x, 9)) - powl(a, 10) * si / 3628800.0L;
    const long double t2 = -(a * (y2 - 2.0L) * cy + (y2 - 6.0L) * sy / x) / (24.0L * x * x * x) + powl(a, 4) * si / 24.0L;
    return (double)(8.0L * eps * s6 * (s6 * t1 - t2) / (long double)p);
}

// Coarse-graining kernel W(p) = 3 (sin x - x cos x)/x^3, x = 2 pi p h
// (initcg.f90:32-49), 1 at p = 0 or where the numerator is below 1e-10.
inline double wcg(double p, double h) {
    if (!(p > 0.0)) return 1.0;
    const double twopi = 2.0 * (4.0 * std::atan(1.0));
    const double arg = twopi * p * h;
    const double aux = std::sin(arg) - arg * std::cos(arg);
    return std::fabs(aux) < 1e-10 ? 1.0 : 3.0 * aux / (arg * arg * arg);
}

// Reference p-grid (DFT4He3d.f90:320-349): p(k) = hp*k for k <= n/2, and
// -pmax + hp*(k + 1 - (n/2 + 1)) above (for odd n this is hp*(k - n + 1/2)).
inline double pgrid(int k, int n, double h) {
    const double hp = 1.0 / (n * h), pmax = 1.0 / (2.0 * h);
    return (k + 1 <= n / 2 + 1) ? hp * k : -pmax + hp * (k + 1 - (n / 2 + 1));
}

template <typename F> inline void parallel_for(int n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if ((int)nt > n) nt = (unsigned)n;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=]() { for (int i = (int)t; i < n; i += (int)nt) f(i); });
    for (auto& x : th) x.join();
}

}  // namespace setup
}  // namespace hetd
