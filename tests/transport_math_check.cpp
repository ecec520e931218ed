// Host-side accuracy check of the FP64 log and exp used by the transport
// coefficients (ns_log, ns_exp in navierstokes_amd/csrc/ns_device.h) against
// the C library evaluated in long double.  Over random arguments covering the
// transport's range (log: x in [1e-3, 1e6]; exp: a in [-60, 60]) and the
// edges of the reduction intervals, prints the largest errors in ulp of the
// double result, and the largest difference (ulp) when the reciprocal
// estimate the Newton steps start from is perturbed by up to 2^-12 (the
// device starts from the hardware estimate).  Exits non-zero beyond 1.5 ulp.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>

#include "ns_device.h"

using namespace nsamd;

static double ulp_err(double got, long double ref)
{
    const double r = (double)ref;
    const double u = std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
    return (double)(std::fabs((long double)got - ref) / (long double)u);
}

// ns_log with the reciprocal estimate perturbed by rel before the Newton steps
static double log_perturbed(double x, double rel)
{
    int e;
    double m = std::frexp(x, &e);
    if (m < 0.70710678118654752) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0, d = m + 1.0, dl = m - (d - 1.0);
    double r = (1.0 / d) * (1.0 + rel);
    double t = std::fma(-d, r, 1.0);
    r = std::fma(r, t, r);
    t = std::fma(-d, r, 1.0);
    r = std::fma(r, t, r);
    double s = f * r;
    s = std::fma(std::fma(-s, dl, std::fma(-s, d, f)), r, s);
    const double z = s * s;
    double p = 1.0 / 19.0;
    p = std::fma(p, z, 1.0 / 17.0);
    p = std::fma(p, z, 1.0 / 15.0);
    p = std::fma(p, z, 1.0 / 13.0);
    p = std::fma(p, z, 1.0 / 11.0);
    p = std::fma(p, z, 1.0 / 9.0);
    p = std::fma(p, z, 1.0 / 7.0);
    p = std::fma(p, z, 1.0 / 5.0);
    p = std::fma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double lm = std::fma(s2 * z, p, s2);
    const double de = (double)e;
    return std::fma(de, NS_LN2_HI, std::fma(de, NS_LN2_LO, lm));
}

int main()
{
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> ul(std::log(1e-3), std::log(1e6)), ue(-60.0, 60.0), up(-1.0, 1.0);
    double lmax = 0, emax = 0, pmax = 0;
    auto check_log = [&](double x) {
        lmax = std::fmax(lmax, ulp_err(ns_log(x), logl((long double)x)));
        const double a = log_perturbed(x, std::ldexp(up(rng), -12));
        pmax = std::fmax(pmax, std::fabs(a - ns_log(x)) / (std::nextafter(std::fabs(a), INFINITY) - std::fabs(a)));
    };
    auto check_exp = [&](double a) { emax = std::fmax(emax, ulp_err(ns_exp(a), expl((long double)a))); };
    for (int i = 0; i < 2000000; ++i) {
        check_log(std::exp(ul(rng)));
        check_exp(ue(rng));
    }
    // reduction edges: m near sqrt(1/2) and 1 (log), a near (k + 1/2) ln 2 (exp)
    for (int k = -30; k <= 30; ++k) {
        for (int j = -64; j <= 64; ++j) {
            check_log(std::ldexp(0.70710678118654752, k) * (1.0 + j * 1e-16));
            check_log(std::ldexp(1.0, k) * (1.0 + j * 1e-16));
            check_exp((k + 0.5) * 0.69314718055994531 + j * 1e-16);
        }
    }
    std::printf("log %.3f exp %.3f perturbed %.3f\n", lmax, emax, pmax);
    return (lmax <= 1.5 && emax <= 1.5 && pmax <= 1.0) ? 0 : 1;
}
