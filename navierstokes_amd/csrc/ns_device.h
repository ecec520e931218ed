// ns_device.h -- coefficient tables and device building blocks shared by the
// gfx950 kernels of the time-advance hot path.
//
// Tables are returned by constexpr functions and always indexed with
// compile-time constants (unrolled loops), so they fold into instruction
// immediates: no constant-memory loads, no scratch.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace nsamd {

// Tam-Webb DRP interior stencil, init.f90:288-290
constexpr double DRP_A1 = 0.770882380518;
constexpr double DRP_A2 = -0.166705904415;
constexpr double DRP_A3 = 0.0208431427703;

// One-sided 7-point DRP stencils, init.f90:292-338.  Rows 1,2,3 use
// a_dudx6, a_dudx5, a_dudx4 on u(1..7); rows imax-2, imax-1, imax use
// a_dudx4b, a_dudx5b, a_dudx6b on u(imax-6..imax) (diff.f90:49-125).
// drp_left(r, m): row r in {0,1,2}, coefficient of u(m), m = 0..6.
__host__ __device__ constexpr double drp_left(int r, int m)
{
    constexpr double t[3][7] = {
        {-2.19228033900, 4.74861140100, -5.10885191500, 4.46156710400, -2.83349874100, 1.12832886100,
         -0.20387637100},
        {-0.20933762200, -1.08487567600, 2.14777605000, -1.38892832200, 0.76894976600, -0.28181465000,
         0.048230454000},
        {0.049041958000, -0.46884035700, -0.47476091400, 1.27327473700, -0.51848452600, 0.16613853300,
         -0.026369431000}};
    return t[r][m];
}
// drp_right(r, m): row n-3+r, coefficient of u(n-7+m), m = 0..6 (a_dudx*b(m+1))
__host__ __device__ constexpr double drp_right(int r, int m)
{
    constexpr double t[3][7] = {
        {0.026369431000, -0.16613853300, 0.51848452600, -1.27327473700, 0.47476091400, 0.46884035700,
         -0.049041958000},
        {-0.048230454000, 0.28181465000, -0.76894976600, 1.38892832200, -2.14777605000, 1.08487567600,
         0.20933762200},
        {0.20387637100, -1.12832886100, 2.83349874100, -4.46156710400, 5.10885191500, -4.74861140100,
         2.19228033900}};
    return t[r][m];
}

// Bogey-Bailly 13-point selective filter, filtering.f90:54-66 (index s+6)
__host__ __device__ constexpr double filt_d(int s)
{
    constexpr double t[13] = {0.001254597714, -0.008520738659, 0.029662754736, -0.069975429105,
                              0.123632891797, -0.171503832236, 0.190899511506, -0.171503832236,
                              0.123632891797, -0.069975429105, 0.029662754736, -0.008520738659,
                              0.001254597714};
    return t[s];
}
// 7-point off-centre boundary filter dfb(-1..5) (index s+1).  x direction:
// FP64 literals (filtering.f90:25-31).  y direction: filtering.f90:114-120
// writes the same literals without a kind suffix, so the reference's y filter
// runs with their float32 roundings widened to FP64.
template <bool YDIR>
__host__ __device__ constexpr double filt_dfb(int s)
{
    constexpr double tx[7] = {-0.057717512738, 0.199278374994, -0.292668277650, 0.244537361546,
                              -0.134605018019, 0.056184263460, -0.015009191593};
    constexpr double ty[7] = {(double)-0.057717512738f, (double)0.199278374994f,
                              (double)-0.292668277650f, (double)0.244537361546f,
                              (double)-0.134605018019f, (double)0.056184263460f,
                              (double)-0.015009191593f};
    return YDIR ? ty[s] : tx[s];
}

// Berland LDDRK6 2N-storage tables, rk.f90:27-46 (host side)
constexpr double RK_ALPHA[6] = {0.0, -0.737101392796, -1.634740794341,
                                -0.744739003780, -1.469897351522, -2.813971388035};
constexpr double RK_BETA[6] = {0.032918605146, 0.823256998200, 0.381530948900,
                               0.200092213184, 1.718581042715, 0.27};
constexpr double RK_C[6] = {0.0, 0.032918605146, 0.249351723343,
                            0.466911705055, 0.582030414044, 0.847252983783};

template <int ROW, typename R, typename F>
__device__ __forceinline__ R drp_left_row(const F &f, R inv_h)
{
    R s = (R(drp_left(ROW, 0)) * f(0)) * inv_h;
#pragma unroll
    for (int m = 1; m < 7; ++m) s = s + (R(drp_left(ROW, m)) * f(m)) * inv_h;
    return s;
}

template <int ROW, typename R, typename F>
__device__ __forceinline__ R drp_right_row(const F &f, int n, R inv_h)
{
    // reference order: terms m = 7..1 (descending)
    R s = (R(drp_right(ROW, 6)) * f(n - 1)) * inv_h;
#pragma unroll
    for (int m = 5; m >= 0; --m) s = s + (R(drp_right(ROW, m)) * f(n - 7 + m)) * inv_h;
    return s;
}

// Interior DRP first derivative at q (3 <= q <= n-4), diff.f90:40-47.
template <typename R, typename F>
__device__ __forceinline__ R drp_interior(const F &f, int q, R inv_h)
{
    const R t1 = R(DRP_A1) * (f(q + 1) - f(q - 1));
    const R t2 = R(DRP_A2) * (f(q + 2) - f(q - 2));
    const R t3 = R(DRP_A3) * (f(q + 3) - f(q - 3));
    return (t1 * inv_h + t2 * inv_h) + t3 * inv_h;
}

// First derivative at position q of a line of n points (0-based), the line
// read through accessor f(pos).  Interior: DRP (diff.f90:40-47); positions
// 0..2 and n-3..n-1: one-sided 7-point stencils (diff.f90:49-125).  The
// caller multiplies by the metric.
template <typename R, typename F>
__device__ __forceinline__ R drp_deriv(const F &f, int q, int n, R inv_h)
{
    if (q >= 3 && q <= n - 4) return drp_interior<R>(f, q, inv_h);
    if (q == 0) return drp_left_row<0, R>(f, inv_h);
    if (q == 1) return drp_left_row<1, R>(f, inv_h);
    if (q == 2) return drp_left_row<2, R>(f, inv_h);
    if (q == n - 3) return drp_right_row<0, R>(f, n, inv_h);
    if (q == n - 2) return drp_right_row<1, R>(f, n, inv_h);
    return drp_right_row<2, R>(f, n, inv_h);
}

// Natural logarithm and exponential for the transport coefficients (FP64).
// The arguments are bounded (a positive normal temperature times Tref; an
// exponent of a few tens), so neither needs the special-case handling of a
// general library routine, which costs the FP64 log about 100 and the exp
// about 45 instructions.  Both are textbook methods with explicit FMAs,
// within 1.5 ulp (log) and 1 ulp (exp) of the correctly rounded result
// (tests/transport_math_check.cpp measures them against the C library).
//
// ln2 split (Cody & Waite): LN2_HI keeps a 32-bit significand, so e * LN2_HI
// is exact for |e| < 2^21; LN2_HI + LN2_LO = ln 2 to 1e-26.
constexpr double NS_LN2_HI = 0.6931471803691238;
constexpr double NS_LN2_LO = 1.9082149292705877e-10;
constexpr double NS_INV_LN2 = 1.4426950408889634;

// 1 / d for d in [1.7, 2.5): a hardware estimate (the host: the division)
// refined by two Newton steps
__host__ __device__ inline double ns_recip(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
#else
    double r = 1.0 / d;
#endif
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// log x for a positive normal x: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.1716, and
// 2 atanh(s) = 2s + 2s (s^2/3 + s^4/5 + ... + s^18/19) (the next term is
// below 3e-17 of the sum)
__host__ __device__ inline double ns_log(double x)
{
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.70710678118654752) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;         // exact
    const double d = m + 1.0;         // rounded: m + 1 = d + dl exactly,
    const double dl = m - (d - 1.0);  // both differences exact (Sterbenz)
    const double r = ns_recip(d);
    double s = f * r;
    s = fma(fma(-s, dl, fma(-s, d, f)), r, s);  // the residual f - s (d + dl) folded back
    const double z = s * s;
    double p = 1.0 / 19.0;
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double lm = fma(s2 * z, p, s2);
    const double de = (double)e;
    return fma(de, NS_LN2_HI, fma(de, NS_LN2_LO, lm));
}

// exp a for |a| < 700: a = n ln2 + r, |r| <= ln2/2, e^r by its Taylor
// polynomial through r^13 (the remainder is below 5e-18 of e^r)
__host__ __device__ inline double ns_exp(double a)
{
    const double n = rint(a * NS_INV_LN2);
    double r = fma(-n, NS_LN2_HI, a);
    r = fma(-n, NS_LN2_LO, r);
    double p = 1.0 / 6227020800.0;          // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);       // 1/12!
    p = fma(p, r, 1.0 / 39916800.0);        // 1/11!
    p = fma(p, r, 1.0 / 3628800.0);         // 1/10!
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// FP32 (the perturbation form): the hardware base-2 log and exp.  Their
// result differs from a correctly rounded one by a few parts in 1e7 for the
// exponents the coefficients take (|a| of order 10), far inside the form's
// 1e-4 error bar; mu and k only scale the viscous terms.
template <typename R>
__device__ __forceinline__ R tr_log(R x)
{
    if constexpr (sizeof(R) == 8) return ns_log(x);
    else return __builtin_amdgcn_logf(x) * 0.69314718f;
}
template <typename R>
__device__ __forceinline__ R tr_exp(R x)
{
    if constexpr (sizeof(R) == 8) return ns_exp(x);
    else return __builtin_amdgcn_exp2f(x * 1.44269504f);
}

// Transport coefficients, coefficients.f90:35-42 and :50-58:
//   mu = exp(A ln(Tref T) + B/(Tref T) + C/(Tref T)^2 + D) / MI0 (same for k).
// B/Tref, C/Tref^2 and 1/MI0, 1/K0 are folded on the host; the result agrees
// with the reference's operation order to a few ulp.
template <typename R>
struct Transport {
    R A_mi, Bt_mi, Ct_mi, D_mi, inv_mi0;
    R A_k, Bt_k, Ct_k, D_k, inv_k0;
    R Tref;
};

template <typename R>
__device__ __forceinline__ void transport_eval(const Transport<R> &tc, R T, R &mi, R &kk)
{
    const R lt = tr_log<R>(tc.Tref * T);
    R it;
    if constexpr (sizeof(R) == 8) it = R(1) / T;
    else it = __builtin_amdgcn_rcpf(T);  // the hardware reciprocal, 1 ulp
    mi = tc.inv_mi0 * tr_exp<R>(tc.A_mi * lt + tc.Bt_mi * it + tc.Ct_mi * it * it + tc.D_mi);
    kk = tc.inv_k0 * tr_exp<R>(tc.A_k * lt + tc.Bt_k * it + tc.Ct_k * it * it + tc.D_k);
}

}  // namespace nsamd
