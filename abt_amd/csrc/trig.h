// trig.h -- sin/cos kernels for the diffraction kernels (host + device).
//
// sincos_f64: Cody-Waite reduction by pi/2 with two fma steps (error ~2e-16
//   absolute for |x| up to ~1e8, which covers the k*|r| ~ 1.2e7 rad far-field
//   phase), then polynomials on |x| <= pi/4.  Used by the FP64 kernel.
// sincos_rev_f32: sin/cos(2*pi*f) for a reduced revolution count f in
//   [-0.5, 0.5] (from an FP64 phase reduction), FP32 polynomials on |r| <= 1/8.
//   Used by the mixed-precision kernel.
// Coefficients come from gen_trig_coeffs.py (least squares in 60-digit
// arithmetic on Chebyshev nodes).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

// fp64 S(z), z = x^2, x^3 .. x^13
#define ABT_S1 -1.6666666666666665774e-1
#define ABT_S2 8.3333333333314736334e-3
#define ABT_S3 -1.9841269837307909411e-4
#define ABT_S4 2.7557316322996411462e-6
#define ABT_S5 -2.5051169473163993038e-8
#define ABT_S6 1.5920429339605812411e-10
// fp64 C(z), x^4 .. x^14
#define ABT_C1 4.1666666666666666112e-2
#define ABT_C2 -1.3888888888887729934e-3
#define ABT_C3 2.4801587299115424459e-5
#define ABT_C4 -2.7557317412550660125e-7
#define ABT_C5 2.0876170365636353732e-9
#define ABT_C6 -1.1384108242074115052e-11
// fp32 sin(2 pi r) = r * s(r^2)
#define ABT_SR0 6.2831853071666184464f
#define ABT_SR1 -4.1341702168545227836e+1f
#define ABT_SR2 8.1605208488268080076e+1f
#define ABT_SR3 -7.6698201365296159018e+1f
#define ABT_SR4 4.1482575541750033494e+1f
// fp32 cos(2 pi r) = c(r^2)
#define ABT_CR0 9.9999999997439937953e-1f
#define ABT_CR1 -1.9739208666567780209e+1f
#define ABT_CR2 6.4939319231019713357e+1f
#define ABT_CR3 -8.5443057786058814865e+1f
#define ABT_CR4 5.9225039097985419551e+1f

// fp32 sin(2 pi r) = r * h(r^2), cos(2 pi r) = k(r^2), |r| <= 1/4
// (max abs error in exact arithmetic 6.1e-9 and 8.3e-10)
#define ABT_HR0 6.2831852918370438308f
#define ABT_HR1 -4.1341682323056890963e+1f
#define ABT_HR2 8.1602533454522807081e+1f
#define ABT_HR3 -7.658169219540134564e+1f
#define ABT_HR4 3.9759611649540012305e+1f
#define ABT_KR0 9.9999999955408172259e-1f
#define ABT_KR1 -1.973920819234015311e+1f
#define ABT_KR2 6.4939300053282044217e+1f
#define ABT_KR3 -8.5451541679337708511e+1f
#define ABT_KR4 6.010815923321459004e+1f
#define ABT_KR5 -2.4760335289685230034e+1f

#define ABT_2_OVER_PI 0.63661977236758138243
#define ABT_PIO2_HI 1.5707963267948965580      // double nearest pi/2
#define ABT_PIO2_LO 6.1232339957367658e-17     // pi/2 - ABT_PIO2_HI

namespace abt {

__host__ __device__ inline void sincos_f64(double x, double* s, double* c)
{
    double n = rint(x * ABT_2_OVER_PI);
    double r = fma(-n, ABT_PIO2_HI, x);
    r = fma(-n, ABT_PIO2_LO, r);
    double z = r * r;
    double ps = fma(z, fma(z, fma(z, fma(z, fma(z, ABT_S6, ABT_S5), ABT_S4), ABT_S3), ABT_S2), ABT_S1);
    double sn = fma(r * z, ps, r);
    double pc = fma(z, fma(z, fma(z, fma(z, fma(z, ABT_C6, ABT_C5), ABT_C4), ABT_C3), ABT_C2), ABT_C1);
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double cs = w + (((1.0 - w) - hz) + z * z * pc);
    long q = (long)n & 3;
    double so = (q & 1) ? cs : sn;
    double co = (q & 1) ? sn : cs;
    if (q == 1 || q == 2) co = -co;
    if (q >= 2) so = -so;
    *s = so;
    *c = co;
}

// sin(2 pi f), cos(2 pi f) for |f| <= 0.5 (any f works; the octant split assumes a
// reduced input for full accuracy)
__host__ __device__ inline void sincos_rev_f32(float f, float* s, float* c)
{
    float qf = rintf(4.0f * f);
    float r = fmaf(-0.25f, qf, f);   // exact: |r| <= 1/8
    float z = r * r;
    float ps = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_SR4, ABT_SR3), ABT_SR2), ABT_SR1), ABT_SR0);
    float sn = r * ps;
    float cs = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_CR4, ABT_CR3), ABT_CR2), ABT_CR1), ABT_CR0);
    int q = ((int)qf) & 3;
    float so = (q & 1) ? cs : sn;
    float co = (q & 1) ? sn : cs;
    if (q == 1 || q == 2) co = -co;
    if (q >= 2) so = -so;
    *s = so;
    *c = co;
}

// reduce an FP64 revolution count to its fractional part in [-0.5, 0.5]
__host__ __device__ inline float frac_rev(double t)
{
    return (float)(t - rint(t));
}

}  // namespace abt
