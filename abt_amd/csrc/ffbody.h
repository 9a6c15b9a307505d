// ffbody.h -- per (beam, bin) bodies of the far-field kernels (far_field.hip).
// __host__ __device__ so that the same code also runs in the CPU emulator
// used by the tests (tests/emu); on the device abt_rsqf is the hardware
// reciprocal square root.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "records.h"
#include "trig.h"

namespace abt {

__host__ __device__ __forceinline__ float abt_rsqf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

struct BinIn {
    double x, y, z;      // bin position
    double xt, yt, zt;   // bin (same phi, theta_i): forward fallback of the hc vector
};

// ---------------------------------------------------------------------------
// FP64 plain kernel body: one beam, one bin.  acc: re/im of (1,1) (2,1) (1,2) (2,2)
__host__ __device__ __forceinline__ void beam_bin_fp64(const double* __restrict__ r,
                                              const double* __restrict__ vrecs,
                                              const DevGrid& g, const BinIn& bi, double cp,
                                              double sp, int fast_diff, double acc[8])
{
#pragma clang fp contract(off)
    const double x1 = bi.x - r[R_COM], y1 = bi.y - r[R_COM + 1], z1 = bi.z - r[R_COM + 2];
    const double r1 = sqrt(x1 * x1 + y1 * y1 + z1 * z1);
    if (fast_diff) {
        double vx = x1 / r1, vy = y1 / r1, vz = z1 / r1;
        if (vx * r[R_PROP0] + vy * r[R_PROP0 + 1] + vz * r[R_PROP0 + 2] < r[R_COSFOV]) return;
    }
    const double* R = r + R_RR;
    const double x3 = R[0] * x1 + R[1] * y1 + R[2] * z1;
    const double y3 = R[3] * x1 + R[4] * y1 + R[5] * z1;
    const double z3 = R[6] * x1 + R[7] * y1 + R[8] * z1;
    // karczewski (diff_mod.f90:145-181)
    double k1 = x3 / r1, k2 = y3 / r1, k3 = z3 / r1;
    if (fabs(k2) > K2CLAMP) k2 = copysign(K2CLAMP, k2);
    const double Kx = r[R_PROP2], Ky = r[R_PROP2 + 1], Kz = r[R_PROP2 + 2];
    const double sq = sqrt(1 - k2 * k2);
    const double m1 = -(k1 * k2 / sq), m2 = sq, m3 = -(k2 * k3 / sq);
    const double frac = sqrt((1 - k2 * k2) / (1 - Ky * Ky));
    const double a1m = -(Kz * frac);
    const double b2m = -(k3 / frac);
    const double b1m = -(k1 * k2 / frac) + Kx * Ky * frac;
    const double d11 = 0.5 * (a1m + b2m), d12 = 0.5 * b1m, d21 = 0.5 * (-b1m),
                 d22 = 0.5 * (b2m + a1m);
    // scattering-plane vector hc (diff_mod.f90:630-635)
    const double i1 = r[R_INC2], i2 = r[R_INC2 + 1], i3 = r[R_INC2 + 2];
    double kk1 = k1, kk2 = k2, kk3 = k3;
    if (!(fabs(i1 * k1 + i2 * k2 + i3 * k3) < FWD_DOT)) {
        const double xt1 = bi.xt - r[R_COM], yt1 = bi.yt - r[R_COM + 1], zt1 = bi.zt - r[R_COM + 2];
        const double rt = sqrt(xt1 * xt1 + yt1 * yt1 + zt1 * zt1);
        kk1 = (R[0] * xt1 + R[1] * yt1 + R[2] * zt1) / rt;
        kk2 = (R[3] * xt1 + R[4] * yt1 + R[5] * zt1) / rt;
        kk3 = (R[6] * xt1 + R[7] * yt1 + R[8] * zt1) / rt;
    }
    double h1 = i2 * kk3 - i3 * kk2, h2 = i3 * kk1 - i1 * kk3, h3 = i1 * kk2 - i2 * kk1;
    const double nh = sqrt(h1 * h1 + h2 * h2 + h3 * h3);
    h1 = h1 / nh; h2 = h2 / nh; h3 = h3 / nh;
    double e1 = k2 * m3 - k3 * m2, e2 = k3 * m1 - k1 * m3, e3 = k1 * m2 - k2 * m1;
    const double ne = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
    e1 = e1 / ne; e2 = e2 / ne; e3 = e3 / ne;
    const double p = h1 * m1 + h2 * m2 + h3 * m3;     // rot4(1,1) = rot4(2,2)
    const double qh = h1 * e1 + h2 * e2 + h3 * e3;    // rot4(2,1); rot4(1,2) = -qh
    // amplC = rot4 . (diff_ampl . (ampl . temp_rot1)), temp_rot1 = [[-s,-c],[c,-s]]
    const double* A = r + R_AMPL;
    const double a11r = A[0], a11i = A[1], a21r = A[2], a21i = A[3];
    const double a12r = A[4], a12i = A[5], a22r = A[6], a22i = A[7];
    const double t11 = -sp, t12 = -cp, t21 = cp, t22 = -sp;
    const double at11r = a11r * t11 + a12r * t21, at11i = a11i * t11 + a12i * t21;
    const double at12r = a11r * t12 + a12r * t22, at12i = a11i * t12 + a12i * t22;
    const double at21r = a21r * t11 + a22r * t21, at21i = a21i * t11 + a22i * t21;
    const double at22r = a21r * t12 + a22r * t22, at22i = a21i * t12 + a22i * t22;
    const double dt11r = d11 * at11r + d12 * at21r, dt11i = d11 * at11i + d12 * at21i;
    const double dt12r = d11 * at12r + d12 * at22r, dt12i = d11 * at12i + d12 * at22i;
    const double dt21r = d21 * at11r + d22 * at21r, dt21i = d21 * at11i + d22 * at21i;
    const double dt22r = d21 * at12r + d22 * at22r, dt22i = d21 * at12i + d22 * at22i;
    const double mq = -qh;
    const double c11r = p * dt11r + mq * dt21r, c11i = p * dt11i + mq * dt21i;
    const double c12r = p * dt12r + mq * dt22r, c12i = p * dt12i + mq * dt22i;
    const double c21r = qh * dt11r + p * dt21r, c21i = qh * dt11i + p * dt21i;
    const double c22r = qh * dt12r + p * dt22r, c22i = qh * dt12i + p * dt22i;
    // contour_integral (diff_mod.f90:357-421) for this bin
    const double w = g.waveno, lam = g.lambda;
    const double bvk = w * sqrt(x3 * x3 + y3 * y3 + z3 * z3);
    double kxx = Kx * w - w * w * x3 / bvk;
    double kyy = Ky * w - w * w * y3 / bvk;
    if (fabs(kxx) < KMIN) kxx = KMIN;
    if (fabs(kyy) < KMIN) kyy = KMIN;
    double sb, cb;
    sincos_f64(bvk, &sb, &cb);
    double ar = 0.0, ai = 0.0;
    const int nv = (int)r[R_NV];
    const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
    for (int j = 0; j < nv; j++) {
        double xj, yj, mj, nj, dx, dy;
        if (nv <= 3) {
            xj = r[R_EX + j]; yj = r[R_EY + j]; mj = r[R_EM + j]; nj = r[R_EN + j];
            dx = r[R_EDX + j]; dy = r[R_EDY + j];
        } else {
            const double* v = vr + VREC * j;
            xj = v[V_X]; yj = v[V_Y]; mj = v[V_M]; nj = v[V_N]; dx = v[V_DX]; dy = v[V_DY];
        }
        const double delta = kxx * xj + kyy * yj;
        const double delta1 = kyy * mj + kxx;
        const double delta2 = kxx * nj + kyy;
        const double omega1 = dx * delta1;
        const double omega2 = dy * delta2;
        const double alpha = 1 / (2 * kyy * delta1);
        const double beta = 1 / (2 * kxx * delta2);
        double s0, c0, s1, c1, s2, c2;
        sincos_f64(delta, &s0, &c0);
        sincos_f64(delta + omega1, &s1, &c1);
        sincos_f64(delta + omega2, &s2, &c2);
        const double sumim = alpha * (c0 - c1) - beta * (c0 - c2);
        const double sumre = -(alpha * (s0 - s1)) + beta * (s0 - s2);
        ar = ar + (cb * sumre - sb * sumim) / lam;
        ai = ai + (cb * sumim + sb * sumre) / lam;
    }
    acc[0] = acc[0] + (c11r * ar - c11i * ai); acc[1] = acc[1] + (c11r * ai + c11i * ar);
    acc[2] = acc[2] + (c21r * ar - c21i * ai); acc[3] = acc[3] + (c21r * ai + c21i * ar);
    acc[4] = acc[4] + (c12r * ar - c12i * ai); acc[5] = acc[5] + (c12r * ai + c12i * ar);
    acc[6] = acc[6] + (c22r * ar - c22i * ai); acc[7] = acc[7] + (c22r * ai + c22i * ar);
}

// ---------------------------------------------------------------------------
// Mixed-precision kernel body.

// sin/cos(2 pi t) for an FP64 revolution count t: quadrant reduction in FP64,
// FP32 polynomial on the reduced octant (|r| <= 1/8 rev).
__host__ __device__ __forceinline__ void sincos_rev64(double t, float* s, float* c)
{
    const double q = rint(4.0 * t);
    const float r = (float)fma(-0.25, q, t);
    const float z = r * r;
    const float ps = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_SR4, ABT_SR3), ABT_SR2), ABT_SR1), ABT_SR0);
    const float sn = r * ps;
    const float cs = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_CR4, ABT_CR3), ABT_CR2), ABT_CR1), ABT_CR0);
    const int qi = ((int)q) & 3;
    float so = (qi & 1) ? cs : sn;
    float co = (qi & 1) ? sn : cs;
    co = (qi == 1 || qi == 2) ? -co : co;
    so = (qi >= 2) ? -so : so;
    *s = so;
    *c = co;
}

// sinc(pi*hh) = sin(pi*hh)/(pi*hh) ... expressed in revolutions: for an edge with
// half-phase h (revolutions), returns sin(2 pi h) / (2 pi h), and e^{2 pi i mu}.
__host__ __device__ __forceinline__ float sinc_rev64(double h)
{
    const double q = rint(4.0 * h);
    const float r = (float)fma(-0.25, q, h);
    const float z = r * r;
    const float ps = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_SR4, ABT_SR3), ABT_SR2), ABT_SR1), ABT_SR0);
    const float cs = fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_CR4, ABT_CR3), ABT_CR2), ABT_CR1), ABT_CR0);
    const int qi = ((int)q) & 3;
    float sn = (qi & 1) ? cs : r * ps;
    sn = (qi >= 2) ? -sn : sn;
    const float hf = (float)h;
    // |h| <= 1/8: sin(2 pi r)/(2 pi r) = ps / (2 pi) exactly, no cancellation
    const float small = ps * 0.15915494309189535f;
    const float big = sn / (6.283185307179586f * hf);
    return (q == 0.0) ? small : big;
}

constexpr double SERIES_QA = 1.5;    // |q_perp| * a (rad) below which the FP64 series is used
constexpr int SERIES_N = 24;         // truncation error < 1e-19 relative at |q| a = 1.5

// 2 * sum_n i^n h_n(a,b,c) / (n+2)!  (triangle form factor / area, vertex phases a,b,c rad)
__host__ __device__ __forceinline__ void tri_series(double a, double b, double c, double* re, double* im)
{
    const double e1 = a + b + c;
    const double e2 = a * b + b * c + c * a;
    const double e3 = a * b * c;
    double h0 = 1.0, h1 = e1, h2 = e1 * e1 - e2;
    double inv = 0.5;             // 1/(n+2)! at n = 0
    double sr = h0 * inv, si = 0.0;
    inv = inv / 3.0;              // n = 1
    si += h1 * inv;
    inv = inv / 4.0;              // n = 2
    sr -= h2 * inv;
    double hm3 = h0, hm2 = h1, hm1 = h2;
#pragma unroll
    for (int n = 3; n <= SERIES_N; n++) {
        const double hn = e1 * hm1 - e2 * hm2 + e3 * hm3;
        inv = inv / (double)(n + 2);
        const double t = hn * inv;
        switch (n & 3) {
            case 0: sr += t; break;
            case 1: si += t; break;
            case 2: sr -= t; break;
            default: si -= t; break;
        }
        hm3 = hm2; hm2 = hm1; hm1 = hn;
    }
    *re = 2.0 * sr;
    *im = 2.0 * si;
}


__host__ __device__ __forceinline__ void beam_bin_mixed(const double* __restrict__ r,
                                               const double* __restrict__ vrecs,
                                               const DevGrid& g, const BinIn& bi,
                                               int fast_diff, double acc[8])
{
    // ---- FP64 geometry ----
    const double dx = bi.x - r[R_COM], dy = bi.y - r[R_COM + 1], dz = bi.z - r[R_COM + 2];
    const double s = fma(dx, dx, fma(dy, dy, dz * dz));
    double y0 = (double)abt_rsqf((float)s);
    double e = fma(-s * y0, y0, 1.0);
    y0 = fma(0.5 * y0, e, y0);
    e = fma(-s * y0, y0, 1.0);
    const double invd = fma(0.5 * y0, e, y0);          // 1/|r - com|
    const double ux = dx * invd, uy = dy * invd, uz = dz * invd;   // bin direction
    if (fast_diff) {
        if (ux * r[R_PROP0] + uy * r[R_PROP0 + 1] + uz * r[R_PROP0 + 2] < r[R_COSFOV]) return;
    }
    const double krev = g.krev;                            // k / (2 pi)
    const double dist = s * invd;
    const double* R = r + R_RR;
    const double k1 = fma(R[0], ux, fma(R[1], uy, R[2] * uz));
    const double k2 = fma(R[3], ux, fma(R[4], uy, R[5] * uz));
    const double k3 = fma(R[6], ux, fma(R[7], uy, R[8] * uz));
    // q in the aperture plane, revolutions per um
    const double qx = krev * (r[R_PROP2] - k1);
    const double qy = krev * (r[R_PROP2 + 1] - k2);
    float fsb, fcb;
    sincos_rev64(krev * dist, &fsb, &fcb);               // e^{i k |r - com|}
    const float fqx = (float)qx, fqy = (float)qy;
    const float q2 = fmaf(fqx, fqx, fqy * fqy);
    const double qa = 6.283185307179586 * sqrt((double)q2) * r[R_AMAX];

    // ---- aperture integral: sum_ref = i F for the clockwise aperture frame ----
    float fr_, fi_;
    const int nv = (int)r[R_NV];
    const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
    if (qa < SERIES_QA) {
        // FP64 series of the form factor about q = 0; sum_ref = -i * (signed-area F)
        const double tw = 6.283185307179586;
        double sr = 0.0, si = 0.0;
        if (nv == 3) {
            const double a = tw * fma(qx, r[R_EX], qy * r[R_EY]);
            const double b = tw * fma(qx, r[R_EX + 1], qy * r[R_EY + 1]);
            const double c = tw * fma(qx, r[R_EX + 2], qy * r[R_EY + 2]);
            double tr, ti;
            tri_series(a, b, c, &tr, &ti);
            sr = r[R_AREA] * tr;
            si = r[R_AREA] * ti;
        } else {
            for (int j = 0; j < nv; j++) {
                const int jp = (j == nv - 1) ? 0 : j + 1;
                const double xj = vr[VREC * j + V_X], yj = vr[VREC * j + V_Y];
                const double xn = vr[VREC * jp + V_X], yn = vr[VREC * jp + V_Y];
                const double a = tw * fma(qx, xj, qy * yj);
                const double b = tw * fma(qx, xn, qy * yn);
                const double sa = 0.5 * (xj * yn - xn * yj);   // signed area of (com, j, j+1)
                double tr, ti;
                tri_series(0.0, a, b, &tr, &ti);
                sr += sa * tr;
                si += sa * ti;
            }
        }
        fr_ = (float)si;
        fi_ = (float)(-sr);
    } else {
        float ssr = 0.f, ssi = 0.f;
        // half vertex phases (revolutions), FP64
        double hp_prev = 0.0, hp_first = 0.0;
        for (int j = 0; j <= nv; j++) {
            const int jj = (j == nv) ? 0 : j;
            double xj, yj;
            if (nv <= 3) { xj = r[R_EX + jj]; yj = r[R_EY + jj]; }
            else { xj = vr[VREC * jj + V_X]; yj = vr[VREC * jj + V_Y]; }
            const double hp = 0.5 * fma(qx, xj, qy * yj);
            if (j == 0) { hp_first = hp; hp_prev = hp; continue; }
            const double hcur = (j == nv) ? hp_first : hp;
            const double mu = hp_prev + hcur;        // (delta_j + delta_{j+1}) / 2
            const double hh = hcur - hp_prev;        // (delta_{j+1} - delta_j) / 2
            const int je = j - 1;
            double exj, eyj;
            if (nv <= 3) { exj = r[R_EDX + je]; eyj = r[R_EDY + je]; }
            else { exj = vr[VREC * je + V_DX]; eyj = vr[VREC * je + V_DY]; }
            const float cj = fmaf(fqx, (float)eyj, -fqy * (float)exj);
            const float sc = sinc_rev64(hh);
            float sm, cm;
            sincos_rev64(mu, &sm, &cm);
            const float t = cj * sc;
            ssr = fmaf(t, cm, ssr);
            ssi = fmaf(t, sm, ssi);
            hp_prev = hcur;
        }
        // sum_ref = -(1/2pi) * sum / |q|^2
        const float sc = -0.15915494309189535f / q2;
        fr_ = ssr * sc;
        fi_ = ssi * sc;
    }
    // area_facs2 = e^{i bvk} * sum_ref / lambda
    const float il = g.inv_lambda_f;
    const float fR = (fcb * fr_ - fsb * fi_) * il;
    const float fI = (fcb * fi_ + fsb * fr_) * il;

    // ---- polarisation, FP32 ----
    // the |k_y| > 0.999999 clamp (diff_mod.f90:151) is decided in FP64, as the reference does
    const bool kclamp = fabs(k2) > K2CLAMP;
    const float fk1 = (float)k1, fk3 = (float)k3;
    const float fk2 = kclamp ? copysignf(0.99999898672103882f, (float)k2) : (float)k2;
    const float s2 = fmaf(fk1, fk1, fk3 * fk3);                // 1 - k2^2 (unclamped)
    const float s2c = kclamp ? 2.02655792e-06f : s2;           // clamped: 1 - 0.999999^2
    const float rs = abt_rsqf(s2c);
    const float sq = s2c * rs;                                   // sqrt(1 - k2^2)
    const float sig = (float)r[R_FRAC];                          // 1/sqrt(1 - Ky^2)
    const float Kx = (float)r[R_PROP2], Ky = (float)r[R_PROP2 + 1], Kz = (float)r[R_PROP2 + 2];
    const float frac = sq * sig;
    const float ifrac = rs / sig;
    const float da = -0.5f * fmaf(Kz, frac, fk3 * ifrac);        // diff_ampl(1,1) = (2,2)
    const float db = 0.5f * fmaf(Kx * Ky, frac, -fk1 * fk2 * ifrac);  // diff_ampl(1,2) = -(2,1)
    // m and evo2 = unit(k x m) in closed form
    const float m1 = -fk1 * fk2 * rs, m2 = sq, m3 = -fk2 * fk3 * rs;
    const float r13 = abt_rsqf(s2);
    const float ev1 = -fk3 * r13, ev3 = fk1 * r13;
    // hc = R . unit(inc x khat), inc = (0,0,-1): global (u_y, -u_x, 0)
    float hx = (float)uy, hy = (float)(-ux);
    if (!(fabs(uz) < FWD_DOT)) {
        hx = (float)(bi.yt - r[R_COM + 1]);
        hy = (float)(-(bi.xt - r[R_COM]));
    }
    const float rh = abt_rsqf(fmaf(hx, hx, hy * hy));
    hx *= rh; hy *= rh;
    const float h1 = fmaf((float)R[0], hx, (float)R[1] * hy);
    const float h2 = fmaf((float)R[3], hx, (float)R[4] * hy);
    const float h3 = fmaf((float)R[6], hx, (float)R[7] * hy);
    const float p = fmaf(h1, m1, fmaf(h2, m2, h3 * m3));
    const float qh = fmaf(h1, ev1, h3 * ev3);
    // rot4 . diff_ampl = gg I + hh J,  J = [[0,1],[-1,0]]
    const float gg = fmaf(p, da, qh * db);
    const float hhh = fmaf(p, db, -qh * da);
    // weights: acc += u A + v (J A)
    const double ur = (double)(fR * gg), ui = (double)(fI * gg);
    const double vr_ = (double)(fR * hhh), vi = (double)(fI * hhh);
    const double* A = r + R_AMPL;
    // A: (1,1)=A0,A1 (2,1)=A2,A3 (1,2)=A4,A5 (2,2)=A6,A7 ; J A: (1,1)=A(2,1), (2,1)=-A(1,1),
    // (1,2)=A(2,2), (2,2)=-A(1,2)
#define CMAC(o, ar_, ai_, br_, bi_)                                                 \
    acc[o] = fma(ur, ar_, fma(-ui, ai_, fma(vr_, br_, fma(-vi, bi_, acc[o]))));     \
    acc[o + 1] = fma(ur, ai_, fma(ui, ar_, fma(vr_, bi_, fma(vi, br_, acc[o + 1]))));
    CMAC(0, A[0], A[1], A[2], A[3])
    CMAC(2, A[2], A[3], -A[0], -A[1])
    CMAC(4, A[4], A[5], A[6], A[7])
    CMAC(6, A[6], A[7], -A[4], -A[5])
#undef CMAC
}

// The mixed kernel sums sum_b f_b G_b A_b per bin; the phi-only rotation
// temp_rot1 = [[-s,-c],[c,-s]] (diff_mod.f90:646-650) is applied once on the right:
// (X T)(r,1) = -s X(r,1) + c X(r,2);  (X T)(r,2) = -c X(r,1) - s X(r,2).
__host__ __device__ __forceinline__ void apply_temp_rot1(double acc[8], double cp, double sp)
{
    double o[8];
    for (int rr = 0; rr < 2; rr++)
        for (int ri = 0; ri < 2; ri++) {
            const double x1 = acc[2 * rr + ri];        // X(r,1)
            const double x2 = acc[4 + 2 * rr + ri];    // X(r,2)
            o[2 * rr + ri] = fma(-sp, x1, cp * x2);
            o[4 + 2 * rr + ri] = fma(-cp, x1, -sp * x2);
        }
    for (int q = 0; q < 8; q++) acc[q] = o[q];
}

}  // namespace abt
