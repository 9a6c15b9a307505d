// ffbody.h -- per (beam, bin) bodies of the far-field kernels (far_field.hip).
// __host__ __device__ so that the same code also runs in the CPU emulator
// used by the tests (tests/emu); on the device abt_rsqf is the hardware
// reciprocal square root.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

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
    double rad;          // |bin position| (FP64 model body only)
};

__host__ __device__ __forceinline__ float abt_rcpf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
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

// Half-period reduction of a revolution count t (FP64): t = r + m/2 with
// |r| <= 1/4, so that e^{2 pi i t} = (-1)^m e^{2 pi i r}.  The parity of m is
// returned as an FP32 sign mask (bit 31), which callers fold into one product.
__host__ __device__ __forceinline__ float half_reduce(double t, unsigned* sgn)
{
    const double m = rint(2.0 * t);
    *sgn = ((unsigned)(int)m) << 31;
    return (float)fma(-0.5, m, t);
}

__host__ __device__ __forceinline__ float flipsign(float x, unsigned sgn)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(__float_as_uint(x) ^ sgn);
#else
    unsigned u;
    memcpy(&u, &x, 4);
    u ^= sgn;
    memcpy(&x, &u, 4);
    return x;
#endif
}

// sin(2 pi r) / r and cos(2 pi r) for |r| <= 1/4
__host__ __device__ __forceinline__ float sin_over_r(float z)
{
    return fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_HR4, ABT_HR3), ABT_HR2), ABT_HR1), ABT_HR0);
}
__host__ __device__ __forceinline__ float cos_poly(float z)
{
    return fmaf(z, fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_KR5, ABT_KR4), ABT_KR3), ABT_KR2), ABT_KR1),
                ABT_KR0);
}

// Polynomials in half revolutions: for |x| <= 1/2 and z = x^2,
//   sinpi_over(z) = sin(pi x) / x,  cospi_poly(z) = cos(pi x),  sincpi_poly(z) = sin(pi x) / (pi x);
// the first two are sin_over_r / cos_poly with the argument doubled (coefficients scaled
// by powers of two, so the values are the same to the last bit).
__host__ __device__ __forceinline__ float sinpi_over(float z)
{
    return fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_HR4 / 512.f, ABT_HR3 / 128.f), ABT_HR2 / 32.f), ABT_HR1 / 8.f),
                ABT_HR0 / 2.f);
}
__host__ __device__ __forceinline__ float cospi_poly(float z)
{
    return fmaf(z, fmaf(z, fmaf(z, fmaf(z, fmaf(z, ABT_KR5 / 1024.f, ABT_KR4 / 256.f), ABT_KR3 / 64.f),
                                ABT_KR2 / 16.f), ABT_KR1 / 4.f), ABT_KR0);
}
__host__ __device__ __forceinline__ float sincpi_poly(float z)
{
    constexpr double c = 0.15915494309189535;   // 1 / 2 pi
    return fmaf(z, fmaf(z, fmaf(z, fmaf(z, (float)(ABT_HR4 * c / 256.), (float)(ABT_HR3 * c / 64.)),
                                (float)(ABT_HR2 * c / 16.)), (float)(ABT_HR1 * c / 4.)), (float)(ABT_HR0 * c));
}

// t (FP64, half revolutions: e^{i pi t}) reduced to |x| <= 1/2: e^{i pi t} = (-1)^m e^{i pi x},
// m = rint(t); the parity of m comes back as an FP32 sign mask (bit 31).
__host__ __device__ __forceinline__ float pi_reduce(double t, unsigned* sgn)
{
    const double m = rint(t);
    *sgn = ((unsigned)(int)m) << 31;
    return (float)(t - m);
}

// e^{i pi t} in FP32 from the FP64 t
__host__ __device__ __forceinline__ void expi_pi(double t, float* c, float* s)
{
    unsigned sg;
    const float x = pi_reduce(t, &sg);
    const float z = x * x;
    *s = flipsign(x * sinpi_over(z), sg);
    *c = flipsign(cospi_poly(z), sg);
}

constexpr float SERIES_QA2 = 2.25f;  // (|q_perp| a)^2 below which the FP64 series is used
constexpr int SERIES_N = 16;         // truncation < 3e-12 relative at |q| a = 1.5

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

// One edge of the rotation-invariant form factor with the far-field phase folded in:
// c_j sinc(pi d) e^{i (mu_j + k|r - com|)}, from the vertex exponentials w = e^{i pi t'}
// (t' = q.v + k|r - com|/2pi, half revolutions) of its two ends; d = t'_{j+1} - t'_j.
__host__ __device__ __forceinline__ void edge_exp(float wr0, float wi0, float wr1, float wi1, float d, float fqx,
                                                  float fqy, float edx, float edy, float& ssr, float& ssi)
{
    const float Er = fmaf(wr0, wr1, -wi0 * wi1);
    const float Ei = fmaf(wr0, wi1, wi0 * wr1);
    const float sx = fmaf(wi1, wr0, -wr1 * wi0);        // sin(pi d) = Im(w1 conj(w0))
    const float S = fabsf(d) < 0.5f ? sincpi_poly(d * d) : sx * abt_rcpf(3.14159265358979f * d);
    const float t = fmaf(fqx, edy, -fqy * edx) * S;
    ssr = fmaf(t, Er, ssr);
    ssi = fmaf(t, Ei, ssi);
}

// Mixed-precision body: one beam, one bin.  Adds the beam's contribution to the FP32
// projector sums accf: S+ (accf[0..3]) and S- (accf[4..7]), see mixed_finish; callers
// add accf into FP64 sums every few beams.  INC: also adds the squared Frobenius norm of
// the beam's amplitude, |f|^2 (gg^2 + hh^2) |A|_F^2 (the incoherent sum), to inc.
template <bool INC>
__host__ __device__ __forceinline__ void beam_bin_mixed_t(const double* __restrict__ r,
                                                          const double* __restrict__ vrecs,
                                                          const DevGrid& g, const BinIn& bi,
                                                          int fast_diff, float accf[8], float& inc)
{
    const float* rf = reinterpret_cast<const float*>(r + R_F32);
    const float* rg = reinterpret_cast<const float*>(r + R_G32);
    // ---- FP64 geometry: bin direction from the aperture centre, |r - com| ----
    const double dx = bi.x - r[R_COM], dy = bi.y - r[R_COM + 1], dz = bi.z - r[R_COM + 2];
    const double s = fma(dx, dx, fma(dy, dy, dz * dz));
    double y0 = (double)abt_rsqf((float)s);
    double e = fma(-s * y0, y0, 1.0);
    y0 = fma(0.5 * y0, e, y0);
    e = fma(-s * y0, y0, 1.0);
    const double invd = fma(0.5 * y0, e, y0);          // 1/|r - com|
    const double ux = dx * invd, uy = dy * invd, uz = dz * invd;
    if (fast_diff) {
        if (ux * r[R_PROP0] + uy * r[R_PROP0 + 1] + uz * r[R_PROP0 + 2] < r[R_COSFOV]) return;
    }
    // q = krev (prop2 - khat) in the aperture plane (revolutions per um); krev khat_x,y
    const double* K = r + R_KRR;
    const double qx = fma(-K[0], ux, fma(-K[1], uy, fma(-K[2], uz, r[R_KP2])));
    const double qy = fma(-K[3], ux, fma(-K[4], uy, fma(-K[5], uz, r[R_KP2 + 1])));
    const double k1r = r[R_KP2] - qx, k2r = r[R_KP2 + 1] - qy;
    const double kd = g.krev * (s * invd);               // k |r - com| / 2 pi
    const float fqx = (float)qx, fqy = (float)qy;
    const float q2 = fmaf(fqx, fqx, fqy * fqy);
    const float il = g.inv_lambda_f;

    // ---- aperture integral with the far-field phase: (fR, fI) = e^{i k|r-com|} sum_ref / lambda ----
    float fR, fI;
    const bool tri = rf[F_NV] == 3.0f;
    if (q2 * rf[F_QA2] < SERIES_QA2) {
        // FP64 series of the form factor about q = 0; sum_ref = -i * (signed-area F)
        const double tw = 6.283185307179586;
        double sr = 0.0, si = 0.0;
        if (tri) {
            const double a = tw * fma(qx, r[R_EX], qy * r[R_EY]);
            const double b = tw * fma(qx, r[R_EX + 1], qy * r[R_EY + 1]);
            const double c = tw * fma(qx, r[R_EX + 2], qy * r[R_EY + 2]);
            double tr, ti;
            tri_series(a, b, c, &tr, &ti);
            sr = r[R_AREA] * tr;
            si = r[R_AREA] * ti;
        } else {
            const int nv = (int)rf[F_NV];
            const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
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
        const float fr_ = (float)si, fi_ = (float)(-sr);
        float cb, sb;
        expi_pi(2.0 * kd, &cb, &sb);
        fR = (cb * fr_ - sb * fi_) * il;
        fI = (cb * fi_ + sb * fr_) * il;
    } else {
        // sum_ref = -(1/2pi) sum_j c_j sinc(pi d_j) e^{i mu_j} / |q|^2 (Fraunhofer, rotation-invariant form)
        float ssr = 0.f, ssi = 0.f;
        if (tri) {
            double t[3];
            float wr[3], wi[3];
#pragma unroll
            for (int l = 0; l < 3; l++) {
                t[l] = fma(qx, r[R_EX + l], fma(qy, r[R_EY + l], kd));
                expi_pi(t[l], &wr[l], &wi[l]);
            }
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int jn = j == 2 ? 0 : j + 1;
                edge_exp(wr[j], wi[j], wr[jn], wi[jn], (float)(t[jn] - t[j]), fqx, fqy, rf[F_EDX + j],
                         rf[F_EDY + j], ssr, ssi);
            }
        } else {
            const int nv = (int)rf[F_NV];
            const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
            const double t0 = fma(qx, vr[V_X], fma(qy, vr[V_Y], kd));
            float wr0, wi0;
            expi_pi(t0, &wr0, &wi0);
            double tp = t0;
            float wrp = wr0, wip = wi0;
            for (int j = 0; j < nv; j++) {
                double tn = t0;
                float wrn = wr0, win = wi0;
                if (j < nv - 1) {
                    tn = fma(qx, vr[VREC * (j + 1) + V_X], fma(qy, vr[VREC * (j + 1) + V_Y], kd));
                    expi_pi(tn, &wrn, &win);
                }
                edge_exp(wrp, wip, wrn, win, (float)(tn - tp), fqx, fqy, (float)vr[VREC * j + V_DX],
                         (float)vr[VREC * j + V_DY], ssr, ssi);
                tp = tn;
                wrp = wrn;
                wip = win;
            }
        }
        const float sc = (-0.15915494309189535f * il) * abt_rcpf(q2);
        fR = ssr * sc;
        fI = ssi * sc;
    }

    // ---- polarisation, FP32 (karczewski + rot4 of diff_mod.f90:124-198, 630-643) ----
    // rot4 . diff_ampl = gg I + hh J, J = [[0,1],[-1,0]]; the |k_y| > 0.999999 clamp
    // (diff_mod.f90:151) is decided in FP64, as the reference does
    const bool kclamp = fabs(k2r) > g.krev_k2clamp;
    const bool fwd = !(fabs(uz) < FWD_DOT);
    const float ik = g.inv_krev_f;
    const float fk1 = (float)k1r * ik, fk2 = (float)k2r * ik;
    // k_z in FP64: near k = +-e_y, 1 - k_y^2 = k_x^2 + k_z^2 is small and needs k_x, k_z to
    // relative (not absolute) FP32 precision
    const double* RR = r + R_RR;
    const float fk3 = (float)fma(RR[6], ux, fma(RR[7], uy, RR[8] * uz));
    const float s2 = fmaf(fk1, fk1, fk3 * fk3);                   // 1 - k_y^2 (unclamped)
    float gg, hh;
    if (!kclamp && !fwd) {
        // hc = unit(inc x khat) is normal to khat, so with m and k x m of karczewski:
        // hc.m = h_y / sqrt(1 - k_y^2), hc.unit(k x m) = (hc x khat)_y / sqrt(1 - k_y^2), where
        // |inc x khat| h_y = krow2 . (u_y, -u_x, 0) and |inc x khat| (hc x khat)_y = rr_23 - u_z k_y
        // (aperture frame, inc = (0,0,-1)); FP64 here, both vanish at k = +-e_y
        const double P0 = fma(K[3], uy, -K[4] * ux);       // krev |inc x khat| h_y
        const double Q0 = fma(-uz, k2r, K[5]);             // krev |inc x khat| (hc x khat)_y
        const float fux = (float)ux, fuy = (float)uy;
        const float rh = abt_rsqf(fmaf(fux, fux, fuy * fuy)) * abt_rcpf(s2);
        const float P = (float)P0 * rh, Q = (float)Q0 * rh;
        // diff_ampl(1,1) = (2,2) = A / sqrt(1 - k_y^2), diff_ampl(1,2) = -(2,1) = B / sqrt(1 - k_y^2)
        const float A = fmaf(rg[G_C1], s2, rg[G_C2] * fk3);
        const float B = fmaf(rg[G_C3], s2, rg[G_C2] * (fk1 * fk2));
        gg = fmaf(P, A, Q * B);
        hh = fmaf(P, B, -Q * A);
    } else {
        // clamped k_y, or the forward fallback of hc (diff_mod.f90:630-635): the bin
        // (same phi, theta_i); m and evo2 = unit(k x m) in closed form
        const float fk2c = kclamp ? copysignf(0.99999898672103882f, fk2) : fk2;
        const float s2c = kclamp ? 2.02655679e-06f : s2;        // clamped: 1 - 0.999999^2
        const float rs = abt_rsqf(s2c);
        const float sq = s2c * rs;                                 // sqrt(1 - k2^2)
        const float frac = sq * rf[F_SIG];
        const float ifrac = rs * rf[F_ISIG];
        float da = -0.5f * fmaf(rf[F_KZ], frac, fk3 * ifrac);          // diff_ampl(1,1) = (2,2)
        float db = 0.5f * fmaf(rf[F_KXY], frac, -fk1 * fk2c * ifrac);  // diff_ampl(1,2) = -(2,1)
        {
            // diff_ampl vanishes where k is the mirror image of prop2 in the aperture plane,
            // (k_x, k_y, k_z) = (K_x, K_y, -K_z): for every externally reflected beam that is
            // the forward direction, so thousands of beams share the zero there.  Through
            // dk = (k_x - K_x, k_y - K_y, k_z + K_z) (the first two from q, the third in FP64)
            // and u2 = k_y^2 - K_y^2 = dk_y (k_y + K_y), w2 = K_x dk_y + dk_x k_y:
            //   K_z (1 - k_y^2) + k_z (1 - K_y^2) = -K_z u2 + (1 - K_y^2) dk_z,
            //   K_x K_y (1 - k_y^2) - k_x k_y (1 - K_y^2) = -K_x K_y u2 - (1 - K_y^2) w2,
            // whose terms vanish with dk.  Used where |k_y| < 1/sqrt(2) (aperture frames put
            // prop2 in the x-z plane, K_y = 0, so that covers the zero); nearer k_y = +-1 the
            // terms of u2 outgrow 1 - k_y^2 and the direct form is kept.
            const float fdk1 = -fqx * ik, fdk2 = -fqy * ik;
            const float fdk3 = (float)fma(RR[6], ux, fma(RR[7], uy, fma(RR[8], uz, r[R_PROP2 + 2])));
            const float u2 = fdk2 * (fk2 + rf[F_KY]);
            const float w2 = fmaf(rf[F_KX], fdk2, fdk1 * fk2);
            const float dan = 0.5f * rs * fmaf(rf[F_KZ] * rf[F_SIG], u2, -rf[F_ISIG] * fdk3);
            const float dbn = -0.5f * rs * fmaf(rf[F_KXY] * rf[F_SIG], u2, rf[F_ISIG] * w2);
            const bool mir = !kclamp && s2 > 0.5f;
            da = mir ? dan : da;
            db = mir ? dbn : db;
        }
        const float m1 = -fk1 * fk2c * rs, m2 = sq, m3 = -fk2c * fk3 * rs;
        const float r13 = kclamp ? abt_rsqf(s2) : rs;
        const float ev1 = -fk3 * r13, ev3 = fk1 * r13;
        float hx = (float)uy, hy = (float)(-ux);
        if (fwd) {
            hx = (float)(bi.yt - r[R_COM + 1]);
            hy = (float)(-(bi.xt - r[R_COM]));
        }
        const float rh = abt_rsqf(fmaf(hx, hx, hy * hy));
        hx *= rh;
        hy *= rh;
        const float h1 = fmaf(rf[F_RR + 0], hx, rf[F_RR + 1] * hy);
        const float h2 = fmaf(rf[F_RR + 3], hx, rf[F_RR + 4] * hy);
        const float h3 = fmaf(rf[F_RR + 6], hx, rf[F_RR + 7] * hy);
        const float p = fmaf(h1, m1, fmaf(h2, m2, h3 * m3));
        const float qh = fmaf(h1, ev1, h3 * ev3);
        gg = fmaf(p, da, qh * db);
        hh = fmaf(p, db, -qh * da);
    }
    // f (gg I + hh J) = w P + w' Q with the projectors P, Q of J:  w = f (gg + i hh),
    // w' = f (gg - i hh);  P A = [r+; i r+],  Q A = [r-; -i r-]
    const float ta = fR * gg, tb = fI * hh, tc = fR * hh, td = fI * gg;
    const float wr = ta - tb, wi = tc + td, vr_ = ta + tb, vi = td - tc;
    if (INC) inc = fmaf(fmaf(fR, fR, fI * fI) * fmaf(gg, gg, hh * hh), rf[F_C2], inc);
    const float* rp = rg + G_RP;
    const float* rm = rg + G_RM;
    accf[0] = fmaf(wr, rp[0], fmaf(-wi, rp[1], accf[0]));
    accf[1] = fmaf(wr, rp[1], fmaf(wi, rp[0], accf[1]));
    accf[2] = fmaf(wr, rp[2], fmaf(-wi, rp[3], accf[2]));
    accf[3] = fmaf(wr, rp[3], fmaf(wi, rp[2], accf[3]));
    accf[4] = fmaf(vr_, rm[0], fmaf(-vi, rm[1], accf[4]));
    accf[5] = fmaf(vr_, rm[1], fmaf(vi, rm[0], accf[5]));
    accf[6] = fmaf(vr_, rm[2], fmaf(-vi, rm[3], accf[6]));
    accf[7] = fmaf(vr_, rm[3], fmaf(vi, rm[2], accf[7]));
}

__host__ __device__ __forceinline__ void beam_bin_mixed(const double* __restrict__ r,
                                                        const double* __restrict__ vrecs,
                                                        const DevGrid& g, const BinIn& bi,
                                                        int fast_diff, float accf[8])
{
    float unused = 0.f;
    beam_bin_mixed_t<false>(r, vrecs, g, bi, fast_diff, accf, unused);
}

// ---------------------------------------------------------------------------
// Restructured FP64 body: the mixed body's algorithm (bin direction from the aperture
// centre, rotation-invariant edge sum from vertex half-phase exponentials, the hc-normal
// polarisation closed form, projector accumulation) with every step in FP64.  Agrees with
// the reference's formula to ~1e-12 of the largest amplitude (the reference's own direct
// phase k |x3| ~ 1e7 rad carries ~1e-9 rad of rounding), at a fraction of its cost: three
// vertex exponentials per beam and bin instead of the split sum's twelve trigonometric
// pairs.  Used where the FP32 body's error would exceed the accuracy gate (the refinement
// pass of the mixed kernels) and as a fast FP64 precision of its own.

constexpr double PI64 = 3.141592653589793;

// reciprocal and reciprocal square root: hardware approximations refined by Newton steps
__host__ __device__ __forceinline__ double rcp64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
#else
    return 1.0 / x;
#endif
}
__host__ __device__ __forceinline__ double rsq64(double x)
{
    double y = (double)abt_rsqf((float)x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// reciprocal and reciprocal square root to ~1e-14: the FP32 hardware approximation and
// one Newton step in FP64 (inputs within the FP32 range)
__host__ __device__ __forceinline__ double rcp64_1(double x)
{
    const double y = (double)abt_rcpf((float)x);
    return fma(y, fma(-x, y, 1.0), y);
}
__host__ __device__ __forceinline__ double rsq64_1(double x)
{
    const double y = (double)abt_rsqf((float)x);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

// The table of e^{i pi j / ETAB_N} (DevGrid::etab, ETAB_GRID doubles), built by the host in x87
// extended precision (64-bit significands): each angle reduced exactly to at most pi/4 (quarter
// turns, and the reflection a -> pi/2 - a within a quarter), so the extended-precision angle is
// within ~1e-19 of the exact one; cosine and sine rounded once to double, and the remainders (the
// low parts, at ETAB_LO).
inline void fill_exp_table(double* tab)
{
    const long double pi_l = 3.14159265358979323846264338327950288L;
    const int quarter = ETAB_N / 2;
    for (int j = 0; j < 2 * ETAB_N; j++) {
        const int q = j / quarter, m = j % quarter;
        const bool refl = 2 * m > quarter;
        const long double a = pi_l * (long double)(refl ? quarter - m : m) / (long double)ETAB_N;
        long double c = refl ? sinl(a) : cosl(a), s = refl ? cosl(a) : sinl(a);
        for (int k = 0; k < q; k++) {   // rotate by a quarter turn: (c, s) -> (-s, c)
            const long double t = c;
            c = -s;
            s = t;
        }
        const int i = 2 * j;
        tab[i] = (double)c;
        tab[i + 1] = (double)s;
        tab[ETAB_LO + i] = (double)(c - (long double)tab[i]);
        tab[ETAB_LO + i + 1] = (double)(s - (long double)tab[i + 1]);
    }
}

// The grid constants the FP64r body reads, stored after the table in its LDS block (read
// there, they take no scalar registers across the beam loop).
__host__ __device__ __forceinline__ void fp64r_constants(const DevGrid& g, double* c)
{
    c[ETAB_KREV - ETAB_DOUBLES] = g.krev;
    c[ETAB_K2CLAMP - ETAB_DOUBLES] = g.krev_k2clamp;
    c[ETAB_INV_KREV - ETAB_DOUBLES] = g.inv_krev;
    c[ETAB_TRI_SCALE - ETAB_DOUBLES] = g.tri_scale;
    c[ETAB_LAMBDA - ETAB_DOUBLES] = g.lambda;
}

// e^{i pi t} for an FP64 t (half revolutions): t = 2m + j / ETAB_N + r with |r| <= 1 / (2 ETAB_N)
// (every step exact), e^{i pi t} = tab[j] e^{i pi r}, the last from its Taylor polynomials
// (|pi r| <= 3.1e-3: truncation below 2e-18).  tab: the exponential table (in LDS on the device).
__host__ __device__ __forceinline__ void expi_pi64(double t, const double* __restrict__ tab, double* c, double* s)
{
    const double m = rint(0.5 * t);
    const double u = fma(-2.0, m, t);                       // [-1, 1]
    const double jf = rint((double)ETAB_N * u);
    const double r = fma(-1.0 / (double)ETAB_N, jf, u);     // |r| <= 1 / (2 ETAB_N)
    const int j = (int)jf & (2 * ETAB_N - 1);
    const double c0 = tab[2 * j], s0 = tab[2 * j + 1];
    const double y = PI64 * r;
    const double z = y * y;
    const double sr = fma(y * z, fma(z, 1.0 / 120.0, -1.0 / 6.0), y);
    const double cr = fma(z, fma(z, 1.0 / 24.0, -0.5), 1.0);
    *c = fma(c0, cr, -s0 * sr);
    *s = fma(s0, cr, c0 * sr);
}

// e^{2 pi i t} for an FP64 t (revolutions, |t| < 2^20), from the same table: t = J / (2 ETAB_N) + r
// with J = rint(2 ETAB_N t) and |r| <= 1 / (4 ETAB_N) (exact: t and J / (2 ETAB_N) are multiples of
// the ulp of t), e^{2 pi i t} = e^{2 pi i J / (2 ETAB_N)} e^{2 pi i r}; the table holds the angle
// 2 pi J / (2 ETAB_N) modulo 2 pi at index J mod 2 ETAB_N.  e^{2 pi i r} from its Taylor
// polynomials in r itself, 2 pi folded into the coefficients (truncation below 2e-18).
constexpr double EXP_S1 = 6.283185307179586;     // 2 pi
constexpr double EXP_S3 = -41.34170224039976;    // -(2 pi)^3 / 6
constexpr double EXP_S5 = 81.60524927607506;     // (2 pi)^5 / 120
constexpr double EXP_C2 = -19.739208802178716;   // -(2 pi)^2 / 2
constexpr double EXP_C4 = 64.9393940226683;      // (2 pi)^4 / 24
__host__ __device__ __forceinline__ void expi_2pi64(double t, const double* __restrict__ tab, double* c, double* s)
{
    const double jf = rint((double)(2 * ETAB_N) * t);
    const double r = fma(-1.0 / (double)(2 * ETAB_N), jf, t);     // |r| <= 1 / (4 ETAB_N)
    const int j = (int)jf & (2 * ETAB_N - 1);
    const double c0 = tab[2 * j], s0 = tab[2 * j + 1];
    const double w = r * r;
    const double sr = r * fma(w, fma(w, EXP_S5, EXP_S3), EXP_S1);
    const double cr = fma(w, fma(w, EXP_C4, EXP_C2), 1.0);
    *c = fma(c0, cr, -s0 * sr);
    *s = fma(s0, cr, c0 * sr);
}

// sin and cos of an FP64 angle x (radians, |x| < 2^40) from the same table with its low parts:
// x = n pi / ETAB_N + r with n = rint(x ETAB_N / pi), formed in the low bits of a 1.5 * 2^52
// shift, and r by a two-part Cody-Waite reduction (|r| <= pi / (2 ETAB_N) + 2^-30, error below
// 1e-18); e^{ix} = tab[n mod 2 ETAB_N] e^{ir}, the last from its Taylor polynomials (truncation
// below 2e-19 relative), the table entry taken as a double-double and the correction summed
// before its high part: ~0.5 ulp, the result rounded nearly correctly.  tab: ETAB_GRID doubles.
constexpr double SINCOS_SHIFT = 6755399441055744.0;                   // 1.5 * 2^52
constexpr double N_OVER_PI = (double)ETAB_N / 3.141592653589793;
constexpr double PI_OVER_N_HI = 3.141592653589793 / (double)ETAB_N;   // exact: a power-of-two scaling
constexpr double PI_OVER_N_LO = 1.2246467991473532e-16 / (double)ETAB_N;
__host__ __device__ __forceinline__ void sincos_tab64(double x, const double* __restrict__ tab, double* s,
                                                      double* c)
{
    const double t = fma(x, N_OVER_PI, SINCOS_SHIFT);
    const double n = t - SINCOS_SHIFT;
    double r = fma(-n, PI_OVER_N_HI, x);
    r = fma(-n, PI_OVER_N_LO, r);
    const int j = (int)(uint32_t)__builtin_bit_cast(uint64_t, t) & (2 * ETAB_N - 1);
    const double c0 = tab[2 * j], s0 = tab[2 * j + 1];
    const double cl = tab[ETAB_LO + 2 * j], sl = tab[ETAB_LO + 2 * j + 1];
    const double z = r * r;
    const double sr = fma(r * z, fma(z, 1.0 / 120.0, -1.0 / 6.0), r);
    const double cm1 = z * fma(z, 1.0 / 24.0, -0.5);                   // cos r - 1
    *s = s0 + fma(c0, sr, fma(s0, cm1, sl));
    *c = c0 + fma(-s0, sr, fma(c0, cm1, cl));
}

// a / b given y = 1 / b correctly rounded: q = a y, r = a - b q (exact), q + r y.  Equal to the
// IEEE quotient (Markstein's correction; no overflow or underflow in the kernels' ranges), at
// three operations per quotient once y is known.
__host__ __device__ __forceinline__ double div_by(double a, double b, double y)
{
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}

// ---------------------------------------------------------------------------
// FP64 plain kernel body: one beam, one bin (diffraction, diff_mod.f90:576-670; karczewski
// :124-198; contour_integral :357-421).  Every input of a sine or cosine -- the bin offset
// and its rotation into the aperture frame, bvk = k |x3|, kxx and kyy with their clamps, and per
// edge delta, delta1, delta2, omega1, omega2 and the sums delta + omega -- is formed with the
// reference's operations in its order, without contraction, so the phases are the reference's
// doubles bit for bit (the quotients through div_by equal the divisions).  The amplitude algebra
// around them is reorganised; each step there changes a result by a few ulp at most:
//   * sines and cosines from the table (sincos_tab64);
//   * the Karczewski direction normalised by reciprocal square roots, the ratio
//     sqrt((1 - k_y^2) / (1 - K_y^2)) from the record's 1 / sqrt(1 - K_y^2);
//   * rot4 . diff_ampl = g I + h J (J = [[0,1],[-1,0]]: both factors have that form), applied
//     to the amplitude matrix; temp_rot1 depends on phi only and is applied to the bin's sum
//     after the beam loop (apply_temp_rot1).
// The edge terms of contour_integral cancel one another, so they keep the reference's
// operations as well (the divisions included), in its order.
// acc: re/im of (1,1) (2,1) (1,2) (2,2), before temp_rot1.  tab: the exponential table (LDS).
// incoherent: acc[0] += |amplC|_F^2 instead (temp_rot1 is orthogonal: the norm is final).
__host__ __device__ __forceinline__ void beam_bin_fp64(const double* __restrict__ r,
                                                       const double* __restrict__ vrecs,
                                                       const DevGrid& g, const BinIn& bi,
                                                       const double* __restrict__ tab, int fast_diff,
                                                       double acc[8], bool incoherent = false)
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
    const double Kx = r[R_PROP2], Ky = r[R_PROP2 + 1], Kz = r[R_PROP2 + 2];
    // the phase inputs of contour_integral (diff_mod.f90:357-375)
    const double w = g.waveno;
    const double bvk = w * sqrt(x3 * x3 + y3 * y3 + z3 * z3);
    const double ibvk = 1.0 / bvk;
    const double ww = w * w;
    double kxx = Kx * w - div_by(ww * x3, bvk, ibvk);
    double kyy = Ky * w - div_by(ww * y3, bvk, ibvk);
    if (fabs(kxx) < KMIN) kxx = KMIN;
    if (fabs(kyy) < KMIN) kyy = KMIN;
    // karczewski (diff_mod.f90:145-181): k = x3 / r1 as the reference divides, since k also
    // decides the k_y clamp and the forward test below
    const double ir1 = 1.0 / r1;
    const double k1 = div_by(x3, r1, ir1), k3 = div_by(z3, r1, ir1);
    double k2 = div_by(y3, r1, ir1);
    if (fabs(k2) > K2CLAMP) k2 = copysign(K2CLAMP, k2);
    const double omk = fma(-k2, k2, 1.0);
    const double isq = rsq64(omk);                       // 1 / sqrt(1 - k2^2)
    const double k12 = k1 * k2;
    const double m1 = -(k12 * isq), m2 = omk * isq, m3 = -(k2 * k3 * isq);
    const double frac = m2 * r[R_FRAC];                  // sqrt((1 - k2^2) / (1 - Ky^2))
    const double ifrac = isq * r[R_ISIG];                // its reciprocal
    const double a1m = -(Kz * frac), b2m = -(k3 * ifrac);
    const double b1m = fma(Kx * Ky, frac, -(k12 * ifrac));
    const double d11 = 0.5 * (a1m + b2m), d12 = 0.5 * b1m;   // diff_ampl = d11 I + d12 J
    // scattering-plane vector hc (diff_mod.f90:630-635) and rot4 = p I - qh J
    const double i1 = r[R_INC2], i2 = r[R_INC2 + 1], i3 = r[R_INC2 + 2];
    double kk1 = k1, kk2 = k2, kk3 = k3;
    if (!(fabs(i1 * k1 + i2 * k2 + i3 * k3) < FWD_DOT)) {
        const double xt1 = bi.xt - r[R_COM], yt1 = bi.yt - r[R_COM + 1], zt1 = bi.zt - r[R_COM + 2];
        const double irt = 1.0 / sqrt(xt1 * xt1 + yt1 * yt1 + zt1 * zt1);
        kk1 = (R[0] * xt1 + R[1] * yt1 + R[2] * zt1) * irt;
        kk2 = (R[3] * xt1 + R[4] * yt1 + R[5] * zt1) * irt;
        kk3 = (R[6] * xt1 + R[7] * yt1 + R[8] * zt1) * irt;
    }
    const double h1 = fma(i2, kk3, -(i3 * kk2)), h2 = fma(i3, kk1, -(i1 * kk3)), h3 = fma(i1, kk2, -(i2 * kk1));
    const double e1 = fma(k2, m3, -(k3 * m2)), e2 = fma(k3, m1, -(k1 * m3)), e3 = fma(k1, m2, -(k2 * m1));
    const double hh = fma(h1, h1, fma(h2, h2, h3 * h3));
    const double ee = fma(e1, e1, fma(e2, e2, e3 * e3));
    const double ihh = rsq64(hh);
    const double p = fma(h1, m1, fma(h2, m2, h3 * m3)) * ihh;
    const double qh = fma(h1, e1, fma(h2, e2, h3 * e3)) * (ihh * rsq64(ee));
    const double gg = fma(p, d11, qh * d12), hj = fma(p, d12, -(qh * d11));   // rot4 . diff_ampl
    // contour_integral (diff_mod.f90:357-421), edge by edge as the reference sums: the edge
    // terms cancel, so their roundings are kept too
    double sb, cb;
    sincos_tab64(bvk, tab, &sb, &cb);
    const double lam = g.lambda, ilam = g.inv_lambda;
    const double tkx = 2 * kxx, tky = 2 * kyy;
    double ar = 0.0, ai = 0.0;
    const auto edge = [&](double xj, double yj, double mj, double nj, double dx, double dy) {
#pragma clang fp contract(off)
        const double delta = kxx * xj + kyy * yj;
        const double delta1 = kyy * mj + kxx;
        const double delta2 = kxx * nj + kyy;
        const double omega1 = dx * delta1;
        const double omega2 = dy * delta2;
        const double alpha = 1 / (tky * delta1);
        const double beta = 1 / (tkx * delta2);
        double s0, c0, s1, c1, s2, c2;
        sincos_tab64(delta, tab, &s0, &c0);
        sincos_tab64(delta + omega1, tab, &s1, &c1);
        sincos_tab64(delta + omega2, tab, &s2, &c2);
        const double sumim = alpha * (c0 - c1) - beta * (c0 - c2);
        const double sumre = -(alpha * (s0 - s1)) + beta * (s0 - s2);
        ar = ar + div_by(cb * sumre - sb * sumim, lam, ilam);
        ai = ai + div_by(cb * sumim + sb * sumre, lam, ilam);
    };
    // triangles: edges inline in the record; polygons: the vertex records.  Two
    // loops, so each reads one address space (records staged in LDS, vrecs global).
    const int nv = (int)r[R_NV];
    if (nv <= 3) {
        for (int j = 0; j < nv; j++)
            edge(r[R_EX + j], r[R_EY + j], r[R_EM + j], r[R_EN + j], r[R_EDX + j], r[R_EDY + j]);
    } else {
        const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
        for (int j = 0; j < nv; j++) {
            const double* v = vr + VREC * j;
            edge(v[V_X], v[V_Y], v[V_M], v[V_N], v[V_DX], v[V_DY]);
        }
    }
    // (g I + h J) . ampl, times the aperture integral
    const double* A = r + R_AMPL;
    double o[8];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++) {
            const double a1 = A[4 * q + ri], a2 = A[4 * q + 2 + ri];   // A(1,q), A(2,q)
            o[4 * q + ri] = fma(gg, a1, hj * a2);
            o[4 * q + 2 + ri] = fma(gg, a2, -(hj * a1));
        }
    if (incoherent) {
        // diagnostic: sum over beams of |amplC|_F^2 (conditioning of the coherent sum)
        double n2 = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) n2 = fma(o[q], o[q], n2);
        acc[0] = fma(n2, fma(ar, ar, ai * ai), acc[0]);
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double br = o[2 * q], bim = o[2 * q + 1];
        acc[2 * q] = fma(br, ar, fma(-bim, ai, acc[2 * q]));
        acc[2 * q + 1] = fma(br, ai, fma(bim, ar, acc[2 * q + 1]));
    }
}

// Difference form of one edge of the rotation-invariant form factor, from the full vertex
// phase exponentials W = e^{2 pi i t} of its two ends (t in revolutions):
//   c_j sinc(pi d) e^{i pi (t_j + t_{j+1})} = c_j (W_{j+1} - W_j) / (2 pi i d),  d = t_{j+1} - t_j,
// c_j = qx dy_j - qy dx_j.  Below |d| = 1e-3 (relative error of the difference there ~2e-16 /
// (pi |d|)) the product form with the Taylor polynomials of sinc(pi d) and e^{i pi d}, which
// only edges nearly normal to q reach.
__host__ __device__ __forceinline__ void edge_diff64(double Wr0, double Wi0, double Wr1, double Wi1, double d,
                                                     double qx, double qy, double edx, double edy, double& ssr,
                                                     double& ssi)
{
    const double c = fma(qx, edy, -qy * edx);
    if (fabs(d) >= 1e-3) {
        const double k = c * rcp64_1((2.0 * PI64) * d);
        ssr = fma(k, Wi1 - Wi0, ssr);       // c (W1 - W0) / (2 pi i d)
        ssi = fma(-k, Wr1 - Wr0, ssi);
    } else {
        const double x = PI64 * d, z = x * x;
        const double sinc = fma(z, fma(z, 1.0 / 120.0, -1.0 / 6.0), 1.0);
        const double ch = fma(z, fma(z, 1.0 / 24.0, -0.5), 1.0);          // e^{i pi d}
        const double sh = x * fma(z, fma(z, 1.0 / 120.0, -1.0 / 6.0), 1.0);
        const double k = c * sinc;
        ssr = fma(k, fma(Wr0, ch, -Wi0 * sh), ssr);
        ssi = fma(k, fma(Wr0, sh, Wi0 * ch), ssi);
    }
}

// Restructured FP64 body: one beam, one bin; adds to the projector sums S+ (acc[0..3]) and
// S- (acc[4..7]) that fp64r_finish turns into the amplitude matrix.  bi.rad = |x| is read.
__host__ __device__ __forceinline__ void beam_bin_fp64r(const double* __restrict__ r,
                                                        const double* __restrict__ vrecs,
                                                        const DevGrid& g, const BinIn& bi,
                                                        const double* __restrict__ etab,
                                                        int fast_diff, double acc[8])
{
    // ---- geometry: bin direction from the aperture centre; the far-field phase relative to
    // the bin's own, k (|x - c| - |x|) = k (|c|^2 - 2 x.c) / (|x - c| + |x|), formed without
    // cancellation (~1e-13 rad; the per-bin e^{i k |x|} is applied after the beam sum) ----
    const double dx = bi.x - r[R_COM], dy = bi.y - r[R_COM + 1], dz = bi.z - r[R_COM + 2];
    const double s = fma(dx, dx, fma(dy, dy, dz * dz));
    const double invd = rsq64_1(s);
    const double ux = dx * invd, uy = dy * invd, uz = dz * invd;
    if (fast_diff) {
        if (ux * r[R_PROP0] + uy * r[R_PROP0 + 1] + uz * r[R_PROP0 + 2] < r[R_COSFOV]) return;
    }
    const double* K = r + R_KRR;
    const double qx = fma(-K[0], ux, fma(-K[1], uy, fma(-K[2], uz, r[R_KP2])));
    const double qy = fma(-K[3], ux, fma(-K[4], uy, fma(-K[5], uz, r[R_KP2 + 1])));
    const double k1r = r[R_KP2] - qx, k2r = r[R_KP2 + 1] - qy;
    const double rc = fma(bi.x, r[R_COM], fma(bi.y, r[R_COM + 1], bi.z * r[R_COM + 2]));
    const double kd = etab[ETAB_KREV] * (fma(-2.0, rc, r[R_CC]) * rcp64_1(fma(s, invd, bi.rad)));   // revolutions
    const double q2 = fma(qx, qx, qy * qy);
    const double lam = etab[ETAB_LAMBDA];

    // ---- aperture integral with the far-field phase: (FR, FI) = e^{i k|r-com|} sum_ref / lambda ----
    double FR, FI;
    const double* rf_ = r;   // vertex coordinates of triangles are inline
    const bool tri = r[R_NV] == 3.0;
    const double qa2 = (double)reinterpret_cast<const float*>(r + R_F32)[F_QA2];
    if (q2 * qa2 < (double)SERIES_QA2) {
        const double tw = 6.283185307179586;
        double sr = 0.0, si = 0.0;
        if (tri) {
            const double a = tw * fma(qx, rf_[R_EX], qy * rf_[R_EY]);
            const double b = tw * fma(qx, rf_[R_EX + 1], qy * rf_[R_EY + 1]);
            const double c = tw * fma(qx, rf_[R_EX + 2], qy * rf_[R_EY + 2]);
            double tr, ti;
            tri_series(a, b, c, &tr, &ti);
            sr = r[R_AREA] * tr;
            si = r[R_AREA] * ti;
        } else {
            const int nv = (int)r[R_NV];
            const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
            for (int j = 0; j < nv; j++) {
                const int jp = (j == nv - 1) ? 0 : j + 1;
                const double xj = vr[VREC * j + V_X], yj = vr[VREC * j + V_Y];
                const double xn = vr[VREC * jp + V_X], yn = vr[VREC * jp + V_Y];
                const double a = tw * fma(qx, xj, qy * yj);
                const double b = tw * fma(qx, xn, qy * yn);
                const double sa = 0.5 * (xj * yn - xn * yj);
                double tr, ti;
                tri_series(0.0, a, b, &tr, &ti);
                sr += sa * tr;
                si += sa * ti;
            }
        }
        // sum_ref = -i F = si - i sr
        double cb, sb;
        expi_pi64(2.0 * kd, etab, &cb, &sb);
        const double il = rcp64(lam);
        FR = (cb * si + sb * sr) * il;
        FI = (sb * si - cb * sr) * il;
    } else {
        double ssr = 0.0, ssi = 0.0;
        bool edges = true;
        if (tri) {
            double t[3], Wr[3], Wi[3];
#pragma unroll
            for (int l = 0; l < 3; l++) {
                t[l] = fma(qx, r[R_EX + l], fma(qy, r[R_EY + l], kd));
                expi_2pi64(t[l], etab, &Wr[l], &Wi[l]);
            }
            // One denominator: the form factor of a triangle is 2 A times the second divided
            // difference of e^{i alpha} at its vertex phases alpha_l = 2 pi t_l, divided by i^2,
            //   F = 2 A (W0 d1 + W1 d2 + W2 d0) / ((2 pi)^2 d0 d1 d2),  d_j = t_{j+1} - t_j,
            // and sum_ref / lambda = -i F / lambda.  Its rounding is that of the edge sum while
            // every |d_j| >= 1e-3 (relative to the largest edge term, ~ulp / min |d_j|); below,
            // the edge sum with its product form.
            const double d0 = t[1] - t[0], d1 = t[2] - t[1], d2 = t[0] - t[2];
            if (fmin(fabs(d0), fmin(fabs(d1), fabs(d2))) >= 1e-3) {
                const double nr = fma(Wr[0], d1, fma(Wr[1], d2, Wr[2] * d0));
                const double ni = fma(Wi[0], d1, fma(Wi[1], d2, Wi[2] * d0));
                const double f = r[R_AREA] * etab[ETAB_TRI_SCALE] * rcp64_1(d0 * d1 * d2);
                FR = f * ni;
                FI = -f * nr;
                edges = false;
            } else {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int jn = j == 2 ? 0 : j + 1;
                    edge_diff64(Wr[j], Wi[j], Wr[jn], Wi[jn], t[jn] - t[j], qx, qy, r[R_EDX + j], r[R_EDY + j],
                                ssr, ssi);
                }
            }
        } else {
            const int nv = (int)r[R_NV];
            const double* vr = vrecs + (int64_t)r[R_VOFF] * VREC;
            const double t0 = fma(qx, vr[V_X], fma(qy, vr[V_Y], kd));
            double Wr0, Wi0;
            expi_2pi64(t0, etab, &Wr0, &Wi0);
            double tp = t0, Wrp = Wr0, Wip = Wi0;
            for (int j = 0; j < nv; j++) {
                double tn = t0, Wrn = Wr0, Win = Wi0;
                if (j < nv - 1) {
                    tn = fma(qx, vr[VREC * (j + 1) + V_X], fma(qy, vr[VREC * (j + 1) + V_Y], kd));
                    expi_2pi64(tn, etab, &Wrn, &Win);
                }
                edge_diff64(Wrp, Wip, Wrn, Win, tn - tp, qx, qy, vr[VREC * j + V_DX], vr[VREC * j + V_DY], ssr, ssi);
                tp = tn;
                Wrp = Wrn;
                Wip = Win;
            }
        }
        if (edges) {
            const double sc = -rcp64_1(6.283185307179586 * lam * q2);
            FR = ssr * sc;
            FI = ssi * sc;
        }
    }

    // ---- polarisation (karczewski + rot4, diff_mod.f90:124-198, 630-643) ----
    const bool kclamp = fabs(k2r) > etab[ETAB_K2CLAMP];
    const bool fwd = !(fabs(uz) < FWD_DOT);
    const double ik = etab[ETAB_INV_KREV];
    const double k1 = k1r * ik, k2 = k2r * ik;
    const double* RR = r + R_RR;
    const double k3 = fma(RR[6], ux, fma(RR[7], uy, RR[8] * uz));
    const double s2 = fma(k1, k1, k3 * k3);                     // 1 - k_y^2 (unclamped)
    const double* C = r + R_C64;
    double gg, hh;
    if (!kclamp && !fwd) {
        const double P0 = fma(K[3], uy, -K[4] * ux);
        const double Q0 = fma(-uz, k2r, K[5]);
        const double rh = rsq64_1(fma(ux, ux, uy * uy) * (s2 * s2));   // 1 / (|(ux, uy)| s2)
        const double P = P0 * rh, Q = Q0 * rh;
        const double A = fma(C[0], s2, C[1] * k3);
        const double B = fma(C[2], s2, C[1] * (k1 * k2));
        gg = fma(P, A, Q * B);
        hh = fma(P, B, -Q * A);
    } else {
        const double Kx = r[R_PROP2], Ky = r[R_PROP2 + 1], Kz = r[R_PROP2 + 2];
        const double sig = r[R_FRAC], isig = r[R_ISIG];
        const double k2c = kclamp ? copysign(K2CLAMP, k2) : k2;
        const double s2c = kclamp ? (1.0 - K2CLAMP * K2CLAMP) : s2;
        const double rs = rsq64(s2c);
        const double sq = s2c * rs;
        const double frac = sq * sig;
        const double ifrac = rs * isig;
        double da = -0.5 * fma(Kz, frac, k3 * ifrac);
        double db = 0.5 * fma(Kx * Ky, frac, -k1 * k2c * ifrac);
        {
            // the mirror-zero form (see beam_bin_mixed)
            const double dk1 = -qx * ik, dk2 = -qy * ik;
            const double dk3 = fma(RR[6], ux, fma(RR[7], uy, fma(RR[8], uz, Kz)));
            const double u2 = dk2 * (k2 + Ky);
            const double w2 = fma(Kx, dk2, dk1 * k2);
            const double dan = 0.5 * rs * fma(Kz * sig, u2, -isig * dk3);
            const double dbn = -0.5 * rs * fma(Kx * Ky * sig, u2, isig * w2);
            const bool mir = !kclamp && s2 > 0.5;
            da = mir ? dan : da;
            db = mir ? dbn : db;
        }
        const double m1 = -k1 * k2c * rs, m2 = sq, m3 = -k2c * k3 * rs;
        const double r13 = kclamp ? rsq64(s2) : rs;
        const double ev1 = -k3 * r13, ev3 = k1 * r13;
        double hx = uy, hy = -ux;
        if (fwd) {
            hx = bi.yt - r[R_COM + 1];
            hy = -(bi.xt - r[R_COM]);
        }
        const double rh = rsq64(fma(hx, hx, hy * hy));
        hx *= rh;
        hy *= rh;
        const double h1 = fma(RR[0], hx, RR[1] * hy);
        const double h2 = fma(RR[3], hx, RR[4] * hy);
        const double h3 = fma(RR[6], hx, RR[7] * hy);
        const double p = fma(h1, m1, fma(h2, m2, h3 * m3));
        const double qh = fma(h1, ev1, h3 * ev3);
        gg = fma(p, da, qh * db);
        hh = fma(p, db, -qh * da);
    }
    // f (gg I + hh J) A = w P A + w' Q A, w = f (gg + i hh), w' = f (gg - i hh)
    const double ta = FR * gg, tb = FI * hh, tc = FR * hh, td = FI * gg;
    const double wr = ta - tb, wi = tc + td, vr_ = ta + tb, vi = td - tc;
    const double* rp = r + R_RP64;
    const double* rm = r + R_RM64;
    acc[0] = fma(wr, rp[0], fma(-wi, rp[1], acc[0]));
    acc[1] = fma(wr, rp[1], fma(wi, rp[0], acc[1]));
    acc[2] = fma(wr, rp[2], fma(-wi, rp[3], acc[2]));
    acc[3] = fma(wr, rp[3], fma(wi, rp[2], acc[3]));
    acc[4] = fma(vr_, rm[0], fma(-vi, rm[1], acc[4]));
    acc[5] = fma(vr_, rm[1], fma(vi, rm[0], acc[5]));
    acc[6] = fma(vr_, rm[2], fma(-vi, rm[3], acc[6]));
    acc[7] = fma(vr_, rm[3], fma(vi, rm[2], acc[7]));
}

// Beams per FP32 partial sum of the mixed kernels: every MIXED_GROUP consecutive beams
// (counted from the start of the beam list, or of the orientation) are summed in FP32 and
// the partial is added to the FP64 sums.  Tile sizes and chunk lengths are multiples of it,
// so the groups, and the result, do not depend on them.
constexpr int MIXED_GROUP = 8;

// The per-bin phase factor e^{i k |x|} of the FP64 model body (adjbody.h).
__host__ __device__ __forceinline__ void apply_bin_phase(double acc[8], double kr_rev)
{
#pragma clang fp contract(off)
    double s, c;
    sincos_f64(6.283185307179586 * (kr_rev - rint(kr_rev)), &s, &c);
    for (int q = 0; q < 4; q++) {
        const double re = acc[2 * q], im = acc[2 * q + 1];
        acc[2 * q] = re * c - im * s;
        acc[2 * q + 1] = re * s + im * c;
    }
}

// From the projector sums S+ = acc[0..3], S- = acc[4..7] (columns 1, 2) to the
// amplitude matrix X = [S+ + S-; i (S+ - S-)], then temp_rot1 on the right.
__host__ __device__ __forceinline__ void apply_temp_rot1(double acc[8], double cp, double sp);

__host__ __device__ __forceinline__ void mixed_finish(double acc[8], double cp, double sp)
{
    double x[8];
    // layout of x: (1,1) (2,1) (1,2) (2,2), re/im
    x[0] = acc[0] + acc[4];          x[1] = acc[1] + acc[5];           // X(1,1)
    x[2] = -(acc[1] - acc[5]);       x[3] = acc[0] - acc[4];           // X(2,1) = i (S+1 - S-1)
    x[4] = acc[2] + acc[6];          x[5] = acc[3] + acc[7];           // X(1,2)
    x[6] = -(acc[3] - acc[7]);       x[7] = acc[2] - acc[6];           // X(2,2) = i (S+2 - S-2)
    apply_temp_rot1(x, cp, sp);
    for (int q = 0; q < 8; q++) acc[q] = x[q];
}

// The mixed kernel sums f_b G_b A_b over beams per bin; the phi-only rotation
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

// After the beam sum of beam_bin_fp64r: the projector sums to the amplitude matrix (with
// temp_rot1), then the bin's far-field phase e^{i k |x|} (kr_rev = k |x| / 2 pi).
__host__ __device__ __forceinline__ void fp64r_finish(double acc[8], double cp, double sp, double kr_rev)
{
    mixed_finish(acc, cp, sp);
    apply_bin_phase(acc, kr_rev);
}

}  // namespace abt
