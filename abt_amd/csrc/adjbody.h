// adjbody.h -- per (beam, bin) body of the adjoint of the far-field amplitude
// (host + device, so the tests can run it on the CPU).
//
// Model.  For a triangular aperture with vertices v_l (global frame), centre
// c = mean(v_l), right-handed normal n = N / |N|, N = (v_1 - v_0) x (v_2 - v_0), the
// amplitude `diffraction` (diff_mod.f90:576-670) puts on a bin at x is
//     A = s (g I + h J) C T_phi,     s = e^{i k |x - c|} (i / lambda) F(q),
// where C is the amplitude matrix with the reference's anti-parallel sign
// (diff_mod.f90:561-571), T_phi = temp_rot1 (diff_mod.f90:646-650), J = [[0,1],[-1,0]],
// F(q) = int_T e^{i q.(u - c)} dA the aperture's form factor at q = k (prop - khat),
// khat = (x - c) / |x - c|, and g, h the two numbers of the Karczewski-Wolf matrix
// and the scattering-plane rotation, rot4 . diff_ampl = g I + h J (diff_mod.f90:
// 124-198, 630-643), written with the aperture frame e_z = -n, e_y = unit(e_z x prop),
// e_x = e_y x e_z (the frame get_rotation_matrix2 and rot2 build,
// diff_mod.f90:200-289, 533-547):
//     k_j = khat.e_j, sq = sqrt(1 - k_2^2), m = (e_y - k_2 khat) / sq,
//     evo = (khat x e_y) / sqrt(1 - k_2^2), hc = unit(inc x khat) (or the theta_i bin),
//     p = hc.m, qh = hc.evo, da = -(Kz sq + k_3 / sq) / 2, db = -k_1 k_2 / (2 sq),
//     g = p da + qh db, h = p db - qh da.
// This equals the reference's formula up to rounding (the split contour sum of
// contour_integral is the same F, +i F for the clockwise aperture frame).
//
// Adjoint.  Given Lam = dL/dA (as re/im pairs: dL = Re sum conj(Lam) dA), the body
// adds this bin's part of dL/dC and of the per-beam sums from which dL/dv_l follows
// (adj_finish):
//   * the form factor at fixed q: dF/dv_l = a_l F + q (a_l . grad_q F), a_l =
//     d|N|/dv_l / |N| (= the gradient of the barycentric coordinate l), so
//     sum over bins of Re(nu F) and of q Re(nu grad F) suffice;
//   * c through khat (q and the polarisation) and through the phase k |x - c|;
//   * the frame through e_z = -n.
// Per-edge phase terms are recomputed, never stored.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ffbody.h"
#include "records.h"
#include "trig.h"

namespace abt {

// Per-bin inputs of the adjoint: bin position, the theta_i bin of the same phi (forward
// fallback of hc), |x| and k|x|/(2 pi); Lam' = Lam T_phi^T (re/im, Fortran pair order).
struct AdjBin {
    double x, y, z, xt, yt, rad, kr_rev;
};

// Per-beam constants (from the prepared record).
struct AdjBeam {
    double c[3], cc;            // centre, |c|^2
    double ex[3], ey[3], ez[3]; // aperture frame
    double P[3], Kz;            // propagation vector, P.e_z
    double C[8];                // sigma * ampl
    double xl[3], yl[3];        // in-plane coordinates of v_l - c
    double xr[3], yr[3];        // the same over 2 pi (mixed body: phases in revolutions)
    double area;                // |N| / 2
    double qa2;                 // (aperture radius)^2: series branch when |q|^2 qa2 < SERIES_QA^2
    // FP32 copies for the mixed-precision body
    float fex[3], fey[3], fez[3], fP[3], fKz, fC[8];
    float fxl[3], fyl[3];
    float fso;                  // orientation sign of (xl, yl) for the edge form
};

struct AdjAcc {
    double GC[8];               // dL/dC (re/im pairs)
    double S0;                  // sum Re(nu F)
    double Qx[3], Qy[3];        // sum q Re(nu dF/dq_x), sum q Re(nu dF/dq_y)
    double cb[3];               // dL/dc through khat, the phase and the fallback hc
    double exb[3], eyb[3], ezb[3];
};

__host__ __device__ inline void adj_beam_from_record(const double* __restrict__ r, AdjBeam& B)
{
    for (int i = 0; i < 3; i++) {
        B.c[i] = r[R_COM + i];
        B.ex[i] = r[R_RR + i];
        B.ey[i] = r[R_RR + 3 + i];
        B.ez[i] = r[R_RR + 6 + i];
        B.P[i] = r[R_PROP0 + i];
        B.xl[i] = r[R_EX + i];
        B.yl[i] = r[R_EY + i];
        B.xr[i] = B.xl[i] * 0.15915494309189535;
        B.yr[i] = B.yl[i] * 0.15915494309189535;
    }
    B.cc = r[R_CC];
    B.Kz = B.P[0] * B.ez[0] + B.P[1] * B.ez[1] + B.P[2] * B.ez[2];
    for (int q = 0; q < 8; q++) B.C[q] = r[R_AMPL + q];
    B.area = fabs(r[R_AREA]);
    B.qa2 = r[R_AMAX] * r[R_AMAX];
    for (int i = 0; i < 3; i++) {
        B.fex[i] = (float)B.ex[i];
        B.fey[i] = (float)B.ey[i];
        B.fez[i] = (float)B.ez[i];
        B.fP[i] = (float)B.P[i];
        B.fxl[i] = (float)B.xl[i];
        B.fyl[i] = (float)B.yl[i];
    }
    B.fKz = (float)B.Kz;
    for (int q = 0; q < 8; q++) B.fC[q] = (float)B.C[q];
    const double a2 = (B.xl[1] - B.xl[0]) * (B.yl[2] - B.yl[0]) - (B.xl[2] - B.xl[0]) * (B.yl[1] - B.yl[0]);
    B.fso = a2 < 0.0 ? 1.0f : -1.0f;
}

// FP32 partial sums of the mixed body over a group of bins, added into AdjAcc
struct AdjAccF {
    float GC[8], S0, Qx[3], Qy[3], cb[3], exb[3], eyb[3], ezb[3];
};

__host__ __device__ inline void adj_zero_f(AdjAccF& a)
{
    for (int q = 0; q < 8; q++) a.GC[q] = 0.f;
    a.S0 = 0.f;
    for (int i = 0; i < 3; i++) {
        a.Qx[i] = a.Qy[i] = a.cb[i] = 0.f;
        a.exb[i] = a.eyb[i] = a.ezb[i] = 0.f;
    }
}

__host__ __device__ inline void adj_add(AdjAcc& a, const AdjAccF& f)
{
    for (int q = 0; q < 8; q++) a.GC[q] += (double)f.GC[q];
    a.S0 += (double)f.S0;
    for (int i = 0; i < 3; i++) {
        a.Qx[i] += (double)f.Qx[i];
        a.Qy[i] += (double)f.Qy[i];
        a.cb[i] += (double)f.cb[i];
        a.exb[i] += (double)f.exb[i];
        a.eyb[i] += (double)f.eyb[i];
        a.ezb[i] += (double)f.ezb[i];
    }
}

// bins per FP32 partial sum of the mixed adjoint
constexpr int ADJ_GROUP = 8;

__host__ __device__ inline void adj_zero(AdjAcc& a)
{
    for (int q = 0; q < 8; q++) a.GC[q] = 0.0;
    a.S0 = 0.0;
    for (int i = 0; i < 3; i++) {
        a.Qx[i] = a.Qy[i] = a.cb[i] = 0.0;
        a.exb[i] = a.eyb[i] = a.ezb[i] = 0.0;
    }
}

// sin(x)/x and its derivative
__host__ __device__ __forceinline__ double sinc64(double x)
{
    const double z = x * x;
    if (fabs(x) < 0.5)
        return 1.0 + z * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 + z * (-1.0 / 39916800)))));
    double s, c;
    sincos_f64(x, &s, &c);
    return s / x;
}

__host__ __device__ __forceinline__ double dsinc64(double x)
{
    const double z = x * x;
    if (fabs(x) < 0.5)
        return x * (-1.0 / 3 + z * (1.0 / 30 + z * (-1.0 / 840 + z * (1.0 / 45360 + z * (-1.0 / 3991680 + z * (1.0 / 518918400))))));
    double s, c;
    sincos_f64(x, &s, &c);
    return (c - s / x) / x;
}

// F(q) and grad_q F of the triangle (in-plane coordinates xl, yl about the centre,
// either orientation), q = (qx, qy) rad/um.  Series about q = 0 for |q| a < SERIES_QA
// (complete homogeneous symmetric polynomials of the vertex phases and their
// derivatives), else the edge form F = (i/|q|^2) sum_j c_j sinc(q.e_j/2) e^{i q.m_j}.
__host__ __device__ inline void tri_ff_grad(const AdjBeam& B, double qx, double qy, double F[2],
                                           double Gx[2], double Gy[2])
{
    const double q2 = qx * qx + qy * qy;
    if (q2 * B.qa2 < 2.25) {
        double al[3];
        for (int l = 0; l < 3; l++) al[l] = qx * B.xl[l] + qy * B.yl[l];
        const double e1 = al[0] + al[1] + al[2];
        const double e2 = al[0] * al[1] + al[1] * al[2] + al[2] * al[0];
        const double e3 = al[0] * al[1] * al[2];
        double hm3 = 0.0, hm2 = 0.0, hm1 = 1.0;          // h_{n-3}, h_{n-2}, h_{n-1} at n = 1
        double d[3] = {0.0, 0.0, 0.0};                    // h_{n-1}(al_l, al) at n = 0
        double fr = 0.5, fi = 0.0;                        // n = 0 term: h_0 / 2!
        double dr[3] = {0, 0, 0}, di[3] = {0, 0, 0};
        double inv = 0.5;
        double hprev = 1.0;                               // h_{n-1}
        for (int n = 1; n <= 24; n++) {
            // d^{(l)}_{n-1} = h_{n-1} + al_l d^{(l)}_{n-2}
            for (int l = 0; l < 3; l++) d[l] = hprev + al[l] * d[l];
            const double hn = e1 * hm1 - e2 * hm2 + e3 * hm3;
            inv = inv / (double)(n + 2);
            const double t = hn * inv;
            // i^n
            switch (n & 3) {
                case 0: fr += t; for (int l = 0; l < 3; l++) dr[l] += d[l] * inv; break;
                case 1: fi += t; for (int l = 0; l < 3; l++) di[l] += d[l] * inv; break;
                case 2: fr -= t; for (int l = 0; l < 3; l++) dr[l] -= d[l] * inv; break;
                default: fi -= t; for (int l = 0; l < 3; l++) di[l] -= d[l] * inv; break;
            }
            hm3 = hm2; hm2 = hm1; hm1 = hn;
            hprev = hn;
        }
        const double A2 = 2.0 * B.area;
        F[0] = A2 * fr;
        F[1] = A2 * fi;
        Gx[0] = Gx[1] = Gy[0] = Gy[1] = 0.0;
        for (int l = 0; l < 3; l++) {
            Gx[0] += A2 * dr[l] * B.xl[l];
            Gx[1] += A2 * di[l] * B.xl[l];
            Gy[0] += A2 * dr[l] * B.yl[l];
            Gy[1] += A2 * di[l] * B.yl[l];
        }
        return;
    }
    // orientation of (xl, yl): the edge form below is for counter-clockwise order
    const double a2 = (B.xl[1] - B.xl[0]) * (B.yl[2] - B.yl[0]) - (B.xl[2] - B.xl[0]) * (B.yl[1] - B.yl[0]);
    const double so = a2 < 0.0 ? 1.0 : -1.0;
    double sr = 0, si = 0, gxr = 0, gxi = 0, gyr = 0, gyi = 0;
    for (int j = 0; j < 3; j++) {
        const int jn = j == 2 ? 0 : j + 1;
        const double ex = B.xl[jn] - B.xl[j], ey = B.yl[jn] - B.yl[j];
        const double mx = 0.5 * (B.xl[j] + B.xl[jn]), my = 0.5 * (B.yl[j] + B.yl[jn]);
        const double cj = so * (qx * ey - qy * ex);
        const double K = qx * ex + qy * ey;
        const double mu = qx * mx + qy * my;
        const double S = sinc64(0.5 * K), dS = 0.5 * dsinc64(0.5 * K);
        double es, ec;
        sincos_f64(mu, &es, &ec);
        const double t = cj * S;
        sr += t * ec;
        si += t * es;
        // d/dq of c_j S e^{i mu}: (grad c_j S + c_j dS e) e^{i mu} + i c_j S m e^{i mu}
        const double ax = so * ey * S + cj * dS * ex, ay = -so * ex * S + cj * dS * ey;
        gxr += ax * ec - t * mx * es;
        gxi += ax * es + t * mx * ec;
        gyr += ay * ec - t * my * es;
        gyi += ay * es + t * my * ec;
    }
    const double iq2 = 1.0 / q2;
    // F = i (sr + i si) / q2
    F[0] = -si * iq2;
    F[1] = sr * iq2;
    Gx[0] = -gxi * iq2 - 2.0 * qx * iq2 * F[0];
    Gx[1] = gxr * iq2 - 2.0 * qx * iq2 * F[1];
    Gy[0] = -gyi * iq2 - 2.0 * qy * iq2 * F[0];
    Gy[1] = gyr * iq2 - 2.0 * qy * iq2 * F[1];
}

__host__ __device__ __forceinline__ double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__host__ __device__ __forceinline__ void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// One (beam, bin) of the FP64 adjoint.  lam: Lam' = Lam T_phi^T (8 doubles).
__host__ __device__ inline void adj_bin_fp64(const AdjBeam& B, const AdjBin& bn, const double* __restrict__ lam,
                                             double waveno, double lambda, AdjAcc& acc)
{
    const double k = waveno;
    double d[3] = {bn.x - B.c[0], bn.y - B.c[1], bn.z - B.c[2]};
    const double rho = sqrt(dot3(d, d));
    const double ir = 1.0 / rho;
    double kh[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
    // phase e^{i k rho}: k (rho - |x|) without cancellation, plus k |x| from kr_rev
    const double rc = bn.x * B.c[0] + bn.y * B.c[1] + bn.z * B.c[2];
    const double drho = (B.cc - 2.0 * rc) / (rho + bn.rad);
    const double ph_rev = (bn.kr_rev - rint(bn.kr_rev)) + (k / 6.283185307179586) * drho;
    double Es, Ec;
    sincos_f64(6.283185307179586 * (ph_rev - rint(ph_rev)), &Es, &Ec);
    // q = k (P - khat), in-plane components
    double q3[3] = {k * (B.P[0] - kh[0]), k * (B.P[1] - kh[1]), k * (B.P[2] - kh[2])};
    const double qx = dot3(q3, B.ex), qy = dot3(q3, B.ey);
    double F[2], Gx[2], Gy[2];
    tri_ff_grad(B, qx, qy, F, Gx, Gy);
    // polarisation
    const double k1 = dot3(kh, B.ex), k2 = dot3(kh, B.ey), k3 = dot3(kh, B.ez);
    const bool clamp = fabs(k2) > K2CLAMP;
    const double k2c = clamp ? copysign(K2CLAMP, k2) : k2;
    const double sq = sqrt(1.0 - k2c * k2c);
    const double s13 = sqrt(1.0 - k2 * k2);
    const bool fb = !(fabs(kh[2]) < FWD_DOT);       // |inc . khat|, inc = (0,0,-1)
    double hv[3];
    if (fb) {
        hv[0] = bn.yt - B.c[1];
        hv[1] = -(bn.xt - B.c[0]);
    } else {
        hv[0] = kh[1];
        hv[1] = -kh[0];
    }
    hv[2] = 0.0;
    const double nh = sqrt(hv[0] * hv[0] + hv[1] * hv[1]);
    const double hc[3] = {hv[0] / nh, hv[1] / nh, 0.0};
    double mvec[3], uu[3], evo[3];
    for (int i = 0; i < 3; i++) mvec[i] = (B.ey[i] - k2c * kh[i]) / sq;
    cross3(kh, B.ey, uu);
    for (int i = 0; i < 3; i++) evo[i] = uu[i] / s13;
    const double p = dot3(hc, mvec), qh = dot3(hc, evo);
    const double da = -0.5 * (B.Kz * sq + k3 / sq);
    const double db = -0.5 * k1 * k2c / sq;
    const double g = p * da + qh * db;
    const double h = p * db - qh * da;
    // s = E (i / lambda) F
    const double il = 1.0 / lambda;
    const double EFr = Ec * F[0] - Es * F[1], EFi = Ec * F[1] + Es * F[0];
    const double sr = -EFi * il, si = EFr * il;
    // z1 = <Lam', C>, z2 = <Lam', J C>  (<X,Y> = sum conj(X) Y); J C rows: (C_2., -C_1.)
    double z1r = 0, z1i = 0, z2r = 0, z2i = 0;
    for (int cc = 0; cc < 2; cc++) {
        const int i1 = 2 * (0 + 2 * cc), i2 = 2 * (1 + 2 * cc);   // (1,c), (2,c)
        const double l1r = lam[i1], l1i = lam[i1 + 1], l2r = lam[i2], l2i = lam[i2 + 1];
        const double c1r = B.C[i1], c1i = B.C[i1 + 1], c2r = B.C[i2], c2i = B.C[i2 + 1];
        z1r += l1r * c1r + l1i * c1i + l2r * c2r + l2i * c2i;
        z1i += l1r * c1i - l1i * c1r + l2r * c2i - l2i * c2r;
        // (JC)(1,c) = C(2,c), (JC)(2,c) = -C(1,c)
        z2r += l1r * c2r + l1i * c2i - (l2r * c1r + l2i * c1i);
        z2i += l1r * c2i - l1i * c2r - (l2r * c1i - l2i * c1r);
    }
    // dL/dC += conj(s) (g Lam' - h J^T... ) : G_C = conj(s) (g I - h J) Lam'
    // (g I - h J) Lam': row 1 = g L1 - h L2, row 2 = g L2 + h L1
    for (int cc = 0; cc < 2; cc++) {
        const int i1 = 2 * (0 + 2 * cc), i2 = 2 * (1 + 2 * cc);
        const double r1r = g * lam[i1] - h * lam[i2], r1i = g * lam[i1 + 1] - h * lam[i2 + 1];
        const double r2r = g * lam[i2] + h * lam[i1], r2i = g * lam[i2 + 1] + h * lam[i1 + 1];
        acc.GC[i1] += sr * r1r + si * r1i;
        acc.GC[i1 + 1] += sr * r1i - si * r1r;
        acc.GC[i2] += sr * r2r + si * r2i;
        acc.GC[i2 + 1] += sr * r2i - si * r2r;
    }
    const double gg = sr * z1r - si * z1i;          // Re(s z1)
    const double gh = sr * z2r - si * z2i;          // Re(s z2)
    // nu = (g z1 + h z2) E i / lambda
    const double mur = g * z1r + h * z2r, mui = g * z1i + h * z2i;
    const double mEr = mur * Ec - mui * Es, mEi = mur * Es + mui * Ec;
    const double nur = -mEi * il, nui = mEr * il;
    const double nFr = nur * F[0] - nui * F[1], nFi = nur * F[1] + nui * F[0];
    const double Rx = nur * Gx[0] - nui * Gx[1];
    const double Ry = nur * Gy[0] - nui * Gy[1];
    acc.S0 += nFr;
    for (int i = 0; i < 3; i++) {
        acc.Qx[i] += q3[i] * Rx;
        acc.Qy[i] += q3[i] * Ry;
    }
    // khat adjoint: q = k (P - khat) and the polarisation
    double kb[3];
    for (int i = 0; i < 3; i++) kb[i] = -k * (Rx * B.ex[i] + Ry * B.ey[i]);
    const double pb = gg * da + gh * db, qb = gg * db - gh * da;
    const double dab = gg * p - gh * qh, dbb = gg * qh + gh * p;
    const double Kzb = -0.5 * sq * dab;
    double sqb = -0.5 * (B.Kz - k3 / (sq * sq)) * dab + 0.5 * k1 * k2c / (sq * sq) * dbb;
    const double k3b = -0.5 * dab / sq;
    const double k1b = -0.5 * k2c / sq * dbb;
    double k2cb = -0.5 * k1 / sq * dbb;
    double hcb[3], mb[3];
    for (int i = 0; i < 3; i++) {
        hcb[i] = pb * mvec[i] + qb * evo[i];
        mb[i] = pb * hc[i];
    }
    k2cb += -dot3(mb, kh) / sq;
    sqb += -dot3(mb, mvec) / sq;
    double ub[3];
    for (int i = 0; i < 3; i++) {
        acc.eyb[i] += mb[i] / sq;
        kb[i] += -k2c / sq * mb[i];
        ub[i] = qb * hc[i] / s13;
    }
    const double s13b = -qb * dot3(hc, evo) / s13;
    double t3[3];
    cross3(B.ey, ub, t3);
    for (int i = 0; i < 3; i++) kb[i] += t3[i];
    cross3(ub, kh, t3);
    for (int i = 0; i < 3; i++) acc.eyb[i] += t3[i];
    double k2b = -k2 / s13 * s13b;
    k2cb += -k2c / sq * sqb;
    if (!clamp) k2b += k2cb;
    for (int i = 0; i < 3; i++) {
        kb[i] += k1b * B.ex[i] + k2b * B.ey[i] + k3b * B.ez[i];
        acc.exb[i] += k1b * kh[i];
        acc.eyb[i] += k2b * kh[i];
        acc.ezb[i] += k3b * kh[i] + Kzb * B.P[i];
    }
    // hc = hv / |hv|, hv = inc x khat (or inc x (x_t - c)); d/dy (inc x y).hvb = y.(hvb x inc)
    const double hdot = hcb[0] * hc[0] + hcb[1] * hc[1];
    const double hvbx = (hcb[0] - hdot * hc[0]) / nh, hvby = (hcb[1] - hdot * hc[1]) / nh;
    const double tx = -hvby, ty = hvbx;               // hvb x (0,0,-1)
    if (fb) {
        acc.cb[0] -= tx;
        acc.cb[1] -= ty;
    } else {
        kb[0] += tx;
        kb[1] += ty;
    }
    // khat = d / rho, d = x - c:  dc = -(I - khat khat^T) kb / rho;  phase: k Im(nu F) khat
    const double kk = dot3(kb, kh);
    const double ph = k * nFi;
    for (int i = 0; i < 3; i++) acc.cb[i] += -(kb[i] - kk * kh[i]) * ir + ph * kh[i];
}

// sin(x)/x and its derivative for |x| < 0.5 (FP32 series)
__host__ __device__ __forceinline__ float sinc_ser(float x)
{
    const float z = x * x;
    return fmaf(z, fmaf(z, fmaf(z, -1.98412698e-4f, 8.33333333e-3f), -0.166666667f), 1.0f);
}
__host__ __device__ __forceinline__ float dsinc_ser(float x)
{
    const float z = x * x;
    return x * fmaf(z, fmaf(z, fmaf(z, 2.20458554e-5f, -1.19047619e-3f), 3.33333333e-2f), -0.333333333f);
}

// Forward of the model in FP64 (the function the adjoint differentiates): adds
// s' (g I + h J) C to acc (re/im, Fortran pair order), s' = e^{i k (|x - c| - |x|)}
// (i / lambda) F with the phase formed without cancellation; the caller applies the
// per-bin e^{i k |x|} and T_phi after the beam sum (model_finish).  Agrees with the
// reference's formula to rounding and is smooth to ~1e-15, so finite differences
// with steps of 1e-4 um resolve its derivatives.
__host__ __device__ inline void model_bin_fp64(const AdjBeam& B, double x, double y, double z, double xt,
                                               double yt, double rad, double waveno, double lambda,
                                               double acc[8])
{
    const double k = waveno;
    const double d[3] = {x - B.c[0], y - B.c[1], z - B.c[2]};
    const double rho = sqrt(dot3(d, d));
    const double kh[3] = {d[0] / rho, d[1] / rho, d[2] / rho};
    const double rc = x * B.c[0] + y * B.c[1] + z * B.c[2];
    const double drho = (B.cc - 2.0 * rc) / (rho + rad);
    double Es, Ec;
    sincos_f64(k * drho, &Es, &Ec);
    const double q3[3] = {k * (B.P[0] - kh[0]), k * (B.P[1] - kh[1]), k * (B.P[2] - kh[2])};
    double F[2], Gx[2], Gy[2];
    tri_ff_grad(B, dot3(q3, B.ex), dot3(q3, B.ey), F, Gx, Gy);
    const double k1 = dot3(kh, B.ex), k2 = dot3(kh, B.ey), k3 = dot3(kh, B.ez);
    const double k2c = fabs(k2) > K2CLAMP ? copysign(K2CLAMP, k2) : k2;
    const double sq = sqrt(1.0 - k2c * k2c), s13 = sqrt(1.0 - k2 * k2);
    double hv0 = kh[1], hv1 = -kh[0];
    if (!(fabs(kh[2]) < FWD_DOT)) {
        hv0 = yt - B.c[1];
        hv1 = -(xt - B.c[0]);
    }
    const double nh = sqrt(hv0 * hv0 + hv1 * hv1);
    const double hc[3] = {hv0 / nh, hv1 / nh, 0.0};
    double mvec[3], uu[3];
    for (int i = 0; i < 3; i++) mvec[i] = (B.ey[i] - k2c * kh[i]) / sq;
    cross3(kh, B.ey, uu);
    const double p = dot3(hc, mvec), qh = dot3(hc, uu) / s13;
    const double da = -0.5 * (B.Kz * sq + k3 / sq), db = -0.5 * k1 * k2c / sq;
    const double g = p * da + qh * db, h = p * db - qh * da;
    const double EFr = Ec * F[0] - Es * F[1], EFi = Ec * F[1] + Es * F[0];
    const double sr = -EFi / lambda, si = EFr / lambda;
    // (g I + h J) C: row 1 = g C1 + h C2, row 2 = g C2 - h C1
    for (int cc = 0; cc < 2; cc++) {
        const int i1 = 2 * (0 + 2 * cc), i2 = 2 * (1 + 2 * cc);
        const double r1r = g * B.C[i1] + h * B.C[i2], r1i = g * B.C[i1 + 1] + h * B.C[i2 + 1];
        const double r2r = g * B.C[i2] - h * B.C[i1], r2i = g * B.C[i2 + 1] - h * B.C[i1 + 1];
        acc[i1] += sr * r1r - si * r1i;
        acc[i1 + 1] += sr * r1i + si * r1r;
        acc[i2] += sr * r2r - si * r2i;
        acc[i2 + 1] += sr * r2i + si * r2r;
    }
}

// after the beam sum of model_bin_fp64: x T_phi, then e^{i k |x|}
__host__ __device__ inline void model_finish(double acc[8], double cp, double sp, double kr_rev)
{
    apply_temp_rot1(acc, cp, sp);
    apply_bin_phase(acc, kr_rev);
}

// Mixed-precision adjoint body: the geometry, q and every phase in FP64 (reduced to a
// half period before FP32 polynomials), the form factor's edge terms, the
// polarisation and its adjoint in FP32, the accumulators in FP64.  Near q -> 0 an
// FP32 series of the form factor and its gradient.  Edge phases come from the three vertex half-phase
// exponentials w_l = e^{i alpha_l / 2}: e^{i mu_j} = w_j w_{j+1},
// e^{i K_j / 2} = w_{j+1} conj(w_j); sinc(K_j / 2) from its series for |K_j| < 1.
// Mixed body: FP64 geometry, q and phases; the far-field phase e^{i k rho} is folded into
// the vertex exponentials, so F and dF/dq below carry it (F' = e^{i k rho} F); FP32
// elsewhere; contributions go to FP32 partial sums (the caller adds them into FP64 every
// ADJ_GROUP bins).  lam: Lam' of this bin in FP32.
__host__ __device__ inline void adj_bin_mixed(const AdjBeam& B, const AdjBin& bn, const float* __restrict__ lam,
                                              double waveno, double lambda, AdjAccF& acc)
{
    const double k = waveno;
    const double dx = bn.x - B.c[0], dy = bn.y - B.c[1], dz = bn.z - B.c[2];
    const double s2 = fma(dx, dx, fma(dy, dy, dz * dz));
    double y0 = (double)abt_rsqf((float)s2);
    double e = fma(-s2 * y0, y0, 1.0);
    y0 = fma(0.5 * y0, e, y0);
    e = fma(-s2 * y0, y0, 1.0);
    const double ir = fma(0.5 * y0, e, y0);              // 1 / rho
    const double kh[3] = {dx * ir, dy * ir, dz * ir};
    const double krev = k * 0.15915494309189535;
    const double kd = krev * (s2 * ir);                   // k rho / 2 pi
    const double q3[3] = {k * (B.P[0] - kh[0]), k * (B.P[1] - kh[1]), k * (B.P[2] - kh[2])};
    const double qx = fma(q3[0], B.ex[0], fma(q3[1], B.ex[1], q3[2] * B.ex[2]));
    const double qy = fma(q3[0], B.ey[0], fma(q3[1], B.ey[1], q3[2] * B.ey[2]));
    const float fqx = (float)qx, fqy = (float)qy;
    const float q2 = fmaf(fqx, fqx, fqy * fqy);
    float F[2], Gx[2], Gy[2];
    if ((double)q2 * B.qa2 < 2.25) {
        // FP32 series about q = 0 (|alpha_l| <= 1.5: 14 terms leave < 1e-9)
        float al[3];
        for (int l = 0; l < 3; l++) al[l] = fmaf(fqx, B.fxl[l], fqy * B.fyl[l]);
        const float e1 = al[0] + al[1] + al[2];
        const float e2 = fmaf(al[0], al[1], fmaf(al[1], al[2], al[2] * al[0]));
        const float e3 = al[0] * al[1] * al[2];
        float hm3 = 0.f, hm2 = 0.f, hm1 = 1.f, hprev = 1.f, inv = 0.5f;
        float d[3] = {0.f, 0.f, 0.f}, dr[3] = {0.f, 0.f, 0.f}, di[3] = {0.f, 0.f, 0.f};
        float fr = 0.5f, fi = 0.f;
        for (int n = 1; n <= 14; n++) {
            for (int l = 0; l < 3; l++) d[l] = fmaf(al[l], d[l], hprev);
            const float hn = fmaf(e1, hm1, fmaf(-e2, hm2, e3 * hm3));
            inv = inv / (float)(n + 2);
            const float t = hn * inv;
            switch (n & 3) {
                case 0: fr += t; for (int l = 0; l < 3; l++) dr[l] += d[l] * inv; break;
                case 1: fi += t; for (int l = 0; l < 3; l++) di[l] += d[l] * inv; break;
                case 2: fr -= t; for (int l = 0; l < 3; l++) dr[l] -= d[l] * inv; break;
                default: fi -= t; for (int l = 0; l < 3; l++) di[l] -= d[l] * inv; break;
            }
            hm3 = hm2; hm2 = hm1; hm1 = hn; hprev = hn;
        }
        const float A2 = (float)(2.0 * B.area);
        // times the far-field phase e^{i k rho}
        float Ec, Es;
        expi_pi(2.0 * kd, &Ec, &Es);
        const float Er = A2 * Ec, Ei = A2 * Es;
        const float gxr = dr[0] * B.fxl[0] + dr[1] * B.fxl[1] + dr[2] * B.fxl[2];
        const float gxi = di[0] * B.fxl[0] + di[1] * B.fxl[1] + di[2] * B.fxl[2];
        const float gyr = dr[0] * B.fyl[0] + dr[1] * B.fyl[1] + dr[2] * B.fyl[2];
        const float gyi = di[0] * B.fyl[0] + di[1] * B.fyl[1] + di[2] * B.fyl[2];
        F[0] = Er * fr - Ei * fi;
        F[1] = Er * fi + Ei * fr;
        Gx[0] = Er * gxr - Ei * gxi;
        Gx[1] = Er * gxi + Ei * gxr;
        Gy[0] = Er * gyr - Ei * gyi;
        Gy[1] = Er * gyi + Ei * gyr;
    } else {
        // vertex half phases alpha_l / 2 plus k rho / 2: w_l = e^{i pi t_l}, t_l = alpha_l / 2 pi + k rho / 2 pi
        float wr[3], wi[3];
        double al[3];
        for (int l = 0; l < 3; l++) {
            al[l] = fma(qx, B.xr[l], fma(qy, B.yr[l], kd));
            expi_pi(al[l], &wr[l], &wi[l]);
        }
        float sr = 0.f, si = 0.f, gxr = 0.f, gxi = 0.f, gyr = 0.f, gyi = 0.f;
        for (int j = 0; j < 3; j++) {
            const int jn = j == 2 ? 0 : j + 1;
            const float ex = B.fxl[jn] - B.fxl[j], ey = B.fyl[jn] - B.fyl[j];
            const float mx = 0.5f * (B.fxl[j] + B.fxl[jn]), my = 0.5f * (B.fyl[j] + B.fyl[jn]);
            const float x = 3.14159265358979f * (float)(al[jn] - al[j]);   // K_j / 2
            // e^{i x} = w_{j+1} conj(w_j), e^{i mu} = w_j w_{j+1}
            const float cx = wr[jn] * wr[j] + wi[jn] * wi[j];
            const float sx = wi[jn] * wr[j] - wr[jn] * wi[j];
            const float Er = wr[j] * wr[jn] - wi[j] * wi[jn];
            const float Ei = wr[j] * wi[jn] + wi[j] * wr[jn];
            const bool small = fabsf(x) < 0.5f;
            const float ix = abt_rcpf(x);
            const float Sq = sx * ix;
            const float S = small ? sinc_ser(x) : Sq;
            const float dS = 0.5f * (small ? dsinc_ser(x) : (cx - Sq) * ix);
            const float cj = B.fso * fmaf(fqx, ey, -fqy * ex);
            const float t = cj * S;
            sr = fmaf(t, Er, sr);
            si = fmaf(t, Ei, si);
            const float ax = fmaf(B.fso * ey, S, cj * dS * ex), ay = fmaf(-B.fso * ex, S, cj * dS * ey);
            gxr += ax * Er - t * mx * Ei;
            gxi += ax * Ei + t * mx * Er;
            gyr += ay * Er - t * my * Ei;
            gyi += ay * Ei + t * my * Er;
        }
        const float iq2 = abt_rcpf(q2);
        F[0] = -si * iq2;
        F[1] = sr * iq2;
        const float tx = 2.0f * fqx * iq2, ty = 2.0f * fqy * iq2;
        Gx[0] = -gxi * iq2 - tx * F[0];
        Gx[1] = gxr * iq2 - tx * F[1];
        Gy[0] = -gyi * iq2 - ty * F[0];
        Gy[1] = gyr * iq2 - ty * F[1];
    }
    // polarisation (FP32; the k_y clamp decided in FP64, as the reference does)
    const float kf[3] = {(float)kh[0], (float)kh[1], (float)kh[2]};
    const double k2d = fma(kh[0], B.ey[0], fma(kh[1], B.ey[1], kh[2] * B.ey[2]));
    const bool clamp = fabs(k2d) > K2CLAMP;
    const float k1 = fmaf(kf[0], B.fex[0], fmaf(kf[1], B.fex[1], kf[2] * B.fex[2]));
    const float k3 = fmaf(kf[0], B.fez[0], fmaf(kf[1], B.fez[1], kf[2] * B.fez[2]));
    const float k2 = (float)k2d;
    const float s13q = fmaf(k1, k1, k3 * k3);                    // 1 - k2^2
    const bool fb = !(fabs(kh[2]) < FWD_DOT);
    // Away from the clamp and the forward fallback, hc = unit(inc x khat) is normal to
    // khat, so hc.m = D' P0 and hc.unit(k x m) = D' Q0 with D' = rh / sqrt(1 - k2^2),
    // rh = 1 / |inc x khat|, P0 = e_y . (kh_y, -kh_x, 0), Q0 = e_y . ((inc.khat) khat - inc),
    // and g = D (P0 A + Q0 B), h = D (P0 B - Q0 A) with D = rh / (1 - k2^2),
    // A = -(K_z (1 - k2^2) + k3) / 2, B = -k1 k2 / 2.  P0, Q0 in FP64: both vanish at khat = +-e_y.
    const bool fast = !clamp && !fb;
    float g, h;
    float fP0 = 0.f, fQ0 = 0.f, fA = 0.f, fB = 0.f, fD = 0.f, rh = 0.f, is2 = 0.f;
    float k2c = 0.f, r13 = 0.f, s13 = 0.f, rsq = 0.f, sq = 0.f, inh = 0.f, hc0 = 0.f, hc1 = 0.f;
    float mv[3] = {0.f, 0.f, 0.f}, ev[3] = {0.f, 0.f, 0.f}, p = 0.f, qh = 0.f, da = 0.f, db = 0.f;
    if (fast) {
        fP0 = (float)fma(B.ey[0], kh[1], -B.ey[1] * kh[0]);
        fQ0 = (float)fma(-kh[2], k2d, B.ey[2]);
        rh = abt_rsqf(fmaf(kf[0], kf[0], kf[1] * kf[1]));
        is2 = abt_rcpf(s13q);
        fD = rh * is2;
        fA = -0.5f * fmaf(B.fKz, s13q, k3);
        fB = -0.5f * k1 * k2;
        g = fD * fmaf(fP0, fA, fQ0 * fB);
        h = fD * fmaf(fP0, fB, -fQ0 * fA);
    } else {
        k2c = clamp ? copysignf(0.99999898672103882f, k2) : k2;
        r13 = abt_rsqf(s13q);
        s13 = s13q * r13;
        const float sqq = clamp ? 2.02655679e-06f : s13q;
        rsq = clamp ? abt_rsqf(sqq) : r13;
        sq = sqq * rsq;
        float hvx, hvy;
        if (fb) {
            hvx = (float)(bn.yt - B.c[1]);
            hvy = (float)(-(bn.xt - B.c[0]));
        } else {
            hvx = kf[1];
            hvy = -kf[0];
        }
        inh = abt_rsqf(fmaf(hvx, hvx, hvy * hvy));
        hc0 = hvx * inh;
        hc1 = hvy * inh;
        const float uu[3] = {kf[1] * B.fey[2] - kf[2] * B.fey[1], kf[2] * B.fey[0] - kf[0] * B.fey[2],
                             kf[0] * B.fey[1] - kf[1] * B.fey[0]};
        for (int i = 0; i < 3; i++) {
            mv[i] = (B.fey[i] - k2c * kf[i]) * rsq;
            ev[i] = uu[i] * r13;
        }
        p = fmaf(hc0, mv[0], hc1 * mv[1]);
        qh = fmaf(hc0, ev[0], hc1 * ev[1]);
        da = -0.5f * fmaf(B.fKz, sq, k3 * rsq);
        db = -0.5f * k1 * k2c * rsq;
        g = fmaf(p, da, qh * db);
        h = fmaf(p, db, -qh * da);
    }
    // s = E (i / lambda) F  (F carries E)
    const float il = (float)(1.0 / lambda);
    const float sr = -F[1] * il, si = F[0] * il;
    float L[8];
    for (int q = 0; q < 8; q++) L[q] = lam[q];
    float z1r = 0.f, z1i = 0.f, z2r = 0.f, z2i = 0.f;
    for (int cc = 0; cc < 2; cc++) {
        const int i1 = 2 * (0 + 2 * cc), i2 = 2 * (1 + 2 * cc);
        const float l1r = L[i1], l1i = L[i1 + 1], l2r = L[i2], l2i = L[i2 + 1];
        const float c1r = B.fC[i1], c1i = B.fC[i1 + 1], c2r = B.fC[i2], c2i = B.fC[i2 + 1];
        z1r += l1r * c1r + l1i * c1i + l2r * c2r + l2i * c2i;
        z1i += l1r * c1i - l1i * c1r + l2r * c2i - l2i * c2r;
        z2r += l1r * c2r + l1i * c2i - (l2r * c1r + l2i * c1i);
        z2i += l1r * c2i - l1i * c2r - (l2r * c1i - l2i * c1r);
    }
    for (int cc = 0; cc < 2; cc++) {
        const int i1 = 2 * (0 + 2 * cc), i2 = 2 * (1 + 2 * cc);
        const float r1r = g * L[i1] - h * L[i2], r1i = g * L[i1 + 1] - h * L[i2 + 1];
        const float r2r = g * L[i2] + h * L[i1], r2i = g * L[i2 + 1] + h * L[i1 + 1];
        acc.GC[i1] += sr * r1r + si * r1i;
        acc.GC[i1 + 1] += sr * r1i - si * r1r;
        acc.GC[i2] += sr * r2r + si * r2i;
        acc.GC[i2 + 1] += sr * r2i - si * r2r;
    }
    const float gg = sr * z1r - si * z1i, gh = sr * z2r - si * z2i;
    const float mur = g * z1r + h * z2r, mui = g * z1i + h * z2i;
    // nu = mu E (i / lambda); E is in F and dF/dq
    const float nur = -mui * il, nui = mur * il;
    const float nFr = nur * F[0] - nui * F[1], nFi = nur * F[1] + nui * F[0];
    const float Rx = nur * Gx[0] - nui * Gx[1];
    const float Ry = nur * Gy[0] - nui * Gy[1];
    acc.S0 += nFr;
    for (int i = 0; i < 3; i++) {
        const float fq = (float)q3[i];
        acc.Qx[i] = fmaf(fq, Rx, acc.Qx[i]);
        acc.Qy[i] = fmaf(fq, Ry, acc.Qy[i]);
    }
    const float kw = (float)k;
    float kb[3];
    for (int i = 0; i < 3; i++) kb[i] = -kw * (Rx * B.fex[i] + Ry * B.fey[i]);
    float k1b, k2b, k3b, Kzb, eyb[3];
    float cbf[3] = {0.f, 0.f, 0.f};
    if (fast) {
        // reverse of g, h: the 1 - k2^2 inside A and D differentiated as a function of k2
        const float Pb = fD * fmaf(gg, fA, gh * fB);
        const float Qb = fD * fmaf(gg, fB, -gh * fA);
        const float Ab = fD * fmaf(gg, fP0, -gh * fQ0);
        const float Bb = fD * fmaf(gg, fQ0, gh * fP0);
        const float w = fmaf(gg, g, gh * h);                 // (dL/dD) D
        const float rw = w * rh * rh;
        Kzb = -0.5f * s13q * Ab;
        k3b = -0.5f * Ab;
        k1b = -0.5f * k2 * Bb;
        k2b = B.fKz * k2 * Ab - 0.5f * k1 * Bb + 2.f * k2 * w * is2 - kf[2] * Qb;
        kb[0] += -rw * kf[0] - B.fey[1] * Pb;
        kb[1] += -rw * kf[1] + B.fey[0] * Pb;
        kb[2] += -k2 * Qb;
        eyb[0] = kf[1] * Pb;
        eyb[1] = -kf[0] * Pb;
        eyb[2] = Qb;
    } else {
        const float pb = gg * da + gh * db, qb = gg * db - gh * da;
        const float dab = gg * p - gh * qh, dbb = gg * qh + gh * p;
        Kzb = -0.5f * sq * dab;
        float sqb = -0.5f * (B.fKz - k3 * rsq * rsq) * dab + 0.5f * k1 * k2c * rsq * rsq * dbb;
        k3b = -0.5f * dab * rsq;
        k1b = -0.5f * k2c * rsq * dbb;
        float k2cb = -0.5f * k1 * rsq * dbb;
        const float hcb0 = pb * mv[0] + qb * ev[0], hcb1 = pb * mv[1] + qb * ev[1];
        const float mb[3] = {pb * hc0, pb * hc1, 0.f};
        k2cb += -(mb[0] * kf[0] + mb[1] * kf[1]) * rsq;
        sqb += -(mb[0] * mv[0] + mb[1] * mv[1]) * rsq;
        const float ub[3] = {qb * hc0 * r13, qb * hc1 * r13, 0.f};
        const float s13b = -qb * (hc0 * ev[0] + hc1 * ev[1]) * r13;
        for (int i = 0; i < 3; i++) {
            eyb[i] = mb[i] * rsq;
            kb[i] += -k2c * rsq * mb[i];
        }
        // kb += e_y x ub; eyb += ub x khat  (ub has no z component)
        kb[0] += -B.fey[2] * ub[1];
        kb[1] += B.fey[2] * ub[0];
        kb[2] += B.fey[0] * ub[1] - B.fey[1] * ub[0];
        eyb[0] += ub[1] * kf[2];
        eyb[1] += -ub[0] * kf[2];
        eyb[2] += ub[0] * kf[1] - ub[1] * kf[0];
        k2b = -k2 * r13 * s13b;
        k2cb += -k2c * rsq * sqb;
        if (!clamp) k2b += k2cb;
        (void)s13;
        const float hdot = hcb0 * hc0 + hcb1 * hc1;
        const float hvbx = (hcb0 - hdot * hc0) * inh, hvby = (hcb1 - hdot * hc1) * inh;
        const float tx = -hvby, ty = hvbx;
        if (fb) {
            cbf[0] = -tx;
            cbf[1] = -ty;
        } else {
            kb[0] += tx;
            kb[1] += ty;
        }
    }
    for (int i = 0; i < 3; i++) {
        kb[i] += k1b * B.fex[i] + k2b * B.fey[i] + k3b * B.fez[i];
        acc.exb[i] += k1b * kf[i];
        acc.eyb[i] += eyb[i] + k2b * kf[i];
        acc.ezb[i] += k3b * kf[i] + Kzb * B.fP[i];
    }
    const float kk = kb[0] * kf[0] + kb[1] * kf[1] + kb[2] * kf[2];
    const float irf = (float)ir;
    const float ph = kw * nFi;
    for (int i = 0; i < 3; i++) acc.cb[i] += cbf[i] - (kb[i] - kk * kf[i]) * irf + ph * kf[i];
}

// Per-beam gradients from the accumulated sums: gv[9] = dL/dv_l (x,y,z per vertex, the
// record's vertex order), ga[8] = dL/d ampl (re/im pairs, the beam's own sign).
// u: the centre-relative vertices (record R_V20), P the propagation vector.
__host__ __device__ inline void adj_finish(const AdjBeam& B, const double* __restrict__ u, double sigma,
                                           const AdjAcc& acc, double* gv, double* ga)
{
    const double v0[3] = {u[0], u[1], u[2]}, v1[3] = {u[3], u[4], u[5]}, v2[3] = {u[6], u[7], u[8]};
    double a[3], b[3], N[3];
    for (int i = 0; i < 3; i++) { a[i] = v1[i] - v0[i]; b[i] = v2[i] - v0[i]; }
    cross3(a, b, N);
    const double nN = sqrt(dot3(N, N));
    const double n[3] = {N[0] / nN, N[1] / nN, N[2] / nN};
    // frame: e_x = e_y x e_z, e_y = w / |w|, w = e_z x P, e_z = -n
    double eyb[3], ezb[3], t[3];
    cross3(B.ez, acc.exb, t);
    for (int i = 0; i < 3; i++) eyb[i] = acc.eyb[i] + t[i];
    cross3(acc.exb, B.ey, t);
    for (int i = 0; i < 3; i++) ezb[i] = acc.ezb[i] + t[i];
    double w[3];
    cross3(B.ez, B.P, w);
    const double nw = sqrt(dot3(w, w));
    const double eye = dot3(eyb, B.ey);
    double wb[3];
    for (int i = 0; i < 3; i++) wb[i] = (eyb[i] - eye * B.ey[i]) / nw;
    cross3(B.P, wb, t);
    for (int i = 0; i < 3; i++) ezb[i] += t[i];
    const double ezn = dot3(ezb, n);
    double Nb[3];
    for (int i = 0; i < 3; i++) Nb[i] = -(ezb[i] - ezn * n[i]) / nN;
    // N = a x b
    double ab[3], bb[3];
    cross3(b, Nb, ab);
    cross3(Nb, a, bb);
    const double* V[3] = {v0, v1, v2};
    for (int l = 0; l < 3; l++) {
        const double* p1 = V[l == 2 ? 0 : l + 1];
        const double* p2 = V[l == 0 ? 2 : l - 1];
        double e[3], al[3];
        for (int i = 0; i < 3; i++) e[i] = p1[i] - p2[i];
        cross3(e, n, al);
        for (int i = 0; i < 3; i++) al[i] /= nN;
        const double ax = dot3(al, B.ex), ay = dot3(al, B.ey);
        for (int i = 0; i < 3; i++)
            gv[3 * l + i] = acc.S0 * al[i] + acc.Qx[i] * ax + acc.Qy[i] * ay + acc.cb[i] / 3.0;
    }
    for (int i = 0; i < 3; i++) {
        gv[3 + i] += ab[i];
        gv[6 + i] += bb[i];
        gv[i] -= ab[i] + bb[i];
    }
    for (int q = 0; q < 8; q++) ga[q] = sigma * acc.GC[q];
}

}  // namespace abt
