// frame.h -- per-beam set-up of `diffraction` (diff_mod.f90:515-574 and the
// aperture part of contour_integral, diff_mod.f90:334-354), shared by the host
// preparation (glibc libm, bit-reproducible against a CPU FP64 run) and the
// device preparation kernel.  Operation order follows the reference; FMA
// contraction is off so host and device round each operation alike.  Cosine and
// sine of one angle come from one sincos call: that is what gfortran -O3 makes of
// the reference's pairs, and glibc's sincos differs from separate cos and sin in
// the last bit for a few arguments in a thousand -- enough to change the frame of
// an aperture at normal incidence, whose in-plane angle is atan2 of rounding noise.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>
#include "records.h"

namespace abt {

struct M3 { double a[3][3]; };

__host__ __device__ inline void m3_mul(const M3& x, const M3& y, M3& out)
{
#pragma clang fp contract(off)
    M3 t;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
    out = t;
}

__host__ __device__ inline void m3_vec(const M3& x, const double v[3], double out[3])
{
#pragma clang fp contract(off)
    double t0 = x.a[0][0] * v[0] + x.a[0][1] * v[1] + x.a[0][2] * v[2];
    double t1 = x.a[1][0] * v[0] + x.a[1][1] * v[1] + x.a[1][2] * v[2];
    double t2 = x.a[2][0] * v[0] + x.a[2][1] * v[1] + x.a[2][2] * v[2];
    out[0] = t0; out[1] = t1; out[2] = t2;
}

// get_rotation_matrix2 (diff_mod.f90:200-289); a1, b1: first two com-relative vertices
__host__ __device__ inline void aperture_rotation(const double a1[3], const double b1[3], M3& rot)
{
#pragma clang fp contract(off)
    double th1 = atan(a1[0] / a1[1]), c1, s1;
    sincos(th1, &s1, &c1);
    M3 r1 = {{{c1, -s1, 0}, {s1, c1, 0}, {0, 0, 1}}};
    double a2[3], b2[3];
    m3_vec(r1, a1, a2); m3_vec(r1, b1, b2);
    double th2 = -atan(a2[2] / a2[1]), c2, s2;
    sincos(th2, &s2, &c2);
    M3 r2 = {{{1, 0, 0}, {0, c2, -s2}, {0, s2, c2}}};
    double a3[3], b3[3];
    m3_vec(r2, a2, a3); m3_vec(r2, b2, b3);
    double th3 = atan(b3[2] / b3[0]), c3, s3;
    sincos(th3, &s3, &c3);
    M3 r3 = {{{c3, 0, s3}, {0, 1, 0}, {-s3, 0, c3}}};
    double a4[3], b4[3];
    m3_vec(r3, a3, a4); m3_vec(r3, b3, b4);
    M3 t21, t321;
    m3_mul(r2, r1, t21);
    m3_mul(r3, t21, t321);
    if (a4[0] * b4[1] - a4[1] * b4[0] > 0) {
        M3 r4 = {{{-1, 0, 0}, {0, 1, 0}, {0, 0, -1}}};
        m3_mul(r4, t321, rot);
    } else {
        rot = t321;
    }
}

// Fill one beam record (and, for nv > 3, its vertex records).
// verts: nv*3 doubles; rec: REC doubles; vrec: nv*VREC doubles or nullptr.
__host__ __device__ inline void make_beam_record(const double* ampl, const double* vk7,
                                                 const double* prop, double fov, int64_t nv,
                                                 const double* verts, double lambda,
                                                 int64_t voff, double* rec, double* vrec)
{
#pragma clang fp contract(off)
    for (int q = 0; q < REC; q++) rec[q] = 0.0;
    double com[3];
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int64_t i = 0; i < nv; i++) s = s + verts[3 * i + r];
        com[r] = s / (double)nv;
    }
    double a1[3], b1[3];
    for (int r = 0; r < 3; r++) {
        a1[r] = verts[r] - com[r];
        b1[r] = verts[3 + r] - com[r];
    }
    M3 rot;
    aperture_rotation(a1, b1, rot);
    double prop0[3] = {prop[0], prop[1], prop[2]};
    double perp0[3] = {vk7[0], vk7[1], vk7[2]};
    double prop1[3], perp1[3];
    m3_vec(rot, prop0, prop1);
    m3_vec(rot, perp0, perp1);
    double angle = -atan2(prop1[1], prop1[0]);
    double ca, sa;
    sincos(angle, &sa, &ca);
    M3 rot2 = {{{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}}};
    double prop2[3], perp2[3];
    m3_vec(rot2, prop1, prop2);
    m3_vec(rot2, perp1, perp2);
    // cross(perp2, prop2, e_par2, .true.)  (misc_submod.f90:2367-2390)
    double e0 = perp2[1] * prop2[2] - perp2[2] * prop2[1];
    double e1 = perp2[2] * prop2[0] - perp2[0] * prop2[2];
    double e2 = perp2[0] * prop2[1] - perp2[1] * prop2[0];
    double nf = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    (void)e0; (void)e1;
    bool anti = (e2 / nf) > ANTI_THR;
    double inc[3] = {0.0, 0.0, -1.0}, t[3], inc2[3];
    m3_vec(rot, inc, t);
    m3_vec(rot2, t, inc2);
    M3 rr;
    m3_mul(rot2, rot, rr);

    for (int r = 0; r < 3; r++) rec[R_COM + r] = com[r];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rec[R_RR + 3 * i + j] = rr.a[i][j];
    for (int r = 0; r < 3; r++) {
        rec[R_PROP2 + r] = prop2[r];
        rec[R_INC2 + r] = inc2[r];
        rec[R_PROP0 + r] = prop0[r];
    }
    for (int q = 0; q < 8; q++) rec[R_AMPL + q] = anti ? -ampl[q] : ampl[q];
    rec[R_SIGMA] = anti ? -1.0 : 1.0;
    rec[R_COSFOV] = cos(fov);
    rec[R_NV] = (double)nv;
    rec[R_VOFF] = (double)voff;
    double waveno = 2.0 * PI_R4 / lambda;
    double krev = waveno / 6.283185307179586476925;
    for (int r = 0; r < 3; r++) rec[R_KPROP + r] = prop0[r] * krev;
    rec[R_CC] = com[0] * com[0] + com[1] * com[1] + com[2] * com[2];
    rec[R_KY1] = 1.0 - prop2[1] * prop2[1];
    rec[R_FRAC] = 1.0 / sqrt(rec[R_KY1]);

    // aperture edges in the aperture frame (contour_integral, diff_mod.f90:334-354)
    double amax = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < nv; j++) {
        int64_t jp = (j == nv - 1) ? 0 : j + 1;
        double v20[3], v20n[3];
        for (int r = 0; r < 3; r++) {
            v20[r] = verts[3 * j + r] - com[r];
            v20n[r] = verts[3 * jp + r] - com[r];
        }
        double v1[3], v1n[3];
        m3_vec(rr, v20, v1);
        m3_vec(rr, v20n, v1n);
        double m = (v1n[1] - v1[1]) / (v1n[0] - v1[0]);
        double n = 1 / m;
        if (fabs(m) > DBL_MAX) m = 1e6;
        if (fabs(n) > DBL_MAX) n = 1e6;
        if (fabs(m) < 1e-9) n = 1e6;
        if (fabs(n) < 1e-9) m = 1e6;
        double dx = v1n[0] - v1[0], dy = v1n[1] - v1[1];
        double rr2 = sqrt(v20[0] * v20[0] + v20[1] * v20[1] + v20[2] * v20[2]);
        amax = rr2 > amax ? rr2 : amax;
        // vector area about com, for the forward series branch
        double cx = v20[1] * v20n[2] - v20[2] * v20n[1];
        double cy = v20[2] * v20n[0] - v20[0] * v20n[2];
        double cz = v20[0] * v20n[1] - v20[1] * v20n[0];
        ax += cx; ay += cy; az += cz;
        if (nv <= 3) {
            rec[R_EX + j] = v1[0];
            rec[R_EY + j] = v1[1];
            rec[R_EM + j] = m;
            rec[R_EN + j] = n;
            rec[R_EDX + j] = dx;
            rec[R_EDY + j] = dy;
            for (int r = 0; r < 3; r++) rec[R_V20 + 3 * j + r] = v20[r];
        }
        if (vrec) {
            double* w = vrec + VREC * j;
            for (int q = 0; q < VREC; q++) w[q] = 0.0;
            w[V_X] = v1[0]; w[V_Y] = v1[1]; w[V_M] = m; w[V_N] = n;
            w[V_DX] = dx; w[V_DY] = dy;
            for (int r = 0; r < 3; r++) w[V_V20 + r] = v20[r];
        }
    }
    rec[R_AMAX] = amax;
    // signed area in the aperture frame: the frame normal is rr^T (0,0,1)
    rec[R_AREA] = 0.5 * (rr.a[2][0] * ax + rr.a[2][1] * ay + rr.a[2][2] * az);

    // constants of the mixed-precision kernel
    for (int j = 0; j < 3; j++) {
        rec[R_KRR + j] = krev * rr.a[0][j];
        rec[R_KRR + 3 + j] = krev * rr.a[1][j];
    }
    rec[R_KP2] = krev * prop2[0];
    rec[R_KP2 + 1] = krev * prop2[1];
    const double* A = rec + R_AMPL;   // sign-flipped amplitude matrix
    // r+ = (A(1,:) - i A(2,:))/2 and r- = (A(1,:) + i A(2,:))/2, columns 1 and 2
    float* rg = reinterpret_cast<float*>(rec + R_G32);
    rg[G_RP + 0] = (float)(0.5 * (A[0] + A[3])); rg[G_RP + 1] = (float)(0.5 * (A[1] - A[2]));
    rg[G_RP + 2] = (float)(0.5 * (A[4] + A[7])); rg[G_RP + 3] = (float)(0.5 * (A[5] - A[6]));
    rg[G_RM + 0] = (float)(0.5 * (A[0] - A[3])); rg[G_RM + 1] = (float)(0.5 * (A[1] + A[2]));
    rg[G_RM + 2] = (float)(0.5 * (A[4] - A[7])); rg[G_RM + 3] = (float)(0.5 * (A[5] + A[6]));
    rec[R_RP64 + 0] = 0.5 * (A[0] + A[3]); rec[R_RP64 + 1] = 0.5 * (A[1] - A[2]);
    rec[R_RP64 + 2] = 0.5 * (A[4] + A[7]); rec[R_RP64 + 3] = 0.5 * (A[5] - A[6]);
    rec[R_RM64 + 0] = 0.5 * (A[0] - A[3]); rec[R_RM64 + 1] = 0.5 * (A[1] + A[2]);
    rec[R_RM64 + 2] = 0.5 * (A[4] - A[7]); rec[R_RM64 + 3] = 0.5 * (A[5] + A[6]);
    // Karczewski coefficients over krev (the krev-scaled polarisation of the restructured bodies)
    const double isig = sqrt(rec[R_KY1]);
    rec[R_C64 + 0] = -0.5 * prop2[2] * rec[R_FRAC] / krev;
    rec[R_C64 + 1] = -0.5 * isig / krev;
    rec[R_C64 + 2] = 0.5 * prop2[0] * prop2[1] * rec[R_FRAC] / krev;
    rec[R_ISIG] = isig;
    rg[G_C1] = (float)rec[R_C64 + 0];
    rg[G_C2] = (float)rec[R_C64 + 1];
    rg[G_C3] = (float)rec[R_C64 + 2];
    float* rf = reinterpret_cast<float*>(rec + R_F32);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rf[F_RR + 3 * i + j] = (float)rr.a[i][j];
    rf[F_KX] = (float)prop2[0];
    rf[F_KY] = (float)prop2[1];
    rf[F_KZ] = (float)prop2[2];
    rf[F_KXY] = (float)(prop2[0] * prop2[1]);
    rf[F_SIG] = (float)rec[R_FRAC];
    rf[F_ISIG] = (float)sqrt(rec[R_KY1]);
    const double tpa = 6.283185307179586 * amax;
    rf[F_QA2] = (float)(tpa * tpa);
    rf[F_NV] = (float)nv;
    double c2 = 0.0;
    for (int q = 0; q < 8; q++) c2 += A[q] * A[q];
    rf[F_C2] = (float)c2;
    for (int j = 0; j < 3 && j < nv; j++) {
        rf[F_EDX + j] = (float)rec[R_EDX + j];
        rf[F_EDY + j] = (float)rec[R_EDY + j];
    }
}

}  // namespace abt
