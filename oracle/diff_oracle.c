/*
 * diff_oracle.c -- CPU FP64 restatement of PBT's far-field diffraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker (or the timed CPU baseline), never as the product.
 * The product path is the HIP library under abt_amd/ and fails loudly when
 * that library is missing.
 *
 * It restates, in the reference's operation order and with its REAL*4
 * constants widened to double, the following reference routines:
 *   make_far_field_bins   diff_mod.f90:827-855 (+ meshgrid_real misc_submod.f90:1743-1779)
 *   get_rotation_matrix2  diff_mod.f90:200-289
 *   karczewski            diff_mod.f90:124-198
 *   contour_integral      diff_mod.f90:291-444
 *   diffraction           diff_mod.f90:446-672
 *   diff_main             diff_mod.f90:674-825 (sequential, -mt 0 order)
 *   cross                 misc_submod.f90:2367-2390 (normalises by default)
 *   simpne                misc_submod.f90:2392-2531
 *   ampl_to_mueller       outputs_mod.f90:473-549
 *   get_1d_mueller        outputs_mod.f90:280-329
 *   finalise              outputs_mod.f90:63-278 (numerical part)
 *
 * Build with -ffp-contract=off (no FMA contraction, like the reference's
 * baseline x86-64 build) and glibc libm.  Where the reference takes the cosine
 * and the sine of one argument, gfortran -O3 (the reference's Makefiles) turns
 * the pair into one glibc sincos call, which differs from separate sin and cos
 * calls in the last bit for a few arguments in a thousand; this file calls
 * sincos explicitly at those places so that its results do not depend on the
 * optimisation level it is built with.
 *
 * Array conventions (all match the Fortran column-major layouts):
 *   grid arrays x,y,z     (nphi, ntheta)        index i + nphi*j
 *   amplitude arrays      (nphi, ntheta, 2, 2)  complex, stored as (re, im) pairs,
 *                         element (r,c) (0-based) of bin b at pair index b + nb*(r + 2c)
 *   beam ampl             8 doubles per beam: re/im of ampl(1,1), ampl(2,1), ampl(1,2), ampl(2,2)
 *   beam vertices         ragged: vert_off[nbeam+1], verts[3*total] (x,y,z per vertex)
 *   mueller               (nphi, ntheta, 16)
 *   mueller_1d            (ntheta, 16)
 *
 * Parity pins (tests/test_oracle_pins.py): make_far_field_bins (closed form),
 * get_rotation_matrix2 (orthonormality, det +1, planarity, flip convention),
 * karczewski (closed form and forward limit), contour_integral (Green's-theorem
 * reduction to the polygon form factor checked against brute-force quadrature
 * and the triangle closed form), ampl_to_mueller (Jones-to-Mueller Kronecker
 * form), simpne (exact on quadratics), diffraction (Babinet cross-section,
 * forward-direction non-depolarisation, shift theorem).  The per-beam sign
 * flip (diff_mod.f90:561-571) is pinned by its consequence that reversing an
 * aperture's vertex order negates the diffracted amplitude, and the theta_i
 * forward fallback (diff_mod.f90:630-635) by the smoothness of the far field
 * across the edge of the fallback cone; each check also fails on an oracle
 * rebuilt with the corresponding mistake planted.  The absolute sign of the
 * flip (which of the two orientations is negated) is not observable in any
 * Mueller matrix and stays pinned by code reading only.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* REAL*4 literals of the reference, widened to real(8) */
#define PI_R4     3.1415927410125732   /* misc_submod.f90:12 */
#define K2CLAMP   0.9999989867210388   /* diff_mod.f90:151, 0.999999 */
#define KMIN      9.999999974752427e-07 /* diff_mod.f90:364-365, 1e-6 */
#define FWD_DOT   0.9990000128746033   /* diff_mod.f90:630, 0.999 */
#define ANTI_THR  9.999999747378752e-05 /* diff_mod.f90:563, 0.0001 */

/* --- small vector helpers (Fortran semantics) ------------------------------ */

/* matmul of two 3x3 (row-major m[i][j] = fortran m(i+1,j+1)) */
static void mat3_mul(const double a[3][3], const double b[3][3], double c[3][3])
{
    double t[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    memcpy(c, t, sizeof(t));
}

static void mat3_vec(const double a[3][3], const double v[3], double out[3])
{
    double t[3];
    for (int i = 0; i < 3; i++)
        t[i] = a[i][0] * v[0] + a[i][1] * v[1] + a[i][2] * v[2];
    out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}

/* misc_submod.f90:2367-2390: c = a x b, normalised unless normalise == 0 */
static void cross3(const double a[3], const double b[3], double c[3], int normalise)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    double nf = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (!normalise) nf = 1.0;
    c[0] = c[0] / nf; c[1] = c[1] / nf; c[2] = c[2] / nf;
}

static double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* --- make_far_field_bins: diff_mod.f90:827-855 ----------------------------- */
void oracle_make_far_field_bins(int64_t nphi, int64_t ntheta, const double *theta,
                                const double *phi, double *x, double *y, double *z)
{
    const double r = 1e6; /* diff_mod.f90:848 */
    for (int64_t j = 0; j < ntheta; j++)
        for (int64_t i = 0; i < nphi; i++) {
            int64_t b = i + nphi * j;
            double st, ct, sp, cp;
            sincos(theta[j], &st, &ct);
            sincos(phi[i], &sp, &cp);
            x[b] = r * st * cp;
            y[b] = r * st * sp;
            z[b] = -(r * ct);
        }
}

/* --- get_rotation_matrix2: diff_mod.f90:200-289 ---------------------------- */
/* v0: com-relative vertices, v0[3*j + r] = fortran v0(r+1, j+1) */
void oracle_get_rotation_matrix2(const double *v0, double rot_out[9])
{
    double a1[3] = {v0[0], v0[1], v0[2]};
    double b1[3] = {v0[3], v0[4], v0[5]};
    double th1 = atan(a1[0] / a1[1]), s1, c1;
    sincos(th1, &s1, &c1);
    double r1[3][3] = {{c1, -s1, 0}, {s1, c1, 0}, {0, 0, 1}};
    double a2[3], b2[3];
    mat3_vec(r1, a1, a2); mat3_vec(r1, b1, b2);
    double th2 = -atan(a2[2] / a2[1]), s2, c2;
    sincos(th2, &s2, &c2);
    double r2[3][3] = {{1, 0, 0}, {0, c2, -s2}, {0, s2, c2}};
    double a3[3], b3[3];
    mat3_vec(r2, a2, a3); mat3_vec(r2, b2, b3);
    double th3 = atan(b3[2] / b3[0]), s3, c3;
    sincos(th3, &s3, &c3);
    double r3[3][3] = {{c3, 0, s3}, {0, 1, 0}, {-s3, 0, c3}};
    double a4[3], b4[3];
    mat3_vec(r3, a3, a4); mat3_vec(r3, b3, b4);
    double t21[3][3], t321[3][3], rot[3][3];
    mat3_mul(r2, r1, t21);
    mat3_mul(r3, t21, t321);
    if (a4[0] * b4[1] - a4[1] * b4[0] > 0) {
        double r4[3][3] = {{-1, 0, 0}, {0, 1, 0}, {0, 0, -1}};
        mat3_mul(r4, t321, rot);
    } else {
        memcpy(rot, t321, sizeof(rot));
    }
    memcpy(rot_out, rot, sizeof(rot));
}

/* --- karczewski: diff_mod.f90:124-198 -------------------------------------- */
/* diff_ampl row-major: d[0]=(1,1) d[1]=(1,2) d[2]=(2,1) d[3]=(2,2) */
void oracle_karczewski(const double prop2[3], double x3, double y3, double z3, double r,
                       double d[4], double m[3], double k[3])
{
    k[0] = x3 / r;
    k[1] = y3 / r;
    k[2] = z3 / r;
    if (fabs(k[1]) > K2CLAMP) k[1] = copysign(K2CLAMP, k[1]);
    double bkx = prop2[0], bky = prop2[1], bkz = prop2[2];
    m[0] = -(k[0] * k[1] / sqrt(1 - k[1] * k[1]));
    m[1] = sqrt(1 - k[1] * k[1]);
    m[2] = -(k[1] * k[2] / sqrt(1 - k[1] * k[1]));
    double frac = sqrt((1 - k[1] * k[1]) / (1 - bky * bky));
    double a1m = -(bkz * frac);
    double b2m = -(k[2] / frac);
    double a1e = b2m;
    double b2e = a1m;
    double b1m = -(k[0] * k[1] / frac) + bkx * bky * frac;
    double a2e = -b1m;
    d[0] = 0.5 * (a1m + a1e); /* a1em */
    d[1] = 0.5 * b1m;         /* b1em */
    d[2] = 0.5 * a2e;         /* a2em */
    d[3] = 0.5 * (b2m + b2e); /* b2em */
}

/* --- per-beam frame: diff_mod.f90:515-574 ---------------------------------- */
typedef struct {
    int64_t nv;
    double com[3];
    double rot[3][3];   /* get_rotation_matrix2 */
    double rot2[3][3];  /* rotation about z, diff_mod.f90:547-555 */
    double rr[3][3];    /* matmul(rot2, rot) */
    double prop0[3], prop2[3], perp2[3], incidence2[3];
    double ampl[8];     /* sign-flipped when anti-parallel */
    int anti_parallel;
    double cos_fov;
    double *v20;        /* 3*nv, com-relative vertices */
} beam_frame;

void oracle_beam_frame(const double *ampl, const double *vk7, const double *prop, double fov,
                       int64_t nv, const double *verts, double out_rot[9], double out_rot2[9],
                       double out_prop2[3], double out_perp2[3], double out_inc2[3],
                       double out_com[3], int *out_anti);

static void beam_frame_setup(beam_frame *bf, const double *ampl, const double *vk7,
                             const double *prop, double fov, int64_t nv, const double *verts)
{
    bf->nv = nv;
    bf->cos_fov = cos(fov);
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int64_t i = 0; i < nv; i++) s = s + verts[3 * i + r];
        bf->com[r] = s / (double)nv;
    }
    bf->v20 = (double *)malloc(sizeof(double) * 3 * (size_t)nv);
    for (int64_t i = 0; i < nv; i++)
        for (int r = 0; r < 3; r++) bf->v20[3 * i + r] = verts[3 * i + r] - bf->com[r];
    double rotf[9];
    oracle_get_rotation_matrix2(bf->v20, rotf);
    memcpy(bf->rot, rotf, sizeof(rotf));
    double prop0[3] = {prop[0], prop[1], prop[2]};
    double perp0[3] = {vk7[0], vk7[1], vk7[2]};
    double prop1[3], perp1[3];
    mat3_vec(bf->rot, prop0, prop1);
    mat3_vec(bf->rot, perp0, perp1);
    double angle = -atan2(prop1[1], prop1[0]), sa, ca;
    sincos(angle, &sa, &ca);
    double r2[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    memcpy(bf->rot2, r2, sizeof(r2));
    mat3_vec(bf->rot2, prop1, bf->prop2);
    mat3_vec(bf->rot2, perp1, bf->perp2);
    double e_par2[3];
    cross3(bf->perp2, bf->prop2, e_par2, 1);
    bf->anti_parallel = (e_par2[2] > ANTI_THR);
    for (int q = 0; q < 8; q++) bf->ampl[q] = bf->anti_parallel ? -ampl[q] : ampl[q];
    double inc[3] = {0.0, 0.0, -1.0}, t[3];
    mat3_vec(bf->rot, inc, t);
    mat3_vec(bf->rot2, t, bf->incidence2);
    mat3_mul(bf->rot2, bf->rot, bf->rr);
    memcpy(bf->prop0, prop0, sizeof(prop0));
}

/* Exposed for pinning tests: the per-beam frame quantities. */
void oracle_beam_frame(const double *ampl, const double *vk7, const double *prop, double fov,
                       int64_t nv, const double *verts, double out_rot[9], double out_rot2[9],
                       double out_prop2[3], double out_perp2[3], double out_inc2[3],
                       double out_com[3], int *out_anti)
{
    beam_frame bf;
    beam_frame_setup(&bf, ampl, vk7, prop, fov, nv, verts);
    memcpy(out_rot, bf.rot, sizeof(double) * 9);
    memcpy(out_rot2, bf.rot2, sizeof(double) * 9);
    memcpy(out_prop2, bf.prop2, sizeof(double) * 3);
    memcpy(out_perp2, bf.perp2, sizeof(double) * 3);
    memcpy(out_inc2, bf.incidence2, sizeof(double) * 3);
    memcpy(out_com, bf.com, sizeof(double) * 3);
    *out_anti = bf.anti_parallel;
    free(bf.v20);
}

/* --- contour_integral for one bin: diff_mod.f90:291-444 -------------------- */
typedef struct {
    int64_t nv;
    double *x, *y, *m, *n; /* aperture-frame vertex coords, slopes (after clamps) */
    double waveno, kinc[2], lambda;
} contour_ctx;

static void contour_setup(contour_ctx *c, const beam_frame *bf, double lambda)
{
    int64_t nv = bf->nv;
    c->nv = nv;
    c->lambda = lambda;
    c->waveno = 2.0 * PI_R4 / lambda;
    c->x = (double *)malloc(sizeof(double) * 4 * (size_t)nv);
    c->y = c->x + nv; c->m = c->y + nv; c->n = c->m + nv;
    for (int64_t j = 0; j < nv; j++) {
        double v1[3];
        mat3_vec(bf->rr, bf->v20 + 3 * j, v1);
        c->x[j] = v1[0];
        c->y[j] = v1[1];
    }
    for (int64_t j = 0; j < nv; j++) {
        int64_t jp = (j == nv - 1) ? 0 : j + 1;
        c->m[j] = (c->y[jp] - c->y[j]) / (c->x[jp] - c->x[j]);
    }
    for (int64_t j = 0; j < nv; j++) c->n[j] = 1 / c->m[j];
    c->kinc[0] = bf->prop2[0] * c->waveno;
    c->kinc[1] = bf->prop2[1] * c->waveno;
}

static void contour_bin(const contour_ctx *c, double x3, double y3, double z3,
                        double *are, double *aim)
{
    double w = c->waveno;
    double bvk = w * sqrt(x3 * x3 + y3 * y3 + z3 * z3);
    double kxx = c->kinc[0] - w * w * x3 / bvk;
    double kyy = c->kinc[1] - w * w * y3 / bvk;
    if (fabs(kxx) < KMIN) kxx = KMIN;
    if (fabs(kyy) < KMIN) kyy = KMIN;
    double cb, sb;
    sincos(bvk, &sb, &cb);
    double ar = 0.0, ai = 0.0;
    int64_t nv = c->nv;
    for (int64_t j = 0; j < nv; j++) {
        double mj = c->m[j], nj = c->n[j], xj = c->x[j], yj = c->y[j];
        if (fabs(mj) > DBL_MAX) mj = 1e6;
        if (fabs(nj) > DBL_MAX) nj = 1e6;
        if (fabs(mj) < 1e-9) nj = 1e6;
        if (fabs(nj) < 1e-9) mj = 1e6;
        int64_t jp = (j == nv - 1) ? 0 : j + 1;
        double dx = c->x[jp] - xj;
        double dy = c->y[jp] - yj;
        double delta = kxx * xj + kyy * yj;
        double delta1 = kyy * mj + kxx;
        double delta2 = kxx * nj + kyy;
        double omega1 = dx * delta1;
        double omega2 = dy * delta2;
        double alpha = 1 / (2 * kyy * delta1);
        double beta = 1 / (2 * kxx * delta2);
        double s0, c0, s1, c1, s2, c2;
        sincos(delta, &s0, &c0);
        sincos(delta + omega1, &s1, &c1);
        sincos(delta + omega2, &s2, &c2);
        double sumim = alpha * (c0 - c1) - beta * (c0 - c2);
        double sumre = -(alpha * (s0 - s1)) + beta * (s0 - s2);
        double pr = cb * sumre - sb * sumim;
        double pi = cb * sumim + sb * sumre;
        ar = ar + pr / c->lambda;
        ai = ai + pi / c->lambda;
    }
    *are = ar;
    *aim = ai;
}

/* Exposed for pinning: area_facs2 of contour_integral for given rotated bins.
 * rot (get_rotation_matrix2) and rot2 (about z) as row-major 3x3, v0 com-relative
 * vertices (3*nv), prop2 aperture-frame propagation.  out: 2*nbins doubles. */
void oracle_contour_integral(double lambda, const double *rot, const double *rot2, const double *v0,
                             int64_t nv, const double *prop2, int64_t nbins, const double *x3,
                             const double *y3, const double *z3, double *out)
{
    beam_frame bf;
    memset(&bf, 0, sizeof(bf));
    bf.nv = nv;
    memcpy(bf.rot, rot, sizeof(double) * 9);
    memcpy(bf.rot2, rot2, sizeof(double) * 9);
    mat3_mul(bf.rot2, bf.rot, bf.rr);
    memcpy(bf.prop2, prop2, sizeof(double) * 3);
    bf.v20 = (double *)v0;
    contour_ctx c;
    contour_setup(&c, &bf, lambda);
    for (int64_t b = 0; b < nbins; b++) contour_bin(&c, x3[b], y3[b], z3[b], out + 2 * b, out + 2 * b + 1);
    free(c.x);
}

/* --- diffraction for one beam, all bins: diff_mod.f90:446-672 -------------- */
/* cplx helpers on (re,im) */
typedef struct { double r, i; } cpx;

static void diffraction_bins(const beam_frame *bf, const contour_ctx *cc, int64_t nphi,
                             int64_t ntheta, const double *x, const double *y, const double *z,
                             const double *phi_vals, const double *cphi, const double *sphi,
                             int64_t theta_i, int is_fast_diff, int64_t j0, int64_t j1,
                             double *amplC)
{
    const int64_t nbins = nphi * ntheta;
    for (int64_t j = j0; j < j1; j++) {
        for (int64_t i = 0; i < nphi; i++) {
            int64_t b = i + nphi * j;
            double x1 = x[b] - bf->com[0], y1 = y[b] - bf->com[1], z1 = z[b] - bf->com[2];
            double r1 = sqrt(x1 * x1 + y1 * y1 + z1 * z1);
            int in_fov = 1;
            if (is_fast_diff) {
                double vec[3] = {x1 / r1, y1 / r1, z1 / r1};
                if (dot3(vec, bf->prop0) < bf->cos_fov) in_fov = 0;
            }
            if (!in_fov) {
                for (int q = 0; q < 4; q++) { amplC[2 * (b + nbins * q)] = 0.0; amplC[2 * (b + nbins * q) + 1] = 0.0; }
                continue;
            }
            const double (*R)[3] = bf->rr;
            double x3 = R[0][0] * x1 + R[0][1] * y1 + R[0][2] * z1;
            double y3 = R[1][0] * x1 + R[1][1] * y1 + R[1][2] * z1;
            double z3 = R[2][0] * x1 + R[2][1] * y1 + R[2][2] * z1;
            double d[4], m[3], k[3];
            oracle_karczewski(bf->prop2, x3, y3, z3, r1, d, m, k);
            double hc[3];
            if (fabs(dot3(bf->incidence2, k)) < FWD_DOT) {
                cross3(bf->incidence2, k, hc, 1);
            } else {
                int64_t bt = i + nphi * theta_i;
                double xt = x[bt] - bf->com[0], yt = y[bt] - bf->com[1], zt = z[bt] - bf->com[2];
                double rt = sqrt(xt * xt + yt * yt + zt * zt);
                double my_k[3];
                my_k[0] = (R[0][0] * xt + R[0][1] * yt + R[0][2] * zt) / rt;
                my_k[1] = (R[1][0] * xt + R[1][1] * yt + R[1][2] * zt) / rt;
                my_k[2] = (R[2][0] * xt + R[2][1] * yt + R[2][2] * zt) / rt;
                cross3(bf->incidence2, my_k, hc, 1);
            }
            double evo2[3];
            cross3(k, m, evo2, 1);
            double r4[4];
            r4[0] = dot3(hc, m);
            r4[3] = dot3(hc, m);
            r4[1] = -dot3(hc, evo2);
            r4[2] = dot3(hc, evo2);
            (void)phi_vals;
            double cp = cphi[i], sp = sphi[i];
            /* temp_rot1 after the transpose bodge (diff_mod.f90:646-650) */
            double T[4] = {-sp, -cp, cp, -sp}; /* row-major (1,1),(1,2),(2,1),(2,2) */
            /* ampl (column-major storage) -> row-major complex A[r][c] */
            cpx A[2][2] = {{{bf->ampl[0], bf->ampl[1]}, {bf->ampl[4], bf->ampl[5]}},
                           {{bf->ampl[2], bf->ampl[3]}, {bf->ampl[6], bf->ampl[7]}}};
            cpx AT[2][2], DAT[2][2], R4[2][2];
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) {
                    AT[r][c].r = A[r][0].r * T[0 * 2 + c] + A[r][1].r * T[1 * 2 + c];
                    AT[r][c].i = A[r][0].i * T[0 * 2 + c] + A[r][1].i * T[1 * 2 + c];
                }
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) {
                    DAT[r][c].r = d[r * 2 + 0] * AT[0][c].r + d[r * 2 + 1] * AT[1][c].r;
                    DAT[r][c].i = d[r * 2 + 0] * AT[0][c].i + d[r * 2 + 1] * AT[1][c].i;
                }
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) {
                    R4[r][c].r = r4[r * 2 + 0] * DAT[0][c].r + r4[r * 2 + 1] * DAT[1][c].r;
                    R4[r][c].i = r4[r * 2 + 0] * DAT[0][c].i + r4[r * 2 + 1] * DAT[1][c].i;
                }
            double ar, ai;
            contour_bin(cc, x3, y3, z3, &ar, &ai);
            for (int r = 0; r < 2; r++)
                for (int c = 0; c < 2; c++) {
                    int64_t o = 2 * (b + nbins * (r + 2 * c));
                    amplC[o] = R4[r][c].r * ar - R4[r][c].i * ai;
                    amplC[o + 1] = R4[r][c].r * ai + R4[r][c].i * ar;
                }
        }
    }
}

static int64_t find_theta_i(int64_t ntheta, const double *theta)
{
    /* diff_mod.f90:583-591: first theta > 1 degree (REAL*4 pi).  The reference
     * leaves theta_i undefined when none exists; here it falls back to the
     * last column. */
    for (int64_t j = 0; j < ntheta; j++)
        if (theta[j] > 1 * PI_R4 / 180) return j;
    return ntheta - 1;
}

static void phi_tables(int64_t nphi, const double *phi, double **cphi, double **sphi)
{
    *cphi = (double *)malloc(sizeof(double) * 2 * (size_t)nphi);
    *sphi = *cphi + nphi;
    for (int64_t i = 0; i < nphi; i++) sincos(phi[i], &(*sphi)[i], &(*cphi)[i]);
}

/* One call of `diffraction` for a single beam. amplC fully overwritten. */
void oracle_diffraction(int64_t nphi, int64_t ntheta, const double *x, const double *y,
                        const double *z, double lambda, const double *phi_vals,
                        const double *theta_vals, int is_fast_diff, const double *ampl,
                        const double *vk7, const double *prop, double fov, int64_t nv,
                        const double *verts, double *amplC)
{
    beam_frame bf;
    beam_frame_setup(&bf, ampl, vk7, prop, fov, nv, verts);
    contour_ctx cc;
    contour_setup(&cc, &bf, lambda);
    double *cphi, *sphi;
    phi_tables(nphi, phi_vals, &cphi, &sphi);
    int64_t ti = find_theta_i(ntheta, theta_vals);
    diffraction_bins(&bf, &cc, nphi, ntheta, x, y, z, phi_vals, cphi, sphi, ti, is_fast_diff, 0,
                     ntheta, amplC);
    free(cphi);
    free(cc.x);
    free(bf.v20);
}

/* --- diff_main: diff_mod.f90:674-825 ---------------------------------------
 * Sums beams in tree order (the -mt 0 order).  `nthreads` > 1 splits the
 * theta columns of each beam across OpenMP threads; each bin's sum over
 * beams keeps the same order, so results do not depend on nthreads.
 * Ext-diff beams use fov = pi/2 (diff_mod.f90:797). */
static void sum_tree(int64_t nphi, int64_t ntheta, const double *x, const double *y,
                     const double *z, double lambda, const double *theta, const double *phi,
                     int is_fast_diff, int64_t nb, const double *ampl, const double *vk7,
                     const double *prop, const double *fov, int fov_override, double fov_value,
                     const int64_t *vert_off, const double *verts, double *out, int nthreads)
{
    const int64_t nbins = nphi * ntheta;
    memset(out, 0, sizeof(double) * 8 * (size_t)nbins);
    if (nb == 0) return;
    double *amplC = (double *)malloc(sizeof(double) * 8 * (size_t)nbins);
    double *cphi, *sphi;
    phi_tables(nphi, phi, &cphi, &sphi);
    int64_t ti = find_theta_i(ntheta, theta);
    for (int64_t bi = 0; bi < nb; bi++) {
        beam_frame bf;
        double f = fov_override ? fov_value : fov[bi];
        int64_t nv = vert_off[bi + 1] - vert_off[bi];
        beam_frame_setup(&bf, ampl + 8 * bi, vk7 + 3 * bi, prop + 3 * bi, f, nv,
                         verts + 3 * vert_off[bi]);
        contour_ctx cc;
        contour_setup(&cc, &bf, lambda);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int64_t j = 0; j < ntheta; j++) {
            diffraction_bins(&bf, &cc, nphi, ntheta, x, y, z, phi, cphi, sphi, ti, is_fast_diff,
                             j, j + 1, amplC);
            for (int q = 0; q < 4; q++)
                for (int64_t i = 0; i < nphi; i++) {
                    int64_t o = 2 * ((i + nphi * j) + nbins * q);
                    out[o] = out[o] + amplC[o];
                    out[o + 1] = out[o + 1] + amplC[o + 1];
                }
        }
        free(cc.x);
        free(bf.v20);
    }
    (void)nthreads;
    free(cphi);
    free(amplC);
}

void oracle_diff_main(int64_t nphi, int64_t ntheta, const double *theta, const double *phi,
                      double lambda, int is_fast_diff,
                      int64_t nb, const double *ampl, const double *vk7, const double *prop,
                      const double *fov, const int64_t *vert_off, const double *verts,
                      int64_t ne, const double *e_ampl, const double *e_vk7, const double *e_prop,
                      const int64_t *e_vert_off, const double *e_verts,
                      double *ampl_far_beam, double *ampl_far_ext_diff, int nthreads)
{
    const int64_t nbins = nphi * ntheta;
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)nbins);
    double *y = x + nbins, *z = y + nbins;
    oracle_make_far_field_bins(nphi, ntheta, theta, phi, x, y, z);
    sum_tree(nphi, ntheta, x, y, z, lambda, theta, phi, is_fast_diff, nb, ampl, vk7, prop, fov,
             0, 0.0, vert_off, verts, ampl_far_beam, nthreads);
    sum_tree(nphi, ntheta, x, y, z, lambda, theta, phi, is_fast_diff, ne, e_ampl, e_vk7, e_prop,
             NULL, 1, PI_R4 / 2.0, e_vert_off, e_verts, ampl_far_ext_diff, nthreads);
    free(x);
}

/* --- ampl_to_mueller: outputs_mod.f90:473-549 ------------------------------ */
static inline cpx cmul_conj(cpx a, cpx b) /* a * conjg(b) */
{
    cpx r = {a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i};
    return r;
}

void oracle_ampl_to_mueller(int64_t nbins, const double *ampl, double *mueller)
{
    for (int64_t b = 0; b < nbins; b++) {
        cpx s11 = {ampl[2 * (b)], ampl[2 * (b) + 1]};
        cpx s21 = {ampl[2 * (b + nbins)], ampl[2 * (b + nbins) + 1]};
        cpx s12 = {ampl[2 * (b + 2 * nbins)], ampl[2 * (b + 2 * nbins) + 1]};
        cpx s22 = {ampl[2 * (b + 3 * nbins)], ampl[2 * (b + 3 * nbins) + 1]};
        cpx a11 = cmul_conj(s11, s11), a12 = cmul_conj(s12, s12);
        cpx a21 = cmul_conj(s21, s21), a22 = cmul_conj(s22, s22);
        cpx p1112 = cmul_conj(s11, s12), p2221 = cmul_conj(s22, s21);
        cpx p1121 = cmul_conj(s11, s21), p2212 = cmul_conj(s22, s12);
        cpx p1122 = cmul_conj(s11, s22), p1221 = cmul_conj(s12, s21);
        cpx p2112 = cmul_conj(s21, s12);
        cpx p2111 = cmul_conj(s21, s11), p2211 = cmul_conj(s22, s11);
        double *M = mueller;
        M[b + nbins * 0] = 0.5 * (((a11.r + a12.r) + a21.r) + a22.r);
        M[b + nbins * 1] = 0.5 * (((a11.r - a12.r) + a21.r) - a22.r);
        M[b + nbins * 2] = p1112.r + p2221.r;
        M[b + nbins * 3] = p1112.i - p2221.i;
        M[b + nbins * 4] = 0.5 * (((a11.r + a12.r) - a21.r) - a22.r);
        M[b + nbins * 5] = 0.5 * (((a11.r - a12.r) - a21.r) + a22.r);
        M[b + nbins * 6] = p1112.r - p2221.r;
        M[b + nbins * 7] = p1112.i + p2221.i;
        M[b + nbins * 8] = p1121.r + p2212.r;
        M[b + nbins * 9] = p1121.r - p2212.r;
        M[b + nbins * 10] = p1122.r + p1221.r;
        M[b + nbins * 11] = p1122.i + p2112.i;
        M[b + nbins * 12] = p2111.i + cmul_conj(s22, s12).i;
        M[b + nbins * 13] = p2111.i - cmul_conj(s22, s12).i;
        M[b + nbins * 14] = p2211.i - p1221.i;
        M[b + nbins * 15] = p2211.r - p1221.r;
    }
}

/* --- simpne: misc_submod.f90:2392-2531 ------------------------------------- */
double oracle_simpne(int64_t ntab, const double *x, const double *y, int64_t ystride)
{
    double result = 0.0;
    if (ntab <= 2) return NAN; /* the reference stops with 'SIMPNE - Fatal error!' */
#define Y(k) y[(k) * ystride]
    int64_t n = 1;
    for (;;) {
        double x1 = x[n - 1], x2 = x[n], x3 = x[n + 1];
        double e = x3 * x3 - x1 * x1;
        double f = x3 * x3 * x3 - x1 * x1 * x1;
        double feints = x3 - x1;
        double del[3] = {x3 - x2, x1 - x3, x2 - x1};
        double g[3] = {x2 + x3, x1 + x3, x1 + x2};
        double pi1[3] = {x2 * x3, x1 * x3, x1 * x2};
        double sum1 = 0.0;
        for (int i = 0; i < 3; i++)
            sum1 = sum1 + Y(n - 1 + i) * del[i] * (f / 3.0 - g[i] * 0.5 * e + pi1[i] * feints);
        result = result - sum1 / (del[0] * del[1] * del[2]);
        n = n + 2;
        if (ntab <= n + 1) break;
    }
    if (ntab % 2 != 0) return result;
    n = ntab - 2;
    double x3 = x[ntab - 1], x2 = x[ntab - 2], x1 = x[ntab - 3];
    double e = x3 * x3 - x2 * x2;
    double f = x3 * x3 * x3 - x2 * x2 * x2;
    double feints = x3 - x2;
    double del[3] = {x3 - x2, x1 - x3, x2 - x1};
    double g[3] = {x2 + x3, x1 + x3, x1 + x2};
    double pi1[3] = {x2 * x3, x1 * x3, x1 * x2};
    double sum1 = 0.0;
    for (int i = 0; i < 3; i++)
        sum1 = sum1 + Y(n - 1 + i) * del[i] * (f / 3.0 - g[i] * 0.5 * e + pi1[i] * feints);
    result = result - sum1 / (del[0] * del[1] * del[2]);
#undef Y
    return result;
}

/* --- get_1d_mueller: outputs_mod.f90:280-329 ------------------------------- */
void oracle_get_1d_mueller(int64_t nphi, int64_t ntheta, const double *mueller,
                           const double *phi, double *mueller_1d)
{
    const int64_t nbins = nphi * ntheta;
    for (int64_t i = 0; i < ntheta; i++)
        for (int e = 0; e < 16; e++)
            mueller_1d[i + ntheta * e] = oracle_simpne(nphi, phi, mueller + nphi * i + nbins * e, 1);
}

/* --- finalise: outputs_mod.f90:63-278 (numerical part) ---------------------
 * Inputs are the diff_main outputs (without the k factor); they are scaled in
 * place by k like the reference (outputs_mod.f90:104-105).  Outputs:
 *   mueller/mueller_beam/mueller_ext (nbins*16), *_1d (ntheta*16)
 *   params[0..8] = scatt, scatt_beam, scatt_ext_diff, asymmetry, asymmetry_beam,
 *                  asymmetry_ext_diff, back_scatt, (scaling factors beam, ext or 0)
 * scaling: when nonzero, beam_energy_out / ext_energy_out rescale the amplitudes
 * (outputs_mod.f90:146-184). */
void oracle_finalise(int64_t nphi, int64_t ntheta, const double *theta, const double *phi,
                     double la, double *ampl_beam, double *ampl_ext, int scaling,
                     double beam_energy_out, double ext_energy_out, double *mueller,
                     double *mueller_beam, double *mueller_ext, double *m1d, double *m1d_beam,
                     double *m1d_ext, double *params)
{
    const int64_t nbins = nphi * ntheta;
    double waveno = 2.0 * PI_R4 / la;
    double *tot = (double *)malloc(sizeof(double) * 8 * (size_t)nbins);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)ntheta);
    for (int64_t q = 0; q < 8 * nbins; q++) {
        ampl_beam[q] = ampl_beam[q] * waveno;
        ampl_ext[q] = ampl_ext[q] * waveno;
    }
    double scatt = 0, scatt_b = 0, scatt_e = 0;
    for (int pass = 0; pass < (scaling ? 2 : 1); pass++) {
        if (pass == 1) {
            double fb = sqrt(scatt_b / beam_energy_out), fe = sqrt(scatt_e / ext_energy_out);
            for (int64_t q = 0; q < 8 * nbins; q++) {
                ampl_beam[q] = ampl_beam[q] / fb;
                ampl_ext[q] = ampl_ext[q] / fe;
            }
            params[7] = fb;
            params[8] = fe;
        }
        for (int64_t q = 0; q < 8 * nbins; q++) tot[q] = ampl_beam[q] + ampl_ext[q];
        oracle_ampl_to_mueller(nbins, tot, mueller);
        oracle_ampl_to_mueller(nbins, ampl_beam, mueller_beam);
        oracle_ampl_to_mueller(nbins, ampl_ext, mueller_ext);
        oracle_get_1d_mueller(nphi, ntheta, mueller, phi, m1d);
        oracle_get_1d_mueller(nphi, ntheta, mueller_beam, phi, m1d_beam);
        oracle_get_1d_mueller(nphi, ntheta, mueller_ext, phi, m1d_ext);
        double w2 = waveno * waveno;
        for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d[j] * sin(theta[j]) / w2;
        scatt = oracle_simpne(ntheta, theta, tmp, 1);
        for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d_beam[j] * sin(theta[j]) / w2;
        scatt_b = oracle_simpne(ntheta, theta, tmp, 1);
        for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d_ext[j] * sin(theta[j]) / w2;
        scatt_e = oracle_simpne(ntheta, theta, tmp, 1);
    }
    if (!scaling) { params[7] = 0; params[8] = 0; }
    double w2 = waveno * waveno;
    for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d[j] * sin(theta[j]) * cos(theta[j]) / scatt / w2;
    double g = oracle_simpne(ntheta, theta, tmp, 1);
    for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d_beam[j] * sin(theta[j]) * cos(theta[j]) / scatt_b / w2;
    double gb = oracle_simpne(ntheta, theta, tmp, 1);
    for (int64_t j = 0; j < ntheta; j++) tmp[j] = m1d_ext[j] * sin(theta[j]) * cos(theta[j]) / scatt_e / w2;
    double ge = oracle_simpne(ntheta, theta, tmp, 1);
    /* minloc(abs(theta_vals - 2*pi)): first index of the minimum */
    int64_t ib = 0;
    double best = fabs(theta[0] - 2 * PI_R4);
    for (int64_t j = 1; j < ntheta; j++) {
        double d = fabs(theta[j] - 2 * PI_R4);
        if (d < best) { best = d; ib = j; }
    }
    double back = mueller[0 + nphi * ib] * 4 * PI_R4 / w2;
    params[0] = scatt; params[1] = scatt_b; params[2] = scatt_e;
    params[3] = g; params[4] = gb; params[5] = ge; params[6] = back;
    free(tot);
    free(tmp);
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
