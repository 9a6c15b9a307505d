/*
 * vis_oracle.c -- CPU FP64 restatement of PBT's facet-visibility step of the
 * near-field tracer (the facets a propagating beam illuminates).
 *
 * TEST INFRASTRUCTURE ONLY.  Like diff_oracle.c, this file is the parity
 * oracle: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker (or the timed CPU baseline), never as
 * the product.  The product path is the HIP library under abt_amd/.
 *
 * Restated reference routines (operation order kept, REAL*4 literals widened
 * to double, built with -ffp-contract=off):
 *   find_vis_int                  beam_loop_mod.f90:2037-2247
 *   find_vis_ext                  beam_loop_mod.f90:2249-2459
 *   find_vis_inc                  beam_loop_mod.f90:2461-2609
 *   beam_aligned_bounding_boxes   beam_loop_mod.f90:2668-2743
 *   compute_geometry_midpoints    misc_submod.f90:199-209 (for the boxes)
 *   rotate                        beam_loop_mod.f90:1467-1526 (the beam frame)
 *
 * Inputs are the fields of geometry_type (types_mod.f90:156-190) that these
 * routines read, in the beam frame (propagation along -z):
 *   v        (nv, 3) column-major: x[nv], y[nv], z[nv]            geometry%v
 *   f_off    [nf+1] offsets into f_vi (facet i has f_off[i+1]-f_off[i] vertices)
 *   f_vi     1-based vertex ids, facet after facet                 geometry%f(:)%vi
 *   f_mid    [nf][3] facet midpoints                               geometry%f(:)%mid
 *   f_nz     [nf] z component of the facet normal                  geometry%n(f(i)%ni, 3)
 *   f_ap     [nf] 1-based aperture id                              geometry%f(:)%ap
 *   ap_nz    [na] z component of the aperture normal               geometry%ap(:)%n(3)
 *   field_in [nf_in] 1-based facet ids of the illuminating beam    beam%field_in(:)%fi
 *   beam_mid_z  beam_geometry%f(1)%mid(3) (find_vis_inc only)
 * Outputs per facet: in_beam, is_vis (a local of find_vis_int/_ext, an output
 * of find_vis_inc), dist_beam, id_beam (1-based position in field_in),
 * is_shad.  counts (optional) receives the reference's work: [0] candidate
 * pairs that reach the depth test, [1] blocker edges evaluated.
 *
 * Parity rests on checks that do not restate this file: closed forms on convex
 * bodies, brute-force point-in-polygon occlusion on layered scenes, the box
 * construction against its closed form, and planted mistakes
 * (tests/test_vis_oracle.py).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define VIS_NB 16                                  /* 4 x 4 boxes */
static const double VIS_FAC = 1.100000023841858;   /* real(8), parameter :: fac = 1.1 (REAL*4 literal) */
static const double VIS_0P01 = 0.009999999776482582;   /* 0.01 (REAL*4 literal) */

/* beam_aligned_bounding_boxes (beam_loop_mod.f90:2668-2743) followed by
 * compute_geometry_midpoints on its 16 quadrilaterals (misc_submod.f90:199-209). */
void oracle_vis_boxes(int64_t nv, const double* v, double* bbx, double* bby)
{
    const double* x = v;
    const double* y = v + nv;
    double max_x = x[0], min_x = x[0], max_y = y[0], min_y = y[0];
    for (int64_t i = 1; i < nv; i++) {
        if (x[i] > max_x) max_x = x[i];
        if (x[i] < min_x) min_x = x[i];
        if (y[i] > max_y) max_y = y[i];
        if (y[i] < min_y) min_y = y[i];
    }
    double vx[25], vy[25];
    int c = 0;
    for (int i = 1; i <= 5; i++)
        for (int j = 1; j <= 5; j++) {
            vx[c] = VIS_FAC * (min_x + (max_x - min_x) / 4.0 * (double)(i - 1));
            vy[c] = VIS_FAC * (min_y + (max_y - min_y) / 4.0 * (double)(j - 1));
            c++;
        }
    int f = 0;
    for (int i = 1; i <= 4; i++)
        for (int j = 1; j <= 4; j++) {
            const int q[4] = {(i - 1) * 5 + j, (i - 1) * 5 + j + 1, i * 5 + j + 1, i * 5 + j};
            double sx = 0.0, sy = 0.0;
            for (int k = 0; k < 4; k++) {
                sx = sx + vx[q[k] - 1];
                sy = sy + vy[q[k] - 1];
            }
            bbx[f] = sx / 4.0;
            bby[f] = sy / 4.0;
            f++;
        }
}

/* minloc over the 16 box midpoints of sqrt((bbx - px)**2 + (bby - py)**2) (first minimum) */
static int vis_nearest(const double* bbx, const double* bby, double px, double py)
{
    int best = 0;
    double dbest = 0.0;
    for (int b = 0; b < VIS_NB; b++) {
        const double dx = bbx[b] - px, dy = bby[b] - py;
        const double d = sqrt(dx * dx + dy * dy);
        if (b == 0 || d < dbest) {
            dbest = d;
            best = b;
        }
    }
    return best;
}

/* The centroid (mx, my) of facet m against the projected edges of facet j
 * (beam_loop_mod.f90:2149-2162 int, :2361-2374 ext, :2558-2571 inc). */
static int vis_within(const double* x, const double* y, const int64_t* vi, int64_t nvj, double mx,
                      double my, int flip, int64_t* edges)
{
    int within = 1;
    for (int64_t k = 0; k < nvj; k++) {
        const int64_t a = vi[k] - 1;
        const int64_t b = (k == nvj - 1) ? vi[0] - 1 : vi[k + 1] - 1;
        const double en2 = -x[b] + x[a];
        const double en1 = y[b] - y[a];
        const double vb1 = mx - x[a];
        const double vb2 = my - y[a];
        const double chk = flip ? -vb1 * en1 - vb2 * en2 : vb1 * en1 + vb2 * en2;
        if (chk > 0) within = 0;
    }
    if (edges) *edges += nvj;
    return within;
}

void oracle_find_vis(int mode, int64_t nv, const double* v, int64_t nf, const int64_t* f_off,
                     const int64_t* f_vi, const double* f_mid, const double* f_nz, const int64_t* f_ap,
                     int64_t na, const double* ap_nz, int64_t nf_in, const int64_t* field_in,
                     double beam_mid_z, uint8_t* in_beam, uint8_t* is_vis, double* dist_beam,
                     int64_t* id_beam, uint8_t* is_shad, int64_t* counts)
{
    (void)na;
    const double* x = v;
    const double* y = v + nv;
    const int inc = mode == 2, ext = mode == 1;
    double bbx[VIS_NB], bby[VIS_NB];
    oracle_vis_boxes(nv, v, bbx, bby);

    /* f3: box of each facet midpoint; f5: boxes nearest to any of its first 3 vertices */
    int* f3 = (int*)malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1));
    uint8_t* f5 = (uint8_t*)calloc((size_t)(nf > 0 ? nf : 1) * VIS_NB, 1);
    uint8_t* is_beam = (uint8_t*)calloc((size_t)(nf > 0 ? nf : 1), 1);
    int64_t* mapping = (int64_t*)calloc((size_t)(nf > 0 ? nf : 1), sizeof(int64_t));
    for (int64_t i = 0; i < nf; i++) {
        f3[i] = vis_nearest(bbx, bby, f_mid[3 * i], f_mid[3 * i + 1]);
        for (int j = 0; j < 3; j++) {
            const int64_t vj = f_vi[f_off[i] + j] - 1;
            f5[i * VIS_NB + vis_nearest(bbx, bby, x[vj], y[vj])] = 1;
        }
    }
    for (int64_t i = 0; i < nf; i++) {
        is_vis[i] = 1;
        is_shad[i] = 0;
        in_beam[i] = inc ? 1 : 0;
        id_beam[i] = 0;
        dist_beam[i] = 0.0;
    }
    for (int64_t i = 0; i < nf_in; i++) {
        is_beam[field_in[i] - 1] = 1;
        mapping[field_in[i] - 1] = i + 1;
    }
    if (!inc && nf_in == 0) {
        /* no illuminating facet: nothing is lit (the reference reads field_in(1) out of
         * range here; the C-ABI documents this outcome) */
        for (int64_t i = 0; i < nf; i++) is_vis[i] = 0;
    }
    if (!inc && nf_in > 0) {
        /* beam extent in x-y (beam_loop_mod.f90:2124-2143) and the midpoint cull (:2145-2155) */
        const int64_t f0 = field_in[0] - 1;
        const int64_t v0 = f_vi[f_off[f0]] - 1;
        double bxmax = x[v0], bxmin = x[v0], bymax = y[v0], bymin = y[v0];
        for (int64_t i = 0; i < nf_in; i++) {
            const int64_t fi = field_in[i] - 1;
            for (int64_t k = f_off[fi]; k < f_off[fi + 1]; k++) {
                const int64_t vk = f_vi[k] - 1;
                if (x[vk] > bxmax) bxmax = x[vk];
                if (x[vk] < bxmin) bxmin = x[vk];
                if (y[vk] > bymax) bymax = y[vk];
                if (y[vk] < bymin) bymin = y[vk];
            }
        }
        for (int64_t i = 0; i < nf; i++) {
            if (f_mid[3 * i] < bxmin) is_vis[i] = 0;
            else if (f_mid[3 * i] > bxmax) is_vis[i] = 0;
            else if (f_mid[3 * i + 1] < bymin) is_vis[i] = 0;
            else if (f_mid[3 * i + 1] > bymax) is_vis[i] = 0;
        }
    }
    for (int64_t i = 0; i < nf; i++) {
        if (mode == 0) { if (f_nz[i] >= 0) is_shad[i] = 1; }
        else { if (f_nz[i] <= 0) is_shad[i] = 1; }
    }
    int64_t npairs = 0, nedges = 0;
    for (int64_t m = 0; m < nf; m++) {
        if (!inc && !is_vis[m]) continue;
        const double apn = ap_nz[f_ap[m] - 1];
        const int down = mode == 0 ? apn > -VIS_0P01 : apn < VIS_0P01;
        if (down) {
            is_vis[m] = 0;
            if (inc) in_beam[m] = 0;
            continue;
        }
        const int bb = f3[m];
        const double mx = f_mid[3 * m], my = f_mid[3 * m + 1], mz = f_mid[3 * m + 2];
        for (int64_t j = 0; j < nf; j++) {
            if (!f5[j * VIS_NB + bb]) continue;
            if (j == m) continue;
            if (inc && f_nz[j] < VIS_0P01) continue;
            npairs++;
            const double jz = f_mid[3 * j + 2];
            if (mz > jz) continue;
            if (!vis_within(x, y, f_vi + f_off[j], f_off[j + 1] - f_off[j], mx, my, ext, &nedges)) continue;
            if (inc) {
                if (f_ap[m] == f_ap[j]) is_shad[m] = 1;
                else {
                    is_vis[m] = 0;
                    in_beam[m] = 0;
                }
                continue;
            }
            if (is_beam[j]) {
                if (in_beam[m]) {
                    const int64_t pf = field_in[id_beam[m] - 1] - 1;
                    if (jz < f_mid[3 * pf + 2]) {
                        in_beam[m] = 1;
                        id_beam[m] = mapping[j];
                        dist_beam[m] = jz - mz;
                    }
                } else {
                    in_beam[m] = 1;
                    id_beam[m] = mapping[j];
                    dist_beam[m] = jz - mz;
                }
            } else {
                if (f_ap[m] == f_ap[j]) continue;
                if (in_beam[m]) {
                    const int64_t pf = field_in[id_beam[m] - 1] - 1;
                    if (jz > f_mid[3 * pf + 2]) continue;
                }
                is_vis[m] = 0;
                in_beam[m] = 0;
                break;
            }
        }
    }
    if (inc) {
        for (int64_t m = 0; m < nf; m++)
            if (in_beam[m]) {
                dist_beam[m] = beam_mid_z - f_mid[3 * m + 2];
                id_beam[m] = 1;
            }
    }
    if (counts) {
        counts[0] = npairs;
        counts[1] = nedges;
    }
    free(f3);
    free(f5);
    free(is_beam);
    free(mapping);
}

/* rotate (beam_loop_mod.f90:1467-1526): the rotation taking the beam's propagation
 * direction prop to -z, rot = rot2 . rot1, as R[3 i + k] = rot(i+1, k+1).  pi is the
 * reference's REAL*4 literal (misc_submod.f90:12); cosine and sine of one angle are taken
 * together (sincos), as gfortran does. */
void oracle_vis_rotation(const double* prop, double* R)
{
    const double pi = 3.1415927410125732;
    double r1[9] = {0}, r2[9] = {0}, c, s;
    const double th1 = 2 * pi - atan2(prop[1], prop[0]);
    sincos(th1, &s, &c);
    r1[0] = c;
    r1[1] = -s;
    r1[3] = s;
    r1[4] = c;
    r1[8] = 1;
    double t[3];
    for (int i = 0; i < 3; i++) {          /* matmul(rot1, prop) */
        double a = 0.0;
        for (int k = 0; k < 3; k++) a = a + r1[3 * i + k] * prop[k];
        t[i] = a;
    }
    const double nrm = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);   /* normalise_vec, misc_submod.f90:2216-2227 */
    for (int i = 0; i < 3; i++) t[i] = t[i] / nrm;
    const double th2 = acos(-t[2]);
    sincos(th2, &s, &c);
    r2[0] = c;
    r2[2] = s;
    r2[6] = -s;
    r2[8] = c;
    r2[4] = 1;
    for (int i = 0; i < 3; i++)            /* matmul(rot2, rot1) */
        for (int j = 0; j < 3; j++) {
            double a = 0.0;
            for (int k = 0; k < 3; k++) a = a + r2[3 * i + k] * r1[3 * k + j];
            R[3 * i + j] = a;
        }
}

/* matmul(rot, p) for one 3-vector, the reference's summation order */
static void vis_rot3(const double* R, double px, double py, double pz, double* o)
{
    for (int i = 0; i < 3; i++) {
        double a = 0.0;
        a = a + R[3 * i] * px;
        a = a + R[3 * i + 1] * py;
        a = a + R[3 * i + 2] * pz;
        o[i] = a;
    }
}

/* The geometry fields the visibility routines read, rotated by R as rotate() does
 * (vertices, facet normals, midpoints and aperture normals each rotated, nothing
 * recomputed): v0 (nv,3) column-major, mid0 [nf][3], n0 [nf][3] (each facet's normal),
 * apn0 [na][3] -> v (nv,3) column-major, mid [nf][3], nz [nf], ap_nz [na]. */
void oracle_vis_rotate(const double* R, int64_t nv, const double* v0, int64_t nf, const double* mid0,
                       const double* n0, int64_t na, const double* apn0, double* v, double* mid, double* nz,
                       double* ap_nz)
{
    double o[3];
    for (int64_t i = 0; i < nv; i++) {
        vis_rot3(R, v0[i], v0[nv + i], v0[2 * nv + i], o);
        v[i] = o[0];
        v[nv + i] = o[1];
        v[2 * nv + i] = o[2];
    }
    for (int64_t i = 0; i < nf; i++) {
        vis_rot3(R, mid0[3 * i], mid0[3 * i + 1], mid0[3 * i + 2], mid + 3 * i);
        vis_rot3(R, n0[3 * i], n0[3 * i + 1], n0[3 * i + 2], o);
        nz[i] = o[2];
    }
    for (int64_t i = 0; i < na; i++) {
        vis_rot3(R, apn0[3 * i], apn0[3 * i + 1], apn0[3 * i + 2], o);
        ap_nz[i] = o[2];
    }
}

/* Many beams over one facet topology, OpenMP over beams (the CPU baseline).  Per beam b:
 * v + 3*nv*b, f_mid + 3*nf*b, f_nz + nf*b, ap_nz + na*b, field_in[in_off[b] .. in_off[b+1]),
 * beam_mid_z[b], outputs + nf*b, counts + 2*b. */
void oracle_find_vis_batch(int mode, int64_t nbeam, int64_t nv, const double* v, int64_t nf,
                           const int64_t* f_off, const int64_t* f_vi, const double* f_mid,
                           const double* f_nz, const int64_t* f_ap, int64_t na, const double* ap_nz,
                           const int64_t* in_off, const int64_t* field_in, const double* beam_mid_z,
                           uint8_t* in_beam, uint8_t* is_vis, double* dist_beam, int64_t* id_beam,
                           uint8_t* is_shad, int64_t* counts, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < nbeam; b++) {
        oracle_find_vis(mode, nv, v + 3 * nv * b, nf, f_off, f_vi, f_mid + 3 * nf * b, f_nz + nf * b, f_ap,
                        na, ap_nz + na * b, in_off[b + 1] - in_off[b], field_in + in_off[b],
                        beam_mid_z ? beam_mid_z[b] : 0.0, in_beam + nf * b, is_vis + nf * b,
                        dist_beam + nf * b, id_beam + nf * b, is_shad + nf * b,
                        counts ? counts + 2 * b : NULL);
    }
}
