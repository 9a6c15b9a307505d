/*
 * abt_vis.h -- C-ABI of the facet-visibility step of the near-field tracer
 * (libabt_diff.so): which facets a propagating beam illuminates.
 *
 * Drop-in boundary for the three visibility subroutines of beam_loop_mod
 * (src/beam_loop_mod.f90 of hballington12/abt):
 *   find_vis_int  beam_loop_mod.f90:2037-2247  (internal beam, called at :676)
 *   find_vis_ext  beam_loop_mod.f90:2249-2459  (external beam, called at :876)
 *   find_vis_inc  beam_loop_mod.f90:2461-2609  (incident beam, called at :487)
 * with their helper beam_aligned_bounding_boxes (:2668-2743).  Results are
 * bit-identical to the reference's for the same inputs, including the entries
 * the reference leaves behind for blocked facets (dist_beam and id_beam keep
 * the last beam facet found before the blocker).
 *
 * Conventions as in abt_diff.h (0 = success, abt_last_error() for the message;
 * host entry points are synchronous; "_dev" entry points take device pointers
 * and a hipStream_t as void*).  Indices are 1-based where the reference's are
 * (vertex ids, aperture ids, facet ids in field_in, id_beam).
 */
#ifndef ABT_VIS_H
#define ABT_VIS_H

#include <stddef.h>
#include <stdint.h>

#include "abt_diff.h"

#ifdef __cplusplus
extern "C" {
#endif

enum abt_vis_mode {
    ABT_VIS_INT = 0,   /* find_vis_int: beam inside the particle */
    ABT_VIS_EXT = 1,   /* find_vis_ext: beam outside the particle */
    ABT_VIS_INC = 2    /* find_vis_inc: the incident beam */
};

/* The fields of geometry_type (types_mod.f90:156-190) the visibility routines read,
 * in the beam frame (propagation along -z; the caller's rotate() for internal and
 * external beams, the particle frame for the incident beam). */
typedef struct abt_vis_geometry {
    int64_t nv;            /* vertices */
    const double* v;       /* (nv, 3) column-major, geometry%v: x[nv], then y[nv], then z[nv] */
    int64_t nf;            /* facets */
    const int64_t* f_off;  /* [nf+1] facet i lists f_vi[f_off[i] .. f_off[i+1]-1] (>= 3 vertices) */
    const int64_t* f_vi;   /* 1-based vertex ids, geometry%f(i)%vi(:) facet after facet */
    const double* f_mid;   /* [nf][3] geometry%f(i)%mid */
    const double* f_nz;    /* [nf] geometry%n(geometry%f(i)%ni, 3) */
    const int64_t* f_ap;   /* [nf] 1-based aperture id, geometry%f(i)%ap */
    int64_t na;            /* apertures */
    const double* ap_nz;   /* [na] geometry%ap(:)%n(3) */
} abt_vis_geometry;

/* One call of find_vis_int / find_vis_ext / find_vis_inc (mode), host arrays.
 * field_in: the 1-based facet ids beam%field_in(1:nf_in)%fi of the illuminating beam
 * (at least one for ABT_VIS_INT/EXT; not read for ABT_VIS_INC).  beam_mid_z:
 * beam_geometry%f(1)%mid(3) (ABT_VIS_INC only).  Outputs, nf entries each:
 * in_beam, dist_beam, id_beam, is_shad (all modes) and is_vis (an output of
 * find_vis_inc, the local of the same name for the other two; may be NULL). */
int abt_find_vis(int mode, const abt_vis_geometry* geo, int64_t nf_in, const int64_t* field_in,
                 double beam_mid_z, uint8_t* in_beam, uint8_t* is_vis, double* dist_beam,
                 int64_t* id_beam, uint8_t* is_shad);

/* Many beams in one launch sequence (no reference counterpart: the tracer's loops at
 * beam_loop_mod.f90:676 and :876 call find_vis_* once per beam).  The beams share
 * one facet topology (f_off, f_vi, f_ap); the coordinates are per beam.  All arrays
 * are device arrays:
 *   v      [nbeam][3][nv]  (each beam's geometry%v, column-major as above)
 *   f_mid  [nbeam][nf][3],  f_nz [nbeam][nf],  ap_nz [nbeam][na]
 *   in_off [nbeam+1]: beam b's field_in is field_in[in_off[b] .. in_off[b+1]-1]
 *          (for ABT_VIS_INT/EXT a beam with none illuminates nothing)
 *   beam_mid_z [nbeam] (ABT_VIS_INC; else may be NULL)
 *   outputs [nbeam][nf]; is_vis may be NULL.
 * counts: NULL, or 2 device uint64 the launch adds its work to: [0] candidate
 * (facet, blocker) pairs that reach the depth test, [1] blocker edges the
 * reference evaluates for them (nv of the blocker per pair passing the test).
 * phases: 0 (or 3) runs the whole evaluation; 1 runs only the set-up (boxes, flags,
 * candidate lists, blocker records), 2 only the blocker scan, which then reads the
 * set-up that the last phase-1 call left in the same workspace for the same inputs.
 * The workspace (abt_find_vis_workspace bytes, 256-byte aligned) is the caller's;
 * launches that share one must not overlap. */
typedef struct abt_vis_batch {
    int mode;
    int64_t nbeam, nv, nf, na;
    const int64_t* f_off;
    const int64_t* f_vi;
    const int64_t* f_ap;
    const double* v;
    const double* f_mid;
    const double* f_nz;
    const double* ap_nz;
    const int64_t* in_off;
    const int64_t* field_in;
    const double* beam_mid_z;
    uint8_t* in_beam;
    uint8_t* is_vis;
    double* dist_beam;
    int64_t* id_beam;
    uint8_t* is_shad;
    unsigned long long* counts;
    int phases;
} abt_vis_batch;

/* The beam frames of the tracer's rotate() (beam_loop_mod.f90:1467-1526; no counterpart
 * as a separate routine).  abt_vis_rotations: for nbeam propagation directions prop
 * [nbeam][3] (host), the rotations R [nbeam][9] (host, R[9 b + 3 i + k] = rot(i+1, k+1))
 * taking each to -z, computed with the C library as the reference does.
 * abt_vis_rotate_dev: the particle geometry rotated into each beam's frame (device arrays):
 * v0 (nv,3) column-major, mid0 [nf][3], n0 [nf][3] (each facet's normal), apn0 [na][3]
 * -> v [nbeam][3][nv], mid [nbeam][nf][3], nz [nbeam][nf], ap_nz [nbeam][na], the inputs
 * of abt_find_vis_dev; bit-identical to rotate(). */
int abt_vis_rotations(int64_t nbeam, const double* prop, double* R);
int abt_vis_rotate_dev(int64_t nbeam, const double* R, int64_t nv, const double* v0, int64_t nf,
                       const double* mid0, const double* n0, int64_t na, const double* apn0, double* v,
                       double* mid, double* nz, double* ap_nz, void* stream);

/* Device workspace bytes for nbeam beams of nf facets with nvref = f_off[nf] vertex references. */
size_t abt_find_vis_workspace(int64_t nbeam, int64_t nf, int64_t nvref);
/* nvref must equal f_off[nf] (the host copy of the last offset). */
int abt_find_vis_dev(const abt_vis_batch* batch, int64_t nvref, void* workspace, size_t bytes,
                     void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ABT_VIS_H */
