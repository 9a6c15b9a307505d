/*
 * abt_diff.h -- C-ABI of the MI355X far-field diffraction library (libabt_diff.so).
 *
 * Drop-in boundary for PBT's far-field diffraction module `diff_mod`
 * (src/diff_mod.f90 of hballington12/abt).  The Fortran module
 * abt_amd/fortran/diff_mod.f90 keeps the reference's procedure interfaces and
 * reaches this library through iso_c_binding; the same entry points serve the
 * Python/PyTorch harness.
 *
 * Conventions
 *  - Return value: 0 on success, nonzero on error; abt_last_error() then holds a
 *    thread-local message.
 *  - Host entry points take host pointers and are synchronous.  "_dev" entry
 *    points take device pointers and a HIP stream (hipStream_t passed as void*;
 *    NULL = the default stream) and are asynchronous on that stream.
 *  - Amplitude arrays use the Fortran layout complex(8) a(nphi, ntheta, 2, 2):
 *    interleaved (re, im) doubles, bin b = i_phi + nphi * j_theta, element (r, c)
 *    (0-based) at complex index b + nbins * (r + 2 c).
 *  - Mueller arrays: real(8) m(nphi, ntheta, 16) and m1d(ntheta, 16).
 *  - Angles in radians.  Lengths in micrometres (the caller's unit).
 */
#ifndef ABT_DIFF_H
#define ABT_DIFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ABT_DIFF_ABI_VERSION 1

enum abt_status {
    ABT_OK = 0,
    ABT_ERR_ARG = 1,      /* invalid argument or shape */
    ABT_ERR_HIP = 2,      /* HIP runtime error */
    ABT_ERR_NODEV = 3,    /* no usable gfx950 device */
    ABT_ERR_NOMEM = 4     /* device allocation failed */
};

enum abt_precision {
    ABT_FP64 = 0,         /* plain FP64 kernel, the reference's formulas */
    ABT_MIXED = 1,        /* FP32 body, FP64 phases, FP64 series near the forward limit */
    ABT_FP64_ACCURATE = 2,/* batched plans only: ABT_FP64 with the far-field phase k|x - com|
                             formed without cancellation (k(|x - com| - |x|) plus the per-bin
                             k|x|); for finite-difference gradient checks */
    ABT_FP64_FAST = 3,    /* FP64 throughout in the mixed kernel's restructured algorithm (three
                             vertex exponentials per beam and bin instead of the split sum's twelve
                             trigonometric pairs); agrees with ABT_FP64 to ~1e-12 of the largest
                             amplitude */
    ABT_MIXED_GATED = 4   /* batched plans only: ABT_MIXED, plus a per-bin bound of the FP32 body's
                             error over the orientation sum; every bin whose bound exceeds the
                             accuracy gate (|dS11| <= 1e-6 S11, |dSij| <= 1e-5 max(|Sij|, 1e-3 S11))
                             is recomputed over all orientations with ABT_FP64_FAST.  The
                             per-orientation integrals and amplitudes keep the ABT_MIXED values
                             outside the recomputed bins */
};

/* One tree of outgoing beams (the flattened outbeamtype records of
 * types_mod.f90:8-17 plus the facet vertices `diffraction` looks up through
 * `fi`, diff_mod.f90:529-531). */
typedef struct abt_beams {
    int64_t n;                /* number of beams */
    const double* ampl;       /* [n*8]  re/im of ampl(1,1), ampl(2,1), ampl(1,2), ampl(2,2) */
    const double* vk7;        /* [n*3]  e-perp vector */
    const double* prop;       /* [n*3]  outgoing propagation vector (prop_out) */
    const double* fov;        /* [n]    field of view, rad (NULL: every beam uses fov_value) */
    const int64_t* vert_off;  /* [n+1]  vertex offsets; vertices of beam b: vert_off[b]..vert_off[b+1]-1 */
    const double* verts;      /* [3*vert_off[n]]  x,y,z per vertex */
} abt_beams;

/* Angular grid: theta_vals / phi_vals of job_parameters_type (types_mod.f90:73-74). */
typedef struct abt_grid {
    int64_t n_theta;
    int64_t n_phi;
    const double* theta;      /* [n_theta] rad */
    const double* phi;        /* [n_phi]   rad */
} abt_grid;

typedef struct abt_dev_grid abt_dev_grid;   /* opaque: far-field bins resident on the device */

/* ---- library ------------------------------------------------------------ */
int abt_version(void);
const char* abt_last_error(void);
/* Select the device for the calling process (lazy: host entry points call it
 * with device 0 when it was never called). */
int abt_init(int device);
/* Free every device buffer of the library (grids created by the caller stay
 * valid until abt_grid_destroy). */
int abt_finalize(void);

/* ---- reference interfaces ------------------------------------------------ */

/* make_far_field_bins (diff_mod.f90:827-855): x,y,z of shape (n_phi, n_theta). */
int abt_make_far_field_bins(int64_t n_theta, const double* theta, int64_t n_phi,
                            const double* phi, double* x, double* y, double* z);

/* diff_main (diff_mod.f90:674-825).  Sums every beam of `beams` into
 * ampl_far_beam and every beam of `ext` (fov forced to pi/2, diff_mod.f90:797)
 * into ampl_far_ext_diff, both overwritten.  Beams are summed in tree order
 * for every bin (deterministic).  is_fast_diff: FOV culling (diff_mod.f90:593-603).
 * One of the two outputs may be NULL: that tree is then neither computed nor copied
 * back, and its beam list is not read (callers that time the two trees apart). */
int abt_diff_main(const abt_grid* grid, double lambda, int is_fast_diff,
                  const abt_beams* beams, const abt_beams* ext, int precision,
                  double* ampl_far_beam, double* ampl_far_ext_diff);

/* diffraction (diff_mod.f90:446-672) for one beam on caller-given bins
 * xfar/yfar/zfar (n_phi, n_theta); amplC overwritten.  Re-entrant: callable
 * concurrently from several host threads (the reference's -memory mode calls it
 * inside OpenMP regions, beam_loop_mod.f90:37-51, :1935-1967). */
int abt_diffraction(int64_t n_phi, int64_t n_theta, const double* xfar, const double* yfar,
                    const double* zfar, double lambda, const double* phi_vals,
                    const double* theta_vals, int is_fast_diff, const double* ampl,
                    const double* vk7, const double* prop, double fov, int64_t nv,
                    const double* verts, int precision, double* amplC);

/* finalise (outputs_mod.f90:63-278), numerical part, on the device.  Inputs are
 * diff_main's outputs without the k factor; they are scaled by k in place like
 * the reference.  Any output pointer may be NULL.  params[9] = scatt, scatt_beam,
 * scatt_ext_diff, asymmetry, asymmetry_beam, asymmetry_ext_diff, back_scatt,
 * beam scale, ext scale (the last two 0 unless scaling). */
int abt_finalise(const abt_grid* grid, double lambda, double* ampl_far_beam,
                 double* ampl_far_ext_diff, int scaling, double beam_energy_out,
                 double ext_energy_out, double* mueller, double* mueller_beam,
                 double* mueller_ext, double* mueller_1d, double* mueller_1d_beam,
                 double* mueller_1d_ext, double* params);

/* ---- device-resident interface (no reference counterpart: the resident form of
 * diff_main for callers whose beams already live in HBM) ------------------------ */

/* Doubles per prepared beam record / per vertex record (see abt_amd/csrc/records.h). */
int64_t abt_record_doubles(void);
int64_t abt_vertex_record_doubles(void);

/* Per-beam frame set-up of `diffraction` (diff_mod.f90:515-574) on the host, in
 * the reference's operation order with glibc libm (bit-reproducible against a
 * CPU FP64 run, including at normal incidence where the frame is set by
 * rounding).  fov_value replaces beams->fov when that is NULL.  recs: n *
 * abt_record_doubles(); vrecs: vert_off[n] * abt_vertex_record_doubles() (may be
 * NULL when every aperture is a triangle). */
int abt_prepare_beams_host(const abt_beams* beams, double fov_value, double lambda,
                           double* recs, double* vrecs);

/* The same set-up on the device from device-resident beams (FP64, device libm). */
int abt_prepare_beams_dev(const abt_beams* dev_beams, double fov_value, double lambda,
                          double* dev_recs, double* dev_vrecs, void* stream);

/* Upload a grid: bins built on the host by make_far_field_bins. */
int abt_grid_create(const abt_grid* grid, double lambda, abt_dev_grid** out);
int abt_grid_destroy(abt_dev_grid* g);
int64_t abt_grid_bins(const abt_dev_grid* g);

/* Workspace bytes abt_far_field_dev needs for nb beams (0 when none): 64 bytes per bin
 * per beam chunk, at most 32 chunks of >= 512 beams (2048 bytes per bin). */
size_t abt_far_field_workspace(const abt_dev_grid* g, int64_t nb, int precision);

/* Coherent far-field sum of nb prepared beams: dev_out (complex, Fortran layout,
 * 8*nbins doubles) overwritten.  dev_work: abt_far_field_workspace bytes (NULL:
 * the library allocates). */
int abt_far_field_dev(const abt_dev_grid* g, const double* dev_recs, const double* dev_vrecs,
                      int64_t nb, int precision, int is_fast_diff, double* dev_out,
                      void* dev_work, size_t work_bytes, void* stream);

/* abt_finalise_dev without any allocation.  No reference counterpart: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The epilogue's scratch (the total amplitude, 8 doubles per bin, and the
 * theta integrands, 3 per theta value) is a 64 MiB array resident on each GPU the library
 * runs on, for the life of the process, instead of a workspace argument or a library
 * allocation: grids up to abt_finalise_static_max_bins(n_theta) bins (about 1.05e6).
 * Larger grids are rejected with ABT_ERR_ARG.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams or threads) race and must be serialised by
 * the caller.  Arguments and outputs as abt_finalise_dev. */
int abt_finalise_dev_static(const abt_dev_grid* g, double* dev_ampl_beam, double* dev_ampl_ext,
                            int scaling, double beam_energy_out, double ext_energy_out,
                            double* dev_mueller, double* dev_mueller_1d, double* dev_params,
                            void* stream);
int64_t abt_finalise_static_max_bins(int64_t n_theta);

/* Conditioning diagnostic (no reference counterpart): dev_intensity[nbins] =
 * sum over beams of the squared Frobenius norm of each beam's FP64 amplitude
 * (the incoherent sum).  sqrt of it over |coherent sum| bounds how much the
 * coherent sum amplifies per-beam rounding. */
int abt_far_field_incoherent_dev(const abt_dev_grid* g, const double* dev_recs,
                                 const double* dev_vrecs, int64_t nb, int is_fast_diff,
                                 double* dev_intensity, void* stream);

/* Device form of abt_finalise: amplitudes resident on the device (scaled by k in
 * place).  dev_mueller: 3*16*nbins (total, beam, ext, each (nphi, ntheta, 16));
 * dev_mueller_1d: 3*16*n_theta; dev_params: 9 doubles (as abt_finalise).
 * dev_work: abt_finalise_workspace bytes (NULL: the library allocates). */
size_t abt_finalise_workspace(const abt_dev_grid* g);
int abt_finalise_dev(const abt_dev_grid* g, double* dev_ampl_beam, double* dev_ampl_ext,
                     int scaling, double beam_energy_out, double ext_energy_out,
                     double* dev_mueller, double* dev_mueller_1d, double* dev_params,
                     void* dev_work, size_t work_bytes, void* stream);

/* ---- orientation-batched far field ------------------------------------------
 * Replaces the orientation loop of the reference drivers (pbt_seq_mod.f90:146-269;
 * pbt_mpi_mod.f90 spreads it over MPI ranks): for every orientation diff_main of
 * the beam tree and of the ext tree, finalise (x k, total = beam + ext, Mueller
 * matrix, integrals; outputs_mod.f90:63-278) and summation (outputs_mod.f90:407-471),
 * as one persistent kernel over all orientations.
 *
 * Records: the prepared records (abt_prepare_beams_*) of every orientation's
 * beam tree back to back, and likewise for the ext trees (prepared with
 * fov_value pi/2); orientation o owns beam records beam_off[o]..beam_off[o+1]-1
 * and ext records ext_off[o]..ext_off[o+1]-1 (host arrays of norient+1 offsets).
 *
 * Outputs (device):
 *   dev_mueller[16*nbins]     sum over orientations of the total Mueller matrix
 *                             (mueller_total before divide_by_num_orientations);
 *   dev_mueller_1d[16*ntheta] its phi integral (may be NULL) = the sum of the
 *                             orientations' mueller_1d up to rounding;
 *   dev_integrals[6*norient]  per orientation: scatt, I2, scatt_beam, I2_beam,
 *                             scatt_ext_diff, I2_ext_diff, where asymmetry = I2 / scatt
 *                             (outputs_mod.f90:133-135, 198-200);
 *   dev_ampl[8*nbins*norient] optional (NULL: not written): per orientation the total
 *                             amplitude k (ampl_far_beam + ampl_far_ext_diff)
 *                             (outputs_mod.f90:104-108), Fortran layout per orientation.
 * A plan owns its workspace: runs of one plan must not overlap.  The grid must
 * outlive the plan.  abt_batch_run only launches kernels (no allocation, no
 * synchronisation, no memset), so callers may capture it in a HIP graph;
 * abt_batch_capture / abt_batch_launch do that inside the library. */
typedef struct abt_batch_plan abt_batch_plan;
int abt_batch_plan_create(const abt_dev_grid* g, int32_t norient, const int64_t* beam_off,
                          const int64_t* ext_off, int precision, int is_fast_diff,
                          abt_batch_plan** out);
int abt_batch_plan_destroy(abt_batch_plan* plan);
/* Device bytes the plan holds (partial sums: 128 bytes per bin per orientation chunk, up to
 * 128 chunks; weights, tables). */
int abt_batch_workspace(const abt_batch_plan* plan, size_t* bytes);
/* Tile shape: beams staged per LDS tile (8, 16, 32) and bins per lane (1, 2).
 * Default 16 x 1 until set or autotuned (ABT_FP64_ACCURATE: 16 x 1 only; ABT_FP64_FAST:
 * 2 bins per lane with 8 or 16 beams per tile only). */
int abt_batch_set_tiles(abt_batch_plan* plan, int32_t beams_per_tile, int32_t bins_per_thread);
int abt_batch_get_tiles(const abt_batch_plan* plan, int32_t* beams_per_tile,
                        int32_t* bins_per_thread, int32_t* chunks, int32_t* workgroups);
/* Time every tile shape on the first sample_orient orientations (0: 32) and keep
 * the fastest; ms_out (optional) receives the 6 timings, bins_per_thread major.
 * Synchronises the stream. */
int abt_batch_autotune(abt_batch_plan* plan, const double* dev_recs_beam,
                       const double* dev_vrecs_beam, const double* dev_recs_ext,
                       const double* dev_vrecs_ext, int32_t sample_orient, void* stream,
                       double* ms_out);
int abt_batch_run(const abt_batch_plan* plan, const double* dev_recs_beam,
                  const double* dev_vrecs_beam, const double* dev_recs_ext,
                  const double* dev_vrecs_ext, double* dev_mueller, double* dev_mueller_1d,
                  double* dev_integrals, double* dev_ampl, void* stream);
/* Capture abt_batch_run with these buffers into a HIP graph held by the plan
 * (stream must not be NULL), then replay it with abt_batch_launch. */
int abt_batch_capture(abt_batch_plan* plan, const double* dev_recs_beam,
                      const double* dev_vrecs_beam, const double* dev_recs_ext,
                      const double* dev_vrecs_ext, double* dev_mueller, double* dev_mueller_1d,
                      double* dev_integrals, double* dev_ampl, void* stream);
int abt_batch_launch(const abt_batch_plan* plan, void* stream);
/* ABT_MIXED_GATED plans: the number of bins the last run recomputed in FP64 (0 for other
 * precisions).  Synchronises the stream. */
int abt_batch_gate_stats(const abt_batch_plan* plan, int32_t* refined_bins, void* stream);
/* -scaling of finalise (outputs_mod.f90:146-184) in the batched plan: in every orientation o
 * the beam amplitudes are divided by sqrt(scatt_beam / beam_energy_out[o]) and the ext
 * amplitudes by sqrt(scatt_ext_diff / ext_energy_out[o]), scatt_* being that orientation's
 * own cross sections, before the Mueller sum; the integrals are those of the divided
 * amplitudes.  beam_energy_out, ext_energy_out: device arrays of norient doubles, read by
 * every run (a captured graph reads their values at replay); NULL, NULL turns scaling off.
 * The plan then keeps every orientation's amplitudes: 128 bytes per bin per orientation,
 * allocated here (not in abt_batch_run).  Drops a captured graph.  Not for ABT_MIXED_GATED. */
int abt_batch_set_scaling(abt_batch_plan* plan, const double* beam_energy_out,
                          const double* ext_energy_out);

/* ---- adjoint: gradients for shape and roughness retrieval ---------------------
 * No reference counterpart (the reference has no derivatives); the forward it
 * differentiates is the batched plan's per-orientation total amplitude
 * X_o = k (ampl_far_beam + ampl_far_ext_diff) (dev_ampl of abt_batch_run).
 * Given dev_dldx = dL/dX_o for a real loss L (re/im pairs, the same layout as
 * dev_ampl: [norient][8][nbins]), abt_adjoint_run writes for every beam of both
 * trees dL/d(aperture vertices) (9 doubles per beam: x,y,z of each vertex in the
 * record's vertex order, global frame) and dL/d(amplitude matrix) (8 doubles per
 * beam, re/im in the ampl layout), any output may be NULL.  Beams hold their own
 * direction, e-perp vector and field of view fixed; the aperture frame, centre,
 * polarisation, phases and form factor all follow the vertices.  Triangular
 * apertures only (others get NaN).  Per-edge phase terms are recomputed, not
 * stored.  A plan owns its workspace: runs of one plan must not overlap. */
typedef struct abt_adjoint_plan abt_adjoint_plan;
int abt_adjoint_plan_create(const abt_dev_grid* g, int32_t norient, const int64_t* beam_off,
                            const int64_t* ext_off, int precision, abt_adjoint_plan** out);
int abt_adjoint_plan_destroy(abt_adjoint_plan* plan);
int abt_adjoint_run(const abt_adjoint_plan* plan, const double* dev_recs_beam,
                    const double* dev_recs_ext, const double* dev_dldx,
                    double* dev_grad_verts_beam, double* dev_grad_ampl_beam,
                    double* dev_grad_verts_ext, double* dev_grad_ampl_ext, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ABT_DIFF_H */
