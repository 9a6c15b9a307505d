// records.h -- device layout of prepared beam records and the far-field grid.
//
// One beam record per outgoing beam: the per-beam frame quantities of
// `diffraction` (diff_mod.f90:515-574) computed once in FP64, followed by the
// aperture edges of triangular apertures inline.  Apertures with more than
// three vertices keep their edges in a separate vertex-record array.
#pragma once
#include <stdint.h>

namespace abt {

// ---- beam record: REC doubles per beam --------------------------------------
enum : int {
    R_COM = 0,      // [3] aperture centre (vertex mean), global frame
    R_RR = 3,       // [9] matmul(rot2, rot), row-major
    R_PROP2 = 12,   // [3] propagation vector in the aperture frame
    R_INC2 = 15,    // [3] incidence (0,0,-1) in the aperture frame
    R_AMPL = 18,    // [8] amplitude matrix, sign-flipped when anti-parallel:
                    //     re/im of (1,1) (2,1) (1,2) (2,2)
    R_PROP0 = 26,   // [3] propagation vector, global frame
    R_COSFOV = 29,  // cos(fov)
    R_NV = 30,      // number of vertices
    R_VOFF = 31,    // first vertex record of this beam (apertures with nv > 3)
    R_KPROP = 32,   // [3] prop0 * k / (2 pi)   (revolutions per um)
    R_CC = 35,      // |com|^2
    R_AMAX = 36,    // max_j |v20_j|, aperture radius about com
    R_AREA = 37,    // signed aperture area in the aperture frame (negative: clockwise)
    R_FRAC = 38,    // 1 / sqrt(1 - prop2_y^2)
    R_KY1 = 39,     // 1 - prop2_y^2
    // inline edges for nv <= 3 (unused slots zero)
    R_EX = 40,      // [3] aperture-frame x of vertex j
    R_EY = 43,      // [3] aperture-frame y of vertex j
    R_EM = 46,      // [3] slope m_j after the clamps of diff_mod.f90:379-382
    R_EN = 49,      // [3] inverse slope n_j after the clamps
    R_EDX = 52,     // [3] x_{j+1} - x_j
    R_EDY = 55,     // [3] y_{j+1} - y_j
    R_V20 = 58,     // [9] com-relative global vertices (x,y,z per vertex)
    // mixed-precision kernel constants (krev = k / 2 pi, revolutions per um)
    R_KRR = 68,     // [6] krev * rows 1 and 2 of matmul(rot2, rot)
    R_KP2 = 74,     // [2] krev * prop2_x, krev * prop2_y
    R_G32 = 76,     // second FP32 block (float index G_* below), 6 doubles
    R_F32 = 82,     // FP32 block (float index F_* below), 12 doubles
    R_SIGMA = 94,   // -1 when the amplitude matrix was sign-flipped (anti-parallel), else +1
    // FP64 copies of the mixed kernel's per-beam constants, for the restructured FP64 body
    R_RP64 = 96,    // [4] r+ = (A(1,:) - i A(2,:)) / 2: re/im of columns 1, 2
    R_RM64 = 100,   // [4] r- = (A(1,:) + i A(2,:)) / 2
    R_C64 = 104,    // [3] the Karczewski coefficients G_C1, G_C2, G_C3 in FP64
    R_ISIG = 107,   // sqrt(1 - prop2_y^2)
    REC = 108       // padded to a multiple of 4 doubles (32 bytes)
};

// float offsets inside the FP32 block of a beam record
enum : int {
    F_RR = 0,       // [9] rot2.rot
    F_KX = 9, F_KY = 10, F_KZ = 11,  // prop2
    F_KXY = 12,     // prop2_x * prop2_y
    F_SIG = 13,     // 1 / sqrt(1 - prop2_y^2)
    F_ISIG = 14,    // sqrt(1 - prop2_y^2)
    F_QA2 = 15,     // (2 pi a)^2, a = aperture radius about com
    F_NV = 16,      // number of vertices
    F_EDX = 17,     // [3] edge x_{j+1} - x_j
    F_EDY = 20,     // [3] edge y_{j+1} - y_j
    F_C2 = 23,      // squared Frobenius norm of the amplitude matrix
    F_END = 24
};

// float offsets inside the second FP32 block
enum : int {
    G_RP = 0,       // [4] r+ = (A(1,:) - i A(2,:)) / 2: re/im of columns 1, 2
    G_RM = 4,       // [4] r- = (A(1,:) + i A(2,:)) / 2
    G_C1 = 8,       // -prop2_z / (2 sqrt(1 - prop2_y^2) krev)
    G_C2 = 9,       // -sqrt(1 - prop2_y^2) / (2 krev)
    G_C3 = 10,      // prop2_x prop2_y / (2 sqrt(1 - prop2_y^2) krev)
    G_END = 11
};

// ---- vertex record (nv > 3): VREC doubles per vertex -------------------------
enum : int {
    V_X = 0, V_Y = 1, V_M = 2, V_N = 3, V_DX = 4, V_DY = 5, V_V20 = 6, VREC = 12
};

// ---- device grid ------------------------------------------------------------
struct DevGrid {
    int32_t nphi, ntheta, nbins, theta_i;   // theta_i: 0-based forward-fallback column
    double lambda, waveno;                  // waveno = 2*pi_r4/lambda
    double krev;                            // waveno / (2 pi), revolutions per um
    double krev_k2clamp;                    // krev * K2CLAMP (mixed kernel's k_y clamp test)
    float inv_lambda_f;                     // 1/lambda (FP32 body of the mixed kernel)
    float inv_krev_f;                       // 1/krev
    double inv_krev;                        // 1/krev (restructured FP64 body)
    double tri_scale;                       // 1 / (2 pi^2 lambda) (its triangle form factor)
    double inv_lambda;                      // 1 / lambda (plain FP64 body)
    const double* x;                        // [nbins] far-field bin positions
    const double* y;
    const double* z;
    const double* cphi;                     // [nphi] cos(phi), glibc
    const double* sphi;                     // [nphi] sin(phi)
    const double* rad;                      // [nbins] |x|
    const double* kr_rev;                   // [nbins] k|x|/(2 pi)
    const double* theta;                    // [ntheta] theta_vals, rad
    const double* phi;                      // [nphi] phi_vals, rad
    const double* etab;                     // [ETAB_GRID] e^{i pi j / ETAB_N}, then its low parts (FP64 bodies)
};

// Table of the FP64 bodies' exponentials: cos and sin of pi j / ETAB_N for j = 0 .. 2 ETAB_N - 1
// (one turn, so an angle's index is its multiple of pi / ETAB_N modulo 2 ETAB_N), interleaved and
// correctly rounded, followed by the low parts (the values less the doubles, same layout) that
// make each entry a double-double; built on the host in extended precision.
constexpr int ETAB_N = 512;
constexpr int ETAB_DOUBLES = 2 * (2 * ETAB_N);
constexpr int ETAB_LO = ETAB_DOUBLES;       // offset of the low parts
constexpr int ETAB_GRID = 2 * ETAB_DOUBLES;  // doubles of DevGrid::etab, and the plain body's LDS block
// The FP64r body's LDS block: the table, then its grid constants (fp64r_constants).
constexpr int ETAB_KREV = ETAB_DOUBLES, ETAB_K2CLAMP = ETAB_DOUBLES + 1, ETAB_INV_KREV = ETAB_DOUBLES + 2,
              ETAB_TRI_SCALE = ETAB_DOUBLES + 3, ETAB_LAMBDA = ETAB_DOUBLES + 4, ETAB_LDS = ETAB_DOUBLES + 5;

// REAL*4 literals of the reference, widened to double
constexpr double PI_R4 = 3.1415927410125732;
constexpr double K2CLAMP = 0.9999989867210388;
constexpr double KMIN = 9.999999974752427e-07;
constexpr double FWD_DOT = 0.9990000128746033;
constexpr double ANTI_THR = 9.999999747378752e-05;

// The wavelength-derived constants of a grid: waveno = 2 pi_r4 / lambda (diff_mod.f90:332,
// with the REAL*4 pi) and the scaled forms the kernel bodies read.
inline void devgrid_constants(DevGrid& g, double lambda)
{
    g.lambda = lambda;
    g.waveno = 2.0 * PI_R4 / lambda;
    g.krev = g.waveno / 6.283185307179586476925;
    g.krev_k2clamp = g.krev * K2CLAMP;
    g.inv_lambda_f = (float)(1.0 / lambda);
    g.inv_krev_f = (float)(1.0 / g.krev);
    g.inv_krev = 1.0 / g.krev;
    g.tri_scale = 1.0 / (2.0 * 9.869604401089358 * lambda);
    g.inv_lambda = 1.0 / lambda;
}

}  // namespace abt
