// CPU emulator of the far-field kernels: runs the kernel bodies of
// abt_amd/csrc/ffbody.h on the host, bin by bin, in the kernel's beam order.
// Test infrastructure: lets the CPU suite check the mixed-precision algorithm
// against the oracle without a GPU.  (Host FP32 arithmetic and 1/sqrtf stand
// in for the device's; the algorithmic error budget is the same.)
#include <stdint.h>
#include <string.h>
#include "../../abt_amd/csrc/ffbody.h"
#include "../../abt_amd/csrc/adjbody.h"

extern "C" void emu_far_field(int prec, int nphi, int ntheta, int theta_i, double lambda,
                              const double* x, const double* y, const double* z,
                              const double* cphi, const double* sphi, const double* recs,
                              const double* vrecs, int64_t nb, int fast_diff, double* out)
{
    using namespace abt;
    DevGrid g;
    memset(&g, 0, sizeof(g));
    g.nphi = nphi; g.ntheta = ntheta; g.nbins = nphi * ntheta; g.theta_i = theta_i;
    devgrid_constants(g, lambda);
    static double etab[ETAB_GRID], blk4[ETAB_LDS];   // the grid's table; the FP64r body's LDS block
    fill_exp_table(etab);
    memcpy(blk4, etab, sizeof(double) * ETAB_DOUBLES);
    fp64r_constants(g, blk4 + ETAB_DOUBLES);
    g.etab = etab;
    const int nbins = g.nbins;
#pragma omp parallel for schedule(dynamic, 64)
    for (int bin = 0; bin < nbins; bin++) {
        const int iphi = bin % nphi;
        const int bt = iphi + nphi * theta_i;
        const double rad = sqrt(x[bin] * x[bin] + y[bin] * y[bin] + z[bin] * z[bin]);
        BinIn bi{x[bin], y[bin], z[bin], x[bt], y[bt], z[bt], rad};
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float accf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t b = 0; b < nb; b++) {
            if (prec == 0)
                beam_bin_fp64(recs + b * REC, vrecs, g, bi, etab, fast_diff, acc);
            else if (prec == 4)     // the restructured FP64 body
                beam_bin_fp64r(recs + b * REC, vrecs, g, bi, blk4, fast_diff, acc);
            else if (prec == 3) {   // the FP64 model (adjbody.h)
                AdjBeam B;
                adj_beam_from_record(recs + b * REC, B);
                model_bin_fp64(B, bi.x, bi.y, bi.z, bi.xt, bi.yt, rad, g.waveno, g.lambda, acc);
            }
            else {   // FP32 sums over groups of MIXED_GROUP beams, as the kernels form them
                beam_bin_mixed(recs + b * REC, vrecs, g, bi, fast_diff, accf);
                if ((b + 1) % MIXED_GROUP == 0 || b + 1 == nb)
                    for (int q = 0; q < 8; q++) {
                        acc[q] += (double)accf[q];
                        accf[q] = 0.f;
                    }
            }
        }
        if (prec == 0) apply_temp_rot1(acc, cphi[iphi], sphi[iphi]);
        if (prec == 1) mixed_finish(acc, cphi[iphi], sphi[iphi]);
        if (prec == 4) fp64r_finish(acc, cphi[iphi], sphi[iphi], g.krev * rad);
        if (prec == 3) model_finish(acc, cphi[iphi], sphi[iphi], g.krev * rad);
        for (int q = 0; q < 4; q++) {
            out[2 * (bin + (int64_t)nbins * q)] = acc[2 * q];
            out[2 * (bin + (int64_t)nbins * q) + 1] = acc[2 * q + 1];
        }
    }
}

// Adjoint bodies (abt_amd/csrc/adjbody.h) on the CPU: per beam, the sum over bins.
// bins: 7 doubles per bin (x, y, z, xt, yt, |x|, k|x|/2pi); lam: 8 doubles per bin
// (Lam' = Lam T_phi^T); gv: 9 per beam, ga: 8 per beam.
extern "C" void emu_adjoint(const double* recs, int64_t nb, const double* bins, const double* lam,
                            int64_t nbins, double lambda, double* gv, double* ga)
{
    using namespace abt;
    const double waveno = 2.0 * PI_R4 / lambda;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; b++) {
        const double* r = recs + b * REC;
        AdjBeam B;
        adj_beam_from_record(r, B);
        AdjAcc acc;
        adj_zero(acc);
        for (int64_t i = 0; i < nbins; i++) {
            const double* bb = bins + 7 * i;
            AdjBin bn{bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], bb[6]};
            adj_bin_fp64(B, bn, lam + 8 * i, waveno, lambda, acc);
        }
        adj_finish(B, r + R_V20, r[R_SIGMA], acc, gv + 9 * b, ga + 8 * b);
    }
}

// Mixed-precision adjoint bodies on the CPU (host FP32 and 1/sqrtf for the device's).
extern "C" void emu_adjoint_mixed(const double* recs, int64_t nb, const double* bins, const double* lam,
                                  int64_t nbins, double lambda, double* gv, double* ga)
{
    using namespace abt;
    const double waveno = 2.0 * PI_R4 / lambda;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; b++) {
        const double* r = recs + b * REC;
        AdjBeam B;
        adj_beam_from_record(r, B);
        AdjAcc acc;
        adj_zero(acc);
        AdjAccF af;
        adj_zero_f(af);
        for (int64_t i = 0; i < nbins; i++) {
            const double* bb = bins + 7 * i;
            AdjBin bn{bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], bb[6]};
            float lf[8];
            for (int q = 0; q < 8; q++) lf[q] = (float)lam[8 * i + q];
            adj_bin_mixed(B, bn, lf, waveno, lambda, af);
            if ((i + 1) % ADJ_GROUP == 0 || i + 1 == nbins) {   // FP32 groups of bins, as the kernel forms them
                adj_add(acc, af);
                adj_zero_f(af);
            }
        }
        adj_finish(B, r + R_V20, r[R_SIGMA], acc, gv + 9 * b, ga + 8 * b);
    }
}

// The plain FP64 body's helpers on the CPU: the table sine and cosine (sincos_tab64 over the
// grid's table) and the quotient by a correctly rounded reciprocal (div_by).
extern "C" void emu_sincos_tab(int64_t n, const double* x, double* s, double* c)
{
    using namespace abt;
    static double tab[ETAB_GRID];
    static bool filled = false;
    if (!filled) {
        fill_exp_table(tab);
        filled = true;
    }
    for (int64_t i = 0; i < n; i++) sincos_tab64(x[i], tab, s + i, c + i);
}

extern "C" void emu_div_by(int64_t n, const double* a, const double* b, double* q)
{
    using namespace abt;
    for (int64_t i = 0; i < n; i++) q[i] = div_by(a[i], b[i], 1.0 / b[i]);
}
