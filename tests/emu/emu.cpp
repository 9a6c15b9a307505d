// CPU emulator of the far-field kernels: runs the kernel bodies of
// abt_amd/csrc/ffbody.h on the host, bin by bin, in the kernel's beam order.
// Test infrastructure: lets the CPU suite check the mixed-precision algorithm
// against the oracle without a GPU.  (Host FP32 arithmetic and 1/sqrtf stand
// in for the device's; the algorithmic error budget is the same.)
#include <stdint.h>
#include <string.h>
#include "../../abt_amd/csrc/ffbody.h"

extern "C" void emu_far_field(int prec, int nphi, int ntheta, int theta_i, double lambda,
                              const double* x, const double* y, const double* z,
                              const double* cphi, const double* sphi, const double* recs,
                              const double* vrecs, int64_t nb, int fast_diff, double* out)
{
    using namespace abt;
    DevGrid g;
    memset(&g, 0, sizeof(g));
    g.nphi = nphi; g.ntheta = ntheta; g.nbins = nphi * ntheta; g.theta_i = theta_i;
    g.lambda = lambda;
    g.waveno = 2.0 * PI_R4 / lambda;
    g.krev = g.waveno / 6.283185307179586476925;
    g.inv_lambda_f = (float)(1.0 / lambda);
    const int nbins = g.nbins;
#pragma omp parallel for schedule(dynamic, 64)
    for (int bin = 0; bin < nbins; bin++) {
        const int iphi = bin % nphi;
        const int bt = iphi + nphi * theta_i;
        BinIn bi{x[bin], y[bin], z[bin], x[bt], y[bt], z[bt]};
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t b = 0; b < nb; b++) {
            if (prec == 0)
                beam_bin_fp64(recs + b * REC, vrecs, g, bi, cphi[iphi], sphi[iphi], fast_diff, acc);
            else
                beam_bin_mixed(recs + b * REC, vrecs, g, bi, fast_diff, acc);
        }
        if (prec != 0) mixed_finish(acc, cphi[iphi], sphi[iphi]);
        for (int q = 0; q < 4; q++) {
            out[2 * (bin + (int64_t)nbins * q)] = acc[2 * q];
            out[2 * (bin + (int64_t)nbins * q) + 1] = acc[2 * q + 1];
        }
    }
}
