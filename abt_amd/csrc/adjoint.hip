// adjoint.hip -- gradients of a loss of the orientation-averaged far field with
// respect to every aperture's vertices and amplitude matrix.
//
// Input: G = dL/dX_o per orientation and bin, X_o = k (ampl_far_beam + ampl_far_ext_diff)
// (the per-orientation total amplitude the batched forward writes).  Every beam b of
// orientation o (both trees) then has dL/dA_b = k G_o, and its gradients are sums over
// bins of per-(beam, bin) terms (adjbody.h).
//
// Work decomposition: one lane per beam, 64 lanes per workgroup, all lanes of a
// workgroup from one orientation and tree, so every lane reads the same bin at the same
// time: the bin's position and Lam' come through scalar loads and stay in scalar
// registers, the beam's constants and its 27 accumulators stay in vector registers, and
// no reduction across lanes is needed.  A unit may cover a chunk of the bins only (when
// there are too few beams to fill the chip); chunk partials are then summed in chunk
// order.  Every sum has a fixed order: results repeat bit for bit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adjbody.h"
#include "adjoint.h"

namespace abt {

constexpr int AL = 64;   // lanes per workgroup (one wavefront)

// Lam' = k G T_phi^T per (orientation, bin), [o][bin][8] (FP64, or FP32 for the mixed
// kernel); G in the batched forward's amplitude layout [o][8][nbins].
// T_phi = [[-s, -c], [c, -s]] (diff_mod.f90:646-650).
template <typename T>
__global__ __launch_bounds__(256) void k_adj_lam(const double* __restrict__ G, const double* __restrict__ cphi,
                                                 const double* __restrict__ sphi, int nphi, int nbins,
                                                 int norient, double k, T* __restrict__ lam)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)norient * nbins) return;
    const int o = (int)(i / nbins), bin = (int)(i - (int64_t)o * nbins);
    const int iphi = bin % nphi;
    const double c = cphi[iphi], s = sphi[iphi];
    const double* g = G + (int64_t)o * 8 * nbins + bin;
    double L[8];
#pragma unroll
    for (int q = 0; q < 8; q++) L[q] = k * g[(int64_t)q * nbins];
    T* out = lam + i * 8;
    // pair (r, col): index 2 (r + 2 col)
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++) {
            const double l1 = L[2 * r + ri], l2 = L[2 * (r + 2) + ri];
            out[2 * r + ri] = (T)(-s * l1 - c * l2);            // Lam'(r,1)
            out[2 * (r + 2) + ri] = (T)(c * l1 - s * l2);       // Lam'(r,2)
        }
}

// per-bin inputs: x, y, z, x_t, y_t (theta_i bin of the same phi), |x|, k|x|/2pi, 0
__global__ __launch_bounds__(256) void k_adj_bins(DevGrid g, double* __restrict__ bins)
{
    const int bin = blockIdx.x * 256 + threadIdx.x;
    if (bin >= g.nbins) return;
    const int bt = bin % g.nphi + g.nphi * g.theta_i;
    double* o = bins + (int64_t)bin * 8;
    o[0] = g.x[bin]; o[1] = g.y[bin]; o[2] = g.z[bin];
    o[3] = g.x[bt]; o[4] = g.y[bt];
    o[5] = g.rad[bin]; o[6] = g.kr_rev[bin]; o[7] = 0.0;
}

// two waves per SIMD: the beam's constants and 27 FP64 accumulators live in registers
template <int PREC>
__global__ __launch_bounds__(AL, 2) void k_adjoint(AdjArgs a)
{
    const AdjUnit U = a.units[blockIdx.x];
    const int lane = threadIdx.x;
    const bool act = lane < U.count;
    const int64_t beam = U.start + (act ? lane : 0);
    const double* recs = U.tree ? a.recs_e : a.recs_b;
    const double* r = recs + beam * REC;
    AdjBeam B;
    adj_beam_from_record(r, B);
    const bool tri = r[R_NV] == 3.0;
    AdjAcc acc;
    adj_zero(acc);
    const int chunk_bins = (a.nbins + a.nchunks - 1) / a.nchunks;
    const int b0 = U.chunk * chunk_bins;
    const int b1 = min(a.nbins, b0 + chunk_bins);
    if (PREC == 0) {
        const double* lam = a.lam + (int64_t)U.o * a.nbins * 8;
        for (int i = b0; i < b1; i++) {
            const double* bb = a.bins + (int64_t)i * 8;
            const AdjBin bn{bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], bb[6]};
            adj_bin_fp64(B, bn, lam + (int64_t)i * 8, a.waveno, a.lambda, acc);
        }
    } else {
        // FP32 partial sums over groups of ADJ_GROUP bins, added into FP64
        const float* lam = reinterpret_cast<const float*>(a.lam) + (int64_t)U.o * a.nbins * 8;
        for (int i0 = b0; i0 < b1; i0 += ADJ_GROUP) {
            AdjAccF af;
            adj_zero_f(af);
            const int i1 = min(b1, i0 + ADJ_GROUP);
            for (int i = i0; i < i1; i++) {
                const double* bb = a.bins + (int64_t)i * 8;
                const AdjBin bn{bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], bb[6]};
                adj_bin_mixed(B, bn, lam + (int64_t)i * 8, a.waveno, a.lambda, af);
            }
            adj_add(acc, af);
        }
    }
    if (!act) return;
    double gv[9], ga[8];
    adj_finish(B, r + R_V20, r[R_SIGMA], acc, gv, ga);
    if (!tri) {   // the adjoint covers triangular apertures only
#pragma unroll
        for (int q = 0; q < 9; q++) gv[q] = __builtin_nan("");
#pragma unroll
        for (int q = 0; q < 8; q++) ga[q] = __builtin_nan("");
    }
    // [chunk][tree offset + beam][17]
    const int64_t row = (U.tree ? a.nb_beam : 0) + beam;
    double* o = a.part + ((int64_t)U.chunk * a.nb_total + row) * 17;
#pragma unroll
    for (int q = 0; q < 9; q++) o[q] = gv[q];
#pragma unroll
    for (int q = 0; q < 8; q++) o[9 + q] = ga[q];
}

// sum the chunk partials in chunk order and split them into the four outputs
__global__ __launch_bounds__(256) void k_adj_reduce(const double* __restrict__ part, int nchunks, int64_t nb_beam,
                                                    int64_t nb_total, double* gvb, double* gab, double* gve,
                                                    double* gae)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb_total * 17) return;
    const int64_t row = i / 17;
    const int q = (int)(i - row * 17);
    double s = part[i];
    for (int c = 1; c < nchunks; c++) s += part[(int64_t)c * nb_total * 17 + i];
    const bool ext = row >= nb_beam;
    const int64_t b = ext ? row - nb_beam : row;
    if (q < 9) {
        double* o = ext ? gve : gvb;
        if (o) o[b * 9 + q] = s;
    } else {
        double* o = ext ? gae : gab;
        if (o) o[b * 8 + q - 9] = s;
    }
}

}  // namespace abt

using namespace abt;

extern "C" int abt_launch_adj_bins(const DevGrid* g, double* bins, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_bins, dim3((unsigned)((g->nbins + 255) / 256)), dim3(256), 0, st, *g, bins);
    return (int)hipGetLastError();
}

extern "C" int abt_launch_adjoint(const AdjArgs* a, int prec, int nunits, const double* G, const DevGrid* g,
                                  int norient, double* gvb, double* gab, double* gve, double* gae,
                                  hipStream_t st)
{
    const int64_t nl = (int64_t)norient * g->nbins;
    if (prec == 0)
        hipLaunchKernelGGL(k_adj_lam<double>, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, G, g->cphi,
                           g->sphi, g->nphi, g->nbins, norient, g->waveno, a->lam);
    else
        hipLaunchKernelGGL(k_adj_lam<float>, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, G, g->cphi,
                           g->sphi, g->nphi, g->nbins, norient, g->waveno, reinterpret_cast<float*>(a->lam));
    if (nunits > 0) {
        if (prec == 0)
            hipLaunchKernelGGL(k_adjoint<0>, dim3(nunits), dim3(AL), 0, st, *a);
        else
            hipLaunchKernelGGL(k_adjoint<1>, dim3(nunits), dim3(AL), 0, st, *a);
    }
    const int64_t n = a->nb_total * 17;
    if (n > 0)
        hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           (const double*)a->part, a->nchunks, a->nb_beam, a->nb_total, gvb, gab, gve, gae);
    return (int)hipGetLastError();
}
