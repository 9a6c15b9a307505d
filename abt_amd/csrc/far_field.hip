// far_field.hip -- coherent far-field sum of prepared beams onto the (theta, phi) bins.
//
// Work decomposition (both kernels): one lane per far-field bin, bins contiguous
// with phi fastest (the Fortran (nphi, ntheta) layout), 256 lanes per workgroup.
// Beam records are staged in LDS one tile at a time (cooperative coalesced
// loads, then broadcast reads) and every lane walks the tile in beam order, so
// each bin's sum is formed in a fixed order and is run-to-run deterministic.
// blockIdx.y splits the beam list into chunks so that a launch has many short
// workgroups (a small last round); chunk partials are then summed in chunk order
// by k_reduce_chunks.
//
// k_ff_fp64  -- the reference's formulas in FP64 (diffraction, diff_mod.f90:576-670;
//               karczewski :124-198; contour_integral :357-421), every phase formed with
//               the reference's operations (bit-identical arguments of every sine and
//               cosine), the amplitude algebra around them reorganised (beam_bin_fp64).
// k_ff_fp64r -- (PREC 4) the mixed kernel's algorithm below with every step in FP64.
// k_ff_mixed -- the same quantity restructured for the vector ALUs:
//               * geometry, phases and their reduction to octants in FP64
//                 (the bin direction, q = k (prop - khat), vertex phases,
//                 the k|r - com| ~ 1.2e7 rad far-field phase);
//               * polarisation (Karczewski-Wolf, scattering-plane rotation)
//                 and the aperture integral body in FP32: the edge sum is the
//                 rotation-invariant polygon form factor
//                   sum_ref = -(1/2pi) sum_j (qx dy_j - qy dx_j)/|q|^2 sinc(K_j/2) e^{i mu_j}
//                 (q in revolutions/um), which equals the reference's split
//                 kxx/kyy sum but is singular only at q -> 0;
//               * near q -> 0 (|q| a < SERIES_QA) an FP64 power series of the
//                 form factor (complete homogeneous symmetric polynomials of
//                 the vertex phases);
//               * the scattering-plane rotation temp_rot1 depends on phi only,
//                 so it is applied once per bin after the beam sum; the
//                 polarisation matrix rot4 . diff_ampl has the form g I + h J
//                 and is folded into two complex weights per beam;
//               * accumulation in FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ffbody.h"

namespace abt {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// The mixed kernel with up to 16-beam tiles is held to 128 VGPRs (4 waves per SIMD), its
// fastest occupancy on gfx950.
template <int PREC, int TB>
__global__ __launch_bounds__(BLOCK, (PREC == 1 && TB <= 16) ? 4 : (PREC == 4 ? 2 : 1))
void k_far_field(DevGrid g, const double* __restrict__ recs, const double* __restrict__ vrecs,
                 int64_t nb, int64_t per_chunk, int fast_diff, double* __restrict__ out,
                 double* __restrict__ work)
{
    // Mixed kernel: two record tiles; the next tile's records are loaded into registers
    // while the current one is summed, then stored to the other buffer (one barrier per
    // tile).  The FP64 bodies need their registers: one tile, loaded between barriers.
    constexpr bool DB = PREC == 1;
    __shared__ double tile[(DB ? 2 : 1) * TB * REC];
    // the exponential table and the FP64r body's constants, or the table with its low parts
    __shared__ double etab[PREC == 4 ? ETAB_LDS : (PREC != 1 ? ETAB_GRID : 1)];
    constexpr int PF = DB ? (TB * REC + BLOCK - 1) / BLOCK : 1;   // doubles per lane per tile
    const int tid = threadIdx.x;
    if (PREC != 1) {   // published by the barriers of the tile loop before any use
        for (int q = tid; q < (PREC == 4 ? ETAB_DOUBLES : ETAB_GRID); q += BLOCK) etab[q] = g.etab[q];
        if (PREC == 4 && tid == 0) fp64r_constants(g, etab + ETAB_DOUBLES);
    }
    const int bin = blockIdx.x * BLOCK + tid;
    const bool active = bin < g.nbins;
    const int bc = active ? bin : g.nbins - 1;
    const int iphi = bc % g.nphi;
    BinIn bi;
    bi.x = g.x[bc]; bi.y = g.y[bc]; bi.z = g.z[bc];
    const int bt = iphi + g.nphi * g.theta_i;
    bi.xt = g.x[bt]; bi.yt = g.y[bt]; bi.zt = g.z[bt];
    bi.rad = PREC == 4 ? g.rad[bc] : 0.0;
    const double cp = g.cphi[iphi], sp = g.sphi[iphi];
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0;
    const int64_t b0 = (int64_t)blockIdx.y * per_chunk;
    const int64_t b1 = min(nb, b0 + per_chunk);
    if (DB) {
        // the bin's inputs must have arrived before the tile loop: a use inside it would make
        // the compiler wait there for the older loads, and with them for the prefetched tile
        asm volatile("" : "+v"(bi.x), "+v"(bi.y), "+v"(bi.z), "+v"(bi.xt), "+v"(bi.yt), "+v"(bi.zt));
    }
    if (DB && b0 < b1) {
        const int nt0 = (int)min((int64_t)TB, b1 - b0);
        const double* src = recs + b0 * REC;
        for (int q = tid; q < nt0 * REC; q += BLOCK) tile[q] = src[q];
    }
    if (DB) __syncthreads();
    int buf = 0;
    for (int64_t t0 = b0; t0 < b1; t0 += TB) {
        const int nt = (int)min((int64_t)TB, b1 - t0);
        const int64_t tn = t0 + TB;
        const int ntn = DB && tn < b1 ? (int)min((int64_t)TB, b1 - tn) : 0;
        const double* srcn = recs + (ntn ? tn : 0) * REC;
        double pre[PF];
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int q = tid + k * BLOCK;
            pre[k] = q < ntn * REC ? srcn[q] : 0.0;
        }
        if (!DB) {
            __syncthreads();
            const double* src = recs + t0 * REC;
            for (int q = tid; q < nt * REC; q += BLOCK) tile[q] = src[q];
            __syncthreads();
        }
        const double* cur = tile + buf * (TB * REC);
        if (active) {
            if (PREC == 1) {
                // FP32 partial sums over groups of MIXED_GROUP beams, added into FP64
                for (int g0 = 0; g0 < nt; g0 += MIXED_GROUP) {
                    float accf[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) accf[q] = 0.f;
                    const int g1 = min(nt, g0 + MIXED_GROUP);
                    for (int bb = g0; bb < g1; bb++) beam_bin_mixed(cur + bb * REC, vrecs, g, bi, fast_diff, accf);
#pragma unroll
                    for (int q = 0; q < 8; q++) acc[q] += (double)accf[q];
                }
            } else if (PREC == 4) {
                for (int bb = 0; bb < nt; bb++) beam_bin_fp64r(cur + bb * REC, vrecs, g, bi, etab, fast_diff, acc);
            } else {
                for (int bb = 0; bb < nt; bb++)
                    beam_bin_fp64(cur + bb * REC, vrecs, g, bi, etab, fast_diff, acc, PREC == 2);
            }
        }
        if (DB) {
            double* nxt = tile + (buf ^ 1) * (TB * REC);
#pragma unroll
            for (int k = 0; k < PF; k++) {
                const int q = tid + k * BLOCK;
                if (q < ntn * REC) nxt[q] = pre[k];
            }
            __syncthreads();
            buf ^= 1;
        }
    }
    if (!active) return;
    if (PREC == 0) apply_temp_rot1(acc, cp, sp);
    if (PREC == 1) mixed_finish(acc, cp, sp);
    if (PREC == 4) fp64r_finish(acc, cp, sp, g.kr_rev[bin]);
    const int nbins = g.nbins;
    if (PREC == 2) {   // incoherent intensity: one double per bin, never chunked
        out[bin] = acc[0];
        return;
    }
    if (work == nullptr) {
        double2* o2 = reinterpret_cast<double2*>(out);
#pragma unroll
        for (int q = 0; q < 4; q++) o2[bin + (int64_t)nbins * q] = make_double2(acc[2 * q], acc[2 * q + 1]);
    } else {
        double* wc = work + (int64_t)blockIdx.y * 8 * nbins;
#pragma unroll
        for (int q = 0; q < 8; q++) wc[(int64_t)q * nbins + bin] = acc[q];
    }
}

// Sum chunk partials in chunk order: work[(c*8 + comp)*nbins + bin] -> out.
__global__ __launch_bounds__(BLOCK) void k_reduce_chunks(const double* __restrict__ work,
                                                         int nchunks, int nbins,
                                                         double* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;   // over 8*nbins
    if (idx >= (int64_t)8 * nbins) return;
    const int comp = (int)(idx / nbins);
    const int bin = (int)(idx - (int64_t)comp * nbins);
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s = s + work[((int64_t)c * 8 + comp) * nbins + bin];
    const int q = comp >> 1, ri = comp & 1;
    out[2 * (bin + (int64_t)nbins * q) + ri] = s;
}

}  // namespace abt

// ---------------------------------------------------------------------------
// launch helpers (called by the runtime)
namespace abt {
__global__ __launch_bounds__(BLOCK) void k_fill_zero(double* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = 0.0;
}
}  // namespace abt

// Zero-fill of n doubles (a kernel, not hipMemsetAsync, so callers may capture it in a graph).
extern "C" int abt_launch_zero(double* out, int64_t n, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(abt::k_fill_zero, dim3((unsigned)((n + abt::BLOCK - 1) / abt::BLOCK)),
                       dim3(abt::BLOCK), 0, st, out, n);
    return (int)hipGetLastError();
}

extern "C" int abt_launch_far_field(const abt::DevGrid* g, const double* recs, const double* vrecs,
                                    int64_t nb, int prec, int tb, int fast_diff, int nchunks,
                                    double* out, double* work, hipStream_t st)
{
    using namespace abt;
    if (g->nbins <= 0) return 0;
    // chunk lengths are multiples of the largest tile (32 beams), so the mixed kernel's
    // FP32 beam groups do not depend on the chunking
    int64_t per_chunk = (nb + nchunks - 1) / nchunks;
    per_chunk = per_chunk < 32 ? 32 : (per_chunk + 31) / 32 * 32;
    nchunks = (int)((nb + per_chunk - 1) / per_chunk);
    if (nchunks < 1) nchunks = 1;
    dim3 grid((g->nbins + BLOCK - 1) / BLOCK, nchunks);
    double* w = nchunks > 1 ? work : nullptr;
#define L(P, T) hipLaunchKernelGGL((k_far_field<P, T>), grid, dim3(BLOCK), 0, st, *g, recs, vrecs, nb, per_chunk, fast_diff, out, w)
    if (prec == 0) {
        if (tb == 8) L(0, 8); else if (tb == 32) L(0, 32); else L(0, 16);
    } else if (prec == 1) {
        if (tb == 8) L(1, 8); else if (tb == 32) L(1, 32); else L(1, 16);
    } else if (prec == 3) {   // ABT_FP64_FAST: the restructured FP64 body
        L(4, 16);
    } else {
        L(2, 16);
    }
#undef L
    if (nchunks > 1) {
        const int64_t n = (int64_t)8 * g->nbins;
        hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                           st, work, nchunks, g->nbins, out);
    }
    return (int)hipGetLastError();
}
