// batch.hip -- orientation-averaged far field of many orientations in one launch.
//
// Replaces the orientation loop of the reference drivers (pbt_seq_mod.f90:146-269,
// pbt_mpi_mod.f90 over MPI ranks): per orientation diff_main (coherent sums of the
// beam tree and of the external-diffraction tree), finalise (amplitudes x k, total =
// beam + ext, Mueller matrices, phi/theta integrals; outputs_mod.f90:63-278) and
// summation (Mueller matrices summed incoherently over orientations,
// outputs_mod.f90:407-471).
//
// Work decomposition.  A work unit is (orientation chunk c, bin tile t): the
// orientations of chunk c (a contiguous range, chunks balanced by beam count) on
// BLOCK * BPT bins.  A persistent grid of workgroups takes units from a queue
// (one global atomic counter).  Inside a unit every lane owns BPT bins and walks
// the orientations of the chunk in order; per orientation it sums the beam tree
// and then the ext tree (beam records staged in LDS, TB at a time, as in the
// single-orientation kernel), applies k, forms the Mueller matrix of the total
// and adds it to the chunk's partial Mueller sum, and adds the bin's weighted S11
// values to the orientation's integrals.  Every output element is written by one
// unit, and every sum has a fixed order (beams in tree order, orientations in
// order within a chunk, chunks in order, tiles in order), so results do not
// depend on scheduling and repeat bit for bit.
//
// Per-orientation integrals (for the asymmetry parameters, which are ratios per
// orientation and so cannot be formed from the summed Mueller matrix):
//   I1 = sum_bins W1 S11,  I2 = sum_bins W2 S11,   for total, beam and ext,
// with W1 = wtheta(theta) sin(theta) wphi(phi) / k^2 and W2 = W1 cos(theta), where
// wtheta, wphi are the simpne weights of the theta and phi nodes: I1 is the
// orientation's scattering cross section (outputs_mod.f90:133-135) and I2 / I1 its
// asymmetry parameter (outputs_mod.f90:198-200).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adjbody.h"
#include "batch.h"
#include "ffbody.h"
#include "mueller.h"
#include "simpne.h"

namespace abt {

constexpr int BB = 256;   // lanes per workgroup

template <int PREC, int TB, int BPT>
__device__ __forceinline__ void sweep(const BatchArgs& a, const double* __restrict__ recs,
                                      const double* __restrict__ vrecs, int64_t b0, int64_t b1,
                                      double* tile, const BinIn* bi, const double* cp,
                                      const double* sp, double (*acc)[8])
{
    const int tid = threadIdx.x;
    for (int64_t t0 = b0; t0 < b1; t0 += TB) {
        const int nt = (int)min((int64_t)TB, b1 - t0);
        __syncthreads();
        const double* src = recs + t0 * REC;
        for (int q = tid; q < nt * REC; q += BB) tile[q] = src[q];
        __syncthreads();
        if (PREC == 1) {
            // FP32 partial sums over groups of MIXED_GROUP beams (counted from the
            // orientation's first beam; TB is a multiple), added into FP64
            for (int g0 = 0; g0 < nt; g0 += MIXED_GROUP) {
                float accf[BPT][8];
#pragma unroll
                for (int j = 0; j < BPT; j++)
#pragma unroll
                    for (int q = 0; q < 8; q++) accf[j][q] = 0.f;
                const int g1 = min(nt, g0 + MIXED_GROUP);
                for (int bb = g0; bb < g1; bb++) {
#pragma unroll
                    for (int j = 0; j < BPT; j++)
                        beam_bin_mixed(tile + bb * REC, vrecs, a.g, bi[j], a.fast_diff, accf[j]);
                }
#pragma unroll
                for (int j = 0; j < BPT; j++)
#pragma unroll
                    for (int q = 0; q < 8; q++) acc[j][q] += (double)accf[j][q];
            }
        } else {
            for (int bb = 0; bb < nt; bb++) {
#pragma unroll
                for (int j = 0; j < BPT; j++) {
                    if (PREC == 0)
                        beam_bin_fp64(tile + bb * REC, vrecs, a.g, bi[j], cp[j], sp[j], a.fast_diff, acc[j]);
                    else {
                        AdjBeam B;
                        adj_beam_from_record(tile + bb * REC, B);
                        model_bin_fp64(B, bi[j].x, bi[j].y, bi[j].z, bi[j].xt, bi[j].yt, bi[j].rad, a.g.waveno,
                                       a.g.lambda, acc[j]);
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Register budget: the per-orientation beam-tree result waits in LDS (xbs) while
// the ext tree is summed, and the phi rotation and integral weights are re-read
// after the sweeps, so the live state across a sweep is the accumulators and the
// bin position only.
template <int PREC, int TB, int BPT>
__global__ __launch_bounds__(BB, PREC == 1 ? 4 : 3) void k_batch(BatchArgs a)
{
    __shared__ double tile[TB * REC];
    __shared__ double xbs[8][BB * BPT];
    __shared__ double red[BB / 64][6];
    __shared__ unsigned s_unit;
    const int tid = threadIdx.x;
    const int nbins = a.g.nbins;
    const double k = a.g.waveno;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_unit = atomicAdd(a.counter, 1u);
        __syncthreads();
        const unsigned u = s_unit;
        if (u >= (unsigned)a.nunits) break;   // uniform over the workgroup
        const int c = (int)(u / (unsigned)a.ntiles);
        const int t = (int)(u - (unsigned)c * (unsigned)a.ntiles);
        const int bin0 = t * BPT * BB + tid;   // bin of slot j: bin0 + j * BB

        BinIn bi[BPT];
        double cp[BPT], sp[BPT];
#pragma unroll
        for (int j = 0; j < BPT; j++) {
            const int bc = min(bin0 + j * BB, nbins - 1);
            const int iphi = bc % a.g.nphi;
            bi[j].x = a.g.x[bc]; bi[j].y = a.g.y[bc]; bi[j].z = a.g.z[bc];
            const int bt = iphi + a.g.nphi * a.g.theta_i;
            bi[j].xt = a.g.x[bt]; bi[j].yt = a.g.y[bt]; bi[j].zt = a.g.z[bt];
            bi[j].rad = PREC == 3 ? a.g.rad[bc] : 0.0;
            cp[j] = PREC != 1 ? a.g.cphi[iphi] : 0.0;   // the mixed body rotates after the sum
            sp[j] = PREC != 1 ? a.g.sphi[iphi] : 0.0;
        }
        const int o0 = a.chunk_o[c], o1 = a.chunk_o[c + 1];
        double* mp = a.mpart + (int64_t)c * 16 * nbins;
        for (int o = o0; o < o1; o++) {
            double acc[BPT][8];
            // segment 0: the beam tree (result parked in LDS), segment 1: the ext tree
#pragma nounroll
            for (int seg = 0; seg < 2; seg++) {
#pragma unroll
                for (int j = 0; j < BPT; j++)
#pragma unroll
                    for (int q = 0; q < 8; q++) acc[j][q] = 0.0;
                const double* recs = seg ? a.recs_e : a.recs_b;
                const double* vrecs = seg ? a.vrecs_e : a.vrecs_b;
                const int64_t* off = seg ? a.eoff : a.boff;
                sweep<PREC, TB, BPT>(a, recs, vrecs, off[o], off[o + 1], tile, bi, cp, sp, acc);
                if (seg == 1) break;
#pragma unroll
                for (int j = 0; j < BPT; j++) {
                    if (PREC == 1) {
                        const int iphi = min(bin0 + j * BB, nbins - 1) % a.g.nphi;
                        mixed_finish(acc[j], a.g.cphi[iphi], a.g.sphi[iphi]);
                    }
                    if (PREC == 3) {
                        const int bc = min(bin0 + j * BB, nbins - 1);
                        model_finish(acc[j], cp[j], sp[j], a.g.kr_rev[bc]);
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        xbs[q][j * BB + tid] = acc[j][q] * k;   // ampl_far_beam * waveno (outputs_mod.f90:104)
                }
            }
            double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < BPT; j++) {
                const int bin = bin0 + j * BB;
                const int bc = min(bin, nbins - 1);
                if (PREC == 1) {
                    const int iphi = bc % a.g.nphi;
                    mixed_finish(acc[j], a.g.cphi[iphi], a.g.sphi[iphi]);
                }
                if (PREC == 3) model_finish(acc[j], cp[j], sp[j], a.g.kr_rev[bc]);
                double xb[8], xe[8], xt[8], m[16];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    xb[q] = xbs[q][j * BB + tid];
                    xe[q] = acc[j][q] * k;      // outputs_mod.f90:105
                    xt[q] = xb[q] + xe[q];      // outputs_mod.f90:108
                }
                mueller16(xt, m);
                if (bin < nbins) {
                    if (a.ampl_out) {
                        double* ao = a.ampl_out + (int64_t)o * 8 * nbins + bin;
#pragma unroll
                        for (int q = 0; q < 8; q++) ao[(int64_t)q * nbins] = xt[q];
                    }
                    double* mo = mp + bin;
                    if (o == o0) {
#pragma unroll
                        for (int e = 0; e < 16; e++) mo[(int64_t)e * nbins] = m[e];
                    } else {
#pragma unroll
                        for (int e = 0; e < 16; e++) mo[(int64_t)e * nbins] += m[e];
                    }
                    const double w1 = a.w1[bin], w2 = a.w2[bin];
                    const double sb = mueller_s11(xb), se = mueller_s11(xe);
                    v[0] = fma(w1, m[0], v[0]);
                    v[1] = fma(w2, m[0], v[1]);
                    v[2] = fma(w1, sb, v[2]);
                    v[3] = fma(w2, sb, v[3]);
                    v[4] = fma(w1, se, v[4]);
                    v[5] = fma(w2, se, v[5]);
                }
            }
#pragma unroll
            for (int q = 0; q < 6; q++) v[q] = wave_sum(v[q]);
            if ((tid & 63) == 0) {
#pragma unroll
                for (int q = 0; q < 6; q++) red[tid >> 6][q] = v[q];
            }
            __syncthreads();
            if (tid == 0) {
                double* ip = a.ipart + ((int64_t)o * a.ntiles + t) * 6;
#pragma unroll
                for (int q = 0; q < 6; q++) ip[q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
            }
            __syncthreads();
        }
    }
}

// mueller[e][bin] = sum over chunks, in chunk order
__global__ __launch_bounds__(BB) void k_batch_reduce_m(const double* __restrict__ mpart, int nchunks,
                                                       int64_t n, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * BB + threadIdx.x;
    if (i >= n) return;
    double s = mpart[i];
    for (int c = 1; c < nchunks; c++) s += mpart[(int64_t)c * n + i];
    out[i] = s;
}

// integrals[o][q] = sum over tiles, in tile order
__global__ __launch_bounds__(BB) void k_batch_reduce_i(const double* __restrict__ ipart, int ntiles,
                                                       int norient, double* __restrict__ out)
{
    const int i = blockIdx.x * BB + threadIdx.x;
    if (i >= norient * 6) return;
    const int o = i / 6, q = i - 6 * o;
    const double* p = ipart + (int64_t)o * ntiles * 6 + q;
    double s = 0.0;
    for (int t = 0; t < ntiles; t++) s += p[(int64_t)t * 6];
    out[i] = s;
}

// mueller_1d(theta, e) = simpne over phi of the summed Mueller matrix (get_1d_mueller,
// outputs_mod.f90:280-329); the reference sums the per-orientation 1-D matrices,
// which is the same linear map applied before the sum.
__global__ __launch_bounds__(BB) void k_batch_phi(const double* __restrict__ M, const double* __restrict__ phi,
                                                  int nphi, int ntheta, double* __restrict__ out)
{
    const int i = blockIdx.x * BB + threadIdx.x;
    if (i >= 16 * ntheta) return;
    const int e = i / ntheta, j = i - e * ntheta;
    out[i] = simpne_eval(nphi, phi, M + (int64_t)nphi * ntheta * e + (int64_t)nphi * j, 1);
}

__global__ void k_zero_u32(unsigned int* p) { if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0u; }

template <int PREC, int TB, int BPT>
const void* batch_kernel() { return reinterpret_cast<const void*>(&k_batch<PREC, TB, BPT>); }

}  // namespace abt

// ---------------------------------------------------------------------------
// launch helpers (called by the runtime)
using abt::BatchArgs;

extern "C" const void* abt_batch_kernel(int prec, int tb, int bpt)
{
    using namespace abt;
#define K(P, T, B) if (prec == P && tb == T && bpt == B) return batch_kernel<P, T, B>();
    K(0, 8, 1) K(0, 16, 1) K(0, 32, 1) K(0, 8, 2) K(0, 16, 2) K(0, 32, 2)
    K(1, 8, 1) K(1, 16, 1) K(1, 32, 1) K(1, 8, 2) K(1, 16, 2) K(1, 32, 2)
    K(3, 16, 1)
#undef K
    return nullptr;
}

extern "C" int abt_launch_batch(const BatchArgs* a, int prec, int tb, int bpt, int wgs, int norient,
                                double* mueller, double* mueller_1d, double* integrals, hipStream_t st)
{
    using namespace abt;
    const void* kern = abt_batch_kernel(prec, tb, bpt);
    if (!kern) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, a->counter);
    BatchArgs args = *a;
    void* params[] = {&args};
    hipError_t e = hipLaunchKernel(kern, dim3(wgs), dim3(BB), params, 0, st);
    if (e != hipSuccess) return (int)e;
    const int64_t n = (int64_t)16 * a->g.nbins;
    hipLaunchKernelGGL(k_batch_reduce_m, dim3((unsigned)((n + BB - 1) / BB)), dim3(BB), 0, st,
                       (const double*)a->mpart, a->nchunks, n, mueller);
    hipLaunchKernelGGL(k_batch_reduce_i, dim3((unsigned)((norient * 6 + BB - 1) / BB)), dim3(BB), 0, st,
                       (const double*)a->ipart, a->ntiles, norient, integrals);
    if (mueller_1d)
        hipLaunchKernelGGL(k_batch_phi, dim3((unsigned)((16 * a->g.ntheta + BB - 1) / BB)), dim3(BB), 0, st,
                           (const double*)mueller, a->g.phi, a->g.nphi, a->g.ntheta, mueller_1d);
    return (int)hipGetLastError();
}
