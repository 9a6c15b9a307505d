// mueller.hip -- far-field epilogue of one orientation on the device:
// finalise (outputs_mod.f90:63-278): amplitudes x k, total = beam + ext,
// ampl_to_mueller for total/beam/ext (outputs_mod.f90:473-549, Bohren & Huffman
// 3.16), phi integration by simpne (get_1d_mueller, outputs_mod.f90:280-329;
// simpne misc_submod.f90:2392-2531), theta integrals (scattering cross
// sections, asymmetry parameters, back-scattering) and the optional energy
// scaling (outputs_mod.f90:146-184).  FP64 throughout, contraction off so the
// epilogue rounds like the reference.
//
// Also the incoherent orientation sum (summation, outputs_mod.f90:407-471):
// mueller_total += mueller in a fixed order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "../../include/abt_diff.h"
#include "mueller.h"
#include "records.h"
#include "rt_internal.h"
#include "simpne.h"

namespace abt {

constexpr int MB = 256;

__global__ __launch_bounds__(MB) void k_scale_sum(double* __restrict__ ab, double* __restrict__ ae,
                                                  double* __restrict__ tot, double fb, double fe,
                                                  const double* __restrict__ dev_f, int64_t n)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * MB + threadIdx.x;
    if (i >= n) return;
    double b = ab[i], e = ae[i];
    if (dev_f) {   // energy scaling: divide by the factors computed on the device
        b = b / dev_f[0];
        e = e / dev_f[1];
    } else {
        b = b * fb;
        e = e * fe;
    }
    ab[i] = b;
    ae[i] = e;
    tot[i] = b + e;
}

__device__ __forceinline__ void mueller_bin(const double* __restrict__ a, int64_t b, int64_t nb,
                                            double* __restrict__ M)
{
    double s[8], m[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        s[2 * q] = a[2 * (b + nb * q)];
        s[2 * q + 1] = a[2 * (b + nb * q) + 1];
    }
    mueller16(s, m);
#pragma unroll
    for (int e = 0; e < 16; e++) M[b + nb * e] = m[e];
}

// three amplitude arrays -> three Mueller arrays (any output may be null)
__global__ __launch_bounds__(MB) void k_mueller3(const double* __restrict__ a0,
                                                 const double* __restrict__ a1,
                                                 const double* __restrict__ a2, double* M0,
                                                 double* M1, double* M2, int64_t nb)
{
    const int64_t b = (int64_t)blockIdx.x * MB + threadIdx.x;
    if (b >= nb) return;
    if (M0) mueller_bin(a0, b, nb, M0);
    if (M1) mueller_bin(a1, b, nb, M1);
    if (M2) mueller_bin(a2, b, nb, M2);
}

__device__ __forceinline__ double simpne_dev(int n, const double* __restrict__ x,
                                             const double* __restrict__ y, int64_t stride)
{
    return simpne_eval(n, x, y, stride);
}

// mueller_1d(theta, e) of up to three matrices: one lane per (matrix, element, theta)
__global__ __launch_bounds__(MB) void k_phi_integrate(const double* M0, const double* M1,
                                                      const double* M2, const double* __restrict__ phi,
                                                      int nphi, int ntheta, double* o0, double* o1,
                                                      double* o2)
{
    const int64_t t = (int64_t)blockIdx.x * MB + threadIdx.x;
    if (t >= (int64_t)3 * 16 * ntheta) return;
    const int mat = (int)(t / (16 * ntheta));
    const int rem = (int)(t - (int64_t)mat * 16 * ntheta);
    const int e = rem / ntheta, j = rem - e * ntheta;
    const double* M = mat == 0 ? M0 : (mat == 1 ? M1 : M2);
    double* o = mat == 0 ? o0 : (mat == 1 ? o1 : o2);
    if (!M || !o) return;
    const int64_t nb = (int64_t)nphi * ntheta;
    o[j + (int64_t)ntheta * e] = simpne_dev(nphi, phi, M + nb * e + (int64_t)nphi * j, 1);
}

// theta integrals (outputs_mod.f90:133-135, 198-205): the integrands of the three
// matrices in parallel (one workgroup), then one lane per matrix runs simpne in the
// reference's order.
// params: scatt, scatt_beam, scatt_ext, g, g_beam, g_ext, back_scatt, f_beam, f_ext;
// tmp: 3*ntheta scratch.  pass 0 with scaling: computes the energy scaling
// factors f = sqrt(scatt_x / energy_x) into params[7..8].
__global__ __launch_bounds__(MB) void k_theta_params(const double* __restrict__ m0, const double* __restrict__ m1,
                                                     const double* __restrict__ m2, const double* __restrict__ M0,
                                                     const double* __restrict__ theta, int ntheta, int nphi,
                                                     double waveno, int scaling_pass, double ebeam, double eext,
                                                     double* __restrict__ tmp, double* __restrict__ params)
{
#pragma clang fp contract(off)
    __shared__ double sc[3];
    const int tid = threadIdx.x;
    const double w2 = waveno * waveno;
    const double* m[3] = {m0, m1, m2};
    for (int i = tid; i < 3 * ntheta; i += MB) {
        const int k = i / ntheta, j = i - k * ntheta;
        tmp[i] = m[k][j] * sin(theta[j]) / w2;
    }
    __threadfence_block();
    __syncthreads();
    if (tid < 3) sc[tid] = simpne_dev(ntheta, theta, tmp + (int64_t)tid * ntheta, 1);
    __syncthreads();
    if (scaling_pass) {
        if (tid == 0) {
            params[7] = sqrt(sc[1] / ebeam);
            params[8] = sqrt(sc[2] / eext);
        }
        return;
    }
    for (int i = tid; i < 3 * ntheta; i += MB) {
        const int k = i / ntheta, j = i - k * ntheta;
        tmp[i] = m[k][j] * sin(theta[j]) * cos(theta[j]) / sc[k] / w2;
    }
    __threadfence_block();
    __syncthreads();
    if (tid < 3) params[3 + tid] = simpne_dev(ntheta, theta, tmp + (int64_t)tid * ntheta, 1);
    if (tid == 0) {
        int ib = 0;
        double best = fabs(theta[0] - 2 * PI_R4);
        for (int j = 1; j < ntheta; j++) {
            const double d = fabs(theta[j] - 2 * PI_R4);
            if (d < best) { best = d; ib = j; }
        }
        params[0] = sc[0]; params[1] = sc[1]; params[2] = sc[2];
        params[6] = M0[(int64_t)nphi * ib] * 4 * PI_R4 / w2;
    }
}

__global__ __launch_bounds__(MB) void k_accumulate(double* __restrict__ total,
                                                   const double* __restrict__ x, int64_t n)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * MB + threadIdx.x;
    if (i < n) total[i] = total[i] + x[i];
}

}  // namespace abt

static inline unsigned nblk(int64_t n) { return (unsigned)((n + abt::MB - 1) / abt::MB); }

// Device epilogue.  ab/ae: amplitudes (8*nbins doubles, scaled by k in place),
// tot: 8*nbins scratch (total amplitude), M: 3*16*nbins (total, beam, ext; may be
// null), m1d: 3*16*ntheta, tmp: 3*ntheta, params: 9 doubles.
extern "C" int abt_launch_finalise(const abt::DevGrid* g, double* ab,
                                   double* ae, double* tot, double* M, double* m1d, double* tmp,
                                   double* params, int scaling, double ebeam, double eext,
                                   hipStream_t st)
{
    using namespace abt;
    const int64_t nb = g->nbins, n8 = 8 * nb;
    const int nth = g->ntheta, nph = g->nphi;
    double *M0 = M, *M1 = M + 16 * nb, *M2 = M + 32 * nb;
    double *o0 = m1d, *o1 = m1d + 16 * nth, *o2 = m1d + 32 * nth;
    hipLaunchKernelGGL(k_scale_sum, dim3(nblk(n8)), dim3(MB), 0, st, ab, ae, tot, g->waveno, g->waveno,
                       (const double*)nullptr, n8);
    for (int pass = 0; pass < (scaling ? 2 : 1); pass++) {
        if (pass == 1)
            hipLaunchKernelGGL(k_scale_sum, dim3(nblk(n8)), dim3(MB), 0, st, ab, ae, tot, 1.0, 1.0,
                               (const double*)(params + 7), n8);
        hipLaunchKernelGGL(k_mueller3, dim3(nblk(nb)), dim3(MB), 0, st, (const double*)tot,
                           (const double*)ab, (const double*)ae, M0, M1, M2, nb);
        hipLaunchKernelGGL(k_phi_integrate, dim3(nblk((int64_t)48 * nth)), dim3(MB), 0, st,
                           (const double*)M0, (const double*)M1, (const double*)M2, g->phi, nph, nth,
                           o0, o1, o2);
        if (scaling && pass == 0)
            hipLaunchKernelGGL(k_theta_params, dim3(1), dim3(MB), 0, st, (const double*)o0,
                               (const double*)o1, (const double*)o2, (const double*)M0, g->theta,
                               nth, nph, g->waveno, 1, ebeam, eext, tmp, params);
    }
    hipLaunchKernelGGL(k_theta_params, dim3(1), dim3(MB), 0, st, (const double*)o0, (const double*)o1,
                       (const double*)o2, (const double*)M0, g->theta, nth, nph, g->waveno, 0,
                       ebeam, eext, tmp, params);
    return (int)hipGetLastError();
}

extern "C" int abt_launch_accumulate(double* total, const double* x, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(abt::k_accumulate, dim3(nblk(n)), dim3(abt::MB), 0, st, total, x, n);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// The epilogue without allocation (abt_finalise_dev_static): its scratch (the total
// amplitude, 8 doubles per bin, and the theta integrands, 3 per theta) lives in a
// file-scope device array, one per GPU for the life of the process.  In host code the
// array's name is not a device address: its address on the current device comes from
// hipGetSymbolAddress, once per device, and only that address reaches the kernels.
namespace {

constexpr int64_t STATIC_SCRATCH_DOUBLES = (int64_t)8 << 20;   // 64 MiB
static __device__ __align__(256) double scratch[STATIC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// The scratch array's address on the library's device (looked up on first use there).
int scratch_on_device(double** out)
{
    int rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return abt_fail(ABT_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return abt_fail(ABT_ERR_NODEV, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return abt_fail(ABT_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" int64_t abt_finalise_static_max_bins(int64_t n_theta)
{
    return n_theta < 0 ? 0 : (STATIC_SCRATCH_DOUBLES - 3 * n_theta) / 8;
}

extern "C" int abt_finalise_dev_static(const abt_dev_grid* g, double* ab, double* ae, int scaling,
                                       double ebeam, double eext, double* dev_mueller, double* dev_m1d,
                                       double* dev_params, void* stream)
{
    if (!g || !ab || !ae || !dev_mueller || !dev_m1d || !dev_params)
        return abt_fail(ABT_ERR_ARG, "abt_finalise_dev_static: null argument");
    if (g->d.nphi < 3 || g->d.ntheta < 3)
        return abt_fail(ABT_ERR_ARG, "finalise needs at least 3 theta and 3 phi values (simpne)");
    if (scaling && !(ebeam > 0 && eext > 0))
        return abt_fail(ABT_ERR_ARG, "energy scaling needs positive beam and ext energies");
    const int64_t need = 8 * (int64_t)g->nbins + 3 * (int64_t)g->d.ntheta;
    if (need > STATIC_SCRATCH_DOUBLES)
        return abt_fail(ABT_ERR_ARG,
                        "abt_finalise_dev_static: %lld bins x %d theta need %lld scratch doubles, more than the "
                        "%lld it holds (at most %lld bins at this theta count); use abt_finalise_dev",
                        (long long)g->nbins, g->d.ntheta, (long long)need, (long long)STATIC_SCRATCH_DOUBLES,
                        (long long)abt_finalise_static_max_bins(g->d.ntheta));
    double* sp = nullptr;
    int rc = scratch_on_device(&sp);
    if (rc) return rc;
    double* tot = sp;
    double* tmp = sp + 8 * g->nbins;
    int e = abt_launch_finalise(&g->d, ab, ae, tot, dev_mueller, dev_m1d, tmp, dev_params, scaling, ebeam, eext,
                                (hipStream_t)stream);
    if (e) return abt_fail(ABT_ERR_HIP, "finalise launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}
