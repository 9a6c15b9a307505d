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

#include "records.h"

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

struct Cx { double r, i; };
__device__ __forceinline__ Cx mc(Cx a, Cx b)  // a * conjg(b)
{
#pragma clang fp contract(off)
    return Cx{a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i};
}

__device__ __forceinline__ void mueller_bin(const double* __restrict__ a, int64_t b, int64_t nb,
                                            double* __restrict__ M)
{
#pragma clang fp contract(off)
    const Cx s11{a[2 * b], a[2 * b + 1]};
    const Cx s21{a[2 * (b + nb)], a[2 * (b + nb) + 1]};
    const Cx s12{a[2 * (b + 2 * nb)], a[2 * (b + 2 * nb) + 1]};
    const Cx s22{a[2 * (b + 3 * nb)], a[2 * (b + 3 * nb) + 1]};
    const Cx a11 = mc(s11, s11), a12 = mc(s12, s12), a21 = mc(s21, s21), a22 = mc(s22, s22);
    const Cx p1112 = mc(s11, s12), p2221 = mc(s22, s21), p1121 = mc(s11, s21), p2212 = mc(s22, s12);
    const Cx p1122 = mc(s11, s22), p1221 = mc(s12, s21), p2112 = mc(s21, s12);
    const Cx p2111 = mc(s21, s11), p2211 = mc(s22, s11);
    M[b + nb * 0] = 0.5 * (((a11.r + a12.r) + a21.r) + a22.r);
    M[b + nb * 1] = 0.5 * (((a11.r - a12.r) + a21.r) - a22.r);
    M[b + nb * 2] = p1112.r + p2221.r;
    M[b + nb * 3] = p1112.i - p2221.i;
    M[b + nb * 4] = 0.5 * (((a11.r + a12.r) - a21.r) - a22.r);
    M[b + nb * 5] = 0.5 * (((a11.r - a12.r) - a21.r) + a22.r);
    M[b + nb * 6] = p1112.r - p2221.r;
    M[b + nb * 7] = p1112.i + p2221.i;
    M[b + nb * 8] = p1121.r + p2212.r;
    M[b + nb * 9] = p1121.r - p2212.r;
    M[b + nb * 10] = p1122.r + p1221.r;
    M[b + nb * 11] = p1122.i + p2112.i;
    M[b + nb * 12] = p2111.i + p2212.i;
    M[b + nb * 13] = p2111.i - p2212.i;
    M[b + nb * 14] = p2211.i - p1221.i;
    M[b + nb * 15] = p2211.r - p1221.r;
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

// simpne (misc_submod.f90:2392-2531) on y[k*stride], k = 0..n-1
__device__ double simpne_dev(int n, const double* __restrict__ x, const double* __restrict__ y,
                             int64_t stride)
{
#pragma clang fp contract(off)
    double result = 0.0;
    if (n <= 2) return __builtin_nan("");
    int m = 1;
    for (;;) {
        const double x1 = x[m - 1], x2 = x[m], x3 = x[m + 1];
        const double e = x3 * x3 - x1 * x1;
        const double f = x3 * x3 * x3 - x1 * x1 * x1;
        const double fe = x3 - x1;
        const double del[3] = {x3 - x2, x1 - x3, x2 - x1};
        const double g[3] = {x2 + x3, x1 + x3, x1 + x2};
        const double p[3] = {x2 * x3, x1 * x3, x1 * x2};
        double sum1 = 0.0;
        for (int i = 0; i < 3; i++)
            sum1 = sum1 + y[(int64_t)(m - 1 + i) * stride] * del[i] * (f / 3.0 - g[i] * 0.5 * e + p[i] * fe);
        result = result - sum1 / (del[0] * del[1] * del[2]);
        m = m + 2;
        if (n <= m + 1) break;
    }
    if (n % 2 != 0) return result;
    m = n - 2;
    const double x3 = x[n - 1], x2 = x[n - 2], x1 = x[n - 3];
    const double e = x3 * x3 - x2 * x2;
    const double f = x3 * x3 * x3 - x2 * x2 * x2;
    const double fe = x3 - x2;
    const double del[3] = {x3 - x2, x1 - x3, x2 - x1};
    const double g[3] = {x2 + x3, x1 + x3, x1 + x2};
    const double p[3] = {x2 * x3, x1 * x3, x1 * x2};
    double sum1 = 0.0;
    for (int i = 0; i < 3; i++)
        sum1 = sum1 + y[(int64_t)(m - 1 + i) * stride] * del[i] * (f / 3.0 - g[i] * 0.5 * e + p[i] * fe);
    return result - sum1 / (del[0] * del[1] * del[2]);
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

// theta integrals (outputs_mod.f90:133-135, 198-205), one lane.
// params: scatt, scatt_beam, scatt_ext, g, g_beam, g_ext, back_scatt, f_beam, f_ext;
// tmp: 3*ntheta scratch.  pass 0 with scaling: computes the energy scaling
// factors f = sqrt(scatt_x / energy_x) into params[7..8].
__global__ void k_theta_params(const double* __restrict__ m0, const double* __restrict__ m1,
                               const double* __restrict__ m2, const double* __restrict__ M0,
                               const double* __restrict__ theta, int ntheta, int nphi,
                               double waveno, int scaling_pass, double ebeam, double eext,
                               double* __restrict__ tmp, double* __restrict__ params)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double w2 = waveno * waveno;
    const double* m[3] = {m0, m1, m2};
    double sc[3];
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < ntheta; j++) tmp[j] = m[k][j] * sin(theta[j]) / w2;
        sc[k] = simpne_dev(ntheta, theta, tmp, 1);
    }
    if (scaling_pass) {
        params[7] = sqrt(sc[1] / ebeam);
        params[8] = sqrt(sc[2] / eext);
        return;
    }
    double gg[3];
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < ntheta; j++) tmp[j] = m[k][j] * sin(theta[j]) * cos(theta[j]) / sc[k] / w2;
        gg[k] = simpne_dev(ntheta, theta, tmp, 1);
    }
    int ib = 0;
    double best = fabs(theta[0] - 2 * PI_R4);
    for (int j = 1; j < ntheta; j++) {
        const double d = fabs(theta[j] - 2 * PI_R4);
        if (d < best) { best = d; ib = j; }
    }
    params[0] = sc[0]; params[1] = sc[1]; params[2] = sc[2];
    params[3] = gg[0]; params[4] = gg[1]; params[5] = gg[2];
    params[6] = M0[(int64_t)nphi * ib] * 4 * PI_R4 / w2;
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
            hipLaunchKernelGGL(k_theta_params, dim3(1), dim3(64), 0, st, (const double*)o0,
                               (const double*)o1, (const double*)o2, (const double*)M0, g->theta,
                               nth, nph, g->waveno, 1, ebeam, eext, tmp, params);
    }
    hipLaunchKernelGGL(k_theta_params, dim3(1), dim3(64), 0, st, (const double*)o0, (const double*)o1,
                       (const double*)o2, (const double*)M0, g->theta, nth, nph, g->waveno, 0,
                       ebeam, eext, tmp, params);
    return (int)hipGetLastError();
}

extern "C" int abt_launch_accumulate(double* total, const double* x, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(abt::k_accumulate, dim3(nblk(n)), dim3(abt::MB), 0, st, total, x, n);
    return (int)hipGetLastError();
}
