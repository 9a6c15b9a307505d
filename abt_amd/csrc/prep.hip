// prep.hip -- per-beam frame set-up on the device for beams resident in HBM.
// One lane per beam runs the same make_beam_record as the host path (FP64,
// contraction off); only the libm (device vs glibc) differs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frame.h"
#include "records.h"

namespace abt {

__global__ __launch_bounds__(128) void k_prepare(int64_t n, const double* __restrict__ ampl,
                                                 const double* __restrict__ vk7,
                                                 const double* __restrict__ prop,
                                                 const double* __restrict__ fov, double fov_value,
                                                 const int64_t* __restrict__ vert_off,
                                                 const double* __restrict__ verts, double lambda,
                                                 double* __restrict__ recs,
                                                 double* __restrict__ vrecs)
{
    const int64_t b = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (b >= n) return;
    const int64_t o0 = vert_off[b], o1 = vert_off[b + 1];
    const double f = fov ? fov[b] : fov_value;
    double rec[REC];
    make_beam_record(ampl + 8 * b, vk7 + 3 * b, prop + 3 * b, f, o1 - o0, verts + 3 * o0, lambda,
                     o0, rec, (vrecs && (o1 - o0) > 3) ? vrecs + VREC * o0 : nullptr);
    double* dst = recs + b * REC;
    for (int q = 0; q < REC; q++) dst[q] = rec[q];
}

}  // namespace abt

extern "C" int abt_launch_prepare(int64_t n, const double* ampl, const double* vk7,
                                  const double* prop, const double* fov, double fov_value,
                                  const int64_t* vert_off, const double* verts, double lambda,
                                  double* recs, double* vrecs, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(abt::k_prepare, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, n, ampl,
                       vk7, prop, fov, fov_value, vert_off, verts, lambda, recs, vrecs);
    return (int)hipGetLastError();
}
