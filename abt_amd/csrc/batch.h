// batch.h -- launch arguments of the orientation-batched far-field kernel
// (batch.hip), shared with its host plan (batch_plan.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "records.h"

namespace abt {

struct BatchArgs {
    DevGrid g;
    const double* recs_b;     // prepared beam-tree records, all orientations
    const double* vrecs_b;
    const double* recs_e;     // prepared ext-tree records
    const double* vrecs_e;
    const int64_t* boff;      // [norient+1] beam-tree record offsets
    const int64_t* eoff;      // [norient+1] ext-tree record offsets
    const int32_t* chunk_o;   // [nchunks+1] orientation range of each chunk
    const double* w1;         // [nbins] integral weights
    const double* w2;
    double* mpart;            // [nchunks][16][nbins] partial Mueller sums
    double* ipart;            // [norient][ntiles][6] partial integrals
    unsigned int* counter;    // work queue
    double* ampl_out;         // [norient][8][nbins] k-scaled total amplitude per orientation, or null
    double* store;            // [norient][16][nbins] k-scaled beam then ext amplitude per orientation
                              // (the -scaling path), or null
    // gated mixed precision (the error scale of the FP32 pass and the bins it leaves to FP64)
    double* epart;            // [nchunks][nbins] partial sums of S11_o * I_o (I_o: incoherent beam sum)
    const int32_t* list;      // [nbins] bins recomputed in FP64 (list mode), ascending
    const int32_t* list_count;// number of listed bins (device)
    int nchunks, ntiles, nunits, fast_diff;
};

// Kernel modes: plain, the mixed pass with its error scale, and the FP64 pass over listed bins.
enum : int { BMODE_PLAIN = 0, BMODE_EST = 1, BMODE_LIST = 2 };

}  // namespace abt

// Kernel instance for (precision, beams per tile, bins per thread, mode); nullptr if none.
extern "C" const void* abt_batch_kernel(int prec, int tb, int bpt, int mode);
// Counter reset, the persistent kernel on `wgs` workgroups, the chunk and tile
// reductions, and (mueller_1d non-null) the phi integral of the summed matrix.
extern "C" int abt_launch_batch(const abt::BatchArgs* a, int prec, int tb, int bpt, int wgs, int norient,
                                double* mueller, double* mueller_1d, double* integrals, hipStream_t st);
// One pass alone: counter reset and the persistent kernel (prec, tb, bpt, mode) on `wgs`
// workgroups, partial sums left in the argument block's buffers (the autotuner's timing).
extern "C" int abt_launch_batch_pass(const abt::BatchArgs* a, int prec, int tb, int bpt, int mode, int wgs,
                                     hipStream_t st);
// -scaling (outputs_mod.f90:146-184): the persistent pass storing every orientation's beam and
// ext amplitudes (a->store), the unscaled integrals reduced into raw[norient][6], the divisors
// sqrt(scatt_beam / beam_energy[o]) and sqrt(scatt_ext / ext_energy[o]) into div[norient][2],
// the Mueller sums and all six integrals recomputed from the divided amplitudes (one lane per
// bin, orientations of a chunk in order), then the chunk and tile reductions and the phi
// integral.  a->ampl_out (non-null) receives the divided totals.
extern "C" int abt_launch_batch_scaled(const abt::BatchArgs* a, int prec, int tb, int bpt, int wgs, int norient,
                                       const double* beam_energy, const double* ext_energy, double* raw,
                                       double* div, double* mueller, double* mueller_1d, double* integrals,
                                       hipStream_t st);
// Gated mixed precision: the mixed pass with its error scale (into est[nbins]), the bins whose
// error bound exceeds the gate listed (list, *count), those bins recomputed in FP64 by the
// restructured body (`wgs2` persistent workgroups of the 16 x 1 list kernel) and written over
// the mixed pass's sums, then the phi integral.  flags: [nbins] int32 scratch.
extern "C" int abt_launch_batch_gated(const abt::BatchArgs* a, int tb, int bpt, int wgs, int wgs2, int norient,
                                      double* est, int32_t* flags, int32_t* list, int32_t* count,
                                      double* mueller, double* mueller_1d, double* integrals, hipStream_t st);
