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
    int nchunks, ntiles, nunits, fast_diff;
};

}  // namespace abt

// Kernel instance for (precision, beams per tile, bins per thread); nullptr if none.
extern "C" const void* abt_batch_kernel(int prec, int tb, int bpt);
// Counter reset, the persistent kernel on `wgs` workgroups, the chunk and tile
// reductions, and (mueller_1d non-null) the phi integral of the summed matrix.
extern "C" int abt_launch_batch(const abt::BatchArgs* a, int prec, int tb, int bpt, int wgs, int norient,
                                double* mueller, double* mueller_1d, double* integrals, hipStream_t st);
