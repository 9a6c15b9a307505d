// adjoint.h -- launch arguments of the adjoint kernel (adjoint.hip), shared with its
// host plan (adjoint_plan.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "records.h"

namespace abt {

// One workgroup's work: up to 64 beams of one tree of one orientation, on one bin chunk.
struct AdjUnit {
    int64_t start;    // first beam (index into the tree's records)
    int32_t o;        // orientation
    int32_t tree;     // 0: beam tree, 1: ext tree
    int32_t count;    // beams in this unit (<= 64)
    int32_t chunk;    // bin chunk
};

struct AdjArgs {
    const double* recs_b;     // prepared records, beam trees of all orientations
    const double* recs_e;     // prepared records, ext trees
    const double* bins;       // [nbins][8]
    double* lam;              // [norient][nbins][8]  Lam' = k dL/dX T_phi^T
    const AdjUnit* units;
    double* part;             // [nchunks][nb_total][17] per-beam gradients (vertices 9, ampl 8)
    int64_t nb_beam, nb_total;
    int nbins, nchunks;
    double waveno, lambda;
};

}  // namespace abt

extern "C" int abt_launch_adj_bins(const abt::DevGrid* g, double* bins, hipStream_t st);
// Lam' from G = dL/dX (batched-forward amplitude layout), the adjoint kernel (prec 0: FP64,
// 1: mixed) over nunits units, and the chunk reduction into the four outputs (any may be null).
extern "C" int abt_launch_adjoint(const abt::AdjArgs* a, int prec, int nunits, const double* G,
                                  const abt::DevGrid* g, int norient, double* gvb, double* gab,
                                  double* gve, double* gae, hipStream_t st);
