// adjoint_plan.cpp -- host side of the adjoint: plans (orientation offsets, work units,
// per-bin table, resident workspace) and the C-ABI entry points.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/abt_diff.h"
#include "adjoint.h"
#include "rt_internal.h"

namespace {
constexpr int LANES = 64;
constexpr int MAX_BIN_CHUNKS = 16;
}  // namespace

struct abt_adjoint_plan {
    const abt_dev_grid* g = nullptr;
    int norient = 0, prec = 0, nchunks = 1, nunits = 0;
    int64_t nb_beam = 0, nb_ext = 0;
    void* mem = nullptr;
    double* bins = nullptr;
    double* lam = nullptr;
    double* part = nullptr;
    abt::AdjUnit* units = nullptr;
};

extern "C" {

int abt_adjoint_plan_create(const abt_dev_grid* g, int32_t norient, const int64_t* beam_off,
                            const int64_t* ext_off, int precision, abt_adjoint_plan** out)
{
    if (!g || !beam_off || !ext_off || !out) return abt_fail(ABT_ERR_ARG, "abt_adjoint_plan_create: null argument");
    if (norient < 1) return abt_fail(ABT_ERR_ARG, "abt_adjoint_plan_create: need at least one orientation");
    if (precision != ABT_FP64 && precision != ABT_MIXED)
        return abt_fail(ABT_ERR_ARG, "unknown precision %d", precision);
    if (beam_off[0] != 0 || ext_off[0] != 0) return abt_fail(ABT_ERR_ARG, "offsets must start at 0");
    for (int o = 0; o < norient; o++)
        if (beam_off[o + 1] < beam_off[o] || ext_off[o + 1] < ext_off[o])
            return abt_fail(ABT_ERR_ARG, "offsets must be non-decreasing (orientation %d)", o);
    int cus = 0;
    int rc = abt_ensure_device(&cus);
    if (rc) return rc;
    abt_adjoint_plan* p = new abt_adjoint_plan();
    p->g = g;
    p->norient = norient;
    p->prec = precision;
    p->nb_beam = beam_off[norient];
    p->nb_ext = ext_off[norient];
    // units: blocks of 64 beams of one tree of one orientation; split the bins into
    // chunks while there are fewer than ~16 waves per SIMD over the launch (about 8 rounds at
    // the kernel's 2 waves per SIMD), so that the partly filled last round is a small share:
    // 64 orientations of the rough column (5427 blocks) run the backward in 171 ms unsplit,
    // 155 ms in 4 chunks (8 and 16 chunks: 155, 157 ms)
    int64_t nblocks = 0;
    for (int o = 0; o < norient; o++)
        nblocks += (beam_off[o + 1] - beam_off[o] + LANES - 1) / LANES + (ext_off[o + 1] - ext_off[o] + LANES - 1) / LANES;
    const int64_t want = (int64_t)cus * 4 * 16;
    int64_t nch = nblocks > 0 ? (want + nblocks - 1) / nblocks : 1;
    nch = std::max<int64_t>(1, std::min<int64_t>(nch, std::min<int64_t>(MAX_BIN_CHUNKS, std::max<int64_t>(1, g->nbins / 256))));
    p->nchunks = (int)nch;
    std::vector<abt::AdjUnit> units;
    units.reserve((size_t)(nblocks * nch));
    for (int o = 0; o < norient; o++)
        for (int tree = 0; tree < 2; tree++) {
            const int64_t* off = tree ? ext_off : beam_off;
            for (int64_t b = off[o]; b < off[o + 1]; b += LANES)
                for (int c = 0; c < p->nchunks; c++) {
                    abt::AdjUnit u;
                    u.start = b;
                    u.o = o;
                    u.tree = tree;
                    u.count = (int32_t)std::min<int64_t>(LANES, off[o + 1] - b);
                    u.chunk = c;
                    units.push_back(u);
                }
        }
    p->nunits = (int)units.size();
    const int64_t nbins = g->nbins, nbt = p->nb_beam + p->nb_ext;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t o_bins = take(sizeof(double) * 8 * nbins);
    const size_t o_lam = take(sizeof(double) * 8 * nbins * (size_t)norient);
    const size_t o_part = take(sizeof(double) * 17 * (size_t)nbt * p->nchunks);
    const size_t o_units = take(sizeof(abt::AdjUnit) * std::max<size_t>(1, units.size()));
    if (hipMalloc(&p->mem, off) != hipSuccess) {
        delete p;
        return abt_fail(ABT_ERR_NOMEM, "adjoint plan: hipMalloc of %zu bytes failed", off);
    }
    char* m = (char*)p->mem;
    p->bins = (double*)(m + o_bins);
    p->lam = (double*)(m + o_lam);
    p->part = (double*)(m + o_part);
    p->units = (abt::AdjUnit*)(m + o_units);
    hipError_t e = hipSuccess;
    if (!units.empty())
        e = hipMemcpy(p->units, units.data(), sizeof(abt::AdjUnit) * units.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        int le = abt_launch_adj_bins(&g->d, p->bins, nullptr);
        if (le) e = (hipError_t)le;
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(p->mem);
        delete p;
        return abt_fail(ABT_ERR_HIP, "adjoint plan set-up: %s", hipGetErrorString(e));
    }
    *out = p;
    return 0;
}

int abt_adjoint_plan_destroy(abt_adjoint_plan* p)
{
    if (!p) return 0;
    if (p->mem) (void)hipFree(p->mem);
    delete p;
    return 0;
}

int abt_adjoint_run(const abt_adjoint_plan* p, const double* dev_recs_beam, const double* dev_recs_ext,
                    const double* dev_dldx, double* dev_grad_verts_beam, double* dev_grad_ampl_beam,
                    double* dev_grad_verts_ext, double* dev_grad_ampl_ext, void* stream)
{
    if (!p || !dev_dldx) return abt_fail(ABT_ERR_ARG, "abt_adjoint_run: null argument");
    if (p->nb_beam > 0 && !dev_recs_beam) return abt_fail(ABT_ERR_ARG, "abt_adjoint_run: null beam records");
    if (p->nb_ext > 0 && !dev_recs_ext) return abt_fail(ABT_ERR_ARG, "abt_adjoint_run: null ext records");
    int rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    abt::AdjArgs a;
    a.recs_b = dev_recs_beam;
    a.recs_e = dev_recs_ext;
    a.bins = p->bins;
    a.lam = p->lam;
    a.units = p->units;
    a.part = p->part;
    a.nb_beam = p->nb_beam;
    a.nb_total = p->nb_beam + p->nb_ext;
    a.nbins = (int)p->g->nbins;
    a.nchunks = p->nchunks;
    a.waveno = p->g->d.waveno;
    a.lambda = p->g->d.lambda;
    int e = abt_launch_adjoint(&a, p->prec == ABT_MIXED ? 1 : 0, p->nunits, dev_dldx, &p->g->d, p->norient,
                               dev_grad_verts_beam, dev_grad_ampl_beam, dev_grad_verts_ext, dev_grad_ampl_ext,
                               (hipStream_t)stream);
    if (e) return abt_fail(ABT_ERR_HIP, "adjoint launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

}  // extern "C"
