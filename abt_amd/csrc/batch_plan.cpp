// batch_plan.cpp -- host side of the orientation-batched far field: plans (orientation
// offsets, chunking, integral weights, resident workspace), the tile-shape
// autotuner, the capture-safe launch sequence and its HIP graph.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/abt_diff.h"
#include "batch.h"
#include "rt_internal.h"
#include "simpne.h"

namespace {

constexpr int LANES = 256;
// Queue depth per persistent workgroup: bounds the tail (the last units finish at different
// times).  On 1000 orientations of the rough column: 16 units per workgroup (32 chunks)
// 3552 ms, 32 (64) 3503 ms, 64 (128) 3473 ms.  Partial sums: 128 B per bin per chunk.
constexpr int MAX_CHUNKS = 128;
constexpr int UNITS_PER_WG = 64;
const int TB_CAND[3] = {8, 16, 32};
const int BPT_CAND[2] = {1, 2};

struct Shape {
    int tb = 16, bpt = 1;
    int wgs = 0;        // persistent workgroups
    int ntiles = 0;
    int nchunks = 0;
};

// Contiguous orientation ranges with about equal beam counts (beam + ext trees).
void make_chunks(const int64_t* boff, const int64_t* eoff, int n, int nch, std::vector<int32_t>& out)
{
    out.assign(1, 0);
    const double total = (double)(boff[n] - boff[0] + eoff[n] - eoff[0]);
    int o = 0;
    for (int c = 1; c < nch; c++) {
        const double target = total * c / nch;
        while (o < n && (double)(boff[o] - boff[0] + eoff[o] - eoff[0]) < target) o++;
        o = std::max(o, out.back() + 1);
        o = std::min(o, n - (nch - c));
        out.push_back(o);
    }
    out.push_back(n);
}

// ABI precision -> kernel variant: ABT_FP64_ACCURATE runs the accurate-phase FP64 body (3),
// ABT_FP64_FAST the restructured FP64 body (4); ABT_MIXED_GATED's tuned pass is the mixed
// body with its error scale
int kernel_prec(int precision)
{
    switch (precision) {
        case ABT_FP64_ACCURATE: return 3;
        case ABT_FP64_FAST: return 4;
        case ABT_MIXED_GATED: return 1;
        default: return precision;
    }
}
int kernel_mode(int precision) { return precision == ABT_MIXED_GATED ? abt::BMODE_EST : abt::BMODE_PLAIN; }

bool shape_allowed(int prec, int tb, int bpt)
{
    if (prec == ABT_FP64_ACCURATE) return tb == 16 && bpt == 1;
    if (prec == ABT_FP64_FAST) return bpt == 1 || tb <= 16;
    return true;
}

int kernel_occupancy(const void* k, int* per_cu)
{
    int n = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, LANES, 0));
    *per_cu = std::max(1, n);
    return 0;
}

int occupancy(int prec, int tb, int bpt, int* per_cu)
{
    const void* k = abt_batch_kernel(kernel_prec(prec), tb, bpt, kernel_mode(prec));
    if (!k) return abt_fail(ABT_ERR_ARG, "no batched kernel for beams_per_tile %d, bins_per_thread %d", tb, bpt);
    return kernel_occupancy(k, per_cu);
}

}  // namespace

struct abt_batch_plan {
    const abt_dev_grid* g = nullptr;
    int norient = 0, prec = 0, fast_diff = 0, cus = 256;
    std::vector<int64_t> boff, eoff;
    Shape shape;
    bool tuned = false;
    // device memory: one allocation
    void* mem = nullptr;
    int64_t* d_boff = nullptr;
    int64_t* d_eoff = nullptr;
    int32_t* d_chunk = nullptr;      // [MAX_CHUNKS + 1] chunk table of the plan's shape
    int32_t* d_chunk_tune = nullptr; // [MAX_CHUNKS + 1] chunk table of the autotuner's sample
    double* d_w1 = nullptr;
    double* d_w2 = nullptr;
    double* d_mpart = nullptr;       // [max_chunks][16][nbins]
    double* d_ipart = nullptr;       // [norient][ntiles(bpt=1)][6]
    unsigned int* d_counter = nullptr;
    // gated mixed precision: error-scale partials and sums, flags, the list of FP64 bins
    double* d_epart = nullptr;       // [max_chunks][nbins]
    double* d_est = nullptr;         // [nbins]
    int32_t* d_flags = nullptr;      // [nbins]
    int32_t* d_list = nullptr;       // [nbins]
    int32_t* d_count = nullptr;
    int wgs_list = 0;                // persistent workgroups of the FP64 list pass
    // -scaling: per-orientation amplitudes, unscaled integrals, divisors; the caller's energies
    void* smem = nullptr;
    double* d_store = nullptr;       // [norient][16][nbins]
    double* d_raw = nullptr;         // [norient][6]
    double* d_div = nullptr;         // [norient][2]
    const double* beam_energy = nullptr;
    const double* ext_energy = nullptr;
    int max_chunks = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int shape_for(const abt_batch_plan* p, int norient, int tb, int bpt, Shape* s)
{
    int per_cu = 1;
    int rc = occupancy(p->prec, tb, bpt, &per_cu);
    if (rc) return rc;
    s->tb = tb;
    s->bpt = bpt;
    s->wgs = per_cu * p->cus;
    s->ntiles = (int)((p->g->nbins + (int64_t)LANES * bpt - 1) / ((int64_t)LANES * bpt));
    int64_t want = (int64_t)UNITS_PER_WG * s->wgs;
    int64_t nch = (want + s->ntiles - 1) / s->ntiles;
    nch = std::min<int64_t>(nch, std::min(norient, MAX_CHUNKS));
    s->nchunks = (int)std::max<int64_t>(1, nch);
    return 0;
}

// Launch arguments of the plan's buffers with chunk table `chunk` and shape s.
void fill_args(const abt_batch_plan* p, const Shape& s, const int32_t* chunk, const double* rb,
               const double* vb, const double* re, const double* ve, abt::BatchArgs* a)
{
    a->g = p->g->d;
    a->recs_b = rb;
    a->vrecs_b = vb ? vb : rb;
    a->recs_e = re;
    a->vrecs_e = ve ? ve : re;
    a->boff = p->d_boff;
    a->eoff = p->d_eoff;
    a->chunk_o = chunk;
    a->w1 = p->d_w1;
    a->w2 = p->d_w2;
    a->mpart = p->d_mpart;
    a->ipart = p->d_ipart;
    a->counter = p->d_counter;
    a->ampl_out = nullptr;
    a->store = p->d_store;
    a->epart = p->d_epart;
    a->list = p->d_list;
    a->list_count = p->d_count;
    a->nchunks = s.nchunks;
    a->ntiles = s.ntiles;
    a->nunits = s.nchunks * s.ntiles;
    a->fast_diff = p->fast_diff;
}

// Upload the chunk table of the first `norient` orientations for `nchunks` chunks.
int upload_chunks(const abt_batch_plan* p, int norient, int nchunks, int32_t* dst)
{
    std::vector<int32_t> ch;
    make_chunks(p->boff.data(), p->eoff.data(), norient, nchunks, ch);
    HIPCHK(hipMemcpy(dst, ch.data(), sizeof(int32_t) * ch.size(), hipMemcpyHostToDevice));
    return 0;
}

int sync_chunks(abt_batch_plan* p) { return upload_chunks(p, p->norient, p->shape.nchunks, p->d_chunk); }

int check_recs(const double* rb, const double* re, const abt_batch_plan* p)
{
    if (p->boff[p->norient] > 0 && !rb) return abt_fail(ABT_ERR_ARG, "batched far field: null beam records");
    if (p->eoff[p->norient] > 0 && !re) return abt_fail(ABT_ERR_ARG, "batched far field: null ext records");
    return 0;
}

void drop_graph(abt_batch_plan* p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

}  // namespace

extern "C" {

int abt_batch_plan_create(const abt_dev_grid* g, int32_t norient, const int64_t* beam_off,
                          const int64_t* ext_off, int precision, int is_fast_diff, abt_batch_plan** out)
{
    if (!g || !beam_off || !ext_off || !out) return abt_fail(ABT_ERR_ARG, "abt_batch_plan_create: null argument");
    if (norient < 1) return abt_fail(ABT_ERR_ARG, "abt_batch_plan_create: need at least one orientation");
    if (precision != ABT_FP64 && precision != ABT_MIXED && precision != ABT_FP64_ACCURATE &&
        precision != ABT_FP64_FAST && precision != ABT_MIXED_GATED)
        return abt_fail(ABT_ERR_ARG, "unknown precision %d", precision);
    if (g->d.nphi < 3 || g->d.ntheta < 3)
        return abt_fail(ABT_ERR_ARG, "the phi and theta integrals (simpne) need at least 3 nodes each");
    if (beam_off[0] != 0 || ext_off[0] != 0) return abt_fail(ABT_ERR_ARG, "offsets must start at 0");
    for (int o = 0; o < norient; o++)
        if (beam_off[o + 1] < beam_off[o] || ext_off[o + 1] < ext_off[o])
            return abt_fail(ABT_ERR_ARG, "offsets must be non-decreasing (orientation %d)", o);
    int cus = 0;
    int rc = abt_ensure_device(&cus);
    if (rc) return rc;
    abt_batch_plan* p = new abt_batch_plan();
    p->g = g;
    p->norient = norient;
    p->prec = precision;
    p->fast_diff = is_fast_diff ? 1 : 0;
    p->cus = cus;
    p->boff.assign(beam_off, beam_off + norient + 1);
    p->eoff.assign(ext_off, ext_off + norient + 1);
    // largest chunk count over the candidate shapes sizes the partial sums
    p->max_chunks = 1;
    for (int tb : TB_CAND)
        for (int bpt : BPT_CAND) {
            if (!shape_allowed(precision, tb, bpt)) continue;
            Shape s;
            rc = shape_for(p, norient, tb, bpt, &s);
            if (rc) { delete p; return rc; }
            p->max_chunks = std::max(p->max_chunks, s.nchunks);
        }
    rc = shape_for(p, norient, 16, 1, &p->shape);
    if (rc) { delete p; return rc; }
    const bool gated = precision == ABT_MIXED_GATED;
    if (gated) {
        int per_cu = 1;
        rc = kernel_occupancy(abt_batch_kernel(4, 16, 1, abt::BMODE_LIST), &per_cu);
        if (rc) { delete p; return rc; }
        p->wgs_list = per_cu * cus;
    }
    const int64_t nbins = g->nbins, ntiles1 = (nbins + LANES - 1) / LANES;
    const int nth = g->d.ntheta, nph = g->d.nphi;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t o_boff = take(sizeof(int64_t) * (norient + 1));
    const size_t o_eoff = take(sizeof(int64_t) * (norient + 1));
    const size_t o_chunk = take(sizeof(int32_t) * (MAX_CHUNKS + 1));
    const size_t o_chunk_t = take(sizeof(int32_t) * (MAX_CHUNKS + 1));
    const size_t o_w1 = take(sizeof(double) * nbins);
    const size_t o_w2 = take(sizeof(double) * nbins);
    const size_t o_mp = take(sizeof(double) * 16 * nbins * (size_t)p->max_chunks);
    const size_t o_ip = take(sizeof(double) * 6 * (size_t)ntiles1 * norient);
    const size_t o_ct = take(sizeof(unsigned int) * 64);
    const size_t o_ep = gated ? take(sizeof(double) * nbins * (size_t)p->max_chunks) : 0;
    const size_t o_est = gated ? take(sizeof(double) * nbins) : 0;
    const size_t o_fl = gated ? take(sizeof(int32_t) * nbins) : 0;
    const size_t o_li = gated ? take(sizeof(int32_t) * nbins) : 0;
    const size_t o_cn = gated ? take(sizeof(int32_t) * 64) : 0;
    if (hipMalloc(&p->mem, off) != hipSuccess) {
        delete p;
        return abt_fail(ABT_ERR_NOMEM, "batched plan: hipMalloc of %zu bytes failed", off);
    }
    char* m = (char*)p->mem;
    p->d_boff = (int64_t*)(m + o_boff);
    p->d_eoff = (int64_t*)(m + o_eoff);
    p->d_chunk = (int32_t*)(m + o_chunk);
    p->d_chunk_tune = (int32_t*)(m + o_chunk_t);
    p->d_w1 = (double*)(m + o_w1);
    p->d_w2 = (double*)(m + o_w2);
    p->d_mpart = (double*)(m + o_mp);
    p->d_ipart = (double*)(m + o_ip);
    p->d_counter = (unsigned int*)(m + o_ct);
    if (gated) {
        p->d_epart = (double*)(m + o_ep);
        p->d_est = (double*)(m + o_est);
        p->d_flags = (int32_t*)(m + o_fl);
        p->d_list = (int32_t*)(m + o_li);
        p->d_count = (int32_t*)(m + o_cn);
    }
    // integral weights: simpne over phi, then x sin(theta) / k^2 (x cos(theta)) and simpne over theta
    std::vector<double> th(nth), ph(nph), wt(nth), wp(nph), w1(nbins), w2(nbins);
    hipError_t e = hipMemcpy(th.data(), g->d.theta, sizeof(double) * nth, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ph.data(), g->d.phi, sizeof(double) * nph, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        abt::simpne_weights(nth, th.data(), wt.data());
        abt::simpne_weights(nph, ph.data(), wp.data());
        const double k2 = g->d.waveno * g->d.waveno;
        for (int j = 0; j < nth; j++) {
            const double st = sin(th[j]) / k2, ct = cos(th[j]);
            for (int i = 0; i < nph; i++) {
                const double w = wt[j] * st * wp[i];
                w1[i + (int64_t)nph * j] = w;
                w2[i + (int64_t)nph * j] = w * ct;
            }
        }
        e = hipMemcpy(p->d_w1, w1.data(), sizeof(double) * nbins, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(p->d_w2, w2.data(), sizeof(double) * nbins, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_boff, beam_off, sizeof(int64_t) * (norient + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_eoff, ext_off, sizeof(int64_t) * (norient + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->mem);
        delete p;
        return abt_fail(ABT_ERR_HIP, "batched plan upload: %s", hipGetErrorString(e));
    }
    rc = sync_chunks(p);
    if (rc) {
        (void)hipFree(p->mem);
        delete p;
        return rc;
    }
    *out = p;
    return 0;
}

int abt_batch_plan_destroy(abt_batch_plan* p)
{
    if (!p) return 0;
    drop_graph(p);
    if (p->mem) (void)hipFree(p->mem);
    if (p->smem) (void)hipFree(p->smem);
    delete p;
    return 0;
}

int abt_batch_set_scaling(abt_batch_plan* p, const double* beam_energy_out, const double* ext_energy_out)
{
    if (!p) return abt_fail(ABT_ERR_ARG, "abt_batch_set_scaling: null plan");
    if (!beam_energy_out != !ext_energy_out)
        return abt_fail(ABT_ERR_ARG, "abt_batch_set_scaling: give both energy arrays, or neither");
    if (beam_energy_out && p->prec == ABT_MIXED_GATED)
        return abt_fail(ABT_ERR_ARG, "abt_batch_set_scaling: not available for the gated mixed plan");
    int rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    drop_graph(p);   // a captured graph holds the other launch sequence
    if (!beam_energy_out) {
        if (p->smem) (void)hipFree(p->smem);
        p->smem = nullptr;
        p->d_store = p->d_raw = p->d_div = nullptr;
        p->beam_energy = p->ext_energy = nullptr;
        return 0;
    }
    if (!p->smem) {
        const size_t nstore = (size_t)16 * p->g->nbins * p->norient;
        const size_t bytes = sizeof(double) * (nstore + 8 * (size_t)p->norient);
        void* m = nullptr;
        hipError_t e = hipMalloc(&m, bytes);
        if (e != hipSuccess)
            return abt_fail(ABT_ERR_HIP, "abt_batch_set_scaling: %zu bytes: %s", bytes, hipGetErrorString(e));
        p->smem = m;
        p->d_store = (double*)m;
        p->d_raw = p->d_store + nstore;
        p->d_div = p->d_raw + 6 * (size_t)p->norient;
    }
    p->beam_energy = beam_energy_out;
    p->ext_energy = ext_energy_out;
    return 0;
}

int abt_batch_set_tiles(abt_batch_plan* p, int32_t beams_per_tile, int32_t bins_per_thread)
{
    if (!p) return abt_fail(ABT_ERR_ARG, "abt_batch_set_tiles: null plan");
    Shape s;
    int rc = shape_for(p, p->norient, beams_per_tile, bins_per_thread, &s);
    if (rc) return rc;
    drop_graph(p);   // a captured graph holds the old shape
    p->shape = s;
    p->tuned = true;
    return sync_chunks(p);
}

int abt_batch_get_tiles(const abt_batch_plan* p, int32_t* beams_per_tile, int32_t* bins_per_thread,
                        int32_t* chunks, int32_t* workgroups)
{
    if (!p) return abt_fail(ABT_ERR_ARG, "abt_batch_get_tiles: null plan");
    if (beams_per_tile) *beams_per_tile = p->shape.tb;
    if (bins_per_thread) *bins_per_thread = p->shape.bpt;
    if (chunks) *chunks = p->shape.nchunks;
    if (workgroups) *workgroups = p->shape.wgs;
    return 0;
}

int abt_batch_autotune(abt_batch_plan* p, const double* recs_b, const double* vrecs_b,
                       const double* recs_e, const double* vrecs_e, int32_t sample_orient,
                       void* stream, double* ms_out)
{
    if (!p) return abt_fail(ABT_ERR_ARG, "abt_batch_autotune: null plan");
    int rc = check_recs(recs_b, recs_e, p);
    if (rc) return rc;
    rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int ns = std::max(1, std::min(p->norient, sample_orient > 0 ? (int)sample_orient : 32));
    const int64_t nbins = p->g->nbins;
    double* out = nullptr;
    const size_t ob = sizeof(double) * (16 * (size_t)nbins + 6 * (size_t)ns);
    HIPCHK(hipMalloc(&out, ob));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    float best = 1e30f;
    Shape best_s = p->shape;
    int idx = 0;
    rc = 0;
    for (int bpt : BPT_CAND) {
        for (int tb : TB_CAND) {
            if (rc || he != hipSuccess) break;
            if (!shape_allowed(p->prec, tb, bpt)) {
                if (ms_out) ms_out[idx] = 0.0;
                idx++;
                continue;
            }
            Shape s;
            rc = shape_for(p, ns, tb, bpt, &s);
            if (rc) break;
            rc = upload_chunks(p, ns, s.nchunks, p->d_chunk_tune);
            if (rc) break;
            abt::BatchArgs a;
            fill_args(p, s, p->d_chunk_tune, recs_b, vrecs_b, recs_e, vrecs_e, &a);
            a.store = nullptr;   // the timing is the far-field pass's
            float ms = 0.f;
            for (int rep = 0; rep < 2 && !rc; rep++) {   // first launch warms up
                he = hipEventRecord(e0, st);
                // gated plans tune their mixed pass (the plain mixed kernel of the same shape
                // plus the error scale: run the pass alone)
                int le = kernel_mode(p->prec) == abt::BMODE_EST
                             ? abt_launch_batch_pass(&a, kernel_prec(p->prec), tb, bpt, kernel_mode(p->prec), s.wgs, st)
                             : abt_launch_batch(&a, kernel_prec(p->prec), tb, bpt, s.wgs, ns, out, nullptr,
                                                out + 16 * nbins, st);
                if (le) rc = abt_fail(ABT_ERR_HIP, "batched launch failed: %s", hipGetErrorString((hipError_t)le));
                if (he == hipSuccess) he = hipEventRecord(e1, st);
                if (he == hipSuccess) he = hipEventSynchronize(e1);
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
            }
            if (ms_out) ms_out[idx] = ms;
            idx++;
            if (!rc && he == hipSuccess && ms < best) {
                best = ms;
                best_s = s;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc) return rc;
    if (he != hipSuccess) return abt_fail(ABT_ERR_HIP, "autotune timing: %s", hipGetErrorString(he));
    Shape s;
    rc = shape_for(p, p->norient, best_s.tb, best_s.bpt, &s);
    if (rc) return rc;
    drop_graph(p);
    p->shape = s;
    p->tuned = true;
    return sync_chunks(p);
}

int abt_batch_run(const abt_batch_plan* p, const double* recs_b, const double* vrecs_b,
                  const double* recs_e, const double* vrecs_e, double* dev_mueller,
                  double* dev_mueller_1d, double* dev_integrals, double* dev_ampl, void* stream)
{
    if (!p || !dev_mueller || !dev_integrals) return abt_fail(ABT_ERR_ARG, "abt_batch_run: null argument");
    int rc = check_recs(recs_b, recs_e, p);
    if (rc) return rc;
    rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    abt::BatchArgs a;
    fill_args(p, p->shape, p->d_chunk, recs_b, vrecs_b, recs_e, vrecs_e, &a);
    a.ampl_out = dev_ampl;
    int e;
    if (p->beam_energy)
        e = abt_launch_batch_scaled(&a, kernel_prec(p->prec), p->shape.tb, p->shape.bpt, p->shape.wgs, p->norient,
                                    p->beam_energy, p->ext_energy, p->d_raw, p->d_div, dev_mueller,
                                    dev_mueller_1d, dev_integrals, (hipStream_t)stream);
    else if (p->prec == ABT_MIXED_GATED)
        e = abt_launch_batch_gated(&a, p->shape.tb, p->shape.bpt, p->shape.wgs, p->wgs_list, p->norient,
                                   p->d_est, p->d_flags, p->d_list, p->d_count, dev_mueller, dev_mueller_1d,
                                   dev_integrals, (hipStream_t)stream);
    else
        e = abt_launch_batch(&a, kernel_prec(p->prec), p->shape.tb, p->shape.bpt, p->shape.wgs, p->norient,
                             dev_mueller, dev_mueller_1d, dev_integrals, (hipStream_t)stream);
    if (e) return abt_fail(ABT_ERR_HIP, "batched launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

int abt_batch_capture(abt_batch_plan* p, const double* recs_b, const double* vrecs_b,
                      const double* recs_e, const double* vrecs_e, double* dev_mueller,
                      double* dev_mueller_1d, double* dev_integrals, double* dev_ampl, void* stream)
{
    if (!p) return abt_fail(ABT_ERR_ARG, "abt_batch_capture: null plan");
    if (!stream) return abt_fail(ABT_ERR_ARG, "abt_batch_capture: capture needs a non-default stream");
    int rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    drop_graph(p);
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = abt_batch_run(p, recs_b, vrecs_b, recs_e, vrecs_e, dev_mueller, dev_mueller_1d, dev_integrals,
                       dev_ampl, stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return abt_fail(ABT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return abt_fail(ABT_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    p->graph = graph;
    p->exec = exec;
    return 0;
}

int abt_batch_launch(const abt_batch_plan* p, void* stream)
{
    if (!p || !p->exec) return abt_fail(ABT_ERR_ARG, "abt_batch_launch: no captured graph (call abt_batch_capture)");
    HIPCHK(hipGraphLaunch(p->exec, (hipStream_t)stream));
    return 0;
}

int abt_batch_workspace(const abt_batch_plan* p, size_t* bytes)
{
    if (!p || !bytes) return abt_fail(ABT_ERR_ARG, "abt_batch_workspace: null argument");
    const int64_t nbins = p->g->nbins, ntiles1 = (nbins + LANES - 1) / LANES;
    *bytes = sizeof(double) * (16 * (size_t)nbins * p->max_chunks + 6 * (size_t)ntiles1 * p->norient +
                               2 * (size_t)nbins);
    if (p->prec == ABT_MIXED_GATED)
        *bytes += sizeof(double) * ((size_t)nbins * p->max_chunks + nbins) + sizeof(int32_t) * 2 * (size_t)nbins;
    if (p->smem) *bytes += sizeof(double) * ((size_t)16 * nbins * p->norient + 8 * (size_t)p->norient);
    return 0;
}

int abt_batch_gate_stats(const abt_batch_plan* p, int32_t* refined_bins, void* stream)
{
    if (!p || !refined_bins) return abt_fail(ABT_ERR_ARG, "abt_batch_gate_stats: null argument");
    if (p->prec != ABT_MIXED_GATED) {
        *refined_bins = 0;
        return 0;
    }
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(refined_bins, p->d_count, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
