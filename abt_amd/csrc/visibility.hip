// visibility.hip -- the facets a propagating beam illuminates (the near-field
// tracer's visibility step), bit-identical to find_vis_int / find_vis_ext /
// find_vis_inc (beam_loop_mod.f90:2037-2609).
//
// The reference, per beam: 16 boxes on a 4 x 4 grid over the x-y extent of
// the vertices (beam_aligned_bounding_boxes, :2668-2743); each facet's box is
// the one whose midpoint is nearest its midpoint (f3), and a facet is a
// candidate blocker in every box nearest to one of its first three vertices
// (f4, f5).  A tested facet m then scans the candidates j of its box in facet
// order: j must lie in front (mid_z(m) <= mid_z(j)) and cover m's midpoint in
// projection (edge checks); the first non-beam blocker of another aperture
// ends the scan, beam facets record the nearest one (id_beam, dist_beam).
//
// Here, for many beams at once (beam = blockIdx.y):
//   k_vis_boxes   one workgroup per beam: x-y extents (block min/max), box
//                 midpoints, the illuminating facets' extent;
//   k_vis_facets  one lane per facet: box, candidate mask, cull and shadow
//                 flags, initial outputs;
//   k_vis_map     field_in -> mapping (last position wins, as the reference's
//                 sequential store);
//   k_vis_lists   one workgroup per beam: per box, the ascending candidate
//                 list and the ascending list of tested facets (block-wide
//                 ballot scans, so list order is facet order);
//   k_vis_main    one wave per 64 tested facets of one box: every lane of the
//                 wave scans the same candidate list in facet order; tiles of
//                 candidates (depth, aperture, projected vertices) are gathered
//                 into LDS by parallel loads and read back as broadcasts; a lane
//                 stops at its blocker, the wave when all lanes have stopped.
// All arithmetic is the reference's FP64 operations with contraction off and
// comparisons in the same sense; min/max reductions are exact, so results do
// not depend on the decomposition.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/abt_vis.h"
#include "rt_internal.h"

namespace abt {
namespace vis {

constexpr int NB = 16;                                // 4 x 4 boxes
constexpr int BLK = 256;
constexpr int WAVE = 64;
constexpr int OFFS = 2 * (NB + 1) + 2;                // per-beam list offsets (padded)
constexpr int TC = 32;                                // candidates per LDS tile of the scan
constexpr int UNR = 4;                                // candidates evaluated together in the scan
constexpr int NVS = 8;                                // vertices per candidate staged in LDS
constexpr int SLOT = 4 + 2 * NVS;                     // doubles per staged candidate
constexpr int64_t BEAM_CHUNK = 65535;                 // beams per launch (grid.y)
constexpr double FAC = 1.100000023841858;             // fac = 1.1, a REAL*4 literal (:2679)
constexpr double C001 = 0.009999999776482582;         // 0.01, a REAL*4 literal

struct BeamBox {
    double bx[NB], by[NB];
    double xmin, xmax, ymin, ymax;   // x-y extent of the illuminating facets (int / ext)
    double pad[4];
};

struct Work {
    BeamBox* boxes;     // [nbeam]
    int32_t* fbox;      // [nbeam][nf] box of a tested facet's midpoint, -1: not tested
    int32_t* fmask;     // [nbeam][nf] bit k: candidate blocker in box k
    int32_t* mapping;   // [nbeam][nf] 1-based position in field_in (the last), 0: not a beam facet
    int32_t* offs;      // [nbeam][OFFS] candidate-list offsets [0..NB], tested-list offsets [NB+1..2NB+1]
    int2* cand;         // [nbeam][3 nf] (blocker facet, f_off of it)
    int32_t* mlist;     // [nbeam][nf]
    int64_t nvref;      // f_off[nf]
    int64_t b0;         // first beam of this launch
};

__device__ __forceinline__ int64_t clampi(int64_t i, int64_t n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// minloc over the box midpoints of sqrt((bx - px)**2 + (by - py)**2), first minimum
// (beam_loop_mod.f90:2087-2093)
// The square root is monotonic, so the smallest distance is the root of the smallest
// squared distance d2[k0] (k0 its first index); an earlier box ties with it only if its
// squared distance rounds to the same root, which needs d2[k] <= d2[k0] (1 + 2^-51): those
// few are decided by the roots themselves, as the reference does.  bx, by: the box
// midpoints in registers; gbx, gby: the same in memory.
__device__ __forceinline__ int nearest_box(const double* bx, const double* by, const double* gbx,
                                           const double* gby, double px, double py)
{
#pragma clang fp contract(off)
    double d2[NB];
    double dmin = 0.0;
    int k0 = 0;
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const double dx = bx[k] - px, dy = by[k] - py;
        d2[k] = dx * dx + dy * dy;
        if (k == 0 || d2[k] < dmin) {
            dmin = d2[k];
            k0 = k;
        }
    }
    const double lim = dmin + ldexp(dmin, -49);
    bool near = false;
#pragma unroll
    for (int k = 0; k < NB; k++) near |= k < k0 && d2[k] <= lim;
    if (!near) return k0;
    // rare: recompute from the boxes in memory (a rolled loop, no register indexing)
    const double s0 = sqrt(dmin);
    for (int k = 0; k < k0; k++) {
        const double dx = gbx[k] - px, dy = gby[k] - py;
        const double d = dx * dx + dy * dy;
        if (d <= lim && sqrt(d) == s0) return k;
    }
    return k0;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_vis_boxes(abt_vis_batch a, Work w)
{
#pragma clang fp contract(off)
    __shared__ double red[BLK / WAVE][8];
    const int64_t b = w.b0 + blockIdx.x;
    const int tid = threadIdx.x;
    const double* x = a.v + 3 * a.nv * b;
    const double* y = x + a.nv;
    double r[8];
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
        r[q] = INFINITY;
        r[q + 1] = -INFINITY;
    }
    for (int64_t i = tid; i < a.nv; i += BLK) {
        const double xi = x[i], yi = y[i];
        r[0] = xi < r[0] ? xi : r[0];
        r[1] = xi > r[1] ? xi : r[1];
        r[2] = yi < r[2] ? yi : r[2];
        r[3] = yi > r[3] ? yi : r[3];
    }
    if (a.mode != ABT_VIS_INC) {
        const int64_t e0 = a.in_off[b], e1 = a.in_off[b + 1];
        for (int64_t e = e0 + tid; e < e1; e += BLK) {
            const int64_t fi = clampi(a.field_in[e] - 1, a.nf);
            for (int64_t k = a.f_off[fi]; k < a.f_off[fi + 1]; k++) {
                const int64_t vk = clampi(a.f_vi[k] - 1, a.nv);
                const double xk = x[vk], yk = y[vk];
                r[4] = xk < r[4] ? xk : r[4];
                r[5] = xk > r[5] ? xk : r[5];
                r[6] = yk < r[6] ? yk : r[6];
                r[7] = yk > r[7] ? yk : r[7];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 8; q++) r[q] = (q & 1) ? wave_max(r[q]) : wave_min(r[q]);
    if ((tid & (WAVE - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) red[tid / WAVE][q] = r[q];
    }
    __syncthreads();
    if (tid != 0) return;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        double v = red[0][q];
        for (int s = 1; s < BLK / WAVE; s++) {
            const double u = red[s][q];
            v = (q & 1) ? (u > v ? u : v) : (u < v ? u : v);
        }
        r[q] = v;
    }
    // box corners fac * (min + (max - min) / 4 * (i - 1)) (:2711-2720), box k = 4 (i-1) + (j-1)
    // with corners (i, j), (i, j+1), (i+1, j+1), (i+1, j) summed in that order, / 4 (:2732-2740,
    // compute_geometry_midpoints misc_submod.f90:199-209)
    const double min_x = r[0], max_x = r[1], min_y = r[2], max_y = r[3];
    BeamBox* B = w.boxes + b;
    for (int i = 1; i <= 4; i++) {
        const double x0 = FAC * (min_x + (max_x - min_x) / 4.0 * (double)(i - 1));
        const double x1 = FAC * (min_x + (max_x - min_x) / 4.0 * (double)i);
        for (int j = 1; j <= 4; j++) {
            const double y0 = FAC * (min_y + (max_y - min_y) / 4.0 * (double)(j - 1));
            const double y1 = FAC * (min_y + (max_y - min_y) / 4.0 * (double)j);
            const int k = 4 * (i - 1) + (j - 1);
            B->bx[k] = (((0.0 + x0) + x0) + x1) + x1;
            B->bx[k] = B->bx[k] / 4.0;
            B->by[k] = (((0.0 + y0) + y1) + y1) + y0;
            B->by[k] = B->by[k] / 4.0;
        }
    }
    B->xmin = r[4];
    B->xmax = r[5];
    B->ymin = r[6];
    B->ymax = r[7];
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_vis_facets(abt_vis_batch a, Work w)
{
#pragma clang fp contract(off)
    const int64_t b = w.b0 + blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= a.nf) return;
    const int64_t nf = a.nf;
    const BeamBox& B = w.boxes[b];
    double bx[NB], by[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        bx[k] = B.bx[k];
        by[k] = B.by[k];
    }
    const double* x = a.v + 3 * a.nv * b;
    const double* y = x + a.nv;
    const double* mid = a.f_mid + 3 * nf * b + 3 * i;
    const double mx = mid[0], my = mid[1];
    const int f3 = nearest_box(bx, by, B.bx, B.by, mx, my);
    const int64_t o0 = a.f_off[i];
    int mask = 0;
    for (int q = 0; q < 3; q++) {
        const int64_t vq = clampi(a.f_vi[min(o0 + q, w.nvref - 1)] - 1, a.nv);
        mask |= 1 << nearest_box(bx, by, B.bx, B.by, x[vq], y[vq]);
    }
    const double nz = a.f_nz[nf * b + i];
    const int64_t api = a.f_ap[i];
    const double apn = a.ap_nz[a.na * b + clampi(api - 1, a.na)];
    bool vis = true;
    if (a.mode != ABT_VIS_INC) {   // midpoint outside the beam's extent (:2145-2155)
        if (mx < B.xmin) vis = false;
        else if (mx > B.xmax) vis = false;
        else if (my < B.ymin) vis = false;
        else if (my > B.ymax) vis = false;
    }
    const bool shad = a.mode == ABT_VIS_INT ? nz >= 0 : nz <= 0;                 // :2158, :2370, :2543
    const bool down = a.mode == ABT_VIS_INT ? apn > -C001 : apn < C001;          // :2165, :2377, :2547
    const bool tested = vis && !down;
    const int64_t o = nf * b + i;
    a.in_beam[o] = (a.mode == ABT_VIS_INC && !down) ? 1 : 0;
    if (a.is_vis) a.is_vis[o] = tested ? 1 : 0;
    a.dist_beam[o] = 0.0;
    a.id_beam[o] = 0;
    a.is_shad[o] = shad ? 1 : 0;
    w.fbox[o] = tested ? f3 : -1;
    // the incident beam's blockers face up (:2553)
    w.fmask[o] = (a.mode == ABT_VIS_INC && nz < C001) ? 0 : mask;
    w.mapping[o] = 0;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_vis_map(abt_vis_batch a, Work w)
{
    const int64_t b = w.b0 + blockIdx.x;
    const int64_t e0 = a.in_off[b], e1 = a.in_off[b + 1];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += BLK) {
        const int64_t fi = a.field_in[e] - 1;
        if (fi < 0 || fi >= a.nf) continue;
        atomicMax(w.mapping + a.nf * b + fi, (int)(e - e0 + 1));
    }
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__global__ __launch_bounds__(BLK) void k_vis_lists(abt_vis_batch a, Work w)
{
    constexpr int NQ = 2 * NB;   // q < NB: candidate list of box q; q >= NB: tested list of box q - NB
    constexpr int NW = BLK / WAVE;
    __shared__ int s_run[NQ];
    __shared__ int s_wt[NW][NQ];
    const int64_t b = w.b0 + blockIdx.x;
    const int tid = threadIdx.x, wv = tid / WAVE;
    const int64_t nf = a.nf;
    const int32_t* fmask = w.fmask + nf * b;
    const int32_t* fbox = w.fbox + nf * b;
    if (tid < NQ) s_run[tid] = 0;
    __syncthreads();
    for (int64_t i = tid; i < nf; i += BLK) {
        const int m = fmask[i], f = fbox[i];
        for (int q = 0; q < NB; q++)
            if ((m >> q) & 1) atomicAdd(&s_run[q], 1);
        if (f >= 0) atomicAdd(&s_run[NB + f], 1);
    }
    __syncthreads();
    int32_t* offs = w.offs + (int64_t)OFFS * b;
    if (tid == 0) {
        int c = 0, t = 0;
        for (int q = 0; q < NB; q++) {
            const int nc = s_run[q], nt = s_run[NB + q];
            offs[q] = c;
            offs[NB + 1 + q] = t;
            s_run[q] = c;
            s_run[NB + q] = t;
            c += nc;
            t += nt;
        }
        offs[NB] = c;
        offs[2 * NB + 1] = t;
    }
    __syncthreads();
    int2* cand = w.cand + 3 * nf * b;
    int32_t* mlist = w.mlist + nf * b;
    const int64_t* f_off = a.f_off;
    for (int64_t base = 0; base < nf; base += BLK) {
        const int64_t i = base + tid;
        const bool ok = i < nf;
        const unsigned bits = ok ? ((unsigned)fmask[i] & 0xffffu) | (fbox[i] >= 0 ? 1u << (NB + fbox[i]) : 0u) : 0u;
        int pre[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const unsigned long long bal = __ballot((bits >> q) & 1u);
            pre[q] = lanes_below(bal);
            if ((tid & (WAVE - 1)) == 0) s_wt[wv][q] = __popcll(bal);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            if ((bits >> q) & 1u) {
                int pos = s_run[q] + pre[q];
                for (int s = 0; s < wv; s++) pos += s_wt[s][q];
                if (q < NB) cand[pos] = make_int2((int)i, (int)f_off[i]);
                else mlist[pos] = (int32_t)i;
            }
        }
        __syncthreads();
        if (tid < NQ) {
            int t = 0;
            for (int s = 0; s < NW; s++) t += s_wt[s][tid];
            s_run[tid] += t;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
template <int MODE, bool COUNT>
__global__ __launch_bounds__(WAVE) void k_vis_main(abt_vis_batch a, Work w)
{
#pragma clang fp contract(off)
    const int64_t b = w.b0 + blockIdx.y;
    const int u = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t nf = a.nf;
    const int32_t* offs = w.offs + (int64_t)OFFS * b;
    // unit u -> (box, 64-facet slice of that box's tested list)
    int bb = -1, t = 0, acc = 0;
    for (int k = 0; k < NB; k++) {
        const int n = offs[NB + 2 + k] - offs[NB + 1 + k];
        const int nu = (n + WAVE - 1) / WAVE;
        if (bb < 0 && u < acc + nu) {
            bb = k;
            t = u - acc;
        }
        acc += nu;
    }
    if (bb < 0) return;   // uniform: no work for this unit
    const int mi = offs[NB + 1 + bb] + t * WAVE + lane;
    const bool has = mi < offs[NB + 2 + bb];
    const int64_t m = has ? w.mlist[nf * b + mi] : 0;
    const double* mid = a.f_mid + 3 * nf * b + 3 * m;
    const double mx = mid[0], my = mid[1], mz = mid[2];
    const int apm = (int)a.f_ap[m];
    const int2* cand = w.cand + 3 * nf * b;
    const int32_t* mapping = w.mapping + nf * b;
    bool done = !has;
    bool vis = true, inb = MODE == ABT_VIS_INC, shad = false;
    int32_t idb = 0;
    double dist = 0.0, zbest = 0.0;
    uint64_t npair = 0, nedge = 0;
    // Candidates are staged TC at a time: lane c < TC gathers candidate c's depth, aperture,
    // vertex count, mapping and first NVS projected vertices into LDS with independent loads;
    // then every lane walks the tile in order, reading each slot at one LDS address (a
    // broadcast) and forming the edge normals per pair as the reference does.
    __shared__ double tile[TC * SLOT];
    const double* vx = a.v + 3 * a.nv * b;
    const double* vy = vx + a.nv;
    const double* zmid = a.f_mid + 3 * nf * b + 2;
    const int64_t* fvi = a.f_vi;
    const int k0 = offs[bb], k1 = offs[bb + 1];
    for (int kt = k0; kt < k1; kt += TC) {
        if (__all(done)) break;
        const int nt = min(TC, k1 - kt);
        if (lane < nt) {
            const int2 cj = cand[kt + lane];
            const int64_t j = cj.x, o0 = cj.y;
            const int nvj = (int)(a.f_off[j + 1] - o0);
            double* sl = tile + lane * SLOT;
            sl[0] = zmid[3 * j];
            sl[1] = __builtin_bit_cast(double, make_int2((int)a.f_ap[j], nvj));
            sl[2] = __builtin_bit_cast(double, make_int2((int)j, mapping[j]));
            sl[3] = __builtin_bit_cast(double, make_int2((int)o0, 0));
            const int ns = min(nvj, NVS);
            for (int e = 0; e < ns; e++) {
                const int64_t vi = clampi(fvi[min(o0 + e, w.nvref - 1)] - 1, a.nv);
                sl[4 + 2 * e] = vx[vi];
                sl[5 + 2 * e] = vy[vi];
            }
        }
        __syncthreads();
        // UNR candidates at a time: their depth and edge tests are independent of the scan state,
        // so they are evaluated together (their LDS reads and FP64 chains overlap); the
        // reference's order-dependent rules are then applied to them in candidate order.
        for (int c = 0; c < nt; c += UNR) {
            if (__all(done)) break;
            bool pr[UNR], fr[UNR], hit[UNR];
            double qz[UNR];
            int qap[UNR], qmap[UNR], qnv[UNR];
#pragma unroll
            for (int q = 0; q < UNR; q++) {
                const bool valid = c + q < nt;
                const double* sl = tile + (valid ? c + q : 0) * SLOT;
                const int2 jm = __builtin_bit_cast(int2, sl[2]);
                const int2 an = __builtin_bit_cast(int2, sl[1]);
                qz[q] = sl[0];
                qap[q] = an.x;
                qnv[q] = an.y;
                qmap[q] = jm.y;
                pr[q] = valid && jm.x != m;
                fr[q] = pr[q] && !(mz > qz[q]);                         // blocker not behind m
                // edge k from vertex k to k+1 (the last back to the first): edge_norm1 =
                // y_{k+1} - y_k, edge_norm2 = -x_{k+1} + x_k, offset (mx - x_k, my - y_k)
                // (:2149-2162)
                auto outside = [&](double xa, double ya, double xb, double yb) {
                    const double en2 = -xb + xa, en1 = yb - ya;
                    const double vb1 = mx - xa, vb2 = my - ya;
                    const double chk = MODE == ABT_VIS_EXT ? -vb1 * en1 - vb2 * en2 : vb1 * en1 + vb2 * en2;
                    return chk > 0;
                };
                const double x0 = sl[4], y0 = sl[5];
                bool within;
                if (an.y == 3) {
                    const double x1 = sl[6], y1 = sl[7], x2 = sl[8], y2 = sl[9];
                    const bool o0 = outside(x0, y0, x1, y1), o1 = outside(x1, y1, x2, y2),
                               o2 = outside(x2, y2, x0, y0);
                    within = !(o0 || o1 || o2);
                } else {
                    within = true;
                    const int64_t o0 = __builtin_bit_cast(int2, sl[3]).x;
                    double xa = x0, ya = y0;
                    for (int e = 0; e < an.y; e++) {
                        const int en = e + 1 == an.y ? 0 : e + 1;
                        double xb, yb;
                        if (en < NVS) {
                            xb = sl[4 + 2 * en];
                            yb = sl[5 + 2 * en];
                        } else {
                            const int64_t vi = clampi(fvi[min(o0 + en, w.nvref - 1)] - 1, a.nv);
                            xb = vx[vi];
                            yb = vy[vi];
                            // keeps the two branches' loads apart: merged, they would become one
                            // load through a generic pointer
                            asm volatile("" : "+v"(xb), "+v"(yb));
                        }
                        if (outside(xa, ya, xb, yb)) {
                            within = false;
                            break;
                        }
                        xa = xb;
                        ya = yb;
                    }
                }
                hit[q] = fr[q] && within;
            }
#pragma unroll
            for (int q = 0; q < UNR; q++) {
                if (done) break;
                if (COUNT) {
                    npair += pr[q] ? 1 : 0;
                    nedge += fr[q] ? (uint64_t)qnv[q] : 0;
                }
                if (!hit[q]) continue;
                if (MODE == ABT_VIS_INC) {                                // :2573-2580
                    if (apm == qap[q]) shad = true;
                    else {
                        vis = false;
                        inb = false;
                    }
                    continue;
                }
                if (qmap[q] > 0) {                                        // a beam facet: keep the nearest (:2166-2180)
                    if (!inb || qz[q] < zbest) {
                        inb = true;
                        idb = qmap[q];
                        dist = qz[q] - mz;
                        zbest = qz[q];
                    }
                } else {                                                  // a blocker (:2181-2197)
                    if (apm == qap[q]) continue;
                    if (inb && qz[q] > zbest) continue;
                    vis = false;
                    inb = false;
                    done = true;
                }
            }
        }
        __syncthreads();
    }
    if (COUNT) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            npair += __shfl_xor(npair, o, 64);
            nedge += __shfl_xor(nedge, o, 64);
        }
        if (lane == 0) {
            atomicAdd(a.counts, (unsigned long long)npair);
            atomicAdd(a.counts + 1, (unsigned long long)nedge);
        }
    }
    if (!has) return;
    const int64_t o = nf * b + m;
    a.in_beam[o] = inb ? 1 : 0;
    if (a.is_vis) a.is_vis[o] = vis ? 1 : 0;
    if (MODE == ABT_VIS_INC) {                                    // :2597-2602
        if (shad) a.is_shad[o] = 1;
        if (inb) {
            a.dist_beam[o] = a.beam_mid_z[b] - mz;
            a.id_beam[o] = 1;
        }
    } else {
        a.dist_beam[o] = dist;
        a.id_beam[o] = idb;
    }
}

// ---------------------------------------------------------------------------
// rotate (beam_loop_mod.f90:1467-1526) applied to many beams: vertices, facet normals,
// midpoints and aperture normals each multiplied by the beam's rotation, every element as
// ((r_i1 p_1 + r_i2 p_2) + r_i3 p_3) like the reference's matmul (contraction off).
__device__ __forceinline__ double rot_row(const double* R, int i, double px, double py, double pz)
{
#pragma clang fp contract(off)
    double a = R[3 * i] * px;
    a = a + R[3 * i + 1] * py;
    a = a + R[3 * i + 2] * pz;
    return a;
}

__global__ __launch_bounds__(BLK) void k_vis_rotate(int64_t b0, const double* __restrict__ rot, int64_t nv,
                                                    const double* __restrict__ v0, int64_t nf,
                                                    const double* __restrict__ mid0, const double* __restrict__ n0,
                                                    int64_t na, const double* __restrict__ apn0,
                                                    double* __restrict__ v, double* __restrict__ mid,
                                                    double* __restrict__ nz, double* __restrict__ ap_nz)
{
    const int64_t b = b0 + blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    double R[9];
#pragma unroll
    for (int k = 0; k < 9; k++) R[k] = rot[9 * b + k];
    if (i < nv) {
        const double x = v0[i], y = v0[nv + i], z = v0[2 * nv + i];
        double* o = v + 3 * nv * b;
        o[i] = rot_row(R, 0, x, y, z);
        o[nv + i] = rot_row(R, 1, x, y, z);
        o[2 * nv + i] = rot_row(R, 2, x, y, z);
    }
    if (i < nf) {
        const double* p = mid0 + 3 * i;
        double* o = mid + 3 * nf * b + 3 * i;
        o[0] = rot_row(R, 0, p[0], p[1], p[2]);
        o[1] = rot_row(R, 1, p[0], p[1], p[2]);
        o[2] = rot_row(R, 2, p[0], p[1], p[2]);
        const double* q = n0 + 3 * i;
        nz[nf * b + i] = rot_row(R, 2, q[0], q[1], q[2]);
    }
    if (i < na) {
        const double* q = apn0 + 3 * i;
        ap_nz[na * b + i] = rot_row(R, 2, q[0], q[1], q[2]);
    }
}

}  // namespace vis
}  // namespace abt

// ---------------------------------------------------------------------------
// host side
namespace {

using abt::vis::Work;

constexpr size_t ALIGN = 256;
size_t up(size_t n) { return (n + ALIGN - 1) / ALIGN * ALIGN; }

struct Layout {
    size_t boxes, fbox, fmask, mapping, offs, cand, mlist, total;
};

Layout layout(int64_t nbeam, int64_t nf, int64_t nvref)
{
    Layout L;
    size_t o = 0;
    const size_t nb = (size_t)nbeam, f = (size_t)nf;
    L.boxes = o;   o += up(nb * sizeof(abt::vis::BeamBox));
    L.fbox = o;    o += up(nb * f * 4);
    L.fmask = o;   o += up(nb * f * 4);
    L.mapping = o; o += up(nb * f * 4);
    L.offs = o;    o += up(nb * abt::vis::OFFS * 4);
    L.cand = o;    o += up(nb * 3 * f * 8);
    L.mlist = o;   o += up(nb * f * 4);
    L.total = o;
    return L;
}

int launch_all(const abt_vis_batch& a, int64_t nvref, char* ws, hipStream_t st)
{
    using namespace abt::vis;
    const Layout L = layout(a.nbeam, a.nf, nvref);
    Work w;
    w.boxes = (BeamBox*)(ws + L.boxes);
    w.fbox = (int32_t*)(ws + L.fbox);
    w.fmask = (int32_t*)(ws + L.fmask);
    w.mapping = (int32_t*)(ws + L.mapping);
    w.offs = (int32_t*)(ws + L.offs);
    w.cand = (int2*)(ws + L.cand);
    w.mlist = (int32_t*)(ws + L.mlist);
    w.nvref = nvref;
    const unsigned ftiles = (unsigned)((a.nf + BLK - 1) / BLK);
    const unsigned units = (unsigned)((a.nf + WAVE - 1) / WAVE + NB);
    const bool setup = a.phases != 2, scan = a.phases != 1;
    for (int64_t b0 = 0; b0 < a.nbeam; b0 += BEAM_CHUNK) {
        const unsigned nb = (unsigned)std::min<int64_t>(BEAM_CHUNK, a.nbeam - b0);
        w.b0 = b0;
        if (setup) {
            hipLaunchKernelGGL(k_vis_boxes, dim3(nb), dim3(BLK), 0, st, a, w);
            hipLaunchKernelGGL(k_vis_facets, dim3(ftiles, nb), dim3(BLK), 0, st, a, w);
            if (a.mode != ABT_VIS_INC) hipLaunchKernelGGL(k_vis_map, dim3(nb), dim3(BLK), 0, st, a, w);
            hipLaunchKernelGGL(k_vis_lists, dim3(nb), dim3(BLK), 0, st, a, w);
        }
        if (!scan) continue;
        const dim3 g(units, nb);
#define VM(M)                                                                           \
    do {                                                                                \
        if (a.counts) hipLaunchKernelGGL((k_vis_main<M, true>), g, dim3(WAVE), 0, st, a, w);  \
        else hipLaunchKernelGGL((k_vis_main<M, false>), g, dim3(WAVE), 0, st, a, w);          \
    } while (0)
        if (a.mode == ABT_VIS_INT) VM(ABT_VIS_INT);
        else if (a.mode == ABT_VIS_EXT) VM(ABT_VIS_EXT);
        else VM(ABT_VIS_INC);
#undef VM
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return abt_fail(ABT_ERR_HIP, "visibility launch failed: %s", hipGetErrorString(e));
    return 0;
}

int check_batch(const abt_vis_batch* a)
{
    if (!a) return abt_fail(ABT_ERR_ARG, "abt_find_vis_dev: null batch");
    if (a->mode != ABT_VIS_INT && a->mode != ABT_VIS_EXT && a->mode != ABT_VIS_INC)
        return abt_fail(ABT_ERR_ARG, "unknown visibility mode %d", a->mode);
    if (a->nbeam < 0 || a->nf < 1 || a->nv < 1 || a->na < 1)
        return abt_fail(ABT_ERR_ARG, "visibility: need nbeam >= 0, nf, nv, na >= 1");
    if (a->nf >= (1LL << 31) / 3) return abt_fail(ABT_ERR_ARG, "visibility: too many facets (%lld)", (long long)a->nf);
    if (!a->f_off || !a->f_vi || !a->f_ap || !a->v || !a->f_mid || !a->f_nz || !a->ap_nz || !a->in_off ||
        !a->in_beam || !a->dist_beam || !a->id_beam || !a->is_shad)
        return abt_fail(ABT_ERR_ARG, "visibility: null array");
    if (a->mode != ABT_VIS_INC && !a->field_in) return abt_fail(ABT_ERR_ARG, "visibility: null field_in");
    if (a->mode == ABT_VIS_INC && !a->beam_mid_z) return abt_fail(ABT_ERR_ARG, "find_vis_inc needs beam_mid_z");
    if (a->phases < 0 || a->phases > 3) return abt_fail(ABT_ERR_ARG, "visibility: phases must be 0..3");
    return 0;
}

}  // namespace

extern "C" size_t abt_find_vis_workspace(int64_t nbeam, int64_t nf, int64_t nvref)
{
    if (nbeam < 0 || nf < 0 || nvref < 0) return 0;
    return layout(nbeam, nf, nvref).total;
}

extern "C" int abt_find_vis_dev(const abt_vis_batch* a, int64_t nvref, void* workspace, size_t bytes,
                                void* stream)
{
    int rc = check_batch(a);
    if (rc) return rc;
    if (a->nbeam == 0) return 0;
    if (nvref < 3 * a->nf) return abt_fail(ABT_ERR_ARG, "visibility: nvref %lld < 3 nf", (long long)nvref);
    if (nvref >= (1LL << 31)) return abt_fail(ABT_ERR_ARG, "visibility: too many vertex references");
    const size_t need = abt_find_vis_workspace(a->nbeam, a->nf, nvref);
    if (!workspace || bytes < need)
        return abt_fail(ABT_ERR_ARG, "visibility workspace of %zu bytes is smaller than %zu", bytes, need);
    if ((uintptr_t)workspace % ALIGN) return abt_fail(ABT_ERR_ARG, "visibility workspace is not 256-byte aligned");
    rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    return launch_all(*a, nvref, (char*)workspace, (hipStream_t)stream);
}

extern "C" int abt_find_vis(int mode, const abt_vis_geometry* geo, int64_t nf_in, const int64_t* field_in,
                            double beam_mid_z, uint8_t* in_beam, uint8_t* is_vis, double* dist_beam,
                            int64_t* id_beam, uint8_t* is_shad)
{
    if (!geo) return abt_fail(ABT_ERR_ARG, "abt_find_vis: null geometry");
    if (mode != ABT_VIS_INT && mode != ABT_VIS_EXT && mode != ABT_VIS_INC)
        return abt_fail(ABT_ERR_ARG, "unknown visibility mode %d", mode);
    const int64_t nv = geo->nv, nf = geo->nf, na = geo->na;
    if (nv < 1 || nf < 1 || na < 1) return abt_fail(ABT_ERR_ARG, "abt_find_vis: need nv, nf, na >= 1");
    if (!geo->v || !geo->f_off || !geo->f_vi || !geo->f_mid || !geo->f_nz || !geo->f_ap || !geo->ap_nz ||
        !in_beam || !dist_beam || !id_beam || !is_shad)
        return abt_fail(ABT_ERR_ARG, "abt_find_vis: null array");
    if (nf >= (1LL << 31) / 3) return abt_fail(ABT_ERR_ARG, "abt_find_vis: too many facets (%lld)", (long long)nf);
    if (geo->f_off[0] != 0) return abt_fail(ABT_ERR_ARG, "abt_find_vis: f_off[0] must be 0");
    for (int64_t i = 0; i < nf; i++) {
        if (geo->f_off[i + 1] - geo->f_off[i] < 3)
            return abt_fail(ABT_ERR_ARG, "abt_find_vis: facet %lld has fewer than 3 vertices", (long long)(i + 1));
        if (geo->f_ap[i] < 1 || geo->f_ap[i] > na)
            return abt_fail(ABT_ERR_ARG, "abt_find_vis: facet %lld aperture id %lld out of 1..%lld",
                            (long long)(i + 1), (long long)geo->f_ap[i], (long long)na);
    }
    const int64_t nvref = geo->f_off[nf];
    if (nvref >= (1LL << 31)) return abt_fail(ABT_ERR_ARG, "abt_find_vis: too many vertex references");
    for (int64_t k = 0; k < nvref; k++)
        if (geo->f_vi[k] < 1 || geo->f_vi[k] > nv)
            return abt_fail(ABT_ERR_ARG, "abt_find_vis: vertex id %lld out of 1..%lld", (long long)geo->f_vi[k],
                            (long long)nv);
    if (mode != ABT_VIS_INC) {
        if (nf_in < 1 || !field_in)
            return abt_fail(ABT_ERR_ARG, "abt_find_vis: the beam needs at least one illuminating facet");
        for (int64_t k = 0; k < nf_in; k++)
            if (field_in[k] < 1 || field_in[k] > nf)
                return abt_fail(ABT_ERR_ARG, "abt_find_vis: field_in facet %lld out of 1..%lld",
                                (long long)field_in[k], (long long)nf);
    } else {
        nf_in = 0;
    }
    // one staging buffer: inputs, outputs, workspace
    size_t o = 0;
    auto take = [&](size_t n) { const size_t at = o; o += up(n); return at; };
    const size_t o_v = take(sizeof(double) * 3 * nv), o_off = take(8 * (nf + 1)), o_vi = take(8 * nvref),
                 o_mid = take(sizeof(double) * 3 * nf), o_nz = take(8 * nf), o_ap = take(8 * nf),
                 o_apnz = take(8 * na), o_inoff = take(16), o_fin = take(8 * (nf_in > 0 ? nf_in : 1)),
                 o_bz = take(8), o_inb = take(nf), o_vis = take(nf), o_dist = take(8 * nf), o_id = take(8 * nf),
                 o_shad = take(nf);
    const size_t o_ws = o;
    o += layout(1, nf, nvref).total;
    hipStream_t st;
    void* p;
    int rc = abt_thread_staging(8, o, &st, &p);
    if (rc) return rc;
    char* d = (char*)p;
    const int64_t inoff[2] = {0, nf_in};
    HIPCHK(hipMemcpyAsync(d + o_v, geo->v, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_off, geo->f_off, 8 * (nf + 1), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_vi, geo->f_vi, 8 * nvref, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_mid, geo->f_mid, sizeof(double) * 3 * nf, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_nz, geo->f_nz, 8 * nf, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_ap, geo->f_ap, 8 * nf, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_apnz, geo->ap_nz, 8 * na, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_inoff, inoff, 16, hipMemcpyHostToDevice, st));
    if (nf_in > 0) HIPCHK(hipMemcpyAsync(d + o_fin, field_in, 8 * nf_in, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + o_bz, &beam_mid_z, 8, hipMemcpyHostToDevice, st));
    abt_vis_batch a;
    a.mode = mode;
    a.nbeam = 1;
    a.nv = nv;
    a.nf = nf;
    a.na = na;
    a.f_off = (const int64_t*)(d + o_off);
    a.f_vi = (const int64_t*)(d + o_vi);
    a.f_ap = (const int64_t*)(d + o_ap);
    a.v = (const double*)(d + o_v);
    a.f_mid = (const double*)(d + o_mid);
    a.f_nz = (const double*)(d + o_nz);
    a.ap_nz = (const double*)(d + o_apnz);
    a.in_off = (const int64_t*)(d + o_inoff);
    a.field_in = (const int64_t*)(d + o_fin);
    a.beam_mid_z = (const double*)(d + o_bz);
    a.in_beam = (uint8_t*)(d + o_inb);
    a.is_vis = (uint8_t*)(d + o_vis);
    a.dist_beam = (double*)(d + o_dist);
    a.id_beam = (int64_t*)(d + o_id);
    a.is_shad = (uint8_t*)(d + o_shad);
    a.counts = nullptr;
    a.phases = 0;
    rc = launch_all(a, nvref, d + o_ws, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(in_beam, d + o_inb, nf, hipMemcpyDeviceToHost, st));
    if (is_vis) HIPCHK(hipMemcpyAsync(is_vis, d + o_vis, nf, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dist_beam, d + o_dist, 8 * nf, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(id_beam, d + o_id, 8 * nf, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(is_shad, d + o_shad, nf, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// rotate()'s rotation for each beam (beam_loop_mod.f90:1488-1503), on the host with the C
// library's atan2, acos, sqrt and sincos like the reference (pi its REAL*4 literal,
// misc_submod.f90:12): R[9 b + 3 i + k] = rot(i+1, k+1).
extern "C" int abt_vis_rotations(int64_t nbeam, const double* prop, double* R)
{
#pragma clang fp contract(off)
    if (nbeam < 0 || (nbeam > 0 && (!prop || !R))) return abt_fail(ABT_ERR_ARG, "abt_vis_rotations: bad arguments");
    const double pi = 3.1415927410125732;
    for (int64_t b = 0; b < nbeam; b++) {
        const double* p = prop + 3 * b;
        double r1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, r2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, c, sn;
        const double th1 = 2 * pi - atan2(p[1], p[0]);
        sincos(th1, &sn, &c);
        r1[0] = c;
        r1[1] = -sn;
        r1[3] = sn;
        r1[4] = c;
        r1[8] = 1;
        double t[3];
        for (int i = 0; i < 3; i++) {
            double a = 0.0;
            for (int k = 0; k < 3; k++) a = a + r1[3 * i + k] * p[k];
            t[i] = a;
        }
        const double nrm = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);   // normalise_vec (misc_submod.f90:2216-2227)
        for (int i = 0; i < 3; i++) t[i] = t[i] / nrm;
        sincos(acos(-t[2]), &sn, &c);
        r2[0] = c;
        r2[2] = sn;
        r2[6] = -sn;
        r2[8] = c;
        r2[4] = 1;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double a = 0.0;
                for (int k = 0; k < 3; k++) a = a + r2[3 * i + k] * r1[3 * k + j];
                R[9 * b + 3 * i + j] = a;
            }
    }
    return 0;
}

extern "C" int abt_vis_rotate_dev(int64_t nbeam, const double* rot, int64_t nv, const double* v0, int64_t nf,
                                  const double* mid0, const double* n0, int64_t na, const double* apn0, double* v,
                                  double* mid, double* nz, double* ap_nz, void* stream)
{
    using namespace abt::vis;
    if (nbeam < 0 || nv < 1 || nf < 1 || na < 1) return abt_fail(ABT_ERR_ARG, "abt_vis_rotate_dev: bad sizes");
    if (nbeam == 0) return 0;
    if (!rot || !v0 || !mid0 || !n0 || !apn0 || !v || !mid || !nz || !ap_nz)
        return abt_fail(ABT_ERR_ARG, "abt_vis_rotate_dev: null array");
    int rc = abt_ensure_device(nullptr);
    if (rc) return rc;
    const int64_t n = std::max(nv, std::max(nf, na));
    const unsigned gx = (unsigned)((n + BLK - 1) / BLK);
    for (int64_t b0 = 0; b0 < nbeam; b0 += BEAM_CHUNK) {
        const unsigned nb = (unsigned)std::min<int64_t>(BEAM_CHUNK, nbeam - b0);
        hipLaunchKernelGGL(k_vis_rotate, dim3(gx, nb), dim3(BLK), 0, (hipStream_t)stream, b0, rot, nv, v0, nf, mid0,
                           n0, na, apn0, v, mid, nz, ap_nz);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return abt_fail(ABT_ERR_HIP, "abt_vis_rotate_dev launch failed: %s", hipGetErrorString(e));
    return 0;
}
