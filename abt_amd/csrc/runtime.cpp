// runtime.cpp -- host side of libabt_diff.so: the C-ABI of include/abt_diff.h.
//
// Device context (lazy init, one device per process), grids resident on the
// device, host preparation of beam records (reference operation order, glibc),
// per-thread staging buffers and streams so that the host entry points are
// re-entrant, and the launch policy (beam chunking so that every launch has
// enough workgroups for 256 CUs).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/abt_diff.h"
#include "ffbody.h"
#include "frame.h"
#include "records.h"
#include "rt_internal.h"

extern "C" int abt_launch_far_field(const abt::DevGrid* g, const double* recs, const double* vrecs,
                                    int64_t nb, int prec, int tb, int fast_diff, int nchunks,
                                    double* out, double* work, hipStream_t st);
extern "C" int abt_launch_zero(double* out, int64_t n, hipStream_t st);
extern "C" int abt_launch_finalise(const abt::DevGrid* g, double* ab, double* ae, double* tot,
                                   double* M, double* m1d, double* tmp, double* params,
                                   int scaling, double ebeam, double eext, hipStream_t st);
extern "C" int abt_launch_prepare(int64_t n, const double* ampl, const double* vk7,
                                  const double* prop, const double* fov, double fov_value,
                                  const int64_t* vert_off, const double* verts, double lambda,
                                  double* recs, double* vrecs, hipStream_t st);

using abt::DevGrid;

// ---------------------------------------------------------------------------
// errors
static thread_local char t_err[512];

int abt_fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_err, sizeof(t_err), fmt, ap);
    va_end(ap);
    return code;
}
#define fail abt_fail


// ---------------------------------------------------------------------------
// context
namespace {

struct Context {
    std::mutex mu;
    int device = -1;
    int cus = 256;
};
Context g_ctx;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

// per-thread staging: the host entry points may run concurrently on several threads
struct ThreadState {
    hipStream_t st = nullptr;
    DevBuf bufs[9];   // [8]: the visibility entry points
    std::vector<double> hrec;
};

std::mutex g_reg_mu;
std::vector<ThreadState*> g_threads;

ThreadState& tstate()
{
    static thread_local ThreadState* ts = nullptr;
    if (!ts) {
        ts = new ThreadState();
        std::lock_guard<std::mutex> lk(g_reg_mu);
        g_threads.push_back(ts);
    }
    return *ts;
}

int ensure_device()
{
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    if (g_ctx.device >= 0) {
        HIPCHK(hipSetDevice(g_ctx.device));
        return 0;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(ABT_ERR_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, 0));
    g_ctx.device = 0;
    g_ctx.cus = prop.multiProcessorCount;
    return 0;
}

int buf_ensure(DevBuf& b, size_t bytes)
{
    if (b.cap >= bytes) return 0;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t cap = std::max(bytes, (size_t)4096);
    if (hipMalloc(&b.p, cap) != hipSuccess) {
        b.p = nullptr;
        return fail(ABT_ERR_NOMEM, "hipMalloc of %zu bytes failed", cap);
    }
    b.cap = cap;
    return 0;
}

int stream_of(ThreadState& ts)
{
    if (!ts.st) HIPCHK(hipStreamCreateWithFlags(&ts.st, hipStreamNonBlocking));
    return 0;
}

int check_beams(const abt_beams* b, const char* what)
{
    if (!b) return fail(ABT_ERR_ARG, "%s: null beam set", what);
    if (b->n < 0) return fail(ABT_ERR_ARG, "%s: negative beam count", what);
    if (b->n == 0) return 0;
    if (!b->ampl || !b->vk7 || !b->prop || !b->vert_off || !b->verts)
        return fail(ABT_ERR_ARG, "%s: null array", what);
    return 0;
}

int check_beams_host(const abt_beams* b, const char* what)
{
    int rc = check_beams(b, what);
    if (rc) return rc;
    if (b->n == 0) return 0;
    if (b->vert_off[0] != 0) return fail(ABT_ERR_ARG, "%s: vert_off[0] must be 0", what);
    for (int64_t i = 0; i < b->n; i++)
        if (b->vert_off[i + 1] - b->vert_off[i] < 3)
            return fail(ABT_ERR_ARG, "%s: beam %lld has fewer than 3 vertices", what, (long long)i);
    return 0;
}

// theta_i (diff_mod.f90:583-591): first theta > 1 degree (REAL*4 pi), 0-based.
// The reference leaves it undefined when there is none; here the last column.
int theta_i_of(int64_t nth, const double* th)
{
    for (int64_t j = 0; j < nth; j++)
        if (th[j] > 1 * abt::PI_R4 / 180) return (int)j;
    return (int)(nth - 1);
}

int choose_chunks(int64_t nbins, int64_t nb, int cus)
{
    // ~320 workgroups per CU (80 rounds at 4 resident per CU): the last round's partly filled
    // CUs, and the slower forward-fallback rows, are then a small share of the launch.  Measured
    // on the 200,000-aperture stress test (2533 bin tiles, mixed kernel): 4 chunks run 9 % faster
    // than 1, 16 a further 3 %, 32 and 64 a further 0.7 % and 1 %.  Chunks keep >= 512 beams;
    // at most 32 (the workspace is 64 B per bin per chunk).
    const int64_t tiles = (nbins + 255) / 256;
    const int64_t want = (int64_t)cus * 320;
    int64_t c = (want + tiles - 1) / tiles;
    c = std::min<int64_t>(c, std::max<int64_t>(1, nb / 512));
    c = std::max<int64_t>(1, std::min<int64_t>(c, 32));
    return (int)c;
}

int tile_beams(int prec) { return prec == ABT_FP64 ? 16 : 16; }

}  // namespace

int abt_ensure_device(int* cus)
{
    int rc = ensure_device();
    if (!rc && cus) *cus = g_ctx.cus;
    return rc;
}

int abt_thread_staging(int slot, size_t bytes, hipStream_t* st, void** p)
{
    if (slot < 0 || slot >= 9) return fail(ABT_ERR_ARG, "staging slot %d out of range", slot);
    int rc = ensure_device();
    if (rc) return rc;
    ThreadState& ts = tstate();
    rc = stream_of(ts);
    if (!rc) rc = buf_ensure(ts.bufs[slot], bytes);
    if (rc) return rc;
    *st = ts.st;
    *p = ts.bufs[slot].p;
    return 0;
}

// ---------------------------------------------------------------------------
// device grid

static void host_bins(int64_t nth, const double* th, int64_t nph, const double* ph, double* x,
                      double* y, double* z)
{
    // make_far_field_bins (diff_mod.f90:827-855), meshgrid layout (nphi, ntheta); the
    // sine and cosine of one angle from one sincos call, as gfortran -O3 computes them
    const double r = 1e6;
    for (int64_t j = 0; j < nth; j++)
        for (int64_t i = 0; i < nph; i++) {
            const int64_t b = i + nph * j;
            double st, ct, sp, cp;
            sincos(th[j], &st, &ct);
            sincos(ph[i], &sp, &cp);
            x[b] = r * st * cp;
            y[b] = r * st * sp;
            z[b] = -(r * ct);
        }
}

static int grid_build(int64_t nth, const double* th, int64_t nph, const double* ph,
                      const double* xs, const double* ys, const double* zs, double lambda,
                      abt_dev_grid** out)
{
    if (nth <= 0 || nph <= 0) return fail(ABT_ERR_ARG, "empty grid (%lld x %lld)", (long long)nth, (long long)nph);
    if ((double)nth * (double)nph > 2.0e9) return fail(ABT_ERR_ARG, "grid too large");
    if (!(lambda > 0)) return fail(ABT_ERR_ARG, "wavelength must be positive");
    int rc = ensure_device();
    if (rc) return rc;
    const int64_t nb = nth * nph;
    std::vector<double> h((size_t)(5 * nb + 3 * nph + nth + abt::ETAB_GRID));
    double *x = h.data(), *y = x + nb, *z = y + nb, *rad = z + nb, *krr = rad + nb;
    double *cp = krr + nb, *sp = cp + nph, *phv = sp + nph, *thv = phv + nph, *etab = thv + nth;
    abt::fill_exp_table(etab);
    memcpy(phv, ph, sizeof(double) * nph);
    memcpy(thv, th, sizeof(double) * nth);
    if (xs) {
        memcpy(x, xs, sizeof(double) * nb);
        memcpy(y, ys, sizeof(double) * nb);
        memcpy(z, zs, sizeof(double) * nb);
    } else {
        host_bins(nth, th, nph, ph, x, y, z);
    }
    const double waveno = 2.0 * abt::PI_R4 / lambda;
    const double krev = waveno / 6.283185307179586476925;
    for (int64_t b = 0; b < nb; b++) {
        const double r = sqrt(x[b] * x[b] + y[b] * y[b] + z[b] * z[b]);
        rad[b] = r;
        krr[b] = krev * r;
    }
    for (int64_t i = 0; i < nph; i++) sincos(ph[i], &sp[i], &cp[i]);
    abt_dev_grid* g = new abt_dev_grid();
    if (hipMalloc(&g->mem, sizeof(double) * h.size()) != hipSuccess) {
        delete g;
        return fail(ABT_ERR_NOMEM, "grid allocation failed");
    }
    if (hipMemcpy(g->mem, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(g->mem);
        delete g;
        return fail(ABT_ERR_HIP, "grid upload failed");
    }
    double* d = (double*)g->mem;
    g->d.nphi = (int32_t)nph;
    g->d.ntheta = (int32_t)nth;
    g->d.nbins = (int32_t)nb;
    g->d.theta_i = theta_i_of(nth, th);
    abt::devgrid_constants(g->d, lambda);
    g->d.x = d;
    g->d.y = d + nb;
    g->d.z = d + 2 * nb;
    g->d.rad = d + 3 * nb;
    g->d.kr_rev = d + 4 * nb;
    g->d.cphi = d + 5 * nb;
    g->d.sphi = d + 5 * nb + nph;
    g->d.phi = d + 5 * nb + 2 * nph;
    g->d.theta = d + 5 * nb + 3 * nph;
    g->d.etab = d + 5 * nb + 3 * nph + nth;
    g->nbins = nb;
    *out = g;
    return 0;
}

// ---------------------------------------------------------------------------
// Grid cache of the host entry points.  PBT calls diff_main once per orientation and,
// in -memory mode, diffraction once per beam from several OpenMP threads, always on the
// same bins: the device grid is built on the first call and kept, keyed by the exact
// contents of its inputs (angles, wavelength and, for diffraction, the caller's bin
// positions), so later calls neither rebuild nor allocate it.  Entries in use are never
// evicted; when every slot is busy a call builds a private grid and frees it.
namespace {

struct GridEntry {
    int64_t nth = 0, nph = 0;
    double lambda = 0.0;
    std::vector<double> th, ph, xyz;   // xyz empty: bins from make_far_field_bins
    abt_dev_grid* g = nullptr;
    int users = 0;
    uint64_t last = 0;
};

constexpr size_t GRID_CACHE_SLOTS = 8;
std::mutex g_grid_mu;
std::list<GridEntry> g_grids;
uint64_t g_grid_clock = 0;

bool same(const std::vector<double>& v, const double* p, int64_t n)
{
    return (int64_t)v.size() == n && (n == 0 || memcmp(v.data(), p, sizeof(double) * (size_t)n) == 0);
}

// A cached grid for these inputs (xs, ys, zs may be null), with a use held: release it
// with grid_release.  *priv is set when the grid is private to the call (cache full).
int grid_acquire(int64_t nth, const double* th, int64_t nph, const double* ph, const double* xs,
                 const double* ys, const double* zs, double lambda, abt_dev_grid** out, bool* priv)
{
    const int64_t nb = nth * nph;
    *priv = false;
    {
        std::lock_guard<std::mutex> lk(g_grid_mu);
        for (GridEntry& e : g_grids) {
            if (e.nth != nth || e.nph != nph || e.lambda != lambda) continue;
            if (!same(e.th, th, nth) || !same(e.ph, ph, nph)) continue;
            if (xs ? (e.xyz.size() != (size_t)(3 * nb) || memcmp(e.xyz.data(), xs, sizeof(double) * nb) != 0 ||
                      memcmp(e.xyz.data() + nb, ys, sizeof(double) * nb) != 0 ||
                      memcmp(e.xyz.data() + 2 * nb, zs, sizeof(double) * nb) != 0)
                   : !e.xyz.empty())
                continue;
            e.users++;
            e.last = ++g_grid_clock;
            *out = e.g;
            return 0;
        }
    }
    abt_dev_grid* g = nullptr;
    int rc = grid_build(nth, th, nph, ph, xs, ys, zs, lambda, &g);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_grid_mu);
    if (g_grids.size() >= GRID_CACHE_SLOTS) {
        // evict the least recently used idle entry
        auto victim = g_grids.end();
        for (auto it = g_grids.begin(); it != g_grids.end(); ++it)
            if (it->users == 0 && (victim == g_grids.end() || it->last < victim->last)) victim = it;
        if (victim == g_grids.end()) {
            *priv = true;
            *out = g;
            return 0;
        }
        abt_grid_destroy(victim->g);
        g_grids.erase(victim);
    }
    GridEntry e;
    e.nth = nth;
    e.nph = nph;
    e.lambda = lambda;
    e.th.assign(th, th + nth);
    e.ph.assign(ph, ph + nph);
    if (xs) {
        e.xyz.resize((size_t)(3 * nb));
        memcpy(e.xyz.data(), xs, sizeof(double) * nb);
        memcpy(e.xyz.data() + nb, ys, sizeof(double) * nb);
        memcpy(e.xyz.data() + 2 * nb, zs, sizeof(double) * nb);
    }
    e.g = g;
    e.users = 1;
    e.last = ++g_grid_clock;
    g_grids.push_back(std::move(e));
    *out = g;
    return 0;
}

void grid_release(abt_dev_grid* g, bool priv)
{
    if (priv) {
        abt_grid_destroy(g);
        return;
    }
    std::lock_guard<std::mutex> lk(g_grid_mu);
    for (GridEntry& e : g_grids)
        if (e.g == g) {
            e.users--;
            return;
        }
}

void grid_cache_clear()
{
    std::lock_guard<std::mutex> lk(g_grid_mu);
    for (auto it = g_grids.begin(); it != g_grids.end();) {
        if (it->users == 0) {
            abt_grid_destroy(it->g);
            it = g_grids.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host preparation of beam records
static void prep_range(const abt_beams* b, double fov_value, double lambda, int64_t i0, int64_t i1,
                       double* recs, double* vrecs)
{
    for (int64_t i = i0; i < i1; i++) {
        const int64_t o0 = b->vert_off[i], o1 = b->vert_off[i + 1];
        const double f = b->fov ? b->fov[i] : fov_value;
        abt::make_beam_record(b->ampl + 8 * i, b->vk7 + 3 * i, b->prop + 3 * i, f, o1 - o0,
                              b->verts + 3 * o0, lambda, o0, recs + abt::REC * i,
                              (vrecs && o1 - o0 > 3) ? vrecs + abt::VREC * o0 : nullptr);
    }
}

static void prep_host(const abt_beams* b, double fov_value, double lambda, double* recs, double* vrecs)
{
    const int64_t n = b->n;
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (n < 4096 || nt == 1) {
        prep_range(b, fov_value, lambda, 0, n, recs, vrecs);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const int64_t i0 = t * per, i1 = std::min(n, i0 + per);
        if (i0 >= i1) break;
        th.emplace_back(prep_range, b, fov_value, lambda, i0, i1, recs, vrecs);
    }
    for (auto& t : th) t.join();
}

static bool any_polygon(const abt_beams* b)
{
    for (int64_t i = 0; i < b->n; i++)
        if (b->vert_off[i + 1] - b->vert_off[i] > 3) return true;
    return false;
}

// Stage a host beam set on the device and run the far-field sum into dev_out.
static int run_host_tree(const DevGrid& g, const abt_beams* b, double fov_value, int prec,
                         int fast_diff, double* dev_out, ThreadState& ts, int slot)
{
    const int64_t n = b->n;
    hipStream_t st = ts.st;
    if (n == 0) {
        int e = abt_launch_zero(dev_out, 8 * (int64_t)g.nbins, st);
        if (e) return fail(ABT_ERR_HIP, "zero-fill launch failed: %s", hipGetErrorString((hipError_t)e));
        return 0;
    }
    const bool poly = any_polygon(b);
    const int64_t nvert = b->vert_off[n];
    const size_t rec_bytes = sizeof(double) * abt::REC * (size_t)n;
    const size_t vrec_bytes = poly ? sizeof(double) * abt::VREC * (size_t)nvert : 0;
    ts.hrec.resize((rec_bytes + vrec_bytes) / sizeof(double) + 1);
    double* hrec = ts.hrec.data();
    double* hvrec = poly ? hrec + abt::REC * n : nullptr;
    prep_host(b, fov_value, g.lambda, hrec, hvrec);
    int rc = buf_ensure(ts.bufs[slot], rec_bytes + vrec_bytes + 64);
    if (rc) return rc;
    double* drec = (double*)ts.bufs[slot].p;
    double* dvrec = poly ? drec + abt::REC * n : drec;
    HIPCHK(hipMemcpyAsync(drec, hrec, rec_bytes + vrec_bytes, hipMemcpyHostToDevice, st));
    const int nch = choose_chunks(g.nbins, n, g_ctx.cus);
    double* work = nullptr;
    if (nch > 1) {
        rc = buf_ensure(ts.bufs[slot + 1], sizeof(double) * 8 * (size_t)g.nbins * nch);
        if (rc) return rc;
        work = (double*)ts.bufs[slot + 1].p;
    }
    int e = abt_launch_far_field(&g, drec, dvrec, n, prec, tile_beams(prec), fast_diff, nch, dev_out,
                                 work, st);
    if (e) return fail(ABT_ERR_HIP, "far-field launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
extern "C" {

int abt_version(void) { return ABT_DIFF_ABI_VERSION; }

const char* abt_last_error(void) { return t_err; }

int abt_init(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(ABT_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= n) return fail(ABT_ERR_ARG, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ABT_ERR_NODEV, "device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
    g_ctx.device = device;
    g_ctx.cus = prop.multiProcessorCount;
    return 0;
}

int abt_finalize(void)
{
    grid_cache_clear();
    std::lock_guard<std::mutex> lk(g_reg_mu);
    for (ThreadState* ts : g_threads) {
        for (auto& b : ts->bufs) {
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
        }
        if (ts->st) (void)hipStreamDestroy(ts->st);
        ts->st = nullptr;
    }
    return 0;
}

int abt_make_far_field_bins(int64_t n_theta, const double* theta, int64_t n_phi, const double* phi,
                            double* x, double* y, double* z)
{
    if (n_theta <= 0 || n_phi <= 0 || !theta || !phi || !x || !y || !z)
        return fail(ABT_ERR_ARG, "abt_make_far_field_bins: bad arguments");
    host_bins(n_theta, theta, n_phi, phi, x, y, z);
    return 0;
}

int64_t abt_record_doubles(void) { return abt::REC; }
int64_t abt_vertex_record_doubles(void) { return abt::VREC; }

int abt_prepare_beams_host(const abt_beams* beams, double fov_value, double lambda, double* recs,
                           double* vrecs)
{
    int rc = check_beams_host(beams, "abt_prepare_beams_host");
    if (rc) return rc;
    if (!recs && beams->n) return fail(ABT_ERR_ARG, "abt_prepare_beams_host: null recs");
    if (!(lambda > 0)) return fail(ABT_ERR_ARG, "wavelength must be positive");
    if (!vrecs && any_polygon(beams))
        return fail(ABT_ERR_ARG, "abt_prepare_beams_host: apertures with more than 3 vertices need vrecs");
    prep_host(beams, fov_value, lambda, recs, vrecs);
    return 0;
}

int abt_prepare_beams_dev(const abt_beams* b, double fov_value, double lambda, double* dev_recs,
                          double* dev_vrecs, void* stream)
{
    int rc = check_beams(b, "abt_prepare_beams_dev");
    if (rc) return rc;
    if (!(lambda > 0)) return fail(ABT_ERR_ARG, "wavelength must be positive");
    if (b->n == 0) return 0;
    rc = ensure_device();
    if (rc) return rc;
    int e = abt_launch_prepare(b->n, b->ampl, b->vk7, b->prop, b->fov, fov_value, b->vert_off,
                               b->verts, lambda, dev_recs, dev_vrecs, (hipStream_t)stream);
    if (e) return fail(ABT_ERR_HIP, "prepare launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

int abt_grid_create(const abt_grid* grid, double lambda, abt_dev_grid** out)
{
    if (!grid || !out || !grid->theta || !grid->phi) return fail(ABT_ERR_ARG, "abt_grid_create: bad arguments");
    return grid_build(grid->n_theta, grid->theta, grid->n_phi, grid->phi, nullptr, nullptr, nullptr,
                      lambda, out);
}

int abt_grid_destroy(abt_dev_grid* g)
{
    if (!g) return 0;
    if (g->mem) (void)hipFree(g->mem);
    delete g;
    return 0;
}

int64_t abt_grid_bins(const abt_dev_grid* g) { return g ? g->nbins : 0; }

size_t abt_far_field_workspace(const abt_dev_grid* g, int64_t nb, int precision)
{
    (void)precision;
    if (!g) return 0;
    if (ensure_device()) return 0;
    const int nch = choose_chunks(g->nbins, nb, g_ctx.cus);
    return nch > 1 ? sizeof(double) * 8 * (size_t)g->nbins * nch : 0;
}

int abt_far_field_dev(const abt_dev_grid* g, const double* dev_recs, const double* dev_vrecs,
                      int64_t nb, int precision, int is_fast_diff, double* dev_out, void* dev_work,
                      size_t work_bytes, void* stream)
{
    if (!g || !dev_out) return fail(ABT_ERR_ARG, "abt_far_field_dev: null argument");
    if (precision != ABT_FP64 && precision != ABT_MIXED && precision != ABT_FP64_FAST)
        return fail(ABT_ERR_ARG, "unknown precision %d", precision);
    if (nb < 0) return fail(ABT_ERR_ARG, "negative beam count");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (nb == 0) {
        int e = abt_launch_zero(dev_out, 8 * (int64_t)g->nbins, st);
        if (e) return fail(ABT_ERR_HIP, "zero-fill launch failed: %s", hipGetErrorString((hipError_t)e));
        return 0;
    }
    if (!dev_recs) return fail(ABT_ERR_ARG, "abt_far_field_dev: null records");
    const int nch = choose_chunks(g->nbins, nb, g_ctx.cus);
    const size_t need = nch > 1 ? sizeof(double) * 8 * (size_t)g->nbins * nch : 0;
    void* work = dev_work;
    if (need && (!work || work_bytes < need)) {
        if (work) return fail(ABT_ERR_ARG, "workspace of %zu bytes is smaller than the %zu needed", work_bytes, need);
        ThreadState& ts = tstate();
        rc = buf_ensure(ts.bufs[7], need);
        if (rc) return rc;
        work = ts.bufs[7].p;
    }
    int e = abt_launch_far_field(&g->d, dev_recs, dev_vrecs ? dev_vrecs : dev_recs, nb, precision,
                                 tile_beams(precision), is_fast_diff, nch, dev_out, (double*)work, st);
    if (e) return fail(ABT_ERR_HIP, "far-field launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

int abt_far_field_incoherent_dev(const abt_dev_grid* g, const double* dev_recs,
                                 const double* dev_vrecs, int64_t nb, int is_fast_diff,
                                 double* dev_intensity, void* stream)
{
    if (!g || !dev_intensity) return fail(ABT_ERR_ARG, "abt_far_field_incoherent_dev: null argument");
    if (nb < 0) return fail(ABT_ERR_ARG, "negative beam count");
    int rc = ensure_device();
    if (rc) return rc;
    if (nb == 0) {
        int e = abt_launch_zero(dev_intensity, g->nbins, (hipStream_t)stream);
        if (e) return fail(ABT_ERR_HIP, "zero-fill launch failed: %s", hipGetErrorString((hipError_t)e));
        return 0;
    }
    if (!dev_recs) return fail(ABT_ERR_ARG, "abt_far_field_incoherent_dev: null records");
    int e = abt_launch_far_field(&g->d, dev_recs, dev_vrecs ? dev_vrecs : dev_recs, nb, 2, 16,
                                 is_fast_diff, 1, dev_intensity, nullptr, (hipStream_t)stream);
    if (e) return fail(ABT_ERR_HIP, "far-field launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

int abt_diff_main(const abt_grid* grid, double lambda, int is_fast_diff, const abt_beams* beams,
                  const abt_beams* ext, int precision, double* ampl_far_beam, double* ampl_far_ext_diff)
{
    const bool want_b = ampl_far_beam != nullptr, want_e = ampl_far_ext_diff != nullptr;
    if (!grid || (!want_b && !want_e)) return fail(ABT_ERR_ARG, "abt_diff_main: null argument");
    if (precision != ABT_FP64 && precision != ABT_MIXED && precision != ABT_FP64_FAST)
        return fail(ABT_ERR_ARG, "unknown precision %d", precision);
    int rc = want_b ? check_beams_host(beams, "abt_diff_main(beams)") : 0;
    if (rc) return rc;
    rc = want_e ? check_beams_host(ext, "abt_diff_main(ext)") : 0;
    if (rc) return rc;
    if (!grid->theta || !grid->phi) return fail(ABT_ERR_ARG, "abt_diff_main: null angles");
    abt_dev_grid* g = nullptr;
    bool priv = false;
    rc = grid_acquire(grid->n_theta, grid->theta, grid->n_phi, grid->phi, nullptr, nullptr, nullptr, lambda, &g,
                      &priv);
    if (rc) return rc;
    ThreadState& ts = tstate();
    rc = stream_of(ts);
    const size_t ob = sizeof(double) * 8 * (size_t)g->nbins;
    if (!rc) rc = buf_ensure(ts.bufs[0], 2 * ob);
    if (!rc) {
        double* d0 = (double*)ts.bufs[0].p;
        double* d1 = d0 + 8 * g->nbins;
        if (want_b) rc = run_host_tree(g->d, beams, 0.0, precision, is_fast_diff, d0, ts, 1);
        if (!rc && want_b && want_e) {
            // the ext tree's staging must not overwrite the beam tree's still in flight
            hipError_t e = hipStreamSynchronize(ts.st);
            if (e != hipSuccess) rc = fail(ABT_ERR_HIP, "stream sync: %s", hipGetErrorString(e));
        }
        // ext-diff beams: fov = pi/2 (diff_mod.f90:797)
        if (!rc && want_e) {
            abt_beams e2 = *ext;
            e2.fov = nullptr;
            rc = run_host_tree(g->d, &e2, abt::PI_R4 / 2.0, precision, is_fast_diff, d1, ts, 1);
        }
        if (!rc) {
            hipError_t e = hipSuccess;
            if (want_b) e = hipMemcpyAsync(ampl_far_beam, d0, ob, hipMemcpyDeviceToHost, ts.st);
            if (e == hipSuccess && want_e) e = hipMemcpyAsync(ampl_far_ext_diff, d1, ob, hipMemcpyDeviceToHost, ts.st);
            if (e == hipSuccess) e = hipStreamSynchronize(ts.st);
            if (e != hipSuccess) rc = fail(ABT_ERR_HIP, "copy-back: %s", hipGetErrorString(e));
        }
    }
    grid_release(g, priv);
    return rc;
}

int abt_diffraction(int64_t n_phi, int64_t n_theta, const double* xfar, const double* yfar,
                    const double* zfar, double lambda, const double* phi_vals,
                    const double* theta_vals, int is_fast_diff, const double* ampl,
                    const double* vk7, const double* prop, double fov, int64_t nv,
                    const double* verts, int precision, double* amplC)
{
    if (!xfar || !yfar || !zfar || !phi_vals || !theta_vals || !ampl || !vk7 || !prop || !verts || !amplC)
        return fail(ABT_ERR_ARG, "abt_diffraction: null argument");
    if (nv < 3) return fail(ABT_ERR_ARG, "abt_diffraction: aperture needs at least 3 vertices");
    if (precision != ABT_FP64 && precision != ABT_MIXED && precision != ABT_FP64_FAST)
        return fail(ABT_ERR_ARG, "unknown precision %d", precision);
    abt_dev_grid* g = nullptr;
    bool priv = false;
    int rc = grid_acquire(n_theta, theta_vals, n_phi, phi_vals, xfar, yfar, zfar, lambda, &g, &priv);
    if (rc) return rc;
    ThreadState& ts = tstate();
    rc = stream_of(ts);
    const size_t ob = sizeof(double) * 8 * (size_t)g->nbins;
    if (!rc) rc = buf_ensure(ts.bufs[0], ob);
    if (!rc) {
        int64_t off[2] = {0, nv};
        abt_beams b = {1, ampl, vk7, prop, &fov, off, verts};
        rc = run_host_tree(g->d, &b, fov, precision, is_fast_diff, (double*)ts.bufs[0].p, ts, 1);
        if (!rc) {
            hipError_t e = hipMemcpyAsync(amplC, ts.bufs[0].p, ob, hipMemcpyDeviceToHost, ts.st);
            if (e == hipSuccess) e = hipStreamSynchronize(ts.st);
            if (e != hipSuccess) rc = fail(ABT_ERR_HIP, "copy-back: %s", hipGetErrorString(e));
        }
    }
    grid_release(g, priv);
    return rc;
}

size_t abt_finalise_workspace(const abt_dev_grid* g)
{
    if (!g) return 0;
    return sizeof(double) * (8 * (size_t)g->nbins + 3 * (size_t)g->d.ntheta);
}

int abt_finalise_dev(const abt_dev_grid* g, double* ab, double* ae, int scaling, double ebeam,
                     double eext, double* dev_mueller, double* dev_m1d, double* dev_params,
                     void* dev_work, size_t work_bytes, void* stream)
{
    if (!g || !ab || !ae || !dev_mueller || !dev_m1d || !dev_params)
        return fail(ABT_ERR_ARG, "abt_finalise_dev: null argument");
    if (g->d.nphi < 3 || g->d.ntheta < 3)
        return fail(ABT_ERR_ARG, "finalise needs at least 3 theta and 3 phi values (simpne)");
    if (scaling && !(ebeam > 0 && eext > 0))
        return fail(ABT_ERR_ARG, "energy scaling needs positive beam and ext energies");
    int rc = ensure_device();
    if (rc) return rc;
    const size_t need = abt_finalise_workspace(g);
    void* work = dev_work;
    if (!work || work_bytes < need) {
        if (work) return fail(ABT_ERR_ARG, "finalise workspace of %zu bytes is smaller than %zu", work_bytes, need);
        ThreadState& ts = tstate();
        rc = buf_ensure(ts.bufs[6], need);
        if (rc) return rc;
        work = ts.bufs[6].p;
    }
    double* tot = (double*)work;
    double* tmp = tot + 8 * g->nbins;
    int e = abt_launch_finalise(&g->d, ab, ae, tot, dev_mueller, dev_m1d, tmp, dev_params, scaling,
                                ebeam, eext, (hipStream_t)stream);
    if (e) return fail(ABT_ERR_HIP, "finalise launch failed: %s", hipGetErrorString((hipError_t)e));
    return 0;
}

int abt_finalise(const abt_grid* grid, double lambda, double* ampl_far_beam, double* ampl_far_ext_diff,
                 int scaling, double beam_energy_out, double ext_energy_out, double* mueller,
                 double* mueller_beam, double* mueller_ext, double* mueller_1d,
                 double* mueller_1d_beam, double* mueller_1d_ext, double* params)
{
    if (!grid || !ampl_far_beam || !ampl_far_ext_diff || !grid->theta || !grid->phi)
        return fail(ABT_ERR_ARG, "abt_finalise: null argument");
    abt_dev_grid* g = nullptr;
    bool priv = false;
    int rc = grid_acquire(grid->n_theta, grid->theta, grid->n_phi, grid->phi, nullptr, nullptr, nullptr, lambda, &g,
                          &priv);
    if (rc) return rc;
    ThreadState& ts = tstate();
    rc = stream_of(ts);
    const int64_t nb = g->nbins, nth = g->d.ntheta;
    const size_t ob = sizeof(double) * 8 * nb;
    const size_t mb = sizeof(double) * 16 * nb, m1b = sizeof(double) * 16 * nth;
    if (!rc) rc = buf_ensure(ts.bufs[0], 2 * ob + 3 * mb + 3 * m1b + 16 * sizeof(double));
    if (!rc) {
        double* dab = (double*)ts.bufs[0].p;
        double* dae = dab + 8 * nb;
        double* dM = dae + 8 * nb;
        double* dm1 = dM + 48 * nb;
        double* dpar = dm1 + 48 * nth;
        hipStream_t st = ts.st;
        hipError_t e = hipMemcpyAsync(dab, ampl_far_beam, ob, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dae, ampl_far_ext_diff, ob, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) rc = fail(ABT_ERR_HIP, "upload: %s", hipGetErrorString(e));
        if (!rc) rc = abt_finalise_dev(g, dab, dae, scaling, beam_energy_out, ext_energy_out, dM, dm1,
                                       dpar, nullptr, 0, st);
        if (!rc) {
            struct { void* h; const void* d; size_t n; } cp[] = {
                {ampl_far_beam, dab, ob}, {ampl_far_ext_diff, dae, ob},
                {mueller, dM, mb}, {mueller_beam, dM + 16 * nb, mb}, {mueller_ext, dM + 32 * nb, mb},
                {mueller_1d, dm1, m1b}, {mueller_1d_beam, dm1 + 16 * nth, m1b},
                {mueller_1d_ext, dm1 + 32 * nth, m1b}, {params, dpar, 9 * sizeof(double)}};
            for (auto& c : cp) {
                if (!c.h) continue;
                e = hipMemcpyAsync(c.h, c.d, c.n, hipMemcpyDeviceToHost, st);
                if (e != hipSuccess) break;
            }
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = fail(ABT_ERR_HIP, "copy-back: %s", hipGetErrorString(e));
        }
    }
    grid_release(g, priv);
    return rc;
}

}  // extern "C"
