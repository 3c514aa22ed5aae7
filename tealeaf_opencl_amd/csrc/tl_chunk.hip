// Host side of the chunk: allocation, field transfer, the reference-granularity
// operations and the Fortran-callable drop-in entry points.
//
// Each tl_* operation enqueues the kernels of one reference entry point on the
// chunk's stream; operations that return a scalar wait for it (a pinned
// 32-byte read), the others return at once (in-order stream semantics, as the
// reference's single in-order queue, ocl_init.cpp:424-434).
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tealeaf_mi355x.h"
#include "tl_chunk.h"
#include "tl_deck.h"
#include "tl_kernels.h"

using namespace tl;

static thread_local char g_err[1024];
static tl_chunk *g_default = nullptr;

namespace tl {

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(const char *what, hipError_t e)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return TL_ERR_HIP;
}

Ext make_ext(int j0, int j1, int k0, int k1)
{
    Ext e;
    e.j0 = j0;
    e.j1 = j1;
    e.k0 = k0;
    e.k1 = k1;
    const int nx = j1 - j0 + 1, ny = k1 - k0 + 1;
    e.ntx = nx > 0 ? (nx + TILE_X - 1) / TILE_X : 0;
    e.nty = ny > 0 ? (ny + TILE_Y - 1) / TILE_Y : 0;
    return e;
}

// a kernel compiled with launch bounds (xmin,xmax,ymin,ymax) covers columns
// [hd+xmin, x_max+hd-1+xmax] (findPaddingSize + WITHIN_BOUNDS)
Ext ext_range(const tl_chunk *c, int xmin, int xmax, int ymin, int ymax)
{
    const int hd = c->g.hd;
    return make_ext(hd + xmin, c->g.x_max + hd - 1 + xmax, hd + ymin, c->g.y_max + hd - 1 + ymax);
}

RedArgs red_args(tl_chunk *c, int mode, bool with_state)
{
    RedArgs ra;
    ra.partials = c->partials;
    // without a device state the totals go to the host (read_result)
    ra.result = with_state ? c->result : c->d_result;
    ra.flag = with_state ? nullptr : c->d_flag;
    ra.seq = with_state ? 0u : ++c->seq;
    ra.st = with_state ? c->st : nullptr;
    ra.hist_a = with_state ? c->hist_a : nullptr;
    ra.hist_b = with_state ? c->hist_b : nullptr;
    ra.hist_e = with_state ? c->hist_e : nullptr;
    ra.mode = mode;
    return ra;
}

void prof_begin(tl_chunk *c)
{
    if (c->profiler_on) (void)hipEventRecord(c->pe0, c->s);
}

void prof_add(tl_chunk *c, const char *name, double ms)
{
    auto &e = c->prof[name];
    e.first += ms;
    e.second += 1;
}

void prof_end_split(tl_chunk *c, const char *name, const char *second)
{
    if (!c->profiler_on) return;
    (void)hipEventRecord(c->pe1, c->s);
    (void)hipEventSynchronize(c->pe1);
    float a = 0.f, b = 0.f;
    (void)hipEventElapsedTime(&a, c->pe0, c->pe_mid);
    (void)hipEventElapsedTime(&b, c->pe_mid, c->pe1);
    prof_add(c, name, a);
    prof_add(c, second, b);
}

void prof_end(tl_chunk *c, const char *name)
{
    if (!c->profiler_on) return;
    (void)hipEventRecord(c->pe1, c->s);
    (void)hipEventSynchronize(c->pe1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->pe0, c->pe1);
    auto &e = c->prof[name];
    e.first += ms;
    e.second += 1;
}

int ensure_coef_capacity(tl_chunk *c, int n)
{
    if (n <= c->ch_cap) return TL_OK;
    const int cap = std::max(n, 64);
    hipError_t e;
    if (c->ch_a) (void)hipFree(c->ch_a);
    if (c->ch_b) (void)hipFree(c->ch_b);
    c->ch_a = c->ch_b = nullptr;
    if ((e = hipMalloc(&c->ch_a, sizeof(double) * cap)) != hipSuccess) return hip_fail("hipMalloc coefs", e);
    if ((e = hipMalloc(&c->ch_b, sizeof(double) * cap)) != hipSuccess) return hip_fail("hipMalloc coefs", e);
    c->ch_cap = cap;
    // captured batches hold the old coefficient pointers
    for (auto *gs : {&c->g_cheby, &c->g_ppcg})
        if (gs->exec) {
            (void)hipGraphExecDestroy(gs->exec);
            gs->exec = nullptr;
            gs->key = -1;
        }
    return TL_OK;
}

void drop_graphs(tl_chunk *c)
{
    tl_chunk::GraphSlot *slots[] = {&c->g_cg, &c->g_cheby, &c->g_ppcg, &c->g_jacobi, &c->rf[0].graph,
                                    &c->rf[0].graph_ppcg, &c->rf[1].graph, &c->rf[1].graph_ppcg};
    for (auto *gs : slots)
        if (gs->exec) {
            (void)hipGraphExecDestroy(gs->exec);
            gs->exec = nullptr;
            gs->key = -1;
        }
}

int capture_graph(tl_chunk *c, tl_chunk::GraphSlot &gs, int key, const std::function<void()> &body)
{
    if (gs.exec && gs.key == key) return TL_OK;
    if (gs.exec) {
        (void)hipGraphExecDestroy(gs.exec);
        gs.exec = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail("hipStreamBeginCapture", e);
    body();
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(c->s, &graph);
    if (e != hipSuccess) return hip_fail("hipStreamEndCapture", e);
    e = hipGraphInstantiate(&gs.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        gs.exec = nullptr;
        return hip_fail("hipGraphInstantiate", e);
    }
    gs.key = key;
    return TL_OK;
}

int ensure_hist_capacity(tl_chunk *c, int n)
{
    if (n <= c->hist_len) return TL_OK;
    const int cap = std::max(n, 1024);
    hipError_t e;
    for (double **p : {&c->hist_a, &c->hist_b, &c->hist_e}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        if ((e = hipMalloc(p, sizeof(double) * cap)) != hipSuccess) return hip_fail("hipMalloc history", e);
        TL_HIP_TRY(hipMemsetAsync(*p, 0, sizeof(double) * cap, c->s), "hipMemsetAsync history");
    }
    c->hist_len = cap;
    for (auto *gs : {&c->g_cg, &c->g_cheby, &c->g_ppcg, &c->g_jacobi})
        if (gs->exec) {
            (void)hipGraphExecDestroy(gs->exec);
            gs->exec = nullptr;
            gs->key = -1;
        }
    return TL_OK;
}

// Totals of the latest reduction launched with red_args(c, mode, false): its
// finisher wrote them to pinned host memory and then stored its sequence
// number into the pinned flag, so the host polls the flag instead of
// copying and synchronising (the round trip of each blocking scalar the
// reference's Fortran driver asks for). After 20 ms without the flag (a long
// kernel queue ahead, or an error) it waits for the stream, which also
// reports any error.
int read_result(tl_chunk *c, double *out, int nq)
{
    const unsigned int want = c->seq;
    bool seen = __atomic_load_n(c->h_flag, __ATOMIC_ACQUIRE) == want;
    if (!seen) {
        const auto t0 = std::chrono::steady_clock::now();
        for (long spin = 1; !seen; ++spin) {
            seen = __atomic_load_n(c->h_flag, __ATOMIC_ACQUIRE) == want;
            if ((spin & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
        }
    }
    if (!seen) {
        hipError_t e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) return hip_fail("hipStreamSynchronize", e);
    }
    for (int q = 0; q < nq; ++q) out[q] = c->h_result[q];
    return TL_OK;
}

}  // namespace tl

#define TL_CHECK_LAUNCH(c, name)                                                  \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) return hip_fail("launch " name, e_);                \
    } while (0)

// ---------------------------------------------------------------------------
// chunk lifecycle
// ---------------------------------------------------------------------------
extern "C" const char *tl_last_error(void) { return g_err; }

extern "C" int tl_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail("hipGetDeviceCount", e);
    }
    *count = n;
    return TL_OK;
}

extern "C" int tl_chunk_destroy(tl_chunk *c)
{
    if (!c) return TL_OK;
    // teardown: every resource is released whatever the earlier calls return
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (auto *gs : {&c->g_cg, &c->g_cheby, &c->g_ppcg, &c->g_jacobi})
        if (gs->exec) (void)hipGraphExecDestroy(gs->exec);
    refine_release(c);
    for (int i = 0; i < NUM_FIELDS_2D; ++i)
        if (c->base2[i]) (void)hipFree(c->base2[i]);
    for (int i = 0; i < NUM_FIELDS_1D; ++i)
        if (c->f1[i]) (void)hipFree(c->f1[i]);
    for (void *p : {(void *)c->partials, (void *)c->result, (void *)c->st, (void *)c->range_bad,
                    (void *)c->hist_a, (void *)c->hist_b, (void *)c->hist_e, (void *)c->ch_a,
                    (void *)c->ch_b})
        if (p) (void)hipFree(p);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    if (c->h_st) (void)hipHostFree(c->h_st);
    for (hipEvent_t ev : {c->pe0, c->pe1, c->pe_mid, c->ev_batch[0], c->ev_batch[1]})
        if (ev) (void)hipEventDestroy(ev);
    for (auto &set : c->ev_t)
        for (hipEvent_t ev : set)
            if (ev) (void)hipEventDestroy(ev);
    if (c->s) (void)hipStreamDestroy(c->s);
    if (g_default == c) g_default = nullptr;
    delete c;
    return TL_OK;
}

// Process-wide default of the tuning keys (tl_set_tuning with no chunk),
// taken by chunks created later.
static int g_ppcg_steps_default = 2, g_cheby_steps_default = 2;

extern "C" int tl_chunk_create(int x_max, int y_max, int halo_depth, int preconditioner, int profiler_on,
                               tl_chunk **out)
{
    *out = nullptr;
    if (x_max < 1 || y_max < 1 || halo_depth < 1 || preconditioner < PREC_NONE || preconditioner > PREC_BLOCK) {
        set_error("tl_chunk_create: invalid size %dx%d, halo_depth %d or preconditioner %d", x_max, y_max,
                  halo_depth, preconditioner);
        return TL_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("tl_chunk_create: no HIP device available");
        return TL_ERR_HIP;
    }
    tl_chunk *c = new tl_chunk();
    c->g.x_max = x_max;
    c->g.y_max = y_max;
    c->g.hd = halo_depth;
    if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
    c->prec = preconditioner;
    c->ppcg_steps_per_launch = g_ppcg_steps_default;
    c->cheby_steps_per_launch = g_cheby_steps_default;
    c->profiler_on = profiler_on;
    // pitch: padded row plus one slack column, rounded to 16 doubles (128 B)
    const int w = x_max + 2 * halo_depth + 2;
    c->g.pitch = (w + 15) / 16 * 16;
    const int margin_rows = 8;
    const long colpad = (16 - (halo_depth % 16)) % 16;
    c->off = (long)margin_rows * c->g.pitch + colpad;
    const long rows = (long)y_max + 2 * halo_depth + 2 * margin_rows;
    c->elems2d = rows * c->g.pitch + colpad + 64;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking)) != hipSuccess) {
        tl_chunk_destroy(c);
        return hip_fail("hipStreamCreate", e);
    }
    for (int i = 0; i < NUM_FIELDS_2D; ++i) {
        if ((e = hipMalloc(&c->base2[i], sizeof(double) * c->elems2d)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipMalloc field", e);
        }
        // every buffer starts at zero, halos included (ocl_buffers.cpp:3-19)
        if ((e = hipMemsetAsync(c->base2[i], 0, sizeof(double) * c->elems2d, c->s)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipMemsetAsync field", e);
        }
        c->f2[i] = c->base2[i] + c->off;
    }
    c->len1d = std::max(x_max, y_max) + 2 * halo_depth + 8;
    for (int i = 0; i < NUM_FIELDS_1D; ++i) {
        if ((e = hipMalloc(&c->f1[i], sizeof(double) * c->len1d)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipMalloc 1d", e);
        }
        if ((e = hipMemsetAsync(c->f1[i], 0, sizeof(double) * c->len1d, c->s)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipMemsetAsync 1d", e);
        }
    }
    if ((e = hipMalloc(&c->partials, sizeof(double) * MAX_RED * c->cap)) != hipSuccess ||
        (e = hipMalloc(&c->result, sizeof(double) * MAX_RED)) != hipSuccess ||
        (e = hipMalloc(&c->st, sizeof(SolveState))) != hipSuccess ||
        (e = hipMalloc(&c->range_bad, sizeof(int))) != hipSuccess ||
        (e = hipHostMalloc(&c->h_result, sizeof(double) * MAX_RED, hipHostMallocCoherent)) != hipSuccess ||
        (e = hipHostMalloc(&c->h_flag, 64, hipHostMallocCoherent)) != hipSuccess ||
        (e = hipHostGetDevicePointer((void **)&c->d_result, c->h_result, 0)) != hipSuccess ||
        (e = hipHostGetDevicePointer((void **)&c->d_flag, c->h_flag, 0)) != hipSuccess ||
        (e = hipHostMalloc(&c->h_st, sizeof(SolveState) * 2, hipHostMallocDefault)) != hipSuccess) {
        tl_chunk_destroy(c);
        return hip_fail("hipMalloc scratch", e);
    }
    if ((e = hipMemsetAsync(c->st, 0, sizeof(SolveState), c->s)) != hipSuccess) {
        tl_chunk_destroy(c);
        return hip_fail("hipMemsetAsync state", e);
    }
    *c->h_flag = 0u;
    for (hipEvent_t *ev : {&c->pe0, &c->pe1, &c->pe_mid})
        if ((e = hipEventCreate(ev)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipEventCreate", e);
        }
    for (auto &set : c->ev_t)
        for (hipEvent_t &ev : set)
            if ((e = hipEventCreate(&ev)) != hipSuccess) {
                tl_chunk_destroy(c);
                return hip_fail("hipEventCreate", e);
            }
    for (hipEvent_t *ev : {&c->ev_batch[0], &c->ev_batch[1]})
        if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) {
            tl_chunk_destroy(c);
            return hip_fail("hipEventCreate", e);
        }
    if (ensure_hist_capacity(c, 1024) != TL_OK) {
        tl_chunk_destroy(c);
        return TL_ERR_HIP;
    }
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) {
        tl_chunk_destroy(c);
        return hip_fail("hipStreamSynchronize", e);
    }
    *out = c;
    return TL_OK;
}

extern "C" int tl_chunk_set_default(tl_chunk *c)
{
    g_default = c;
    return TL_OK;
}

extern "C" tl_chunk *tl_chunk_default(void) { return g_default; }

extern "C" int tl_chunk_info(const tl_chunk *c, int *x_max, int *y_max, int *halo_depth, int *preconditioner)
{
    if (!c) return TL_ERR_ARG;
    *x_max = c->g.x_max;
    *y_max = c->g.y_max;
    *halo_depth = c->g.hd;
    *preconditioner = c->prec;
    return TL_OK;
}

extern "C" int tl_synchronize(tl_chunk *c)
{
    hipError_t e = hipStreamSynchronize(c->s);
    return e == hipSuccess ? TL_OK : hip_fail("hipStreamSynchronize", e);
}

// ---------------------------------------------------------------------------
// field transfer (reference layout on the host)
// ---------------------------------------------------------------------------
extern "C" int tl_field_get(tl_chunk *c, int field, double *host)
{
    if (!c || field < 0 || field >= NUM_PUBLIC_2D || !host) {
        set_error("tl_field_get: bad field %d", field);
        return TL_ERR_ARG;
    }
    const int w = c->g.x_max + 2 * c->g.hd, h = c->g.y_max + 2 * c->g.hd;
    hipError_t e = hipMemcpy2DAsync(host, sizeof(double) * w, c->f2[field], sizeof(double) * c->g.pitch,
                                    sizeof(double) * w, h, hipMemcpyDeviceToHost, c->s);
    if (e != hipSuccess) return hip_fail("hipMemcpy2D", e);
    return tl_synchronize(c);
}

extern "C" int tl_field_set(tl_chunk *c, int field, const double *host)
{
    if (!c || field < 0 || field >= NUM_PUBLIC_2D || !host) {
        set_error("tl_field_set: bad field %d", field);
        return TL_ERR_ARG;
    }
    const int w = c->g.x_max + 2 * c->g.hd, h = c->g.y_max + 2 * c->g.hd;
    if (field == F_KX || field == F_KY || field == F_CP || field == F_BFP) c->factors_current = false;
    if (field == F_KX || field == F_KY) c->faces_current = false;
    hipError_t e = hipMemcpy2DAsync(c->f2[field], sizeof(double) * c->g.pitch, host, sizeof(double) * w,
                                    sizeof(double) * w, h, hipMemcpyHostToDevice, c->s);
    if (e != hipSuccess) return hip_fail("hipMemcpy2D", e);
    return tl_synchronize(c);
}

extern "C" int tl_field1d_get(tl_chunk *c, int field, double *host, int n)
{
    if (!c || field < 0 || field >= NUM_FIELDS_1D || !host || n < 0 || n > c->len1d) {
        set_error("tl_field1d_get: bad field %d or length %d", field, n);
        return TL_ERR_ARG;
    }
    hipError_t e = hipMemcpyAsync(host, c->f1[field], sizeof(double) * n, hipMemcpyDeviceToHost, c->s);
    if (e != hipSuccess) return hip_fail("hipMemcpy", e);
    return tl_synchronize(c);
}

// ---------------------------------------------------------------------------
// reference-granularity operations
// ---------------------------------------------------------------------------

// initialise_chunk_kernel_ocl.cpp:9-27
extern "C" int tl_initialise_chunk(tl_chunk *c, double xmin, double ymin, double dx, double dy)
{
    c->dx = dx;
    c->dy = dy;
    prof_begin(c);
    launch_init_chunk_1d(c->s, c->g, xmin, ymin, dx, dy, c->f1);
    TL_CHECK_LAUNCH(c, "initialise_chunk_first");
    prof_end(c, "initialise_chunk_first");
    prof_begin(c);
    launch_init_chunk_2d(c->s, c->g, ext_range(c, -1, 1, -1, 1), c->cap, dx, dy, c->f2[F_VOLUME],
                         c->f2[F_XAREA], c->f2[F_YAREA]);
    TL_CHECK_LAUNCH(c, "initialise_chunk_second");
    prof_end(c, "initialise_chunk_second");
    return tl_synchronize(c);
}

// generate_chunk_kernel_ocl.cpp:25-91 (all states painted in one pass)
namespace {

// the state painting of generate_chunk_kernel_ocl.cpp:25-91 from state arrays
// already on the device (dev: density, energy, xmin, xmax, ymin, ymax, radius,
// nstates each), then the wait the reference ends with
int generate_staged(tl_chunk *c, int nstates, const double *dev, const int *dgeom, int g_rect, int g_circ,
                    int g_point)
{
    const double *sarr[7];
    for (int a = 0; a < 7; ++a) sarr[a] = dev + a * nstates;
    prof_begin(c);
    launch_generate(c->s, c->g, ext_interior(c), c->cap, c->f1, c->f2[F_DENSITY], c->f2[F_ENERGY0], c->f2[F_U],
                    c->f2[F_U0], nstates, sarr, dgeom, g_rect, g_circ, g_point);
    TL_CHECK_LAUNCH(c, "generate_chunk");
    prof_end(c, "generate_chunk");
    return tl_synchronize(c);
}

}  // namespace

extern "C" int tl_generate_chunk(tl_chunk *c, int nstates, const double *density, const double *energy,
                                 const double *xmin, const double *xmax, const double *ymin,
                                 const double *ymax, const double *radius, const int *geometry,
                                 int g_rect, int g_circ, int g_point)
{
    if (nstates < 1) {
        set_error("generate_chunk: no states");
        return TL_ERR_ARG;
    }
    const double *src[7] = {density, energy, xmin, xmax, ymin, ymax, radius};
    double *dev = nullptr;
    int *dgeom = nullptr;
    hipError_t e;
    if ((e = hipMalloc(&dev, sizeof(double) * 7 * nstates)) != hipSuccess) return hip_fail("hipMalloc states", e);
    if ((e = hipMalloc(&dgeom, sizeof(int) * nstates)) != hipSuccess) {
        (void)hipFree(dev);
        return hip_fail("hipMalloc states", e);
    }
    for (int a = 0; a < 7 && e == hipSuccess; ++a)
        e = hipMemcpyAsync(dev + a * nstates, src[a], sizeof(double) * nstates, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(dgeom, geometry, sizeof(int) * nstates, hipMemcpyHostToDevice, c->s);
    const int rc = e == hipSuccess ? generate_staged(c, nstates, dev, dgeom, g_rect, g_circ, g_point)
                                   : hip_fail("generate_chunk: state copy", e);
    // the copies and the kernel are done before the arrays are freed
    // (hipFree waits for the device)
    (void)hipFree(dev);
    (void)hipFree(dgeom);
    return rc;
}

// ---------------------------------------------------------------------------
// generate_chunk without allocation: the state arrays are staged in `scratch`,
// a file-scope device array (one copy per device, held until the process
// ends), instead of device buffers allocated and freed per call. Calls that
// overlap on one device share it and race.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_DOUBLES = 7 * (size_t)TL_NOALLOC_MAX_STATES;
constexpr size_t NOALLOC_BYTES = sizeof(double) * NOALLOC_DOUBLES + sizeof(int) * TL_NOALLOC_MAX_STATES;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

// device address of `scratch` on each device, looked up once per device with
// that device current (in host code the array's name is only a stand-in)
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_of(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("tl_generate_chunk_noalloc: device %d outside 0..%d", dev, NOALLOC_MAX_DEVICES - 1);
        return TL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return hip_fail("hipGetDevice", e);
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return hip_fail("hipSetDevice", e);
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return hip_fail("hipGetSymbolAddress(scratch)", e);
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return TL_OK;
}

}  // namespace

extern "C" int tl_generate_chunk_noalloc(tl_chunk *c, int nstates, const double *density, const double *energy,
                                         const double *xmin, const double *xmax, const double *ymin,
                                         const double *ymax, const double *radius, const int *geometry,
                                         int g_rect, int g_circ, int g_point)
{
    if (!c || nstates < 1 || nstates > TL_NOALLOC_MAX_STATES) {
        set_error("tl_generate_chunk_noalloc: %d states outside 1..%d", nstates, TL_NOALLOC_MAX_STATES);
        return TL_ERR_ARG;
    }
    unsigned char *ptr = nullptr;
    const int rs = scratch_of(c->device, &ptr);
    if (rs != TL_OK) return rs;
    double *dev = reinterpret_cast<double *>(ptr);
    int *dgeom = reinterpret_cast<int *>(ptr + sizeof(double) * NOALLOC_DOUBLES);
    const double *src[7] = {density, energy, xmin, xmax, ymin, ymax, radius};
    hipError_t e = hipSuccess;
    for (int a = 0; a < 7 && e == hipSuccess; ++a)
        e = hipMemcpyAsync(dev + a * nstates, src[a], sizeof(double) * nstates, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(dgeom, geometry, sizeof(int) * nstates, hipMemcpyHostToDevice, c->s);
    if (e != hipSuccess) return hip_fail("tl_generate_chunk_noalloc: state copy", e);
    return generate_staged(c, nstates, dev, dgeom, g_rect, g_circ, g_point);
}

// set_field_kernel_ocl.cpp:3-7
extern "C" int tl_set_field(tl_chunk *c)
{
    prof_begin(c);
    launch_copy(c->s, c->g, ext_interior(c), c->cap, c->f2[F_ENERGY0], c->f2[F_ENERGY1]);
    TL_CHECK_LAUNCH(c, "set_field");
    prof_end(c, "set_field");
    return TL_OK;
}

// field_summary_kernel_ocl.cpp:10-20
extern "C" int tl_field_summary(tl_chunk *c, double *vol, double *mass, double *ie, double *temp)
{
    prof_begin(c);
    launch_field_summary(c->s, c->g, ext_interior(c), c->cap, c->f2[F_VOLUME], c->f2[F_DENSITY],
                         c->f2[F_ENERGY1], c->f2[F_U], red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "field_summary");
    prof_end(c, "field_summary");
    double t[4];
    int rc = read_result(c, t, 4);
    *vol = t[0];
    *mass = t[1];
    *ie = t[2];
    *temp = t[3];
    return rc;
}

// update_halo_kernel_ocl.cpp:14-62: density, energy0, energy1 only; faces in
// the order bottom, top, left, right
extern "C" int tl_update_halo(tl_chunk *c, const int *chunk_neighbours, const int *fields, int depth)
{
    if (depth < 1) return TL_OK;
    if (depth > c->g.hd) depth = c->g.hd;  // the allocated halo bounds every fill
    const int ids[3] = {F_DENSITY, F_ENERGY0, F_ENERGY1};
    for (int q = 0; q < 3; ++q) {
        if (fields[q] != 1) continue;
        double *a = c->f2[ids[q]];
        const int order[4] = {2, 3, 0, 1};
        for (int f : order) {
            // chunk_neighbours: 1 left, 2 right, 3 bottom, 4 top (0-based here)
            if (chunk_neighbours[f] != -1) continue;
            prof_begin(c);
            launch_halo(c->s, c->g, a, depth, f);
            TL_CHECK_LAUNCH(c, "update_halo");
            prof_end(c, f == 0 ? "update_halo_left" : f == 1 ? "update_halo_right"
                        : f == 2 ? "update_halo_bottom" : "update_halo_top");
        }
    }
    return TL_OK;
}

// tea_leaf_kernel_ocl.cpp:8-32 (calcrxry) and :254-289
extern "C" int tl_common_init(tl_chunk *c, int coefficient, double dt, double *rx, double *ry,
                              const int *zero_boundary, int reflective_boundary)
{
    if (coefficient != COEF_CONDUCTIVITY && coefficient != COEF_RECIP_CONDUCTIVITY) {
        set_error("Unknown coefficient %d passed to tea leaf", coefficient);
        return TL_ERR_ARG;
    }
    // celldx/celldy[hd+1]: every entry holds dx, kept on the host since
    // initialise_chunk (the reference reads it back each time)
    *rx = dt / (c->dx * c->dx);
    *ry = dt / (c->dy * c->dy);
    c->rx = *rx;
    c->ry = *ry;
    const int hd = c->g.hd;
    prof_begin(c);
    launch_init_common(c->s, c->g, ext_range(c, 1 - hd, hd, 1 - hd, hd), c->cap, c->f2[F_DENSITY], c->f2[F_KX],
                       c->f2[F_KY], *rx, *ry, coefficient);
    TL_CHECK_LAUNCH(c, "tea_leaf_init_common");
    prof_end(c, "tea_leaf_init_common");
    // the same coefficients in recomputable form for the fused kernels
    c->factors_current = false;
    TL_HIP_TRY(hipMemsetAsync(c->range_bad, 0, sizeof(int), c->s), "tl_common_init");
    launch_conductivity(c->s, c->g, c->f2[F_DENSITY], c->f2[F_COND], coefficient, c->range_bad);
    TL_CHECK_LAUNCH(c, "conductivity");
    // the iteration kernels' division (div_nr) is the IEEE quotient for the
    // divisors that densities in [2^-499, 2^499] give; others are refused
    // (the reference would go on with infinite or NaN coefficients)
    {
        int bad = 0;
        hipError_t e = hipMemcpyAsync(&bad, c->range_bad, sizeof(int), hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) return hip_fail("tl_common_init", e);
        if (bad) {
            set_error("tl_common_init: a density outside [2^-499, 2^499] or not a number: the iteration kernels' "
                      "division is exact only for the divisors such densities give");
            return TL_ERR_ARG;
        }
    }
    Faces nf;
    nf.rx = *rx;
    nf.ry = *ry;
    nf.zl = !reflective_boundary && zero_boundary[0];
    nf.zr = !reflective_boundary && zero_boundary[1];
    nf.zb = !reflective_boundary && zero_boundary[2];
    nf.zt = !reflective_boundary && zero_boundary[3];
    // the captured batches were recorded with the previous faces as kernel
    // arguments: replaying them after a change of dt or walls would apply the
    // old operator
    if (memcmp(&nf.rx, &c->faces.rx, sizeof(double)) != 0 || memcmp(&nf.ry, &c->faces.ry, sizeof(double)) != 0 ||
        nf.zl != c->faces.zl || nf.zr != c->faces.zr || nf.zb != c->faces.zb || nf.zt != c->faces.zt) {
        hipError_t e = hipStreamSynchronize(c->s);
        if (e != hipSuccess) return hip_fail("tl_common_init", e);
        drop_graphs(c);
    }
    c->faces = nf;
    if (!reflective_boundary) {
        prof_begin(c);
        launch_zero_boundary(c->s, c->g, ext_range(c, -hd, hd, -hd, hd), c->cap, c->f2[F_KX], c->f2[F_KY],
                             zero_boundary);
        TL_CHECK_LAUNCH(c, "tea_leaf_zero_boundary");
        prof_end(c, "tea_leaf_zero_boundary");
    }
    prof_begin(c);
    launch_init_u(c->s, c->g, ext_interior(c), c->cap, c->f2[F_DENSITY], c->f2[F_ENERGY1], c->f2[F_U],
                  c->f2[F_U0]);
    TL_CHECK_LAUNCH(c, "generate_chunk_init_u");
    prof_end(c, "generate_chunk_init_u");
    c->faces_current = true;
    return TL_OK;
}

extern "C" int tl_calc_residual(tl_chunk *c)
{
    prof_begin(c);
    launch_calc_residual(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U], c->f2[F_U0], c->f2[F_R], c->f2[F_KX],
                         c->f2[F_KY]);
    TL_CHECK_LAUNCH(c, "tea_leaf_calc_residual");
    prof_end(c, "tea_leaf_calc_residual");
    return TL_OK;
}

// tea_leaf_kernel_ocl.cpp:56-96
extern "C" int tl_calc_2norm(tl_chunk *c, int norm_array, double *norm)
{
    const double *a, *b;
    if (norm_array == 0) {
        a = b = c->f2[F_U0];
    } else if (norm_array == 1) {
        a = b = c->f2[F_R];
    } else if (norm_array == 2) {
        a = c->f2[F_R];
        b = (c->prec == PREC_NONE) ? c->f2[F_R] : c->f2[F_Z];
    } else {
        set_error("Invalid value '%d' for norm_array passed, should be 0 for u0, 1 for r, 2 for r*z", norm_array);
        return TL_ERR_ARG;
    }
    prof_begin(c);
    launch_dot(c->s, c->g, ext_interior(c), c->cap, a, b, red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_calc_2norm");
    prof_end(c, "tea_leaf_calc_2norm");
    return read_result(c, norm, 1);
}

extern "C" int tl_finalise(tl_chunk *c)
{
    prof_begin(c);
    launch_finalise(c->s, c->g, ext_interior(c), c->cap, c->f2[F_DENSITY], c->f2[F_U], c->f2[F_ENERGY1]);
    TL_CHECK_LAUNCH(c, "tea_leaf_finalise");
    prof_end(c, "tea_leaf_finalise");
    return TL_OK;
}

extern "C" int tl_block_init(tl_chunk *c)
{
    prof_begin(c);
    launch_block_init(c->s, c->g, ext_interior(c), c->cap, c->f2[F_CP], c->f2[F_BFP], c->f2[F_KX], c->f2[F_KY]);
    TL_CHECK_LAUNCH(c, "tea_leaf_block_init");
    c->factors_current = true;
    prof_end(c, "tea_leaf_block_init");
    return TL_OK;
}

extern "C" int tl_block_solve(tl_chunk *c)
{
    prof_begin(c);
    launch_block_solve(c->s, c->g, ext_interior(c), c->cap, c->f2[F_R], c->f2[F_Z], c->f2[F_CP], c->f2[F_BFP],
                       c->f2[F_KX], c->f2[F_KY]);
    TL_CHECK_LAUNCH(c, "tea_leaf_block_solve");
    prof_end(c, "tea_leaf_block_solve");
    return TL_OK;
}

int tl::precondition_r(tl_chunk *c)
{
    if (c->prec == PREC_BLOCK) return tl_block_solve(c);
    if (c->prec == PREC_DIAG) {
        launch_diag_apply(c->s, c->g, ext_interior(c), c->cap, c->f2[F_R], c->f2[F_MI], c->f2[F_Z]);
        TL_CHECK_LAUNCH(c, "diag_apply");
    }
    return TL_OK;
}

extern "C" int tl_init_jac_diag(tl_chunk *c)
{
    const int hd = c->g.hd;
    prof_begin(c);
    launch_init_jac_diag(c->s, c->g, ext_range(c, -hd, hd, -hd, hd), c->cap, c->f2[F_MI], c->f2[F_KX],
                         c->f2[F_KY]);
    TL_CHECK_LAUNCH(c, "tea_leaf_init_jac_diag");
    prof_end(c, "tea_leaf_init_jac_diag");
    return TL_OK;
}

// preconditioner setup + init_p; mode selects host-returned or device-state rro
static int cg_init_impl(tl_chunk *c, int mode, bool with_state)
{
    int rc = TL_OK;
    if (c->prec == PREC_BLOCK) {
        if ((rc = tl_block_init(c)) != TL_OK) return rc;
        if ((rc = tl_block_solve(c)) != TL_OK) return rc;
    } else if (c->prec == PREC_DIAG) {
        if ((rc = tl_init_jac_diag(c)) != TL_OK) return rc;
    }
    prof_begin(c);
    launch_cg_init_p(c->s, c->g, ext_interior(c), c->cap, c->prec, c->f2[F_P], c->f2[F_R], c->f2[F_Z], c->f2[F_MI],
                     red_args(c, mode, with_state));
    TL_CHECK_LAUNCH(c, "tea_leaf_cg_solve_init_p");
    prof_end(c, "tea_leaf_cg_solve_init_p");
    return TL_OK;
}

int tl_cg_init_device(tl_chunk *c) { return cg_init_impl(c, RED_INIT_RRO, true); }

// tea_leaf_kernel_ocl.cpp:167-187
extern "C" int tl_cg_init(tl_chunk *c, double *rro)
{
    int rc = cg_init_impl(c, RED_STORE, false);
    if (rc != TL_OK) return rc;
    return read_result(c, rro, 1);
}

// block-Jacobi factors for the per-iteration drop-in kernels: null (recompute
// from Kx, Ky in registers, bitwise the stored values) while the stored ones
// belong to the current Kx, Ky; the stored arrays otherwise, as the
// reference reads them (tea_block_jacobi.cl:44-76)
static const double *stored_cp(tl_chunk *c) { return c->factors_current ? nullptr : c->f2[F_CP]; }
static const double *stored_bfp(tl_chunk *c) { return c->factors_current ? nullptr : c->f2[F_BFP]; }
// conductivity for the drop-in CG kernels while Kx, Ky are its faces, else null
static const double *faces_cond(tl_chunk *c) { return c->faces_current ? c->f2[F_COND] : nullptr; }

extern "C" int tl_cg_calc_w(tl_chunk *c, double *pw)
{
    prof_begin(c);
    launch_cg_calc_w(c->s, c->g, ext_interior(c), c->cap, c->f2[F_P], c->f2[F_W], c->f2[F_KX], c->f2[F_KY],
                     faces_cond(c), c->faces, red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_cg_solve_calc_w");
    prof_end(c, "tea_leaf_cg_solve_calc_w");
    return read_result(c, pw, 1);
}


extern "C" int tl_cg_calc_ur(tl_chunk *c, double alpha, double *rrn)
{
    prof_begin(c);
    launch_cg_calc_ur(c->s, c->g, ext_interior(c), c->cap, c->prec, alpha, c->f2[F_U], c->f2[F_P], c->f2[F_R],
                      c->f2[F_W], c->f2[F_Z], stored_cp(c), stored_bfp(c), c->f2[F_MI], c->f2[F_KX], c->f2[F_KY],
                      faces_cond(c), c->faces, red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_cg_solve_calc_ur");
    prof_end(c, "tea_leaf_cg_solve_calc_ur");
    return read_result(c, rrn, 1);
}

extern "C" int tl_cg_calc_p(tl_chunk *c, double beta)
{
    prof_begin(c);
    launch_cg_calc_p(c->s, c->g, ext_interior(c), c->cap, c->prec, beta, c->f2[F_P], c->f2[F_R], c->f2[F_Z]);
    TL_CHECK_LAUNCH(c, "tea_leaf_cg_solve_calc_p");
    prof_end(c, "tea_leaf_cg_solve_calc_p");
    return TL_OK;
}

// tea_leaf_kernel_ocl.cpp:98-129. The coefficients live in global memory
// (no constant-buffer size limit applies).
extern "C" int tl_cheby_init(tl_chunk *c, const double *ch_alphas, const double *ch_betas, int n_coefs,
                             double rx, double ry, double theta)
{
    if (n_coefs < 1) {
        set_error("tl_cheby_init: n_coefs %d", n_coefs);
        return TL_ERR_ARG;
    }
    int rc = ensure_coef_capacity(c, n_coefs);
    if (rc != TL_OK) return rc;
    TL_HIP_TRY(hipMemcpyAsync(c->ch_a, ch_alphas, sizeof(double) * n_coefs, hipMemcpyHostToDevice, c->s),
               "tl_cheby_init: coefficient copy");
    TL_HIP_TRY(hipMemcpyAsync(c->ch_b, ch_betas, sizeof(double) * n_coefs, hipMemcpyHostToDevice, c->s),
               "tl_cheby_init: coefficient copy");
    c->n_coefs = n_coefs;
    c->rx = rx;
    c->ry = ry;
    c->theta = theta;
    prof_begin(c);
    launch_cheby_p(c->s, c->g, ext_interior(c), c->cap, c->prec, 1, c->f2[F_U], c->f2[F_U0], c->f2[F_P],
                   c->f2[F_R], c->f2[F_W], stored_cp(c), stored_bfp(c), c->f2[F_MI], c->f2[F_KX], c->f2[F_KY],
                   faces_cond(c), c->faces, c->ch_a, c->ch_b, theta, 0, nullptr);
    TL_CHECK_LAUNCH(c, "tea_leaf_cheby_solve_init_p");
    prof_end(c, "tea_leaf_cheby_solve_init_p");
    prof_begin(c);
    launch_cheby_u(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U], c->f2[F_P], c->f2[F_R],
                   red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_cheby_solve_calc_u");
    prof_end(c, "tea_leaf_cheby_solve_calc_u");
    return tl_synchronize(c);
}

extern "C" int tl_cheby_iterate(tl_chunk *c, int cheby_calc_step)
{
    const int step = cheby_calc_step - 1;
    if (step < 0 || step >= c->n_coefs) {
        set_error("tl_cheby_iterate: step %d outside the %d coefficients", cheby_calc_step, c->n_coefs);
        return TL_ERR_ARG;
    }
    prof_begin(c);
    launch_cheby_p(c->s, c->g, ext_interior(c), c->cap, c->prec, 0, c->f2[F_U], c->f2[F_U0], c->f2[F_P],
                   c->f2[F_R], c->f2[F_W], stored_cp(c), stored_bfp(c), c->f2[F_MI], c->f2[F_KX], c->f2[F_KY],
                   faces_cond(c), c->faces, c->ch_a, c->ch_b, c->theta, step, nullptr);
    TL_CHECK_LAUNCH(c, "tea_leaf_cheby_solve_calc_p");
    prof_end(c, "tea_leaf_cheby_solve_calc_p");
    prof_begin(c);
    launch_cheby_u(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U], c->f2[F_P], c->f2[F_R],
                   red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_cheby_solve_calc_u");
    prof_end(c, "tea_leaf_cheby_solve_calc_u");
    return TL_OK;
}

// tea_leaf_kernel_ocl.cpp:327-344
extern "C" int tl_ppcg_init(tl_chunk *c, const double *ch_alphas, const double *ch_betas, double theta,
                            int n_inner_steps)
{
    if (n_inner_steps < 1) {
        set_error("tl_ppcg_init: n_inner_steps %d", n_inner_steps);
        return TL_ERR_ARG;
    }
    int rc = ensure_coef_capacity(c, n_inner_steps);
    if (rc != TL_OK) return rc;
    TL_HIP_TRY(hipMemcpyAsync(c->ch_a, ch_alphas, sizeof(double) * n_inner_steps, hipMemcpyHostToDevice, c->s),
               "tl_ppcg_init: coefficient copy");
    TL_HIP_TRY(hipMemcpyAsync(c->ch_b, ch_betas, sizeof(double) * n_inner_steps, hipMemcpyHostToDevice, c->s),
               "tl_ppcg_init: coefficient copy");
    c->n_coefs = n_inner_steps;
    c->theta = theta;
    return tl_synchronize(c);
}

extern "C" int tl_ppcg_init_sd(tl_chunk *c)
{
    prof_begin(c);
    launch_ppcg_init_sd(c->s, c->g, ext_interior(c), c->cap, c->prec, c->f2[F_R], c->f2[F_SD], c->f2[F_CP],
                        c->f2[F_BFP], c->f2[F_MI], c->f2[F_KX], c->f2[F_KY], c->theta);
    TL_CHECK_LAUNCH(c, "tea_leaf_ppcg_solve_init_sd");
    prof_end(c, "tea_leaf_ppcg_solve_init_sd");
    return TL_OK;
}

// matrix-powers extent of tea_leaf_kernel_ocl.cpp:352-391
static Ext ppcg_extent(const tl_chunk *c, int bounds_extra, const int *nb)
{
    const int hd = c->g.hd;
    const int step_depth = hd - bounds_extra;
    int off[2] = {step_depth, step_depth};
    int gsz[2] = {c->g.x_max + (hd - step_depth) * 2, c->g.y_max + (hd - step_depth) * 2};
    int bex = bounds_extra, bey = bounds_extra;
    if (nb[0] == -1) {
        off[0] = hd;
        gsz[0] -= (hd - step_depth);
    }
    if (nb[1] == -1) {
        gsz[0] -= (hd - step_depth);
        bex = 0;
    }
    if (nb[2] == -1) {
        off[1] = hd;
        gsz[1] -= (hd - step_depth);
    }
    if (nb[3] == -1) {
        gsz[1] -= (hd - step_depth);
        bey = 0;
    }
    gsz[0] -= gsz[0] % 32;
    gsz[0] += 32;
    gsz[1] -= gsz[1] % 4;
    gsz[1] += 4;
    const int j1 = std::min(c->g.x_max + hd - 1 + bex, off[0] + gsz[0] - 1);
    const int k1 = std::min(c->g.y_max + hd - 1 + bey, off[1] + gsz[1] - 1);
    return make_ext(off[0], j1, off[1], k1);
}

extern "C" int tl_ppcg_inner(tl_chunk *c, int inner_step, int bounds_extra, const int *chunk_neighbours)
{
    const int step = inner_step - 1;
    if (step < 0 || step >= c->n_coefs || bounds_extra < 0 || bounds_extra >= c->g.hd) {
        set_error("tl_ppcg_inner: inner step %d / bounds_extra %d out of range", inner_step, bounds_extra);
        return TL_ERR_ARG;
    }
    const Ext e = ppcg_extent(c, bounds_extra, chunk_neighbours);
    prof_begin(c);
    launch_ppcg_update_r(c->s, c->g, e, c->cap, c->f2[F_U], c->f2[F_R], c->f2[F_KX], c->f2[F_KY], faces_cond(c),
                         c->faces, c->f2[F_SD]);
    TL_CHECK_LAUNCH(c, "tea_leaf_ppcg_solve_update_r");
    prof_end(c, "tea_leaf_ppcg_solve_update_r");
    prof_begin(c);
    launch_ppcg_calc_sd(c->s, c->g, e, c->cap, c->prec, c->f2[F_R], c->f2[F_SD], stored_cp(c), stored_bfp(c),
                        c->f2[F_MI], c->f2[F_KX], c->f2[F_KY], faces_cond(c), c->faces, c->ch_a, c->ch_b, step);
    TL_CHECK_LAUNCH(c, "tea_leaf_ppcg_solve_calc_sd");
    prof_end(c, "tea_leaf_ppcg_solve_calc_sd");
    return TL_OK;
}

// tea_leaf_kernel_ocl.cpp:222-229 (un is the Mi buffer)
extern "C" int tl_jacobi_solve(tl_chunk *c, double *error)
{
    prof_begin(c);
    launch_copy(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U], c->f2[F_MI]);
    TL_CHECK_LAUNCH(c, "tea_leaf_jacobi_copy_u");
    prof_end(c, "tea_leaf_jacobi_copy_u");
    prof_begin(c);
    launch_jacobi(c->s, c->g, ext_interior(c), c->cap, c->f2[F_KX], c->f2[F_KY], faces_cond(c), c->faces,
                  c->f2[F_U0], c->f2[F_U], c->f2[F_MI], red_args(c, RED_STORE, false));
    TL_CHECK_LAUNCH(c, "tea_leaf_jacobi_solve");
    prof_end(c, "tea_leaf_jacobi_solve");
    return read_result(c, error, 1);
}

// The iteration kernels' division (div_nr) and the IEEE quotient of n operand
// pairs, both computed on the device, copied back for comparison.
extern "C" int tl_check_division(const double *num, const double *den, double *quot, double *fast, long n)
{
    if (!num || !den || !quot || !fast || n < 1 || n > (1L << 26)) {
        set_error("tl_check_division: null array or n %ld outside 1..2^26", n);
        return TL_ERR_ARG;
    }
    const size_t bytes = sizeof(double) * (size_t)n;
    double *d[4] = {nullptr, nullptr, nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (double *&p : d)
        if (e == hipSuccess) e = hipMalloc(&p, bytes);
    if (e == hipSuccess) e = hipMemcpy(d[0], num, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d[1], den, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_check_division(nullptr, d[0], d[1], d[2], d[3], n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(quot, d[2], bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(fast, d[3], bytes, hipMemcpyDeviceToHost);
    for (double *p : d)
        if (p) (void)hipFree(p);
    return e == hipSuccess ? TL_OK : hip_fail("tl_check_division", e);
}

// Attainable streaming bandwidth on the chunk's layout: `reps` launches of a
// copy (kind 0: 8 B/cell read + 8 B/cell written) or of the 4-read/3-write
// mix of the CG update kernel (kind 1: 56 B/cell) over the interior, on
// temporary fields, timed with events on the chunk's stream.
extern "C" int tl_stream_bandwidth(tl_chunk *c, int kind, int reps, double *bytes_per_s)
{
    if (!c || (kind != 0 && kind != 1) || reps < 1 || !bytes_per_s) {
        set_error("tl_stream_bandwidth: bad kind %d or reps %d", kind, reps);
        return TL_ERR_ARG;
    }
    const int nf = kind == 0 ? 2 : 5;
    double *base[5] = {}, *f[5] = {};
    hipError_t e = hipSuccess;
    int rc = TL_OK;
    for (int i = 0; i < nf && e == hipSuccess; ++i) {
        e = hipMalloc(&base[i], sizeof(double) * c->elems2d);
        if (e == hipSuccess) e = hipMemsetAsync(base[i], 0, sizeof(double) * c->elems2d, c->s);
        f[i] = base[i] + c->off;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) {
        const Ext in = ext_interior(c);
        auto once = [&]() {
            if (kind == 0) launch_copy(c->s, c->g, in, c->cap, f[0], f[1]);
            else launch_stream_mix(c->s, c->g, in, c->cap, f[0], f[1], f[2], f[3], f[4]);
        };
        once();
        e = hipEventRecord(t0, c->s);
        for (int r = 0; r < reps && e == hipSuccess; ++r) once();
        if (e == hipSuccess) e = hipEventRecord(t1, c->s);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        const double cells = (double)c->g.x_max * c->g.y_max;
        *bytes_per_s = cells * (kind == 0 ? 16.0 : 56.0) * reps / (1e-3 * ms);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail("tl_stream_bandwidth", e);
    (void)hipStreamSynchronize(c->s);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    for (double *b : base)
        if (b) (void)hipFree(b);
    return rc;
}

extern "C" int tl_set_tuning(tl_chunk *c, int key, int value)
{
    if ((key != TL_TUNE_PPCG_STEPS_PER_LAUNCH && key != TL_TUNE_CHEBY_STEPS_PER_LAUNCH) || (value != 1 && value != 2)) {
        set_error("tl_set_tuning: bad key %d or value %d", key, value);
        return TL_ERR_ARG;
    }
    int &slot = key == TL_TUNE_PPCG_STEPS_PER_LAUNCH ? (c ? c->ppcg_steps_per_launch : g_ppcg_steps_default)
                                                     : (c ? c->cheby_steps_per_launch : g_cheby_steps_default);
    slot = value;
    return TL_OK;
}

// ocl_errors.cpp:203-232
extern "C" int tl_print_profiling(tl_chunk *c)
{
    if (!c || !c->profiler_on) return TL_OK;
    fprintf(stdout, "@@@@@ HIP Profiling information @@@@@\n");
    for (auto &kv : c->prof) {
        if (!kv.second.second) continue;
        fprintf(stdout, "%30s : %10.3f ms (%.2f μs avg. over %ld calls)\n", kv.first.c_str(), kv.second.first,
                1e3 * kv.second.first / kv.second.second, kv.second.second);
    }
    fflush(stdout);
    return TL_OK;
}

// ===========================================================================
// drop-in entry points (Fortran ABI): same semantics, abort on error
// ===========================================================================

// cloverDie (ocl_errors.cpp:175-194): report, then MPI_Abort when the host
// program runs under MPI, else exit(1)
[[noreturn]] static void die(int line, const char *what)
{
    fprintf(stderr, "@@@@@\n");
    fprintf(stderr, "\x1b[31m");
    fprintf(stderr, "Fatal error at line %d in %s:", line, __FILE__);
    fprintf(stderr, "\x1b[0m");
    fprintf(stderr, "\n%s\n%s\n", what, g_err);
    fprintf(stderr, "\nExiting\n");
    fflush(stderr);
    typedef int (*init_fn)(int *);
    typedef int (*abort_fn)(long, int);
    init_fn mpi_initialized = (init_fn)dlsym(RTLD_DEFAULT, "MPI_Initialized");
    if (mpi_initialized) {
        int flag = 0;
        mpi_initialized(&flag);
        void *comm_world = dlsym(RTLD_DEFAULT, "ompi_mpi_comm_world");
        abort_fn mpi_abort = (abort_fn)dlsym(RTLD_DEFAULT, "MPI_Abort");
        // MPICH encodes MPI_COMM_WORLD as the integer 0x44000000
        if (flag && mpi_abort) mpi_abort(comm_world ? (long)comm_world : 0x44000000L, 1);
    }
    exit(1);
}

#define TL_ABI(call)                                   \
    do {                                               \
        if (!g_default) {                              \
            set_error("chunk not initialised");        \
            die(__LINE__, #call);                      \
        }                                              \
        if ((call) != TL_OK) die(__LINE__, #call);     \
    } while (0)

extern "C" void initialise_ocl_(int *x_min, int *x_max, int *y_min, int *y_max)
{
    tl::SideChannel sc;
    if (tl::read_side_channel("tea.in", &sc) != 0) {
        set_error("Input file not found");
        die(__LINE__, "initialise_ocl_");
    }
    if (sc.halo_depth < 1) {
        set_error("Halo exchange depth unspecified or was too small");
        die(__LINE__, "initialise_ocl_");
    }
    const int nx = *x_max - *x_min + 1, ny = *y_max - *y_min + 1;
    if (g_default) tl_chunk_destroy(g_default);
    tl_chunk *c = nullptr;
    if (tl_chunk_create(nx, ny, sc.halo_depth, sc.preconditioner, sc.profiler_on, &c) != TL_OK)
        die(__LINE__, "initialise_ocl_");
    c->solver = sc.solver;
    g_default = c;
}

extern "C" void initialise_chunk_kernel_ocl_(double *xmin, double *ymin, double *dx, double *dy)
{
    TL_ABI(tl_initialise_chunk(g_default, *xmin, *ymin, *dx, *dy));
}

extern "C" void generate_chunk_kernel_ocl_(const int *number_of_states, const double *state_density,
                                           const double *state_energy, const double *state_xmin,
                                           const double *state_xmax, const double *state_ymin,
                                           const double *state_ymax, const double *state_radius,
                                           const int *state_geometry, const int *g_rect, const int *g_circ,
                                           const int *g_point)
{
    TL_ABI(tl_generate_chunk(g_default, *number_of_states, state_density, state_energy, state_xmin, state_xmax,
                             state_ymin, state_ymax, state_radius, state_geometry, *g_rect, *g_circ, *g_point));
}

extern "C" void set_field_kernel_ocl_(void) { TL_ABI(tl_set_field(g_default)); }

extern "C" void field_summary_kernel_ocl_(double *vol, double *mass, double *ie, double *temp)
{
    TL_ABI(tl_field_summary(g_default, vol, mass, ie, temp));
}

extern "C" void update_halo_kernel_ocl_(const int *chunk_neighbours, const int *fields, const int *depth)
{
    TL_ABI(tl_update_halo(g_default, chunk_neighbours, fields, *depth));
}

extern "C" void tea_leaf_common_init_kernel_ocl_(const int *coefficient, double *dt, double *rx, double *ry,
                                                 int *zero_boundary, int *reflective_boundary)
{
    TL_ABI(tl_common_init(g_default, *coefficient, *dt, rx, ry, zero_boundary, *reflective_boundary));
}

extern "C" void tea_leaf_calc_residual_ocl_(void) { TL_ABI(tl_calc_residual(g_default)); }

extern "C" void tea_leaf_calc_2norm_kernel_ocl_(int *norm_array, double *norm)
{
    TL_ABI(tl_calc_2norm(g_default, *norm_array, norm));
}

extern "C" void tea_leaf_common_finalise_kernel_ocl_(void) { TL_ABI(tl_finalise(g_default)); }

extern "C" void tea_leaf_cg_init_kernel_ocl_(double *rro) { TL_ABI(tl_cg_init(g_default, rro)); }

extern "C" void tea_leaf_cg_calc_w_kernel_ocl_(double *pw) { TL_ABI(tl_cg_calc_w(g_default, pw)); }

extern "C" void tea_leaf_cg_calc_ur_kernel_ocl_(double *alpha, double *rrn)
{
    TL_ABI(tl_cg_calc_ur(g_default, *alpha, rrn));
}

extern "C" void tea_leaf_cg_calc_p_kernel_ocl_(double *beta) { TL_ABI(tl_cg_calc_p(g_default, *beta)); }

extern "C" void tea_leaf_cheby_init_kernel_ocl_(const double *ch_alphas, const double *ch_betas, int *n_coefs,
                                                const double *rx, const double *ry, const double *theta)
{
    TL_ABI(tl_cheby_init(g_default, ch_alphas, ch_betas, *n_coefs, *rx, *ry, *theta));
}

extern "C" void tea_leaf_cheby_iterate_kernel_ocl_(const int *cheby_calc_step)
{
    TL_ABI(tl_cheby_iterate(g_default, *cheby_calc_step));
}

extern "C" void tea_leaf_ppcg_init_ocl_(const double *ch_alphas, const double *ch_betas, double *theta,
                                        int *n_inner_steps)
{
    TL_ABI(tl_ppcg_init(g_default, ch_alphas, ch_betas, *theta, *n_inner_steps));
}

extern "C" void tea_leaf_ppcg_init_sd_kernel_ocl_(void) { TL_ABI(tl_ppcg_init_sd(g_default)); }

extern "C" void tea_leaf_ppcg_inner_kernel_ocl_(int *inner_step, int *bounds_extra, const int *chunk_neighbours)
{
    TL_ABI(tl_ppcg_inner(g_default, *inner_step, *bounds_extra, chunk_neighbours));
}

extern "C" void tea_leaf_jacobi_solve_kernel_ocl_(double *error) { TL_ABI(tl_jacobi_solve(g_default, error)); }

extern "C" void ocl_pack_buffers_(int fields[7], int offsets[7], int *depth, int *face, double *host_buffer)
{
    (void)fields; (void)offsets; (void)depth; (void)face; (void)host_buffer;
    set_error("multi-chunk halo packing is not part of the one-chunk-per-GPU build");
    die(__LINE__, "ocl_pack_buffers_");
}

extern "C" void ocl_unpack_buffers_(int fields[7], int offsets[7], int *depth, int *face, double *host_buffer)
{
    (void)fields; (void)offsets; (void)depth; (void)face; (void)host_buffer;
    set_error("multi-chunk halo packing is not part of the one-chunk-per-GPU build");
    die(__LINE__, "ocl_unpack_buffers_");
}

extern "C" void print_opencl_profiling_info_(void)
{
    if (g_default) tl_print_profiling(g_default);
}
