// Private definition of the chunk: the device buffers of one TeaLeaf chunk
// (types.hpp:42-367 of the reference), its stream, the reduction scratch and
// the device-resident solver state.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <utility>

#include "tl_common.h"

struct tl_chunk {
    tl::Geom g{};
    int prec = tl::PREC_NONE;
    int device = 0;                  // the device current at creation
    int profiler_on = 0;
    int solver = tl::SOLVER_JACOBI;  // side-channel solver (informational)
    hipStream_t s = nullptr;

    long elems2d = 0;                // doubles per 2-D allocation
    long off = 0;                    // offset of element (0,0) in an allocation
    double *base2[tl::NUM_FIELDS_2D] = {};
    double *f2[tl::NUM_FIELDS_2D] = {};
    double *f1[tl::NUM_FIELDS_1D] = {};
    int len1d = 0;

    int cap = 8192;                  // max blocks per launch
    double *partials = nullptr;
    double *result = nullptr;
    tl::SolveState *st = nullptr;
    int *range_bad = nullptr;        // k_conductivity's flag: a density outside the fast division's range
    double *hist_a = nullptr, *hist_b = nullptr, *hist_e = nullptr;
    int hist_len = 0;
    double *ch_a = nullptr, *ch_b = nullptr;
    int ch_cap = 0;
    int n_coefs = 0;

    double *h_result = nullptr;      // pinned, coherent: totals of host-returned reductions
    double *d_result = nullptr;      // h_result's device address
    unsigned int *h_flag = nullptr;  // pinned, coherent: sequence number of the latest totals
    unsigned int *d_flag = nullptr;
    unsigned int seq = 0;            // sequence number of the latest host-returned reduction
    tl::SolveState *h_st = nullptr;  // pinned, 2 slots

    // host copies of scalars the reference keeps in kernel arguments
    double rx = 0, ry = 0, theta = 0, dx = 0, dy = 0;
    // face parameters of the latest common_init (F_COND holds its conductivity)
    tl::Faces faces{};
    // progress callback of long calls (tl_set_progress)
    void (*progress)(void *, const char *, int, int) = nullptr;
    void *progress_user = nullptr;

    // profiler_on: per-kernel device time (ms) and calls (ocl_errors.cpp:64-121)
    std::map<std::string, std::pair<double, long>> prof;
    hipEvent_t pe0 = nullptr, pe1 = nullptr;
    hipEvent_t pe_mid = nullptr;     // between a reduction kernel and its finisher

    // captured iteration batches of the native driver
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int key = -1;
    };
    GraphSlot g_cg, g_cheby, g_ppcg, g_jacobi;
    // PPCG inner steps per fused launch in the native driver (1 or 2;
    // tl_set_tuning TL_TUNE_PPCG_STEPS_PER_LAUNCH)
    int ppcg_steps_per_launch = 2;
    // PPCG forms r.z with z = M^-1 r of the final r of the inner steps (deck
    // keyword tl_ppcg_fresh_z) instead of the reference's stale z
    bool ppcg_fresh_z = false;
    // cp, bfp hold the block-Jacobi factors of the current Kx, Ky (set by
    // tl_block_init, cleared by whatever rewrites Kx, Ky, cp or bfp): the
    // drop-in kernels then recompute the factors in registers (bitwise the
    // stored ones) instead of reading them
    bool factors_current = false;
    // Kx, Ky are the faces of the last tl_common_init (cleared by an upload of
    // Kx or Ky): the drop-in CG kernels then form them from the conductivity
    // (bitwise the stored values) instead of reading both arrays
    bool faces_current = false;
    // Chebyshev iterations per fused launch (1 or 2; TL_TUNE_CHEBY_STEPS_PER_LAUNCH)
    int cheby_steps_per_launch = 2;
    // events around the eagerly launched first CG iterations of each batch
    // (an odd and an even one; one set per pipeline slot)
    hipEvent_t ev_t[2][8] = {};
    hipEvent_t ev_batch[2] = {nullptr, nullptr};
    bool timing = false;
    double t_k1_ms = 0, t_k2_ms = 0;
    long t_samples = 0;

    // refined solve (tl_refine.hip): [0] FP64 inner fields, [1] FP32 inner
    // fields, allocated on first use
    struct RefineWork {
        int bytes = 0;                    // 8 or 4: storage of the inner fields
        tl::Geom g{};                     // layout of the inner fields
        long off = 0, elems = 0;
        void *base[10] = {};              // owned allocations
        void *f[10] = {};                 // faces (FP64: conductivity w; FP32: h = 1/(2w) or Kx), r, z,
                                          // p, p2, d, sd, sd2, r2, FP32 Ky
        int faces = 0;                    // face representation of the solve in progress
        tl::RefineState *st = nullptr;
        tl::RefineState *h_st = nullptr;  // pinned, 2 slots
        GraphSlot graph;                  // CG iterations
        GraphSlot graph_ppcg;             // PPCG outer steps
        double *hist = nullptr;           // CG alphas, betas of the first iterations
        double *coef = nullptr;           // PPCG inner Chebyshev alphas, betas
    } rf[2];
};

namespace tl {

// interior and launch-bound extents (ocl_kernels.cpp:123-134)
Ext make_ext(int j0, int j1, int k0, int k1);
Ext ext_range(const tl_chunk *c, int xmin, int xmax, int ymin, int ymax);
inline Ext ext_interior(const tl_chunk *c) { return ext_range(c, 0, 0, 0, 0); }
RedArgs red_args(tl_chunk *c, int mode, bool with_state);
void set_error(const char *fmt, ...);
int hip_fail(const char *what, hipError_t e);
// Runs a HIP call in a function that returns a TL_ status: on failure the
// error is recorded and the function returns TL_ERR_HIP. Calls whose result
// is deliberately dropped (teardown, and profiling events whose errors the
// next launch check reports) are cast to void instead.
#define TL_HIP_TRY(call, what)                                                   \
    do {                                                                         \
        const hipError_t e_try_ = (call);                                        \
        if (e_try_ != hipSuccess) return ::tl::hip_fail(what, e_try_);           \
    } while (0)
// profiling wrapper bracket (no-ops unless profiler_on)
void prof_begin(tl_chunk *c);
void prof_end(tl_chunk *c, const char *name);
// ... for a launch that records c->pe_mid between its two kernels: the time
// up to pe_mid goes to `name`, the rest to `second`
void prof_end_split(tl_chunk *c, const char *name, const char *second);
void prof_add(tl_chunk *c, const char *name, double ms);
int ensure_coef_capacity(tl_chunk *c, int n);
int ensure_hist_capacity(tl_chunk *c, int n);
int read_result(tl_chunk *c, double *out, int nq);
// z = M^-1 r over the interior with the chunk's preconditioner (stored block
// factors or Mi; nothing without one): the fresh-z PPCG option in the
// reference call sequence
int precondition_r(tl_chunk *c);
// frees the refined-solve workspaces (tl_refine.hip)
void refine_release(tl_chunk *c);
// Whether the CG iterations' w kernel reads z with nontemporal loads: when an
// FP64 field is at most 256 MB (the size of MI355X's last-level cache, so
// that p_old and the conductivity, which the hint leaves cached, span about
// twice it). Measured: at 4096^2 (134 MB fields) the hint makes k_fused_cg_w
// 7 % faster, at 8192^2 (537 MB) 8 % slower.
inline bool stream_z_pays(const tl_chunk *c) { return sizeof(double) * c->elems2d <= (256L << 20); }
// destroys every captured iteration batch (native and refined): they hold the
// face parameters (rx, ry, walls) and buffer addresses by value
void drop_graphs(tl_chunk *c);
// captures the launches `body` enqueues on the chunk's stream into `gs`,
// unless gs already holds a batch captured under the same key
int capture_graph(tl_chunk *c, tl_chunk::GraphSlot &gs, int key, const std::function<void()> &body);

}  // namespace tl
