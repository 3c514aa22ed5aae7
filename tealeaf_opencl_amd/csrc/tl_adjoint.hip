// Adjoint of the implicit diffusion timesteps with respect to the per-cell
// density. The reference has no adjoint; the forward model differentiated
// here is its timestep (tea_solve.f90:70-101 and :352-420 around the solve):
//
//   u^0 = rho * e^0,   u^{n+1} = A(rho)^{-1} u^n  (n = 0 .. T-1),
//   A = I + sum_f K_f (e_a - e_b)(e_a - e_b)^T over the cell faces f = (a, b),
//   K_f = r (w_a + w_b) / (2 w_a w_b)   (tea_leaf_common_cl.cl:56-90),
//   w = rho (conductivity) or 1/rho (reciprocal), faces on the walls zero,
//   L = sum over the interior of (u^T - u_obs)^2.
//
// Backward pass (A is symmetric, so its adjoint solve uses the same kernels):
//   g^T = 2 (u^T - u_obs);  for n = T-1 .. 0:  A lambda_n = g^{n+1},
//   dL/dK_f += -(lambda_a - lambda_b)(u^{n+1}_a - u^{n+1}_b),  g^n = lambda_n;
//   dL/drho_c = sum over the faces f of c of dL/dK_f dK_f/drho_c + g^0_c e^0_c,
//   dK_f/dw_c = -r / (2 w_c^2), so dK_f/drho_c = -r / (2 rho_c^2) for w = rho
//   and r / 2 for w = 1/rho.
//
// The forward states u^n are kept only every `checkpoint_every` steps; the
// backward pass recomputes each segment from its checkpoint before walking
// it backwards. Every solve is the refined FP64 CG (tl_solve_refined) with
// the chunk's preconditioner.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/tealeaf_mi355x.h"
#include "tl_chunk.h"
#include "tl_device.h"
#include "tl_kernels.h"

namespace tl {

// kernels of the adjoint (a named namespace: kernels with internal linkage are
// not visible to the profiling tools)
namespace adj {

// g = 2 (u - u_obs) over the interior, partial sums of (u - u_obs)^2
__global__ void __launch_bounds__(BLOCK_THREADS) k_adj_seed(Geom g, Ext e, const double *u, const double *u_obs,
                                                            double *gv, double *partials)
{
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double uv[STRIP], ov[STRIP];
    col4(u, i0, P, uv);
    col4(u_obs, i0, P, ov);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const double d = uv[q] - ov[q];
        gv[i0 + q * P] = 2.0 * d;
        acc[0] += d * d;
    }
    TL_TILE_LOOP_END
    block_partial<1>(acc, partials);
}

// the loss to the host as the drop-in reductions return theirs (read_result)
__global__ void __launch_bounds__(FINISH_THREADS) k_adj_sum(const double *partials, int nblocks, RedArgs ra)
{
    double tot[MAX_RED];
    sum_partials<1>(partials, nblocks, tot);
    if (threadIdx.x == 0) {
        ra.result[0] = tot[0];
        __hip_atomic_store(ra.flag, ra.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// whether tea_leaf_zero_boundary leaves the face zero (the conditions of
// kx_at / ky_at)
__device__ __forceinline__ bool kx_fixed(const Geom &g, const Faces &f, int j, int k)
{
    const int hd = g.hd;
    return outside_init(g, j, k) || (f.zl && j <= hd) || (f.zr && j > g.x_max - 1 + hd) || (f.zb && k < hd) ||
           (f.zt && k > g.y_max - 1 + hd);
}

__device__ __forceinline__ bool ky_fixed(const Geom &g, const Faces &f, int j, int k)
{
    const int hd = g.hd;
    return outside_init(g, j, k) || (f.zl && j < hd) || (f.zr && j > g.x_max - 1 + hd) || (f.zb && k <= hd) ||
           (f.zt && k > g.y_max - 1 + hd);
}

// grad_c += sum over the four faces of c of
//   -(lambda_c - lambda_nb)(u_c - u_nb) * dK_f/drho_c
// (a gather per cell: every face term that involves rho_c is formed by c's
// own thread, so no two threads write the same cell)
__global__ void __launch_bounds__(BLOCK_THREADS) k_adj_face_grad(Geom g, Ext e, const double *lam, const double *u,
                                                                 const double *w, Faces fc, int coef, double *grad)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double lc[STRIP + 2], lw[STRIP], le[STRIP], uc[STRIP + 2], uw[STRIP], ue[STRIP], wc[STRIP], gr[STRIP];
    col6(lam, i0, P, lc);
    col4(lam, i0 - 1, P, lw);
    col4(lam, i0 + 1, P, le);
    col6(u, i0, P, uc);
    col4(u, i0 - 1, P, uw);
    col4(u, i0 + 1, P, ue);
    col4(w, i0, P, wc);
    col4(grad, i0, P, gr);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        const int k = ks + q;
        if (k > e.k1) continue;
        const double l0 = lc[q + 1], u0 = uc[q + 1];
        // dK/drho_c without the face's own scale factor r
        const double dkx = (coef == COEF_CONDUCTIVITY) ? -fc.rx / (2.0 * wc[q] * wc[q]) : 0.5 * fc.rx;
        const double dky = (coef == COEF_CONDUCTIVITY) ? -fc.ry / (2.0 * wc[q] * wc[q]) : 0.5 * fc.ry;
        double s = 0.0;
        if (!kx_fixed(g, fc, j, k)) s -= (l0 - lw[q]) * (u0 - uw[q]) * dkx;
        if (!kx_fixed(g, fc, j + 1, k)) s -= (l0 - le[q]) * (u0 - ue[q]) * dkx;
        if (!ky_fixed(g, fc, j, k)) s -= (l0 - lc[q]) * (u0 - uc[q]) * dky;
        if (!ky_fixed(g, fc, j, k + 1)) s -= (l0 - lc[q + 2]) * (u0 - uc[q + 2]) * dky;
        grad[i0 + q * P] = gr[q] + s;
    }
    TL_TILE_LOOP_END
}

// grad += lambda_0 * e^0 (the dependence of u^0 = rho e^0 on rho)
__global__ void __launch_bounds__(BLOCK_THREADS) k_adj_initial(Geom g, Ext e, const double *lam, const double *e0,
                                                               double *grad)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double lv[STRIP], ev[STRIP], gr[STRIP];
    col4(lam, i0, P, lv);
    col4(e0, i0, P, ev);
    col4(grad, i0, P, gr);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) grad[i0 + q * P] = gr[q] + lv[q] * ev[q];
    TL_TILE_LOOP_END
}

#define RC(x)                          \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != TL_OK) return rc_;  \
    } while (0)

// device fields owned by one adjoint call
struct Work {
    tl_chunk *c;
    std::vector<double *> bases;
    ~Work()
    {
        (void)hipStreamSynchronize(c->s);
        for (double *b : bases) (void)hipFree(b);
    }
    int alloc(double **out)
    {
        double *b = nullptr;
        hipError_t e = hipMalloc(&b, sizeof(double) * c->elems2d);
        if (e != hipSuccess) return hip_fail("hipMalloc adjoint", e);
        bases.push_back(b);
        e = hipMemsetAsync(b, 0, sizeof(double) * c->elems2d, c->s);
        if (e != hipSuccess) return hip_fail("hipMemset adjoint", e);
        *out = b + c->off;
        return TL_OK;
    }
};

int upload_padded(tl_chunk *c, double *dst, const double *host)
{
    const int w = c->g.x_max + 2 * c->g.hd, h = c->g.y_max + 2 * c->g.hd;
    hipError_t e = hipMemcpy2DAsync(dst, sizeof(double) * c->g.pitch, host, sizeof(double) * w, sizeof(double) * w,
                                    h, hipMemcpyHostToDevice, c->s);
    return e == hipSuccess ? TL_OK : hip_fail("hipMemcpy2D", e);
}

int download_padded(tl_chunk *c, double *host, const double *src)
{
    const int w = c->g.x_max + 2 * c->g.hd, h = c->g.y_max + 2 * c->g.hd;
    hipError_t e = hipMemcpy2DAsync(host, sizeof(double) * w, src, sizeof(double) * c->g.pitch, sizeof(double) * w,
                                    h, hipMemcpyDeviceToHost, c->s);
    if (e != hipSuccess) return hip_fail("hipMemcpy2D", e);
    e = hipStreamSynchronize(c->s);
    return e == hipSuccess ? TL_OK : hip_fail("hipStreamSynchronize", e);
}

void copy_interior(tl_chunk *c, const double *src, double *dst)
{
    launch_copy(c->s, c->g, ext_interior(c), c->cap, src, dst);
}

// counters of the solves of one call
struct Counts {
    long forward = 0, recompute = 0, adjoint = 0;
    int stalled = 0;
};

// One FP64 refined solve; a solve that stops at the rounding floor of the
// true residual above tol (stalled) is accepted and counted, one that runs
// out of iterations is an error.
int solve(tl_chunk *c, double tol, int max_iters, long *iterations, Counts *cnt, const char *phase, int step)
{
    tl_refine_result rr;
    RC(tl_solve_refined(c, TL_PRECISION_FP64, tol, max_iters, 0.0, &rr));
    if (!rr.converged && !rr.stalled) {
        set_error("adjoint: a solve did not reach %g in %d iterations (residual %g)", tol, max_iters,
                  rr.rel_residual);
        return TL_ERR_SOLVER;
    }
    cnt->stalled += rr.stalled;
    *iterations += rr.iterations;
    if (c->progress) c->progress(c->progress_user, phase, step, rr.iterations);
    return TL_OK;
}

// the forward timesteps from the chunk's energy1; with `ck`, u^n = rho e^n is
// copied into ck[n / every] at every multiple of `every`
int forward(tl_chunk *c, int coef, double dt, int nsteps, double tol, int max_iters, double *const *ck, int every,
            Counts *cnt)
{
    const int zb[4] = {1, 1, 1, 1};
    for (int n = 0; n < nsteps; ++n) {
        double rx, ry;
        RC(tl_common_init(c, coef, dt, &rx, &ry, zb, 0));
        if (ck && n % every == 0) copy_interior(c, c->f2[F_U0], ck[n / every]);
        RC(solve(c, tol, max_iters, &cnt->forward, cnt, "forward", n + 1));
        RC(tl_finalise(c));
    }
    return TL_OK;
}

int loss_and_seed(tl_chunk *c, const double *u_obs, double *gvec, double *loss)
{
    const Ext e = ext_interior(c);
    const int grid = grid_for(e, c->cap);
    hipLaunchKernelGGL(k_adj_seed, dim3(grid), dim3(BLOCK_THREADS), 0, c->s, c->g, e, c->f2[F_U], u_obs, gvec,
                       c->partials);
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(FINISH_THREADS), 0, c->s, c->partials, grid,
                       red_args(c, RED_STORE, false));
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return hip_fail("launch adjoint seed", err);
    return read_result(c, loss, 1);
}

int check_args(tl_chunk *c, int coef, double dt, int nsteps, double tol, int max_iters, const double *u_obs)
{
    if (!c || !u_obs || (coef != COEF_CONDUCTIVITY && coef != COEF_RECIP_CONDUCTIVITY) || !(dt > 0.0) ||
        nsteps < 1 || !(tol > 0.0) || max_iters < 1) {
        set_error("adjoint: invalid chunk, coefficient %d, dt %g, steps %d, tol %g or max_iters %d", coef, dt, nsteps,
                  tol, max_iters);
        return TL_ERR_ARG;
    }
    return TL_OK;
}

}  // namespace adj
}  // namespace tl

using namespace tl;

extern "C" int tl_forward_loss(tl_chunk *c, int coefficient, double dt, int nsteps, double tol, int max_iters,
                               const double *u_obs, double *loss)
{
    int rc = adj::check_args(c, coefficient, dt, nsteps, tol, max_iters, u_obs);
    if (rc != TL_OK) return rc;
    if (!loss) {
        set_error("tl_forward_loss: null loss");
        return TL_ERR_ARG;
    }
    adj::Work wk{c, {}};
    double *uo = nullptr, *gv = nullptr;
    if ((rc = wk.alloc(&uo)) != TL_OK || (rc = wk.alloc(&gv)) != TL_OK) return rc;
    if ((rc = adj::upload_padded(c, uo, u_obs)) != TL_OK) return rc;
    adj::Counts cnt;
    if ((rc = adj::forward(c, coefficient, dt, nsteps, tol, max_iters, nullptr, 1, &cnt)) != TL_OK) return rc;
    return adj::loss_and_seed(c, uo, gv, loss);
}

extern "C" int tl_adjoint_density(tl_chunk *c, int coefficient, double dt, int nsteps, double tol, int max_iters,
                                  int checkpoint_every, const double *u_obs, double *grad,
                                  tl_adjoint_result *out)
{
    int rc = adj::check_args(c, coefficient, dt, nsteps, tol, max_iters, u_obs);
    if (rc != TL_OK) return rc;
    if (!grad || checkpoint_every < 1) {
        set_error("tl_adjoint_density: null gradient or checkpoint interval %d", checkpoint_every);
        return TL_ERR_ARG;
    }
    const int every = checkpoint_every < nsteps ? checkpoint_every : nsteps;
    const int nck = (nsteps + every - 1) / every;
    adj::Work wk{c, {}};
    std::vector<double *> ck(nck), seg(every);
    double *uo, *gv, *e0, *gr;
    for (auto &p : ck) RC(wk.alloc(&p));
    for (auto &p : seg) RC(wk.alloc(&p));
    RC(wk.alloc(&uo));
    RC(wk.alloc(&gv));
    RC(wk.alloc(&e0));
    RC(wk.alloc(&gr));
    RC(adj::upload_padded(c, uo, u_obs));
    adj::copy_interior(c, c->f2[F_ENERGY1], e0);

    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    for (auto &x : ev)
        if (hipEventCreate(&x) != hipSuccess) {
            for (auto &y : ev)
                if (y) (void)hipEventDestroy(y);
            set_error("tl_adjoint_density: hipEventCreate failed");
            return TL_ERR_HIP;
        }
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard()
        {
            for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};

    adj::Counts cnt;
    TL_HIP_TRY(hipEventRecord(ev[0], c->s), "tl_adjoint_density");
    RC(adj::forward(c, coefficient, dt, nsteps, tol, max_iters, ck.data(), every, &cnt));
    TL_HIP_TRY(hipEventRecord(ev[1], c->s), "tl_adjoint_density");
    double loss = 0.0;
    RC(adj::loss_and_seed(c, uo, gv, &loss));

    const Ext e = ext_interior(c);
    for (int s = ((nsteps - 1) / every) * every; s >= 0; s -= every) {
        const int end = (s + every < nsteps) ? s + every : nsteps;
        // recompute u^{s+1} .. u^{end} from the checkpoint u^s
        const double *prev = ck[s / every];
        for (int m = s; m < end; ++m) {
            adj::copy_interior(c, prev, c->f2[F_U0]);
            adj::copy_interior(c, prev, c->f2[F_U]);
            RC(adj::solve(c, tol, max_iters, &cnt.recompute, &cnt, "recompute", m + 1));
            adj::copy_interior(c, c->f2[F_U], seg[m - s]);
            prev = seg[m - s];
        }
        // walk the segment backwards
        for (int n = end - 1; n >= s; --n) {
            adj::copy_interior(c, gv, c->f2[F_U0]);
            TL_HIP_TRY(hipMemsetAsync(c->base2[F_U], 0, sizeof(double) * c->elems2d, c->s), "tl_adjoint_density");
            RC(adj::solve(c, tol, max_iters, &cnt.adjoint, &cnt, "adjoint", n + 1));
            hipLaunchKernelGGL(adj::k_adj_face_grad, dim3(grid_for(e, c->cap)), dim3(BLOCK_THREADS), 0, c->s,
                               c->g, e, c->f2[F_U], seg[n - s], c->f2[F_COND], c->faces, coefficient, gr);
            adj::copy_interior(c, c->f2[F_U], gv);
        }
    }
    hipLaunchKernelGGL(adj::k_adj_initial, dim3(grid_for(e, c->cap)), dim3(BLOCK_THREADS), 0, c->s, c->g, e, gv,
                       e0, gr);
    TL_HIP_TRY(hipEventRecord(ev[2], c->s), "tl_adjoint_density");
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return hip_fail("launch adjoint", err);
    RC(adj::download_padded(c, grad, gr));
    if (out) {
        memset(out, 0, sizeof *out);
        float a = 0.f, b = 0.f;
        TL_HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]), "tl_adjoint_density");
        TL_HIP_TRY(hipEventElapsedTime(&b, ev[1], ev[2]), "tl_adjoint_density");
        out->loss = loss;
        out->forward_s = 1e-3 * a;
        out->adjoint_s = 1e-3 * b;
        out->work_bytes = (double)wk.bases.size() * sizeof(double) * c->elems2d;
        double chunk = (double)NUM_FIELDS_2D * sizeof(double) * c->elems2d;
        for (const auto &w : c->rf)
            for (void *buf : w.base)
                if (buf) chunk += (double)w.bytes * w.elems;
        out->chunk_bytes = chunk;
        out->forward_iterations = cnt.forward;
        out->recompute_iterations = cnt.recompute;
        out->adjoint_iterations = cnt.adjoint;
        out->checkpoints = nck;
        out->stalled_solves = cnt.stalled;
    }
    return TL_OK;
}

extern "C" int tl_set_progress(tl_chunk *c, tl_progress_fn fn, void *user)
{
    if (!c) {
        set_error("tl_set_progress: null chunk");
        return TL_ERR_ARG;
    }
    c->progress = fn;
    c->progress_user = user;
    return TL_OK;
}
