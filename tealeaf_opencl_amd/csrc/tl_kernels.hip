// HIP kernels for gfx950 (MI355X): the TeaLeaf implicit diffusion solve in
// FP64. Tiling, reductions and the strip helpers are in tl_device.h.
//
// Two families:
//  * reference-granularity kernels (one per kernel_files/*.cl kernel, same
//    inputs and outputs) behind the drop-in entry points;
//  * fused iteration kernels for the device-resident driver: the CG search
//    direction update is folded into the next matvec (p is double-buffered so
//    neighbours are read before they are overwritten), w = A p is recomputed
//    instead of stored, the face coefficients are recomputed from the
//    conductivity (bitwise those tea_leaf_init_common stores), and the
//    diagonal / block-Jacobi factors are recomputed in registers with the same
//    expressions as tea_leaf_init_jac_diag and tea_leaf_block_init. The PPCG
//    smoothing steps and the Chebyshev iterations run two per launch, with
//    the tile's ring recomputed (k_fused_ppcg_inner2, k_fused_cheby2).
//
// A reduction kernel stores one partial per block; the one-block finisher
// k_finish that follows it sums them and runs the solver's scalar algebra on
// the device-resident SolveState.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "tl_common.h"
#include "tl_device.h"
#include "tl_kernels.h"

namespace tl {

// tea_solve.f90:322-331: error = SQRT(error), exit test, old_error, then the
// solver-switch test at the top of the next iteration (:151-165)
__device__ __forceinline__ void iteration_tail(SolveState *s, double *hist_e)
{
    s->error = sqrt(s->error);
    if (hist_e) hist_e[s->n] = s->error;
    if (fabs(s->error) < s->eps * s->initial_residual) {
        s->done = 1;
        s->halt = 1;
        s->exit_n = s->n;
        return;
    }
    s->old_error = s->error;
    if (s->n >= s->max_iters) {
        s->done = 1;
        s->halt = 1;
        s->exit_n = s->max_iters + 1;
        return;
    }
    s->n += 1;
    if (s->cheby_calc_steps == 0 && (s->solver == SOLVER_CHEBY || s->solver == SOLVER_PPCG)) {
        const bool sw = (!s->first && s->solver == SOLVER_PPCG && s->n > 1) ||
                        (fabs(s->old_error) <= s->epslim && s->n >= s->presteps);
        if (sw) {
            s->halt = 1;
            s->switch_n = s->n;
        }
    }
}

__device__ __forceinline__ bool cheby_norm_due(const SolveState *s)
{
    return s->n >= s->est_itc && (s->n % 10) == 0;
}

__device__ __forceinline__ void finish_results(const RedArgs &ra, const double *tot)
{
#pragma unroll
    for (int q = 0; q < MAX_RED; ++q) ra.result[q] = tot[q];
    // the release orders the totals before the flag for the host's reader
    if (ra.flag) __hip_atomic_store(ra.flag, ra.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The scalar algebra of the mode on s, a copy of the solver state in
// registers that the caller writes back
__device__ __forceinline__ void finish_state(const RedArgs &ra, const double *tot, SolveState *s)
{
    switch (ra.mode) {
    case RED_INIT_RESID:
        s->old_error = tot[0];
        s->initial_residual = sqrt(tot[0]);
        s->n = 1;
        break;
    case RED_INIT_RRO: {
        s->rro = tot[0];
        s->beta = 0.0;
        s->n = 1;
        if (s->solver == SOLVER_CHEBY || s->solver == SOLVER_PPCG) {
            const bool sw = fabs(s->old_error) <= s->epslim && s->n >= s->presteps;
            if (sw) {
                s->halt = 1;
                s->switch_n = 1;
            }
        }
        break;
    }
    case RED_CG_PW:
        s->pw = tot[0];
        s->alpha_prev = s->alpha;
        s->alpha = s->rro / s->pw;
        if (ra.hist_a) ra.hist_a[s->n] = s->alpha;
        break;
    case RED_CG_RRN:
        s->rrn = tot[0];
        s->beta = s->rrn / s->rro;
        if (ra.hist_b) ra.hist_b[s->n] = s->beta;
        s->error = s->rrn;
        s->rro = s->rrn;
        s->cg_calc_steps += 1;
        iteration_tail(s, ra.hist_e);
        break;
    case RED_PPCG_PW:
        s->pw = tot[0];
        s->alpha = s->rro / s->pw;
        break;
    case RED_PPCG_RRN:
        s->rrn = tot[0];
        s->beta = s->rrn / s->rro;
        s->error = s->rrn;
        s->rro = s->rrn;
        s->cheby_calc_steps += 1;
        iteration_tail(s, ra.hist_e);
        break;
    case RED_JACOBI:
        s->error = tot[0];
        iteration_tail(s, ra.hist_e);
        break;
    case RED_CHEBY:
        if (cheby_norm_due(s)) s->error = tot[0];
        s->cheby_calc_steps += 1;
        iteration_tail(s, ra.hist_e);
        break;
    case RED_CHEBY2:
        // two iterations; the first (odd n) never forms the norm
        s->cheby_calc_steps += 1;
        iteration_tail(s, ra.hist_e);
        if (s->halt) break;
        if (cheby_norm_due(s)) s->error = tot[0];
        s->cheby_calc_steps += 1;
        iteration_tail(s, ra.hist_e);
        break;
    case RED_EXACT:
        s->exact_error = sqrt(tot[0]);
        break;
    default:
        break;
    }
}

__device__ __forceinline__ bool iteration_mode(int mode)
{
    return mode == RED_CG_PW || mode == RED_CG_RRN || mode == RED_PPCG_PW || mode == RED_PPCG_RRN ||
           mode == RED_JACOBI || mode == RED_CHEBY || mode == RED_CHEBY2;
}

// Sums the partials of the preceding reduction kernel in block order and runs
// the scalar algebra of its mode. An iteration-mode finisher of a halted
// solver changes nothing, like the kernel before it. The state is copied into
// registers as the partials are loaded, so that its loads overlap theirs and
// the algebra makes no round trips to memory; the copy is stored back whole.
template <int NQ>
__global__ void __launch_bounds__(FINISH_THREADS) k_finish(RedArgs ra, int nblocks)
{
    double tot[MAX_RED];
    // (the four-quantity reduction, the field summary, carries no solver state)
    if (NQ == MAX_RED || !ra.st) {
        sum_partials<NQ>(ra.partials, nblocks, tot);
        if (threadIdx.x == 0) finish_results(ra, tot);
        return;
    }
    SolveState loc = *ra.st;
    sum_partials<NQ>(ra.partials, nblocks, tot);
    if (threadIdx.x != 0 || (iteration_mode(ra.mode) && (loc.halt || loc.fault))) return;
    finish_results(ra, tot);
    finish_state(ra, tot, &loc);
    *ra.st = loc;
}

// the finisher of a reduction kernel with nq sums per block
void launch_finish(hipStream_t s, const RedArgs &ra, int nblocks, int nq)
{
    if (nq <= 1) hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(FINISH_THREADS), 0, s, ra, nblocks);
    else if (nq == 2) hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(FINISH_THREADS), 0, s, ra, nblocks);
    else if (nq == 3) hipLaunchKernelGGL(k_finish<3>, dim3(1), dim3(FINISH_THREADS), 0, s, ra, nblocks);
    else hipLaunchKernelGGL(k_finish<MAX_RED>, dim3(1), dim3(FINISH_THREADS), 0, s, ra, nblocks);
}

__device__ __forceinline__ bool halted(const SolveState *st) { return st && st->halt; }

// ===========================================================================
// reference-granularity kernels
// ===========================================================================

// initialise_chunk_first (initialise_chunk_cl.cl:3-49): 1-D coordinate arrays
__global__ void k_init_chunk_1d(Geom g, double xmin, double ymin, double dx, double dy,
                                double *vx, double *vdx, double *vy, double *vdy, double *cx,
                                double *cdx, double *cy, double *cdy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int hd = g.hd;
    if (i <= g.x_max + hd) {
        vx[i] = xmin + dx * (double)(((i)-1) - hd + 1);
        vdx[i] = dx;
        const double vplus = xmin + dx * (double)((i)-hd + 1);
        cx[i] = 0.5 * (vx[i] + vplus);
        cdx[i] = dx;
    }
    if (i <= g.y_max + hd) {
        vy[i] = ymin + dy * (double)(((i)-1) - hd + 1);
        vdy[i] = dy;
        const double vplus = ymin + dy * (double)((i)-hd + 1);
        cy[i] = 0.5 * (vy[i] + vplus);
        cdy[i] = dy;
    }
}

// initialise_chunk_second (initialise_chunk_cl.cl:51-68)
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_chunk_2d(Geom g, Ext e, double dx, double dy,
                                                                 double *vol, double *xa, double *ya)
{
    TL_TILE_LOOP(e)
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = IX(g, j, ks + q);
        vol[i] = dx * dy;
        xa[i] = dy;  // xarea lives in its own pitched buffer
        ya[i] = dx;
    }
    TL_TILE_LOOP_END
}

// generate_chunk_init (generate_chunk_cl.cl:82-95) + every later state in
// order (:3-65, one pass per cell instead of one launch per state) +
// generate_chunk_init_u from energy0 (:67-80)
__global__ void __launch_bounds__(BLOCK_THREADS) k_generate(
    Geom g, Ext e, const double *vx, const double *vy, const double *cx, const double *cy,
    double *rho, double *e0, double *u, double *u0, int nstates, const double *sdens,
    const double *sener, const double *sxmin, const double *sxmax, const double *symin,
    const double *symax, const double *srad, const int *sgeom, int g_rect, int g_circ, int g_point)
{
    TL_TILE_LOOP(e)
    for (int q = 0; q < STRIP; ++q) {
        const int k = ks + q;
        if (k > e.k1) break;
        double en = sener[0], de = sdens[0];
        for (int s = 1; s < nstates; ++s) {
            bool hit = false;
            if (g_rect == sgeom[s]) {
                hit = (vx[1 + j] >= sxmin[s] && vx[j] < sxmax[s] && vy[1 + k] >= symin[s] &&
                       vy[k] < symax[s]);
            } else if (sgeom[s] == g_circ) {
                const double xp = cx[j] - sxmin[s], yp = cy[k] - symin[s];
                const double rad = sqrt(xp * xp + yp * yp);
                hit = (rad <= srad[s]);
            } else if (sgeom[s] == g_point) {
                hit = (vx[j] == sxmin[s] && vy[k] == symin[s]);
            }
            if (hit) {
                en = sener[s];
                de = sdens[s];
            }
        }
        const int i = IX(g, j, k);
        e0[i] = en;
        rho[i] = de;
        u[i] = en * de;
        u0[i] = en * de;
    }
    TL_TILE_LOOP_END
}

// set_field (set_field_cl.cl:3-13) and generic interior copy
__global__ void __launch_bounds__(BLOCK_THREADS) k_copy(Geom g, Ext e, const double *src, double *dst)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks);
    double v[STRIP];
    col4(src, i0, g.pitch, v);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) dst[i0 + q * g.pitch] = v[q];
    TL_TILE_LOOP_END
}

// Streaming reference with the stream mix of k_fused_cg_ur (4 fields read,
// 3 written, no stencil): a += s b, c -= s d, e = c d. Used only to measure
// the attainable bandwidth of that mix.
__global__ void __launch_bounds__(BLOCK_THREADS) k_stream_mix(Geom g, Ext e, double *a, const double *b, double *c,
                                                              const double *d, double *out, double s)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double av[STRIP], bv[STRIP], cv[STRIP], dv[STRIP];
    col4(a, i0, P, av);
    col4(b, i0, P, bv);
    col4(c, i0, P, cv);
    col4(d, i0, P, dv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const double cn = cv[q] - s * dv[q];
        a[i0 + q * P] = av[q] + s * bv[q];
        c[i0 + q * P] = cn;
        out[i0 + q * P] = cn * dv[q];
    }
    TL_TILE_LOOP_END
}

// field_summary (field_summary_cl.cl:3-43), energy1 bound to `energy`
__global__ void __launch_bounds__(BLOCK_THREADS) k_field_summary(Geom g, Ext e, const double *vol,
                                                                 const double *rho, const double *en,
                                                                 const double *u, RedArgs ra)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double v[STRIP], d[STRIP], x[STRIP], y[STRIP];
    col4(vol, i0, P, v);
    col4(rho, i0, P, d);
    col4(en, i0, P, x);
    col4(u, i0, P, y);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const double cm = v[q] * d[q];
        acc[0] += v[q];
        acc[1] += cm;
        acc[2] += cm * x[q];
        acc[3] += cm * y[q];
    }
    TL_TILE_LOOP_END
    block_partial<4>(acc, ra.partials);
}

// update_halo_{bottom,top} (update_halo_cl.cl:43-69): one thread per column
// of the depth rows; face 2 = bottom, 3 = top
__global__ void k_halo_bt(Geom g, double *a, int depth, int face)
{
    const int j = (g.hd - depth) + blockIdx.x * blockDim.x + threadIdx.x;
    if (j > g.x_max + g.hd - 1 + depth) return;
    for (int k = g.hd - depth; k <= g.hd - 1; ++k) {
        if (face == 2) a[IX(g, j, k)] = 1 * a[IX(g, j, 2 * g.hd - 1 - k)];
        else a[IX(g, j, g.y_max + 2 * g.hd - 1 - k)] = a[IX(g, j, g.y_max + k)];
    }
}

// update_halo_{left,right} (update_halo_cl.cl:3-41); face 0 = left, 1 = right
__global__ void k_halo_lr(Geom g, double *a, int depth, int face)
{
    const int k = (g.hd - depth) + blockIdx.x * blockDim.x + threadIdx.x;
    if (k > g.y_max + g.hd - 1 + depth) return;
    for (int j = g.hd - depth; j <= g.hd - 1; ++j) {
        if (face == 0) a[IX(g, j, k)] = 1 * a[IX(g, 2 * g.hd - 1 - j, k)];
        else a[IX(g, g.x_max + 2 * g.hd - 1 - j, k)] = a[IX(g, g.x_max + j, k)];
    }
}

// tea_leaf_init_common (tea_leaf_common_cl.cl:56-90), launch bounds (1-hd, hd)
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_common(Geom g, Ext e, const double *rho,
                                                               double *Kx, double *Ky, double rx,
                                                               double ry, int coef)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double c5[STRIP + 1], w4[STRIP];
#pragma unroll
    for (int q = 0; q < STRIP + 1; ++q) c5[q] = rho[i0 + (q - 1) * P];  // rows ks-1 .. ks+3
    col4(rho, i0 - 1, P, w4);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        double dc, dl, dd;
        if (coef == COEF_CONDUCTIVITY) {
            dc = c5[q + 1];
            dl = w4[q];
            dd = c5[q];
        } else {
            dc = 1.0 / c5[q + 1];
            dl = 1.0 / w4[q];
            dd = 1.0 / c5[q];
        }
        double kx = (dl + dc) / (2.0 * dl * dc);
        kx *= rx;
        double ky = (dd + dc) / (2.0 * dd * dc);
        ky *= ry;
        Kx[i0 + q * P] = kx;
        Ky[i0 + q * P] = ky;
    }
    TL_TILE_LOOP_END
}

// conductivity of tea_leaf_init_common (tea_leaf_common_cl.cl:56-90: density,
// or its reciprocal) over the whole padded index space, kept for the fused
// iteration kernels, which recompute Kx, Ky from it
// (`bad`: set when an interior density lies outside [2^-499, 2^499] or is not
// a number. Inside that range every divisor the iteration kernels form from
// the conductivities -- face denominators 2 w_a w_b, and Thomas pivots and
// diagonals >= 1 -- has a binary exponent within [-999, 999], where div_nr is
// the IEEE quotient.)
__global__ void __launch_bounds__(256) k_conductivity(Geom g, const double *rho, double *w, int coef, int *bad)
{
    const int nx = g.x_max + 2 * g.hd, ny = g.y_max + 2 * g.hd;
    const long n = (long)nx * ny;
    bool out = false;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t % nx), k = (int)(t / nx);
        const double d = rho[IX(g, j, k)];
        w[IX(g, j, k)] = (coef == COEF_CONDUCTIVITY) ? d : 1.0 / d;
        const bool interior = j >= g.hd && j < g.hd + g.x_max && k >= g.hd && k < g.hd + g.y_max;
        out |= interior && !(d >= 0x1p-499 && d <= 0x1p499);
    }
    if (out) *bad = 1;
}

// tea_leaf_zero_boundary (tea_leaf_common_cl.cl:92-125), launch bounds (-hd, hd)
__global__ void __launch_bounds__(BLOCK_THREADS) k_zero_boundary(Geom g, Ext e, double *Kx, double *Ky,
                                                                 int zl, int zr, int zb, int zt)
{
    TL_TILE_LOOP(e)
    for (int q = 0; q < STRIP; ++q) {
        const int k = ks + q;
        if (k > e.k1) break;
        const int i = IX(g, j, k);
        const int hd = g.hd;
        if (zl && j <= hd) {
            Kx[i] = 0;
            if (j < hd) Ky[i] = 0;
        }
        if (zr && j > (g.x_max - 1) + hd) {
            Kx[i] = 0;
            Ky[i] = 0;
        }
        if (zb && k <= hd) {
            Ky[i] = 0;
            if (k < hd) Kx[i] = 0;
        }
        if (zt && k > (g.y_max - 1) + hd) {
            Ky[i] = 0;
            Kx[i] = 0;
        }
    }
    TL_TILE_LOOP_END
}

// generate_chunk_init_u with energy1 (tea_leaf_kernel_ocl.cpp:287-288)
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_u(Geom g, Ext e, const double *rho,
                                                          const double *en, double *u, double *u0)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double d[STRIP], x[STRIP];
    col4(rho, i0, P, d);
    col4(en, i0, P, x);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        u[i0 + q * P] = x[q] * d[q];
        u0[i0 + q * P] = x[q] * d[q];
    }
    TL_TILE_LOOP_END
}

// tea_leaf_init_jac_diag (tea_leaf_common_cl.cl:127-142), launch bounds (-hd, hd)
// z = Mi r (the z that tea_leaf_cg_cl.cl:8-41 forms with the diagonal
// preconditioner), for the fresh-z PPCG option
__global__ void __launch_bounds__(BLOCK_THREADS) k_diag_apply(Geom g, Ext e, const double *r, const double *Mi,
                                                              double *z)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) z[i0 + q * P] = r[i0 + q * P] * Mi[i0 + q * P];
    TL_TILE_LOOP_END
}

__global__ void __launch_bounds__(BLOCK_THREADS) k_init_jac_diag(Geom g, Ext e, double *Mi,
                                                                 const double *Kx, const double *Ky)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1];
    strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) Mi[i0 + q * P] = 1.0 / diag_of(kx0[q], kx1[q], ky[q], ky[q + 1]);
    TL_TILE_LOOP_END
}

// tea_leaf_calc_residual (tea_leaf_common_cl.cl:17-36)
__global__ void __launch_bounds__(BLOCK_THREADS) k_calc_residual(Geom g, Ext e, const double *u,
                                                                 const double *u0, double *r,
                                                                 const double *Kx, const double *Ky)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], w[STRIP], b[STRIP];
    strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    col4(u0, i0, P, b);
    strip_matvec(u, kx0, kx1, ky, i0, P, w, nullptr);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) r[i0 + q * P] = b[q] - w[q];
    TL_TILE_LOOP_END
}

// tea_leaf_calc_2norm (tea_leaf_common_cl.cl:38-54)
__global__ void __launch_bounds__(BLOCK_THREADS) k_dot(Geom g, Ext e, const double *a, const double *b,
                                                       RedArgs ra)
{
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double x[STRIP], y[STRIP];
    col4(a, i0, P, x);
    col4(b, i0, P, y);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) acc[0] += x[q] * y[q];
    TL_TILE_LOOP_END
    block_partial<1>(acc, ra.partials);
}

// tea_leaf_finalise (tea_leaf_common_cl.cl:4-15)
__global__ void __launch_bounds__(BLOCK_THREADS) k_finalise(Geom g, Ext e, const double *rho,
                                                            const double *u, double *en)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double d[STRIP], x[STRIP];
    col4(rho, i0, P, d);
    col4(u, i0, P, x);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) en[i0 + q * P] = x[q] / d[q];
    TL_TILE_LOOP_END
}

// tea_leaf_block_init (tea_block_jacobi.cl:78-105)
__global__ void __launch_bounds__(BLOCK_THREADS) k_block_init(Geom g, Ext e, double *cp, double *bfp,
                                                              const double *Kx, const double *Ky)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], a[STRIP], b[STRIP], c[STRIP], cpv[STRIP], bfv[STRIP];
    strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    strip_abc(kx0, kx1, ky, a, b, c);
    const int upper = block_top(g, ks);
    strip_factors(upper, b, a, c, cpv, bfv);
    cp[i0] = cpv[0];
#pragma unroll
    for (int q = 1; q < STRIP; ++q)
        if (q < upper) {
            bfp[i0 + q * P] = bfv[q];
            cp[i0 + q * P] = cpv[q];
        }
    TL_TILE_LOOP_END
}

// tea_leaf_block_solve (tea_block_jacobi.cl:44-76)
__global__ void __launch_bounds__(BLOCK_THREADS) k_block_solve(Geom g, Ext e, const double *r,
                                                               double *z, const double *cp,
                                                               const double *bfp, const double *Kx,
                                                               const double *Ky)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], rv[STRIP], zv[STRIP];
    strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    col4(r, i0, P, rv);
    mask_rows(ks, e.k1, rv);
    strip_block(g, ks, i0, kx0, kx1, ky, cp, bfp, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) z[i0 + q * P] = zv[q];
    TL_TILE_LOOP_END
}

// tea_leaf_cg_solve_init_p (tea_leaf_cg_cl.cl:8-41)
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_cg_init_p(Geom g, Ext e, double *p, const double *r,
                                                             double *z, const double *Mi, RedArgs ra)
{
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double rv[STRIP], x[STRIP];
    col4(r, i0, P, rv);
    if (PREC == PREC_BLOCK) col4(z, i0, P, x);
    else if (PREC == PREC_DIAG) col4(Mi, i0, P, x);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        double pv;
        if (PREC == PREC_BLOCK) {
            pv = x[q];
        } else if (PREC == PREC_DIAG) {
            pv = rv[q] * x[q];
            z[i] = pv;
        } else {
            pv = rv[q];
        }
        p[i] = pv;
        acc[0] += rv[q] * pv;
    }
    TL_TILE_LOOP_END
    block_partial<1>(acc, ra.partials);
}

// tea_leaf_cg_solve_calc_w (tea_leaf_cg_cl.cl:45-69)
__global__ void __launch_bounds__(BLOCK_THREADS) k_cg_calc_w(Geom g, Ext e, const double *p, double *w,
                                                             const double *Kx, const double *Ky,
                                                             const double *W, Faces fc, RedArgs ra)
{
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], wv[STRIP], pc[STRIP];
    if (W) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    strip_matvec(p, kx0, kx1, ky, i0, P, wv, pc);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        store_nt(w + i0 + q * P, wv[q]);
        acc[0] += pc[q] * wv[q];
    }
    TL_TILE_LOOP_END
    block_partial<1>(acc, ra.partials);
}

// tea_leaf_cg_solve_calc_ur (tea_leaf_cg_cl.cl:73-137)
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_cg_calc_ur(Geom g, Ext e, double alpha, double *u,
                                                              const double *p, double *r,
                                                              const double *w, double *z,
                                                              const double *cp, const double *bfp,
                                                              const double *Mi, const double *Kx,
                                                              const double *Ky, const double *W, Faces fc,
                                                              RedArgs ra)
{
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double uv[STRIP], pv[STRIP], rv[STRIP], wv[STRIP], x[STRIP], zv[STRIP];
    col4_nt(u, i0, P, uv);
    col4_nt(p, i0, P, pv);
    col4_nt(r, i0, P, rv);
    col4_nt(w, i0, P, wv);
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1];
    if (PREC == PREC_BLOCK) {
        if (W) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
        else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    }
    if (PREC == PREC_DIAG) col4_nt(Mi, i0, P, x);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        uv[q] += alpha * pv[q];
        rv[q] -= alpha * wv[q];
    }
    mask_rows(ks, e.k1, rv);
    if (PREC == PREC_BLOCK) strip_block(g, ks, i0, kx0, kx1, ky, cp, bfp, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        store_nt(u + i, uv[q]);
        store_nt(r + i, rv[q]);
        if (PREC == PREC_BLOCK) {
            store_nt(z + i, zv[q]);
            acc[0] += rv[q] * zv[q];
        } else if (PREC == PREC_DIAG) {
            const double zz = rv[q] * x[q];
            store_nt(z + i, zz);
            acc[0] += rv[q] * zz;
        } else {
            acc[0] += rv[q] * rv[q];
        }
    }
    TL_TILE_LOOP_END
    block_partial<1>(acc, ra.partials);
}

// tea_leaf_cg_solve_calc_p (tea_leaf_cg_cl.cl:141-160)
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_cg_calc_p(Geom g, Ext e, double beta, double *p,
                                                             const double *r, const double *z)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double x[STRIP], pv[STRIP];
    col4_nt(PREC != PREC_NONE ? z : r, i0, P, x);
    col4_nt(p, i0, P, pv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) store_nt(p + i0 + q * P, x[q] + beta * pv[q]);
    TL_TILE_LOOP_END
}

// tea_leaf_cheby_solve_init_p (INIT=1, tea_leaf_cheby_cl.cl:4-64) and
// tea_leaf_cheby_solve_calc_p (INIT=0, :78-145). The coefficient index comes
// from `step` or, when st is set, from the device-resident step counter; the
// factors come from cp/bfp, or are recomputed when cp is null.
template <int PREC, int INIT>
__global__ void __launch_bounds__(BLOCK_THREADS) k_cheby_p(Geom g, Ext e, const double *u,
                                                           const double *u0, double *p, double *r,
                                                           double *w, const double *cp,
                                                           const double *bfp, const double *Mi,
                                                           const double *Kx, const double *Ky,
                                                           const double *W, Faces fc,
                                                           const double *ch_a, const double *ch_b,
                                                           double theta, int step,
                                                           const SolveState *st)
{
    if (halted(st)) return;
    if (st) step = st->cheby_calc_steps - 1;
    const double ca = INIT ? 0.0 : ch_a[step], cb = INIT ? 0.0 : ch_b[step];
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], wv[STRIP], b[STRIP], pv[STRIP], x[STRIP], rv[STRIP];
    if (W) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    col4_nt(u0, i0, P, b);
    if (!INIT) col4_nt(p, i0, P, pv);
    if (PREC == PREC_DIAG) col4_nt(Mi, i0, P, x);
    strip_matvec(u, kx0, kx1, ky, i0, P, wv, nullptr);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] = b[q] - wv[q];
    mask_rows(ks, e.k1, rv);
    double zv[STRIP];
    if (PREC == PREC_BLOCK) strip_block(g, ks, i0, kx0, kx1, ky, cp, bfp, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        if (w) store_nt(w + i, wv[q]);
        store_nt(r + i, rv[q]);
        double np;
        if (PREC == PREC_BLOCK) np = INIT ? zv[q] / theta : ca * pv[q] + cb * zv[q];
        else if (PREC == PREC_DIAG) np = INIT ? (x[q] * rv[q]) / theta : ca * pv[q] + cb * x[q] * rv[q];
        else np = INIT ? rv[q] / theta : ca * pv[q] + cb * rv[q];
        store_nt(p + i, np);
    }
    TL_TILE_LOOP_END
}

// tea_leaf_cheby_solve_calc_u (tea_leaf_cheby_cl.cl:66-76). With st set this
// is the last kernel of a device-resident Chebyshev iteration: it also sums
// r.r when tea_solve.f90:223 asks for it; its finisher runs the iteration tail.
__global__ void __launch_bounds__(BLOCK_THREADS) k_cheby_u(Geom g, Ext e, double *u, const double *p,
                                                           const double *r, RedArgs ra)
{
    if (halted(ra.st)) return;
    const bool want = ra.st && cheby_norm_due(ra.st);
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double uv[STRIP], pv[STRIP], rv[STRIP];
    col4_nt(u, i0, P, uv);
    col4_nt(p, i0, P, pv);
    if (want) col4_nt(r, i0, P, rv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        store_nt(u + i0 + q * P, uv[q] + pv[q]);
        if (want) acc[0] += rv[q] * rv[q];
    }
    TL_TILE_LOOP_END
    if (ra.st) block_partial<1>(acc, ra.partials);
}

// tea_leaf_ppcg_solve_init_sd (tea_leaf_ppcg_cl.cl:4-60)
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_ppcg_init_sd(Geom g, Ext e, const double *r,
                                                                double *sd, const double *cp,
                                                                const double *bfp, const double *Mi,
                                                                const double *Kx, const double *Ky,
                                                                double theta)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double rv[STRIP], x[STRIP], zv[STRIP], kx0[STRIP], kx1[STRIP], ky[STRIP + 1];
    col4(r, i0, P, rv);
    if (PREC == PREC_DIAG) col4(Mi, i0, P, x);
    if (PREC == PREC_BLOCK) {
        strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
        mask_rows(ks, e.k1, rv);
        strip_block(g, ks, i0, kx0, kx1, ky, cp, bfp, rv, zv);
    }
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        if (PREC == PREC_BLOCK) sd[i] = zv[q] / theta;
        else if (PREC == PREC_DIAG) sd[i] = rv[q] * x[q] / theta;
        else sd[i] = rv[q] / theta;
    }
    TL_TILE_LOOP_END
}

// tea_leaf_ppcg_solve_update_r (tea_leaf_ppcg_cl.cl:62-88) on the extent e
__global__ void __launch_bounds__(BLOCK_THREADS) k_ppcg_update_r(Geom g, Ext e, double *u, double *r,
                                                                 const double *Kx, const double *Ky,
                                                                 const double *W, Faces fc, const double *sd)
{
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], res[STRIP], sc[STRIP], uv[STRIP], rv[STRIP];
    if (W) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    col4_nt(u, i0, P, uv);
    col4_nt(r, i0, P, rv);
    strip_matvec(sd, kx0, kx1, ky, i0, P, res, sc);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        store_nt(u + i0 + q * P, uv[q] + sc[q]);
        store_nt(r + i0 + q * P, rv[q] - res[q]);
    }
    TL_TILE_LOOP_END
}

// tea_leaf_ppcg_solve_calc_sd (tea_leaf_ppcg_cl.cl:90-155) on the extent e
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_ppcg_calc_sd(Geom g, Ext e, const double *r,
                                                                double *sd, const double *cp,
                                                                const double *bfp, const double *Mi,
                                                                const double *Kx, const double *Ky,
                                                                const double *W, Faces fc,
                                                                const double *ch_a,
                                                                const double *ch_b, int step)
{
    const double ca = ch_a[step], cb = ch_b[step];
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double rv[STRIP], sv[STRIP], x[STRIP], zv[STRIP], kx0[STRIP], kx1[STRIP], ky[STRIP + 1];
    col4_nt(r, i0, P, rv);
    col4(sd, i0, P, sv);
    if (PREC == PREC_DIAG) col4_nt(Mi, i0, P, x);
    if (PREC == PREC_BLOCK) {
        if (W) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
        else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
        mask_rows(ks, e.k1, rv);
        strip_block(g, ks, i0, kx0, kx1, ky, cp, bfp, rv, zv);
    }
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        if (PREC == PREC_BLOCK) sd[i] = ca * sv[q] + cb * zv[q];
        else if (PREC == PREC_DIAG) sd[i] = ca * sv[q] + cb * rv[q] * x[q];
        else sd[i] = ca * sv[q] + cb * rv[q];
    }
    TL_TILE_LOOP_END
}

// tea_leaf_jacobi_solve (tea_leaf_jacobi_cl.cl:15-42); `un` holds the
// previous iterate. The device-resident driver passes the two buffers of a
// pair and st: iteration n reads buffer (n-1)&1 (this replaces the
// reference's copy of u into the Mi buffer, tea_leaf_jacobi_cl.cl:3-13).
// COND=1 (device-resident driver) forms the faces from the conductivity W
// (bitwise the stored Kx, Ky) instead of reading Kx, Ky: 32 instead of 40
// B/cell.
template <int COND>
__global__ void __launch_bounds__(BLOCK_THREADS) k_jacobi(Geom g, Ext e, const double *Kx,
                                                          const double *Ky, const double *W, Faces fc,
                                                          const double *u0, double *uA, double *uB, RedArgs ra)
{
    if (halted(ra.st)) return;
    const double *un = uB;
    double *u1 = uA;
    if (ra.st && (ra.st->n & 1)) {
        un = uA;
        u1 = uB;
    }
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], b[STRIP], c6[STRIP + 2], w4[STRIP], e4[STRIP];
    if (COND) strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    else strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    col4(u0, i0, P, b);
    col6(un, i0, P, c6);
    col4(un, i0 - 1, P, w4);
    col4(un, i0 + 1, P, e4);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const double v = (b[q] + kx1[q] * e4[q] + kx0[q] * w4[q] + ky[q + 1] * c6[q + 2] + ky[q] * c6[q]) /
                         (1.0 + (kx0[q] + kx1[q]) + (ky[q] + ky[q + 1]));
        u1[i0 + q * P] = v;
        acc[0] += fabs(v - c6[q + 1]);
    }
    TL_TILE_LOOP_END
    block_partial<1>(acc, ra.partials);
}

// ===========================================================================
// fused iteration kernels (device-resident driver)
// ===========================================================================

// Fused CG search direction and matvec: p = z + beta p_old (the previous
// iteration's tea_leaf_cg_solve_calc_p, tea_leaf_cg_cl.cl:141-160) folded into
// w = A p and p.w (calc_w, :45-69); p is double-buffered by the iteration
// counter and w is not stored. A band march (TL_MARCH_LOOP): each row of
// z, p_old and the conductivity is loaded once per band, the new p of the
// rows above and below and the face below are kept in registers. NTZ: z is
// read with nontemporal loads, so that the caches keep more of p_old and the
// conductivity, whose halo rows and overlap columns neighbouring segments
// and bands read again: in CG iterations (at 4096^2) 105.6 -> 98.6 us per
// launch, and the update kernel that follows 0.152 -> 0.149 ms; the same hint
// on p_old or on the conductivity as well measured slower, and in the PPCG
// outer step, after the inner steps, the hint made the kernel slower (113 ->
// 122 us), as it did at 8192^2 (457 -> 495 us): those launches read z with
// plain loads.
// Algorithmic bytes: R z, p_old, cond, W p = 32 B/cell.
template <bool NTZ>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_cg_w(Geom g, Ext e, March m, const double *src,
                                                              double *pA, double *pB, const double *W, Faces fc,
                                                              RedArgs ra)
{
    const SolveState *st = ra.st;
    if (st->halt || st->fault) return;
    const double beta = st->beta;
    const bool odd = (st->n & 1) != 0;
    const double *p_old = odd ? pA : pB;
    double *p_new = odd ? pB : pA;
    const int lane = threadIdx.x & 63;
    const int P = g.pitch, hd = g.hd;
    double acc = 0.0;
    TL_MARCH_LOOP(m, e)
    const int j = jb - 1 + lane;
    const bool out_lane = lane >= 1 && lane <= MARCH_COLS && j <= e.j1;
    // every face the band's outputs use lies strictly inside the chunk in x
    const bool inner_x = jb - 1 > hd && jb + MARCH_COLS < hd + g.x_max;
    // rows k-1 (pa) and k (pb) of the new p, conductivity of row k, the face
    // between them (kyb)
    int i = IX(g, j, k_begin - 1);
    double pa = ld_nt<NTZ>(src + i) + beta * p_old[i];
    const double ca = W[i];
    i += P;
    double pb = ld_nt<NTZ>(src + i) + beta * p_old[i];
    double cb = W[i];
    double kyb = ky_at(g, fc, j, k_begin, ca, cb);
    // loads of the next two rows in flight
    constexpr int PF = 2;
    double zr[PF], pr[PF], cr[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int ii = i + (q + 1) * P;
        zr[q] = ld_nt<NTZ>(src + ii);
        pr[q] = p_old[ii];
        cr[q] = W[ii];
    }
    for (int k = k_begin; k <= k_end; k += PF) {
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int kk = k + q;
            const double pc = zr[q] + beta * pr[q];  // row kk+1
            const double cc = cr[q];
            {
                const int ii = IX(g, j, kk + 1 + PF);
                zr[q] = ld_nt<NTZ>(src + ii);
                pr[q] = p_old[ii];
                cr[q] = W[ii];
            }
            if (kk > k_end) continue;
            const bool inner = inner_x && kk > hd && kk + 1 < hd + g.y_max;
            const double kyn = march_ky(g, fc, inner, j, kk + 1, cb, cc);
            const double kx0 = march_kx(g, fc, inner, j, kk, dpp_shr1(cb, 0.0), cb);
            const double kx1 = dpp_shl1(kx0, 0.0);
            const double wv = opA(kx0, kx1, kyb, kyn, pb, dpp_shr1(pb, 0.0), dpp_shl1(pb, 0.0), pa, pc);
            if (out_lane) {
                store_nt(p_new + IX(g, j, kk), pb);
                acc += pb * wv;
            }
            pa = pb;
            pb = pc;
            cb = cc;
            kyb = kyn;
        }
    }
    TL_MARCH_LOOP_END
    double a1[1] = {acc};
    block_partial<1>(a1, ra.partials);
}

// CG / PPCG outer step, second half (tea_leaf_cg_cl.cl:73-137): w = A p
// recomputed, u += alpha p, r -= alpha w, z = M^-1 r with the diagonal or
// block factors recomputed in registers, r.z. PPCG=1 also writes
// sd = z/theta (tea_leaf_ppcg_cl.cl:4-60 on the same r) and skips the unused
// r.z sum. r_out may be r (in place) or a second buffer (the PPCG outer step
// when the inner steps that follow flip r between two buffers).
// For CG (PPCG=0) the update of u, which nothing in the iteration reads, is
// made every second iteration: iteration n odd leaves u alone, iteration n
// even adds alpha_{n-1} p_{n-1} (still in the other p buffer) and then
// alpha_n p_n, each as one FMA and in that order, so u is bitwise what one
// fused update per iteration gives; a CG phase that ends on an odd iteration
// is completed by k_cg_u_fix. This saves u's read and write on every other
// iteration (8 of 88 B/cell per iteration on average).
// PAR: 0 takes the parity of n from the state; 1 / 2 compile the odd / even
// form only (the driver captures each for iterations of that parity; run on
// the other parity it sets the state's fault flag instead, which stops every
// later kernel and the solve).
template <int PREC, int PPCG, int PAR = 0>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_cg_ur(Geom g, Ext e, double *u,
                                                               const double *pA, const double *pB,
                                                               const double *r, double *r_out, double *z,
                                                               double *sd, const double *W, Faces fc,
                                                               RedArgs ra)
{
    SolveState *st = ra.st;
    if (st->halt || st->fault) return;
    const bool odd = PAR ? PAR == 1 : (st->n & 1) != 0;
    if (PAR && odd != ((st->n & 1) != 0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fault = 1;
        return;
    }
    const double *p = odd ? pB : pA;
    const double *p_prev = odd ? pA : pB;
    const double alpha = st->alpha, alpha_prev = st->alpha_prev;
    const double theta = PPCG ? st->theta : 1.0;
    const bool upd = PPCG || !odd;
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], wv[STRIP], pc[STRIP], uv[STRIP], rv[STRIP], zv[STRIP];
    strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    if (upd) col4_nt(u, i0, P, uv);
    col4_nt(r, i0, P, rv);
    strip_matvec(p, kx0, kx1, ky, i0, P, wv, pc);
    if (!PPCG && upd) {
        double pp[STRIP];
        col4_nt(p_prev, i0, P, pp);
#pragma unroll
        for (int q = 0; q < STRIP; ++q) uv[q] = fma(alpha_prev, pp[q], uv[q]);
    }
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (upd) uv[q] = fma(alpha, pc[q], uv[q]);
        rv[q] -= alpha * wv[q];
    }
    mask_rows(ks, e.k1, rv);
    if (PREC == PREC_BLOCK) strip_block(g, ks, i0, kx0, kx1, ky, nullptr, nullptr, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        if (upd) store_nt(u + i, uv[q]);
        store_nt(r_out + i, rv[q]);
        double zz = rv[q];
        if (PREC == PREC_BLOCK) zz = zv[q];
        if (PREC == PREC_DIAG) zz = rv[q] * div_nr(1.0, diag_of(kx0[q], kx1[q], ky[q], ky[q + 1]));
        if (PREC != PREC_NONE) store_nt(z + i, zz);
        if (PPCG) store_nt(sd + i, zz / theta);
        else acc[0] += rv[q] * zz;
    }
    TL_TILE_LOOP_END
    if (!PPCG) block_partial<1>(acc, ra.partials);
}

// Completes u after a CG phase of the fused kernels that ended on an odd
// iteration L: u += alpha_L p_L (p_L is in the second p buffer).
__global__ void __launch_bounds__(BLOCK_THREADS) k_cg_u_fix(Geom g, Ext e, double *u, const double *p,
                                                            const SolveState *st)
{
    const double alpha = st->alpha;
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        if (ks + q <= e.k1) u[i0 + q * P] = fma(alpha, p[i0 + q * P], u[i0 + q * P]);
    TL_TILE_LOOP_END
}

// One PPCG smoothing step on a strip once A sd is known (res = A sd, sc = sd):
// r -= A sd, then sd' = ca sd + cb M^-1 r (tea_leaf_ppcg_cl.cl:62-155); zf
// gets M^-1 r of the new r (r itself without a preconditioner). Shared by the
// one-step and two-step kernels, with the FMA placed explicitly, so that both
// round identically.
template <int PREC>
__device__ __forceinline__ void ppcg_step(const Geom &g, int ks, int k1, const double *kx0, const double *kx1,
                                          const double *ky, const double *res, const double *sc, double ca,
                                          double cb, double *rv, double *sdn, double *zf)
{
    double zv[STRIP];
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] -= res[q];
    mask_rows(ks, k1, rv);
    if (PREC == PREC_BLOCK) strip_block(g, ks, 0, kx0, kx1, ky, nullptr, nullptr, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (PREC == PREC_BLOCK) {
            sdn[q] = fma(ca, sc[q], cb * zv[q]);
            zf[q] = zv[q];
        } else if (PREC == PREC_DIAG) {
            const double mi = div_nr(1.0, diag_of(kx0[q], kx1[q], ky[q], ky[q + 1]));
            sdn[q] = fma(ca, sc[q], cb * rv[q] * mi);
            zf[q] = rv[q] * mi;
        } else {
            sdn[q] = fma(ca, sc[q], cb * rv[q]);
            zf[q] = rv[q];
        }
    }
}

// r.z of the last smoothing step: with the outer step's z (LAST = 1, the
// reference's stale z: tea_leaf_ppcg.f90:41-60 via calc_2norm(2) reads a z
// that calc_sd never writes), or with the fresh M^-1 r of this step, which is
// then stored as z for the next outer step (LAST = 2, the fresh-z option);
// r.r without a preconditioner
template <int PREC, int LAST>
__device__ __forceinline__ double ppcg_rz(double rv, double zo, double zf)
{
    return rv * (PREC == PREC_NONE ? rv : LAST == 2 ? zf : zo);
}

// PPCG inner step: update_r then calc_sd (tea_leaf_ppcg_cl.cl:62-155) fused;
// sd is double-buffered. LAST = 1, 2 also form r.z (ppcg_rz); the finisher
// runs the tail.
template <int PREC, int LAST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_ppcg_inner(
    Geom g, Ext e, double *u, double *r, const double *sd_in, double *sd_out, const double *W, Faces fc,
    double *z, const double *ch_a, const double *ch_b, int step, RedArgs ra)
{
    const SolveState *st = ra.st;
    if (st->halt) return;
    const double ca = ch_a[step], cb = ch_b[step];
    double acc[1] = {0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], res[STRIP], sc[STRIP], uv[STRIP], rv[STRIP], sdn[STRIP],
        zo[STRIP] = {}, zf[STRIP];
    strip_k_cond(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
    col4_nt(u, i0, P, uv);
    col4_nt(r, i0, P, rv);
    if (LAST == 1 && PREC != PREC_NONE) col4_nt(z, i0, P, zo);
    strip_matvec(sd_in, kx0, kx1, ky, i0, P, res, sc);
    ppcg_step<PREC>(g, ks, e.k1, kx0, kx1, ky, res, sc, ca, cb, rv, sdn, zf);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        store_nt(u + i, uv[q] + sc[q]);
        store_nt(r + i, rv[q]);
        store_nt(sd_out + i, sdn[q]);
        if (LAST == 2 && PREC != PREC_NONE) store_nt(z + i, zf[q]);
        if (LAST) acc[0] += ppcg_rz<PREC, LAST>(rv[q], zo[q], zf[q]);
    }
    TL_TILE_LOOP_END
    if (LAST) block_partial<1>(acc, ra.partials);
}

// sd at ring cell (j, k) when it is not computed in the block: the value the
// one-step kernels would read there, i.e. the halo cell of the sd buffer
// (j0-1, j1+1, k0-1 or k1+1; no interior kernel writes it); 0 for cells
// beyond the halo, which no interior cell reads.
__device__ __forceinline__ double ring_outside(const Geom &g, const Ext &e, const double *sd, int j, int k)
{
    const bool in = j >= e.j0 - 1 && j <= e.j1 + 1 && k >= e.k0 - 1 && k <= e.k1 + 1;
    return in ? sd[IX(g, j, k)] : 0.0;
}

// Two PPCG inner steps (steps `step` and `step`+1 of the smoother) in one
// pass over the fields. Step 1 is computed on the block's 64x16 tile and on
// the ring around it that step 2 reads: the 4-row strips directly above and
// below the tile (their block-Jacobi solve needs the whole strip) and the
// tile's strips in the columns left and right of it. The ring's new sd goes
// through LDS; the tile keeps r, u and the faces in registers. Every value is
// formed by the same expressions as the one-step kernel, so two one-step
// launches and one two-step launch give bitwise the same u, r and sd.
// The face coefficients are read from Kx, Ky (bitwise what the one-step
// kernels recompute from the conductivity): with the ring's extra strips this
// kernel is bound by arithmetic as much as by memory, and reading the faces
// instead of dividing measured 5 % faster per PPCG outer step at 4096^2.
// Algorithmic bytes per two steps: R u, r, sd, Kx, Ky  W u, r, sd (64 B/cell)
// instead of 112. Neighbouring blocks read r and sd of this block's tile in
// their step 1, so r and sd are both read from one buffer and written to
// another.
template <int PREC, int LAST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_ppcg_inner2(
    Geom g, Ext e, double *u, const double *r_in, double *r_out, const double *sd_in, double *sd_out,
    const double *Kx, const double *Ky, double *z, const double *ch_a, const double *ch_b, int step, RedArgs ra)
{
    const SolveState *st = ra.st;
    if (st->halt) return;
    constexpr int TH = STRIPS_PER_TILE * STRIP;  // tile rows
    constexpr int RW = TILE_X + 2, RH = TH + 2;   // ring-padded tile
    __shared__ double ring[RH][RW];
    const double ca1 = ch_a[step], cb1 = ch_b[step], ca2 = ch_a[step + 1], cb2 = ch_b[step + 1];
    const int P = g.pitch;
    const int lane = threadIdx.x & (TILE_X - 1), wv = threadIdx.x / TILE_X;
    const int nt = e.ntx * e.nty;
    double acc[1] = {0.0};
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int jt = e.j0 + (t % e.ntx) * TILE_X, kt = e.k0 + (t / e.ntx) * TH;
        const int j = jt + lane, ks = kt + wv * STRIP;
        const bool act = j <= e.j1 && ks <= e.k1;
        const int i0 = IX(g, j, ks);
        // step 1 on the ring first, so that its registers are free again
        // before the own strip's state (kept for step 2) goes live: strip
        // above (wave 0), strip below (wave 1), the tile's strips in the
        // columns either side (wave 2, lanes 0-7)
        {
            int hj = -1, hk = 0, take = -1, col = 0, row0 = 0;
            if (wv == 0) {
                hj = j; hk = kt - STRIP; take = STRIP - 1; col = 1 + lane; row0 = 0;
            } else if (wv == 1) {
                hj = j; hk = kt + TH; take = 0; col = 1 + lane; row0 = RH - 1;
            } else if (wv == 2 && lane < 2 * STRIPS_PER_TILE) {
                const int side = lane / STRIPS_PER_TILE, qs = lane % STRIPS_PER_TILE;
                hj = side ? jt + TILE_X : jt - 1; hk = kt + qs * STRIP; take = STRIP; col = side ? RW - 1 : 0;
                row0 = 1 + qs * STRIP;
            }
            if (hj >= 0) {
                const bool inside = hj >= e.j0 && hj <= e.j1 && hk >= e.k0 && hk <= e.k1;
                double sdn[STRIP];
                if (inside) {
                    const int h0 = IX(g, hj, hk);
                    double hx0[STRIP], hx1[STRIP], hy[STRIP + 1], hr[STRIP], res[STRIP], sc[STRIP], zf[STRIP];
                    strip_k(Kx, Ky, h0, P, hx0, hx1, hy);
                    col4(r_in, h0, P, hr);
                    strip_matvec(sd_in, hx0, hx1, hy, h0, P, res, sc);
                    ppcg_step<PREC>(g, hk, e.k1, hx0, hx1, hy, res, sc, ca1, cb1, hr, sdn, zf);
                }
                if (take < STRIP) {
                    // one row of a strip above / below the tile
                    const int k = hk + take;
                    if (hj <= e.j1 + 1) ring[row0][col] = (inside && k <= e.k1) ? sdn[take] : ring_outside(g, e, sd_out, hj, k);
                } else {
#pragma unroll
                    for (int q = 0; q < STRIP; ++q)
                        ring[row0 + q][col] = (inside && hk + q <= e.k1) ? sdn[q] : ring_outside(g, e, sd_out, hj, hk + q);
                }
            }
        }
        double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], rv[STRIP], uv[STRIP];
        // step 1 on the own strip
        if (act) {
            double res[STRIP], sc[STRIP], sdn[STRIP], zf[STRIP];
            strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
            col4_nt(u, i0, P, uv);
            col4_nt(r_in, i0, P, rv);
            strip_matvec(sd_in, kx0, kx1, ky, i0, P, res, sc);
            ppcg_step<PREC>(g, ks, e.k1, kx0, kx1, ky, res, sc, ca1, cb1, rv, sdn, zf);
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                uv[q] += sc[q];
                ring[1 + wv * STRIP + q][1 + lane] = (ks + q <= e.k1) ? sdn[q] : ring_outside(g, e, sd_out, j, ks + q);
            }
        } else {
#pragma unroll
            for (int q = 0; q < STRIP; ++q) ring[1 + wv * STRIP + q][1 + lane] = ring_outside(g, e, sd_out, j, ks + q);
        }
        __syncthreads();
        // step 2 on the own strip, sd from the ring
        if (act) {
            double c6[STRIP + 2], w4[STRIP], e4[STRIP], res[STRIP], sc[STRIP], sdn[STRIP], zo[STRIP] = {},
                zf[STRIP];
#pragma unroll
            for (int q = 0; q < STRIP + 2; ++q) c6[q] = ring[wv * STRIP + q][1 + lane];
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                w4[q] = ring[1 + wv * STRIP + q][lane];
                e4[q] = ring[1 + wv * STRIP + q][lane + 2];
                sc[q] = c6[q + 1];
            }
            if (LAST == 1 && PREC != PREC_NONE) col4_nt(z, i0, P, zo);
            strip_apply(c6, w4, e4, kx0, kx1, ky, res);
            ppcg_step<PREC>(g, ks, e.k1, kx0, kx1, ky, res, sc, ca2, cb2, rv, sdn, zf);
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                if (ks + q > e.k1) continue;
                const int i = i0 + q * P;
                store_nt(u + i, uv[q] + sc[q]);
                store_nt(r_out + i, rv[q]);
                store_nt(sd_out + i, sdn[q]);
                if (LAST == 2 && PREC != PREC_NONE) store_nt(z + i, zf[q]);
                if (LAST) acc[0] += ppcg_rz<PREC, LAST>(rv[q], zo[q], zf[q]);
            }
        }
        __syncthreads();
    }
    if (LAST) block_partial<1>(acc, ra.partials);
}

// One Chebyshev iteration on a strip once A u is known (wv = A u, uc = u):
// r = u0 - A u, z = M^-1 r, p' = ca p + cb z, u' = u + p'
// (tea_leaf_cheby_cl.cl:66-145). Shared by the one- and two-iteration kernels,
// with the FMA placed explicitly, so that both round identically.
template <int PREC>
__device__ __forceinline__ void cheby_step(const Geom &g, int ks, int k1, const double *kx0, const double *kx1,
                                           const double *ky, const double *b, const double *wv, const double *uc,
                                           const double *pv, double ca, double cb, double *rv, double *pn,
                                           double *un)
{
    double zv[STRIP];
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] = b[q] - wv[q];
    mask_rows(ks, k1, rv);
    if (PREC == PREC_BLOCK) strip_block(g, ks, 0, kx0, kx1, ky, nullptr, nullptr, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (PREC == PREC_BLOCK) pn[q] = fma(ca, pv[q], cb * zv[q]);
        else if (PREC == PREC_DIAG)
            pn[q] = fma(ca, pv[q], cb * div_nr(1.0, diag_of(kx0[q], kx1[q], ky[q], ky[q + 1])) * rv[q]);
        else pn[q] = fma(ca, pv[q], cb * rv[q]);
        un[q] = uc[q] + pn[q];
    }
}

// Chebyshev iteration: tea_leaf_cheby_solve_calc_p (tea_leaf_cheby_cl.cl:78-145)
// then tea_leaf_cheby_solve_calc_u (:66-76) in one pass: r = u0 - A u,
// z = M^-1 r, p = ca p + cb z, u' = u + p, and r.r when tea_solve.f90:223 asks
// for it; the finisher runs the iteration tail. u is double-buffered so that
// no neighbour is read after it is overwritten: iteration n reads buffer
// (n - n_fused0)&1 and writes the other. r lives in registers only (the
// reference stores it, but nothing in the iteration reads it back). A band
// march with the two-step kernel's geometry (its r.r summed in the same
// order); faces from the conductivity.
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_cheby(Geom g, Ext e, March m, const double *u0,
                                                               double *uA, double *uB, double *p,
                                                               const double *W, Faces fc,
                                                               const double *ch_a, const double *ch_b,
                                                               RedArgs ra)
{
    const SolveState *st = ra.st;
    if (st->halt) return;
    const int step = st->cheby_calc_steps - 1;
    const double ca = ch_a[step], cb = ch_b[step];
    const bool odd = ((st->n - st->n_fused0) & 1) != 0;
    const double *u = odd ? uB : uA;
    double *un_out = odd ? uA : uB;
    const bool want = cheby_norm_due(st);
    const int lane = threadIdx.x & 63;
    const int P = g.pitch, hd = g.hd;
    double acc = 0.0;
    TL_MARCH_LOOP(m, e)
    const int j = jb - 2 + lane;
    const bool out_lane = lane >= 2 && lane <= 61 && j <= e.j1;
    const bool inner_x = jb - 1 > hd && jb + MARCH2_COLS + 1 < hd + g.x_max;
    double u6[STRIP + 2], c6[STRIP + 2];
    const int ib = IX(g, j, k_begin - 1);
    u6[0] = u[ib];
    u6[1] = u[ib + P];
    c6[0] = W[ib];
    c6[1] = W[ib + P];
    double ky0 = ky_at(g, fc, j, k_begin, c6[0], c6[1]);
    for (int ks = k_begin; ks <= k_end; ks += STRIP) {
        const int i0 = IX(g, j, ks);
        double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], w4[STRIP], e4[STRIP], b[STRIP], pv[STRIP], wv[STRIP],
            uc[STRIP], rv[STRIP], pn[STRIP], uv[STRIP];
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            u6[q + 2] = u[i0 + (q + 1) * P];
            c6[q + 2] = W[i0 + (q + 1) * P];
        }
        col4_nt(u0, i0, P, b);
        col4_nt(p, i0, P, pv);
        const bool inner = inner_x && ks > hd && ks + STRIP < hd + g.y_max;
        ky[0] = ky0;
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            kx0[q] = march_kx(g, fc, inner, j, ks + q, dpp_shr1(c6[q + 1], 0.0), c6[q + 1]);
            kx1[q] = dpp_shl1(kx0[q], 0.0);
            ky[q + 1] = march_ky(g, fc, inner, j, ks + q + 1, c6[q + 1], c6[q + 2]);
            w4[q] = dpp_shr1(u6[q + 1], 0.0);
            e4[q] = dpp_shl1(u6[q + 1], 0.0);
            uc[q] = u6[q + 1];
        }
        strip_apply(u6, w4, e4, kx0, kx1, ky, wv);
        cheby_step<PREC>(g, ks, e.k1, kx0, kx1, ky, b, wv, uc, pv, ca, cb, rv, pn, uv);
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            if (ks + q > e.k1 || !out_lane) continue;
            const int i = i0 + q * P;
            store_nt(p + i, pn[q]);
            store_nt(un_out + i, uv[q]);
            if (want) acc += rv[q] * rv[q];
        }
        u6[0] = u6[STRIP];
        u6[1] = u6[STRIP + 1];
        c6[0] = c6[STRIP];
        c6[1] = c6[STRIP + 1];
        ky0 = ky[STRIP];
    }
    TL_MARCH_LOOP_END
    double a1[1] = {acc};
    block_partial<1>(a1, ra.partials);
}

// Two Chebyshev iterations n, n+1 per launch, n odd (tea_solve.f90:223 forms
// the residual norm only on iterations divisible by 10, so only the second one
// can need it), as a band march with a one-strip lag: iteration n on 4-row
// strip S, then iteration n+1 on strip
// S-1 from iteration n's u of the rows it spans plus one row either side
// (kept in registers with S-1's new p, u0 and faces).
// u and p are read from one buffer and written to the other. When the tail of
// iteration n will stop the solve (max_iters, or the stale-error exit test of
// tea_solve.f90:322-328, both known from the state before the launch) only
// iteration n is done. Faces are read from Kx, Ky (bitwise what the
// one-iteration kernel recomputes); cells outside the extent take what the
// one-iteration kernel would read there. The finisher runs both iterations'
// tails. Algorithmic bytes per two iterations: R u, u0, p, Kx, Ky  W u, p
// (56 B/cell) instead of 96.
template <int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fused_cheby2(Geom g, Ext e, March m, const double *u0,
                                                                const double *u_in, double *u_out, const double *p_in,
                                                                double *p_out, const double *Kx, const double *Ky,
                                                                const double *ch_a, const double *ch_b, RedArgs ra)
{
    const SolveState *st = ra.st;
    if (st->halt) return;
    const int n = st->n, step = st->cheby_calc_steps - 1;
    const bool two = n < st->max_iters && !(fabs(sqrt(st->error)) < st->eps * st->initial_residual);
    const bool want = two && n + 1 >= st->est_itc && ((n + 1) % 10) == 0;
    const double ca1 = ch_a[step], cb1 = ch_b[step];
    const double ca2 = two ? ch_a[step + 1] : 0.0, cb2 = two ? ch_b[step + 1] : 0.0;
    const int lane = threadIdx.x & 63;
    const int P = g.pitch;
    double acc = 0.0;
    TL_MARCH_LOOP(m, e)
    const int j = jb - 2 + lane;
    const bool col_in = j >= e.j0 && j <= e.j1;
    const bool out_lane = lane >= 2 && lane <= 61 && j <= e.j1;
    const int ks_last = k_begin + (k_end - k_begin) / STRIP * STRIP;
    // iteration n input window: u rows ks-1 .. ks+4 (the first two carried)
    double c6[STRIP + 2];
    int ks = two ? k_begin - STRIP : k_begin;
    {
        const int i = IX(g, j, ks - 1);
        c6[0] = u_in[i];
        c6[1] = u_in[i + P];
    }
    double ky0 = Ky[IX(g, j, ks)];
    // iteration n's results for strip S-1 (u row above it in u1p[0]), and
    // that strip's faces
    // (its bottom face is ky0, which the next strip's top face shares)
    double u1p[STRIP + 1] = {}, p1p[STRIP] = {}, bp[STRIP] = {}, kxp[STRIP] = {}, kyq[STRIP] = {};
    for (; ks <= ks_last + (two ? STRIP : 0); ks += STRIP) {
        const int i0 = IX(g, j, ks);
        // ---- iteration n on strip ks
        double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], b[STRIP], pv[STRIP], w4[STRIP], e4[STRIP], wv[STRIP],
            uc[STRIP], rv[STRIP], pn[STRIP], un[STRIP], u1[STRIP];
        ky[0] = ky0;
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            c6[q + 2] = u_in[i0 + (q + 1) * P];
            kx0[q] = Kx[i0 + q * P];
            ky[q + 1] = Ky[i0 + (q + 1) * P];
        }
        col4(u0, i0, P, b);
        col4(p_in, i0, P, pv);
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            kx1[q] = dpp_shl1(kx0[q], 0.0);
            w4[q] = dpp_shr1(c6[q + 1], 0.0);
            e4[q] = dpp_shl1(c6[q + 1], 0.0);
            uc[q] = c6[q + 1];
        }
        strip_apply(c6, w4, e4, kx0, kx1, ky, wv);
        cheby_step<PREC>(g, ks, e.k1, kx0, kx1, ky, b, wv, uc, pv, ca1, cb1, rv, pn, un);
        if (!two) {
            // iteration n only: its results are the launch's
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                if (ks + q > e.k1 || !out_lane) continue;
                store_nt(p_out + i0 + q * P, pn[q]);
                store_nt(u_out + i0 + q * P, un[q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                const int k = ks + q;
                u1[q] = (col_in && k >= e.k0 && k <= e.k1) ? un[q] : ring_outside(g, e, u_out, j, k);
            }
            // ---- iteration n+1 on strip ks-4
            if (ks - STRIP >= k_begin) {
                const int ks2 = ks - STRIP, i2 = IX(g, j, ks2);
                double w6[STRIP + 2], w4b[STRIP], e4b[STRIP], wv2[STRIP], uc2[STRIP], rv2[STRIP], pn2[STRIP],
                    un2[STRIP], kx0p[STRIP], kx1p[STRIP], kyp[STRIP + 1];
#pragma unroll
                for (int q = 0; q < STRIP; ++q) kx0p[q] = kxp[q];
#pragma unroll
                for (int q = 0; q < STRIP; ++q) kyp[q] = kyq[q];
                kyp[STRIP] = ky0;
#pragma unroll
                for (int q = 0; q < STRIP + 1; ++q) w6[q] = u1p[q];
                w6[STRIP + 1] = u1[0];
#pragma unroll
                for (int q = 0; q < STRIP; ++q) {
                    kx1p[q] = dpp_shl1(kx0p[q], 0.0);
                    w4b[q] = dpp_shr1(w6[q + 1], 0.0);
                    e4b[q] = dpp_shl1(w6[q + 1], 0.0);
                    uc2[q] = w6[q + 1];
                }
                strip_apply(w6, w4b, e4b, kx0p, kx1p, kyp, wv2);
                cheby_step<PREC>(g, ks2, e.k1, kx0p, kx1p, kyp, bp, wv2, uc2, p1p, ca2, cb2, rv2, pn2, un2);
#pragma unroll
                for (int q = 0; q < STRIP; ++q) {
                    if (ks2 + q > e.k1 || !out_lane) continue;
                    const int i = i2 + q * P;
                    store_nt(p_out + i, pn2[q]);
                    store_nt(u_out + i, un2[q]);
                    if (want) acc += rv2[q] * rv2[q];
                }
            }
            // ---- carry strip ks to the next trip
            u1p[0] = u1p[STRIP];
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                u1p[q + 1] = u1[q];
                p1p[q] = pn[q];
                bp[q] = b[q];
                kxp[q] = kx0[q];
                kyq[q] = ky[q];
            }
        }
        c6[0] = c6[STRIP];
        c6[1] = c6[STRIP + 1];
        ky0 = ky[STRIP];
    }
    TL_MARCH_LOOP_END
    double a1[1] = {acc};
    block_partial<1>(a1, ra.partials);
}

// ===========================================================================
// host launchers. They only enqueue: a failed launch or event record is
// reported by the caller's hipGetLastError check after the launcher returns.
// ===========================================================================
March make_march(const Ext &e, int cols)
{
    March m;
    const int nx = e.j1 - e.j0 + 1, ny = e.k1 - e.k0 + 1;
    m.cols = cols;
    m.nbands = nx > 0 ? (nx + cols - 1) / cols : 0;
    m.ngroups = (m.nbands + 3) / 4;
    m.H = MARCH_ROWS;
    m.nseg = ny > 0 ? (ny + m.H - 1) / m.H : 0;
    m.segs_per = (m.nseg + 7) / 8;
    return m;
}

void set_march_rows(March &m, const Ext &e, int rows)
{
    const int ny = e.k1 - e.k0 + 1;
    m.H = rows;
    m.nseg = ny > 0 ? (ny + rows - 1) / rows : 0;
    m.segs_per = (m.nseg + 7) / 8;
}

// blocks of a march launch: eight per slab item, capped (a multiple of 8)
// at the partials' capacity; blocks then take several items each
int march_grid(const March &m, int cap)
{
    const long items = 8L * m.segs_per * m.ngroups;
    const int c8 = cap / 8 * 8;
    return items < 8 ? 8 : (items < c8 ? (int)items : c8);
}

#define TL_LAUNCH(kern, e, cap, stream, ...)                                                  \
    hipLaunchKernelGGL(kern, dim3(grid_for(e, cap)), dim3(BLOCK_THREADS), 0, stream, __VA_ARGS__)

// One block per tile, for the PPCG kernels that write no per-block partials
// (the cap bounds the partials buffer and the finisher's loads, which these
// launches do not use). At 4096^2 this gave 2.6 % more PPCG iterations per
// second than blocks looping over two tiles each, with bitwise the same
// results.
#define TL_LAUNCH_PER_TILE(kern, e, stream, ...)                                                  \
    hipLaunchKernelGGL(kern, dim3(grid_for(e, INT_MAX)), dim3(BLOCK_THREADS), 0, stream, __VA_ARGS__)

// reduction kernel + the finisher that sums its per-block partials
#define TL_LAUNCH_RED(kern, e, cap, stream, nq, ra, ...) TL_LAUNCH_RED_EV(kern, e, cap, stream, nq, ra, nullptr, __VA_ARGS__)

// ... with an optional event recorded between the reduction kernel and its
// finisher (per-kernel timing of the iteration kernels)
#define TL_LAUNCH_RED_EV(kern, e, cap, stream, nq, ra, ev, ...)                                    \
    do {                                                                                           \
        const int grid_ = grid_for(e, cap);                                                        \
        hipLaunchKernelGGL(kern, dim3(grid_), dim3(BLOCK_THREADS), 0, stream, __VA_ARGS__);         \
        if (ev) (void)hipEventRecord(ev, stream);                                                  \
        launch_finish(stream, ra, grid_, nq);                                                      \
    } while (0)

void launch_init_chunk_1d(hipStream_t s, const Geom &g, double xmin, double ymin, double dx, double dy,
                          double *const *f1)
{
    const int n = (g.x_max > g.y_max ? g.x_max : g.y_max) + 2 * g.hd + 1;
    hipLaunchKernelGGL(k_init_chunk_1d, dim3((n + 255) / 256), dim3(256), 0, s, g, xmin, ymin, dx, dy,
                       f1[F_VERTEXX], f1[F_VERTEXDX], f1[F_VERTEXY], f1[F_VERTEXDY], f1[F_CELLX],
                       f1[F_CELLDX], f1[F_CELLY], f1[F_CELLDY]);
}

void launch_init_chunk_2d(hipStream_t s, const Geom &g, const Ext &e, int cap, double dx, double dy,
                          double *vol, double *xa, double *ya)
{
    TL_LAUNCH(k_init_chunk_2d, e, cap, s, g, e, dx, dy, vol, xa, ya);
}

void launch_generate(hipStream_t s, const Geom &g, const Ext &e, int cap, double *const *f1, double *rho,
                     double *e0, double *u, double *u0, int nstates, const double *const *sarr,
                     const int *sgeom, int g_rect, int g_circ, int g_point)
{
    TL_LAUNCH(k_generate, e, cap, s, g, e, f1[F_VERTEXX], f1[F_VERTEXY], f1[F_CELLX], f1[F_CELLY], rho,
              e0, u, u0, nstates, sarr[0], sarr[1], sarr[2], sarr[3], sarr[4], sarr[5], sarr[6], sgeom,
              g_rect, g_circ, g_point);
}

void launch_copy(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *src, double *dst)
{
    TL_LAUNCH(k_copy, e, cap, s, g, e, src, dst);
}

void launch_stream_mix(hipStream_t s, const Geom &g, const Ext &e, int cap, double *a, const double *b, double *c,
                       const double *d, double *out)
{
    TL_LAUNCH(k_stream_mix, e, cap, s, g, e, a, b, c, d, out, 1e-3);
}

void launch_field_summary(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *vol,
                          const double *rho, const double *en, const double *u, const RedArgs &ra)
{
    TL_LAUNCH_RED(k_field_summary, e, cap, s, 4, ra, g, e, vol, rho, en, u, ra);
}

void launch_halo(hipStream_t s, const Geom &g, double *a, int depth, int face)
{
    if (face >= 2) {
        const int n = g.x_max + 2 * depth;
        hipLaunchKernelGGL(k_halo_bt, dim3((n + 255) / 256), dim3(256), 0, s, g, a, depth, face);
    } else {
        const int n = g.y_max + 2 * depth;
        hipLaunchKernelGGL(k_halo_lr, dim3((n + 255) / 256), dim3(256), 0, s, g, a, depth, face);
    }
}

void launch_init_common(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho, double *Kx,
                        double *Ky, double rx, double ry, int coef)
{
    TL_LAUNCH(k_init_common, e, cap, s, g, e, rho, Kx, Ky, rx, ry, coef);
}

void launch_conductivity(hipStream_t s, const Geom &g, const double *rho, double *w, int coef, int *bad)
{
    hipLaunchKernelGGL(k_conductivity, dim3(1024), dim3(256), 0, s, g, rho, w, coef, bad);
}

void launch_zero_boundary(hipStream_t s, const Geom &g, const Ext &e, int cap, double *Kx, double *Ky,
                          const int *zb)
{
    TL_LAUNCH(k_zero_boundary, e, cap, s, g, e, Kx, Ky, zb[0], zb[1], zb[2], zb[3]);
}

void launch_init_u(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho, const double *en,
                   double *u, double *u0)
{
    TL_LAUNCH(k_init_u, e, cap, s, g, e, rho, en, u, u0);
}

void launch_init_jac_diag(hipStream_t s, const Geom &g, const Ext &e, int cap, double *Mi, const double *Kx,
                          const double *Ky)
{
    TL_LAUNCH(k_init_jac_diag, e, cap, s, g, e, Mi, Kx, Ky);
}

// num / den by the compiler's division and by div_nr (tl_check_division)
__global__ void __launch_bounds__(256) k_check_division(const double *num, const double *den, double *quot,
                                                        double *fast, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        quot[i] = num[i] / den[i];
        fast[i] = div_nr(num[i], den[i]);
    }
}

void launch_check_division(hipStream_t s, const double *num, const double *den, double *quot, double *fast, long n)
{
    hipLaunchKernelGGL(k_check_division, dim3(1024), dim3(256), 0, s, num, den, quot, fast, n);
}

void launch_diag_apply(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *r, const double *Mi,
                       double *z)
{
    TL_LAUNCH(k_diag_apply, e, cap, s, g, e, r, Mi, z);
}

void launch_calc_residual(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *u,
                          const double *u0, double *r, const double *Kx, const double *Ky)
{
    TL_LAUNCH(k_calc_residual, e, cap, s, g, e, u, u0, r, Kx, Ky);
}

void launch_dot(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *a, const double *b,
                const RedArgs &ra)
{
    TL_LAUNCH_RED(k_dot, e, cap, s, 1, ra, g, e, a, b, ra);
}

void launch_finalise(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho,
                     const double *u, double *en)
{
    TL_LAUNCH(k_finalise, e, cap, s, g, e, rho, u, en);
}

void launch_block_init(hipStream_t s, const Geom &g, const Ext &e, int cap, double *cp, double *bfp,
                       const double *Kx, const double *Ky)
{
    TL_LAUNCH(k_block_init, e, cap, s, g, e, cp, bfp, Kx, Ky);
}

void launch_block_solve(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *r, double *z,
                        const double *cp, const double *bfp, const double *Kx, const double *Ky)
{
    TL_LAUNCH(k_block_solve, e, cap, s, g, e, r, z, cp, bfp, Kx, Ky);
}

void launch_cg_init_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double *p,
                      const double *r, double *z, const double *Mi, const RedArgs &ra)
{
    if (prec == PREC_BLOCK) TL_LAUNCH_RED(k_cg_init_p<PREC_BLOCK>, e, cap, s, 1, ra, g, e, p, r, z, Mi, ra);
    else if (prec == PREC_DIAG) TL_LAUNCH_RED(k_cg_init_p<PREC_DIAG>, e, cap, s, 1, ra, g, e, p, r, z, Mi, ra);
    else TL_LAUNCH_RED(k_cg_init_p<PREC_NONE>, e, cap, s, 1, ra, g, e, p, r, z, Mi, ra);
}

void launch_cg_calc_w(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *p, double *w,
                      const double *Kx, const double *Ky, const double *W, const Faces &fc, const RedArgs &ra)
{
    TL_LAUNCH_RED(k_cg_calc_w, e, cap, s, 1, ra, g, e, p, w, Kx, Ky, W, fc, ra);
}

void launch_cg_calc_ur(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double alpha,
                       double *u, const double *p, double *r, const double *w, double *z,
                       const double *cp, const double *bfp, const double *Mi, const double *Kx,
                       const double *Ky, const double *W, const Faces &fc, const RedArgs &ra)
{
#define TL_CUR(P) \
    TL_LAUNCH_RED(k_cg_calc_ur<P>, e, cap, s, 1, ra, g, e, alpha, u, p, r, w, z, cp, bfp, Mi, Kx, Ky, W, fc, ra)
    if (prec == PREC_BLOCK) TL_CUR(PREC_BLOCK);
    else if (prec == PREC_DIAG) TL_CUR(PREC_DIAG);
    else TL_CUR(PREC_NONE);
#undef TL_CUR
}

void launch_cg_calc_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double beta,
                      double *p, const double *r, const double *z)
{
    if (prec == PREC_NONE) TL_LAUNCH(k_cg_calc_p<PREC_NONE>, e, cap, s, g, e, beta, p, r, z);
    else TL_LAUNCH(k_cg_calc_p<PREC_DIAG>, e, cap, s, g, e, beta, p, r, z);
}

void launch_cheby_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int init,
                    const double *u, const double *u0, double *p, double *r, double *w,
                    const double *cp, const double *bfp, const double *Mi, const double *Kx,
                    const double *Ky, const double *W, const Faces &fc, const double *ch_a,
                    const double *ch_b, double theta, int step, const SolveState *st)
{
#define TL_CHP(P, I)                                                                                      \
    TL_LAUNCH((k_cheby_p<P, I>), e, cap, s, g, e, u, u0, p, r, w, cp, bfp, Mi, Kx, Ky, W, fc, ch_a, ch_b, \
              theta, step, st)
    if (init) {
        if (prec == PREC_BLOCK) TL_CHP(PREC_BLOCK, 1);
        else if (prec == PREC_DIAG) TL_CHP(PREC_DIAG, 1);
        else TL_CHP(PREC_NONE, 1);
    } else {
        if (prec == PREC_BLOCK) TL_CHP(PREC_BLOCK, 0);
        else if (prec == PREC_DIAG) TL_CHP(PREC_DIAG, 0);
        else TL_CHP(PREC_NONE, 0);
    }
#undef TL_CHP
}

void launch_cheby_u(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, const double *p,
                    const double *r, const RedArgs &ra)
{
    if (ra.st) TL_LAUNCH_RED(k_cheby_u, e, cap, s, 1, ra, g, e, u, p, r, ra);
    else TL_LAUNCH(k_cheby_u, e, cap, s, g, e, u, p, r, ra);
}

void launch_ppcg_init_sd(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *r,
                         double *sd, const double *cp, const double *bfp, const double *Mi,
                         const double *Kx, const double *Ky, double theta)
{
    if (prec == PREC_BLOCK)
        TL_LAUNCH(k_ppcg_init_sd<PREC_BLOCK>, e, cap, s, g, e, r, sd, cp, bfp, Mi, Kx, Ky, theta);
    else if (prec == PREC_DIAG)
        TL_LAUNCH(k_ppcg_init_sd<PREC_DIAG>, e, cap, s, g, e, r, sd, cp, bfp, Mi, Kx, Ky, theta);
    else
        TL_LAUNCH(k_ppcg_init_sd<PREC_NONE>, e, cap, s, g, e, r, sd, cp, bfp, Mi, Kx, Ky, theta);
}

void launch_ppcg_update_r(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, double *r,
                          const double *Kx, const double *Ky, const double *W, const Faces &fc, const double *sd)
{
    TL_LAUNCH(k_ppcg_update_r, e, cap, s, g, e, u, r, Kx, Ky, W, fc, sd);
}

void launch_ppcg_calc_sd(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *r,
                         double *sd, const double *cp, const double *bfp, const double *Mi,
                         const double *Kx, const double *Ky, const double *W, const Faces &fc,
                         const double *ch_a, const double *ch_b, int step)
{
#define TL_CSD(P) TL_LAUNCH(k_ppcg_calc_sd<P>, e, cap, s, g, e, r, sd, cp, bfp, Mi, Kx, Ky, W, fc, ch_a, ch_b, step)
    if (prec == PREC_BLOCK) TL_CSD(PREC_BLOCK);
    else if (prec == PREC_DIAG) TL_CSD(PREC_DIAG);
    else TL_CSD(PREC_NONE);
#undef TL_CSD
}

void launch_jacobi(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *Kx, const double *Ky,
                   const double *W, const Faces &fc, const double *u0, double *u1, double *un, const RedArgs &ra)
{
    if (W) TL_LAUNCH_RED(k_jacobi<1>, e, cap, s, 1, ra, g, e, nullptr, nullptr, W, fc, u0, u1, un, ra);
    else TL_LAUNCH_RED(k_jacobi<0>, e, cap, s, 1, ra, g, e, Kx, Ky, nullptr, Faces{}, u0, u1, un, ra);
}

void launch_fused_cg_w(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *src,
                       double *pA, double *pB, const double *W, const Faces &fc, const RedArgs &ra,
                       hipEvent_t ev_main, bool stream_z)
{
    const March m = make_march(e, MARCH_COLS);
    const int grid = march_grid(m, cap);
    if (stream_z)
        hipLaunchKernelGGL(k_fused_cg_w<true>, dim3(grid), dim3(BLOCK_THREADS), 0, s, g, e, m, src, pA, pB, W, fc, ra);
    else
        hipLaunchKernelGGL(k_fused_cg_w<false>, dim3(grid), dim3(BLOCK_THREADS), 0, s, g, e, m, src, pA, pB, W, fc,
                           ra);
    if (ev_main) (void)hipEventRecord(ev_main, s);
    launch_finish(s, ra, grid, 1);
}

void launch_cg_u_fix(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, const double *p,
                     const SolveState *st)
{
    TL_LAUNCH(k_cg_u_fix, e, cap, s, g, e, u, p, st);
}

void launch_fused_jacobi(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *W, const Faces &fc,
                         const double *u0, double *uA, double *uB, const RedArgs &ra,
                         hipEvent_t ev_main)
{
    TL_LAUNCH_RED_EV(k_jacobi<1>, e, cap, s, 1, ra, ev_main, g, e, nullptr, nullptr, W, fc, u0, uA, uB, ra);
}

void launch_fused_cg_ur(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int ppcg, double *u,
                        const double *pA, const double *pB, const double *r, double *r_out, double *z,
                        double *sd, const double *W, const Faces &fc, const RedArgs &ra, hipEvent_t ev_main,
                        int parity)
{
    if (!ppcg && (parity == 1 || parity == 2)) {
#define TL_FUP(P, PAR) TL_LAUNCH_RED_EV((k_fused_cg_ur<P, 0, PAR>), e, cap, s, 1, ra, ev_main, g, e, u, pA, pB, r, \
                                        r_out, z, sd, W, fc, ra)
        if (parity == 1) {
            if (prec == PREC_BLOCK) TL_FUP(PREC_BLOCK, 1);
            else if (prec == PREC_DIAG) TL_FUP(PREC_DIAG, 1);
            else TL_FUP(PREC_NONE, 1);
        } else {
            if (prec == PREC_BLOCK) TL_FUP(PREC_BLOCK, 2);
            else if (prec == PREC_DIAG) TL_FUP(PREC_DIAG, 2);
            else TL_FUP(PREC_NONE, 2);
        }
#undef TL_FUP
        return;
    }
#define TL_FUR(P, M)                                                                                           \
    do {                                                                                                       \
        if (M) TL_LAUNCH_PER_TILE((k_fused_cg_ur<P, M>), e, s, g, e, u, pA, pB, r, r_out, z, sd, W, fc, ra);        \
        else TL_LAUNCH_RED_EV((k_fused_cg_ur<P, M>), e, cap, s, 1, ra, ev_main, g, e, u, pA, pB, r, r_out, z,   \
                              sd, W, fc, ra);                                                                  \
    } while (0)
    if (ppcg) {
        if (prec == PREC_BLOCK) TL_FUR(PREC_BLOCK, 1);
        else if (prec == PREC_DIAG) TL_FUR(PREC_DIAG, 1);
        else TL_FUR(PREC_NONE, 1);
    } else {
        if (prec == PREC_BLOCK) TL_FUR(PREC_BLOCK, 0);
        else if (prec == PREC_DIAG) TL_FUR(PREC_DIAG, 0);
        else TL_FUR(PREC_NONE, 0);
    }
#undef TL_FUR
}

void launch_fused_ppcg_inner(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int last,
                             double *u, double *r, const double *sd_in, double *sd_out,
                             const double *W, const Faces &fc, double *z, const double *ch_a,
                             const double *ch_b, int step, const RedArgs &ra, hipEvent_t ev_main)
{
#define TL_FPI(P, L)                                                                                         \
    do {                                                                                                     \
        if (L) TL_LAUNCH_RED_EV((k_fused_ppcg_inner<P, L>), e, cap, s, 1, ra, ev_main, g, e, u, r, sd_in, sd_out, W, fc, z, \
                             ch_a, ch_b, step, ra);                                                          \
        else TL_LAUNCH_PER_TILE((k_fused_ppcg_inner<P, L>), e, s, g, e, u, r, sd_in, sd_out, W, fc, z, ch_a, ch_b,  \
                       step, ra);                                                                            \
    } while (0)
    if (last == 2) {
        if (prec == PREC_BLOCK) TL_FPI(PREC_BLOCK, 2);
        else if (prec == PREC_DIAG) TL_FPI(PREC_DIAG, 2);
        else TL_FPI(PREC_NONE, 1);
    } else if (last) {
        if (prec == PREC_BLOCK) TL_FPI(PREC_BLOCK, 1);
        else if (prec == PREC_DIAG) TL_FPI(PREC_DIAG, 1);
        else TL_FPI(PREC_NONE, 1);
    } else {
        if (prec == PREC_BLOCK) TL_FPI(PREC_BLOCK, 0);
        else if (prec == PREC_DIAG) TL_FPI(PREC_DIAG, 0);
        else TL_FPI(PREC_NONE, 0);
    }
#undef TL_FPI
}

void launch_fused_ppcg_inner2(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int last, double *u,
                              const double *r_in, double *r_out, const double *sd_in, double *sd_out,
                              const double *Kx, const double *Ky, double *z, const double *ch_a,
                              const double *ch_b, int step, const RedArgs &ra, hipEvent_t ev_main)
{
#define TL_FPI2(P, L)                                                                                           \
    do {                                                                                                        \
        if (L) TL_LAUNCH_RED_EV((k_fused_ppcg_inner2<P, L>), e, cap, s, 1, ra, ev_main, g, e, u, r_in, r_out, sd_in, sd_out, \
                             Kx, Ky, z, ch_a, ch_b, step, ra);                                                  \
        else TL_LAUNCH_PER_TILE((k_fused_ppcg_inner2<P, L>), e, s, g, e, u, r_in, r_out, sd_in, sd_out, Kx, Ky, z,   \
                       ch_a, ch_b, step, ra);                                                                   \
    } while (0)
    if (last == 2) {
        if (prec == PREC_BLOCK) TL_FPI2(PREC_BLOCK, 2);
        else if (prec == PREC_DIAG) TL_FPI2(PREC_DIAG, 2);
        else TL_FPI2(PREC_NONE, 1);
    } else if (last) {
        if (prec == PREC_BLOCK) TL_FPI2(PREC_BLOCK, 1);
        else if (prec == PREC_DIAG) TL_FPI2(PREC_DIAG, 1);
        else TL_FPI2(PREC_NONE, 1);
    } else {
        if (prec == PREC_BLOCK) TL_FPI2(PREC_BLOCK, 0);
        else if (prec == PREC_DIAG) TL_FPI2(PREC_DIAG, 0);
        else TL_FPI2(PREC_NONE, 0);
    }
#undef TL_FPI2
}

void launch_fused_cheby2(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *u0,
                         const double *u_in, double *u_out, const double *p_in, double *p_out, const double *Kx,
                         const double *Ky, const double *ch_a, const double *ch_b, const RedArgs &ra,
                         hipEvent_t ev_main)
{
    March m = make_march(e, MARCH2_COLS);
    set_march_rows(m, e, MARCH2_ROWS);
    const int grid = march_grid(m, cap);
#define TL_FC2(P)                                                                                                   \
    hipLaunchKernelGGL(k_fused_cheby2<P>, dim3(grid), dim3(BLOCK_THREADS), 0, s, g, e, m, u0, u_in, u_out, p_in, p_out, \
                       Kx, Ky, ch_a, ch_b, ra)
    if (prec == PREC_BLOCK) TL_FC2(PREC_BLOCK);
    else if (prec == PREC_DIAG) TL_FC2(PREC_DIAG);
    else TL_FC2(PREC_NONE);
#undef TL_FC2
    if (ev_main) (void)hipEventRecord(ev_main, s);
    launch_finish(s, ra, grid, 1);
}

void launch_fused_cheby(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *u0,
                        double *uA, double *uB, double *p, const double *W, const Faces &fc,
                        const double *ch_a, const double *ch_b, const RedArgs &ra, hipEvent_t ev_main)
{
    March m = make_march(e, MARCH2_COLS);
    set_march_rows(m, e, MARCH2_ROWS);
    const int grid = march_grid(m, cap);
#define TL_FC1(P)                                                                                                \
    hipLaunchKernelGGL(k_fused_cheby<P>, dim3(grid), dim3(BLOCK_THREADS), 0, s, g, e, m, u0, uA, uB, p, W, fc, ch_a, \
                       ch_b, ra)
    if (prec == PREC_BLOCK) TL_FC1(PREC_BLOCK);
    else if (prec == PREC_DIAG) TL_FC1(PREC_DIAG);
    else TL_FC1(PREC_NONE);
#undef TL_FC1
    if (ev_main) (void)hipEventRecord(ev_main, s);
    launch_finish(s, ra, grid, 1);
}

}  // namespace tl
