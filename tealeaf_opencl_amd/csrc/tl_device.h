// Device-side building blocks shared by the FP64 solver kernels
// (tl_kernels.hip) and the mixed-precision refinement kernels (tl_refine.hip).
//
// Work decomposition (every kernel): a 256-thread block is four wavefronts;
// a tile is 64 columns x 16 rows; wavefront w of the block owns the 4-row
// strip w of the tile and lane l owns column l. A thread therefore handles one
// column of one 4-row strip, which is exactly the y-strip of the strip-4
// block-Jacobi preconditioner (kernel_files/tea_block_jacobi.cl), so the
// tridiagonal solve of a strip stays in one thread's registers. Each row of a
// strip is one fully coalesced load per wavefront (512 bytes in FP64, 256 in
// FP32). Blocks stride over tiles; the grid is capped so reductions produce
// few partials.
//
// Field loads are templated on the storage type (double or float); all
// arithmetic on loaded values is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "tl_common.h"

namespace tl {

// ---------------------------------------------------------------------------
// tiling
// ---------------------------------------------------------------------------
#define TL_TILE_LOOP(e)                                                          \
    const int tx_ = threadIdx.x & (TILE_X - 1);                                  \
    const int ty_ = threadIdx.x / TILE_X;                                        \
    const int nt_ = (e).ntx * (e).nty;                                           \
    for (int t_ = blockIdx.x; t_ < nt_; t_ += gridDim.x) {                       \
        const int j = (e).j0 + (t_ % (e).ntx) * TILE_X + tx_;                    \
        const int ks = (e).k0 + ((t_ / (e).ntx) * STRIPS_PER_TILE + ty_) * STRIP; \
        if (j > (e).j1 || ks > (e).k1) continue;
#define TL_TILE_LOOP_END }

__device__ __forceinline__ int IX(const Geom &g, int j, int k) { return k * g.pitch + j; }

// 5-point operator of tea_leaf_common_cl.cl:28-32 (same association). The
// products are fused explicitly: left to the compiler, which of two products
// summed together becomes the FMA can depend on how the surrounding kernel is
// laid out in basic blocks, and kernels that apply the operator to the same
// values would then differ in the last bit.
__device__ __forceinline__ double opA(double kx0, double kx1, double ky0, double ky1, double vc,
                                      double vw, double ve, double vs, double vn)
{
    const double ty = fma(ky1, vn, ky0 * vs);
    const double tx = fma(kx1, ve, kx0 * vw);
    return fma(1.0 + (ky1 + ky0) + (kx1 + kx0), vc, -ty) - tx;
}

__device__ __forceinline__ double diag_of(double kx0, double kx1, double ky0, double ky1)
{
    return (1.0 + (ky1 + ky0) + (kx1 + kx0));
}

// BLOCK_TOP of macros_cl.cl:6-10 (BLOCK_TOP_CHECK only when y_max % 4 != 0)
__device__ __forceinline__ int block_top(const Geom &g, int row)
{
    if (g.y_max % STRIP) {
        const int t = g.y_max + 2 - row;
        return t < STRIP ? t : STRIP;
    }
    return STRIP;
}

// 1/x from the FP32 hardware reciprocal (about 1 ulp of FP32), for the
// preconditioner of the FP32-inner solve, whose vectors are rounded to FP32
// anyway. Used for pivots x >= 1 (diagonally dominant blocks), well inside
// the FP32 range.
__device__ __forceinline__ double rcp_fp32(double x)
{
    return (double)__builtin_amdgcn_rcpf((float)x);
}

// a / b for the operands of the iteration kernels: divisors that are positive
// normal numbers of moderate exponent (face denominators 2 w_a w_b of
// positive conductivities, Thomas pivots and diagonals >= 1) and finite
// dividends. The hardware reciprocal, two Newton steps and one residual
// correction (Markstein's method): the operations of the compiler's IEEE FP64
// division less its range scaling (v_div_scale / v_div_fmas) and special-value
// fix-up (v_div_fixup), which leave such operands unchanged, so the quotient
// is the correctly rounded one bit for bit, in 8 instructions instead of 11.
// The recomputed faces and block factors built on it are compared bitwise
// with the stored ones, which tea_leaf_init_common / tea_leaf_block_init form
// with the plain division.
__device__ __forceinline__ double div_nr(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    const double r = fma(-b, q, a);
    return fma(r, y, q);
}

// divisions of the block-Jacobi factors and solve: the plain IEEE division,
// div_nr (bitwise the same for these operands), or the FP32 reciprocal of
// the FP32-inner solve (rcp_fp32)
enum { DIV_IEEE = 0, DIV_NR = 1, DIV_RCP32 = 2 };

template <int DIV>
__device__ __forceinline__ double quot(double a, double b)
{
    return DIV == DIV_NR ? div_nr(a, b) : a / b;
}

// block_solve_func (tea_block_jacobi.cl:6-42) on one strip held in registers.
// b[k] = COEF_B, a[k] = COEF_A of strip row k; cp/bfp as produced by
// tea_leaf_block_init; rows >= upper are left at zero. DIV_RCP32: bfp[0]
// holds 1/b[0] (strip_factors<DIV_RCP32>), used in place of the division.
template <int DIV = DIV_IEEE>
__device__ __forceinline__ void strip_solve(int upper, const double *r, const double *b,
                                            const double *a, const double *cp,
                                            const double *bfp, double *z)
{
    double dp[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k) z[k] = 0.0;
    if (upper <= 0) return;
    dp[0] = DIV == DIV_RCP32 ? r[0] * bfp[0] : quot<DIV>(r[0], b[0]);
#pragma unroll
    for (int k = 1; k < STRIP; ++k)
        if (k < upper) dp[k] = (r[k] - a[k] * dp[k - 1]) * bfp[k];
    // back substitution from row upper-1
#pragma unroll
    for (int k = STRIP - 1; k >= 0; --k) {
        if (k > upper - 1) continue;
        if (k == upper - 1) z[k] = dp[k];
        else z[k] = dp[k] - cp[k] * z[k + 1];
    }
}

// Thomas factors of tea_leaf_block_init (tea_block_jacobi.cl:78-105).
// DIV_RCP32: the pivots' reciprocals from rcp_fp32 (no FP64 division) and
// bfp[0] = 1/b[0]. The factorisation stays exact for the matrix it defines:
// its off-diagonals are (1/bfp[k]) * cp[k] = c[k] and a[k] as before (still
// symmetric), its diagonal differs from b[k] by the reciprocal's error times
// the pivot, about 1e-7 relative.
template <int DIV = DIV_IEEE>
__device__ __forceinline__ void strip_factors(int upper, const double *b, const double *a,
                                              const double *c, double *cp, double *bfp)
{
    if (DIV == DIV_RCP32) {
        bfp[0] = rcp_fp32(b[0]);
        cp[0] = c[0] * bfp[0];
    } else {
        cp[0] = quot<DIV>(c[0], b[0]);
        bfp[0] = 0.0;
    }
#pragma unroll
    for (int k = 1; k < STRIP; ++k) {
        if (k < upper) {
            bfp[k] = DIV == DIV_RCP32 ? rcp_fp32(b[k] - a[k] * cp[k - 1]) : quot<DIV>(1.0, b[k] - a[k] * cp[k - 1]);
            cp[k] = c[k] * bfp[k];
        } else {
            bfp[k] = 0.0;
            cp[k] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// reductions: wavefront shuffle sum, then the four wavefronts of the block in
// fixed order, one partial per block (plain stores); a one-block finisher
// kernel launched after the reduction kernel sums the partials in block order.
// Deterministic for a given grid.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block reduction of NR sums into partials[q * gridDim.x + blockIdx.x]
// (quantity-major, so that the finisher of a one-quantity reduction reads one
// contiguous array). Must be reached by every thread of the block.
template <int NR>
__device__ __forceinline__ void block_partial(const double (&v)[NR], double *partials)
{
    __shared__ double lds[STRIPS_PER_TILE][MAX_RED];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) lds[w][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < NR) {
        const int q = threadIdx.x;
        partials[q * gridDim.x + blockIdx.x] = ((lds[0][q] + lds[1][q]) + lds[2][q]) + lds[3][q];
    }
}

constexpr int FINISH_THREADS = 1024;

// Sum of the first NQ partials of nblocks blocks, in a fixed order, by one
// block of FINISH_THREADS threads; every thread gets the totals (tot[q] for
// q < NQ, 0 beyond). Thread t adds blocks t, t + FINISH_THREADS, ... in that
// order whatever the unroll, so the totals do not depend on it. The loads of
// U blocks' partials are issued before any is added: a grid of up to
// U * FINISH_THREADS blocks costs one memory round trip (8192 blocks for one
// or two quantities, 4096 for more).
template <int NQ>
__device__ __forceinline__ void sum_partials(const double *partials, int nblocks, double *tot)
{
    static_assert(NQ >= 1 && NQ <= MAX_RED, "quantities");
    __shared__ double lds[FINISH_THREADS / 64][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    constexpr int U = NQ <= 2 ? 8 : 4;
    for (int b0 = threadIdx.x; b0 < nblocks; b0 += U * FINISH_THREADS) {
        double v[U][NQ];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * FINISH_THREADS;
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[u][q] = b < nblocks ? partials[q * nblocks + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] += v[u][q];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double s = wave_sum(acc[q]);
        if (lane == 0) lds[w][q] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAX_RED; ++q) {
        double t = 0.0;
        if (q < NQ)
            for (int i = 0; i < FINISH_THREADS / 64; ++i) t += lds[i][q];
        tot[q] = t;
    }
}

// ---------------------------------------------------------------------------
// strip helpers. Every load of a strip is issued before any of it is used;
// rows past the extent are read too (the allocation margins hold them) and
// only stores and sums are predicated on the row being inside the extent.
// ---------------------------------------------------------------------------

// rows ks .. ks+3 of column j (v[q] = row ks+q)
template <typename T>
__device__ __forceinline__ void col4(const T *f, int i0, int P, double *v)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q) v[q] = (double)f[i0 + q * P];
}

// Streaming accesses with the nontemporal hint, for fields that an iteration
// kernel reads or writes exactly once per cell (no stencil neighbours): they
// are not kept in the caches, which stay with the stencil fields. Measured on
// the CG kernels at 4096^2: 339 -> 303 us per iteration at the same grid.
template <typename T>
__device__ __forceinline__ void col4_nt(const T *f, int i0, int P, double *v)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q) v[q] = (double)__builtin_nontemporal_load(f + i0 + q * P);
}

template <typename T>
__device__ __forceinline__ void store_nt(T *p, double v)
{
    __builtin_nontemporal_store((T)v, p);
}

// rows ks-1 .. ks+4 of column j (v[q] = row ks+q-1)
template <typename T>
__device__ __forceinline__ void col6(const T *f, int i0, int P, double *v)
{
#pragma unroll
    for (int q = 0; q < STRIP + 2; ++q) v[q] = (double)f[i0 + (q - 1) * P];
}

// face coefficients of a strip: Kx at the cell and its east neighbour, Ky at
// rows ks .. ks+4
template <typename T>
__device__ __forceinline__ void strip_k(const T *Kx, const T *Ky, int i0, int P, double *kx0, double *kx1,
                                        double *ky)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        kx0[q] = (double)Kx[i0 + q * P];
        kx1[q] = (double)Kx[i0 + q * P + 1];
    }
#pragma unroll
    for (int q = 0; q < STRIP + 1; ++q) ky[q] = (double)Ky[i0 + q * P];
}

// Kx at face (j,k) (between cells j-1 and j of row k) and Ky at face (j,k)
// (between rows k-1 and k of column j) as tea_leaf_init_common computes them
// (same expression and association, so the values are bitwise those it
// stores) over its launch extent [1, n+2hd-1] (ocl_kernels.cpp:112), zero
// outside it, then zeroed like tea_leaf_zero_boundary.
__device__ __forceinline__ double face_coef(double wa, double wb, double scale)
{
    double v = (wa + wb) / (2.0 * wa * wb);
    v *= scale;
    return v;
}

// the same for two positive conductivities (faces strictly inside the
// chunk), with div_nr: bitwise face_coef
__device__ __forceinline__ double face_coef_nr(double wa, double wb, double scale)
{
    double v = div_nr(wa + wb, 2.0 * wa * wb);
    v *= scale;
    return v;
}

// The same face from h = 1/(2w) stored per cell: scale * (h_a + h_b), equal
// to face_coef in exact arithmetic, with no division (the CG kernels of the
// FP32-inner solve, whose operator is an FP32-rounded one anyway).
__device__ __forceinline__ double face_coef_h(double ha, double hb, double scale)
{
    return scale * (ha + hb);
}

__device__ __forceinline__ bool outside_init(const Geom &g, int j, int k)
{
    return j < 1 || j > g.x_max + 2 * g.hd - 1 || k < 1 || k > g.y_max + 2 * g.hd - 1;
}

template <bool HALF_INV = false>
__device__ __forceinline__ double kx_at(const Geom &g, const Faces &f, int j, int k, double wl, double wc)
{
    const int hd = g.hd;
    const bool zero = outside_init(g, j, k) || (f.zl && j <= hd) || (f.zr && j > g.x_max - 1 + hd) ||
                      (f.zb && k < hd) || (f.zt && k > g.y_max - 1 + hd);
    const double v = HALF_INV ? face_coef_h(wl, wc, f.rx) : face_coef(wl, wc, f.rx);
    return zero ? 0.0 : v;
}

template <bool HALF_INV = false>
__device__ __forceinline__ double ky_at(const Geom &g, const Faces &f, int j, int k, double wd, double wc)
{
    const int hd = g.hd;
    const bool zero = outside_init(g, j, k) || (f.zl && j < hd) || (f.zr && j > g.x_max - 1 + hd) ||
                      (f.zb && k <= hd) || (f.zt && k > g.y_max - 1 + hd);
    const double v = HALF_INV ? face_coef_h(wd, wc, f.ry) : face_coef(wd, wc, f.ry);
    return zero ? 0.0 : v;
}

// face coefficients of the strip (same outputs as strip_k) recomputed from a
// per-cell field stored as T: the conductivity w (bitwise the init_common
// faces), or with HALF_INV its 1/(2w)
template <bool HALF_INV = false, typename T>
__device__ __forceinline__ void strip_k_cond(const Geom &g, const Faces &f, const T *w, int i0, int P, int j, int ks,
                                             double *kx0, double *kx1, double *ky)
{
    double wc[STRIP + 2], ww[STRIP], we[STRIP];
    col6(w, i0, P, wc);
    col4(w, i0 - 1, P, ww);
    col4(w, i0 + 1, P, we);
    const int hd = g.hd;
    // every face of the strip lies strictly inside the chunk: no zeroing and
    // no extent test can apply (the common case; only strips on a wall take
    // the checked path)
    const bool inner = j > hd && j + 1 < hd + g.x_max && ks > hd && ks + STRIP < hd + g.y_max;
    if (inner) {
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            kx0[q] = HALF_INV ? face_coef_h(ww[q], wc[q + 1], f.rx) : face_coef_nr(ww[q], wc[q + 1], f.rx);
            kx1[q] = HALF_INV ? face_coef_h(wc[q + 1], we[q], f.rx) : face_coef_nr(wc[q + 1], we[q], f.rx);
        }
#pragma unroll
        for (int q = 0; q < STRIP + 1; ++q)
            ky[q] = HALF_INV ? face_coef_h(wc[q], wc[q + 1], f.ry) : face_coef_nr(wc[q], wc[q + 1], f.ry);
        return;
    }
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        kx0[q] = kx_at<HALF_INV>(g, f, j, ks + q, ww[q], wc[q + 1]);
        kx1[q] = kx_at<HALF_INV>(g, f, j + 1, ks + q, wc[q + 1], we[q]);
    }
#pragma unroll
    for (int q = 0; q < STRIP + 1; ++q) ky[q] = ky_at<HALF_INV>(g, f, j, ks + q, wc[q], wc[q + 1]);
}

// w = A v on the strip: c6 = centre column rows ks-1..ks+4, w4/e4 = west/east
__device__ __forceinline__ void strip_apply(const double *c6, const double *w4, const double *e4,
                                            const double *kx0, const double *kx1, const double *ky, double *out)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q)
        out[q] = opA(kx0[q], kx1[q], ky[q], ky[q + 1], c6[q + 1], w4[q], e4[q], c6[q], c6[q + 2]);
}

// A v for a field read straight from memory
template <typename T>
__device__ __forceinline__ void strip_matvec(const T *v, const double *kx0, const double *kx1,
                                             const double *ky, int i0, int P, double *out, double *centre)
{
    double c6[STRIP + 2], w4[STRIP], e4[STRIP];
    col6(v, i0, P, c6);
    col4(v, i0 - 1, P, w4);
    col4(v, i0 + 1, P, e4);
    strip_apply(c6, w4, e4, kx0, kx1, ky, out);
    if (centre)
#pragma unroll
        for (int q = 0; q < STRIP; ++q) centre[q] = c6[q + 1];
}

// tridiagonal coefficients of the strip from its face coefficients
__device__ __forceinline__ void strip_abc(const double *kx0, const double *kx1, const double *ky, double *a,
                                          double *b, double *c)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        a[q] = -ky[q];
        c[q] = -ky[q + 1];
        b[q] = (1.0 + (ky[q + 1] + ky[q]) + (kx1[q] + kx0[q]));
    }
}

// z = M_blk^-1 r on the strip, with factors either read (cp/bfp non-null)
// or recomputed from the coefficients
__device__ __forceinline__ void strip_block(const Geom &g, int ks, int i0, const double *kx0,
                                            const double *kx1, const double *ky, const double *cp,
                                            const double *bfp, const double *rv, double *zv)
{
    double a[STRIP], b[STRIP], c[STRIP], cpv[STRIP], bfv[STRIP];
    strip_abc(kx0, kx1, ky, a, b, c);
    const int upper = block_top(g, ks);
    if (cp) {
        col4(cp, i0, g.pitch, cpv);
        col4(bfp, i0, g.pitch, bfv);
    } else {
        strip_factors<DIV_NR>(upper, b, a, c, cpv, bfv);
    }
    strip_solve<DIV_NR>(upper, rv, b, a, cpv, bfv, zv);
}

// the strip's right-hand side with rows past k1 zeroed (the reference's
// work-group copy only holds in-bounds rows)
__device__ __forceinline__ void mask_rows(int ks, int k1, double *v)
{
#pragma unroll
    for (int q = 0; q < STRIP; ++q) v[q] = (ks + q <= k1) ? v[q] : 0.0;
}

// z = M^-1 r for one strip with the preconditioner recomputed from Kx, Ky
// (rv already masked past the extent). FAST: reciprocals from rcp_fp32 (the
// FP32-inner solve's preconditioner).
template <int PREC, bool FAST = false>
__device__ __forceinline__ void strip_precondition(const Geom &g, int ks, int i0, const double *kx0,
                                                   const double *kx1, const double *ky, const double *rv,
                                                   double *zv)
{
    if (PREC == PREC_BLOCK) {
        if (FAST) {
            double a[STRIP], b[STRIP], c[STRIP], cpv[STRIP], bfv[STRIP];
            strip_abc(kx0, kx1, ky, a, b, c);
            const int upper = block_top(g, ks);
            strip_factors<DIV_RCP32>(upper, b, a, c, cpv, bfv);
            strip_solve<DIV_RCP32>(upper, rv, b, a, cpv, bfv, zv);
        } else {
            strip_block(g, ks, i0, kx0, kx1, ky, nullptr, nullptr, rv, zv);
        }
    } else if (PREC == PREC_DIAG) {
#pragma unroll
        for (int q = 0; q < STRIP; ++q) {
            const double dg = diag_of(kx0[q], kx1[q], ky[q], ky[q + 1]);
            zv[q] = rv[q] * (FAST ? rcp_fp32(dg) : div_nr(1.0, dg));
        }
    } else {
#pragma unroll
        for (int q = 0; q < STRIP; ++q) zv[q] = rv[q];
    }
}

// ---------------------------------------------------------------------------
// Band march (k_fused_cg_w, the fused Chebyshev kernels and the refined
// solve's k_rf_w). A
// wavefront owns a band of 64 lanes and walks the rows of one segment of M.H
// rows downwards with the rows above and below the current one in
// registers: one load per row and field. The west and east neighbours'
// values come from lanes -1 / +1 by DPP (wave_shr:1 / wave_shl:1), and each
// face coefficient is formed once, by the lane east of the face, and passed to
// the lane west of it. Bands overlap: MARCH_COLS output columns (lanes
// 1..62 = columns jb .. jb+61; lanes 0 and 63 only supply neighbours), or
// MARCH2_COLS for the two-step kernels (lanes 2..61).
// Segments are cut into eight slabs of consecutive rows; blocks
// b with b % 8 == x (workgroups are dispatched round-robin to the eight
// XCDs) walk slab x, four bands side by side per block, so the halo rows and
// columns that neighbouring items share stay in one L2.
// ---------------------------------------------------------------------------
constexpr int MARCH_COLS = 62;
// segments of 16 rows for the one-step kernels (measured at 4096^2: 16 rows
// 110 us per cg_w launch, 32 rows 117 us, 64 rows 125 us; a taller segment
// reuses more rows but leaves fewer wavefronts to hide latency)
constexpr int MARCH_ROWS = 16;
// two-step kernels: 60 output columns per band (two exchange lanes either
// side) and taller segments (each segment also runs step 1 on the strip
// above and below it)
constexpr int MARCH2_COLS = 60;
constexpr int MARCH2_ROWS = 32;

struct March {
    int cols;      // output columns per band (MARCH_COLS, or MARCH2_COLS for two steps)
    int nbands;    // bands across the interior
    int ngroups;   // blocks of four bands across the interior
    int H;         // rows per segment
    int nseg;      // segments
    int segs_per;  // segments per slab
};

// a load with the nontemporal hint when NT is set
template <bool NT>
__device__ __forceinline__ double ld_nt(const double *p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

// lane l gets lane l-1's v (lane 0 keeps `old`)
__device__ __forceinline__ double dpp_shr1(double v, double old)
{
    const int2 a = __builtin_bit_cast(int2, v), o = __builtin_bit_cast(int2, old);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(o.x, a.x, 0x138, 0xF, 0xF, false);
    r.y = __builtin_amdgcn_update_dpp(o.y, a.y, 0x138, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}

// lane l gets lane l+1's v (lane 63 keeps `old`)
__device__ __forceinline__ double dpp_shl1(double v, double old)
{
    const int2 a = __builtin_bit_cast(int2, v), o = __builtin_bit_cast(int2, old);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(o.x, a.x, 0x130, 0xF, 0xF, false);
    r.y = __builtin_amdgcn_update_dpp(o.y, a.y, 0x130, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}

// The items of this block: band (of this wavefront) and segment. Every
// wavefront of the block runs the same number of loop trips, so a block
// reduction after the loop is reached by all of them.
#define TL_MARCH_LOOP(m, e)                                                                    \
    {                                                                                          \
        const int xcd_ = blockIdx.x & 7, per_ = gridDim.x >> 3;                                \
        const int items_ = (m).segs_per * (m).ngroups;                                         \
        for (int t_ = blockIdx.x >> 3; t_ < items_; t_ += per_) {                              \
            const int seg_ = xcd_ * (m).segs_per + t_ / (m).ngroups;                           \
            const int band_ = (t_ % (m).ngroups) * 4 + (int)(threadIdx.x >> 6);                \
            if (seg_ >= (m).nseg || band_ >= (m).nbands) continue;                             \
            const int jb = (e).j0 + band_ * (m).cols;                                          \
            const int k_begin = (e).k0 + seg_ * (m).H;                                         \
            const int k_end = min((e).k1, k_begin + (m).H - 1);
#define TL_MARCH_LOOP_END \
    }                     \
    }

// Kx of the face west of column j in row k from the conductivities either
// side (bitwise tea_leaf_init_common's value, zeroed as tea_leaf_zero_boundary
// zeroes it); `inner`: the face is known to lie strictly inside the chunk
__device__ __forceinline__ double march_kx(const Geom &g, const Faces &f, bool inner, int j, int k, double wl,
                                           double wc)
{
    return inner ? face_coef_nr(wl, wc, f.rx) : kx_at(g, f, j, k, wl, wc);
}
__device__ __forceinline__ double march_ky(const Geom &g, const Faces &f, bool inner, int j, int k, double wd,
                                           double wc)
{
    return inner ? face_coef_nr(wd, wc, f.ry) : ky_at(g, f, j, k, wd, wc);
}

}  // namespace tl
