// Refined solve of A u = u0 for one implicit timestep (after common_init has
// formed Kx, Ky, the conductivity and u = u0): a preconditioned CG whose
// vectors (r, z, p, the solution increment d) are stored in FP32 or FP64.
// FP64 inner fields: the operator's faces are recomputed from the
// conductivity (bitwise those of tea_leaf_init_common) and the
// preconditioner's pivots divided exactly. FP32 inner fields: the operator's
// faces come from FP32 fields (1/(2w) in a CG solve, Kx and Ky rounded to
// FP32 in a PPCG solve) and the preconditioner's pivot reciprocals from the
// FP32 hardware reciprocal. Every other arithmetic operation, every dot
// product and every scalar is FP64. The inner solve is
// wrapped in FP64 reliable updates: whenever the recursive residual has
// dropped by `delta` since the last update, or has reached the tolerance, the
// increment is added into the FP64 solution (u += d, d = 0) and the residual
// is replaced by the true FP64 residual u0 - A u (residual replacement,
// van der Vorst & Ye, SIAM J. Sci. Comput. 22(5), 2000). The search direction
// is kept across updates while the recursive residual still tracks the true
// one, so the Krylov recurrence is not restarted; it is restarted from the
// true residual when the two have drifted apart by more than a factor 2.
// Convergence is declared on the true FP64 residual only:
// ||u0 - A u||_2 <= tol * ||u0 - A u_init||_2.
//
// The CG recurrence is the one of tea_leaf_cg_cl.cl:45-160 with the same
// operator (tea_leaf_common_cl.cl:17-36) and the same preconditioners
// (tea_leaf_common_cl.cl:127-142, tea_block_jacobi.cl:1-105); the reference
// has no mixed-precision path, so the refinement loop itself has no
// counterpart there.
//
// Kernels per inner iteration (same fusion as the FP64 driver): k_rf_w
// (p = z + beta p_old, w = A p, p.w) and k_rf_ur (d += alpha p,
// r -= alpha A p, z = M^-1 r, r.z and r.r), each followed by a one-block
// finisher that runs the scalar algebra on the device-resident RefineState.
// Batches of iterations are captured in a hipGraph; the host only acts
// between batches (reliable updates and the final test).
#include <hip/hip_runtime.h>

#include <limits.h>
#include <math.h>
#include <string.h>

#include <functional>
#include <vector>

#include "../../include/tealeaf_mi355x.h"
#include "tl_chunk.h"
#include "tl_device.h"
#include "tl_eigen.h"
#include "tl_kernels.h"

namespace tl {

// kernels and helpers of the refined solve (a named namespace: kernels with
// internal linkage are not visible to the profiling tools)
namespace rf {

enum { RW_W = 0, RW_R, RW_Z, RW_P, RW_P2, RW_D, RW_SD, RW_SD2, RW_R2, RW_KY, RW_N };

struct RfArgs {
    double *partials;
    RefineState *st;
    int mode;
    double *hist_a;       // CG alphas / betas by iteration (0-based), or null
    double *hist_b;
    int hist_len;
};

// replacements in a row without progress after which the solve stops
constexpr int STALL_LIMIT = 3;

__device__ __forceinline__ void rf_scalars(RefineState *s, const RfArgs &ra, const double *tot)
{
    const double rz = tot[0], rr = tot[1];
    switch (ra.mode) {
    case RF_UPDATE: {
        if (s->updates == 0) {
            s->rr0 = rr;
            s->beta = 0.0;
            s->stalls = 0;
        } else {
            // keep the search direction while the recursive residual tracked
            // the true one (within a factor 2 in norm); otherwise restart the
            // recurrence from the true residual
            s->beta = (rr > 4.0 * s->rr) ? 0.0 : rz / s->rz_prev;
            // the true residual has reached the rounding floor of the FP64
            // residual when replacement after replacement no longer halves it
            s->stalls = (rr > 0.25 * s->rr_prev) ? s->stalls + 1 : 0;
        }
        s->rr_prev = rr;
        s->d_lag = 0;  // d is cleared after the update
        s->rro = rz;
        s->rr_true = rr;
        s->rr = rr;
        s->rmax = sqrt(rr);
        s->updates += 1;
        s->want_update = 0;
        if (sqrt(rr) <= s->tol * sqrt(s->rr0)) {
            s->converged = 1;
            s->halt = 1;
        } else if (s->n >= s->max_iters || !isfinite(rr) || !isfinite(rz) || s->stalls >= STALL_LIMIT) {
            s->halt = 1;
        } else {
            s->halt = 0;
        }
        break;
    }
    case RF_PW:
        s->pw = tot[0];
        s->alpha = s->rro / s->pw;
        if (ra.hist_a && s->n < ra.hist_len) ra.hist_a[s->n] = s->alpha;
        break;
    case RF_RRN_LAG:
        // k_rf_ur adds alpha p into d every second iteration (both steps at
        // once); the PPCG outer step updates d every time (RF_RRN)
        s->alpha_prev = s->alpha;
        s->d_lag ^= 1;
        [[fallthrough]];
    case RF_RRN: {
        s->rz_prev = s->rro;
        s->beta = rz / s->rro;
        if (ra.hist_b && s->n < ra.hist_len) ra.hist_b[s->n] = s->beta;
        s->rro = rz;
        s->rr = rr;
        s->n += 1;
        const double rn = sqrt(rr);
        if (rn > s->rmax) s->rmax = rn;
        if (rn <= s->tol * sqrt(s->rr0) || rn < s->delta * s->rmax || s->n >= s->max_iters || !isfinite(rz)) {
            s->want_update = 1;
            s->halt = 1;
        } else if (s->stop_n > 0 && s->n >= s->stop_n) {
            s->halt = 1;
        }
        break;
    }
    default:
        break;
    }
}

// The state is copied into registers as the partials are loaded and stored
// back whole (as k_finish does); a halted inner iteration changes nothing.
__global__ void __launch_bounds__(FINISH_THREADS) k_rf_finish(RfArgs ra, int nblocks)
{
    RefineState loc;
    if (threadIdx.x == 0) loc = *ra.st;
    double tot[MAX_RED];
    sum_partials<2>(ra.partials, nblocks, tot);
    if (threadIdx.x != 0 || (ra.mode != RF_UPDATE && loc.halt)) return;
    rf_scalars(&loc, ra, tot);
    *ra.st = loc;
}

template <typename T>
__device__ __forceinline__ double rnd(double v)
{
    return (double)(T)v;
}

// Face coefficients of the inner operator, in the representation the solve
// uses in every one of its kernels (so that its operator is one fixed
// matrix): FP64 inner fields recompute them from the conductivity (bitwise
// the faces tea_leaf_init_common stores); an FP32-inner CG solve forms them
// from h = 1/(2w) rounded to FP32 (one field and no division: its kernels are
// bound by memory traffic); an FP32-inner PPCG solve reads the chunk's Kx, Ky
// rounded to FP32 (two fields, but no face formula and no wall test: its
// smoothing kernels are bound by arithmetic).
enum { RF_FACES_COND = 0, RF_FACES_HALF = 1, RF_FACES_K32 = 2 };

template <typename T, int FACES>
__device__ __forceinline__ void rf_strip_faces(const Geom &g, const Faces &fc, const T *W, const T *Wy, int i0,
                                               int P, int j, int ks, double *kx0, double *kx1, double *ky)
{
    if (FACES == RF_FACES_K32) strip_k(W, Wy, i0, P, kx0, kx1, ky);
    else strip_k_cond<FACES == RF_FACES_HALF>(g, fc, W, i0, P, j, ks, kx0, kx1, ky);
}

// the representation of a CG solve's and of a PPCG solve's faces
template <typename T>
constexpr int cg_faces()
{
    return sizeof(T) == 4 ? RF_FACES_HALF : RF_FACES_COND;
}
template <typename T>
constexpr int ppcg_faces()
{
    return sizeof(T) == 4 ? RF_FACES_K32 : RF_FACES_COND;
}

// h = 1/(2w) over the whole padded index space, rounded to FP32
__global__ void __launch_bounds__(256) k_rf_narrow_half(Geom g8, Geom g4, const double *w8, float *h4)
{
    const int w = g8.x_max + 2 * g8.hd, h = g8.y_max + 2 * g8.hd;
    const long n = (long)w * h;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t % w), k = (int)(t / w);
        h4[IX(g4, j, k)] = (float)(0.5 / w8[IX(g8, j, k)]);
    }
}

// The chunk's Kx, Ky rounded to FP32 over the whole padded index space
__global__ void __launch_bounds__(256) k_rf_narrow_faces(Geom g8, Geom g4, const double *kx8, const double *ky8,
                                                         float *kx4, float *ky4)
{
    const int w = g8.x_max + 2 * g8.hd, h = g8.y_max + 2 * g8.hd;
    const long n = (long)w * h;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t % w), k = (int)(t / w);
        kx4[IX(g4, j, k)] = (float)kx8[IX(g8, j, k)];
        ky4[IX(g4, j, k)] = (float)ky8[IX(g8, j, k)];
    }
}

// Preconditioner reciprocals of the inner operator: FP64 divisions for FP64
// inner fields, the FP32 hardware reciprocal for FP32 inner fields (whose
// M^-1 r is rounded to FP32 anyway). Every kernel of one precision uses the
// same one, so the preconditioner is one fixed symmetric operator.
template <typename T>
constexpr bool inner_fast_rcp()
{
    return sizeof(T) == 4;
}

// Reliable update: u_out = u_in + d (FP64), r = u0 - A u_out with the FP64
// face coefficients (bitwise those of tea_leaf_init_common), then r rounded
// to the inner storage and z = M^-1 r with the inner operator's
// preconditioner (faces formed like the inner kernels form them). Sums:
// r.z of the stored values, and the FP64 r.r. d is read at stencil
// neighbours, so it is cleared after the kernel, not in it.
template <typename T, int PREC, int FACES>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_update(Geom g8, Geom g4, Ext e, const double *u_in,
                                                             double *u_out, const T *d, const double *u0,
                                                             const double *cond8, const T *cond_in,
                                                             const T *cond_in_y, Faces fc, T *r, T *z, RfArgs ra)
{
    double acc[2] = {0.0, 0.0};
    TL_TILE_LOOP(e)
    const int i8 = IX(g8, j, ks), P8 = g8.pitch;
    const int i4 = IX(g4, j, ks), P4 = g4.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1];
    double uc[STRIP + 2], uw[STRIP], ue[STRIP], dc[STRIP + 2], dw[STRIP], de[STRIP], b[STRIP];
    strip_k_cond(g8, fc, cond8, i8, P8, j, ks, kx0, kx1, ky);
    col6(u_in, i8, P8, uc);
    col4(u_in, i8 - 1, P8, uw);
    col4(u_in, i8 + 1, P8, ue);
    col6(d, i4, P4, dc);
    col4(d, i4 - 1, P4, dw);
    col4(d, i4 + 1, P4, de);
    col4(u0, i8, P8, b);
    double c6[STRIP + 2], w4[STRIP], e4[STRIP], av[STRIP], rv[STRIP], rt[STRIP], zv[STRIP];
#pragma unroll
    for (int q = 0; q < STRIP + 2; ++q) c6[q] = uc[q] + dc[q];
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        w4[q] = uw[q] + dw[q];
        e4[q] = ue[q] + de[q];
    }
    strip_apply(c6, w4, e4, kx0, kx1, ky, av);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] = b[q] - av[q];
    mask_rows(ks, e.k1, rv);
    double kx0t[STRIP], kx1t[STRIP], kyt[STRIP + 1];
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rt[q] = rnd<T>(rv[q]);
    rf_strip_faces<T, FACES>(g4, fc, cond_in, cond_in_y, i4, P4, j, ks, kx0t, kx1t, kyt);
    strip_precondition<PREC, inner_fast_rcp<T>()>(g4, ks, i4, kx0t, kx1t, kyt, rt, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const T zt = (T)zv[q];
        u_out[i8 + q * P8] = c6[q + 1];
        r[i4 + q * P4] = (T)rt[q];
        z[i4 + q * P4] = zt;
        acc[0] += rt[q] * (double)zt;
        acc[1] += rv[q] * rv[q];
    }
    TL_TILE_LOOP_END
    block_partial<2>(acc, ra.partials);
}

// Faces of the band-march kernels in the solve's representation (see
// rf_strip_faces): from the conductivity or from h = 1/(2w), the face west of
// column j in row k (wl, wc the per-cell values either side), or the face
// between rows k-1 and k (wd, wc). K32 faces are read, not formed.
template <int FACES>
__device__ __forceinline__ double rf_march_kx(const Geom &g, const Faces &f, bool inner, int j, int k, double wl,
                                              double wc)
{
    if (FACES == RF_FACES_HALF) return inner ? face_coef_h(wl, wc, f.rx) : kx_at<true>(g, f, j, k, wl, wc);
    return march_kx(g, f, inner, j, k, wl, wc);
}
template <int FACES>
__device__ __forceinline__ double rf_march_ky(const Geom &g, const Faces &f, bool inner, int j, int k, double wd,
                                              double wc)
{
    if (FACES == RF_FACES_HALF) return inner ? face_coef_h(wd, wc, f.ry) : ky_at<true>(g, f, j, k, wd, wc);
    return march_ky(g, f, inner, j, k, wd, wc);
}

// z in k_rf_w with NTZ: nontemporal loads for FP64 fields (as k_fused_cg_w:
// the caches then keep more of p_old and the faces; 0.104 -> 0.098 ms per CG
// iteration's launch at 4096^2), plain loads for FP32 fields (the hint
// measured 2 % slower there). The PPCG outer step launches it without the
// hint (with it: 115 -> 120 us after the inner steps), and so do the CG
// iterations on fields larger than stream_z_pays allows.
template <bool NTZ, typename T>
__device__ __forceinline__ T ld_z(const T *p)
{
    if (NTZ && sizeof(T) == 8) return __builtin_nontemporal_load(p);
    return *p;
}

// p = z + beta p_old (rounded to the storage), w = A p, p.w. p is
// double-buffered by the parity of the iteration counter. A band march
// (TL_MARCH_LOOP) as k_fused_cg_w: one load per row of z, p_old and the face
// field(s), neighbours by lane exchange, each face formed once.
template <typename T, int FACES, bool NTZ>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_w(Geom g, Ext e, March m, const T *src, T *pA, T *pB,
                                                        const T *W, const T *Wy, Faces fc, RfArgs ra)
{
    const RefineState *st = ra.st;
    if (st->halt) return;
    const double beta = st->beta;
    const bool odd = (st->n & 1) != 0;
    const T *p_old = odd ? pA : pB;
    T *p_new = odd ? pB : pA;
    constexpr bool K32 = FACES == RF_FACES_K32;
    const int lane = threadIdx.x & 63;
    const int P = g.pitch, hd = g.hd;
    // beta == 0 (first iteration) ignores p_old entirely
    auto pn = [&](double zv, double pv) { return rnd<T>(beta != 0.0 ? zv + beta * pv : zv); };
    double acc = 0.0;
    TL_MARCH_LOOP(m, e)
    const int j = jb - 1 + lane;
    const bool out_lane = lane >= 1 && lane <= MARCH_COLS && j <= e.j1;
    const bool inner_x = jb - 1 > hd && jb + MARCH_COLS < hd + g.x_max;
    int i = IX(g, j, k_begin - 1);
    double pa = pn((double)ld_z<NTZ>(src + i), (double)p_old[i]);
    const double ca = (double)W[i];
    i += P;
    double pb = pn((double)ld_z<NTZ>(src + i), (double)p_old[i]);
    double cb = (double)W[i];  // K32: Kx of row k; else the per-cell face field of row k
    double kyb = K32 ? (double)Wy[i] : rf_march_ky<FACES>(g, fc, false, j, k_begin, ca, cb);
    constexpr int PF = 2;
    double zr[PF], pr[PF], cr[PF], yr[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int ii = i + (q + 1) * P;
        zr[q] = (double)ld_z<NTZ>(src + ii);
        pr[q] = (double)p_old[ii];
        cr[q] = (double)W[ii];
        yr[q] = K32 ? (double)Wy[ii] : 0.0;
    }
    for (int k = k_begin; k <= k_end; k += PF) {
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int kk = k + q;
            const double pc = pn(zr[q], pr[q]);  // row kk+1
            const double cc = cr[q], yc = yr[q];
            {
                const int ii = IX(g, j, kk + 1 + PF);
                zr[q] = (double)ld_z<NTZ>(src + ii);
                pr[q] = (double)p_old[ii];
                cr[q] = (double)W[ii];
                yr[q] = K32 ? (double)Wy[ii] : 0.0;
            }
            if (kk > k_end) continue;
            const bool inner = inner_x && kk > hd && kk + 1 < hd + g.y_max;
            const double kyn = K32 ? yc : rf_march_ky<FACES>(g, fc, inner, j, kk + 1, cb, cc);
            const double kx0 = K32 ? cb : rf_march_kx<FACES>(g, fc, inner, j, kk, dpp_shr1(cb, 0.0), cb);
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

// d += alpha p, r -= alpha A p (A p recomputed from the stored p),
// z = M^-1 r with the preconditioner recomputed from the stored
// coefficients; r.z and r.r of the stored values. The update of d, which
// nothing in the iteration reads, is made every second iteration: with
// d_lag clear the kernel leaves d alone, with d_lag set it adds
// alpha_prev p_prev (still in the other p buffer) and then alpha p, each as
// one FMA rounded to the storage, so d is what per-iteration updates give.
// A reliable update or the switch to PPCG first completes d (k_rf_d_fix).
template <typename T, int PREC, int FACES>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_ur(Geom g, Ext e, T *d, const T *pA, const T *pB, T *r,
                                                         T *z, const T *W, const T *Wy, Faces fc, RfArgs ra)
{
    const RefineState *st = ra.st;
    if (st->halt) return;
    const bool odd = (st->n & 1) != 0;
    const T *p = odd ? pB : pA;
    const T *p_prev = odd ? pA : pB;
    const double alpha = st->alpha, alpha_prev = st->alpha_prev;
    const bool upd = st->d_lag != 0;
    double acc[2] = {0.0, 0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], wv[STRIP], pc[STRIP], dv[STRIP], rv[STRIP], zv[STRIP];
    rf_strip_faces<T, FACES>(g, fc, W, Wy, i0, P, j, ks, kx0, kx1, ky);
    if (upd) col4_nt(d, i0, P, dv);
    col4_nt(r, i0, P, rv);
    strip_matvec(p, kx0, kx1, ky, i0, P, wv, pc);
    if (upd) {
        double pp[STRIP];
        col4_nt(p_prev, i0, P, pp);
#pragma unroll
        for (int q = 0; q < STRIP; ++q) dv[q] = fma(alpha, pc[q], rnd<T>(fma(alpha_prev, pp[q], dv[q])));
    }
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] = rnd<T>(rv[q] - alpha * wv[q]);
    mask_rows(ks, e.k1, rv);
    strip_precondition<PREC, inner_fast_rcp<T>()>(g, ks, i0, kx0, kx1, ky, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        const T zt = (T)zv[q];
        if (upd) store_nt(d + i, dv[q]);
        store_nt(r + i, rv[q]);
        if (PREC != PREC_NONE) store_nt(z + i, (double)zt);
        acc[0] += rv[q] * (double)zt;
        acc[1] += rv[q] * rv[q];
    }
    TL_TILE_LOOP_END
    block_partial<2>(acc, ra.partials);
}

// Completes d after CG iterations of k_rf_ur that left it lagging:
// d += alpha_prev p of the last iteration (written into buffer (n-1) & 1).
// Nothing to do when d_lag is clear; the update that follows clears it.
template <typename T>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_d_fix(Geom g, Ext e, T *d, const T *pA, const T *pB,
                                                            const RefineState *st)
{
    if (!st->d_lag) return;
    const T *p = ((st->n - 1) & 1) ? pB : pA;
    const double alpha = st->alpha_prev;
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double dv[STRIP], pc[STRIP];
    col4_nt(d, i0, P, dv);
    col4_nt(p, i0, P, pc);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        store_nt(d + i0 + q * P, fma(alpha, pc[q], dv[q]));
    }
    TL_TILE_LOOP_END
}

// PPCG outer step, second half (the refined form of k_fused_cg_ur in PPCG
// mode, tea_leaf_cg_cl.cl:73-137 + tea_leaf_ppcg_cl.cl:4-60): d += alpha p,
// r -= alpha A p, z = M^-1 r, sd = z / theta; no reduction (the outer step's
// r.z is formed after the inner steps with this z, as the reference does).
template <typename T, int PREC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_ur_ppcg(Geom g, Ext e, T *d, const T *pA, const T *pB,
                                                              const T *r, T *r_out, T *z, T *sd, const T *W,
                                                              const T *Wy, Faces fc, RfArgs ra)
{
    const RefineState *st = ra.st;
    if (st->halt) return;
    const T *p = (st->n & 1) ? pB : pA;
    const double alpha = st->alpha, theta = st->theta;
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], wv[STRIP], pc[STRIP], dv[STRIP], rv[STRIP], zv[STRIP];
    rf_strip_faces<T, ppcg_faces<T>()>(g, fc, W, Wy, i0, P, j, ks, kx0, kx1, ky);
    col4_nt(d, i0, P, dv);
    col4_nt(r, i0, P, rv);
    strip_matvec(p, kx0, kx1, ky, i0, P, wv, pc);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        dv[q] += alpha * pc[q];
        rv[q] = rnd<T>(rv[q] - alpha * wv[q]);
    }
    mask_rows(ks, e.k1, rv);
    strip_precondition<PREC, inner_fast_rcp<T>()>(g, ks, i0, kx0, kx1, ky, rv, zv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        const double zt = rnd<T>(zv[q]);
        store_nt(d + i, dv[q]);
        store_nt(r_out + i, rv[q]);
        store_nt(z + i, zt);
        store_nt(sd + i, zt / theta);
    }
    TL_TILE_LOOP_END
}

// One PPCG smoothing step on a strip once A sd is known (res = A sd, sc = sd):
// r -= A sd rounded to the storage, sd' = a_i sd + b_i M^-1 r rounded to the
// storage (tea_leaf_ppcg_cl.cl:62-155). Shared by the one-step and two-step
// kernels, with the FMA placed explicitly, so that both round identically.
template <typename T, int PREC>
__device__ __forceinline__ void rf_ppcg_step(const Geom &g, int ks, int k1, int i0, const double *kx0,
                                             const double *kx1, const double *ky, const double *res,
                                             const double *sc, double ca, double cb, double *rv, double *sdn)
{
    double mv[STRIP];
#pragma unroll
    for (int q = 0; q < STRIP; ++q) rv[q] = rnd<T>(rv[q] - res[q]);
    mask_rows(ks, k1, rv);
    strip_precondition<PREC, inner_fast_rcp<T>()>(g, ks, i0, kx0, kx1, ky, rv, mv);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) sdn[q] = rnd<T>(fma(ca, sc[q], cb * mv[q]));
}

// PPCG inner step (tea_leaf_ppcg_cl.cl:62-155, fused update_r + calc_sd):
// d += sd, r -= A sd, sd' = a_i sd + b_i M^-1 r; sd double-buffered. LAST
// also sums r.z with the z of the outer step and r.r.
template <typename T, int PREC, int LAST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_inner(Geom g, Ext e, T *d, T *r, const T *sd_in, T *sd_out,
                                                            const T *W, const T *Wy, Faces fc, const T *z,
                                                            const double *ch_a, const double *ch_b, int step,
                                                            RfArgs ra)
{
    const RefineState *st = ra.st;
    if (st->halt) return;
    const double ca = ch_a[step], cb = ch_b[step];
    double acc[2] = {0.0, 0.0};
    TL_TILE_LOOP(e)
    const int i0 = IX(g, j, ks), P = g.pitch;
    double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], res[STRIP], sc[STRIP], dv[STRIP], rv[STRIP], sdn[STRIP],
        zo[STRIP];
    rf_strip_faces<T, ppcg_faces<T>()>(g, fc, W, Wy, i0, P, j, ks, kx0, kx1, ky);
    col4_nt(d, i0, P, dv);
    col4_nt(r, i0, P, rv);
    if (LAST) col4_nt(z, i0, P, zo);
    strip_matvec(sd_in, kx0, kx1, ky, i0, P, res, sc);
    rf_ppcg_step<T, PREC>(g, ks, e.k1, i0, kx0, kx1, ky, res, sc, ca, cb, rv, sdn);
#pragma unroll
    for (int q = 0; q < STRIP; ++q) {
        if (ks + q > e.k1) continue;
        const int i = i0 + q * P;
        store_nt(d + i, dv[q] + sc[q]);
        store_nt(r + i, rv[q]);
        store_nt(sd_out + i, sdn[q]);
        if (LAST) {
            acc[0] += rv[q] * (PREC != PREC_NONE ? zo[q] : rv[q]);
            acc[1] += rv[q] * rv[q];
        }
    }
    TL_TILE_LOOP_END
    if (LAST) block_partial<2>(acc, ra.partials);
}

// face coefficients of a strip for the two-step kernel: read from the
// chunk's Kx, Ky for FP64 inner fields (bitwise what the one-step kernel
// recomputes from the conductivity, without its divisions), from their FP32
// copies for FP32 inner fields (what the one-step kernel reads)
template <typename T>
__device__ __forceinline__ void rf_faces(const T *W, const T *Wy, const double *Kx, const double *Ky, int i0, int P,
                                         double *kx0, double *kx1, double *ky)
{
    if (sizeof(T) == 8) strip_k(Kx, Ky, i0, P, kx0, kx1, ky);
    else strip_k(W, Wy, i0, P, kx0, kx1, ky);
}

// sd at a ring cell the block does not compute: the sd buffer's halo cell
// (what the one-step kernel reads there), 0 beyond the halo
template <typename T>
__device__ __forceinline__ double rf_ring_outside(const Geom &g, const Ext &e, const T *sd, int j, int k)
{
    const bool in = j >= e.j0 - 1 && j <= e.j1 + 1 && k >= e.k0 - 1 && k <= e.k1 + 1;
    return in ? (double)sd[IX(g, j, k)] : 0.0;
}

// Two PPCG inner steps per launch, as k_fused_ppcg_inner2 does for the FP64
// driver: step 1 on the 64x16 tile and on the ring step 2 reads (the 4-row
// strips above and below, the tile's strips in the columns either side), the
// ring's sd through LDS, step 2 on the tile. Values are rounded to the
// storage exactly where the one-step kernel stores and reloads them, so two
// one-step launches and one two-step launch give bitwise the same d, r, sd.
// r and sd are read from one buffer and written to another (neighbouring
// blocks read this tile's r and sd in their step 1).
template <typename T, int PREC, int LAST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_rf_inner2(Geom g, Ext e, T *d, const T *r_in, T *r_out,
                                                             const T *sd_in, T *sd_out, const T *W, const T *Wy,
                                                             const double *Kx, const double *Ky, const T *z,
                                                             const double *ch_a, const double *ch_b, int step,
                                                             RfArgs ra)
{
    const RefineState *st = ra.st;
    if (st->halt) return;
    constexpr int TH = STRIPS_PER_TILE * STRIP;
    constexpr int RW = TILE_X + 2, RH = TH + 2;
    __shared__ double ring[RH][RW];
    const double ca1 = ch_a[step], cb1 = ch_b[step], ca2 = ch_a[step + 1], cb2 = ch_b[step + 1];
    const int P = g.pitch;
    const int lane = threadIdx.x & (TILE_X - 1), wv = threadIdx.x / TILE_X;
    const int nt = e.ntx * e.nty;
    double acc[2] = {0.0, 0.0};
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int jt = e.j0 + (t % e.ntx) * TILE_X, kt = e.k0 + (t / e.ntx) * TH;
        const int j = jt + lane, ks = kt + wv * STRIP;
        const bool act = j <= e.j1 && ks <= e.k1;
        const int i0 = IX(g, j, ks);
        // step 1 on the ring: strip above (wave 0), strip below (wave 1), the
        // tile's strips in the columns either side (wave 2, lanes 0-7)
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
                    double hx0[STRIP], hx1[STRIP], hy[STRIP + 1], hr[STRIP], res[STRIP], sc[STRIP];
                    rf_faces<T>(W, Wy, Kx, Ky, h0, P, hx0, hx1, hy);
                    col4(r_in, h0, P, hr);
                    strip_matvec(sd_in, hx0, hx1, hy, h0, P, res, sc);
                    rf_ppcg_step<T, PREC>(g, hk, e.k1, h0, hx0, hx1, hy, res, sc, ca1, cb1, hr, sdn);
                }
                if (take < STRIP) {
                    const int k = hk + take;
                    if (hj <= e.j1 + 1)
                        ring[row0][col] = (inside && k <= e.k1) ? sdn[take] : rf_ring_outside(g, e, sd_out, hj, k);
                } else {
#pragma unroll
                    for (int q = 0; q < STRIP; ++q)
                        ring[row0 + q][col] =
                            (inside && hk + q <= e.k1) ? sdn[q] : rf_ring_outside(g, e, sd_out, hj, hk + q);
                }
            }
        }
        // step 1 on the own strip; its faces, r and d are kept for step 2
        double kx0[STRIP], kx1[STRIP], ky[STRIP + 1], rv[STRIP], dv[STRIP];
        if (act) {
            double res[STRIP], sc[STRIP], sdn[STRIP];
            rf_faces<T>(W, Wy, Kx, Ky, i0, P, kx0, kx1, ky);
            col4_nt(d, i0, P, dv);
            col4_nt(r_in, i0, P, rv);
            strip_matvec(sd_in, kx0, kx1, ky, i0, P, res, sc);
            rf_ppcg_step<T, PREC>(g, ks, e.k1, i0, kx0, kx1, ky, res, sc, ca1, cb1, rv, sdn);
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                dv[q] = rnd<T>(dv[q] + sc[q]);
                ring[1 + wv * STRIP + q][1 + lane] =
                    (ks + q <= e.k1) ? sdn[q] : rf_ring_outside(g, e, sd_out, j, ks + q);
            }
        } else {
#pragma unroll
            for (int q = 0; q < STRIP; ++q)
                ring[1 + wv * STRIP + q][1 + lane] = rf_ring_outside(g, e, sd_out, j, ks + q);
        }
        __syncthreads();
        // step 2 on the own strip, sd from the ring
        if (act) {
            double c6[STRIP + 2], w4[STRIP], e4[STRIP], res[STRIP], sc[STRIP], sdn[STRIP], zo[STRIP];
#pragma unroll
            for (int q = 0; q < STRIP + 2; ++q) c6[q] = ring[wv * STRIP + q][1 + lane];
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                w4[q] = ring[1 + wv * STRIP + q][lane];
                e4[q] = ring[1 + wv * STRIP + q][lane + 2];
                sc[q] = c6[q + 1];
            }
            if (LAST) col4_nt(z, i0, P, zo);
            strip_apply(c6, w4, e4, kx0, kx1, ky, res);
            rf_ppcg_step<T, PREC>(g, ks, e.k1, i0, kx0, kx1, ky, res, sc, ca2, cb2, rv, sdn);
#pragma unroll
            for (int q = 0; q < STRIP; ++q) {
                if (ks + q > e.k1) continue;
                const int i = i0 + q * P;
                store_nt(d + i, dv[q] + sc[q]);
                store_nt(r_out + i, rv[q]);
                store_nt(sd_out + i, sdn[q]);
                if (LAST) {
                    acc[0] += rv[q] * (PREC != PREC_NONE ? zo[q] : rv[q]);
                    acc[1] += rv[q] * rv[q];
                }
            }
        }
        __syncthreads();
    }
    if (LAST) block_partial<2>(acc, ra.partials);
}

// history of the first CG iterations kept for the eigenvalue estimate, and
// the largest number of PPCG inner steps
constexpr int RF_HIST = 1024;
constexpr int RF_MAX_INNER = 256;

template <typename T, int FACES>
void launch_update_f(tl_chunk *c, tl_chunk::RefineWork &w, const double *u_in, double *u_out)
{
    const Ext e = ext_interior(c);
    const int grid = grid_for(e, c->cap);
    const RfArgs ra{c->partials, w.st, RF_UPDATE, nullptr, nullptr, 0};
    T *const *f = reinterpret_cast<T *const *>(w.f);
    hipLaunchKernelGGL(k_rf_d_fix<T>, dim3(grid), dim3(BLOCK_THREADS), 0, c->s, w.g, e, f[RW_D], f[RW_P], f[RW_P2],
                       w.st);
#define TL_RFU(P)                                                                                           \
    hipLaunchKernelGGL((k_rf_update<T, P, FACES>), dim3(grid), dim3(BLOCK_THREADS), 0, c->s, c->g, w.g, e, u_in, \
                       u_out, f[RW_D], c->f2[F_U0], c->f2[F_COND], f[RW_W], f[RW_KY], c->faces, f[RW_R], f[RW_Z], \
                       ra)
    if (c->prec == PREC_BLOCK) TL_RFU(PREC_BLOCK);
    else if (c->prec == PREC_DIAG) TL_RFU(PREC_DIAG);
    else TL_RFU(PREC_NONE);
#undef TL_RFU
    hipLaunchKernelGGL(k_rf_finish, dim3(1), dim3(FINISH_THREADS), 0, c->s, ra, grid);
    // d = 0 (never inside a captured graph: updates are launched eagerly; a
    // failure is returned by the callers' hipGetLastError check)
    (void)hipMemsetAsync(w.base[RW_D], 0, (size_t)w.bytes * w.elems, c->s);
}

// with the faces of the solve in progress (w.faces)
template <typename T>
void launch_update(tl_chunk *c, tl_chunk::RefineWork &w, const double *u_in, double *u_out)
{
    if (w.faces == RF_FACES_K32) launch_update_f<T, ppcg_faces<T>()>(c, w, u_in, u_out);
    else launch_update_f<T, cg_faces<T>()>(c, w, u_in, u_out);
}

// One inner iteration; with `ev`, ev[0..1] bracket k_rf_w and ev[2..3]
// k_rf_ur (their finishers outside the brackets).
template <typename T, int FACES>
void launch_iteration_f(tl_chunk *c, tl_chunk::RefineWork &w, hipEvent_t *ev)
{
    const Ext e = ext_interior(c);
    const int grid = grid_for(e, c->cap);
    const March m = make_march(e, MARCH_COLS);
    const int grid_m = march_grid(m, c->cap);
    T *const *f = reinterpret_cast<T *const *>(w.f);
    const T *src = (c->prec == PREC_NONE) ? f[RW_R] : f[RW_Z];
    const RfArgs ra_w{c->partials, w.st, RF_PW, w.hist, w.hist + RF_HIST, RF_HIST},
        ra_u{c->partials, w.st, RF_RRN_LAG, w.hist, w.hist + RF_HIST, RF_HIST};
    if (ev) (void)hipEventRecord(ev[0], c->s);
    if (stream_z_pays(c))
        hipLaunchKernelGGL((k_rf_w<T, FACES, true>), dim3(grid_m), dim3(BLOCK_THREADS), 0, c->s, w.g, e, m, src,
                           f[RW_P], f[RW_P2], f[RW_W], f[RW_KY], c->faces, ra_w);
    else
        hipLaunchKernelGGL((k_rf_w<T, FACES, false>), dim3(grid_m), dim3(BLOCK_THREADS), 0, c->s, w.g, e, m, src,
                           f[RW_P], f[RW_P2], f[RW_W], f[RW_KY], c->faces, ra_w);
    if (ev) (void)hipEventRecord(ev[1], c->s);
    hipLaunchKernelGGL(k_rf_finish, dim3(1), dim3(FINISH_THREADS), 0, c->s, ra_w, grid_m);
    if (ev) (void)hipEventRecord(ev[2], c->s);
#define TL_RFI(P)                                                                                       \
    hipLaunchKernelGGL((k_rf_ur<T, P, FACES>), dim3(grid), dim3(BLOCK_THREADS), 0, c->s, w.g, e, f[RW_D], f[RW_P], \
                       f[RW_P2], f[RW_R], f[RW_Z], f[RW_W], f[RW_KY], c->faces, ra_u)
    if (c->prec == PREC_BLOCK) TL_RFI(PREC_BLOCK);
    else if (c->prec == PREC_DIAG) TL_RFI(PREC_DIAG);
    else TL_RFI(PREC_NONE);
#undef TL_RFI
    if (ev) (void)hipEventRecord(ev[3], c->s);
    hipLaunchKernelGGL(k_rf_finish, dim3(1), dim3(FINISH_THREADS), 0, c->s, ra_u, grid);
}

template <typename T>
void launch_iteration(tl_chunk *c, tl_chunk::RefineWork &w, hipEvent_t *ev = nullptr)
{
    if (w.faces == RF_FACES_K32) launch_iteration_f<T, ppcg_faces<T>()>(c, w, ev);
    else launch_iteration_f<T, cg_faces<T>()>(c, w, ev);
}

constexpr int RF_BATCH = 16;

void release_work(tl_chunk::RefineWork &w)
{
    for (auto *gs : {&w.graph, &w.graph_ppcg})
        if (gs->exec) (void)hipGraphExecDestroy(gs->exec);
    for (void *p : w.base)
        if (p) (void)hipFree(p);
    for (void *p : {(void *)w.st, (void *)w.hist, (void *)w.coef})
        if (p) (void)hipFree(p);
    if (w.h_st) (void)hipHostFree(w.h_st);
    w = tl_chunk::RefineWork();
}

int rf_alloc_impl(tl_chunk *c, int which)
{
    tl_chunk::RefineWork &w = c->rf[which];
    hipError_t e;
    w.bytes = which ? 4 : 8;
    w.g = c->g;
    if (which == 1) {
        // FP32 layout: pitch a multiple of 32 floats and column hd on a
        // 128-byte boundary, margins as for the FP64 fields
        const int hd = c->g.hd;
        w.g.pitch = (c->g.x_max + 2 * hd + 2 + 31) / 32 * 32;
        const long colpad = (32 - (hd % 32)) % 32;
        const long rows = (long)c->g.y_max + 2 * hd + 16;
        w.off = 8L * w.g.pitch + colpad;
        w.elems = rows * w.g.pitch + colpad + 64;
        for (int i = 0; i < RW_N; ++i) {
            if ((e = hipMalloc(&w.base[i], sizeof(float) * w.elems)) != hipSuccess) return hip_fail("hipMalloc refine", e);
            TL_HIP_TRY(hipMemsetAsync(w.base[i], 0, sizeof(float) * w.elems, c->s), "refine alloc");
            w.f[i] = static_cast<float *>(w.base[i]) + w.off;
        }
    } else {
        // FP64: the chunk's own r, z, p fields and conductivity, plus d
        w.off = c->off;
        w.elems = c->elems2d;
        if ((e = hipMalloc(&w.base[RW_D], sizeof(double) * w.elems)) != hipSuccess) return hip_fail("hipMalloc refine", e);
        TL_HIP_TRY(hipMemsetAsync(w.base[RW_D], 0, sizeof(double) * w.elems, c->s), "refine alloc");
        w.f[RW_D] = static_cast<double *>(w.base[RW_D]) + w.off;
        w.f[RW_W] = c->f2[F_COND];
        w.f[RW_R] = c->f2[F_R];
        w.f[RW_Z] = c->f2[F_Z];
        w.f[RW_P] = c->f2[F_P];
        w.f[RW_P2] = c->f2[F_P2];
        w.f[RW_SD] = c->f2[F_SD];
        w.f[RW_SD2] = c->f2[F_SD2];
        w.f[RW_R2] = c->f2[F_R2];
    }
    if ((e = hipMalloc(&w.hist, sizeof(double) * 2 * RF_HIST)) != hipSuccess ||
        (e = hipMalloc(&w.coef, sizeof(double) * 2 * RF_MAX_INNER)) != hipSuccess)
        return hip_fail("hipMalloc refine history", e);
    TL_HIP_TRY(hipMemsetAsync(w.hist, 0, sizeof(double) * 2 * RF_HIST, c->s), "refine alloc");
    TL_HIP_TRY(hipMemsetAsync(w.coef, 0, sizeof(double) * 2 * RF_MAX_INNER, c->s), "refine alloc");
    if ((e = hipMalloc(&w.st, sizeof(RefineState))) != hipSuccess) return hip_fail("hipMalloc refine state", e);
    if ((e = hipHostMalloc(&w.h_st, sizeof(RefineState) * 2, hipHostMallocDefault)) != hipSuccess)
        return hip_fail("hipHostMalloc refine state", e);
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("refine alloc", e);
    return TL_OK;
}

int rf_alloc(tl_chunk *c, int which)
{
    if (c->rf[which].st) return TL_OK;
    const int rc = rf_alloc_impl(c, which);
    if (rc != TL_OK) release_work(c->rf[which]);
    return rc;
}

int rf_state_get(tl_chunk *c, tl_chunk::RefineWork &w, RefineState *h)
{
    hipError_t e = hipMemcpyAsync(&w.h_st[0], w.st, sizeof(RefineState), hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    if (e != hipSuccess) return hip_fail("refine state download", e);
    *h = w.h_st[0];
    return TL_OK;
}

// One PPCG outer step: the CG half-step pair with the PPCG form of the
// second kernel, then `inner` smoothing steps whose last one finishes the
// step's scalars (tea_solve.f90:231-267 with tea_leaf_run_ppcg_inner_steps,
// :422-475).
template <typename T>
void launch_ppcg_outer(tl_chunk *c, tl_chunk::RefineWork &w, int inner)
{
    const Ext e = ext_interior(c);
    const int grid = grid_for(e, c->cap);
    // one block per tile for the kernels that write no partials (2.6 % faster
    // FP64 refined PPCG at 4096^2 than blocks looping over two tiles each)
    const int grid_nr = grid_for(e, INT_MAX);
    const March m = make_march(e, MARCH_COLS);
    const int grid_m = march_grid(m, c->cap);
    T *const *f = reinterpret_cast<T *const *>(w.f);
    const T *src = (c->prec == PREC_NONE) ? f[RW_R] : f[RW_Z];
    const RfArgs ra_w{c->partials, w.st, RF_PW, nullptr, nullptr, 0},
        ra_u{c->partials, w.st, RF_RRN, nullptr, nullptr, 0};
    hipLaunchKernelGGL((k_rf_w<T, ppcg_faces<T>(), false>), dim3(grid_m), dim3(BLOCK_THREADS), 0, c->s, w.g, e, m, src,
                       f[RW_P], f[RW_P2], f[RW_W], f[RW_KY], c->faces, ra_w);
    hipLaunchKernelGGL(k_rf_finish, dim3(1), dim3(FINISH_THREADS), 0, c->s, ra_w, grid_m);
    const double *ca = w.coef, *cb = w.coef + RF_MAX_INNER;
    // two inner steps per launch: r and sd flip between their two buffers at
    // every pair; the outer kernel writes r to the second buffer when the
    // number of pairs is odd, so that every outer step ends with r in RW_R
    const int pairs = c->ppcg_steps_per_launch == 2 ? inner / 2 : 0;
    T *r_cur = (pairs & 1) ? f[RW_R2] : f[RW_R], *r_oth = (pairs & 1) ? f[RW_R] : f[RW_R2];
    T *sd_cur = f[RW_SD], *sd_oth = f[RW_SD2];
    const double *Kx = c->f2[F_KX], *Ky = c->f2[F_KY];
#define TL_RFO(P)                                                                                              \
    do {                                                                                                       \
        hipLaunchKernelGGL((k_rf_ur_ppcg<T, P>), dim3(grid_nr), dim3(BLOCK_THREADS), 0, c->s, w.g, e, f[RW_D],     \
                           f[RW_P], f[RW_P2], f[RW_R], r_cur, f[RW_Z], f[RW_SD], f[RW_W], f[RW_KY], c->faces, \
                           ra_u);                                                                              \
        int i = 0;                                                                                             \
        for (int k = 0; k < pairs; ++k, i += 2) {                                                              \
            if (i + 2 == inner)                                                                                \
                hipLaunchKernelGGL((k_rf_inner2<T, P, 1>), dim3(grid), dim3(BLOCK_THREADS), 0, c->s, w.g, e,    \
                                   f[RW_D], r_cur, r_oth, sd_cur, sd_oth, f[RW_W], f[RW_KY], Kx, Ky, f[RW_Z],   \
                                   ca, cb, i, ra_u);                                                           \
            else                                                                                               \
                hipLaunchKernelGGL((k_rf_inner2<T, P, 0>), dim3(grid_nr), dim3(BLOCK_THREADS), 0, c->s, w.g, e,    \
                                   f[RW_D], r_cur, r_oth, sd_cur, sd_oth, f[RW_W], f[RW_KY], Kx, Ky, f[RW_Z],   \
                                   ca, cb, i, ra_u);                                                           \
            std::swap(r_cur, r_oth);                                                                           \
            std::swap(sd_cur, sd_oth);                                                                         \
        }                                                                                                      \
        for (; i < inner; ++i) {                                                                               \
            if (i == inner - 1)                                                                                \
                hipLaunchKernelGGL((k_rf_inner<T, P, 1>), dim3(grid), dim3(BLOCK_THREADS), 0, c->s, w.g, e,     \
                                   f[RW_D], r_cur, sd_cur, sd_oth, f[RW_W], f[RW_KY], c->faces, f[RW_Z], ca, cb, \
                                   i, ra_u);                                                                   \
            else                                                                                               \
                hipLaunchKernelGGL((k_rf_inner<T, P, 0>), dim3(grid_nr), dim3(BLOCK_THREADS), 0, c->s, w.g, e,     \
                                   f[RW_D], r_cur, sd_cur, sd_oth, f[RW_W], f[RW_KY], c->faces, f[RW_Z], ca, cb, \
                                   i, ra_u);                                                                   \
            std::swap(sd_cur, sd_oth);                                                                         \
        }                                                                                                      \
    } while (0)
    if (c->prec == PREC_BLOCK) TL_RFO(PREC_BLOCK);
    else if (c->prec == PREC_DIAG) TL_RFO(PREC_DIAG);
    else TL_RFO(PREC_NONE);
#undef TL_RFO
    hipLaunchKernelGGL(k_rf_finish, dim3(1), dim3(FINISH_THREADS), 0, c->s, ra_u, grid);
}

// CG iterations: the first two iterations of each batch are launched eagerly
// (timed), the other RF_BATCH - 2 replayed
template <typename T>
int rf_capture(tl_chunk *c, tl_chunk::RefineWork &w)
{
    return capture_graph(c, w.graph, c->prec * 4 + w.faces, [&]() {
        for (int b = 0; b < RF_BATCH - 2; ++b) launch_iteration<T>(c, w);
    });
}

int ppcg_batch(int inner) { return inner + 2 >= 32 ? 1 : 32 / (inner + 2); }

template <typename T>
int rf_capture_ppcg(tl_chunk *c, tl_chunk::RefineWork &w, int inner)
{
    return capture_graph(c, w.graph_ppcg, (c->prec * 1000 + inner) * 4 + c->ppcg_steps_per_launch, [&]() {
        for (int b = 0; b < ppcg_batch(inner); ++b) launch_ppcg_outer<T>(c, w, inner);
    });
}

struct KernelTimes {
    double k1_ms = 0, k2_ms = 0;
    int samples = 0;
};

// Replays `exec` until the device state halts; batch i+1 is queued before the
// state after batch i is examined. With `timed`, each batch starts with two
// eagerly launched CG iterations, each kernel bracketed by events; the events
// of a batch are read once it has finished, and only when it did not halt
// (then both eager iterations ran).
template <typename T>
int rf_batches(tl_chunk *c, tl_chunk::RefineWork &w, hipGraphExec_t exec, bool timed, long max_batches,
               KernelTimes *kt)
{
    hipError_t e;
    auto enqueue = [&](int sl) -> hipError_t {
        if (timed) {
            // two iterations, so that both forms of k_rf_ur (with and without
            // the lagged update of d) are timed
            launch_iteration<T>(c, w, c->ev_t[sl]);
            launch_iteration<T>(c, w, c->ev_t[sl] + 4);
        }
        hipError_t ee = hipGraphLaunch(exec, c->s);
        if (ee != hipSuccess) return ee;
        ee = hipMemcpyAsync(&w.h_st[sl], w.st, sizeof(RefineState), hipMemcpyDeviceToHost, c->s);
        if (ee != hipSuccess) return ee;
        return hipEventRecord(c->ev_batch[sl], c->s);
    };
    int slot = 0;
    if ((e = enqueue(0)) != hipSuccess) return hip_fail("hipGraphLaunch", e);
    for (long guard = 0;; ++guard) {
        if ((e = enqueue(1 - slot)) != hipSuccess) return hip_fail("hipGraphLaunch", e);
        if ((e = hipEventSynchronize(c->ev_batch[slot])) != hipSuccess) return hip_fail("batch wait", e);
        if (w.h_st[slot].halt) break;
        // a batch that did not halt ran both timed iterations
        float a = 0.f, b = 0.f, a2 = 0.f, b2 = 0.f;
        if (timed && hipEventElapsedTime(&a, c->ev_t[slot][0], c->ev_t[slot][1]) == hipSuccess &&
            hipEventElapsedTime(&b, c->ev_t[slot][2], c->ev_t[slot][3]) == hipSuccess &&
            hipEventElapsedTime(&a2, c->ev_t[slot][4], c->ev_t[slot][5]) == hipSuccess &&
            hipEventElapsedTime(&b2, c->ev_t[slot][6], c->ev_t[slot][7]) == hipSuccess) {
            kt->k1_ms += a + a2;
            kt->k2_ms += b + b2;
            kt->samples += 2;
        }
        if (guard > max_batches) {
            (void)hipStreamSynchronize(c->s);
            set_error("refined solve did not halt after %ld batches", guard);
            return TL_ERR_SOLVER;
        }
        slot = 1 - slot;
    }
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("batch drain", e);
    (void)hipGetLastError();
    return TL_OK;
}

// Runs batches of `exec` with reliable updates between them until the solve
// has converged, stalled, run out of iterations, or halted at stop_n.
template <typename T>
int rf_phase(tl_chunk *c, tl_chunk::RefineWork &w, hipGraphExec_t exec, bool timed, long max_batches,
             KernelTimes *kt, double **ubuf, int *cur, RefineState *h)
{
    int rc;
    hipError_t e;
    while (!h->converged && h->n < h->max_iters && isfinite(h->rr) && h->stalls < STALL_LIMIT) {
        if ((rc = rf_batches<T>(c, w, exec, timed, max_batches, kt)) != TL_OK) return rc;
        if ((rc = rf_state_get(c, w, h)) != TL_OK) return rc;
        if (!h->want_update) break;
        launch_update<T>(c, w, ubuf[*cur], ubuf[1 - *cur]);
        *cur = 1 - *cur;
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail("launch refine update", e);
        if ((rc = rf_state_get(c, w, h)) != TL_OK) return rc;
    }
    return TL_OK;
}

// The refined solve. inner == 0: CG. inner > 0: PPCG with `inner` smoothing
// steps per outer step after `presteps` CG iterations, whose coefficients
// give the eigenvalue bounds (tea_leaf_cheby.f90:96-172, widened by
// 0.95 / 1.05 as tea_solve.f90:175-176 does).
template <typename T>
int rf_solve(tl_chunk *c, int which, double tol, int max_iters, double delta, int presteps, int inner,
             tl_refine_result *out)
{
    int rc = rf_alloc(c, which);
    if (rc != TL_OK) return rc;
    tl_chunk::RefineWork &w = c->rf[which];
    hipError_t e;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if ((e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess) {
        if (t0) (void)hipEventDestroy(t0);
        return hip_fail("hipEventCreate", e);
    }
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{t0, t1};
    w.faces = inner > 0 ? ppcg_faces<T>() : cg_faces<T>();
    if ((rc = rf_capture<T>(c, w)) != TL_OK) return rc;
    if (inner > 0 && (rc = rf_capture_ppcg<T>(c, w, inner)) != TL_OK) return rc;

    RefineState h;
    memset(&h, 0, sizeof h);
    h.tol = tol;
    h.delta = delta;
    h.max_iters = max_iters;
    h.stop_n = inner > 0 ? presteps : 0;
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("refine start", e);
    w.h_st[0] = h;
    TL_HIP_TRY(hipEventRecord(t0, c->s), "refine start");
    TL_HIP_TRY(hipMemcpyAsync(w.st, &w.h_st[0], sizeof(RefineState), hipMemcpyHostToDevice, c->s), "refine start");
    TL_HIP_TRY(hipMemsetAsync(w.base[RW_D], 0, (size_t)w.bytes * w.elems, c->s), "refine start");
    if (w.faces == RF_FACES_HALF)
        hipLaunchKernelGGL(k_rf_narrow_half, dim3(1024), dim3(256), 0, c->s, c->g, w.g, c->f2[F_COND],
                           static_cast<float *>(w.f[RW_W]));
    else if (w.faces == RF_FACES_K32)
        hipLaunchKernelGGL(k_rf_narrow_faces, dim3(1024), dim3(256), 0, c->s, c->g, w.g, c->f2[F_KX],
                           c->f2[F_KY], static_cast<float *>(w.f[RW_W]), static_cast<float *>(w.f[RW_KY]));
    double *ubuf[2] = {c->f2[F_U], c->f2[F_U2]};
    int cur = 0;
    launch_update<T>(c, w, ubuf[cur], ubuf[1 - cur]);
    cur = 1 - cur;
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail("launch refine update", e);
    if ((rc = rf_state_get(c, w, &h)) != TL_OK) return rc;
    KernelTimes kt;
    if ((rc = rf_phase<T>(c, w, w.graph.exec, true, max_iters / RF_BATCH + 4, &kt, ubuf, &cur, &h)) != TL_OK)
        return rc;
    int outer0 = h.n;
    if (inner > 0 && !h.converged && h.halt && !h.want_update && h.n >= h.stop_n && h.n < max_iters &&
        h.stalls < STALL_LIMIT) {
        // eigenvalue bounds from the CG coefficients, Chebyshev coefficients
        const int ns = h.n < RF_HIST ? h.n : RF_HIST;
        std::vector<double> a(ns), b(ns);
        TL_HIP_TRY(hipMemcpy(a.data(), w.hist, sizeof(double) * ns, hipMemcpyDeviceToHost), "refine history copy");
        TL_HIP_TRY(hipMemcpy(b.data(), w.hist + RF_HIST, sizeof(double) * ns, hipMemcpyDeviceToHost),
                   "refine history copy");
        double eigmin = 0, eigmax = 0, theta = 0;
        if (calc_eigenvalues(a.data(), b.data(), ns, &eigmin, &eigmax) != 0) {
            set_error("refined PPCG: eigenvalue estimate from %d CG steps failed", ns);
            return TL_ERR_SOLVER;
        }
        eigmin *= 0.95;
        eigmax *= 1.05;
        std::vector<double> ca, cb;
        calc_ch_coefs(ca, cb, eigmin, eigmax, &theta, inner);
        TL_HIP_TRY(hipMemcpy(w.coef, ca.data(), sizeof(double) * inner, hipMemcpyHostToDevice), "refine coefficients");
        TL_HIP_TRY(hipMemcpy(w.coef + RF_MAX_INNER, cb.data(), sizeof(double) * inner, hipMemcpyHostToDevice),
                   "refine coefficients");
        // the outer steps update d at every step: complete the CG steps' d first
        {
            T *const *f = reinterpret_cast<T *const *>(w.f);
            hipLaunchKernelGGL(k_rf_d_fix<T>, dim3(grid_for(ext_interior(c), c->cap)), dim3(BLOCK_THREADS), 0, c->s,
                               w.g, ext_interior(c), f[RW_D], f[RW_P], f[RW_P2], w.st);
        }
        h.d_lag = 0;
        h.theta = theta;
        h.stop_n = 0;
        h.halt = 0;
        w.h_st[0] = h;
        TL_HIP_TRY(hipMemcpyAsync(w.st, &w.h_st[0], sizeof(RefineState), hipMemcpyHostToDevice, c->s),
                   "refine state copy");
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail("launch refine d fix", e);
        outer0 = h.n;
        const long mb = (max_iters - h.n) / ppcg_batch(inner) + 4;
        if ((rc = rf_phase<T>(c, w, w.graph_ppcg.exec, false, mb, &kt, ubuf, &cur, &h)) != TL_OK) return rc;
    }
    if (cur == 1) launch_copy(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U2], c->f2[F_U]);
    TL_HIP_TRY(hipEventRecord(t1, c->s), "refine finish");
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("refine finish", e);
    float ms = 0.f;
    TL_HIP_TRY(hipEventElapsedTime(&ms, t0, t1), "refine finish");
    if (out) {
        memset(out, 0, sizeof *out);
        out->seconds = 1e-3 * ms;
        out->initial_residual = sqrt(h.rr0);
        out->rel_residual = h.rr0 > 0 ? sqrt(h.rr_true / h.rr0) : 0.0;
        out->iterations = h.n;
        out->updates = h.updates;
        out->converged = h.converged;
        out->stalled = !h.converged && h.stalls >= STALL_LIMIT;
        out->k_samples = kt.samples;
        out->k1_ms = kt.samples ? kt.k1_ms / kt.samples : -1.0;
        out->k2_ms = kt.samples ? kt.k2_ms / kt.samples : -1.0;
        out->inner_iterations = inner > 0 ? (h.n - outer0) * inner : 0;
    }
    if (!h.converged && !(h.n >= max_iters) && h.stalls < STALL_LIMIT) {
        set_error("refined solve broke down after %d iterations (residual %g)", h.n, sqrt(h.rr));
        return TL_ERR_SOLVER;
    }
    return TL_OK;
}

}  // namespace rf

void refine_release(tl_chunk *c)
{
    for (auto &w : c->rf) rf::release_work(w);
}

}  // namespace tl

using namespace tl;

extern "C" int tl_solve_refined(tl_chunk *c, int precision, double tol, int max_iters, double delta,
                                tl_refine_result *out)
{
    if (!c) {
        set_error("tl_solve_refined: null chunk");
        return TL_ERR_ARG;
    }
    if ((precision != TL_PRECISION_FP64 && precision != TL_PRECISION_MIXED) || !(tol > 0.0) || max_iters < 1 ||
        delta < 0.0 || delta >= 1.0) {
        set_error("tl_solve_refined: invalid precision %d, tol %g, max_iters %d or delta %g", precision, tol,
                  max_iters, delta);
        return TL_ERR_ARG;
    }
    if (precision == TL_PRECISION_MIXED) return rf::rf_solve<float>(c, 1, tol, max_iters, delta, 0, 0, out);
    return rf::rf_solve<double>(c, 0, tol, max_iters, delta, 0, 0, out);
}

extern "C" int tl_solve_refined_ppcg(tl_chunk *c, int precision, double tol, int max_iters, double delta,
                                     int presteps, int inner_steps, tl_refine_result *out)
{
    if (!c) {
        set_error("tl_solve_refined_ppcg: null chunk");
        return TL_ERR_ARG;
    }
    if ((precision != TL_PRECISION_FP64 && precision != TL_PRECISION_MIXED) || !(tol > 0.0) || max_iters < 1 ||
        delta < 0.0 || delta >= 1.0 || presteps < 2 || presteps > rf::RF_HIST || inner_steps < 1 ||
        inner_steps > rf::RF_MAX_INNER) {
        set_error("tl_solve_refined_ppcg: invalid precision %d, tol %g, max_iters %d, delta %g, presteps %d or "
                  "inner steps %d",
                  precision, tol, max_iters, delta, presteps, inner_steps);
        return TL_ERR_ARG;
    }
    if (precision == TL_PRECISION_MIXED)
        return rf::rf_solve<float>(c, 1, tol, max_iters, delta, presteps, inner_steps, out);
    return rf::rf_solve<double>(c, 0, tol, max_iters, delta, presteps, inner_steps, out);
}
