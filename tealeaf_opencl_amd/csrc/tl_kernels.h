// Host-side launchers of the kernels in tl_kernels.hip. Every launcher
// enqueues on the given stream and returns immediately; `cap` bounds the
// number of blocks (and so the number of reduction partials).
#pragma once

#include <hip/hip_runtime.h>

#include "tl_common.h"
#include "tl_device.h"

namespace tl {

// blocks of a tile launch (TL_TILE_LOOP): one per 64x16 tile, at most `cap`
// (blocks then loop over several tiles); INT_MAX for one block per tile
inline int grid_for(const Ext &e, int cap)
{
    const int tiles = e.ntx * e.nty;
    return tiles < cap ? (tiles > 0 ? tiles : 1) : cap;
}

void launch_init_chunk_1d(hipStream_t s, const Geom &g, double xmin, double ymin, double dx, double dy,
                          double *const *f1);
void launch_init_chunk_2d(hipStream_t s, const Geom &g, const Ext &e, int cap, double dx, double dy,
                          double *vol, double *xa, double *ya);
// sarr: density, energy, xmin, xmax, ymin, ymax, radius (device arrays)
void launch_generate(hipStream_t s, const Geom &g, const Ext &e, int cap, double *const *f1, double *rho,
                     double *e0, double *u, double *u0, int nstates, const double *const *sarr,
                     const int *sgeom, int g_rect, int g_circ, int g_point);
void launch_copy(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *src, double *dst);
// band-march geometry of an extent (TL_MARCH_LOOP): bands of `cols` output
// columns, segments of MARCH_ROWS rows (set_march_rows: another height), and
// the grid of a launch, capped at the partials' capacity
March make_march(const Ext &e, int cols);
void set_march_rows(March &m, const Ext &e, int rows);
int march_grid(const March &m, int cap);
// 4 fields read, 3 written (the stream mix of the CG update kernel), no stencil
void launch_stream_mix(hipStream_t s, const Geom &g, const Ext &e, int cap, double *a, const double *b, double *c,
                       const double *d, double *out);
void launch_field_summary(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *vol,
                          const double *rho, const double *en, const double *u, const RedArgs &ra);
// face: 0 left, 1 right, 2 bottom, 3 top
void launch_halo(hipStream_t s, const Geom &g, double *a, int depth, int face);
void launch_init_common(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho, double *Kx,
                        double *Ky, double rx, double ry, int coef);
// conductivity (density or 1/density) over the whole padded index space
void launch_conductivity(hipStream_t s, const Geom &g, const double *rho, double *w, int coef, int *bad);
void launch_zero_boundary(hipStream_t s, const Geom &g, const Ext &e, int cap, double *Kx, double *Ky,
                          const int *zb);
void launch_init_u(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho, const double *en,
                   double *u, double *u0);
void launch_init_jac_diag(hipStream_t s, const Geom &g, const Ext &e, int cap, double *Mi, const double *Kx,
                          const double *Ky);
// quot = num / den (IEEE) and fast = div_nr(num, den) for i < n
void launch_check_division(hipStream_t s, const double *num, const double *den, double *quot, double *fast, long n);
// z = Mi r over e
void launch_diag_apply(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *r, const double *Mi,
                       double *z);
void launch_calc_residual(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *u,
                          const double *u0, double *r, const double *Kx, const double *Ky);
void launch_dot(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *a, const double *b,
                const RedArgs &ra);
void launch_finalise(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *rho,
                     const double *u, double *en);
void launch_block_init(hipStream_t s, const Geom &g, const Ext &e, int cap, double *cp, double *bfp,
                       const double *Kx, const double *Ky);
void launch_block_solve(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *r, double *z,
                        const double *cp, const double *bfp, const double *Kx, const double *Ky);
void launch_cg_init_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double *p,
                      const double *r, double *z, const double *Mi, const RedArgs &ra);
void launch_cg_calc_w(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *p, double *w,
                      const double *Kx, const double *Ky, const double *W, const Faces &fc, const RedArgs &ra);
void launch_cg_calc_ur(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double alpha,
                       double *u, const double *p, double *r, const double *w, double *z,
                       const double *cp, const double *bfp, const double *Mi, const double *Kx,
                       const double *Ky, const double *W, const Faces &fc, const RedArgs &ra);
void launch_cg_calc_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, double beta,
                      double *p, const double *r, const double *z);
void launch_cheby_p(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int init,
                    const double *u, const double *u0, double *p, double *r, double *w,
                    const double *cp, const double *bfp, const double *Mi, const double *Kx,
                    const double *Ky, const double *W, const Faces &fc, const double *ch_a,
                    const double *ch_b, double theta, int step, const SolveState *st);
void launch_cheby_u(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, const double *p,
                    const double *r, const RedArgs &ra);
void launch_ppcg_init_sd(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *r,
                         double *sd, const double *cp, const double *bfp, const double *Mi,
                         const double *Kx, const double *Ky, double theta);
void launch_ppcg_update_r(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, double *r,
                          const double *Kx, const double *Ky, const double *W, const Faces &fc, const double *sd);
void launch_ppcg_calc_sd(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *r,
                         double *sd, const double *cp, const double *bfp, const double *Mi,
                         const double *Kx, const double *Ky, const double *W, const Faces &fc,
                         const double *ch_a, const double *ch_b, int step);
// without a solver state: u1 <- sweep(un); with one: the pair (u1, un) is
// double-buffered by the iteration counter
// W non-null: faces from the conductivity (bitwise the stored Kx, Ky)
void launch_jacobi(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *Kx, const double *Ky,
                   const double *W, const Faces &fc, const double *u0, double *u1, double *un, const RedArgs &ra);
// p is double-buffered (pA, pB); the iteration counter in the state picks
// which buffer is read and which is written. ev_main, when given, is recorded
// after the iteration kernel and before its finisher. The fused kernels
// recompute Kx, Ky from the conductivity W and the face parameters.
// stream_z: read z (src) with nontemporal loads (CG iterations on fields that
// stream_z_pays accepts; the PPCG outer step, which follows the inner steps,
// and larger fields measured faster without).
void launch_fused_cg_w(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *src,
                       double *pA, double *pB, const double *W, const Faces &fc, const RedArgs &ra,
                       hipEvent_t ev_main = nullptr, bool stream_z = true);
// parity 1 / 2: the launch is known to run an odd / even iteration (CG: the
// update kernel compiled for that parity, the odd one leaving u alone); 0:
// either parity (decided by the kernel from the state)
void launch_fused_cg_ur(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int ppcg, double *u,
                        const double *pA, const double *pB, const double *r, double *r_out, double *z,
                        double *sd, const double *W, const Faces &fc, const RedArgs &ra,
                        hipEvent_t ev_main = nullptr, int parity = 0);
// u += alpha p_L after a fused CG phase that ended on an odd iteration
void launch_cg_u_fix(hipStream_t s, const Geom &g, const Ext &e, int cap, double *u, const double *p,
                     const SolveState *st);
void launch_fused_jacobi(hipStream_t s, const Geom &g, const Ext &e, int cap, const double *W, const Faces &fc,
                         const double *u0, double *uA, double *uB, const RedArgs &ra,
                         hipEvent_t ev_main = nullptr);
// last: 0 a smoothing step, 1 the last one forming r.z with the outer step's
// z (the reference's stale z), 2 the last one storing the fresh z = M^-1 r and
// forming r.z with it (the fresh-z option)
void launch_fused_ppcg_inner(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int last,
                             double *u, double *r, const double *sd_in, double *sd_out,
                             const double *W, const Faces &fc, double *z, const double *ch_a,
                             const double *ch_b, int step, const RedArgs &ra, hipEvent_t ev_main = nullptr);
// two PPCG inner steps (`step`, `step`+1) per launch; r and sd are read from
// one buffer and written to another
void launch_fused_ppcg_inner2(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, int last, double *u,
                              const double *r_in, double *r_out, const double *sd_in, double *sd_out,
                              const double *Kx, const double *Ky, double *z, const double *ch_a,
                              const double *ch_b, int step, const RedArgs &ra, hipEvent_t ev_main = nullptr);
// two Chebyshev iterations (n odd) per launch; u and p are read from one
// buffer and written to the other
void launch_fused_cheby2(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *u0,
                         const double *u_in, double *u_out, const double *p_in, double *p_out, const double *Kx,
                         const double *Ky, const double *ch_a, const double *ch_b, const RedArgs &ra,
                         hipEvent_t ev_main = nullptr);
void launch_fused_cheby(hipStream_t s, const Geom &g, const Ext &e, int cap, int prec, const double *u0,
                        double *uA, double *uB, double *p, const double *W, const Faces &fc,
                        const double *ch_a, const double *ch_b, const RedArgs &ra, hipEvent_t ev_main = nullptr);

}  // namespace tl
