// Shared host/device definitions for the MI355X TeaLeaf solve library.
//
// Field layout in HBM: every 2-D field is one allocation; element (j,k) of the
// reference's padded index space (j,k in [0, n+2hd), interior [hd, hd+n),
// kernel_files/macros_cl.cl:20-22) lives at F[k*pitch + j], where F is the
// field's base pointer already advanced by a fixed offset. The offset and the
// pitch (a multiple of 16 doubles) are chosen so that column hd of every row
// starts on a 128-byte boundary: a wavefront's 64 consecutive interior cells
// are then exactly four aligned 128-byte lines. Margins of more than one row
// before and after the padded index space keep every +-1 stencil access of
// every kernel inside the allocation.
#pragma once

#include <stdint.h>

namespace tl {

// preconditioner / solver / coefficient enums: kernel_files/definitions.hpp:14-56
enum { PREC_NONE = 1, PREC_DIAG = 2, PREC_BLOCK = 3 };
enum { SOLVER_JACOBI = 1, SOLVER_CG = 2, SOLVER_CHEBY = 3, SOLVER_PPCG = 4 };
enum { COEF_CONDUCTIVITY = 1, COEF_RECIP_CONDUCTIVITY = 2 };

// 2-D fields (ids are part of the C-ABI, see include/tealeaf_mi355x.h)
enum Field2D {
    F_DENSITY = 0, F_ENERGY0, F_ENERGY1, F_VOLUME, F_XAREA, F_YAREA,
    F_U, F_U0, F_P, F_R, F_W, F_MI, F_KX, F_KY, F_SD, F_Z, F_CP, F_BFP,
    NUM_PUBLIC_2D,
    // second buffers of the double-buffered iteration kernels
    F_P2 = NUM_PUBLIC_2D, F_SD2, F_U2, F_R2,
    // conductivity w (density or 1/density) of the current timestep, from
    // which the fused iteration kernels recompute Kx, Ky
    F_COND,
    NUM_FIELDS_2D
};
enum Field1D {
    F_CELLX = 0, F_CELLDX, F_VERTEXX, F_VERTEXDX,
    F_CELLY, F_CELLDY, F_VERTEXY, F_VERTEXDY, NUM_FIELDS_1D
};

constexpr int TILE_X = 64;          // columns per tile = one wavefront
constexpr int STRIP = 4;            // rows per thread = JACOBI_BLOCK_SIZE (definitions.hpp:4)
constexpr int STRIPS_PER_TILE = 4;  // one strip per wavefront
constexpr int BLOCK_THREADS = TILE_X * STRIPS_PER_TILE;
constexpr int TILE_Y = STRIP * STRIPS_PER_TILE;
constexpr int MAX_RED = 4;          // widest reduction (field_summary)

// inclusive cell extent of one launch plus its tiling
struct Ext {
    int j0, j1, k0, k1;
    int ntx, nty;
};

struct Geom {
    int x_max, y_max, hd;
    int pitch;
};

// What tea_leaf_init_common + tea_leaf_zero_boundary (tea_leaf_common_cl.cl:
// 56-125) leave in Kx, Ky, as parameters for recomputing any face
// coefficient from the conductivity field: the scale factors and which walls
// are zeroed (none when the boundary is reflective).
struct Faces {
    double rx, ry;
    int zl, zr, zb, zt;
};

// Device-resident solver scalars: the per-iteration algebra of
// tea_solve.f90:147-332 runs in the finishing block of each iteration's last
// reduction, so the host never waits on a dot product inside the loop.
struct SolveState {
    // inputs (host-written before the loop)
    double eps;
    double epslim;
    int max_iters;
    int presteps;
    int solver;           // SOLVER_*
    int first;            // SAVE'd `first` of tea_solve.f90:57
    int est_itc;
    int n_fused0;         // first iteration run by a double-buffered kernel
    // running scalars
    double rro, pw, alpha, rrn, beta;
    double alpha_prev;    // alpha of the previous CG iteration (the fused CG update of u lags one iteration)
    double error, old_error, initial_residual, bb, exact_error;
    double theta;         // Chebyshev / least-squares centre, PPCG init_sd scale
    int n;                // iteration being executed (Fortran loop variable)
    int halt;             // 1: iteration kernels must not run (done or switch)
    int done;             // loop exited
    int exit_n;           // value of n after the loop (n-1 is printed)
    int switch_n;         // iteration at which ch_switch_check became true
    int cheby_calc_steps;
    int cg_calc_steps;
    int ppcg_inner_iters;
    int fault;            // 1: a kernel captured for one iteration parity met the other (all kernels stop)
};

// reduction finish modes: which scalar algebra the last block runs
enum RedMode {
    RED_STORE = 0,        // totals -> result[] only
    RED_INIT_RESID,       // initial_residual (tea_solve.f90:103-112)
    RED_INIT_RRO,         // rro after cg_init (:118-125) + switch test for n=1
    RED_CG_PW,            // pw -> alpha (:276-283)
    RED_CG_RRN,           // rrn -> beta, error, exit and switch tests (:287-330)
    RED_PPCG_PW,          // pw -> alpha in a PPCG outer step (:239-245)
    RED_PPCG_RRN,         // zrnorm -> beta, error, exit test (:255-266)
    RED_JACOBI,           // error (:301-307)
    RED_CHEBY,            // optional ||r||^2 then iteration tail (:216-229)
    RED_CHEBY2,           // two Chebyshev iterations (n odd): two tails
    RED_EXACT             // tl_check_result residual (:334-350)
};

struct RedArgs {
    double *partials;     // gridDim.x * MAX_RED
    double *result;       // MAX_RED totals
    SolveState *st;       // may be null (drop-in calls)
    double *hist_a;       // cg_alphas (1-based by iteration)
    double *hist_b;       // cg_betas
    double *hist_e;       // residual trajectory
    int mode;
    // host-returned totals (drop-in calls): result points at pinned host
    // memory and, once it holds them, the finisher stores seq into *flag
    unsigned int *flag;
    unsigned int seq;
};

// Device-resident scalars of the refined solve (tl_refine.hip): an inner
// preconditioned CG in the storage precision of the inner fields (FP32 or
// FP64) with FP64 scalars, whose recursively updated residual is replaced by
// the true FP64 residual u0 - A u at every reliable update.
struct RefineState {
    // inputs
    double tol;           // target ||u0 - A u||_2 / ||u0 - A u_init||_2
    double delta;         // update when ||r|| < delta * max ||r|| since the last one
    int max_iters;
    int pad0_;
    // running scalars
    double rr0;           // ||r||^2 of the first (initial) true residual
    double rr_true;       // ||r||^2 of the latest true residual
    double rr;            // ||r||^2 of the recursive residual
    double rmax;          // max recursive ||r|| since the last update
    double rro;           // r.z of the current residual
    double rz_prev;       // r.z before the latest inner step
    double pw, alpha, beta;
    int n;                // inner iterations done
    int halt;             // 1: inner iteration kernels must not run
    int want_update;      // 1: a reliable update is due
    int converged;
    int updates;          // true residuals formed (the initial one included)
    int stalls;           // consecutive updates whose true residual fell by less than half
    double rr_prev;       // ||r||^2 of the previous true residual
    double theta;         // PPCG: centre of the eigenvalue interval (sd = z / theta)
    int stop_n;           // > 0: halt (without an update) once n reaches it
    int d_lag;            // 1: d lacks the last iteration's alpha p (added by the next one)
    double alpha_prev;    // alpha of the last iteration (what d lacks while d_lag)
};

enum RefineMode {
    RF_UPDATE = 0,        // true residual r.z, r.r after u += d
    RF_PW,                // p.w -> alpha
    RF_RRN,               // recursive r.z, r.r -> beta, update / exit tests
    RF_RRN_LAG            // the same after a CG iteration whose d update lags one iteration
};

}  // namespace tl
