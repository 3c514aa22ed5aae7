/*
 * TeaLeaf CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 C/OpenMP restatement of the reference TeaLeaf_OpenCL solve path
 * (kernel_files/ *.cl, tea_leaf_kernel_ocl.cpp, tea_solve.f90 and the Fortran
 * driver around it). It is the checker for the HIP library in
 * tealeaf_opencl_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. Nothing shipped calls into it.
 *
 * Memory layout follows the reference exactly (kernel_files/macros_cl.cl:20-22):
 * idx(j,k) = j + k*(x_max + 2*hd), j,k in [0, n+2hd), interior [hd, hd+n).
 * Halo cells start at zero (ocl_buffers.cpp:3-19) and are never written by the
 * solver kernels.
 *
 * Reductions are summed per row, then rows in order, so results do not depend
 * on the OpenMP thread count.
 */
#ifndef TL_ORACLE_H
#define TL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* enums: kernel_files/definitions.hpp:14-56, data.f90:38-75 */
enum { TLO_PREC_NONE = 1, TLO_PREC_JAC_DIAG = 2, TLO_PREC_JAC_BLOCK = 3 };
enum { TLO_SOLVER_JACOBI = 1, TLO_SOLVER_CG = 2, TLO_SOLVER_CHEBY = 3, TLO_SOLVER_PPCG = 4 };
enum { TLO_COEF_CONDUCTIVITY = 1, TLO_COEF_RECIP_CONDUCTIVITY = 2 };
enum { TLO_G_RECT = 1, TLO_G_CIRC = 2, TLO_G_POINT = 3 };

/* 2-D fields of the chunk (types.hpp:42-367, ocl_buffers.cpp:36-69) */
enum {
    TLO_DENSITY = 0, TLO_ENERGY0, TLO_ENERGY1, TLO_VOLUME, TLO_XAREA, TLO_YAREA,
    TLO_U, TLO_U0, TLO_P, TLO_R, TLO_W, TLO_MI, TLO_KX, TLO_KY, TLO_SD, TLO_Z,
    TLO_CP, TLO_BFP, TLO_NUM_FIELDS2D
};
/* 1-D coordinate arrays */
enum {
    TLO_CELLX = 0, TLO_CELLDX, TLO_VERTEXX, TLO_VERTEXDX,
    TLO_CELLY, TLO_CELLDY, TLO_VERTEXY, TLO_VERTEXDY, TLO_NUM_FIELDS1D
};

typedef struct tlo_chunk tlo_chunk;

#define TLO_MAX_STATES 64

/* Deck contents: read_input.f90:38-91 defaults + keywords, plus the C++
 * side-channel keys of ocl_init.cpp:131-220. */
typedef struct {
    double xmin, ymin, xmax, ymax;
    int x_cells, y_cells;
    double end_time;
    int end_step;
    double dtinit;
    int summary_frequency, visit_frequency;
    int max_iters;
    double eps;
    int coefficient;
    int tl_ch_cg_presteps;
    double tl_ch_cg_epslim;
    int tl_check_result;
    int tl_ppcg_inner_steps;
    int reflective_boundary;
    int solver;               /* Fortran flags: last tl_use_* keyword wins */
    int fortran_preconditioner;
    int verbose_on, profiler_on;
    int halo_depth;           /* Fortran halo_depth keyword */
    int test_problem;
    /* tl_ppcg_fresh_z (an option of this library, not a reference keyword):
     * PPCG forms r.z with z = M^-1 r of the final r of the inner steps, as the
     * Fortran TeaLeaf's calc_zrnorm does, instead of the OpenCL reference's
     * stale z (tea_leaf_ppcg_cl.cl:52,145 never write z) */
    int ppcg_fresh_z;
    int nstates;
    double state_density[TLO_MAX_STATES], state_energy[TLO_MAX_STATES];
    double state_xmin[TLO_MAX_STATES], state_xmax[TLO_MAX_STATES];
    double state_ymin[TLO_MAX_STATES], state_ymax[TLO_MAX_STATES];
    double state_radius[TLO_MAX_STATES];
    int state_geometry[TLO_MAX_STATES];
    /* C++ side channel (ocl_init.cpp:144-220) */
    int sc_preconditioner;    /* TLO_PREC_* */
    int sc_halo_depth;        /* -1 when absent or unparsable */
    int sc_solver;            /* TLO_SOLVER_* */
    int sc_profiler_on;
} tlo_config;

/* ---- chunk lifecycle and raw field access ---- */
tlo_chunk *tlo_create(int x_max, int y_max, int halo_depth, int preconditioner);
void tlo_destroy(tlo_chunk *c);
double *tlo_field2d(tlo_chunk *c, int id);
double *tlo_field1d(tlo_chunk *c, int id);
int tlo_pitch(const tlo_chunk *c);
long tlo_field2d_len(const tlo_chunk *c);
void tlo_set_preconditioner(tlo_chunk *c, int preconditioner);

/* ---- kernels (one per reference entry point) ---- */
void tlo_initialise_chunk(tlo_chunk *c, double xmin, double ymin, double dx, double dy);
void tlo_generate_chunk(tlo_chunk *c, int nstates, const double *dens, const double *ener,
                        const double *xmin, const double *xmax, const double *ymin,
                        const double *ymax, const double *radius, const int *geom,
                        int g_rect, int g_circ, int g_point);
void tlo_set_field(tlo_chunk *c);
void tlo_field_summary(tlo_chunk *c, double *vol, double *mass, double *ie, double *temp);
void tlo_update_halo(tlo_chunk *c, const int *chunk_neighbours, const int *fields, int depth);
void tlo_common_init(tlo_chunk *c, int coefficient, double dt, double *rx, double *ry,
                     const int *zero_boundary, int reflective_boundary);
void tlo_calc_residual(tlo_chunk *c);
double tlo_calc_2norm(tlo_chunk *c, int norm_array);
void tlo_finalise(tlo_chunk *c);
double tlo_cg_init(tlo_chunk *c);
double tlo_cg_calc_w(tlo_chunk *c);
double tlo_cg_calc_ur(tlo_chunk *c, double alpha);
void tlo_cg_calc_p(tlo_chunk *c, double beta);
void tlo_cheby_init(tlo_chunk *c, const double *ch_alphas, const double *ch_betas, int n_coefs,
                    double rx, double ry, double theta);
void tlo_cheby_iterate(tlo_chunk *c, int cheby_calc_step);
void tlo_ppcg_init(tlo_chunk *c, const double *ch_alphas, const double *ch_betas,
                   double theta, int n_inner_steps);
void tlo_ppcg_init_sd(tlo_chunk *c);
void tlo_ppcg_inner(tlo_chunk *c, int inner_step, int bounds_extra, const int *chunk_neighbours);
double tlo_jacobi_solve(tlo_chunk *c);
/* single building blocks exposed for invariant tests */
void tlo_block_init(tlo_chunk *c);
void tlo_block_solve(tlo_chunk *c);
/* z = M^-1 r over the interior with the chunk's preconditioner (the
 * fresh-z PPCG option) */
void tlo_precondition_r(tlo_chunk *c);
void tlo_init_jac_diag(tlo_chunk *c);

/* ---- host-side scalar routines (tea_leaf_cheby.f90) ---- */
int tlo_tqli(double *d, double *e, int n);
int tlo_calc_eigenvalues(const double *cg_alphas, const double *cg_betas, int nsteps,
                         double *eigmin, double *eigmax);
void tlo_calc_ch_coefs(double *ch_alphas, double *ch_betas, double eigmin, double eigmax,
                       double *theta, int n);

/* ---- deck reader and driver (parse.f90, read_input.f90, start.f90, diffuse.f90,
 *      tea_solve.f90, field_summary.f90) ---- */
int tlo_parse_deck(const char *path, tlo_config *cfg);
/* The parsed deck in the fixed arrays of the library's tl_parse_deck:
 * ints[23] = {x_cells, y_cells, end_step, summary_frequency,
 * visit_frequency, max_iters, coefficient, tl_ch_cg_presteps, tl_check_result,
 * tl_ppcg_inner_steps, reflective_boundary, solver, fortran_preconditioner,
 * verbose_on, profiler_on, halo_depth, test_problem, ppcg_fresh_z, nstates,
 * side-channel preconditioner, halo_depth, solver, profiler_on}, reals[8] =
 * {xmin, ymin, xmax, ymax, end_time, dtinit, eps, tl_ch_cg_epslim}, and for
 * each of the first max_states states states[8 s ..] = {density, energy, xmin,
 * xmax, ymin, ymax, radius, geometry}. */
int tlo_parse_deck_values(const char *path, int *ints, double *reals, double *states, int max_states);
/* Runs the whole deck; writes a line-oriented log (SUMMARY/SOLVE/...).
 * max_steps > 0 caps the number of timesteps (for bounded CPU samples).
 * Returns 0 on success, nonzero on a fatal error (message on stderr). */
int tlo_run_deck(const char *deck_path, const char *log_path, int max_steps);
int tlo_run_config(const tlo_config *cfg, const char *log_path, int max_steps);

/* Bounded CPU baseline sample: builds the deck's chunk, runs one timestep's
 * init and then exactly n_iters CG iterations (no exit test) and returns the
 * seconds spent in those iterations. */
double tlo_time_cg_iterations(const char *deck_path, int n_iters, int *iters_done);

int tlo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
