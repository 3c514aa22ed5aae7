/*
 * tealeaf_mi355x.h -- C-ABI of libtealeaf_mi355x.so, the MI355X (gfx950) HIP
 * implementation of TeaLeaf's implicit diffusion solve.
 *
 * Part 1 is the drop-in boundary: the Fortran-callable entry points that the
 * reference's Fortran driver calls (lower case, trailing underscore, every
 * argument by reference). Each one replaces the reference definition cited
 * beside it and keeps its argument meaning, its synchronous/asynchronous
 * behaviour (functions that return a scalar block until it is on the host)
 * and its error behaviour (print "@@@@@ Fatal error ..." to stderr and abort
 * the job: MPI_Abort when the process has MPI initialised, else exit(1)).
 * They act on one chunk created by initialise_ocl_, which reads ./tea.in for
 * the solver, tl_preconditioner_type=, halo_depth= and profiler_on with the
 * reference's own matching rules.
 *
 * Part 2 is a superset API used by the native driver, the Python package and
 * the tests: explicit chunk handles, status codes instead of aborts, field
 * upload/download in the reference layout, and the device-resident driver.
 *
 * Reference layout of a 2-D field, as exchanged by tl_field_get/tl_field_set:
 * (x_max + 2*hd) columns by (y_max + 2*hd) rows, row-major, interior cells at
 * [hd, hd+x_max) x [hd, hd+y_max) (kernel_files/macros_cl.cl:20-22).
 */
#ifndef TEALEAF_MI355X_H
#define TEALEAF_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================= */
/* Part 1: reference entry points                                          */
/* ======================================================================= */

/* ocl_init.cpp:11-17 -- create the chunk; reads ./tea.in (side channel). */
void initialise_ocl_(int *x_min, int *x_max, int *y_min, int *y_max);
/* initialise_chunk_kernel_ocl.cpp:3-7 -- coordinates, volume, face areas. */
void initialise_chunk_kernel_ocl_(double *xmin, double *ymin, double *dx, double *dy);
/* generate_chunk_kernel_ocl.cpp:3-23 -- paint states 1..n, then u = u0 = e0*density. */
void generate_chunk_kernel_ocl_(const int *number_of_states, const double *state_density,
                                const double *state_energy, const double *state_xmin,
                                const double *state_xmax, const double *state_ymin,
                                const double *state_ymax, const double *state_radius,
                                const int *state_geometry, const int *g_rect, const int *g_circ,
                                const int *g_point);
/* set_field_kernel_ocl.cpp:3-7 -- energy1 = energy0. */
void set_field_kernel_ocl_(void);
/* field_summary_kernel_ocl.cpp:4-8 -- blocking; volume, mass, mass*e1, mass*u sums. */
void field_summary_kernel_ocl_(double *vol, double *mass, double *ie, double *temp);
/* update_halo_kernel_ocl.cpp:6-12 -- reflective fill of density/energy0/energy1. */
void update_halo_kernel_ocl_(const int *chunk_neighbours, const int *fields, const int *depth);
/* tea_leaf_kernel_ocl.cpp:233-239 -- rx, ry are outputs; Kx, Ky; walls; u = u0 = e1*density.
 * Aborts (as on a bad coefficient) when an interior density lies outside
 * [2^-499, 2^499] or is not a number (see tl_common_init). */
void tea_leaf_common_init_kernel_ocl_(const int *coefficient, double *dt, double *rx, double *ry,
                                      int *zero_boundary, int *reflective_boundary);
/* tea_leaf_kernel_ocl.cpp:248-252 -- r = u0 - A u. */
void tea_leaf_calc_residual_ocl_(void);
/* tea_leaf_kernel_ocl.cpp:34-38 -- blocking; 0: u0.u0, 1: r.r, 2: r.z (r.r unpreconditioned). */
void tea_leaf_calc_2norm_kernel_ocl_(int *norm_array, double *norm);
/* tea_leaf_kernel_ocl.cpp:242-246 -- energy1 = u / density. */
void tea_leaf_common_finalise_kernel_ocl_(void);
/* tea_leaf_kernel_ocl.cpp:143-147 -- blocking; preconditioner setup, p, rro. */
void tea_leaf_cg_init_kernel_ocl_(double *rro);
/* tea_leaf_kernel_ocl.cpp:149-153 -- blocking; w = A p, pw = p.w. */
void tea_leaf_cg_calc_w_kernel_ocl_(double *pw);
/* tea_leaf_kernel_ocl.cpp:154-158 -- blocking; u += a p, r -= a w, z = M^-1 r, rrn. */
void tea_leaf_cg_calc_ur_kernel_ocl_(double *alpha, double *rrn);
/* tea_leaf_kernel_ocl.cpp:159-163 -- p = z + beta p (r + beta p unpreconditioned). */
void tea_leaf_cg_calc_p_kernel_ocl_(double *beta);
/* tea_leaf_kernel_ocl.cpp:42-48 -- copies n_coefs coefficients; runs init_p + calc_u. */
void tea_leaf_cheby_init_kernel_ocl_(const double *ch_alphas, const double *ch_betas, int *n_coefs,
                                     const double *rx, const double *ry, const double *theta);
/* tea_leaf_kernel_ocl.cpp:50-54 -- one Chebyshev step with coefficient index step-1. */
void tea_leaf_cheby_iterate_kernel_ocl_(const int *cheby_calc_step);
/* tea_leaf_kernel_ocl.cpp:306-311 -- PPCG coefficients and theta. */
void tea_leaf_ppcg_init_ocl_(const double *ch_alphas, const double *ch_betas, double *theta,
                             int *n_inner_steps);
/* tea_leaf_kernel_ocl.cpp:313-317 -- sd = M^-1 r / theta. */
void tea_leaf_ppcg_init_sd_kernel_ocl_(void);
/* tea_leaf_kernel_ocl.cpp:319-325 -- update_r + calc_sd on the matrix-powers extent. */
void tea_leaf_ppcg_inner_kernel_ocl_(int *inner_step, int *bounds_extra, const int *chunk_neighbours);
/* tea_leaf_kernel_ocl.cpp:216-220 -- blocking; one Jacobi sweep, returns sum |u - un|. */
void tea_leaf_jacobi_solve_kernel_ocl_(double *error);
/* ocl_pack.cpp:4-16 -- multi-chunk halo staging. One chunk per GPU never
 * reaches these (tea.f90:310 returns first); they abort if called. */
void ocl_pack_buffers_(int fields[7], int offsets[7], int *depth, int *face, double *host_buffer);
void ocl_unpack_buffers_(int fields[7], int offsets[7], int *depth, int *face, double *host_buffer);
/* ocl_errors.cpp:196-200 -- per-kernel time table when profiler_on. */
void print_opencl_profiling_info_(void);

/* ======================================================================= */
/* Part 2: superset API (status codes; no aborts)                          */
/* ======================================================================= */

enum {
    TL_OK = 0,
    TL_ERR_ARG = 1,     /* bad argument */
    TL_ERR_HIP = 2,     /* HIP runtime failure (no device, out of memory, launch) */
    TL_ERR_CONFIG = 3,  /* deck or side-channel error */
    TL_ERR_SOLVER = 4   /* solver failure (eigenvalue estimate) */
};

/* preconditioners, solvers, coefficients: kernel_files/definitions.hpp:14-56 */
enum { TL_PREC_NONE = 1, TL_PREC_JAC_DIAG = 2, TL_PREC_JAC_BLOCK = 3 };
enum { TL_SOLVER_JACOBI = 1, TL_SOLVER_CG = 2, TL_SOLVER_CHEBY = 3, TL_SOLVER_PPCG = 4 };
enum { TL_COEF_CONDUCTIVITY = 1, TL_COEF_RECIP_CONDUCTIVITY = 2 };

/* 2-D field ids */
enum {
    TL_F_DENSITY = 0, TL_F_ENERGY0, TL_F_ENERGY1, TL_F_VOLUME, TL_F_XAREA, TL_F_YAREA,
    TL_F_U, TL_F_U0, TL_F_P, TL_F_R, TL_F_W, TL_F_MI, TL_F_KX, TL_F_KY, TL_F_SD, TL_F_Z,
    TL_F_CP, TL_F_BFP, TL_NUM_FIELDS2D
};
/* 1-D field ids */
enum {
    TL_F_CELLX = 0, TL_F_CELLDX, TL_F_VERTEXX, TL_F_VERTEXDX,
    TL_F_CELLY, TL_F_CELLDY, TL_F_VERTEXY, TL_F_VERTEXDY, TL_NUM_FIELDS1D
};

/* driver modes of tl_run_deck */
enum {
    TL_MODE_DROPIN = 0, /* reference call granularity, scalars on the host */
    TL_MODE_NATIVE = 1  /* fused kernels, device-resident scalars, hipGraph batches */
};

typedef struct tl_chunk tl_chunk;

const char *tl_last_error(void);
int tl_device_count(int *count);

int tl_chunk_create(int x_max, int y_max, int halo_depth, int preconditioner, int profiler_on,
                    tl_chunk **out);
int tl_chunk_destroy(tl_chunk *c);
/* chunk the Part 1 entry points act on (initialise_ocl_ sets it too) */
int tl_chunk_set_default(tl_chunk *c);
tl_chunk *tl_chunk_default(void);
int tl_chunk_info(const tl_chunk *c, int *x_max, int *y_max, int *halo_depth, int *preconditioner);

int tl_field_get(tl_chunk *c, int field, double *host);
int tl_field_set(tl_chunk *c, int field, const double *host);
int tl_field1d_get(tl_chunk *c, int field, double *host, int n);

/* the bodies of the Part 1 entry points, one per reference kernel call */
int tl_initialise_chunk(tl_chunk *c, double xmin, double ymin, double dx, double dy);
int tl_generate_chunk(tl_chunk *c, int nstates, const double *density, const double *energy,
                      const double *xmin, const double *xmax, const double *ymin, const double *ymax,
                      const double *radius, const int *geometry, int g_rect, int g_circ, int g_point);
/* generate_chunk for callers that cannot allocate device memory (code under
 * HIP graph capture, tight loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference. The same painting as tl_generate_chunk, whose
 * state arrays are staged in device buffers allocated and freed on every
 * call; here they are staged in a file-scope device array of
 * TL_NOALLOC_MAX_STATES states (about 30 KB, one per device, held until the
 * process ends). TL_ERR_ARG for more states. Calls that overlap on one device
 * share that array and race. */
enum { TL_NOALLOC_MAX_STATES = 512 };
int tl_generate_chunk_noalloc(tl_chunk *c, int nstates, const double *density, const double *energy,
                              const double *xmin, const double *xmax, const double *ymin, const double *ymax,
                              const double *radius, const int *geometry, int g_rect, int g_circ, int g_point);
int tl_set_field(tl_chunk *c);
int tl_field_summary(tl_chunk *c, double *vol, double *mass, double *ie, double *temp);
int tl_update_halo(tl_chunk *c, const int *chunk_neighbours, const int *fields, int depth);
/* TL_ERR_ARG for an unknown coefficient, or for an interior density outside
 * [2^-499, 2^499] or not a number: inside that range every divisor the
 * iteration kernels form lies where their reciprocal-Newton division is the
 * IEEE quotient (the reference would go on with infinite or NaN faces). */
int tl_common_init(tl_chunk *c, int coefficient, double dt, double *rx, double *ry,
                   const int *zero_boundary, int reflective_boundary);
int tl_calc_residual(tl_chunk *c);
int tl_calc_2norm(tl_chunk *c, int norm_array, double *norm);
int tl_finalise(tl_chunk *c);
int tl_cg_init(tl_chunk *c, double *rro);
int tl_cg_calc_w(tl_chunk *c, double *pw);
int tl_cg_calc_ur(tl_chunk *c, double alpha, double *rrn);
int tl_cg_calc_p(tl_chunk *c, double beta);
int tl_cheby_init(tl_chunk *c, const double *ch_alphas, const double *ch_betas, int n_coefs,
                  double rx, double ry, double theta);
int tl_cheby_iterate(tl_chunk *c, int cheby_calc_step);
int tl_ppcg_init(tl_chunk *c, const double *ch_alphas, const double *ch_betas, double theta,
                 int n_inner_steps);
int tl_ppcg_init_sd(tl_chunk *c);
int tl_ppcg_inner(tl_chunk *c, int inner_step, int bounds_extra, const int *chunk_neighbours);
int tl_jacobi_solve(tl_chunk *c, double *error);
/* building blocks of tl_cg_init, exposed for preconditioner tests */
int tl_block_init(tl_chunk *c);
int tl_block_solve(tl_chunk *c);
int tl_init_jac_diag(tl_chunk *c);
int tl_synchronize(tl_chunk *c);
int tl_print_profiling(tl_chunk *c);
/* Measurement only (no reference counterpart): attainable streaming bandwidth
 * in bytes/s on this chunk's layout, from `reps` event-timed launches of a
 * copy (kind 0, 16 B/cell) or of a 4-read/3-write stream with the mix of the
 * CG update kernel (kind 1, 56 B/cell), over temporary fields. */
int tl_stream_bandwidth(tl_chunk *c, int kind, int reps, double *bytes_per_s);
/* The deck as tl_run_deck reads it (read_input.f90 keywords and defaults,
 * parse.f90 line handling, the ocl_init.cpp side channel), without a GPU; no
 * reference counterpart: the settings come back in fixed arrays, ints[23] = {x_cells, y_cells, end_step, summary_frequency,
 * visit_frequency, max_iters, coefficient, tl_ch_cg_presteps, tl_check_result,
 * tl_ppcg_inner_steps, reflective_boundary, solver, fortran_preconditioner,
 * verbose_on, profiler_on, halo_depth, test_problem, ppcg_fresh_z, nstates,
 * side-channel preconditioner, halo_depth, solver, profiler_on}, reals[8] =
 * {xmin, ymin, xmax, ymax, end_time, dtinit, eps, tl_ch_cg_epslim}, and for
 * each of the first max_states states states[8 s ..] = {density, energy, xmin,
 * xmax, ymin, ymax, radius, geometry}.
 * TL_ERR_CONFIG with tl_last_error() set when the deck is rejected. */
int tl_parse_deck(const char *deck, int *ints, double *reals, double *states, int max_states);
/* The host routines the drivers use between the CG presteps and the
 * Chebyshev / PPCG phase (tea_calc_eigenvalues and tea_calc_ch_coefs of
 * tea_leaf_cheby.f90:96-172; Fortran routines, no C entry point in the
 * reference): the eigenvalue bounds of the Lanczos matrix of n CG steps
 * (alphas/betas 0-based; status TL_ERR_SOLVER when tqli does not converge or
 * a bound is negative), and n Chebyshev coefficients with theta. No GPU. */
int tl_calc_eigenvalues(const double *cg_alphas, const double *cg_betas, int n, double *eigmin, double *eigmax);
int tl_calc_ch_coefs(double eigmin, double eigmax, int n, double *ch_alphas, double *ch_betas, double *theta);
/* The Fortran edit descriptors tea.out is written with (tl_run_deck_out; no
 * reference counterpart): kind 0 Ew.d, 1 1PEw.d, 2 Fw.d, 3 a list-directed
 * REAL(8) item (w, d unused); the text of v goes to buf (len bytes, NUL
 * terminated). Ew.d, 1PEw.d and Fw.d follow the Fortran standard's
 * editing; a list-directed item's form is compiler-defined (a fixed E25.16
 * style here). */
int tl_format_fortran(int kind, double v, int w, int d, char *buf, int len);
/* Self-check of the division the iteration kernels use for faces, block
 * factors and diagonals (no reference counterpart): for i < n, quot[i] =
 * num[i] / den[i] by the compiler's IEEE FP64 division and fast[i] by the
 * kernels' reciprocal-Newton sequence, both on the GPU; host arrays. */
int tl_check_division(const double *num, const double *den, double *quot, double *fast, long n);

/* Tuning of the native driver's fused kernels. No reference counterpart.
 * TL_TUNE_PPCG_STEPS_PER_LAUNCH: PPCG inner steps per fused launch, 1 or 2
 * (default 2); TL_TUNE_CHEBY_STEPS_PER_LAUNCH: Chebyshev iterations per fused
 * launch, 1 or 2 (default 2). Either way the fields are bitwise the same.
 * With c null the value becomes the default of chunks created afterwards
 * (tl_run_deck creates its own chunk). */
enum { TL_TUNE_PPCG_STEPS_PER_LAUNCH = 1, TL_TUNE_CHEBY_STEPS_PER_LAUNCH = 2 };
int tl_set_tuning(tl_chunk *c, int key, int value);
/* PPCG smoothing steps first_step .. first_step+nsteps-1 (1-based, as the
 * inner_step of tl_ppcg_inner) with the native
 * driver's fused kernels (one or two per launch as tuned) on the chunk's u, r
 * and sd, over the interior: the same fields as nsteps tl_ppcg_inner calls
 * with bounds_extra 0 (tea_leaf_ppcg_cl.cl:62-155), with the face
 * coefficients of the last tl_common_init and the coefficients of the last
 * tl_ppcg_init. No reference counterpart (the reference launches one step per
 * call). */
int tl_ppcg_inner_fused(tl_chunk *c, int first_step, int nsteps);

/* Whole-deck driver (start.f90 + diffuse.f90 + tea_solve.f90 restated).
 * Writes a line-oriented log: SUMMARY step vol mass density ie temp /
 * SOLVE step iters conduction_err exact_err ppcg_inner cg_steps cheby_steps
 * switch_n est_itc / SWITCH / ESTITC / RESID / QA lines. max_steps > 0 caps
 * the number of timesteps. */
int tl_run_deck(const char *deck_path, const char *log_path, int mode, int max_steps);
/* tl_run_deck that also writes the reference's tea.out lines around the solve
 * to tea_out_path (may be null): the step line (timestep.f90:58), the solve
 * summary (tea_solve.f90:201-203, 356-371), the field-summary block and the
 * known-answer verdict (field_summary.f90:41-104) with the reference's
 * Fortran edit descriptors, and the per-step wall clock. The reference's
 * input echo, decomposition and profiler tables are not written. */
int tl_run_deck_out(const char *deck_path, const char *log_path, const char *tea_out_path, int mode,
                    int max_steps);

/* Timed run for benchmarks: builds the deck's chunk, runs `warmup` then
 * `steps` timesteps with the given mode and reports wall seconds of the timed
 * steps, iterations done in them, and the average device duration (ms) of the
 * two CG iteration kernels measured with HIP events on the library's stream
 * (-1 when the mode/solver has none). */
typedef struct {
    double seconds;
    long iterations;        /* outer iterations summed over the timed steps */
    long inner_iterations;  /* PPCG inner steps summed over the timed steps */
    double k1_ms, k2_ms;    /* average durations of the two CG iteration kernels */
    long k_samples;
    double final_summary[5];/* vol, mass, density, ie, temp after the last step */
} tl_bench_result;
int tl_bench_deck(const char *deck_path, int mode, int warmup_steps, int steps, tl_bench_result *out);

/* Refined solve of A u = u0 for the current timestep (call after
 * tl_common_init; u holds the initial guess, the reference's u = u0), with the
 * chunk's preconditioner. No reference counterpart: the reference solves in
 * FP64 only (tea_solve.f90:37-420).
 *   precision TL_PRECISION_FP64: inner CG with FP64 vectors;
 *             TL_PRECISION_MIXED: inner CG with FP32 vectors and the
 *             conductivity held as 1/(2w) in FP32 (face coefficients formed
 *             from it in FP64), FP64 arithmetic, dot products and scalars.
 * Either way the increment is folded into the FP64 solution and the residual
 * is replaced by the true FP64 residual u0 - A u whenever the recursive
 * residual norm has fallen below delta times its largest value since the last
 * replacement (0 < delta < 1; 0 replaces it only at the tolerance), and the
 * solve ends when ||u0 - A u||_2 <= tol * ||u0 - A u_init||_2 holds for the
 * true residual, after max_iters inner iterations, or when three replacements
 * in a row have not halved the true residual (it has reached its FP64
 * rounding floor above tol: stalled = 1); the last two return TL_OK with
 * converged = 0. The result is left in field TL_F_U. */
enum {
    TL_PRECISION_FP64 = 0,
    TL_PRECISION_MIXED = 1
};
typedef struct {
    double seconds;           /* device time of the whole solve, setup included */
    double initial_residual;  /* ||u0 - A u_init||_2 */
    double rel_residual;      /* final true ||u0 - A u||_2 / initial_residual */
    double k1_ms, k2_ms;      /* average durations of the two inner iteration
                                 kernels (HIP events, first iteration of each
                                 batch), -1 when none was sampled */
    int iterations;           /* inner iterations */
    int updates;              /* true residuals formed, the initial one included */
    int converged;
    int stalled;              /* stopped at the rounding floor above tol */
    int k_samples;
    int inner_iterations;     /* PPCG inner steps (tl_solve_refined_ppcg) */
} tl_refine_result;
int tl_solve_refined(tl_chunk *c, int precision, double tol, int max_iters, double delta,
                     tl_refine_result *out);
/* The same refined solve with PPCG (tea_solve.f90:231-267 recurrence, the
 * outer r.z formed with the z of the outer step as the reference does):
 * `presteps` CG iterations (2 .. 1024) give the eigenvalue bounds of the
 * Chebyshev smoother (tea_leaf_cheby.f90:96-172, widened by 0.95 / 1.05), then
 * PPCG outer steps of `inner_steps` (1 .. 256) smoothing steps each, all in the
 * storage precision of the inner fields, with the same residual replacement
 * and stopping rules. iterations counts CG + PPCG outer steps;
 * inner_iterations the smoothing steps. */
int tl_solve_refined_ppcg(tl_chunk *c, int precision, double tol, int max_iters, double delta,
                          int presteps, int inner_steps, tl_refine_result *out);

/* Adjoint of the implicit timesteps with respect to the per-cell density. No
 * reference counterpart (the reference has no derivatives); the forward model
 * is its timestep with zero-flux walls: from the chunk's energy1 e^0 and
 * density rho, nsteps times { tl_common_init(coefficient, dt), solve
 * A(rho) u = u0 (tl_solve_refined, FP64, true relative residual tol),
 * tl_finalise }, and the loss is L = sum over the interior of (u - u_obs)^2
 * after the last step.
 *
 * tl_forward_loss runs that forward model and returns L.
 * tl_adjoint_density runs it keeping u every checkpoint_every steps, then the
 * backward pass: each segment is recomputed from its checkpoint, the adjoint
 * system A lambda = dL/du is solved with the same solver, dL/dK at the faces
 * is formed from -(lambda_a - lambda_b)(u_a - u_b) and chained through the
 * harmonic-mean face coefficient to rho, plus the term of u^0 = rho e^0.
 * A solve that stalls at its rounding floor above tol is accepted (counted in
 * stalled_solves); one that exhausts max_iters is TL_ERR_SOLVER.
 * u_obs and grad are host arrays in the layout of tl_field_get (halo cells
 * of grad are zero). On return energy1 holds the final forward energy;
 * u, u0, p, r, z are overwritten. */
typedef struct {
    double loss;
    double forward_s;           /* device time of the forward timesteps */
    double adjoint_s;           /* loss, recompute, adjoint solves, gradient kernels */
    double work_bytes;          /* device memory of checkpoints, recompute buffers
                                   and adjoint vectors */
    double chunk_bytes;         /* device memory of the chunk's own fields */
    long forward_iterations;
    long recompute_iterations;
    long adjoint_iterations;
    int checkpoints;
    int stalled_solves;         /* solves stopped at their FP64 rounding floor
                                   above tol (accepted) */
} tl_adjoint_result;
int tl_forward_loss(tl_chunk *c, int coefficient, double dt, int nsteps, double tol, int max_iters,
                    const double *u_obs, double *loss);
/* Progress of long calls (no reference counterpart): once set, tl_forward_loss
 * and tl_adjoint_density call fn(user, phase, step, iterations) after every
 * solve, with phase "forward", "recompute" or "adjoint", the timestep and the
 * solve's iterations. fn may be null to switch it off. */
typedef void (*tl_progress_fn)(void *user, const char *phase, int step, int iterations);
int tl_set_progress(tl_chunk *c, tl_progress_fn fn, void *user);
int tl_adjoint_density(tl_chunk *c, int coefficient, double dt, int nsteps, double tol, int max_iters,
                       int checkpoint_every, const double *u_obs, double *grad, tl_adjoint_result *out);

#ifdef __cplusplus
}
#endif
#endif
