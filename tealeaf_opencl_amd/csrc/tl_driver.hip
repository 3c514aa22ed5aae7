// Native driver: start.f90:24-115, diffuse.f90:43-101, tea_solve.f90:37-521
// and field_summary.f90:27-110 restated over the chunk operations.
//
// Two modes share the same host logic:
//  * drop-in: every step calls the reference-granularity operations and does
//    the scalar algebra on the host, exactly as the Fortran does through the
//    entry points (one blocking 8-byte read per dot product);
//  * native: the scalars live on the device (SolveState), each solver
//    iteration is a fixed sequence of fused kernels whose last reduction
//    finishes the iteration's algebra, and batches of iterations are captured
//    once into hipGraphs and replayed. The host only looks at the state after
//    each batch (the next batch is already queued), at the CG -> Chebyshev /
//    PPCG switch and at the end of the solve.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

#include "../../include/tealeaf_mi355x.h"
#include "tl_chunk.h"
#include "tl_deck.h"
#include "tl_eigen.h"
#include "tl_kernels.h"

using namespace tl;

int tl_cg_init_device(tl_chunk *c);

namespace {

struct Save {  // SAVE'd variables of tea_solve.f90:54-57
    int first = 1;
    double eigmin = 0, eigmax = 0, theta = 0, cn = 0;
};

struct Stats {
    long iterations = 0;
    long inner = 0;
};

const double kat_temp[7] = {0.0, 157.550841832793, 116.067951160930, 95.4865103390698,
                            166.838315378708, 116.482111627676, 103.88639125996923};

#define RC(x)                          \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != TL_OK) return rc_;  \
    } while (0)

// ---------------------------------------------------------------------------
// tea.out: the lines the reference writes to unit g_out around the solve, with
// its Fortran edit descriptors (field_summary.f90:41-81, tea_solve.f90:
// 201-203, 356-371, timestep.f90:58, diffuse.f90:69). The library is not
// re-entrant (one in-order stream, like the reference), so the open file of
// the current run is a file-scope handle.
// ---------------------------------------------------------------------------
FILE *g_tout = nullptr;

// ---------------------------------------------------------------------------
// profiler_on: the Fortran driver's phase timers (host seconds), accumulated
// where the reference accumulates them (timestep.f90:55, update_halo.f90:39-83,
// field_summary.f90:77, set_field.f90:44, tea_solve.f90:143/352/387,
// global_mpi.f90:48) and written by diffuse.f90:103-157 at the end of the
// run. One chunk and no MPI: the halo exchange, internal halo update, dot
// product (MPI_Allreduce) and visit timers stay at zero. The run in
// progress owns the file-scope pointer (the library is not re-entrant).
// ---------------------------------------------------------------------------
struct FortranProfile {
    double timestep = 0, halo_exchange = 0, halo_update = 0, internal_halo_update = 0, summary = 0, visit = 0,
           tea_init = 0, dot_product = 0, tea_solve = 0, tea_reset = 0, set_field = 0;
};
FortranProfile *g_fprof = nullptr;

double wtime() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// host time of a device phase: the stream is drained first (profiler_on only)
double phase_time(tl_chunk *c)
{
    // an error here is reported by the phase's own launch checks
    if (g_fprof) (void)hipStreamSynchronize(c->s);
    return wtime();
}

std::string right(const std::string &s, int w)
{
    if ((int)s.size() >= w) return (int)s.size() == w ? s : std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Ew.d (no scale factor): [-]0.d1..ddE+ee, or [-]0.d1..dd+eee for |exp| > 99
// Infinity and NaN in a field of width w, as amdflang writes them under E, 1PE
// and F editing: "NaN", "Inf", "-Inf", right-justified, asterisks when
// narrower
std::string fort_nonfinite(double v, int w)
{
    return right(std::isnan(v) ? "NaN" : v < 0 ? "-Inf" : "Inf", w);
}

std::string fort_E(double v, int w, int d)
{
    if (!std::isfinite(v)) return fort_nonfinite(v, w);
    char buf[128];
    std::string mant, sign = (v < 0 || (v == 0 && std::signbit(v))) ? "-" : "";
    int ex = 0;
    if (v == 0.0) {
        mant = std::string(d, '0');
    } else {
        snprintf(buf, sizeof buf, "%.*e", d - 1, fabs(v));  // m.mmm e+xx
        std::string t(buf);
        const size_t epos = t.find('e');
        if (epos == std::string::npos || epos < 2) return std::string(w, '*');
        mant = t.substr(0, 1) + t.substr(2, epos - 2);
        ex = atoi(t.c_str() + epos + 1) + 1;
    }
    char eb[16];
    if (ex >= -99 && ex <= 99) snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    else snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    return right(sign + "0." + mant + eb, w);
}

// 1PEw.d: [-]d.dd..E+ee, or [-]d.dd..+eee for |exp| > 99
std::string fort_1PE(double v, int w, int d)
{
    if (!std::isfinite(v)) return fort_nonfinite(v, w);
    char buf[128];
    snprintf(buf, sizeof buf, "%.*E", d, v);
    std::string t(buf);
    const size_t epos = t.find('E');
    if (epos == std::string::npos) return std::string(w, '*');
    const int ex = atoi(t.c_str() + epos + 1);
    char eb[16];
    if (ex >= -99 && ex <= 99) snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    else snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    return right(t.substr(0, epos) + eb, w);
}

std::string fort_F(double v, int w, int d)
{
    if (!std::isfinite(v)) return fort_nonfinite(v, w);
    char buf[64];
    snprintf(buf, sizeof buf, "%.*f", d, v);
    return right(buf, w);
}

std::string fort_I(long v, int w) { return right(std::to_string(v), w); }

// list-directed REAL(8) item (the leading blank of the record is written by
// the caller). Its form is the Fortran compiler's choice (the reference's
// output differs between compilers); this one is a fixed E25.16 style.
std::string fort_list(double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%25.16E", v);
    return buf;
}

void tout(const std::string &line)
{
    if (g_tout) fprintf(g_tout, "%s\n", line.c_str());
}

// tea_solve.f90:356-371
void tout_solve(double conduction, bool check, double exact, int iterations, bool ppcg, int inner)
{
    tout("Conduction error " + fort_E(conduction, 14, 7));
    if (check) tout("EXACT error calculated as" + fort_E(exact, 14, 7));
    tout("Iteration count " + fort_I(iterations, 8));
    if (ppcg) tout("PPCG Iteration count" + fort_I(inner, 8) + " (Total " + fort_I(inner + iterations, 8) + ")");
}

// tea_solve.f90:201-203
void tout_switch(int n, double rro, double eigmin, double eigmax, double cn, double old_error)
{
    tout("Switching after " + fort_I(n, 3) + " CG its, error " + fort_E(rro, 15, 7));
    tout("Eigen min" + fort_E(eigmin, 14, 6) + " Eigen max" + fort_E(eigmax, 14, 6) + " Condition number" +
         fort_F(cn, 14, 6) + " Error" + fort_E(old_error, 14, 6));
}

// tea_solve.f90:389-391 (profiler_on): seconds of the whole tea_leaf call and
// the loop counter n at exit
void tout_solve_time(double secs, int n)
{
    char buf[128];
    snprintf(buf, sizeof buf, "%16s%16.10f%7s%7d%16s%16.10f", "Solve Time", secs, "Its", n, "Time Per It",
             n ? secs / n : 0.0);
    tout(buf);
}

// diffuse.f90:103-157 (one task: each total is its own maximum)
void tout_profile(const FortranProfile &p, double wall_clock)
{
    const double total = p.timestep + p.halo_exchange + p.summary + p.visit + p.tea_init + p.set_field +
                         p.tea_solve + p.tea_reset + p.dot_product + p.halo_update + p.internal_halo_update;
    auto line = [&](const char *label, double t) {
        char buf[128];
        snprintf(buf, sizeof buf, "%-23s%s%s", label, fort_F(t, 16, 4).c_str(),
                 fort_F(100.0 * (t / wall_clock), 16, 4).c_str());
        tout(buf);
    };
    tout("");
    tout("Profiler Output                 Time            Percentage");
    line("Timestep              :", p.timestep);
    line("MPI Halo Exchange     :", p.halo_exchange);
    line("Self Halo Update      :", p.halo_update);
    line("Internal Halo Update  :", p.internal_halo_update);
    line("Summary               :", p.summary);
    line("Visit                 :", p.visit);
    line("Tea Init              :", p.tea_init);
    line("Dot Product           :", p.dot_product);
    line("Tea Solve             :", p.tea_solve);
    line("Tea Reset             :", p.tea_reset);
    line("Set Field             :", p.set_field);
    line("Total                 :", total);
    char buf[128];
    snprintf(buf, sizeof buf, "%-23s%s%s", "The Rest              :", fort_F(wall_clock - total, 16, 4).c_str(),
             fort_F(100.0 * (wall_clock - total) / wall_clock, 16, 4).c_str());
    tout(buf);
}

// update_halo (reflective walls) with its time in the Self Halo Update timer
int halo_timed(tl_chunk *c, const int *nb, const int *fields, int depth)
{
    const double t0 = phase_time(c);
    const int rc = tl_update_halo(c, nb, fields, depth);
    if (g_fprof) g_fprof->halo_update += phase_time(c) - t0;
    return rc;
}

void emit_summary(FILE *lg, tl_chunk *c, int step, double time, bool complete, const Config &cfg, double *out5)
{
    double vol, mass, ie, temp;
    const double t_sum = wtime();
    if (tl_field_summary(c, &vol, &mass, &ie, &temp) != TL_OK) return;
    if (g_fprof) g_fprof->summary += wtime() - t_sum;
    if (out5) {
        out5[0] = vol; out5[1] = mass; out5[2] = mass / vol; out5[3] = ie; out5[4] = temp;
    }
    if (lg) fprintf(lg, "SUMMARY %d %.17e %.17e %.17e %.17e %.17e\n", step, vol, mass, mass / vol, ie, temp);
    // field_summary.f90:41-81
    tout("");
    tout(" Time " + fort_list(time));
    tout(std::string(13, ' ') + right("Volume", 26) + right("Mass", 26) + right("Density", 26) + right("Energy", 26) +
         right("U", 26));
    tout(" step:" + fort_I(step, 7) + fort_E(vol, 26, 17) + fort_E(mass, 26, 17) + fort_E(mass / vol, 26, 17) +
         fort_E(ie, 26, 17) + fort_E(temp, 26, 17));
    tout("");
    if (complete && cfg.test_problem >= 1 && cfg.test_problem <= 6) {
        const double qa = fabs((100.0 * (temp / kat_temp[cfg.test_problem])) - 100.0);
        if (lg) fprintf(lg, "QA %d %.7e %s\n", cfg.test_problem, qa, qa < 0.001 ? "PASSED" : "FAILED");
        // field_summary.f90:84-104 (the reference's wording, typo included)
        tout("Test problem" + fort_I(cfg.test_problem, 4) + " is within" + fort_E(qa, 16, 7) +
             "% of the expected solution");
        tout(qa < 0.001 ? " This test is considered PASSED" : " This is test is considered NOT PASSED");
    }
}

// tea_solve.f90:374-391: finalise and the energy halo; with profiler_on the
// reset timer (halo time left out, as there) and the Solve Time line
int reset_phase(tl_chunk *c, const Config &cfg, double t_total, int n_exit)
{
    const int nb[4] = {-1, -1, -1, -1};
    const int fields[7] = {0, 0, 1, 0, 0, 0, 0};
    const double t_reset = phase_time(c), h0 = g_fprof ? g_fprof->halo_update : 0.0;
    RC(tl_finalise(c));
    if (cfg.reflective_boundary) RC(halo_timed(c, nb, fields, 1));
    if (g_fprof) {
        g_fprof->tea_reset += phase_time(c) - t_reset - (g_fprof->halo_update - h0);
        tout_solve_time(wtime() - t_total, n_exit);
    }
    return TL_OK;
}

// ---------------------------------------------------------------------------
// drop-in mode: tea_solve.f90:37-420 over the reference-granularity calls
// ---------------------------------------------------------------------------
int solve_dropin(tl_chunk *c, const Config &cfg, Save &sv, double dt, int step, FILE *lg, Stats *stats)
{
    const int max_iters = cfg.max_iters;
    const bool use_cg = cfg.solver == SOLVER_CG, use_ch = cfg.solver == SOLVER_CHEBY;
    const bool use_ppcg = cfg.solver == SOLVER_PPCG, use_jac = cfg.solver == SOLVER_JACOBI;
    const int nb[4] = {-1, -1, -1, -1};
    std::vector<double> cg_alphas(max_iters + 2, 0.0), cg_betas(max_iters + 2, 0.0);
    std::vector<double> ch_alphas, ch_betas;
    double old_error, error = 0.0, exact_error = 0.0, initial_residual;
    double rro = 0.0, pw, rrn, alpha, beta;
    int est_itc = 0, cheby_calc_steps = 0, max_cheby_iters = 0, ppcg_inner_iters = 0;
    int cg_calc_steps = 0, switch_n = 0;
    bool ch_switch_check = false;
    double rx, ry;

    const double t_total = wtime();
    const double t_init = phase_time(c), h_init = g_fprof ? g_fprof->halo_update : 0.0;
    int fields[7] = {1, 0, 1, 0, 0, 0, 0};
    if (cfg.reflective_boundary) RC(halo_timed(c, nb, fields, cfg.halo_depth));
    const int zb[4] = {1, 1, 1, 1};
    RC(tl_common_init(c, cfg.coefficient, dt, &rx, &ry, zb, cfg.reflective_boundary));
    RC(tl_calc_residual(c));
    RC(tl_calc_2norm(c, 1, &initial_residual));
    old_error = initial_residual;
    initial_residual = sqrt(initial_residual);
    if (lg && cfg.verbose_on) fprintf(lg, "INITRES %d %.17e\n", step, initial_residual);
    if (use_cg || use_ch || use_ppcg) RC(tl_cg_init(c, &rro));
    if (g_fprof) g_fprof->tea_init += phase_time(c) - t_init - (g_fprof->halo_update - h_init);
    const double t_solve = phase_time(c);

    int n;
    for (n = 1; n <= max_iters; n++) {
        if (!ch_switch_check) {
            if (cheby_calc_steps > 0) ch_switch_check = true;
            else if (!sv.first && use_ppcg && n > 1) ch_switch_check = true;
            else if (fabs(old_error) <= cfg.tl_ch_cg_epslim && n >= cfg.tl_ch_cg_presteps) ch_switch_check = true;
            else ch_switch_check = false;
        }
        if ((use_ch || use_ppcg) && ch_switch_check) {
            if (cheby_calc_steps == 0) {
                max_cheby_iters = max_iters - n + 2;
                if (sv.first) {
                    const int info = calc_eigenvalues(&cg_alphas[1], &cg_betas[1], n - 1, &sv.eigmin, &sv.eigmax);
                    sv.first = 0;
                    if (info != 0) {
                        set_error("tea_leaf: Error in calculating eigenvalues");
                        return TL_ERR_SOLVER;
                    }
                    sv.eigmin = sv.eigmin * 0.95;
                    sv.eigmax = sv.eigmax * 1.05;
                }
                calc_ch_coefs(ch_alphas, ch_betas, sv.eigmin, sv.eigmax, &sv.theta,
                              use_ch ? max_cheby_iters : cfg.tl_ppcg_inner_steps);
                sv.cn = sv.eigmax / sv.eigmin;
                switch_n = n;
                if (lg)
                    fprintf(lg, "SWITCH %d %d %.17e %.17e %.17e %.17e %.17e\n", step, n, rro, sv.eigmin, sv.eigmax,
                            sv.cn, old_error);
                tout_switch(n, rro, sv.eigmin, sv.eigmax, sv.cn, old_error);
            }
            if (use_ch) {
                if (cheby_calc_steps == 0) {
                    double bb;
                    RC(tl_calc_2norm(c, 0, &bb));
                    RC(tl_cheby_init(c, ch_alphas.data(), ch_betas.data(), max_cheby_iters, rx, ry, sv.theta));
                    RC(tl_cheby_iterate(c, 1));
                    RC(tl_calc_2norm(c, 1, &old_error));
                    const double it_alpha = cfg.eps * bb / (4.0 * old_error);
                    const double gamm = (sqrt(sv.cn) - 1.0) / (sqrt(sv.cn) + 1.0);
                    est_itc = (int)lround(log(it_alpha) / (2.0 * log(gamm)));
                    if (lg) fprintf(lg, "ESTITC %d %d\n", step, est_itc);
                    cheby_calc_steps = 1;
                } else {
                    RC(tl_cheby_iterate(c, cheby_calc_steps));
                    if (n >= est_itc && n % 10 == 0) RC(tl_calc_2norm(c, 1, &error));
                }
            } else {
                if (cheby_calc_steps == 0)
                    RC(tl_ppcg_init(c, ch_alphas.data(), ch_betas.data(), sv.theta, cfg.tl_ppcg_inner_steps));
                RC(tl_cg_calc_w(c, &pw));
                alpha = rro / pw;
                RC(tl_cg_calc_ur(c, alpha, &rrn));
                RC(tl_ppcg_init_sd(c));
                for (int cur = 1; cur <= cfg.tl_ppcg_inner_steps; cur += cfg.halo_depth) {
                    int inner_step = cur;
                    for (int be = cfg.halo_depth - 1; be >= 0; be--) {
                        RC(tl_ppcg_inner(c, inner_step, be, nb));
                        inner_step++;
                        if (inner_step > cfg.tl_ppcg_inner_steps) break;
                    }
                }
                ppcg_inner_iters += cfg.tl_ppcg_inner_steps;
                if (cfg.ppcg_fresh_z) RC(precondition_r(c));
                RC(tl_calc_2norm(c, 2, &rrn));
                beta = rrn / rro;
                RC(tl_cg_calc_p(c, beta));
                error = rrn;
                rro = rrn;
            }
            cheby_calc_steps++;
        } else if (use_cg || use_ch || use_ppcg) {
            cg_calc_steps++;
            RC(tl_cg_calc_w(c, &pw));
            alpha = rro / pw;
            cg_alphas[n] = alpha;
            RC(tl_cg_calc_ur(c, alpha, &rrn));
            beta = rrn / rro;
            cg_betas[n] = beta;
            RC(tl_cg_calc_p(c, beta));
            error = rrn;
            rro = rrn;
        } else if (use_jac) {
            RC(tl_jacobi_solve(c, &error));
        }
        error = sqrt(error);
        if (lg && cfg.verbose_on) fprintf(lg, "RESID %d %d %.17e\n", step, n, error);
        if (fabs(error) < cfg.eps * initial_residual) break;
        old_error = error;
    }
    if (cfg.tl_check_result) {
        RC(tl_calc_residual(c));
        RC(tl_calc_2norm(c, 1, &exact_error));
        exact_error = sqrt(exact_error);
    }
    if (g_fprof) g_fprof->tea_solve += phase_time(c) - t_solve;
    if (lg)
        fprintf(lg, "SOLVE %d %d %.17e %.17e %d %d %d %d %d\n", step, n - 1, error / initial_residual,
                cfg.tl_check_result ? exact_error / initial_residual : NAN, ppcg_inner_iters, cg_calc_steps,
                cheby_calc_steps, switch_n, est_itc);
    tout_solve(error / initial_residual, cfg.tl_check_result, exact_error / initial_residual, n - 1,
               cfg.solver == SOLVER_PPCG, ppcg_inner_iters);
    if (stats) {
        stats->iterations += n > max_iters ? max_iters : n;
        stats->inner += ppcg_inner_iters;
    }
    return reset_phase(c, cfg, t_total, n);
}

// ---------------------------------------------------------------------------
// native mode
// ---------------------------------------------------------------------------
int state_download(tl_chunk *c, SolveState *h)
{
    hipError_t e = hipMemcpyAsync(&c->h_st[0], c->st, sizeof(SolveState), hipMemcpyDeviceToHost, c->s);
    if (e != hipSuccess) return hip_fail("state download", e);
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("state download", e);
    *h = c->h_st[0];
    return TL_OK;
}

int state_upload(tl_chunk *c, const SolveState &h)
{
    hipError_t e = hipStreamSynchronize(c->s);
    if (e != hipSuccess) return hip_fail("state upload", e);
    c->h_st[0] = h;
    e = hipMemcpyAsync(c->st, &c->h_st[0], sizeof(SolveState), hipMemcpyHostToDevice, c->s);
    if (e != hipSuccess) return hip_fail("state upload", e);
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("state upload", e);
    return TL_OK;
}

// One launch of the device-resident driver. With prof (profiler_on: the
// batches are then launched eagerly instead of replayed) it is bracketed by
// events and its device time goes to the chunk's per-kernel table under the
// HIP kernel's name, its finisher's (fin) under k_finish, as the reference
// times every OpenCL kernel it enqueues (ocl_errors.cpp:64-121).
template <typename F>
void launch_timed(tl_chunk *c, bool prof, const char *name, bool fin, F &&launch)
{
    if (!prof) {
        launch(nullptr);
        return;
    }
    prof_begin(c);
    launch(fin ? c->pe_mid : nullptr);
    if (fin) prof_end_split(c, name, "k_finish");
    else prof_end(c, name);
}

// One CG iteration of the device-resident driver: the fused search-direction
// update + matvec + p.w, then the fused u/r update + preconditioner + r.z.
// With `ev`, ev[0..1] bracket the first iteration kernel and ev[2..3] the
// second (their finishers are outside the brackets). parity 1 / 2: the
// iteration is known to be odd / even (the update kernel's form compiled for
// that parity runs); 0: unknown (the form that reads the parity from the
// state).
void cg_iteration(tl_chunk *c, hipEvent_t *ev, bool prof = false, int parity = 0)
{
    const Ext in = ext_interior(c);
    const double *src = (c->prec == PREC_NONE) ? c->f2[F_R] : c->f2[F_Z];
    if (ev) (void)hipEventRecord(ev[0], c->s);
    launch_timed(c, prof, "k_fused_cg_w", true, [&](hipEvent_t mid) {
        launch_fused_cg_w(c->s, c->g, in, c->cap, src, c->f2[F_P], c->f2[F_P2], c->f2[F_COND], c->faces,
                          red_args(c, RED_CG_PW, true), ev ? ev[1] : mid, stream_z_pays(c));
    });
    if (ev) (void)hipEventRecord(ev[2], c->s);
    launch_timed(c, prof, "k_fused_cg_ur", true, [&](hipEvent_t mid) {
        launch_fused_cg_ur(c->s, c->g, in, c->cap, c->prec, 0, c->f2[F_U], c->f2[F_P], c->f2[F_P2], c->f2[F_R],
                           c->f2[F_R], c->f2[F_Z], c->f2[F_SD], c->f2[F_COND], c->faces,
                           red_args(c, RED_CG_RRN, true), ev ? ev[3] : mid, parity);
    });
}

// profiler_on: the iterations of a phase launched one `body` at a time, each
// followed by a read of the device state, until it halts (so every kernel's
// call count is the number of iterations that ran)
int run_eager(tl_chunk *c, const std::function<void(long)> &body, long max_calls)
{
    SolveState h;
    for (long k = 0;; ++k) {
        body(k);
        RC(state_download(c, &h));
        if (h.halt) return TL_OK;
        if (k > max_calls) {
            set_error("solver state did not halt after %ld eager iterations", k);
            return TL_ERR_SOLVER;
        }
    }
}

// Replays `exec` until the device state says halt; batch i+1 is queued before
// the state of batch i is examined. With `timed`, each batch starts with one
// eagerly launched CG iteration bracketed by events (graph-recorded events
// carry no timestamps); the events of a batch are read once it has finished,
// while the other slot's set is in flight.
int run_batches(tl_chunk *c, hipGraphExec_t exec, bool timed, long max_batches)
{
    hipError_t e;
    int slot = 0;
    auto enqueue = [&](int sl) -> hipError_t {
        if (timed) {
            cg_iteration(c, c->ev_t[sl], false, 1);
            cg_iteration(c, c->ev_t[sl] + 4, false, 2);
        }
        hipError_t ee = hipGraphLaunch(exec, c->s);
        if (ee != hipSuccess) return ee;
        ee = hipMemcpyAsync(&c->h_st[sl], c->st, sizeof(SolveState), hipMemcpyDeviceToHost, c->s);
        if (ee != hipSuccess) return ee;
        return hipEventRecord(c->ev_batch[sl], c->s);
    };
    if ((e = enqueue(0)) != hipSuccess) return hip_fail("hipGraphLaunch", e);
    for (long guard = 0;; ++guard) {
        if ((e = enqueue(1 - slot)) != hipSuccess) return hip_fail("hipGraphLaunch", e);
        if ((e = hipEventSynchronize(c->ev_batch[slot])) != hipSuccess) return hip_fail("batch wait", e);
        const SolveState &h = c->h_st[slot];
        if (h.fault) {
            (void)hipStreamSynchronize(c->s);
            set_error("an iteration kernel captured for one iteration parity ran on the other");
            return TL_ERR_SOLVER;
        }
        if (timed) {
            // every examined batch follows a batch that did not halt, so its
            // eager first iterations ran (unless the pair's first one halted:
            // then the second one's kernels returned at once and the sample is
            // left out); the sample is the mean over the odd and the even one
            float a[4] = {0.f, 0.f, 0.f, 0.f};
            bool ok = true;
            for (int q = 0; q < 4; ++q)
                ok = ok && hipEventElapsedTime(&a[q], c->ev_t[slot][2 * q], c->ev_t[slot][2 * q + 1]) == hipSuccess;
            if (ok && !h.halt) {
                c->t_k1_ms += 0.5 * (a[0] + a[2]);
                c->t_k2_ms += 0.5 * (a[1] + a[3]);
                c->t_samples += 1;
            }
        }
        if (h.halt) break;
        if (guard > max_batches) {
            (void)hipStreamSynchronize(c->s);
            set_error("solver state did not halt after %ld batches", guard);
            return TL_ERR_SOLVER;
        }
        slot = 1 - slot;
    }
    if ((e = hipStreamSynchronize(c->s)) != hipSuccess) return hip_fail("batch drain", e);
    (void)hipGetLastError();
    return TL_OK;
}

// 16 CG iterations per replay; with `timed`, 14 plus the eagerly launched
// odd and even ones of each batch. Every batch starts on an odd iteration
// (the CG phase starts at n = 1), so the update kernel is captured in its
// odd and even forms alternately.
int cg_graph(tl_chunk *c, bool timed)
{
    const int B = timed ? 14 : 16;
    const int key = c->prec * 10 + (timed ? 1 : 0);
    return capture_graph(c, c->g_cg, key, [&]() {
        for (int b = 0; b < B; ++b) cg_iteration(c, nullptr, false, (b & 1) ? 2 : 1);
    });
}

// 16 Chebyshev iterations per replay: one per launch, or (pairs) two per
// launch when the first iteration of the replay is odd, so that the residual
// norm of tea_solve.f90:223 (iterations divisible by 10) always falls on the
// second iteration of a pair. Pairs flip u and p between their two buffers
// at every launch (8 flips per replay).
// Chebyshev launch k of a phase: one iteration, or (pairs) two per launch
// with u and p flipping between their two buffers at every launch
void cheby_launch(tl_chunk *c, bool pairs, long k, bool prof)
{
    const Ext in = ext_interior(c);
    if (pairs) {
        const bool odd = k & 1;
        launch_timed(c, prof, "k_fused_cheby2", true, [&](hipEvent_t mid) {
            launch_fused_cheby2(c->s, c->g, in, c->cap, c->prec, c->f2[F_U0], c->f2[odd ? F_U2 : F_U],
                                c->f2[odd ? F_U : F_U2], c->f2[odd ? F_P2 : F_P], c->f2[odd ? F_P : F_P2],
                                c->f2[F_KX], c->f2[F_KY], c->ch_a, c->ch_b, red_args(c, RED_CHEBY2, true), mid);
        });
        return;
    }
    launch_timed(c, prof, "k_fused_cheby", true, [&](hipEvent_t mid) {
        launch_fused_cheby(c->s, c->g, in, c->cap, c->prec, c->f2[F_U0], c->f2[F_U], c->f2[F_U2], c->f2[F_P],
                           c->f2[F_COND], c->faces, c->ch_a, c->ch_b, red_args(c, RED_CHEBY, true), mid);
    });
}

int cheby_graph(tl_chunk *c, bool pairs)
{
    const int B = 16;
    return capture_graph(c, c->g_cheby, c->prec * 4 + (pairs ? 2 : 1), [&]() {
        for (int b = 0; b < (pairs ? B / 2 : B); ++b) cheby_launch(c, pairs, b, false);
    });
}

// PPCG outer step (tea_solve.f90:231-267) and its inner steps
// (tea_leaf_run_ppcg_inner_steps, :422-475). With two inner steps per launch
// the pairs read r and sd from one buffer and write the other; the outer
// step's update writes r to the second buffer when the number of pairs is
// odd, so that every outer step starts and ends with r in F_R. An odd inner
// count ends with one single-step launch (in place for r).
void ppcg_outer(tl_chunk *c, int inner, bool prof)
{
    const Ext in = ext_interior(c);
    const double *src = (c->prec == PREC_NONE) ? c->f2[F_R] : c->f2[F_Z];
    const int per = c->ppcg_steps_per_launch;
    const int pairs = per == 2 ? inner / 2 : 0;
    const int last = c->ppcg_fresh_z ? 2 : 1;  // the last step's r.z: fresh or stale z
    launch_timed(c, prof, "k_fused_cg_w", true, [&](hipEvent_t mid) {
        launch_fused_cg_w(c->s, c->g, in, c->cap, src, c->f2[F_P], c->f2[F_P2], c->f2[F_COND], c->faces,
                          red_args(c, RED_PPCG_PW, true), mid, false);
    });
    double *r_cur = (pairs & 1) ? c->f2[F_R2] : c->f2[F_R];
    double *r_oth = (pairs & 1) ? c->f2[F_R] : c->f2[F_R2];
    launch_timed(c, prof, "k_fused_cg_ur_ppcg", false, [&](hipEvent_t) {
        launch_fused_cg_ur(c->s, c->g, in, c->cap, c->prec, 1, c->f2[F_U], c->f2[F_P], c->f2[F_P2], c->f2[F_R],
                           r_cur, c->f2[F_Z], c->f2[F_SD], c->f2[F_COND], c->faces, red_args(c, RED_STORE, true));
    });
    double *sd_cur = c->f2[F_SD], *sd_oth = c->f2[F_SD2];
    int i = 1;
    for (int k = 0; k < pairs; ++k, i += 2) {
        const int lst = i + 1 == inner ? last : 0;
        launch_timed(c, prof, "k_fused_ppcg_inner2", lst != 0, [&](hipEvent_t mid) {
            launch_fused_ppcg_inner2(c->s, c->g, in, c->cap, c->prec, lst, c->f2[F_U], r_cur, r_oth, sd_cur, sd_oth,
                                     c->f2[F_KX], c->f2[F_KY], c->f2[F_Z], c->ch_a, c->ch_b, i - 1,
                                     red_args(c, RED_PPCG_RRN, true), mid);
        });
        std::swap(r_cur, r_oth);
        std::swap(sd_cur, sd_oth);
    }
    for (; i <= inner; ++i) {
        const int lst = i == inner ? last : 0;
        launch_timed(c, prof, "k_fused_ppcg_inner", lst != 0, [&](hipEvent_t mid) {
            launch_fused_ppcg_inner(c->s, c->g, in, c->cap, c->prec, lst, c->f2[F_U], r_cur, sd_cur, sd_oth,
                                    c->f2[F_COND], c->faces, c->f2[F_Z], c->ch_a, c->ch_b, i - 1,
                                    red_args(c, RED_PPCG_RRN, true), mid);
        });
        std::swap(sd_cur, sd_oth);
    }
}

int ppcg_graph(tl_chunk *c, int inner)
{
    const int B = std::max(1, 32 / (inner + 2));
    const int per = c->ppcg_steps_per_launch;
    const int last = c->ppcg_fresh_z ? 2 : 1;
    return capture_graph(c, c->g_ppcg, ((c->prec * 100000 + inner) * 4 + per) * 2 + (last - 1), [&]() {
        for (int b = 0; b < B; ++b) ppcg_outer(c, inner, false);
    });
}

// PPCG inner steps first_step .. first_step+nsteps-1 (1-based, like the
// inner_step of tl_ppcg_inner) of the smoother with the
// fused kernels of the native driver (one or two steps per launch, as
// tuned), on the chunk's u, r and sd; results left in u, r, sd like the same
// number of tl_ppcg_inner calls over the interior (tea_leaf_ppcg_cl.cl:62-155).
extern "C" int tl_ppcg_inner_fused(tl_chunk *c, int first_step, int nsteps)
{
    if (!c || first_step < 1 || nsteps < 1 || first_step - 1 + nsteps > c->n_coefs) {
        set_error("tl_ppcg_inner_fused: steps %d..%d outside the %d coefficients", first_step,
                  first_step + nsteps - 1, c ? c->n_coefs : 0);
        return TL_ERR_ARG;
    }
    SolveState h;
    memset(&h, 0, sizeof h);
    RC(state_upload(c, h));  // halt = 0 for the kernels' state test
    const Ext in = ext_interior(c);
    const RedArgs ra = red_args(c, RED_STORE, true);
    double *r_cur = c->f2[F_R], *r_oth = c->f2[F_R2], *sd_cur = c->f2[F_SD], *sd_oth = c->f2[F_SD2];
    int i = 0;
    for (; c->ppcg_steps_per_launch == 2 && i + 1 < nsteps; i += 2) {
        launch_fused_ppcg_inner2(c->s, c->g, in, c->cap, c->prec, 0, c->f2[F_U], r_cur, r_oth, sd_cur, sd_oth,
                                 c->f2[F_KX], c->f2[F_KY], c->f2[F_Z], c->ch_a, c->ch_b, first_step - 1 + i, ra);
        std::swap(r_cur, r_oth);
        std::swap(sd_cur, sd_oth);
    }
    for (; i < nsteps; ++i) {
        launch_fused_ppcg_inner(c->s, c->g, in, c->cap, c->prec, 0, c->f2[F_U], r_cur, sd_cur, sd_oth,
                                c->f2[F_COND], c->faces, c->f2[F_Z], c->ch_a, c->ch_b, first_step - 1 + i, ra);
        std::swap(sd_cur, sd_oth);
    }
    if (r_cur != c->f2[F_R]) launch_copy(c->s, c->g, in, c->cap, r_cur, c->f2[F_R]);
    if (sd_cur != c->f2[F_SD]) launch_copy(c->s, c->g, in, c->cap, sd_cur, c->f2[F_SD]);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    return e == hipSuccess ? TL_OK : hip_fail("tl_ppcg_inner_fused", e);
}

void jacobi_launch(tl_chunk *c, bool prof)
{
    launch_timed(c, prof, "k_jacobi", true, [&](hipEvent_t mid) {
        launch_fused_jacobi(c->s, c->g, ext_interior(c), c->cap, c->f2[F_COND], c->faces, c->f2[F_U0], c->f2[F_U],
                            c->f2[F_U2], red_args(c, RED_JACOBI, true), mid);
    });
}

int jacobi_graph(tl_chunk *c)
{
    const int B = 16;
    return capture_graph(c, c->g_jacobi, 1, [&]() {
        for (int b = 0; b < B; ++b) jacobi_launch(c, false);
    });
}

int solve_native(tl_chunk *c, const Config &cfg, Save &sv, double dt, int step, FILE *lg, Stats *stats)
{
    const int nb[4] = {-1, -1, -1, -1};
    const bool use_ch = cfg.solver == SOLVER_CHEBY, use_ppcg = cfg.solver == SOLVER_PPCG;
    const bool use_jac = cfg.solver == SOLVER_JACOBI;
    const double t_total = wtime();
    const double t_init = phase_time(c), h_init = g_fprof ? g_fprof->halo_update : 0.0;
    // profiler_on: every kernel of the phase launched eagerly and timed
    const bool prof = c->profiler_on != 0;
    const int fields[7] = {1, 0, 1, 0, 0, 0, 0};
    if (cfg.reflective_boundary) RC(halo_timed(c, nb, fields, cfg.halo_depth));
    const int zb[4] = {1, 1, 1, 1};
    double rx, ry;
    RC(tl_common_init(c, cfg.coefficient, dt, &rx, &ry, zb, cfg.reflective_boundary));
    RC(ensure_hist_capacity(c, cfg.max_iters + 2));

    SolveState h;
    memset(&h, 0, sizeof h);
    h.eps = cfg.eps;
    h.epslim = cfg.tl_ch_cg_epslim;
    h.max_iters = cfg.max_iters;
    h.presteps = cfg.tl_ch_cg_presteps;
    h.solver = cfg.solver;
    h.first = sv.first;
    h.n = 1;
    RC(state_upload(c, h));

    RC(tl_calc_residual(c));
    launch_timed(c, prof, "k_dot", false, [&](hipEvent_t) {
        launch_dot(c->s, c->g, ext_interior(c), c->cap, c->f2[F_R], c->f2[F_R], red_args(c, RED_INIT_RESID, true));
    });
    if (!use_jac) RC(tl_cg_init_device(c));
    if (cfg.max_iters < 1) {
        RC(state_download(c, &h));
        h.halt = h.done = 1;
        h.exit_n = cfg.max_iters + 1;
        RC(state_upload(c, h));
    }

    const bool timed = c->timing && !c->profiler_on;
    RC(state_download(c, &h));
    if (g_fprof) g_fprof->tea_init += phase_time(c) - t_init - (g_fprof->halo_update - h_init);
    const double t_solve = phase_time(c);
    if (!h.halt) {
        if (use_jac && prof) {
            RC(run_eager(c, [&](long) { jacobi_launch(c, true); }, cfg.max_iters + 4));
        } else if (use_jac) {
            RC(jacobi_graph(c));
            RC(run_batches(c, c->g_jacobi.exec, false, cfg.max_iters / 16 + 4));
        } else {
            if (prof) {
                RC(run_eager(c, [&](long k) { cg_iteration(c, nullptr, true, (k & 1) ? 2 : 1); }, cfg.max_iters + 4));
            } else {
                RC(cg_graph(c, timed));
                RC(run_batches(c, c->g_cg.exec, timed, cfg.max_iters / 15 + 4));
            }
            RC(state_download(c, &h));
            // the fused CG kernels add alpha p into u on even iterations only
            // (two at a time): complete u when the CG phase ended on an odd one
            const int last = h.done ? std::min(h.exit_n, cfg.max_iters) : (h.switch_n > 0 ? h.switch_n - 1 : 0);
            if (last >= 1 && (last & 1))
                launch_timed(c, prof, "k_cg_u_fix", false, [&](hipEvent_t) {
                    launch_cg_u_fix(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U], c->f2[F_P2], c->st);
                });
        }
        RC(state_download(c, &h));
    }
    int est_itc = 0;
    bool cheby_fused = false, cheby_pairs = false;
    if (!h.done && h.switch_n > 0 && (use_ch || use_ppcg)) {
        const int n = h.switch_n;
        const int max_cheby_iters = cfg.max_iters - n + 2;
        if (sv.first) {
            std::vector<double> a(n > 1 ? n - 1 : 1), b(n > 1 ? n - 1 : 1);
            if (n > 1) {
                TL_HIP_TRY(hipMemcpy(a.data(), c->hist_a + 1, sizeof(double) * (n - 1), hipMemcpyDeviceToHost),
                           "CG history copy");
                TL_HIP_TRY(hipMemcpy(b.data(), c->hist_b + 1, sizeof(double) * (n - 1), hipMemcpyDeviceToHost),
                           "CG history copy");
            }
            const int info = calc_eigenvalues(a.data(), b.data(), n - 1, &sv.eigmin, &sv.eigmax);
            sv.first = 0;
            if (info != 0) {
                set_error("tea_leaf: Error in calculating eigenvalues");
                return TL_ERR_SOLVER;
            }
            sv.eigmin = sv.eigmin * 0.95;
            sv.eigmax = sv.eigmax * 1.05;
        }
        std::vector<double> ca, cb;
        calc_ch_coefs(ca, cb, sv.eigmin, sv.eigmax, &sv.theta, use_ch ? max_cheby_iters : cfg.tl_ppcg_inner_steps);
        sv.cn = sv.eigmax / sv.eigmin;
        if (lg)
            fprintf(lg, "SWITCH %d %d %.17e %.17e %.17e %.17e %.17e\n", step, n, h.rro, sv.eigmin, sv.eigmax, sv.cn,
                    h.old_error);
        tout_switch(n, h.rro, sv.eigmin, sv.eigmax, sv.cn, h.old_error);
        if (use_ch) {
            // tea_leaf_cheby_first_step (tea_solve.f90:477-521)
            double bb, err1;
            RC(tl_calc_2norm(c, 0, &bb));
            RC(tl_cheby_init(c, ca.data(), cb.data(), max_cheby_iters, rx, ry, sv.theta));
            RC(tl_cheby_iterate(c, 1));
            RC(tl_calc_2norm(c, 1, &err1));
            const double it_alpha = cfg.eps * bb / (4.0 * err1);
            const double gamm = (sqrt(sv.cn) - 1.0) / (sqrt(sv.cn) + 1.0);
            est_itc = (int)lround(log(it_alpha) / (2.0 * log(gamm)));
            if (lg) fprintf(lg, "ESTITC %d %d\n", step, est_itc);
            // the rest of iteration n: cheby_calc_steps = 1 then +1, and the
            // loop tail on the stale error (tea_solve.f90:269, 322-331)
            RC(state_download(c, &h));
            h.old_error = err1;
            h.cheby_calc_steps = 2;
            h.est_itc = est_itc;
            h.theta = sv.theta;
            h.error = sqrt(h.error);
            TL_HIP_TRY(hipMemcpy(c->hist_e + n, &h.error, sizeof(double), hipMemcpyHostToDevice), "history copy");
            h.halt = 0;
            if (fabs(h.error) < h.eps * h.initial_residual) {
                h.done = h.halt = 1;
                h.exit_n = n;
            } else {
                h.old_error = h.error;
                if (n >= cfg.max_iters) {
                    h.done = h.halt = 1;
                    h.exit_n = cfg.max_iters + 1;
                } else {
                    h.n = n + 1;
                    h.n_fused0 = n + 1;
                    cheby_fused = true;
                }
            }
            RC(state_upload(c, h));
            if (!h.halt) {
                cheby_pairs = c->cheby_steps_per_launch == 2 && (h.n_fused0 & 1);
                if (prof) {
                    const bool pairs = cheby_pairs;
                    RC(run_eager(c, [&](long k) { cheby_launch(c, pairs, k, true); }, cfg.max_iters + 4));
                } else {
                    RC(cheby_graph(c, cheby_pairs));
                    RC(run_batches(c, c->g_cheby.exec, false, cfg.max_iters / 16 + 4));
                }
            }
        } else {
            RC(tl_ppcg_init(c, ca.data(), cb.data(), sv.theta, cfg.tl_ppcg_inner_steps));
            RC(state_download(c, &h));
            h.theta = sv.theta;
            h.halt = 0;
            RC(state_upload(c, h));
            if (prof) {
                RC(run_eager(c, [&](long) { ppcg_outer(c, cfg.tl_ppcg_inner_steps, true); }, cfg.max_iters + 4));
            } else {
                RC(ppcg_graph(c, cfg.tl_ppcg_inner_steps));
                RC(run_batches(c, c->g_ppcg.exec, false, cfg.max_iters + 4));
            }
        }
        RC(state_download(c, &h));
    }
    const int executed = (h.exit_n > cfg.max_iters) ? cfg.max_iters : h.exit_n;
    // launches of the double-buffered Chebyshev kernels (a pair cut short by
    // the exit test still flipped the buffers once)
    const int cheby_launches =
        cheby_fused ? (cheby_pairs ? (executed - h.n_fused0 + 2) / 2 : executed - h.n_fused0 + 1) : 0;
    if ((use_jac && (executed & 1)) || (cheby_launches & 1)) {
        // the last sweep wrote the second buffer
        launch_copy(c->s, c->g, ext_interior(c), c->cap, c->f2[F_U2], c->f2[F_U]);
    }
    double exact = NAN;
    if (cfg.tl_check_result) {
        RC(tl_calc_residual(c));
        RC(tl_calc_2norm(c, 1, &exact));
        exact = sqrt(exact) / h.initial_residual;
    }
    if (g_fprof) g_fprof->tea_solve += phase_time(c) - t_solve;
    const int ppcg_inner = use_ppcg ? h.cheby_calc_steps * cfg.tl_ppcg_inner_steps : 0;
    if (lg && cfg.verbose_on) {
        fprintf(lg, "INITRES %d %.17e\n", step, h.initial_residual);
        std::vector<double> he(executed + 1, 0.0);
        if (executed > 0)
            TL_HIP_TRY(hipMemcpy(he.data() + 1, c->hist_e + 1, sizeof(double) * executed, hipMemcpyDeviceToHost),
                       "residual history copy");
        for (int i = 1; i <= executed; ++i) fprintf(lg, "RESID %d %d %.17e\n", step, i, he[i]);
    }
    if (lg)
        fprintf(lg, "SOLVE %d %d %.17e %.17e %d %d %d %d %d\n", step, h.exit_n - 1, h.error / h.initial_residual,
                exact, ppcg_inner, h.cg_calc_steps, h.cheby_calc_steps, h.switch_n, est_itc);
    tout_solve(h.error / h.initial_residual, cfg.tl_check_result, exact, h.exit_n - 1, use_ppcg, ppcg_inner);
    if (stats) {
        stats->iterations += executed;
        stats->inner += ppcg_inner;
    }
    return reset_phase(c, cfg, t_total, h.exit_n);
}

// One timestep with the refined solve (tl_solve_refined, or
// tl_solve_refined_ppcg for a PPCG deck with tl_ch_cg_presteps CG steps and
// tl_ppcg_inner_steps smoothing steps) in place of the reference solver: the steps of tea_leaf() around the solve are the same
// (tea_solve.f90:70-101 set-up, :352-420 finalise); tl_eps is the target of
// the true relative residual. The SOLVE line reports the inner iterations and
// the final true relative residual in both error columns; REFINE adds the
// number of true residuals formed, convergence and the device seconds.
int solve_refined(tl_chunk *c, const Config &cfg, double dt, int step, FILE *lg, Stats *stats)
{
    const int nb[4] = {-1, -1, -1, -1};
    const double t_total = wtime();
    const double t_init = phase_time(c), h_init = g_fprof ? g_fprof->halo_update : 0.0;
    const int fields[7] = {1, 0, 1, 0, 0, 0, 0};
    if (cfg.reflective_boundary) RC(halo_timed(c, nb, fields, cfg.halo_depth));
    const int zb[4] = {1, 1, 1, 1};
    double rx, ry;
    RC(tl_common_init(c, cfg.coefficient, dt, &rx, &ry, zb, cfg.reflective_boundary));
    if (g_fprof) g_fprof->tea_init += phase_time(c) - t_init - (g_fprof->halo_update - h_init);
    const double t_solve = phase_time(c);
    const int precision = cfg.refine == 2 ? TL_PRECISION_MIXED : TL_PRECISION_FP64;
    const double delta = cfg.refine_delta >= 0.0 ? cfg.refine_delta : (precision == TL_PRECISION_MIXED ? 0.1 : 0.0);
    tl_refine_result rr;
    if (cfg.solver == SOLVER_PPCG)
        RC(tl_solve_refined_ppcg(c, precision, cfg.eps, cfg.max_iters, delta, cfg.tl_ch_cg_presteps,
                                 cfg.tl_ppcg_inner_steps, &rr));
    else
        RC(tl_solve_refined(c, precision, cfg.eps, cfg.max_iters, delta, &rr));
    if (g_fprof) g_fprof->tea_solve += phase_time(c) - t_solve;
    if (lg) {
        fprintf(lg, "SOLVE %d %d %.17e %.17e %d %d %d %d %d\n", step, rr.iterations, rr.rel_residual, rr.rel_residual,
                rr.inner_iterations, rr.iterations, 0, 0, 0);
        fprintf(lg, "REFINE %d %d %d %d %.17e %.9e\n", step, rr.iterations, rr.updates, rr.converged,
                rr.rel_residual, rr.seconds);
    }
    tout_solve(rr.rel_residual, false, 0.0, rr.iterations, cfg.solver == SOLVER_PPCG, rr.inner_iterations);
    if (stats) {
        stats->iterations += rr.iterations;
        stats->inner += rr.inner_iterations;
    }
    return reset_phase(c, cfg, t_total, rr.iterations);
}

// start.f90:24-115
int start(tl_chunk *c, const Config &cfg)
{
    const double dx = (cfg.xmax - cfg.xmin) / (double)(float)cfg.x_cells;
    const double dy = (cfg.ymax - cfg.ymin) / (double)(float)cfg.y_cells;
    RC(tl_initialise_chunk(c, cfg.xmin, cfg.ymin, dx, dy));
    RC(tl_generate_chunk(c, cfg.nstates, cfg.density.data(), cfg.energy.data(), cfg.sxmin.data(), cfg.sxmax.data(),
                         cfg.symin.data(), cfg.symax.data(), cfg.radius.data(), cfg.geometry.data(), 1, 2, 3));
    const int nb[4] = {-1, -1, -1, -1};
    const int fields[7] = {1, 1, 1, 0, 0, 0, 0};
    if (cfg.reflective_boundary) RC(halo_timed(c, nb, fields, cfg.halo_depth));
    const double t_sf = phase_time(c);
    RC(tl_set_field(c));
    if (g_fprof) g_fprof->set_field += phase_time(c) - t_sf;
    return TL_OK;
}

struct RunOut {
    double seconds = 0;
    Stats stats;
    double summary[5] = {0, 0, 0, 0, 0};
    double k1_ms = 0, k2_ms = 0;  // summed over k_samples batches
    long k_samples = 0;
};

// diffuse.f90:43-101; when warmup >= 0 the first `warmup` steps are untimed
// and the rest are timed into `out`
int run(const Config &cfg, FILE *lg, int mode, int max_steps, int warmup, RunOut *out)
{
    if (cfg.sc.halo_depth < 1) {
        set_error("Halo exchange depth unspecified or was too small");
        return TL_ERR_CONFIG;
    }
    tl_chunk *c = nullptr;
    RC(tl_chunk_create(cfg.x_cells, cfg.y_cells, cfg.sc.halo_depth, cfg.sc.preconditioner, cfg.sc.profiler_on, &c));
    c->solver = cfg.sc.solver;
    c->ppcg_fresh_z = cfg.ppcg_fresh_z != 0;
    c->timing = (warmup >= 0);
    FortranProfile fprof;
    g_fprof = cfg.profiler_on ? &fprof : nullptr;
    int rc = start(c, cfg);
    if (rc == TL_OK) emit_summary(lg, c, 0, 0.0, false, cfg, nullptr);
    Save sv;
    double time = 0.0, dt = cfg.dtinit;
    const double g_small = (double)1.0e-16f;
    int step = 0;
    std::chrono::steady_clock::time_point t0;
    const auto wall0 = std::chrono::steady_clock::now();
    double wall_clock = 0.0;
    Stats stats;
    while (rc == TL_OK) {
        step++;
        if (warmup >= 0 && step == warmup + 1) {
            const hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) {
                rc = hip_fail("hipDeviceSynchronize", e);
                break;
            }
            c->t_k1_ms = c->t_k2_ms = 0;
            c->t_samples = 0;
            stats = Stats();
            t0 = std::chrono::steady_clock::now();
        }
        const double t_dt = wtime();
        if (cfg.dtinit <= dt) dt = cfg.dtinit;
        if (g_fprof) g_fprof->timestep += wtime() - t_dt;
        // timestep.f90:58 (the format's trailing i8 has no item)
        tout(" Step " + fort_I(step, 7) + " time " + fort_F(time, 11, 7) + " timestep  " + fort_1PE(dt, 9, 2));
        rc = cfg.refine ? solve_refined(c, cfg, dt, step, lg, &stats)
             : (mode == TL_MODE_NATIVE) ? solve_native(c, cfg, sv, dt, step, lg, &stats)
                                        : solve_dropin(c, cfg, sv, dt, step, lg, &stats);
        if (rc != TL_OK) break;
        time = time + dt;
        wall_clock = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
        if (g_tout) {
            // diffuse.f90:66-69 (list-directed)
            tout(" Wall clock " + fort_list(wall_clock));
            fflush(g_tout);
        }
        if (lg) fflush(lg);
        bool done = (time + g_small > cfg.end_time || step >= cfg.end_step);
        if (max_steps > 0 && step >= max_steps) done = true;
        if (cfg.summary_frequency != 0 && step % cfg.summary_frequency == 0)
            emit_summary(lg, c, step, time, false, cfg, nullptr);
        if (done) {
            if (warmup >= 0 && out) {
                const hipError_t e = hipDeviceSynchronize();
                if (e != hipSuccess) {
                    rc = hip_fail("hipDeviceSynchronize", e);
                    break;
                }
                out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                out->stats = stats;
            }
            emit_summary(lg, c, step, time, true, cfg, out ? out->summary : nullptr);
            break;
        }
    }
    if (rc == TL_OK && g_fprof) tout_profile(fprof, wall_clock);
    g_fprof = nullptr;
    if (rc == TL_OK && cfg.sc.profiler_on) tl_print_profiling(c);
    if (out) {
        out->k1_ms = c->t_k1_ms;
        out->k2_ms = c->t_k2_ms;
        out->k_samples = c->t_samples;
    }
    tl_chunk_destroy(c);
    return rc;
}

}  // namespace

extern "C" int tl_format_fortran(int kind, double v, int w, int d, char *buf, int len)
{
    // d < w bounds every text (at most 60 + exponent characters), and Ew.d
    // needs at least one mantissa digit
    if (!buf || len < 1 || kind < 0 || kind > 3 ||
        (kind < 3 && (w < 1 || d < 0 || w > 60 || d >= w || (kind == 0 && d < 1)))) {
        set_error("tl_format_fortran: bad kind %d, w %d, d %d or buffer", kind, w, d);
        return TL_ERR_ARG;
    }
    const std::string s = kind == 0 ? fort_E(v, w, d) : kind == 1 ? fort_1PE(v, w, d) : kind == 2 ? fort_F(v, w, d)
                                                                                      : fort_list(v);
    if ((int)s.size() >= len) {
        set_error("tl_format_fortran: buffer of %d bytes too small", len);
        return TL_ERR_ARG;
    }
    memcpy(buf, s.c_str(), s.size() + 1);
    return TL_OK;
}

extern "C" int tl_calc_eigenvalues(const double *cg_alphas, const double *cg_betas, int n, double *eigmin,
                                   double *eigmax)
{
    if (!cg_alphas || !cg_betas || !eigmin || !eigmax || n < 1) {
        set_error("tl_calc_eigenvalues: null argument or n %d < 1", n);
        return TL_ERR_ARG;
    }
    if (calc_eigenvalues(cg_alphas, cg_betas, n, eigmin, eigmax) != 0) {
        set_error("tl_calc_eigenvalues: tqli did not converge or a bound is negative");
        return TL_ERR_SOLVER;
    }
    return TL_OK;
}

extern "C" int tl_calc_ch_coefs(double eigmin, double eigmax, int n, double *ch_alphas, double *ch_betas,
                                double *theta)
{
    if (!ch_alphas || !ch_betas || !theta || n < 1) {
        set_error("tl_calc_ch_coefs: null argument or n %d < 1", n);
        return TL_ERR_ARG;
    }
    std::vector<double> ca, cb;
    calc_ch_coefs(ca, cb, eigmin, eigmax, theta, n);
    std::copy(ca.begin(), ca.begin() + n, ch_alphas);
    std::copy(cb.begin(), cb.begin() + n, ch_betas);
    return TL_OK;
}

extern "C" int tl_parse_deck(const char *deck, int *ints, double *reals, double *states, int max_states)
{
    if (!deck || !ints || !reals || (max_states > 0 && !states)) {
        set_error("tl_parse_deck: null argument");
        return TL_ERR_ARG;
    }
    Config cfg;
    std::string err;
    if (parse_deck(deck, &cfg, &err)) {
        set_error("%s", err.c_str());
        return TL_ERR_CONFIG;
    }
    const int iv[23] = {cfg.x_cells, cfg.y_cells, cfg.end_step, cfg.summary_frequency, cfg.visit_frequency,
                        cfg.max_iters, cfg.coefficient, cfg.tl_ch_cg_presteps, cfg.tl_check_result,
                        cfg.tl_ppcg_inner_steps, cfg.reflective_boundary, cfg.solver, cfg.fortran_preconditioner,
                        cfg.verbose_on, cfg.profiler_on, cfg.halo_depth, cfg.test_problem, cfg.ppcg_fresh_z,
                        cfg.nstates, cfg.sc.preconditioner, cfg.sc.halo_depth, cfg.sc.solver, cfg.sc.profiler_on};
    const double rv[8] = {cfg.xmin, cfg.ymin, cfg.xmax, cfg.ymax, cfg.end_time, cfg.dtinit, cfg.eps,
                          cfg.tl_ch_cg_epslim};
    std::copy(iv, iv + 23, ints);
    std::copy(rv, rv + 8, reals);
    for (int s = 0; s < std::min(max_states, cfg.nstates); ++s) {
        const double sv[8] = {cfg.density[s], cfg.energy[s], cfg.sxmin[s], cfg.sxmax[s],
                              cfg.symin[s],   cfg.symax[s],  cfg.radius[s], (double)cfg.geometry[s]};
        std::copy(sv, sv + 8, states + 8 * s);
    }
    return TL_OK;
}

extern "C" int tl_run_deck_out(const char *deck_path, const char *log_path, const char *tea_out_path, int mode,
                               int max_steps)
{
    Config cfg;
    std::string err;
    if (parse_deck(deck_path, &cfg, &err)) {
        set_error("%s", err.c_str());
        return TL_ERR_CONFIG;
    }
    FILE *lg = nullptr;
    if (log_path && log_path[0]) {
        lg = fopen(log_path, "w");
        if (!lg) {
            set_error("cannot open %s", log_path);
            return TL_ERR_ARG;
        }
    }
    if (tea_out_path && tea_out_path[0]) {
        g_tout = fopen(tea_out_path, "w");
        if (!g_tout) {
            if (lg) fclose(lg);
            set_error("cannot open %s", tea_out_path);
            return TL_ERR_ARG;
        }
        tout("");
        tout(" Starting the calculation");
    }
    const int rc = run(cfg, lg, mode, max_steps, -1, nullptr);
    if (g_tout) {
        if (rc == TL_OK) {
            tout("");
            tout(" Calculation complete");
            tout(" Tea is finishing");
        }
        fclose(g_tout);
        g_tout = nullptr;
    }
    if (lg) fclose(lg);
    return rc;
}

extern "C" int tl_run_deck(const char *deck_path, const char *log_path, int mode, int max_steps)
{
    return tl_run_deck_out(deck_path, log_path, nullptr, mode, max_steps);
}

extern "C" int tl_bench_deck(const char *deck_path, int mode, int warmup_steps, int steps, tl_bench_result *out)
{
    Config cfg;
    std::string err;
    if (parse_deck(deck_path, &cfg, &err)) {
        set_error("%s", err.c_str());
        return TL_ERR_CONFIG;
    }
    if (warmup_steps < 0 || steps < 1) {
        set_error("tl_bench_deck: warmup %d steps %d", warmup_steps, steps);
        return TL_ERR_ARG;
    }
    cfg.end_step = warmup_steps + steps;
    cfg.end_time = 1e300;
    cfg.summary_frequency = 0;
    cfg.verbose_on = 0;
    RunOut ro;
    const int rc = run(cfg, nullptr, mode, 0, warmup_steps, &ro);
    if (rc != TL_OK) return rc;
    memset(out, 0, sizeof *out);
    out->seconds = ro.seconds;
    out->iterations = ro.stats.iterations;
    out->inner_iterations = ro.stats.inner;
    out->k_samples = ro.k_samples;
    out->k1_ms = ro.k_samples ? ro.k1_ms / ro.k_samples : -1.0;
    out->k2_ms = ro.k_samples ? ro.k2_ms / ro.k_samples : -1.0;
    for (int i = 0; i < 5; ++i) out->final_summary[i] = ro.summary[i];
    return TL_OK;
}
