// Deck reader: the tea.in keywords and defaults of read_input.f90:38-327 with
// the line handling of parse.f90:143-354, and the C++ side channel of
// ocl_init.cpp:131-220 / ocl_strings.cpp:7-132 (which is where the reference
// takes the preconditioner, halo_depth and profiler_on from).
#pragma once

#include <string>
#include <vector>

namespace tl {

struct SideChannel {
    int preconditioner = 1;  // PREC_* from "tl_preconditioner_type=..."
    int halo_depth = -1;     // -1 when absent or unparsable
    int solver = 1;          // SOLVER_*; ppcg > chebyshev > cg > jacobi
    int profiler_on = 0;
};

struct Config {
    double xmin = 0.0, ymin = 0.0, xmax = 100.0, ymax = 100.0;
    int x_cells = 10, y_cells = 10;
    double end_time = 10.0;
    int end_step = 640000;  // g_ibig
    double dtinit = (double)0.1f;
    int summary_frequency = 10, visit_frequency = 0;
    int max_iters = 1000;
    double eps = (double)1.0e-10f;
    int coefficient = 1;
    int tl_ch_cg_presteps = 25;
    double tl_ch_cg_epslim = 1.0;
    int tl_check_result = 0;
    int tl_ppcg_inner_steps = -1;
    int reflective_boundary = 0;
    int solver = 1;               // Fortran flags: last tl_use_* wins
    int fortran_preconditioner = 1;
    int verbose_on = 0, profiler_on = 0;
    int halo_depth = 1;           // Fortran keyword
    int test_problem = 0;
    // refined solve instead of the reference solver (library keywords
    // tl_fp64_refined = 1, tl_mixed_precision = 2; tl_refine_delta, < 0 for
    // the default of the precision)
    int refine = 0;
    double refine_delta = -1.0;
    // tl_ppcg_fresh_z (library keyword): PPCG's r.z with z = M^-1 r of the
    // final r of the inner steps, as the Fortran TeaLeaf forms it
    int ppcg_fresh_z = 0;
    int nstates = 0;
    std::vector<double> density, energy, sxmin, sxmax, symin, symax, radius;
    std::vector<int> geometry;
    SideChannel sc;
};

// 0 on success; otherwise nonzero with a message in `err`
int parse_deck(const std::string &path, Config *cfg, std::string *err);
int read_side_channel(const std::string &path, SideChannel *sc);

}  // namespace tl
