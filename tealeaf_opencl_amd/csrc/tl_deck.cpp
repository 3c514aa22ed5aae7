// Deck reader (see tl_deck.h for the reference files it follows).
#include "tl_deck.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>

namespace tl {

namespace {

// parse_getline (parse.f90:143-247): blank control characters, cut '!' and
// ';' comments, lower-case, '=' and ',' become blanks. A line whose first
// word starts with '*' opens the block of that name, or closes one when its
// characters 2-4 are "end" ("*endtea" closes "*tea"; the innermost open block
// must be the one closed). Only lines read while exactly the *tea block is
// open are returned, and reading stops at the *endtea that closes it.
// Returns -1 when the file cannot be opened, -2 for an unmatched */*end pair.
int deck_lines(const std::string &path, std::vector<std::string> *out)
{
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    std::string raw;
    std::vector<std::string> open_blocks;
    auto at_tea = [&]() { return open_blocks.size() == 1 && open_blocks[0] == "*tea"; };
    while (std::getline(f, raw)) {
        if (raw.size() > 100) raw.resize(100);  // READ(... '(a100)')
        for (auto &ch : raw)
            if ((unsigned char)ch < 32 || (unsigned char)ch > 128) ch = ' ';
        size_t p = raw.find('!');
        if (p != std::string::npos) raw.resize(p);
        p = raw.find(';');
        if (p != std::string::npos) raw.resize(p);
        for (auto &ch : raw) {
            ch = (char)std::tolower((unsigned char)ch);
            if (ch == '=' || ch == ',') ch = ' ';
        }
        const size_t b = raw.find_first_not_of(' ');
        if (b == std::string::npos) continue;
        std::string l = raw.substr(b);
        while (!l.empty() && l.back() == ' ') l.pop_back();
        if (l[0] == '*') {
            if (l.compare(0, 8, "*select:") == 0) continue;
            const std::string name = l.substr(0, l.find(' '));
            if (name.compare(1, 3, "end") == 0) {
                const std::string closes = "*" + name.substr(std::min<size_t>(4, name.size()));
                if (open_blocks.empty() || open_blocks.back() != closes) return -2;
                const bool last = at_tea();
                open_blocks.pop_back();
                if (last) break;
            } else {
                open_blocks.push_back(name);
            }
            continue;
        }
        if (at_tea()) out->push_back(l);
    }
    return 0;
}

// parse_getword (parse.f90:249-277) over the processed lines: a word "\"
// continues on the next line; with wrap, an exhausted line continues too
struct Cursor {
    const std::vector<std::string> &lines;
    size_t li = 0;
    std::vector<std::string> words;
    size_t wi = 0;
    explicit Cursor(const std::vector<std::string> &l) : lines(l) {}
    bool next_line()
    {
        if (li >= lines.size()) return false;
        words.clear();
        wi = 0;
        std::istringstream is(lines[li++]);
        std::string s;
        while (is >> s) words.push_back(s);
        return true;
    }
    bool word(bool wrap, std::string *out)
    {
        for (;;) {
            if (wi >= words.size()) {
                if (wrap && next_line()) continue;
                return false;
            }
            const std::string &w = words[wi++];
            if (w == "\\") {
                if (!next_line()) return false;
                continue;
            }
            *out = w;
            return true;
        }
    }
};

// READ(word,'(I7)'): the first seven characters only
bool ival(const std::string &w, int *v)
{
    const std::string s = w.substr(0, 7);
    char *end = nullptr;
    const long x = strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end) return false;
    *v = (int)x;
    return true;
}

bool rval(const std::string &w, double *v)
{
    char *end = nullptr;
    const double x = strtod(w.c_str(), &end);
    if (end == w.c_str() || *end) return false;
    *v = x;
    return true;
}

// matchParam (ocl_strings.cpp:7-37): first raw line without '!' that
// contains the name; the text after its first '=', or the name itself
bool side_match(const std::string &path, const char *name, std::string *val)
{
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.find('!') != std::string::npos) continue;
        if (line.find(name) != std::string::npos) {
            const size_t eq = line.find('=');
            *val = (eq != std::string::npos) ? line.substr(eq + 1) : std::string(name);
            return true;
        }
    }
    val->clear();
    return false;
}

}  // namespace

int read_side_channel(const std::string &path, SideChannel *sc)
{
    {
        std::ifstream f(path);
        if (!f.is_open()) return -1;
    }
    std::string v;
    sc->profiler_on = side_match(path, "profiler_on", &v) && !v.empty();
    sc->halo_depth = -1;
    if (side_match(path, "halo_depth", &v) && !v.empty()) {
        std::stringstream conv(v);
        int d;
        if (conv >> d) sc->halo_depth = d;
    }
    const bool ppcg = side_match(path, "tl_use_ppcg", &v) && !v.empty();
    const bool cheby = side_match(path, "tl_use_chebyshev", &v) && !v.empty();
    const bool cg = side_match(path, "tl_use_cg", &v) && !v.empty();
    sc->solver = ppcg ? 4 : cheby ? 3 : cg ? 2 : 1;
    sc->preconditioner = 1;
    if (side_match(path, "tl_preconditioner_type", &v)) {
        std::transform(v.begin(), v.end(), v.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
        if (v.find("jac_diag") != std::string::npos) sc->preconditioner = 2;
        else if (v.find("jac_block") != std::string::npos) sc->preconditioner = 3;
        else sc->preconditioner = 1;
    }
    return 0;
}

int parse_deck(const std::string &path, Config *cfg, std::string *err)
{
    *cfg = Config();
    std::vector<std::string> lines;
    const int rl = deck_lines(path, &lines);
    if (rl) {
        *err = rl == -2 ? "Unmatched */*end pair in " + path : "Error opening " + path;
        return 1;
    }
    int state_max = 0;
    {
        Cursor cur(lines);
        while (cur.next_line()) {
            std::string w;
            while (cur.word(false, &w))
                if (w == "state") {
                    std::string v;
                    int sn = 0;
                    if (cur.word(true, &v) && ival(v, &sn)) state_max = std::max(state_max, sn);
                    break;
                }
        }
    }
    if (state_max < 1) {
        *err = "read_input: No states defined.";
        return 1;
    }
    cfg->nstates = state_max;
    auto resize = [&](std::vector<double> &v) { v.assign(state_max, 0.0); };
    resize(cfg->density); resize(cfg->energy); resize(cfg->sxmin); resize(cfg->sxmax);
    resize(cfg->symin); resize(cfg->symax); resize(cfg->radius);
    cfg->geometry.assign(state_max, 0);
    std::vector<int> defined(state_max, 0);

    Cursor cur(lines);
    auto getr = [&](double *dst) -> bool {
        std::string s;
        if (!cur.word(true, &s) || !rval(s, dst)) {
            *err = "parse_getrval: Error attempting to convert to real:" + s;
            return false;
        }
        return true;
    };
    auto geti = [&](int *dst) -> bool {
        std::string s;
        if (!cur.word(true, &s) || !ival(s, dst)) {
            *err = "parse_getival: Error attempting to convert to integer:" + s;
            return false;
        }
        return true;
    };
    while (cur.next_line()) {
        std::string k;
        while (cur.word(false, &k)) {
            bool ok = true;
            if (k == "initial_timestep") ok = getr(&cfg->dtinit);
            else if (k == "end_time") ok = getr(&cfg->end_time);
            else if (k == "end_step") ok = geti(&cfg->end_step);
            else if (k == "xmin") ok = getr(&cfg->xmin);
            else if (k == "xmax") ok = getr(&cfg->xmax);
            else if (k == "ymin") ok = getr(&cfg->ymin);
            else if (k == "ymax") ok = getr(&cfg->ymax);
            else if (k == "x_cells") ok = geti(&cfg->x_cells);
            else if (k == "y_cells") ok = geti(&cfg->y_cells);
            else if (k == "visit_frequency") ok = geti(&cfg->visit_frequency);
            else if (k == "summary_frequency") ok = geti(&cfg->summary_frequency);
            else if (k == "tl_ch_cg_presteps") ok = geti(&cfg->tl_ch_cg_presteps);
            else if (k == "tl_ppcg_inner_steps") ok = geti(&cfg->tl_ppcg_inner_steps);
            else if (k == "tl_ch_cg_epslim") ok = getr(&cfg->tl_ch_cg_epslim);
            else if (k == "tl_check_result") cfg->tl_check_result = 1;
            else if (k == "tl_preconditioner_type") {
                std::string v;
                while (cur.word(false, &v)) {
                    if (v == "none") cfg->fortran_preconditioner = 1;
                    else if (v == "jac_diag") cfg->fortran_preconditioner = 2;
                    else if (v == "jac_block") cfg->fortran_preconditioner = 3;
                    else {
                        *err = "read_input: Invalid preconditioner type specified: " + v;
                        return 1;
                    }
                }
            } else if (k == "use_opencl_kernels") {
            } else if (k == "verbose_on") cfg->verbose_on = 1;
            else if (k == "tl_use_jacobi") cfg->solver = 1;
            else if (k == "tl_use_cg") cfg->solver = 2;
            else if (k == "tl_use_ppcg") cfg->solver = 4;
            else if (k == "tl_use_chebyshev") cfg->solver = 3;
            else if (k == "reflective_boundary") cfg->reflective_boundary = 1;
            else if (k == "profiler_on") cfg->profiler_on = 1;
            else if (k == "halo_depth") {
                ok = geti(&cfg->halo_depth);
                if (ok && cfg->halo_depth < 1) {
                    *err = "read_input: Invalid halo exchange depth specified";
                    return 1;
                }
            } else if (k == "tl_max_iters") ok = geti(&cfg->max_iters);
            else if (k == "tl_eps") ok = getr(&cfg->eps);
            else if (k == "tl_coefficient_density") cfg->coefficient = 1;
            else if (k == "tl_coefficient_inverse_density") cfg->coefficient = 2;
            // refined-solve keywords of this library (not in read_input.f90)
            else if (k == "tl_mixed_precision") cfg->refine = 2;
            else if (k == "tl_fp64_refined") cfg->refine = 1;
            else if (k == "tl_refine_delta") ok = getr(&cfg->refine_delta);
            else if (k == "tl_ppcg_fresh_z") cfg->ppcg_fresh_z = 1;
            else if (k == "test_problem") ok = geti(&cfg->test_problem);
            else if (k == "state") {
                int sn = 0;
                if (!geti(&sn)) return 1;
                if (sn < 1 || sn > state_max) {
                    *err = "read_input: bad state number";
                    return 1;
                }
                if (defined[sn - 1]) {
                    *err = "read_input: State defined twice.";
                    return 1;
                }
                defined[sn - 1] = 1;
                const int q = sn - 1;
                std::string kw;
                while (ok && cur.word(false, &kw)) {
                    if (kw == "xmin") ok = getr(&cfg->sxmin[q]);
                    else if (kw == "ymin") ok = getr(&cfg->symin[q]);
                    else if (kw == "xmax") ok = getr(&cfg->sxmax[q]);
                    else if (kw == "ymax") ok = getr(&cfg->symax[q]);
                    else if (kw == "radius") ok = getr(&cfg->radius[q]);
                    else if (kw == "density") ok = getr(&cfg->density[q]);
                    else if (kw == "energy") ok = getr(&cfg->energy[q]);
                    else if (kw == "geometry") {
                        std::string gname;
                        if (cur.word(true, &gname)) {
                            if (gname == "rectangle") cfg->geometry[q] = 1;
                            else if (gname == "circle") cfg->geometry[q] = 2;
                            else if (gname == "point") cfg->geometry[q] = 3;
                        }
                    }
                }
                if (!ok) return 1;
                break;
            }
            if (!ok) return 1;
        }
    }
    // read_input.f90:294-301
    if (cfg->tl_ppcg_inner_steps == -1)
        cfg->tl_ppcg_inner_steps = 4 * (int)(sqrtf(sqrtf((float)(cfg->x_cells * cfg->y_cells))));
    if (cfg->halo_depth > 1 && cfg->fortran_preconditioner == 3 && cfg->solver == 4) {
        *err = "read_input: Unable to use nonstandard halo depth with block jacobi preconditioner";
        return 1;
    }
    // read_input.f90:320-327: states 2..n shrink by a hundredth of a cell
    const double dx = (cfg->xmax - cfg->xmin) / (double)(float)cfg->x_cells;
    const double dy = (cfg->ymax - cfg->ymin) / (double)(float)cfg->y_cells;
    for (int s = 1; s < cfg->nstates; ++s) {
        cfg->sxmin[s] = cfg->sxmin[s] + (dx / 100.0);
        cfg->symin[s] = cfg->symin[s] + (dy / 100.0);
        cfg->sxmax[s] = cfg->sxmax[s] - (dx / 100.0);
        cfg->symax[s] = cfg->symax[s] - (dy / 100.0);
    }
    read_side_channel(path, &cfg->sc);
    return 0;
}

}  // namespace tl
