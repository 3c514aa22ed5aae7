/*
 * TeaLeaf CPU oracle -- TEST INFRASTRUCTURE ONLY (see tl_oracle.h).
 *
 * Every routine below restates one reference kernel or driver routine and
 * cites the file:line it follows. Launch extents follow
 * ocl_kernels.cpp:50-134 (findPaddingSize + WITHIN_BOUNDS): a kernel compiled
 * with launch bounds (xmin,xmax,ymin,ymax) touches columns
 * [hd+xmin, x_max+hd-1+xmax] and rows [hd+ymin, y_max+hd-1+ymax].
 *
 * Parity status of each part is pinned by tests/ (known-answer field summary
 * values of field_summary.f90:90-95, dense solves, operator invariants,
 * LAPACK eigenvalues); see the test files for which check pins what.
 */
#include "tl_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct tlo_chunk {
    int x_max, y_max, hd, pitch;
    int preconditioner;           /* -DPRECONDITIONER (ocl_kernels.cpp:16) */
    long len2d;
    double *f2[TLO_NUM_FIELDS2D];
    double *f1[TLO_NUM_FIELDS1D];
    double *rowbuf;               /* per-row partial sums */
    double *rowbuf4;              /* 4 per row, field_summary */
    /* chebyshev / ppcg coefficients (tea_leaf_kernel_ocl.cpp:98-129,327-344) */
    double *ch_alphas, *ch_betas;
    int n_coefs;
    double theta, rx, ry;
};

#define IDX(c, j, k) ((long)(j) + (long)(k) * (c)->pitch)

int tlo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* chunk lifecycle: ocl_buffers.cpp:3-69 (every buffer zero-filled)          */
/* ------------------------------------------------------------------------ */

tlo_chunk *tlo_create(int x_max, int y_max, int hd, int preconditioner)
{
    if (x_max < 1 || y_max < 1 || hd < 1) return NULL;
    tlo_chunk *c = (tlo_chunk *)calloc(1, sizeof(tlo_chunk));
    c->x_max = x_max;
    c->y_max = y_max;
    c->hd = hd;
    c->pitch = x_max + 2 * hd;
    c->preconditioner = preconditioner;
    /* every 2-D buffer gets the (+1,+1) allocation of the largest reference
     * field (ocl_buffers.cpp:55-66); the pitch stays x_max+2hd */
    c->len2d = (long)(x_max + 2 * hd + 1) * (long)(y_max + 2 * hd + 1);
    for (int i = 0; i < TLO_NUM_FIELDS2D; i++) c->f2[i] = (double *)calloc(c->len2d, sizeof(double));
    for (int i = 0; i < TLO_NUM_FIELDS1D; i++) {
        int n = (i < TLO_CELLY) ? x_max + 2 * hd + 1 : y_max + 2 * hd + 1;
        c->f1[i] = (double *)calloc(n, sizeof(double));
    }
    c->rowbuf = (double *)calloc(y_max + 2 * hd + 1, sizeof(double));
    c->rowbuf4 = (double *)calloc(4 * (y_max + 2 * hd + 1), sizeof(double));
    return c;
}

void tlo_destroy(tlo_chunk *c)
{
    if (!c) return;
    for (int i = 0; i < TLO_NUM_FIELDS2D; i++) free(c->f2[i]);
    for (int i = 0; i < TLO_NUM_FIELDS1D; i++) free(c->f1[i]);
    free(c->rowbuf);
    free(c->rowbuf4);
    free(c->ch_alphas);
    free(c->ch_betas);
    free(c);
}

double *tlo_field2d(tlo_chunk *c, int id) { return (id >= 0 && id < TLO_NUM_FIELDS2D) ? c->f2[id] : NULL; }
double *tlo_field1d(tlo_chunk *c, int id) { return (id >= 0 && id < TLO_NUM_FIELDS1D) ? c->f1[id] : NULL; }
int tlo_pitch(const tlo_chunk *c) { return c->pitch; }
long tlo_field2d_len(const tlo_chunk *c) { return c->len2d; }
void tlo_set_preconditioner(tlo_chunk *c, int p) { c->preconditioner = p; }

/* sum of the per-row partials in row order (deterministic) */
static double sum_rows(const tlo_chunk *c, int k0, int k1)
{
    double s = 0.0;
    for (int k = k0; k <= k1; k++) s += c->rowbuf[k];
    return s;
}

/* 5-point operator (1 + sum K) u_c - K u_nb as written in
 * tea_leaf_common_cl.cl:28-32 / tea_leaf_cg_cl.cl:47-51 */
static inline double smvp(const tlo_chunk *c, const double *Kx, const double *Ky,
                          const double *v, int j, int k)
{
    const long i = IDX(c, j, k);
    const long p = c->pitch;
    return (1.0
            + (Ky[i + p] + Ky[i])
            + (Kx[i + 1] + Kx[i])) * v[i]
           - (Ky[i + p] * v[i + p] + Ky[i] * v[i - p])
           - (Kx[i + 1] * v[i + 1] + Kx[i] * v[i - 1]);
}

/* ------------------------------------------------------------------------ */
/* initialise_chunk: initialise_chunk_cl.cl:3-68, initialise_chunk_kernel_ocl.cpp:9-27 */
/* ------------------------------------------------------------------------ */
void tlo_initialise_chunk(tlo_chunk *c, double xmin, double ymin, double dx, double dy)
{
    const int hd = c->hd;
    double *vx = c->f1[TLO_VERTEXX], *vdx = c->f1[TLO_VERTEXDX];
    double *cx = c->f1[TLO_CELLX], *cdx = c->f1[TLO_CELLDX];
    double *vy = c->f1[TLO_VERTEXY], *vdy = c->f1[TLO_VERTEXDY];
    double *cy = c->f1[TLO_CELLY], *cdy = c->f1[TLO_CELLDY];
    /* initialise_chunk_first: launched from offset 0 over the padded range;
     * row==0 lanes fill x arrays for column <= x_max+hd (:11-31) */
    for (int col = 0; col <= c->x_max + hd; col++) {
        vx[col] = xmin + dx * (double)(((col)-1) - hd + 1);
        vdx[col] = dx;
        const double vplus = xmin + dx * (double)((col) - hd + 1);
        cx[col] = 0.5 * (vx[col] + vplus);
        cdx[col] = dx;
    }
    for (int row = 0; row <= c->y_max + hd; row++) {
        vy[row] = ymin + dy * (double)(((row)-1) - hd + 1);
        vdy[row] = dy;
        const double vplus = ymin + dy * (double)((row) - hd + 1);
        cy[row] = 0.5 * (vy[row] + vplus);
        cdy[row] = dy;
    }
    /* initialise_chunk_second: launch bounds (-1,1,-1,1) (ocl_kernels.cpp:53) */
    double *vol = c->f2[TLO_VOLUME], *xa = c->f2[TLO_XAREA], *ya = c->f2[TLO_YAREA];
    for (int k = hd - 1; k <= c->y_max + hd; k++)
        for (int j = hd - 1; j <= c->x_max + hd; j++) {
            vol[IDX(c, j, k)] = dx * dy;
            xa[(long)j + (long)k * (c->pitch + 1)] = dy; /* THARR2D(0,0,1) */
            ya[IDX(c, j, k)] = dx;
        }
}

/* ------------------------------------------------------------------------ */
/* generate_chunk: generate_chunk_cl.cl:3-95, generate_chunk_kernel_ocl.cpp:25-91 */
/* ------------------------------------------------------------------------ */
void tlo_generate_chunk(tlo_chunk *c, int nstates, const double *dens, const double *ener,
                        const double *sxmin, const double *sxmax, const double *symin,
                        const double *symax, const double *radius, const int *geom,
                        int g_rect, int g_circ, int g_point)
{
    const int hd = c->hd;
    double *rho = c->f2[TLO_DENSITY], *e0 = c->f2[TLO_ENERGY0];
    const double *vx = c->f1[TLO_VERTEXX], *vy = c->f1[TLO_VERTEXY];
    const double *cx = c->f1[TLO_CELLX], *cy = c->f1[TLO_CELLY];
    /* generate_chunk_init: state 1 everywhere in the interior (:82-95) */
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            e0[IDX(c, j, k)] = ener[0];
            rho[IDX(c, j, k)] = dens[0];
        }
    /* one launch per later state, in order (generate_chunk_kernel_ocl.cpp:81-87) */
    for (int s = 1; s < nstates; s++) {
        const double x_cent = sxmin[s], y_cent = symin[s];
#pragma omp parallel for schedule(static)
        for (int k = hd; k < c->y_max + hd; k++)
            for (int j = hd; j < c->x_max + hd; j++) {
                int hit = 0;
                if (g_rect == geom[s]) {
                    hit = (vx[1 + j] >= sxmin[s] && vx[j] < sxmax[s] &&
                           vy[1 + k] >= symin[s] && vy[k] < symax[s]);
                } else if (geom[s] == g_circ) {
                    const double xp = cx[j] - x_cent, yp = cy[k] - y_cent;
                    const double r = sqrt(xp * xp + yp * yp);
                    hit = (r <= radius[s]);
                } else if (geom[s] == g_point) {
                    hit = (vx[j] == x_cent && vy[k] == y_cent);
                }
                if (hit) {
                    e0[IDX(c, j, k)] = ener[s];
                    rho[IDX(c, j, k)] = dens[s];
                }
            }
    }
    /* generate_chunk_init_u with energy0 (:67-80; kernel_ocl.cpp:173-174) */
    double *u = c->f2[TLO_U], *u0 = c->f2[TLO_U0];
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            u[IDX(c, j, k)] = e0[IDX(c, j, k)] * rho[IDX(c, j, k)];
            u0[IDX(c, j, k)] = e0[IDX(c, j, k)] * rho[IDX(c, j, k)];
        }
}

/* set_field_cl.cl:3-13 */
void tlo_set_field(tlo_chunk *c)
{
    const int hd = c->hd;
    double *e0 = c->f2[TLO_ENERGY0], *e1 = c->f2[TLO_ENERGY1];
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) e1[IDX(c, j, k)] = e0[IDX(c, j, k)];
}

/* field_summary_cl.cl:3-43 with energy1 bound to the energy argument
 * (ocl_kernels.cpp:466-469) */
void tlo_field_summary(tlo_chunk *c, double *vol, double *mass, double *ie, double *temp)
{
    const int hd = c->hd;
    const double *V = c->f2[TLO_VOLUME], *rho = c->f2[TLO_DENSITY];
    const double *e1 = c->f2[TLO_ENERGY1], *u = c->f2[TLO_U];
    double *rb = c->rowbuf4;
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++) {
        double sv = 0, sm = 0, si = 0, st = 0;
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            const double cell_vol = V[i];
            const double cell_mass = cell_vol * rho[i];
            sv += cell_vol;
            sm += cell_mass;
            si += cell_mass * e1[i];
            st += cell_mass * u[i];
        }
        rb[4 * k + 0] = sv; rb[4 * k + 1] = sm; rb[4 * k + 2] = si; rb[4 * k + 3] = st;
    }
    double s[4] = {0, 0, 0, 0};
    for (int k = hd; k < c->y_max + hd; k++)
        for (int q = 0; q < 4; q++) s[q] += rb[4 * k + q];
    *vol = s[0]; *mass = s[1]; *ie = s[2]; *temp = s[3];
}

/* ------------------------------------------------------------------------ */
/* update_halo: update_halo_cl.cl:3-69, update_halo_kernel_ocl.cpp:14-62,    */
/* launch sizes ocl_kernels.cpp:356-412. Reflective copy of density,         */
/* energy0 and energy1 only; faces in the order bottom, top, left, right.    */
/* ------------------------------------------------------------------------ */
static void halo_reflect(tlo_chunk *c, double *a, const int *nb, int depth)
{
    const int hd = c->hd, xm = c->x_max, ym = c->y_max;
    if (depth > hd) depth = hd; /* the allocated halo bounds every fill */
    /* chunk_neighbours: 1=left 2=right 3=bottom 4=top, -1 = external */
    if (nb[2] == -1) /* bottom: rows [hd-depth, hd-1], cols [hd-depth, xm+hd-1+depth] */
        for (int k = hd - depth; k <= hd - 1; k++)
            for (int j = hd - depth; j <= xm + hd - 1 + depth; j++)
                a[IDX(c, j, k)] = 1 * a[IDX(c, j, 2 * hd - 1 - k)];
    if (nb[3] == -1) /* top: dst row ym+2hd-1-k from src row ym+k */
        for (int k = hd - depth; k <= hd - 1; k++)
            for (int j = hd - depth; j <= xm + hd - 1 + depth; j++)
                a[IDX(c, j, ym + 2 * hd - 1 - k)] = a[IDX(c, j, ym + k)];
    if (nb[0] == -1) /* left: cols [hd-depth, hd-1], rows [hd-depth, ym+hd-1+depth] */
        for (int k = hd - depth; k <= ym + hd - 1 + depth; k++)
            for (int j = hd - depth; j <= hd - 1; j++)
                a[IDX(c, j, k)] = 1 * a[IDX(c, 2 * hd - 1 - j, k)];
    if (nb[1] == -1)
        for (int k = hd - depth; k <= ym + hd - 1 + depth; k++)
            for (int j = hd - depth; j <= hd - 1; j++)
                a[IDX(c, xm + 2 * hd - 1 - j, k)] = a[IDX(c, xm + j, k)];
}

void tlo_update_halo(tlo_chunk *c, const int *nb, const int *fields, int depth)
{
    /* fields[] indexed by FIELD_* - 1 (definitions.hpp:14-21) */
    if (fields[0] == 1) halo_reflect(c, c->f2[TLO_DENSITY], nb, depth);
    if (fields[1] == 1) halo_reflect(c, c->f2[TLO_ENERGY0], nb, depth);
    if (fields[2] == 1) halo_reflect(c, c->f2[TLO_ENERGY1], nb, depth);
}

/* ------------------------------------------------------------------------ */
/* common init: tea_leaf_kernel_ocl.cpp:8-32 (calcrxry), :254-289;           */
/* kernels tea_leaf_common_cl.cl:56-125, generate_chunk_cl.cl:67-80          */
/* ------------------------------------------------------------------------ */
void tlo_common_init(tlo_chunk *c, int coefficient, double dt, double *rx, double *ry,
                     const int *zero_boundary, int reflective_boundary)
{
    const int hd = c->hd, xm = c->x_max, ym = c->y_max;
    const double dx = c->f1[TLO_CELLDX][1 + hd];
    const double dy = c->f1[TLO_CELLDY][1 + hd];
    *rx = dt / (dx * dx);
    *ry = dt / (dy * dy);
    const double rxv = *rx, ryv = *ry;
    const double *rho = c->f2[TLO_DENSITY];
    double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    /* tea_leaf_init_common: launch bounds (1-hd, hd) (ocl_kernels.cpp:112) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= ym + 2 * hd - 1; k++)
        for (int j = 1; j <= xm + 2 * hd - 1; j++) {
            const long i = IDX(c, j, k);
            double dc, dl, dd;
            if (coefficient == TLO_COEF_CONDUCTIVITY) {
                dc = rho[i]; dl = rho[i - 1]; dd = rho[i - c->pitch];
            } else {
                dc = 1.0 / rho[i]; dl = 1.0 / rho[i - 1]; dd = 1.0 / rho[i - c->pitch];
            }
            double kx = (dl + dc) / (2.0 * dl * dc);
            kx *= rxv;
            double ky = (dd + dc) / (2.0 * dd * dc);
            ky *= ryv;
            Kx[i] = kx;
            Ky[i] = ky;
        }
    if (!reflective_boundary) {
        /* tea_leaf_zero_boundary: launch bounds (-hd, hd) (ocl_kernels.cpp:113) */
        const int zl = zero_boundary[0], zr = zero_boundary[1];
        const int zb = zero_boundary[2], zt = zero_boundary[3];
        for (int k = 0; k <= ym + 2 * hd - 1; k++)
            for (int j = 0; j <= xm + 2 * hd - 1; j++) {
                const long i = IDX(c, j, k);
                if (zl && j <= hd) { Kx[i] = 0; if (j < hd) Ky[i] = 0; }
                if (zr && j > (xm - 1) + hd) { Kx[i] = 0; Ky[i] = 0; }
                if (zb && k <= hd) { Ky[i] = 0; if (k < hd) Kx[i] = 0; }
                if (zt && k > (ym - 1) + hd) { Ky[i] = 0; Kx[i] = 0; }
            }
    }
    /* generate_chunk_init_u with energy1 (tea_leaf_kernel_ocl.cpp:287-288) */
    const double *e1 = c->f2[TLO_ENERGY1];
    double *u = c->f2[TLO_U], *u0 = c->f2[TLO_U0];
    for (int k = hd; k < ym + hd; k++)
        for (int j = hd; j < xm + hd; j++) {
            const long i = IDX(c, j, k);
            u[i] = e1[i] * rho[i];
            u0[i] = e1[i] * rho[i];
        }
}

/* tea_leaf_common_cl.cl:17-36 */
void tlo_calc_residual(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    const double *u = c->f2[TLO_U], *u0 = c->f2[TLO_U0];
    double *r = c->f2[TLO_R];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            const double s = smvp(c, Kx, Ky, u, j, k);
            r[IDX(c, j, k)] = u0[IDX(c, j, k)] - s;
        }
}

/* tea_leaf_common_cl.cl:38-54 with the operands of tea_leaf_kernel_ocl.cpp:56-96 */
double tlo_calc_2norm(tlo_chunk *c, int norm_array)
{
    const int hd = c->hd;
    const double *a, *b;
    if (norm_array == 0) { a = c->f2[TLO_U0]; b = c->f2[TLO_U0]; }
    else if (norm_array == 1) { a = c->f2[TLO_R]; b = c->f2[TLO_R]; }
    else if (norm_array == 2) {
        a = c->f2[TLO_R];
        b = (c->preconditioner == TLO_PREC_NONE) ? c->f2[TLO_R] : c->f2[TLO_Z];
    } else {
        fprintf(stderr, "Invalid value '%d' for norm_array passed, should be 0 for u0, 1 for r, 2 for r*z\n",
                norm_array);
        return NAN;
    }
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++) {
        double s = 0.0;
        for (int j = hd; j < c->x_max + hd; j++) s += a[IDX(c, j, k)] * b[IDX(c, j, k)];
        c->rowbuf[k] = s;
    }
    return sum_rows(c, hd, c->y_max + hd - 1);
}

/* tea_leaf_common_cl.cl:4-15 */
void tlo_finalise(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *rho = c->f2[TLO_DENSITY], *u = c->f2[TLO_U];
    double *e1 = c->f2[TLO_ENERGY1];
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) e1[IDX(c, j, k)] = u[IDX(c, j, k)] / rho[IDX(c, j, k)];
}

/* tea_leaf_common_cl.cl:127-142, launch bounds (-hd, hd) */
void tlo_init_jac_diag(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    double *Mi = c->f2[TLO_MI];
    for (int k = 0; k <= c->y_max + 2 * hd - 1; k++)
        for (int j = 0; j <= c->x_max + 2 * hd - 1; j++) {
            const long i = IDX(c, j, k);
            const double diag = (1.0 + (Ky[i + c->pitch] + Ky[i]) + (Kx[i + 1] + Kx[i]));
            Mi[i] = 1.0 / diag;
        }
}

/* ------------------------------------------------------------------------ */
/* strip-4 block Jacobi: tea_block_jacobi.cl:1-105, BLOCK_TOP macros_cl.cl:6-10 */
/* (BLOCK_TOP_CHECK only when y_max % 4 != 0, ocl_kernels.cpp:26-29).        */
/* Strips start at rows hd, hd+4, ... (work-group rows with offset hd).      */
/* ------------------------------------------------------------------------ */
#define JB 4

static inline int block_top(const tlo_chunk *c, int row)
{
    if (c->y_max % JB) {
        const int t = c->y_max + 2 - row;
        return t < JB ? t : JB;
    }
    return JB;
}

/* COEF_A / COEF_B / COEF_C of tea_block_jacobi.cl:1-3 at strip row k */
static inline double coef_a(const tlo_chunk *c, const double *Ky, long i) { (void)c; return -Ky[i]; }
static inline double coef_b(const tlo_chunk *c, const double *Kx, const double *Ky, long i)
{
    return (1.0 + (Ky[i + c->pitch] + Ky[i]) + (Kx[i + 1] + Kx[i]));
}
static inline double coef_c(const tlo_chunk *c, const double *Ky, long i) { return -Ky[i + c->pitch]; }

/* tea_leaf_block_init (:78-105): only the first row of each strip works */
void tlo_block_init(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    double *cp = c->f2[TLO_CP], *bfp = c->f2[TLO_BFP];
#pragma omp parallel for schedule(static)
    for (int row = hd; row <= c->y_max + hd - 1; row += JB)
        for (int j = hd; j < c->x_max + hd; j++) {
            const int upper = block_top(c, row);
            long i0 = IDX(c, j, row);
            cp[i0] = coef_c(c, Ky, i0) / coef_b(c, Kx, Ky, i0);
            for (int k = 1; k < upper; k++) {
                const long i = i0 + (long)k * c->pitch;
                const long im = i - c->pitch;
                bfp[i] = 1.0 / (coef_b(c, Kx, Ky, i) - coef_a(c, Ky, i) * cp[im]);
                cp[i] = coef_c(c, Ky, i) * bfp[i];
            }
        }
}

/* block_solve_func (:6-42) for the strip starting at `row`, column j.
 * rin(k) is the right-hand side of strip row k (0 outside the kernel's
 * bounds, as the LDS copy is); writes z_out[k] for k < 4 (0 where the strip
 * is truncated). */
static inline void block_solve_strip(const tlo_chunk *c, int j, int row, const double *rin,
                                     double *z_out)
{
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    const double *cp = c->f2[TLO_CP], *bfp = c->f2[TLO_BFP];
    const int upper = block_top(c, row);
    double dp[JB];
    const long i0 = IDX(c, j, row);
    for (int k = 0; k < JB; k++) z_out[k] = 0.0;
    if (upper <= 0) return;
    dp[0] = rin[0] / coef_b(c, Kx, Ky, i0);
    for (int k = 1; k < upper; k++) {
        const long i = i0 + (long)k * c->pitch;
        dp[k] = (rin[k] - coef_a(c, Ky, i) * dp[k - 1]) * bfp[i];
    }
    int k = upper - 1;
    z_out[k] = dp[k];
    for (k = upper - 2; k >= 0; k--) {
        const long i = i0 + (long)k * c->pitch;
        z_out[k] = dp[k] - cp[i] * z_out[k + 1];
    }
}

/* gather the strip's right-hand side from field `src` for rows in bounds */
static inline void strip_rhs(const tlo_chunk *c, const double *src, int j, int row, int row_hi,
                             double *rin)
{
    for (int k = 0; k < JB; k++) {
        const int rr = row + k;
        rin[k] = (rr <= row_hi) ? src[IDX(c, j, rr)] : 0.0;
    }
}

/* tea_leaf_block_solve (:44-76): z = M_blk^-1 r on the interior */
void tlo_block_solve(tlo_chunk *c)
{
    const int hd = c->hd, hi = c->y_max + hd - 1;
    const double *r = c->f2[TLO_R];
    double *z = c->f2[TLO_Z];
#pragma omp parallel for schedule(static)
    for (int row = hd; row <= hi; row += JB)
        for (int j = hd; j < c->x_max + hd; j++) {
            double rin[JB], zo[JB];
            strip_rhs(c, r, j, row, hi, rin);
            block_solve_strip(c, j, row, rin, zo);
            for (int k = 0; k < JB && row + k <= hi; k++) z[IDX(c, j, row + k)] = zo[k];
        }
}

/* z = M^-1 r over the interior: the block solve above, or Mi r (the z of
 * tea_leaf_cg_cl.cl:8-41 for the diagonal preconditioner); nothing without a
 * preconditioner. The fresh-z PPCG option calls it before r.z. */
void tlo_precondition_r(tlo_chunk *c)
{
    const int hd = c->hd;
    if (c->preconditioner == TLO_PREC_JAC_BLOCK) {
        tlo_block_solve(c);
    } else if (c->preconditioner == TLO_PREC_JAC_DIAG) {
        const double *r = c->f2[TLO_R], *Mi = c->f2[TLO_MI];
        double *z = c->f2[TLO_Z];
#pragma omp parallel for schedule(static)
        for (int k = hd; k < c->y_max + hd; k++)
            for (int j = hd; j < c->x_max + hd; j++) {
                const long i = IDX(c, j, k);
                z[i] = r[i] * Mi[i];
            }
    }
}

/* ------------------------------------------------------------------------ */
/* CG: tea_leaf_kernel_ocl.cpp:167-212, tea_leaf_cg_cl.cl:8-160              */
/* ------------------------------------------------------------------------ */
double tlo_cg_init(tlo_chunk *c)
{
    const int hd = c->hd, pre = c->preconditioner;
    if (pre == TLO_PREC_JAC_BLOCK) {
        tlo_block_init(c);
        tlo_block_solve(c);
    } else if (pre == TLO_PREC_JAC_DIAG) {
        tlo_init_jac_diag(c);
    }
    double *p = c->f2[TLO_P], *r = c->f2[TLO_R], *z = c->f2[TLO_Z];
    const double *Mi = c->f2[TLO_MI];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++) {
        double s = 0.0;
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            if (pre == TLO_PREC_JAC_BLOCK) {
                p[i] = z[i];
            } else if (pre == TLO_PREC_JAC_DIAG) {
                z[i] = r[i] * Mi[i];
                p[i] = z[i];
            } else {
                p[i] = r[i];
            }
            s += r[i] * p[i];
        }
        c->rowbuf[k] = s;
    }
    return sum_rows(c, hd, c->y_max + hd - 1);
}

double tlo_cg_calc_w(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY], *p = c->f2[TLO_P];
    double *w = c->f2[TLO_W];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++) {
        double s = 0.0;
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            w[i] = smvp(c, Kx, Ky, p, j, k);
            s += p[i] * w[i];
        }
        c->rowbuf[k] = s;
    }
    return sum_rows(c, hd, c->y_max + hd - 1);
}

double tlo_cg_calc_ur(tlo_chunk *c, double alpha)
{
    const int hd = c->hd, hi = c->y_max + hd - 1, pre = c->preconditioner;
    double *u = c->f2[TLO_U], *r = c->f2[TLO_R], *z = c->f2[TLO_Z];
    const double *p = c->f2[TLO_P], *w = c->f2[TLO_W], *Mi = c->f2[TLO_MI];
#pragma omp parallel for schedule(static)
    for (int k = hd; k <= hi; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            u[i] += alpha * p[i];
            r[i] -= alpha * w[i];
        }
    if (pre == TLO_PREC_JAC_BLOCK) {
#pragma omp parallel for schedule(static)
        for (int row = hd; row <= hi; row += JB) {
            double s = 0.0;
            for (int j = hd; j < c->x_max + hd; j++) {
                double rin[JB], zo[JB];
                strip_rhs(c, r, j, row, hi, rin);
                block_solve_strip(c, j, row, rin, zo);
                for (int k = 0; k < JB && row + k <= hi; k++) {
                    z[IDX(c, j, row + k)] = zo[k];
                }
            }
            (void)s;
        }
    }
#pragma omp parallel for schedule(static)
    for (int k = hd; k <= hi; k++) {
        double s = 0.0;
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            if (pre == TLO_PREC_JAC_BLOCK) {
                s += r[i] * z[i];
            } else if (pre == TLO_PREC_JAC_DIAG) {
                z[i] = r[i] * Mi[i];
                s += r[i] * z[i];
            } else {
                s += r[i] * r[i];
            }
        }
        c->rowbuf[k] = s;
    }
    return sum_rows(c, hd, hi);
}

void tlo_cg_calc_p(tlo_chunk *c, double beta)
{
    const int hd = c->hd;
    double *p = c->f2[TLO_P];
    const double *r = c->f2[TLO_R], *z = c->f2[TLO_Z];
    const int pre = c->preconditioner;
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            if (pre != TLO_PREC_NONE) p[i] = z[i] + beta * p[i];
            else p[i] = r[i] + beta * p[i];
        }
}

/* ------------------------------------------------------------------------ */
/* Chebyshev: tea_leaf_kernel_ocl.cpp:98-138, tea_leaf_cheby_cl.cl:4-145     */
/* ------------------------------------------------------------------------ */
static void store_coefs(tlo_chunk *c, const double *a, const double *b, int n)
{
    free(c->ch_alphas);
    free(c->ch_betas);
    c->ch_alphas = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    c->ch_betas = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (n > 0) {
        memcpy(c->ch_alphas, a, sizeof(double) * n);
        memcpy(c->ch_betas, b, sizeof(double) * n);
    }
    c->n_coefs = n;
}

static void cheby_calc_u(tlo_chunk *c)
{
    const int hd = c->hd;
    double *u = c->f2[TLO_U];
    const double *p = c->f2[TLO_P];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) u[IDX(c, j, k)] += p[IDX(c, j, k)];
}

/* w = A u; r = u0 - w over the interior (shared head of cheby init_p/calc_p) */
static void cheby_residual(tlo_chunk *c)
{
    const int hd = c->hd;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    const double *u = c->f2[TLO_U], *u0 = c->f2[TLO_U0];
    double *w = c->f2[TLO_W], *r = c->f2[TLO_R];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            w[i] = smvp(c, Kx, Ky, u, j, k);
            r[i] = u0[i] - w[i];
        }
}

void tlo_cheby_init(tlo_chunk *c, const double *ch_alphas, const double *ch_betas, int n_coefs,
                    double rx, double ry, double theta)
{
    store_coefs(c, ch_alphas, ch_betas, n_coefs);
    c->rx = rx; c->ry = ry; c->theta = theta;
    const int hd = c->hd, hi = c->y_max + hd - 1, pre = c->preconditioner;
    cheby_residual(c);
    double *p = c->f2[TLO_P];
    const double *r = c->f2[TLO_R], *Mi = c->f2[TLO_MI];
    if (pre == TLO_PREC_JAC_BLOCK) {
#pragma omp parallel for schedule(static)
        for (int row = hd; row <= hi; row += JB)
            for (int j = hd; j < c->x_max + hd; j++) {
                double rin[JB], zo[JB];
                strip_rhs(c, r, j, row, hi, rin);
                block_solve_strip(c, j, row, rin, zo);
                for (int k = 0; k < JB && row + k <= hi; k++) p[IDX(c, j, row + k)] = zo[k] / theta;
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int k = hd; k <= hi; k++)
            for (int j = hd; j < c->x_max + hd; j++) {
                const long i = IDX(c, j, k);
                if (pre == TLO_PREC_JAC_DIAG) p[i] = (Mi[i] * r[i]) / theta;
                else p[i] = r[i] / theta;
            }
    }
    cheby_calc_u(c);
}

void tlo_cheby_iterate(tlo_chunk *c, int cheby_calc_step)
{
    const int step = cheby_calc_step - 1;
    const int hd = c->hd, hi = c->y_max + hd - 1, pre = c->preconditioner;
    const double a = c->ch_alphas[step], b = c->ch_betas[step];
    cheby_residual(c);
    double *p = c->f2[TLO_P];
    const double *r = c->f2[TLO_R], *Mi = c->f2[TLO_MI];
    if (pre == TLO_PREC_JAC_BLOCK) {
#pragma omp parallel for schedule(static)
        for (int row = hd; row <= hi; row += JB)
            for (int j = hd; j < c->x_max + hd; j++) {
                double rin[JB], zo[JB];
                strip_rhs(c, r, j, row, hi, rin);
                block_solve_strip(c, j, row, rin, zo);
                for (int k = 0; k < JB && row + k <= hi; k++) {
                    const long i = IDX(c, j, row + k);
                    p[i] = a * p[i] + b * zo[k];
                }
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int k = hd; k <= hi; k++)
            for (int j = hd; j < c->x_max + hd; j++) {
                const long i = IDX(c, j, k);
                if (pre == TLO_PREC_JAC_DIAG) p[i] = a * p[i] + b * Mi[i] * r[i];
                else p[i] = a * p[i] + b * r[i];
            }
    }
    cheby_calc_u(c);
}

/* ------------------------------------------------------------------------ */
/* PPCG: tea_leaf_kernel_ocl.cpp:306-411, tea_leaf_ppcg_cl.cl:4-155          */
/* ------------------------------------------------------------------------ */
void tlo_ppcg_init(tlo_chunk *c, const double *ch_alphas, const double *ch_betas,
                   double theta, int n_inner_steps)
{
    c->theta = theta;
    store_coefs(c, ch_alphas, ch_betas, n_inner_steps);
}

void tlo_ppcg_init_sd(tlo_chunk *c)
{
    const int hd = c->hd, hi = c->y_max + hd - 1, pre = c->preconditioner;
    const double theta = c->theta;
    double *sd = c->f2[TLO_SD];
    const double *r = c->f2[TLO_R], *Mi = c->f2[TLO_MI];
    if (pre == TLO_PREC_JAC_BLOCK) {
#pragma omp parallel for schedule(static)
        for (int row = hd; row <= hi; row += JB)
            for (int j = hd; j < c->x_max + hd; j++) {
                double rin[JB], zo[JB];
                strip_rhs(c, r, j, row, hi, rin);
                block_solve_strip(c, j, row, rin, zo);
                for (int k = 0; k < JB && row + k <= hi; k++) sd[IDX(c, j, row + k)] = zo[k] / theta;
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int k = hd; k <= hi; k++)
            for (int j = hd; j < c->x_max + hd; j++) {
                const long i = IDX(c, j, k);
                if (pre == TLO_PREC_JAC_DIAG) sd[i] = r[i] * Mi[i] / theta;
                else sd[i] = r[i] / theta;
            }
    }
}

/* matrix-powers extent of tea_leaf_kernel_ocl.cpp:352-391 */
static void ppcg_extent(const tlo_chunk *c, int bounds_extra, const int *nb,
                        int *j0, int *j1, int *k0, int *k1)
{
    const int hd = c->hd;
    const int step_depth = hd - bounds_extra;
    int off[2] = {step_depth, step_depth};
    int gsz[2] = {c->x_max + (hd - step_depth) * 2, c->y_max + (hd - step_depth) * 2};
    int bex = bounds_extra, bey = bounds_extra;
    if (nb[0] == -1) { off[0] = hd; gsz[0] -= (hd - step_depth); }
    if (nb[1] == -1) { gsz[0] -= (hd - step_depth); bex = 0; }
    if (nb[2] == -1) { off[1] = hd; gsz[1] -= (hd - step_depth); }
    if (nb[3] == -1) { gsz[1] -= (hd - step_depth); bey = 0; }
    /* pad to a multiple of the 32x4 work-group, always adding one group */
    gsz[0] -= gsz[0] % 32; gsz[0] += 32;
    gsz[1] -= gsz[1] % 4; gsz[1] += 4;
    *j0 = off[0];
    *k0 = off[1];
    int jm = c->x_max + hd - 1 + bex, km = c->y_max + hd - 1 + bey;
    if (off[0] + gsz[0] - 1 < jm) jm = off[0] + gsz[0] - 1;
    if (off[1] + gsz[1] - 1 < km) km = off[1] + gsz[1] - 1;
    *j1 = jm;
    *k1 = km;
}

void tlo_ppcg_inner(tlo_chunk *c, int inner_step, int bounds_extra, const int *nb)
{
    int j0, j1, k0, k1;
    ppcg_extent(c, bounds_extra, nb, &j0, &j1, &k0, &k1);
    const int pre = c->preconditioner;
    const double *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    double *u = c->f2[TLO_U], *r = c->f2[TLO_R], *sd = c->f2[TLO_SD];
    /* tea_leaf_ppcg_solve_update_r (:62-88) */
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++) {
            const long i = IDX(c, j, k);
            u[i] += sd[i];
            const double result = smvp(c, Kx, Ky, sd, j, k);
            r[i] -= result;
        }
    /* tea_leaf_ppcg_solve_calc_sd (:90-155), coefficient index inner_step-1 */
    const int step = inner_step - 1;
    const double a = c->ch_alphas[step], b = c->ch_betas[step];
    const double *Mi = c->f2[TLO_MI];
    if (pre == TLO_PREC_JAC_BLOCK) {
#pragma omp parallel for schedule(static)
        for (int row = k0; row <= k1; row += JB)
            for (int j = j0; j <= j1; j++) {
                double rin[JB], zo[JB];
                strip_rhs(c, r, j, row, k1, rin);
                block_solve_strip(c, j, row, rin, zo);
                for (int k = 0; k < JB && row + k <= k1; k++) {
                    const long i = IDX(c, j, row + k);
                    sd[i] = a * sd[i] + b * zo[k];
                }
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int k = k0; k <= k1; k++)
            for (int j = j0; j <= j1; j++) {
                const long i = IDX(c, j, k);
                if (pre == TLO_PREC_JAC_DIAG) sd[i] = a * sd[i] + b * r[i] * Mi[i];
                else sd[i] = a * sd[i] + b * r[i];
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Jacobi: tea_leaf_kernel_ocl.cpp:222-229, tea_leaf_jacobi_cl.cl:3-42       */
/* (`un` is the Mi buffer, ocl_kernels.cpp:576-585)                          */
/* ------------------------------------------------------------------------ */
double tlo_jacobi_solve(tlo_chunk *c)
{
    const int hd = c->hd;
    const long P = c->pitch;
    double *u = c->f2[TLO_U], *un = c->f2[TLO_MI];
    const double *u0 = c->f2[TLO_U0], *Kx = c->f2[TLO_KX], *Ky = c->f2[TLO_KY];
    for (int k = hd; k < c->y_max + hd; k++)
        for (int j = hd; j < c->x_max + hd; j++) un[IDX(c, j, k)] = u[IDX(c, j, k)];
#pragma omp parallel for schedule(static)
    for (int k = hd; k < c->y_max + hd; k++) {
        double s = 0.0;
        for (int j = hd; j < c->x_max + hd; j++) {
            const long i = IDX(c, j, k);
            u[i] = (u0[i]
                    + Kx[i + 1] * un[i + 1]
                    + Kx[i] * un[i - 1]
                    + Ky[i + P] * un[i + P]
                    + Ky[i] * un[i - P])
                   / (1.0 + (Kx[i] + Kx[i + 1]) + (Ky[i] + Ky[i + P]));
            s += fabs(u[i] - un[i]);
        }
        c->rowbuf[k] = s;
    }
    return sum_rows(c, hd, c->y_max + hd - 1);
}

/* ------------------------------------------------------------------------ */
/* host scalar routines: tea_leaf_cheby.f90:44-172                           */
/* ------------------------------------------------------------------------ */

/* implicit QL on a symmetric tridiagonal matrix (tea_leaf_cheby.f90:44-94);
 * d[0..n-1] diagonal, e[0..n-1] with e[i] the element below d[i-1]
 * (Fortran offdiag(2..n)), shifted down by one before the sweep. */
int tlo_tqli(double *d, double *e, int n)
{
    int i, iter, l, m;
    double b, cc, dd, f, g, p, r, s;
    for (i = 0; i < n - 1; i++) e[i] = e[i + 1]; /* eoshift(e,1) */
    if (n > 0) e[n - 1] = 0.0;
    for (l = 0; l < n; l++) {
        iter = 0;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (iter == 30) return 1;
            iter++;
            g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            r = sqrt(g * g + 1.0 * 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            s = 1.0;
            cc = 1.0;
            p = 0.0;
            int restart = 0;
            for (i = m - 1; i >= l; i--) {
                f = s * e[i];
                b = cc * e[i];
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    restart = 1;
                    break;
                }
                s = f / r;
                cc = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * cc * b;
                p = s * r;
                d[i + 1] = g + p;
                g = cc * r - b;
            }
            if (restart) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

/* tea_calc_eigenvalues (tea_leaf_cheby.f90:96-143): Lanczos tridiagonal from
 * the first nsteps CG alphas/betas, tqli, ascending sort */
int tlo_calc_eigenvalues(const double *cg_alphas, const double *cg_betas, int nsteps,
                         double *eigmin, double *eigmax)
{
    if (nsteps < 1) return 1;
    double *diag = (double *)calloc(nsteps, sizeof(double));
    double *offdiag = (double *)calloc(nsteps, sizeof(double));
    for (int n = 0; n < nsteps; n++) {
        diag[n] = 1.0 / cg_alphas[n];
        if (n > 0) diag[n] = diag[n] + cg_betas[n - 1] / cg_alphas[n - 1];
        if (n < nsteps - 1) offdiag[n + 1] = sqrt(cg_betas[n]) / cg_alphas[n];
    }
    int info = tlo_tqli(diag, offdiag, nsteps);
    if (info == 0) {
        int swapped;
        do {
            swapped = 0;
            for (int n = 0; n < nsteps - 1; n++)
                if (diag[n] >= diag[n + 1]) {
                    double t = diag[n];
                    diag[n] = diag[n + 1];
                    diag[n + 1] = t;
                    swapped = 1;
                }
            /* the reference's >= comparison swaps equal neighbours; stop once a
             * pass only exchanged equal values */
            if (swapped) {
                int sorted = 1;
                for (int n = 0; n < nsteps - 1; n++)
                    if (diag[n] > diag[n + 1]) { sorted = 0; break; }
                if (sorted) swapped = 0;
            }
        } while (swapped);
        *eigmin = diag[0];
        *eigmax = diag[nsteps - 1];
        if (*eigmin < 0.0 || *eigmax < 0.0) info = 1;
    }
    free(diag);
    free(offdiag);
    return info;
}

/* tea_calc_ch_coefs (tea_leaf_cheby.f90:145-172) */
void tlo_calc_ch_coefs(double *ch_alphas, double *ch_betas, double eigmin, double eigmax,
                       double *theta, int n)
{
    const double th = (eigmax + eigmin) / 2;
    const double delta = (eigmax - eigmin) / 2;
    const double sigma = th / delta;
    double rho_old = 1.0 / sigma;
    for (int i = 0; i < n; i++) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        ch_alphas[i] = rho_new * rho_old;
        ch_betas[i] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
    *theta = th;
}

/* ------------------------------------------------------------------------ */
/* deck reader: parse.f90:143-354 line handling, read_input.f90:38-327       */
/* keywords, ocl_strings.cpp:7-132 + ocl_init.cpp:131-220 side channel       */
/* ------------------------------------------------------------------------ */
#define LINE_MAX_LEN 512

/* Fortran READ(word,'(I7)'): only the first 7 characters are read */
static int fortran_ival(const char *w, int *ok)
{
    char buf[8];
    size_t n = strlen(w);
    if (n > 7) n = 7;
    memcpy(buf, w, n);
    buf[n] = 0;
    char *end;
    long v = strtol(buf, &end, 10);
    *ok = (end != buf);
    while (*end == ' ') end++;
    if (*end) *ok = 0;
    return (int)v;
}

static double fortran_rval(const char *w, int *ok)
{
    char *end;
    double v = strtod(w, &end);
    *ok = (end != w && *end == 0);
    return v;
}

/* the processed deck lines parse_getline returns (parse.f90:143-247:
 * lowercased, '='/',' turned to spaces, comments cut), only those read while
 * the block stack (here, parse.f90:186-224) is exactly "*tea"; the reading
 * ends at the *endtea that closes it (parse_getline returns -1). -1: cannot
 * open, -2: unmatched * / *end pair. */
static int deck_lines(const char *path, char (**out)[LINE_MAX_LEN], int *nout)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int cap = 64, n = 0;
    char here[LINE_MAX_LEN * 2] = "";  /* the open blocks' names, concatenated */
    char (*lines)[LINE_MAX_LEN] = malloc(sizeof(*lines) * cap);
    char l[LINE_MAX_LEN];
    while (fgets(l, sizeof l, f)) {
        /* Fortran reads '(a100)': at most 100 characters per line */
        if (strlen(l) > 100) l[100] = 0;
        for (char *q = l; *q; q++)
            if ((unsigned char)*q < 32 || (unsigned char)*q > 128) *q = ' ';
        char *s = l;
        while (*s == ' ') s++;
        char *bang = strchr(s, '!');
        if (bang) *bang = 0;
        char *semi = strchr(s, ';');
        if (semi) *semi = 0;
        for (char *q = s; *q; q++) {
            *q = (char)tolower((unsigned char)*q);
            if (*q == '=' || *q == ',') *q = ' ';
        }
        /* trim */
        size_t len = strlen(s);
        while (len && s[len - 1] == ' ') s[--len] = 0;
        if (!len) continue;
        if (s[0] == '*') {
            if (strncmp(s, "*select:", 8) == 0) continue;
            char nugget[LINE_MAX_LEN];
            size_t w = strcspn(s, " ");
            memcpy(nugget, s, w);
            nugget[w] = 0;
            if (w >= 4 && strncmp(nugget + 1, "end", 3) == 0) {
                char closes[LINE_MAX_LEN];
                snprintf(closes, sizeof closes, "*%s", nugget + 4);
                char *last = strrchr(here, '*');
                if (!last || strcmp(last, closes) != 0) {
                    fclose(f);
                    free(lines);
                    return -2;
                }
                const int at_mask = strcmp(here, "*tea") == 0;
                *last = 0;
                if (at_mask) break;
            } else if (strlen(here) + w < sizeof here) {
                strcat(here, nugget);
            }
            continue;
        }
        if (strcmp(here, "*tea") != 0) continue;
        if (n == cap) { cap *= 2; lines = realloc(lines, sizeof(*lines) * cap); }
        strncpy(lines[n], s, LINE_MAX_LEN - 1);
        lines[n][LINE_MAX_LEN - 1] = 0;
        n++;
    }
    fclose(f);
    *out = lines;
    *nout = n;
    return 0;
}

/* ocl_strings.cpp:7-37 matchParam: first line (raw file, lines with '!'
 * skipped) containing the name; the text after the first '=' or the name */
static int side_match(const char *path, const char *name, char *val, size_t vlen)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    int found = 0;
    val[0] = 0;
    while (fgets(line, sizeof line, f)) {
        size_t n = strlen(line);
        if (n && line[n - 1] == '\n') line[--n] = 0;
        if (strchr(line, '!')) continue;
        if (strstr(line, name)) {
            const char *eq = strchr(line, '=');
            if (eq) snprintf(val, vlen, "%s", eq + 1);
            else snprintf(val, vlen, "%s", name);
            found = 1;
            break;
        }
    }
    fclose(f);
    return found;
}

static void side_channel(const char *path, tlo_config *cfg)
{
    char v[4096];
    cfg->sc_profiler_on = side_match(path, "profiler_on", v, sizeof v) > 0 && v[0];
    /* readInt: stringstream >> int; -1 when absent or unparsable */
    cfg->sc_halo_depth = -1;
    if (side_match(path, "halo_depth", v, sizeof v) > 0 && v[0]) {
        char *end;
        const char *s = v;
        while (*s == ' ' || *s == '\t') s++;
        long x = strtol(s, &end, 10);
        if (end != s) cfg->sc_halo_depth = (int)x;
    }
    int ppcg = side_match(path, "tl_use_ppcg", v, sizeof v) > 0 && v[0];
    int cheby = side_match(path, "tl_use_chebyshev", v, sizeof v) > 0 && v[0];
    int cg = side_match(path, "tl_use_cg", v, sizeof v) > 0 && v[0];
    cfg->sc_solver = ppcg ? TLO_SOLVER_PPCG : cheby ? TLO_SOLVER_CHEBY : cg ? TLO_SOLVER_CG : TLO_SOLVER_JACOBI;
    cfg->sc_preconditioner = TLO_PREC_NONE;
    if (side_match(path, "tl_preconditioner_type", v, sizeof v) > 0) {
        for (char *q = v; *q; q++) *q = (char)tolower((unsigned char)*q);
        if (strstr(v, "jac_diag")) cfg->sc_preconditioner = TLO_PREC_JAC_DIAG;
        else if (strstr(v, "jac_block")) cfg->sc_preconditioner = TLO_PREC_JAC_BLOCK;
        else cfg->sc_preconditioner = TLO_PREC_NONE;
    }
}

static void config_defaults(tlo_config *cfg)
{
    memset(cfg, 0, sizeof *cfg);
    /* read_input.f90:38-91; single-precision literals as Fortran stores them */
    cfg->xmin = 0.0; cfg->ymin = 0.0; cfg->xmax = 100.0; cfg->ymax = 100.0;
    cfg->x_cells = 10; cfg->y_cells = 10;
    cfg->end_time = 10.0;
    cfg->end_step = 640000; /* g_ibig */
    cfg->visit_frequency = 0;
    cfg->summary_frequency = 10;
    cfg->dtinit = (double)0.1f;
    cfg->max_iters = 1000;
    cfg->eps = (double)1.0e-10f;
    cfg->coefficient = TLO_COEF_CONDUCTIVITY;
    cfg->tl_ch_cg_presteps = 25;
    cfg->tl_ch_cg_epslim = 1.0;
    cfg->tl_check_result = 0;
    cfg->fortran_preconditioner = TLO_PREC_NONE;
    cfg->reflective_boundary = 0;
    cfg->tl_ppcg_inner_steps = -1;
    cfg->solver = TLO_SOLVER_JACOBI;
    cfg->verbose_on = 0;
    cfg->halo_depth = 1;
}

/* parse_getword (parse.f90:249-277) over the processed lines: a word "\"
 * continues on the next line; with wrap, an exhausted line continues too */
typedef struct {
    char (*lines)[LINE_MAX_LEN];
    int nl, li;
    char buf[LINE_MAX_LEN];
    char *cur;
} deck_cursor;

static int cursor_line(deck_cursor *dc)
{
    if (dc->li >= dc->nl) return 0;
    memcpy(dc->buf, dc->lines[dc->li++], LINE_MAX_LEN);
    dc->cur = dc->buf;
    return 1;
}

static const char *get_word(deck_cursor *dc, int wrap)
{
    for (;;) {
        char *s = dc->cur;
        while (*s == ' ') s++;
        if (!*s) {
            dc->cur = s;
            if (wrap && cursor_line(dc)) continue;
            return NULL;
        }
        char *w = s;
        while (*s && *s != ' ') s++;
        if (*s) *s++ = 0;
        dc->cur = s;
        if (w[0] == '\\' && w[1] == 0) {
            if (!cursor_line(dc)) return NULL;
            continue;
        }
        return w;
    }
}

int tlo_parse_deck(const char *path, tlo_config *cfg)
{
    config_defaults(cfg);
    char (*lines)[LINE_MAX_LEN];
    int nl;
    const int rl = deck_lines(path, &lines, &nl);
    if (rl) {
        fprintf(stderr, rl == -2 ? "Unmatched */*end pair in %s\n" : "Error opening %s\n", path);
        return 1;
    }
    int ok, rc = 0;
    int state_max = 0;
    deck_cursor dc = {lines, nl, 0, {0}, NULL};
    while (cursor_line(&dc)) {
        const char *w;
        while ((w = get_word(&dc, 0))) {
            if (!strcmp(w, "state")) {
                const char *v = get_word(&dc, 1);
                int s = v ? fortran_ival(v, &ok) : 0;
                if (s > state_max) state_max = s;
                break;
            }
        }
    }
    if (state_max < 1 || state_max > TLO_MAX_STATES) {
        fprintf(stderr, "read_input: No states defined.\n");
        free(lines);
        return 1;
    }
    cfg->nstates = state_max;
    int defined[TLO_MAX_STATES] = {0};
#define RVAL(dst) do { const char *v_ = get_word(&dc, 1); dst = v_ ? fortran_rval(v_, &ok) : 0; if (!v_ || !ok) { fprintf(stderr, "parse_getrval: Error attempting to convert to real:%s\n", v_ ? v_ : ""); rc = 1; } } while (0)
#define IVAL(dst) do { const char *v_ = get_word(&dc, 1); dst = v_ ? fortran_ival(v_, &ok) : 0; if (!v_ || !ok) { fprintf(stderr, "parse_getival: Error attempting to convert to integer:%s\n", v_ ? v_ : ""); rc = 1; } } while (0)
    dc.li = 0;
    while (!rc && cursor_line(&dc)) {
        const char *w;
        while (!rc && (w = get_word(&dc, 0))) {
            if (!strcmp(w, "initial_timestep")) RVAL(cfg->dtinit);
            else if (!strcmp(w, "end_time")) RVAL(cfg->end_time);
            else if (!strcmp(w, "end_step")) IVAL(cfg->end_step);
            else if (!strcmp(w, "xmin")) RVAL(cfg->xmin);
            else if (!strcmp(w, "xmax")) RVAL(cfg->xmax);
            else if (!strcmp(w, "ymin")) RVAL(cfg->ymin);
            else if (!strcmp(w, "ymax")) RVAL(cfg->ymax);
            else if (!strcmp(w, "x_cells")) IVAL(cfg->x_cells);
            else if (!strcmp(w, "y_cells")) IVAL(cfg->y_cells);
            else if (!strcmp(w, "visit_frequency")) IVAL(cfg->visit_frequency);
            else if (!strcmp(w, "summary_frequency")) IVAL(cfg->summary_frequency);
            else if (!strcmp(w, "tl_ch_cg_presteps")) IVAL(cfg->tl_ch_cg_presteps);
            else if (!strcmp(w, "tl_ppcg_inner_steps")) IVAL(cfg->tl_ppcg_inner_steps);
            else if (!strcmp(w, "tl_ch_cg_epslim")) RVAL(cfg->tl_ch_cg_epslim);
            else if (!strcmp(w, "tl_check_result")) cfg->tl_check_result = 1;
            else if (!strcmp(w, "tl_ppcg_fresh_z")) cfg->ppcg_fresh_z = 1;
            else if (!strcmp(w, "tl_preconditioner_type")) {
                const char *v;
                while ((v = get_word(&dc, 0))) {
                    if (!strcmp(v, "none")) cfg->fortran_preconditioner = TLO_PREC_NONE;
                    else if (!strcmp(v, "jac_diag")) cfg->fortran_preconditioner = TLO_PREC_JAC_DIAG;
                    else if (!strcmp(v, "jac_block")) cfg->fortran_preconditioner = TLO_PREC_JAC_BLOCK;
                    else {
                        fprintf(stderr, "read_input: Invalid preconditioner type specified: %s\n", v);
                        rc = 1;
                        break;
                    }
                }
            }
            else if (!strcmp(w, "use_opencl_kernels")) {}
            else if (!strcmp(w, "verbose_on")) cfg->verbose_on = 1;
            else if (!strcmp(w, "tl_use_jacobi")) cfg->solver = TLO_SOLVER_JACOBI;
            else if (!strcmp(w, "tl_use_cg")) cfg->solver = TLO_SOLVER_CG;
            else if (!strcmp(w, "tl_use_ppcg")) cfg->solver = TLO_SOLVER_PPCG;
            else if (!strcmp(w, "tl_use_chebyshev")) cfg->solver = TLO_SOLVER_CHEBY;
            else if (!strcmp(w, "reflective_boundary")) cfg->reflective_boundary = 1;
            else if (!strcmp(w, "profiler_on")) cfg->profiler_on = 1;
            else if (!strcmp(w, "halo_depth")) {
                IVAL(cfg->halo_depth);
                if (!rc && cfg->halo_depth < 1) {
                    fprintf(stderr, "read_input: Invalid halo exchange depth specified\n");
                    rc = 1;
                }
            }
            else if (!strcmp(w, "tl_max_iters")) IVAL(cfg->max_iters);
            else if (!strcmp(w, "tl_eps")) RVAL(cfg->eps);
            else if (!strcmp(w, "tl_coefficient_density")) cfg->coefficient = TLO_COEF_CONDUCTIVITY;
            else if (!strcmp(w, "tl_coefficient_inverse_density")) cfg->coefficient = TLO_COEF_RECIP_CONDUCTIVITY;
            else if (!strcmp(w, "test_problem")) IVAL(cfg->test_problem);
            else if (!strcmp(w, "state")) {
                int s;
                IVAL(s);
                if (rc) break;
                if (s < 1 || s > state_max) { rc = 1; break; }
                if (defined[s - 1]) {
                    fprintf(stderr, "read_input: State defined twice.\n");
                    rc = 1;
                    break;
                }
                defined[s - 1] = 1;
                const int q = s - 1;
                const char *kw;
                while (!rc && (kw = get_word(&dc, 0))) {
                    if (!strcmp(kw, "xmin")) RVAL(cfg->state_xmin[q]);
                    else if (!strcmp(kw, "ymin")) RVAL(cfg->state_ymin[q]);
                    else if (!strcmp(kw, "xmax")) RVAL(cfg->state_xmax[q]);
                    else if (!strcmp(kw, "ymax")) RVAL(cfg->state_ymax[q]);
                    else if (!strcmp(kw, "radius")) RVAL(cfg->state_radius[q]);
                    else if (!strcmp(kw, "density")) RVAL(cfg->state_density[q]);
                    else if (!strcmp(kw, "energy")) RVAL(cfg->state_energy[q]);
                    else if (!strcmp(kw, "geometry")) {
                        const char *g = get_word(&dc, 1);
                        if (g && !strcmp(g, "rectangle")) cfg->state_geometry[q] = TLO_G_RECT;
                        else if (g && !strcmp(g, "circle")) cfg->state_geometry[q] = TLO_G_CIRC;
                        else if (g && !strcmp(g, "point")) cfg->state_geometry[q] = TLO_G_POINT;
                    }
                }
                break;
            }
        }
    }
#undef RVAL
#undef IVAL
    free(lines);
    if (rc) return rc;
    /* read_input.f90:294-301 */
    if (cfg->tl_ppcg_inner_steps == -1)
        cfg->tl_ppcg_inner_steps = 4 * (int)(sqrtf(sqrtf((float)(cfg->x_cells * cfg->y_cells))));
    if (cfg->halo_depth > 1 && cfg->fortran_preconditioner == TLO_PREC_JAC_BLOCK &&
        cfg->solver == TLO_SOLVER_PPCG) {
        fprintf(stderr, "read_input: Unable to use nonstandard halo depth with block jacobi preconditioner\n");
        return 1;
    }
    /* read_input.f90:320-327: shrink states 2..n by a hundredth of a cell */
    const double dx = (cfg->xmax - cfg->xmin) / (double)(float)cfg->x_cells;
    const double dy = (cfg->ymax - cfg->ymin) / (double)(float)cfg->y_cells;
    for (int s = 1; s < cfg->nstates; s++) {
        cfg->state_xmin[s] = cfg->state_xmin[s] + (dx / 100.0);
        cfg->state_ymin[s] = cfg->state_ymin[s] + (dy / 100.0);
        cfg->state_xmax[s] = cfg->state_xmax[s] - (dx / 100.0);
        cfg->state_ymax[s] = cfg->state_ymax[s] - (dy / 100.0);
    }
    side_channel(path, cfg);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* driver: start.f90:24-115, diffuse.f90:43-101, tea_solve.f90:37-521,       */
/* field_summary.f90:27-110                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int first;                       /* SAVE first (tea_solve.f90:57) */
    double eigmin, eigmax, theta, cn;/* SAVE (tea_solve.f90:54) */
} solve_save;

static const double kat_temp[7] = {0.0, 157.550841832793, 116.067951160930, 95.4865103390698,
                                   166.838315378708, 116.482111627676, 103.88639125996923};

static void emit_summary(FILE *log, tlo_chunk *c, int step, int complete, const tlo_config *cfg)
{
    double vol, mass, ie, temp;
    tlo_field_summary(c, &vol, &mass, &ie, &temp);
    if (log) fprintf(log, "SUMMARY %d %.17e %.17e %.17e %.17e %.17e\n", step, vol, mass, mass / vol, ie, temp);
    if (complete && cfg->test_problem >= 1 && cfg->test_problem <= 6) {
        const double qa = fabs((100.0 * (temp / kat_temp[cfg->test_problem])) - 100.0);
        if (log) fprintf(log, "QA %d %.7e %s\n", cfg->test_problem, qa, qa < 0.001 ? "PASSED" : "FAILED");
    }
}

static int tea_leaf_solve(tlo_chunk *c, const tlo_config *cfg, solve_save *sv, double dt, int step,
                          FILE *lg)
{
    const int hd = c->hd;
    const int max_iters = cfg->max_iters;
    const int use_cg = cfg->solver == TLO_SOLVER_CG, use_ch = cfg->solver == TLO_SOLVER_CHEBY;
    const int use_ppcg = cfg->solver == TLO_SOLVER_PPCG, use_jac = cfg->solver == TLO_SOLVER_JACOBI;
    const int nb[4] = {-1, -1, -1, -1};
    int fields[7] = {0};
    double *cg_alphas = calloc(max_iters + 2, sizeof(double));
    double *cg_betas = calloc(max_iters + 2, sizeof(double));
    double *ch_alphas = calloc(max_iters + 2, sizeof(double));
    double *ch_betas = calloc(max_iters + 2, sizeof(double));
    double old_error, error = 0.0, exact_error = 0.0, initial_residual;
    double rro = 0.0, pw, rrn, alpha, beta;
    int est_itc = 0, cheby_calc_steps = 0, max_cheby_iters = 0, ppcg_inner_iters = 0;
    int cg_calc_steps = 0, ch_switch_check = 0, switch_n = 0;
    double rx, ry;

    /* :86-92 update_halo(density, energy1, hd) -- reflective only */
    fields[0] = 1; fields[2] = 1;
    if (cfg->reflective_boundary) tlo_update_halo(c, nb, fields, cfg->halo_depth);
    /* tea_leaf_common.f90:10-47: all faces external on one chunk */
    const int zero_boundary[4] = {1, 1, 1, 1};
    tlo_common_init(c, cfg->coefficient, dt, &rx, &ry, zero_boundary, cfg->reflective_boundary);
    /* :103-112 */
    tlo_calc_residual(c);
    initial_residual = tlo_calc_2norm(c, 1);
    old_error = initial_residual;
    initial_residual = sqrt(initial_residual);
    if (lg && cfg->verbose_on) fprintf(lg, "INITRES %d %.17e\n", step, initial_residual);

    if (use_cg || use_ch || use_ppcg) rro = tlo_cg_init(c);

    int n;
    for (n = 1; n <= max_iters; n++) {
        if (!ch_switch_check) {
            if (cheby_calc_steps > 0) ch_switch_check = 1;
            else if (!sv->first && use_ppcg && n > 1) ch_switch_check = 1;
            else if (fabs(old_error) <= cfg->tl_ch_cg_epslim && n >= cfg->tl_ch_cg_presteps) ch_switch_check = 1;
            else ch_switch_check = 0;
        }
        if ((use_ch || use_ppcg) && ch_switch_check) {
            if (cheby_calc_steps == 0) {
                max_cheby_iters = max_iters - n + 2;
                if (sv->first) {
                    int info = tlo_calc_eigenvalues(cg_alphas + 1, cg_betas + 1, n - 1, &sv->eigmin, &sv->eigmax);
                    sv->first = 0;
                    if (info != 0) {
                        fprintf(stderr, "tea_leaf: Error in calculating eigenvalues\n");
                        goto fail;
                    }
                    sv->eigmin = sv->eigmin * 0.95;
                    sv->eigmax = sv->eigmax * 1.05;
                }
                if (use_ch)
                    tlo_calc_ch_coefs(ch_alphas, ch_betas, sv->eigmin, sv->eigmax, &sv->theta, max_cheby_iters);
                else
                    tlo_calc_ch_coefs(ch_alphas, ch_betas, sv->eigmin, sv->eigmax, &sv->theta,
                                      cfg->tl_ppcg_inner_steps);
                sv->cn = sv->eigmax / sv->eigmin;
                switch_n = n;
                if (lg) fprintf(lg, "SWITCH %d %d %.17e %.17e %.17e %.17e %.17e\n", step, n, rro,
                                 sv->eigmin, sv->eigmax, sv->cn, old_error);
            }
            if (use_ch) {
                if (cheby_calc_steps == 0) {
                    /* tea_leaf_cheby_first_step (:477-521); its `error` dummy is old_error */
                    const double bb = tlo_calc_2norm(c, 0);
                    tlo_cheby_init(c, ch_alphas, ch_betas, max_cheby_iters, rx, ry, sv->theta);
                    tlo_cheby_iterate(c, 1);
                    old_error = tlo_calc_2norm(c, 1);
                    const double it_alpha = cfg->eps * bb / (4.0 * old_error);
                    const double gamm = (sqrt(sv->cn) - 1.0) / (sqrt(sv->cn) + 1.0);
                    est_itc = (int)lround(log(it_alpha) / (2.0 * log(gamm)));
                    if (lg) fprintf(lg, "ESTITC %d %d\n", step, est_itc);
                    cheby_calc_steps = 1;
                } else {
                    tlo_cheby_iterate(c, cheby_calc_steps);
                    if (n >= est_itc && n % 10 == 0) error = tlo_calc_2norm(c, 1);
                }
            } else {
                if (cheby_calc_steps == 0)
                    tlo_ppcg_init(c, ch_alphas, ch_betas, sv->theta, cfg->tl_ppcg_inner_steps);
                pw = tlo_cg_calc_w(c);
                alpha = rro / pw;
                rrn = tlo_cg_calc_ur(c, alpha);
                /* tea_leaf_run_ppcg_inner_steps (:422-475) */
                tlo_ppcg_init_sd(c);
                for (int cur = 1; cur <= cfg->tl_ppcg_inner_steps; cur += cfg->halo_depth) {
                    int inner_step = cur;
                    for (int be = cfg->halo_depth - 1; be >= 0; be--) {
                        tlo_ppcg_inner(c, inner_step, be, nb);
                        inner_step++;
                        if (inner_step > cfg->tl_ppcg_inner_steps) break;
                    }
                }
                ppcg_inner_iters += cfg->tl_ppcg_inner_steps;
                if (cfg->ppcg_fresh_z) tlo_precondition_r(c);
                rrn = tlo_calc_2norm(c, 2);
                beta = rrn / rro;
                tlo_cg_calc_p(c, beta);
                error = rrn;
                rro = rrn;
            }
            cheby_calc_steps++;
        } else if (use_cg || use_ch || use_ppcg) {
            cg_calc_steps++;
            pw = tlo_cg_calc_w(c);
            alpha = rro / pw;
            cg_alphas[n] = alpha;
            rrn = tlo_cg_calc_ur(c, alpha);
            beta = rrn / rro;
            cg_betas[n] = beta;
            tlo_cg_calc_p(c, beta);
            error = rrn;
            rro = rrn;
        } else if (use_jac) {
            error = tlo_jacobi_solve(c);
        }
        error = sqrt(error);
        if (lg && cfg->verbose_on) fprintf(lg, "RESID %d %d %.17e\n", step, n, error);
        if (fabs(error) < cfg->eps * initial_residual) break;
        old_error = error;
    }
    if (cfg->tl_check_result) {
        tlo_calc_residual(c);
        exact_error = sqrt(tlo_calc_2norm(c, 1));
    }
    if (lg)
        fprintf(lg, "SOLVE %d %d %.17e %.17e %d %d %d %d %d\n", step, n - 1, error / initial_residual,
                cfg->tl_check_result ? exact_error / initial_residual : NAN, ppcg_inner_iters, cg_calc_steps,
                cheby_calc_steps, switch_n, est_itc);
    tlo_finalise(c);
    for (int i = 0; i < 7; i++) fields[i] = 0;
    fields[2] = 1;
    if (cfg->reflective_boundary) tlo_update_halo(c, nb, fields, 1);
    free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
    (void)hd;
    return 0;
fail:
    free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
    return 1;
}

int tlo_run_config(const tlo_config *cfg, const char *log_path, int max_steps)
{
    if (cfg->sc_halo_depth < 1) {
        fprintf(stderr, "Halo exchange depth unspecified or was too small\n");
        return 1;
    }
    FILE *log = NULL;
    if (log_path && log_path[0]) {
        log = fopen(log_path, "w");
        if (!log) return 1;
    }
    const int hd = cfg->sc_halo_depth;
    tlo_chunk *c = tlo_create(cfg->x_cells, cfg->y_cells, hd, cfg->sc_preconditioner);
    if (!c) { if (log) fclose(log); return 1; }
    /* start.f90:24-115 */
    const double dx = (cfg->xmax - cfg->xmin) / (double)(float)cfg->x_cells;
    const double dy = (cfg->ymax - cfg->ymin) / (double)(float)cfg->y_cells;
    tlo_initialise_chunk(c, cfg->xmin, cfg->ymin, dx, dy);
    tlo_generate_chunk(c, cfg->nstates, cfg->state_density, cfg->state_energy, cfg->state_xmin,
                       cfg->state_xmax, cfg->state_ymin, cfg->state_ymax, cfg->state_radius,
                       cfg->state_geometry, TLO_G_RECT, TLO_G_CIRC, TLO_G_POINT);
    const int nb[4] = {-1, -1, -1, -1};
    int fields[7] = {1, 1, 1, 0, 0, 0, 0};
    if (cfg->reflective_boundary) tlo_update_halo(c, nb, fields, cfg->halo_depth);
    tlo_set_field(c);
    emit_summary(log, c, 0, 0, cfg);
    /* diffuse.f90:43-101 */
    solve_save sv = {1, 0, 0, 0, 0};
    double time = 0.0, dt = cfg->dtinit;
    const double g_small = (double)1.0e-16f;
    int step = 0, rc = 0;
    for (;;) {
        step++;
        if (cfg->dtinit <= dt) dt = cfg->dtinit; /* timestep.f90:45-53 */
        if (tea_leaf_solve(c, cfg, &sv, dt, step, log)) { rc = 1; break; }
        time = time + dt;
        int done = (time + g_small > cfg->end_time || step >= cfg->end_step);
        if (max_steps > 0 && step >= max_steps) done = 1;
        if (cfg->summary_frequency != 0 && step % cfg->summary_frequency == 0)
            emit_summary(log, c, step, 0, cfg);
        if (done) {
            emit_summary(log, c, step, 1, cfg);
            break;
        }
    }
    if (log) fclose(log);
    tlo_destroy(c);
    return rc;
}

int tlo_parse_deck_values(const char *path, int *ints, double *reals, double *states, int max_states)
{
    tlo_config cfg;
    const int rc = tlo_parse_deck(path, &cfg);
    if (rc) return rc;
    const int iv[23] = {cfg.x_cells, cfg.y_cells, cfg.end_step, cfg.summary_frequency, cfg.visit_frequency,
                        cfg.max_iters, cfg.coefficient, cfg.tl_ch_cg_presteps, cfg.tl_check_result,
                        cfg.tl_ppcg_inner_steps, cfg.reflective_boundary, cfg.solver, cfg.fortran_preconditioner,
                        cfg.verbose_on, cfg.profiler_on, cfg.halo_depth, cfg.test_problem, cfg.ppcg_fresh_z,
                        cfg.nstates, cfg.sc_preconditioner, cfg.sc_halo_depth, cfg.sc_solver, cfg.sc_profiler_on};
    const double rv[8] = {cfg.xmin, cfg.ymin, cfg.xmax, cfg.ymax, cfg.end_time, cfg.dtinit, cfg.eps,
                          cfg.tl_ch_cg_epslim};
    memcpy(ints, iv, sizeof iv);
    memcpy(reals, rv, sizeof rv);
    for (int s = 0; s < max_states && s < cfg.nstates; s++) {
        const double sv[8] = {cfg.state_density[s], cfg.state_energy[s], cfg.state_xmin[s], cfg.state_xmax[s],
                              cfg.state_ymin[s], cfg.state_ymax[s], cfg.state_radius[s],
                              (double)cfg.state_geometry[s]};
        memcpy(states + 8 * s, sv, sizeof sv);
    }
    return 0;
}

int tlo_run_deck(const char *deck_path, const char *log_path, int max_steps)
{
    tlo_config cfg;
    if (tlo_parse_deck(deck_path, &cfg)) return 1;
    return tlo_run_config(&cfg, log_path, max_steps);
}

double tlo_time_cg_iterations(const char *deck_path, int n_iters, int *iters_done)
{
    tlo_config cfg;
    *iters_done = 0;
    if (tlo_parse_deck(deck_path, &cfg) || cfg.sc_halo_depth < 1) return -1.0;
    tlo_chunk *c = tlo_create(cfg.x_cells, cfg.y_cells, cfg.sc_halo_depth, cfg.sc_preconditioner);
    const double dx = (cfg.xmax - cfg.xmin) / (double)(float)cfg.x_cells;
    const double dy = (cfg.ymax - cfg.ymin) / (double)(float)cfg.y_cells;
    tlo_initialise_chunk(c, cfg.xmin, cfg.ymin, dx, dy);
    tlo_generate_chunk(c, cfg.nstates, cfg.state_density, cfg.state_energy, cfg.state_xmin,
                       cfg.state_xmax, cfg.state_ymin, cfg.state_ymax, cfg.state_radius,
                       cfg.state_geometry, TLO_G_RECT, TLO_G_CIRC, TLO_G_POINT);
    tlo_set_field(c);
    const int zb[4] = {1, 1, 1, 1};
    double rx, ry;
    tlo_common_init(c, cfg.coefficient, cfg.dtinit, &rx, &ry, zb, cfg.reflective_boundary);
    tlo_calc_residual(c);
    (void)tlo_calc_2norm(c, 1);
    double rro = tlo_cg_init(c);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int n = 0; n < n_iters; n++) {
        const double pw = tlo_cg_calc_w(c);
        const double alpha = rro / pw;
        const double rrn = tlo_cg_calc_ur(c, alpha);
        tlo_cg_calc_p(c, rrn / rro);
        rro = rrn;
        (*iters_done)++;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    tlo_destroy(c);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
