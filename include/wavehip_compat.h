/*
 * wavehip_compat.h -- drop-in evolve_wave_half / evolve_wave for a reference
 * program (nilsberglund-orleans/YouTube-simulations).
 *
 * Include it in the program's translation unit after global_pdes.c (for
 * t_neumann_bc, N_NEUMANN_POINTS and the BC_* codes) and after the program's
 * #define block, in place of the program's own evolve_wave_half and
 * evolve_wave (wave_billiard.c:302-631, or wave_3d.c:600-1031 with
 * WAVEHIP_FLAT defined), and link with -lwavehip.  The functions keep the
 * reference signatures and semantics: each call uploads the caller's arrays,
 * advances one (or three) leapfrog steps on the GPU and writes the result
 * back, bit-identical to the reference stepper.
 *
 * The macros the reference stepper reads become runtime parameters of one
 * context per program, created on the first call:
 *   NX, NY, B_COND, TWOSPEEDS, KAPPA, KAPPA_SIDES, GAMMA_SIDES, KAPPA_TOPBOT,
 *   GAMMA_TOPBOT, OSCILLATE_TOPBOT (tb_shift as wave_billiard.c:319, or
 *   wave_3d.c:613-617 with WAVEHIP_FLAT), E_SCALE if defined.
 * Define WAVEHIP_TABLES_FIXED when tcc / tc / tgamma / xy_in do not change
 * between calls (no REFRESH_IOR): they are then uploaded once.
 *
 * BC_NEUMANN (wave_billiard.c:379-389): the program's n_neumann_points
 * active points of neumann_bc are handed to the library on the first call
 * (with WAVEHIP_TABLES_FIXED) or on every call.  OSCILLATE_LEFT
 * (wave_billiard.c:320-341, 397, 585-590): the left-edge cells and corners
 * the reference drives are set, step by step, to the program's own
 * oscillating_bc(time, j), with the step counter of the reference (the
 * flat program drives the whole left edge and its corners, wave_3d.c:651-654,
 * 989-993); the corner (0, NX-1) of the row-pointer reference lies outside
 * row 0 when NX > NY and is then left out.  FLOOR (wave_billiard.c:592-609)
 * clamps to VMAX.
 *
 * PRECOMPUTE_BC (with WAVEHIP_FLAT, wave_3d.c:173): evolve_wave steps with
 * evolve_wave_half_new (wave_3d.c:345-597), which this header also defines,
 * on the neighbour tables the program built with init_laplacian_coords
 * (sub_wave.c:11555): laplace for phi, laplace1 for tmp, laplace2 for psi,
 * as wave_3d.c:1187-1189 builds them.  They are handed to the library as
 * indices (wv_set_laplacian) on the first call with WAVEHIP_TABLES_FIXED, on
 * every call without it.  B_COND may then be BC_LSHAPE.  The left-edge drive
 * follows that stepper: cells between bc_jmin and bc_jmax (BEAM_BC) and the
 * corners (0, 0), (0, NY-1) at AMPLITUDE*cos(time*OMEGA), wave_3d.c:358-376,
 * 575-579.
 *
 * COMPARISON (with WAVEHIP_FLAT, wave_3d.c:736-990): the grid steps as two
 * halves (WV_FLAG_COMPARISON); with OSCILLATE_LEFT the interface cells
 * (0, NY/2-1), (0, NY/2) keep their interface formula where it applies, as
 * in the reference.
 *
 * Faster path for the frame loop (wave_billiard.c:897-899): keep the state on
 * the GPU and call wv_step(wavehip_ctx(), 3*NVID) plus one wv_get_state per
 * frame instead of NVID calls to evolve_wave.
 */
#ifndef WAVEHIP_COMPAT_H
#define WAVEHIP_COMPAT_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wavehip.h"

/* the neighbour-table stepper of wave_3d.c (PRECOMPUTE_BC) */
#if defined(WAVEHIP_FLAT) && defined(PRECOMPUTE_BC)
#define WAVEHIP_TABLE_PATH PRECOMPUTE_BC
#else
#define WAVEHIP_TABLE_PATH 0
#endif

#ifndef E_SCALE
#define WAVEHIP_E_SCALE 1.0
#else
#define WAVEHIP_E_SCALE E_SCALE
#endif

static void wavehip_die(wv_ctx *c, const char *what, int rc)
{
    fprintf(stderr, "wavehip: %s failed (%d): %s\n", what, rc, wv_last_error(c));
    abort();
}

/* the program's context, created from its macros on first use */
static wv_ctx *wavehip_ctx(void)
{
    static wv_ctx *ctx = NULL;
    if (!ctx) {
        wv_params p;
        int rc;
        wv_params_default(&p, NX, NY);
#if defined(COMPARISON) && COMPARISON && defined(WAVEHIP_FLAT)
        p.flags |= WV_FLAG_COMPARISON; /* the flat stepper's two halves, wave_3d.c:736-990 */
#endif
        p.bcond = B_COND;
        p.twospeeds = TWOSPEEDS;
        p.kappa = KAPPA;
        p.kappa_sides = KAPPA_SIDES;
        p.gamma_sides = GAMMA_SIDES;
        p.kappa_topbot = KAPPA_TOPBOT;
        p.gamma_topbot = GAMMA_TOPBOT;
        p.oscillate_topbot = OSCILLATE_TOPBOT;
        p.e_scale = WAVEHIP_E_SCALE;
        p.tb_shift = (int)((XMAX - XMIN) * (double)NX / (XMAX - XMIN));
#if defined(WAVEHIP_FLAT) && !WAVEHIP_TABLE_PATH
        if (p.tb_shift >= NX - 1) p.tb_shift = NY - 2;
#endif
#ifdef WAVEHIP_TABLES_FIXED
        p.flags |= WV_FLAG_TABLES_FIXED;
#endif
        rc = wv_create(&ctx, &p);
        if (rc) wavehip_die(ctx, "wv_create", rc);
#if FLOOR
        rc = wv_set_floor(ctx, VMAX);
        if (rc) wavehip_die(ctx, "wv_set_floor", rc);
#endif
    }
    return ctx;
}

#if BC_NEUMANN
/* the program's active Neumann points (init_neumann_bc, sub_wave.c:13910) */
static void wavehip_neumann(const t_neumann_bc *nb)
{
    static int done = 0;
    int n = 0, q, rc;
    int *a;
    double *w;
#ifdef WAVEHIP_TABLES_FIXED
    if (done) return;
#endif
    a = (int *)malloc(sizeof(int) * 6 * (n_neumann_points > 0 ? n_neumann_points : 1));
    w = (double *)malloc(sizeof(double) * 2 * (n_neumann_points > 0 ? n_neumann_points : 1));
    if (!a || !w) wavehip_die(wavehip_ctx(), "malloc", -1);
    for (q = 0; q < n_neumann_points; q++)
        if (nb[q].active) {
#ifndef WAVEHIP_FLAT
            a[n] = nb[q].i;
            a[n_neumann_points + n] = nb[q].j;
            a[2 * n_neumann_points + n] = nb[q].nbi1;
            a[3 * n_neumann_points + n] = nb[q].nbj1;
            a[4 * n_neumann_points + n] = nb[q].nbi2;
            a[5 * n_neumann_points + n] = nb[q].nbj2;
#else
            /* wave_3d.c:640-644 indexes the flat arrays with k, nbk1, nbk2 */
            a[n] = nb[q].k / NY;
            a[n_neumann_points + n] = nb[q].k % NY;
            a[2 * n_neumann_points + n] = nb[q].nbk1 / NY;
            a[3 * n_neumann_points + n] = nb[q].nbk1 % NY;
            a[4 * n_neumann_points + n] = nb[q].nbk2 / NY;
            a[5 * n_neumann_points + n] = nb[q].nbk2 % NY;
#endif
            w[n] = nb[q].nx;
            w[n_neumann_points + n] = nb[q].ny;
            n++;
        }
    rc = wv_set_neumann(wavehip_ctx(), n, a, a + n_neumann_points, a + 2 * n_neumann_points,
                        a + 3 * n_neumann_points, a + 4 * n_neumann_points, a + 5 * n_neumann_points, w,
                        w + n_neumann_points);
    free(a);
    free(w);
    if (rc) wavehip_die(wavehip_ctx(), "wv_set_neumann", rc);
    done = 1;
    (void)done;
}
#endif

#if WAVEHIP_TABLE_PATH
/* the program's neighbour tables as counts and flat indices: laplace was
 * built for phi, laplace1 for tmp, laplace2 for psi (wave_3d.c:1187-1189);
 * the three must describe the same neighbours */
static void wavehip_laplacian(const t_laplacian *lap, const double *base, const t_laplacian *lap1,
                              const double *base1, const t_laplacian *lap2, const double *base2)
{
    static int done = 0;
    static int *nn = NULL, *nb = NULL;
    long k;
    int q, rc;
#ifdef WAVEHIP_TABLES_FIXED
    if (done) return;
#endif
    if (!nn) {
        nn = (int *)malloc(sizeof(int) * (size_t)NX * NY);
        nb = (int *)malloc(sizeof(int) * 4 * (size_t)NX * NY);
        if (!nn || !nb) wavehip_die(wavehip_ctx(), "malloc", -1);
    }
    for (k = 0; k < (long)NX * NY; k++) {
        const int n = lap[k].nneighb;
        nn[k] = n;
        for (q = 0; q < 4; q++) {
            /* entries the stepper reads point into the field; others may not */
            const double *a = q < n && n <= 4 ? lap[k].nghb[q] : NULL;
            const long o = a ? (long)(a - base) : -1;
            nb[4 * k + q] = (o >= 0 && o < (long)NX * NY) ? (int)o : -1;
            if (lap1 && a) {
                const double *b = lap1[k].nghb[q], *c = lap2[k].nghb[q];
                if (lap1[k].nneighb != n || lap2[k].nneighb != n || b - base1 != o || c - base2 != o) {
                    fprintf(stderr, "wavehip: laplace, laplace1, laplace2 differ at cell %ld\n", k);
                    abort();
                }
            }
        }
    }
    rc = wv_set_laplacian(wavehip_ctx(), nn, nb);
    if (rc) wavehip_die(wavehip_ctx(), "wv_set_laplacian", rc);
    done = 1;
    (void)done;
}
#endif

#if OSCILLATE_LEFT
/* the driven cells of the next nsteps steps, values oscillating_bc(time, j):
 * row-pointer program: (0, j) for bc_jmin < j < bc_jmax, 0 < j < NY-1
 * (wave_billiard.c:320-341, 397), then the corners (0, 0) and (0, NX-1)
 * unless BEAM_BC (585-590); flat program: (0, j) for 0 < j < NY-1 and the
 * corners (0, 0), (0, NY-1) (wave_3d.c:651-654, 989-993); neighbour-table
 * stepper: (0, j) for bc_jmin < j < bc_jmax, 0 < j < NY-1, then the corners
 * (0, 0), (0, NY-1) at AMPLITUDE*cos(time*OMEGA) (wave_3d.c:358-376, 413,
 * 575-579) */
static void wavehip_drive(int nsteps, const short *xy0)
{
    (void)xy0;
    static long time = 0;  /* the reference's step counter, wave_billiard.c:310,315 */
    static int cells[NY + 2], ncell = -1;
    static double vals[(NY + 2) * 3];
    int q, s, rc, bc_jmin = 0, bc_jmax = NY - 1, ij[2];
    int ii[NY + 2];
    if (ncell < 0) {
#if WAVEHIP_TABLE_PATH
        if (BEAM_BC) {
            xy_to_ij(0.0, OSCIL_YMID + OSCIL_YMAX, ij);
            bc_jmax = ij[1];
            xy_to_ij(0.0, OSCIL_YMID - OSCIL_YMAX, ij);
            bc_jmin = ij[1];
        }
        ncell = 0;
        for (q = 1; q < NY - 1; q++)
            if (q > bc_jmin && q < bc_jmax) cells[ncell++] = q;
        cells[ncell++] = 0;
        cells[ncell++] = NY - 1;
#elif defined(WAVEHIP_FLAT)
        ncell = 0;
        for (q = 1; q < NY - 1; q++) {
#if defined(COMPARISON) && COMPARISON
            /* the interface rows overwrite the drive where they apply (wave_3d.c:873-982) */
            if ((q == NY / 2 - 1 || q == NY / 2) && (TWOSPEEDS || xy0[q] != 0)) continue;
#endif
            cells[ncell++] = q;
        }
        cells[ncell++] = 0;
        cells[ncell++] = NY - 1;
        (void)bc_jmin; (void)bc_jmax; (void)ij;
#else
        if (BEAM_BC) {
            xy_to_ij(0.0, OSCIL_YMID + OSCIL_YMAX, ij);
            bc_jmax = ij[1];
            xy_to_ij(0.0, OSCIL_YMID - OSCIL_YMAX, ij);
            bc_jmin = ij[1];
        }
        ncell = 0;
        for (q = 1; q < NY - 1; q++)
            if (q > bc_jmin && q < bc_jmax) cells[ncell++] = q;
        if (!BEAM_BC) {
            cells[ncell++] = 0;
            if (NX - 1 < NY) cells[ncell++] = NX - 1;
        }
#endif
    }
    for (q = 0; q < ncell; q++) {
        ii[q] = 0;
        for (s = 0; s < nsteps; s++) {
#if WAVEHIP_TABLE_PATH
            if (q >= ncell - 2) {
                vals[q * nsteps + s] = AMPLITUDE * cos((double)(time + s + 1) * OMEGA);
                continue;
            }
#endif
            vals[q * nsteps + s] = oscillating_bc((int)(time + s + 1), cells[q]);
        }
    }
    time += nsteps;
    rc = wv_set_drive(wavehip_ctx(), ncell, ii, cells, vals, nsteps);
    if (rc) wavehip_die(wavehip_ctx(), "wv_set_drive", rc);
}
#endif

#ifndef WAVEHIP_FLAT
/* wave_billiard.c:302-304 */
static void evolve_wave_half(double *phi_in[NX], double *psi_in[NX], double *phi_out[NX],
                             short int *xy_in[NX], double *tcc[NX], double *tc[NX],
                             double *tgamma[NX], t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc;
#if BC_NEUMANN
    wavehip_neumann(neumann_bc);
#endif
#if OSCILLATE_LEFT
    wavehip_drive(1, xy_in[0]);
#endif
    rc = wv_evolve_wave_half_rows(wavehip_ctx(), phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave_half", rc);
}

/* wave_billiard.c:613 */
static void evolve_wave(double *phi[NX], double *psi[NX], double *tmp[NX], short int *xy_in[NX],
                        double *tcc_table[NX], double *tc_table[NX], double *tgamma_table[NX],
                        t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc;
#if BC_NEUMANN
    wavehip_neumann(neumann_bc);
#endif
#if OSCILLATE_LEFT
    wavehip_drive(3, xy_in[0]);
#endif
    rc = wv_evolve_wave_rows(wavehip_ctx(), phi, psi, tmp, xy_in, tcc_table, tc_table, tgamma_table);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave", rc);
}
#else
/* wave_3d.c:600 */
static void evolve_wave_half(double phi_in[NX * NY], double psi_in[NX * NY], double phi_out[NX * NY],
                             short int xy_in[NX * NY], double tc[NX * NY], double tcc[NX * NY],
                             double tgamma[NX * NY], t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc;
#if WAVEHIP_TABLE_PATH
    fprintf(stderr, "wavehip: with PRECOMPUTE_BC the program steps with evolve_wave_half_new\n");
    abort();
#endif
#if BC_NEUMANN
    wavehip_neumann(neumann_bc);
#endif
#if OSCILLATE_LEFT
    wavehip_drive(1, xy_in);
#endif
    rc = wv_evolve_wave_half_flat(wavehip_ctx(), phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave_half", rc);
}

#if WAVEHIP_TABLE_PATH
/* wave_3d.c:345 */
static void evolve_wave_half_new(double phi_in[NX * NY], double psi_in[NX * NY], double phi_out[NX * NY],
                                 short int xy_in[NX * NY], double tc[NX * NY], double tcc[NX * NY],
                                 double tgamma[NX * NY], t_laplacian laplace[NX * NY])
{
    int rc;
    wavehip_laplacian(laplace, phi_in, NULL, NULL, NULL, NULL);
#if OSCILLATE_LEFT
    wavehip_drive(1, xy_in);
#endif
    rc = wv_evolve_wave_half_flat(wavehip_ctx(), phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma);
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave_half_new", rc);
}
#endif

/* wave_3d.c:1010 (PRECOMPUTE_BC 0: the t_laplacian tables are not read;
 * PRECOMPUTE_BC 1: the three steps of evolve_wave_half_new) */
static void evolve_wave(double phi[NX * NY], double psi[NX * NY], double tmp[NX * NY],
                        short int xy_in[NX * NY], double tc[NX * NY], double tcc[NX * NY],
                        double tgamma[NX * NY], t_laplacian laplace[NX * NY],
                        t_laplacian laplace1[NX * NY], t_laplacian laplace2[NX * NY],
                        t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc;
#if WAVEHIP_TABLE_PATH
    wavehip_laplacian(laplace, phi, laplace1, tmp, laplace2, psi);
#elif BC_NEUMANN
    wavehip_neumann(neumann_bc);
#endif
#if OSCILLATE_LEFT
    wavehip_drive(3, xy_in);
#endif
    rc = wv_evolve_wave_flat(wavehip_ctx(), phi, psi, tmp, xy_in, tc, tcc, tgamma);
    (void)laplace; (void)laplace1; (void)laplace2; (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave", rc);
}
#endif


#ifdef WAVEHIP_INIT
/* ---- start-up tables on the GPU (define WAVEHIP_INIT) ------------------------
 * Call wavehip_init_wave_flat / wavehip_init_ior_2d where the program calls
 * init_wave_flat (wave_common.c:294-319, wave_billiard.c:797) and init_ior_2d
 * (sub_wave.c:12164, wave_billiard.c:803 and the REFRESH_IOR call at 943), or
 * wavehip_init_wave_flat_mod for init_wave_flat_mod (wave_common.c:1894,
 * wave_3d.c:1212, with WAVEHIP_FLAT); same arguments.  They read the
 * program's #define block (B_DOMAIN, LAMBDA, MU, MU_B, APOLY, NPOLY,
 * WALL_WIDTH, WALL_WIDTH_B, JULIA_SCALE, MDEPTH, MRATIO, MANDELLEVEL,
 * MANDELLIMIT, NGRIDX, NGRIDY, WALL_WIDTH_ASYM, WALL_WIDTH_ASYM_B, OSCIL_YMID,
 * XMIN..YMAX, XYIN_INITIALISED; VARIABLE_IOR,
 * IOR, COURANT, COURANTB, GAMMA, GAMMAB, MANDEL_IOR_SCALE; COMPARISON and
 * B_DOMAIN_B) and its globals (circles, ncircles, circles_b, ncircles_b,
 * julia_x, julia_y, michelson_position; for a list-driven B_DOMAIN, the lists it reads:
 * polyrect, npolyrect, polyrectrot, npolyrect_rot, polygons, polyline,
 * npolyline).  The random wells (IOR 5, 8, 9) are drawn with the
 * program's own gaussian() and init_circle_config_pattern(), in the
 * reference's order, so the draws are the program's.  Masks equal the
 * reference's; the tables too, up to the GPU's exp, tanh and log where the
 * model uses them. */
static const double wavehip_box[4] = {XMIN, XMAX, YMIN, YMAX};

/* the lists of a list-driven B_DOMAIN (wv_build_list_domain_mask): only
 * those the domain reads are taken from the program's globals (polyrect,
 * polyrectrot, polygons with ncircles of them, polyline; global_pdes.c:
 * 533-545), as its init functions left them */
#define WAVEHIP_D_(d) (B_DOMAIN == (d))
#define WAVEHIP_RECTS (WAVEHIP_D_(53) || WAVEHIP_D_(531) || WAVEHIP_D_(54) || WAVEHIP_D_(541) || WAVEHIP_D_(192) || \
                       WAVEHIP_D_(193) || WAVEHIP_D_(194) || WAVEHIP_D_(931) || WAVEHIP_D_(932) || WAVEHIP_D_(933) || \
                       WAVEHIP_D_(941))
#define WAVEHIP_RECTROT (WAVEHIP_D_(95) || WAVEHIP_D_(96) || WAVEHIP_D_(97) || WAVEHIP_D_(932) || WAVEHIP_D_(933) || \
                         WAVEHIP_D_(941))
#define WAVEHIP_POLYGONS WAVEHIP_D_(40)
#define WAVEHIP_POLYLINE (WAVEHIP_D_(41) || WAVEHIP_D_(42) || WAVEHIP_D_(421) || WAVEHIP_D_(27))
#define WAVEHIP_LISTS (WAVEHIP_RECTS || WAVEHIP_RECTROT || WAVEHIP_POLYGONS || WAVEHIP_POLYLINE)

#if WAVEHIP_LISTS && !COMPARISON
static int wavehip_list_mask(const wv_domain_params *dp, int nc, const double *cl, short *flat)
{
    wv_domain_lists L;
    memset(&L, 0, sizeof L);
    L.ncircles = nc;
    L.circles = cl;
    double *r = NULL, *rr = NULL, *pg = NULL, *pl = NULL;
#if WAVEHIP_RECTS
    const int nr = npolyrect;
    r = (double *)malloc(sizeof(double) * 4 * (size_t)(nr > 0 ? nr : 1));
    if (!r) wavehip_die(wavehip_ctx(), "mask (no memory)", WV_ENOMEM);
    for (int k = 0; k < nr; k++) {
        r[k] = polyrect[k].x1; r[nr + k] = polyrect[k].y1; r[2 * nr + k] = polyrect[k].x2; r[3 * nr + k] = polyrect[k].y2;
    }
    L.nrects = nr;
    L.rects = r;
#endif
#if WAVEHIP_RECTROT
    const int nrr = npolyrect_rot;
    rr = (double *)malloc(sizeof(double) * 5 * (size_t)(nrr > 0 ? nrr : 1));
    if (!rr) wavehip_die(wavehip_ctx(), "mask (no memory)", WV_ENOMEM);
    for (int k = 0; k < nrr; k++) {
        rr[k] = polyrectrot[k].x1; rr[nrr + k] = polyrectrot[k].y1; rr[2 * nrr + k] = polyrectrot[k].x2;
        rr[3 * nrr + k] = polyrectrot[k].y2; rr[4 * nrr + k] = polyrectrot[k].width;
    }
    L.nrectrot = nrr;
    L.rectrot = rr;
#endif
#if WAVEHIP_POLYGONS
    const int np = ncircles;
    pg = (double *)malloc(sizeof(double) * 6 * (size_t)(np > 0 ? np : 1));
    if (!pg) wavehip_die(wavehip_ctx(), "mask (no memory)", WV_ENOMEM);
    for (int k = 0; k < np; k++) {
        pg[k] = polygons[k].xc; pg[np + k] = polygons[k].yc; pg[2 * np + k] = polygons[k].radius;
        pg[3 * np + k] = polygons[k].angle; pg[4 * np + k] = (double)polygons[k].nsides;
        pg[5 * np + k] = polygons[k].active ? 1.0 : 0.0;
    }
    L.npolygons = np;
    L.polygons = pg;
#endif
#if WAVEHIP_POLYLINE
    const int npl = npolyline;
    pl = (double *)malloc(sizeof(double) * 2 * (size_t)(npl > 0 ? npl : 1));
    if (!pl) wavehip_die(wavehip_ctx(), "mask (no memory)", WV_ENOMEM);
    for (int k = 0; k < npl; k++) {
        pl[k] = polyline[k].x;
        pl[npl + k] = polyline[k].y;
    }
    L.npolyline = npl;
    L.polyline = pl;
#endif
    const int rc = wv_build_list_domain_mask(wavehip_ctx(), B_DOMAIN, wavehip_box, dp, &L, flat);
    free(r);
    free(rr);
    free(pg);
    free(pl);
    return rc;
}
#endif

static void wavehip_mask(short *flat)
{
    wv_domain_params dp;
    dp.lambda = LAMBDA; dp.mu = MU; dp.mu_b = MU_B; dp.apoly = APOLY; dp.wall_width = WALL_WIDTH;
    dp.wall_width_b = WALL_WIDTH_B; dp.julia_scale = JULIA_SCALE; dp.julia_x = julia_x; dp.julia_y = julia_y;
    dp.mandellimit = MANDELLIMIT; dp.npoly = NPOLY; dp.mdepth = MDEPTH; dp.mratio = MRATIO;
    dp.mandellevel = MANDELLEVEL; dp.ngridx = NGRIDX; dp.ngridy = NGRIDY;
    dp.wall_width_asym = WALL_WIDTH_ASYM; dp.wall_width_asym_b = WALL_WIDTH_ASYM_B; dp.oscil_ymid = OSCIL_YMID;
    dp.michelson_position = michelson_position;
    const int nl = COMPARISON ? 2 : 1;
    const t_circle *cs[2] = {circles, COMPARISON ? circles_b : circles};
    const int nc[2] = {ncircles, COMPARISON ? ncircles_b : 0};
    double *cl[2] = {NULL, NULL};
    for (int q = 0; q < nl; q++) {
        const int n = nc[q] > 0 ? nc[q] : 0;
        cl[q] = (double *)malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
        if (!cl[q]) wavehip_die(wavehip_ctx(), "mask (no memory)", WV_ENOMEM);
        for (int k = 0; k < n; k++) {
            cl[q][k] = cs[q][k].xc;
            cl[q][n + k] = cs[q][k].yc;
            cl[q][2 * n + k] = cs[q][k].radius;
            cl[q][3 * n + k] = cs[q][k].active ? 1.0 : 0.0;
        }
    }
    int rc;
#if COMPARISON
    rc = wv_build_comparison_mask(wavehip_ctx(), B_DOMAIN, B_DOMAIN_B, wavehip_box, &dp, nc[0], cl[0], nc[1], cl[1],
                                  flat);
#elif WAVEHIP_LISTS
    rc = wavehip_list_mask(&dp, nc[0], cl[0], flat);
#else
    rc = wv_build_domain_mask(wavehip_ctx(), B_DOMAIN, wavehip_box, &dp, nc[0], cl[0], flat);
#endif
    free(cl[0]);
    free(cl[1]);
    if (rc) wavehip_die(wavehip_ctx(), "xy_in_billiard", rc);
}

/* init_ior_2d's tables of the flat mask xy into flat tcc, tc, tgamma */
static void wavehip_media(const short *xy, double ior_angle, double *tcc, double *tc, double *tg)
{
    int rc;
#if VARIABLE_IOR && IOR >= 4 && IOR <= 9
    /* the well models, sub_wave.c:12315-12555: the centres (and heights) as
     * the reference places them, the sum on the GPU */
    int nw = NGRIDX * NGRIDY, use_tanh = 0;
    double sigma, *xc, *yc, *h = NULL;
#if IOR == 8 || IOR == 9
    static t_circle wc[NMAXCIRCLES];
    nw = init_circle_config_pattern(wc, IOR == 8 ? C_POISSON_DISC : C_RAND_POISSON);
#endif
    xc = (double *)malloc(sizeof(double) * (size_t)nw);
    yc = (double *)malloc(sizeof(double) * (size_t)nw);
    if (!xc || !yc) wavehip_die(wavehip_ctx(), "init_ior_2d (no memory)", WV_ENOMEM);
#if IOR == 6 || IOR == 7
    rc = wv_rotating_wells(NGRIDX, NGRIDY, IOR == 7, ior_angle, wavehip_box, xc, yc, &sigma);
    if (rc) wavehip_die(wavehip_ctx(), "init_ior_2d", rc);
#elif IOR == 8 || IOR == 9
    h = (double *)malloc(sizeof(double) * (size_t)nw);
    for (int n = 0; n < nw; n++) {
        h[n] = 0.5 + 0.5 * gaussian();
        if (h[n] > 1.0) h[n] = 1.0;
        if (h[n] < 0.0) h[n] = 0.0;
        xc[n] = wc[n].xc;
        yc[n] = wc[n].yc;
    }
    sigma = 0.2 * (XMAX - XMIN) * (YMAX - YMIN) / (double)nw;
    use_tanh = 1;
#else
    {
        const double dx = (XMAX - XMIN) / (double)NGRIDX, dy = (YMAX - YMIN) / (double)NGRIDY;
        sigma = 0.2 * dx * dx;
#if IOR == 5
        h = (double *)malloc(sizeof(double) * (size_t)nw);
        use_tanh = 1;
#endif
        for (int i = 0; i < NGRIDX; i++)
            for (int j = 0; j < NGRIDY; j++) {
                const int n = j * NGRIDX + i;
#if IOR == 5
                xc[n] = XMIN + dx * ((double)i + 0.5 + 0.1 * gaussian());
                yc[n] = YMIN + dy * ((double)j + 0.5 + 0.1 * gaussian());
                h[n] = 0.5 + 0.5 * gaussian();
                if (h[n] > 1.0) h[n] = 1.0;
                if (h[n] < 0.0) h[n] = 0.0;
#else
                xc[n] = XMIN + dx * ((double)i + 0.5);
                yc[n] = YMIN + dy * ((double)j + 0.5);
                if (j % 2 == 1) yc[n] += 0.5 * dx;
#endif
            }
    }
#endif
    rc = wv_set_wells_tables(wavehip_ctx(), xy, nw, xc, yc, h, sigma, use_tanh, COURANT, COURANTB, GAMMA,
                             wavehip_box, tcc, tc, tg);
    free(xc);
    free(yc);
    free(h);
#else
    wv_ior_params ip;
    ip.courant = COURANT; ip.courantb = COURANTB; ip.gamma = GAMMA; ip.gammab = GAMMAB; ip.lambda = LAMBDA;
    ip.mu = MU; ip.mandel_ior_scale = MANDEL_IOR_SCALE; ip.mandellimit = MANDELLIMIT; ip.mandellevel = MANDELLEVEL;
    ip.mdepth = MDEPTH;
    (void)ior_angle;
    rc = wv_set_ior_tables(wavehip_ctx(), xy, VARIABLE_IOR ? IOR : WV_IOR_CONSTANT, wavehip_box, &ip, tcc, tc, tg);
#endif
    if (rc) wavehip_die(wavehip_ctx(), "init_ior_2d", rc);
}

#ifndef WAVEHIP_FLAT
static void wavehip_init_wave_flat(double *phi[NX], double *psi[NX], short int *xy_in[NX])
{
    if (!XYIN_INITIALISED) {
        short *m = (short *)malloc(sizeof(short) * (size_t)NX * NY);
        if (!m) wavehip_die(wavehip_ctx(), "init_wave_flat (no memory)", WV_ENOMEM);
        wavehip_mask(m);
        for (int i = 0; i < NX; i++) memcpy(xy_in[i], m + (size_t)i * NY, sizeof(short) * NY);
        free(m);
    }
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) phi[i][j] = psi[i][j] = 0.0;
}

static void wavehip_init_ior_2d(short int *xy_in[NX], double *tcc_table[NX], double *tc_table[NX],
                                double *tgamma_table[NX], double ior_angle)
{
    const size_t nc = (size_t)NX * NY;
    short *m = (short *)malloc(sizeof(short) * nc);
    double *t = (double *)malloc(sizeof(double) * 3 * nc);
    if (!m || !t) wavehip_die(wavehip_ctx(), "init_ior_2d (no memory)", WV_ENOMEM);
    for (int i = 0; i < NX; i++) memcpy(m + (size_t)i * NY, xy_in[i], sizeof(short) * NY);
    wavehip_media(m, ior_angle, t, t + nc, t + 2 * nc);
    for (int i = 0; i < NX; i++) {
        memcpy(tcc_table[i], t + (size_t)i * NY, sizeof(double) * NY);
        memcpy(tc_table[i], t + nc + (size_t)i * NY, sizeof(double) * NY);
        memcpy(tgamma_table[i], t + 2 * nc + (size_t)i * NY, sizeof(double) * NY);
    }
    free(m);
    free(t);
}
#else
static void wavehip_init_wave_flat_mod(double phi[NX * NY], double psi[NX * NY], short int xy_in[NX * NY])
{
    if (!XYIN_INITIALISED) wavehip_mask(xy_in);
    for (size_t k = 0; k < (size_t)NX * NY; k++) phi[k] = psi[k] = 0.0;
}
#endif
#endif /* WAVEHIP_INIT */

#endif
