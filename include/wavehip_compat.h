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
 * Not provided (the program aborts with a message on the first call):
 * OSCILLATE_LEFT (wave_billiard.c:320-341,397,585-590), BC_NEUMANN
 * (wave_billiard.c:379-389), FLOOR (wave_billiard.c:592-609), COMPARISON
 * (wave_3d.c:805-976).
 *
 * Faster path for the frame loop (wave_billiard.c:897-899): keep the state on
 * the GPU and call wv_step(wavehip_ctx(), 3*NVID) plus one wv_get_state per
 * frame instead of NVID calls to evolve_wave.
 */
#ifndef WAVEHIP_COMPAT_H
#define WAVEHIP_COMPAT_H

#include <stdio.h>
#include <stdlib.h>

#include "wavehip.h"

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
#if OSCILLATE_LEFT
        fprintf(stderr, "wavehip: OSCILLATE_LEFT is not supported by libwavehip\n");
        abort();
#endif
#if BC_NEUMANN
        fprintf(stderr, "wavehip: BC_NEUMANN is not supported by libwavehip\n");
        abort();
#endif
#if FLOOR
        fprintf(stderr, "wavehip: FLOOR is not supported by libwavehip\n");
        abort();
#endif
#if defined(COMPARISON) && COMPARISON
        fprintf(stderr, "wavehip: COMPARISON is not supported by libwavehip\n");
        abort();
#endif
        wv_params_default(&p, NX, NY);
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
#ifdef WAVEHIP_FLAT
        if (p.tb_shift >= NX - 1) p.tb_shift = NY - 2;
#endif
#ifdef WAVEHIP_TABLES_FIXED
        p.flags |= WV_FLAG_TABLES_FIXED;
#endif
        rc = wv_create(&ctx, &p);
        if (rc) wavehip_die(ctx, "wv_create", rc);
    }
    return ctx;
}

#ifndef WAVEHIP_FLAT
/* wave_billiard.c:302-304 */
static void evolve_wave_half(double *phi_in[NX], double *psi_in[NX], double *phi_out[NX],
                             short int *xy_in[NX], double *tcc[NX], double *tc[NX],
                             double *tgamma[NX], t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc = wv_evolve_wave_half_rows(wavehip_ctx(), phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave_half", rc);
}

/* wave_billiard.c:613 */
static void evolve_wave(double *phi[NX], double *psi[NX], double *tmp[NX], short int *xy_in[NX],
                        double *tcc_table[NX], double *tc_table[NX], double *tgamma_table[NX],
                        t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc = wv_evolve_wave_rows(wavehip_ctx(), phi, psi, tmp, xy_in, tcc_table, tc_table, tgamma_table);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave", rc);
}
#else
/* wave_3d.c:600 */
static void evolve_wave_half(double phi_in[NX * NY], double psi_in[NX * NY], double phi_out[NX * NY],
                             short int xy_in[NX * NY], double tc[NX * NY], double tcc[NX * NY],
                             double tgamma[NX * NY], t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc = wv_evolve_wave_half_flat(wavehip_ctx(), phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma);
    (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave_half", rc);
}

/* wave_3d.c:1010 (PRECOMPUTE_BC 0: the t_laplacian tables are not read) */
static void evolve_wave(double phi[NX * NY], double psi[NX * NY], double tmp[NX * NY],
                        short int xy_in[NX * NY], double tc[NX * NY], double tcc[NX * NY],
                        double tgamma[NX * NY], t_laplacian laplace[NX * NY],
                        t_laplacian laplace1[NX * NY], t_laplacian laplace2[NX * NY],
                        t_neumann_bc neumann_bc[N_NEUMANN_POINTS])
{
    int rc = wv_evolve_wave_flat(wavehip_ctx(), phi, psi, tmp, xy_in, tc, tcc, tgamma);
    (void)laplace; (void)laplace1; (void)laplace2; (void)neumann_bc;
    if (rc) wavehip_die(wavehip_ctx(), "evolve_wave", rc);
}
#endif

#endif
