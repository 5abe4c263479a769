/*
 * compat_host.c -- a host program shaped like a reference wave program that
 * steps its field through include/wavehip_compat.h: row-pointer tables
 * (double *phi[NX], one allocation per row, as wave_billiard.c:676-685
 * allocates them) or, with WAVEHIP_FLAT, flat tables k = i*NY + j (wave_3d.c).
 *
 * The parameter macros come from the compile line, as they come from a
 * reference program's #define block.  The two types a reference program takes
 * from global_pdes.c (t_neumann_bc, t_laplacian) are declared in minimal form:
 * the stepper never reads them.
 *
 *   compat_host IN OUT NCALLS MODE
 *   IN    phi, psi, tmp, tcc, tc, tgamma (NX*NY doubles each, k = i*NY + j)
 *         followed by xy_in (NX*NY shorts)
 *   MODE  "wave": NCALLS calls of evolve_wave (3 steps each);
 *         "half": NCALLS calls of evolve_wave_half(phi, psi, tmp), the host
 *         rotating (phi, psi, tmp) <- (tmp, phi, psi) after each
 *   OUT   phi, psi, tmp after the calls, same layout as IN
 *
 *   compat_host --params   prints the macros it was built with
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define XMIN -2.0
#define XMAX 2.0
#define OSCILLATE_LEFT 0
#define BC_NEUMANN 0
#define FLOOR 0
#define N_NEUMANN_POINTS 1

typedef struct {
    double x, y, nx, ny;
} t_neumann_bc;
typedef struct {
    int nneighb;
} t_laplacian;

#include "wavehip_compat.h"

#define NC ((size_t)NX * NY)

static void read_all(FILE *f, void *p, size_t sz, size_t n)
{
    if (fread(p, sz, n, f) != n) {
        fprintf(stderr, "compat_host: short input\n");
        exit(2);
    }
}

static void write_all(FILE *f, const void *p, size_t sz, size_t n)
{
    if (fwrite(p, sz, n, f) != n) {
        fprintf(stderr, "compat_host: short write\n");
        exit(2);
    }
}

#ifdef WAVEHIP_FLAT
#define HOST_FLAT 1
#else
#define HOST_FLAT 0
#endif
#ifdef WAVEHIP_TABLES_FIXED
#define HOST_FIXED 1
#else
#define HOST_FIXED 0
#endif

int main(int argc, char **argv)
{
    if (argc == 2 && strcmp(argv[1], "--params") == 0) {
        printf("nx=%d ny=%d bcond=%d twospeeds=%d oscillate_topbot=%d flat=%d fixed=%d "
               "kappa=%.17g kappa_sides=%.17g gamma_sides=%.17g kappa_topbot=%.17g gamma_topbot=%.17g\n",
               NX, NY, B_COND, TWOSPEEDS, OSCILLATE_TOPBOT, HOST_FLAT, HOST_FIXED, KAPPA, KAPPA_SIDES,
               GAMMA_SIDES, KAPPA_TOPBOT, GAMMA_TOPBOT);
        return 0;
    }
    if (argc != 5) {
        fprintf(stderr, "usage: %s IN OUT NCALLS wave|half\n", argv[0]);
        return 2;
    }
    const int ncalls = atoi(argv[3]);
    const int half = strcmp(argv[4], "half") == 0;
    double *flat[6];
    short *xyf = malloc(sizeof(short) * NC);
    FILE *f = fopen(argv[1], "rb");
    if (!f || !xyf) return 2;
    for (int a = 0; a < 6; a++) {
        flat[a] = malloc(sizeof(double) * NC);
        read_all(f, flat[a], sizeof(double), NC);
    }
    read_all(f, xyf, sizeof(short), NC);
    fclose(f);
    t_neumann_bc nbc[N_NEUMANN_POINTS];
    memset(nbc, 0, sizeof nbc);

#ifndef WAVEHIP_FLAT
    /* row-pointer tables, one malloc per row */
    double **rows[6];
    short *xy[NX];
    for (int a = 0; a < 6; a++) {
        rows[a] = malloc(sizeof(double *) * NX);
        for (int i = 0; i < NX; i++) {
            rows[a][i] = malloc(sizeof(double) * NY);
            memcpy(rows[a][i], flat[a] + (size_t)i * NY, sizeof(double) * NY);
        }
    }
    for (int i = 0; i < NX; i++) {
        xy[i] = malloc(sizeof(short) * NY);
        memcpy(xy[i], xyf + (size_t)i * NY, sizeof(short) * NY);
    }
    double **phi = rows[0], **psi = rows[1], **tmp = rows[2];
    for (int n = 0; n < ncalls; n++) {
        if (half) {
            evolve_wave_half(phi, psi, tmp, xy, rows[3], rows[4], rows[5], nbc);
            double **t = psi;
            psi = phi;
            phi = tmp;
            tmp = t;
        } else {
            evolve_wave(phi, psi, tmp, xy, rows[3], rows[4], rows[5], nbc);
        }
    }
    double **outs[3] = {phi, psi, tmp};
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 2;
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < NX; i++) write_all(g, outs[a][i], sizeof(double), NY);
    fclose(g);
#else
    /* flat tables; evolve_wave_half of wave_3d.c takes (tc, tcc, tgamma) */
    static t_laplacian lap[1];
    double *phi = flat[0], *psi = flat[1], *tmp = flat[2];
    for (int n = 0; n < ncalls; n++) {
        if (half) {
            evolve_wave_half(phi, psi, tmp, xyf, flat[4], flat[3], flat[5], nbc);
            double *t = psi;
            psi = phi;
            phi = tmp;
            tmp = t;
        } else {
            evolve_wave(phi, psi, tmp, xyf, flat[4], flat[3], flat[5], lap, lap, lap, nbc);
        }
    }
    double *outs[3] = {phi, psi, tmp};
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 2;
    for (int a = 0; a < 3; a++) write_all(g, outs[a], sizeof(double), NC);
    fclose(g);
#endif
    wv_destroy(wavehip_ctx());
    return 0;
}
