/*
 * compat_host.c -- a host program shaped like a reference wave program that
 * steps its field through include/wavehip_compat.h: row-pointer tables
 * (double *phi[NX], one allocation per row, as wave_billiard.c:676-685
 * allocates them) or, with WAVEHIP_FLAT, flat tables k = i*NY + j (wave_3d.c).
 *
 * The parameter macros come from the compile line, as they come from a
 * reference program's #define block.  The two types a reference program takes
 * from global_pdes.c (t_neumann_bc, t_laplacian) are declared with the
 * fields the steppers read.  With PRECOMPUTE_BC (flat) the program steps
 * through evolve_wave with neighbour tables for phi, tmp and psi, built from
 * the count/index table in the input as init_laplacian_coords would build
 * them (sub_wave.c:11555).
 *
 *   compat_host IN OUT NCALLS MODE
 *   IN    phi, psi, tmp, tcc, tc, tgamma (NX*NY doubles each, k = i*NY + j)
 *         followed by xy_in (NX*NY shorts), then int32 n and n Neumann points
 *         (int32 i, j, nbi1, nbj1, nbi2, nbj2; doubles nx, ny), then int32 T
 *         and T*NY doubles: oscillating_bc(time, j) for time 1..T; with
 *         PRECOMPUTE_BC then NX*NY int32 counts and NX*NY*4 int32 flat
 *         neighbour indices
 *   MODE  "wave": NCALLS calls of evolve_wave (3 steps each);
 *         "half": NCALLS calls of evolve_wave_half(phi, psi, tmp) (with
 *         PRECOMPUTE_BC evolve_wave_half_new), the host rotating
 *         (phi, psi, tmp) <- (tmp, phi, psi) after each
 *   OUT   phi, psi, tmp after the calls, same layout as IN; with
 *         PRECOMPUTE_BC and OSCILLATE_LEFT then T doubles
 *         AMPLITUDE*cos(time*OMEGA), time 1..T (the corner drive)
 *
 *   compat_host --params   prints the macros it was built with
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define XMIN -2.0
#define XMAX 2.0
#ifndef OSCILLATE_LEFT
#define OSCILLATE_LEFT 0
#endif
#ifndef BC_NEUMANN
#define BC_NEUMANN 0
#endif
#ifndef FLOOR
#define FLOOR 0
#endif
#ifndef VMAX
#define VMAX 2.5
#endif
#define BEAM_BC 0
#define OSCIL_YMID 0.0
#define OSCIL_YMAX 1.0
#define N_NEUMANN_POINTS 4096
#define AMPLITUDE 0.8
#define OMEGA 0.05
#ifndef PRECOMPUTE_BC
#define PRECOMPUTE_BC 0
#endif
#ifndef COMPARISON
#define COMPARISON 0
#endif

/* the fields of global_pdes.c:518-528 the stepper reads */
typedef struct {
    double nx, ny;
    int nbi1, nbj1, nbk1;
    int nbi2, nbj2, nbk2;
    int i, j, k;
    short int active;
} t_neumann_bc;
typedef struct {
    int nneighb;
    double *nghb[4];
} t_laplacian;

int n_neumann_points = 0;

/* the left-edge schedule: a table read from the input (time 1, 2, ...) */
static double *drive_table;
static int drive_steps;
double oscillating_bc(int time, int j)
{
    return (time >= 1 && time <= drive_steps) ? drive_table[(size_t)(time - 1) * NY + j] : 0.0;
}
void xy_to_ij(double x, double y, int ij[2])
{
    ij[0] = (int)((x - XMIN) / (XMAX - XMIN) * NX);
    ij[1] = (int)((y + 1.0) / 2.0 * NY);
}

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
        printf("nx=%d ny=%d bcond=%d twospeeds=%d oscillate_topbot=%d flat=%d fixed=%d neumann=%d "
               "oscillate_left=%d floor=%d vmax=%.17g table=%d comparison=%d "
               "kappa=%.17g kappa_sides=%.17g gamma_sides=%.17g kappa_topbot=%.17g gamma_topbot=%.17g\n",
               NX, NY, B_COND, TWOSPEEDS, OSCILLATE_TOPBOT, HOST_FLAT, HOST_FIXED, BC_NEUMANN, OSCILLATE_LEFT,
               FLOOR, (double)VMAX, PRECOMPUTE_BC, COMPARISON, KAPPA, KAPPA_SIDES, GAMMA_SIDES, KAPPA_TOPBOT,
               GAMMA_TOPBOT);
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
    static t_neumann_bc nbc[N_NEUMANN_POINTS];
    int nn = 0;
    read_all(f, &nn, sizeof(int), 1);
    if (nn < 0 || nn > N_NEUMANN_POINTS) return 2;
    for (int q = 0; q < nn; q++) {
        int v[6];
        read_all(f, v, sizeof(int), 6);
        nbc[q].i = v[0]; nbc[q].j = v[1]; nbc[q].k = v[0] * NY + v[1];
        nbc[q].nbi1 = v[2]; nbc[q].nbj1 = v[3]; nbc[q].nbk1 = v[2] * NY + v[3];
        nbc[q].nbi2 = v[4]; nbc[q].nbj2 = v[5]; nbc[q].nbk2 = v[4] * NY + v[5];
        nbc[q].active = 1;
    }
    for (int q = 0; q < nn; q++) read_all(f, &nbc[q].nx, sizeof(double), 1);
    for (int q = 0; q < nn; q++) read_all(f, &nbc[q].ny, sizeof(double), 1);
    n_neumann_points = nn;
    read_all(f, &drive_steps, sizeof(int), 1);
    if (drive_steps < 0) return 2;
    drive_table = malloc(sizeof(double) * (size_t)(drive_steps > 0 ? drive_steps : 1) * NY);
    read_all(f, drive_table, sizeof(double), (size_t)drive_steps * NY);
#if PRECOMPUTE_BC
    int *tnn = malloc(sizeof(int) * NC), *tnb = malloc(sizeof(int) * 4 * NC);
    if (!tnn || !tnb) return 2;
    read_all(f, tnn, sizeof(int), NC);
    read_all(f, tnb, sizeof(int), 4 * NC);
#endif
    fclose(f);

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
    double *phi = flat[0], *psi = flat[1], *tmp = flat[2];
#if PRECOMPUTE_BC
    /* one table per buffer, as wave_3d.c:1187-1189: laplace for phi,
     * laplace1 for tmp, laplace2 for psi */
    t_laplacian *lap[3];
    double *lbase[3] = {phi, tmp, psi};
    for (int a = 0; a < 3; a++) {
        lap[a] = calloc(NC, sizeof(t_laplacian));
        if (!lap[a]) return 2;
        for (size_t k = 0; k < NC; k++) {
            lap[a][k].nneighb = tnn[k];
            for (int q = 0; q < 4; q++) lap[a][k].nghb[q] = tnb[4 * k + q] >= 0 ? lbase[a] + tnb[4 * k + q] : NULL;
        }
    }
#else
    static t_laplacian lap1[1];
    t_laplacian *lap[3] = {lap1, lap1, lap1};
#endif
    for (int n = 0; n < ncalls; n++) {
        if (half) {
#if PRECOMPUTE_BC
            /* the table built for the buffer that is phi_in this call */
            int a = phi == lbase[0] ? 0 : (phi == lbase[1] ? 1 : 2);
            evolve_wave_half_new(phi, psi, tmp, xyf, flat[4], flat[3], flat[5], lap[a]);
#else
            evolve_wave_half(phi, psi, tmp, xyf, flat[4], flat[3], flat[5], nbc);
#endif
            double *t = psi;
            psi = phi;
            phi = tmp;
            tmp = t;
        } else {
            evolve_wave(phi, psi, tmp, xyf, flat[4], flat[3], flat[5], lap[0], lap[1], lap[2], nbc);
        }
    }
    double *outs[3] = {phi, psi, tmp};
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 2;
    for (int a = 0; a < 3; a++) write_all(g, outs[a], sizeof(double), NC);
#if PRECOMPUTE_BC && OSCILLATE_LEFT
    for (int t = 1; t <= drive_steps; t++) {
        const double c = AMPLITUDE * cos((double)t * OMEGA);
        write_all(g, &c, sizeof(double), 1);
    }
#endif
    fclose(g);
#endif
    wv_destroy(wavehip_ctx());
    return 0;
}
