/*
 * compat_init.c -- the start-up of a reference wave program (wave_billiard.c
 * 780-803, wave_3d.c:1212) through include/wavehip_compat.h with
 * WAVEHIP_INIT: the billiard mask of init_wave_flat / init_wave_flat_mod and
 * the tables of init_ior_2d, built on the GPU by the compat drop-ins.  The
 * program's globals (circles, circles_b, julia_x, julia_y) and #define block
 * are set the way global_pdes.c and the program set them; the circle lists
 * come from the input file.
 *
 *   compat_init IN OUT     IN: int32 n, 4n doubles (xc[], yc[], r[], active[]),
 *                          the same for circles_b (builds with XY_FROM_INPUT:
 *                          then the mask, int16, and the int32 seed of the
 *                          program's srand); OUT: xy_in (int16), phi, psi,
 *                          and (row-pointer builds) tcc, tc, tgamma
 *   compat_init --params   prints the macros it was built with
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define XMIN -2.0
#define XMAX 2.0
#define YMIN -1.125
#define YMAX 1.125
#define OSCILLATE_LEFT 0
#define BC_NEUMANN 0
#define FLOOR 0
#define VMAX 2.5
#define BEAM_BC 0
#define OSCIL_YMID 0.0
#define OSCIL_YMAX 1.0
#define N_NEUMANN_POINTS 16
#define AMPLITUDE 0.8
#define OMEGA 0.05
#define PRECOMPUTE_BC 0
#define B_COND 2
#define OSCILLATE_TOPBOT 0
#define KAPPA 0.0
#define KAPPA_SIDES 0.0
#define GAMMA_SIDES 0.0
#define KAPPA_TOPBOT 0.0
#define GAMMA_TOPBOT 0.0
#ifndef COMPARISON
#define COMPARISON 0
#endif
#ifndef B_DOMAIN_B
#define B_DOMAIN_B 20
#endif
#define D_IMAGE 77
#define XYIN_INITIALISED (B_DOMAIN == D_IMAGE)
#define NMAXCIRCLES 1000
#define LAMBDA 0.8
#define MU 0.3
#define MU_B 0.08
#define APOLY 0.5
#define NPOLY 6
#define WALL_WIDTH 0.1
#define WALL_WIDTH_B 0.1
#define WALL_WIDTH_ASYM 0.75
#define WALL_WIDTH_ASYM_B 0.75
#define JULIA_SCALE 0.8
#define MDEPTH 4
#define MRATIO 3
#define MANDELLEVEL 100
#define MANDELLIMIT 10.0
#ifndef NGRIDX
#define NGRIDX 6
#endif
#ifndef NGRIDY
#define NGRIDY 5
#endif
#define COURANT 0.1
#define COURANTB 0.05
#define GAMMA 1.0e-4
#define GAMMAB 2.0e-4
#define MANDEL_IOR_SCALE -0.05

/* the types and globals of global_pdes.c the compat header reads */
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
typedef struct {
    double xc, yc, radius;
    short int active, top;
} t_circle;
int n_neumann_points = 0;
int ncircles = NMAXCIRCLES, ncircles_b = NMAXCIRCLES;
t_circle circles[NMAXCIRCLES], circles_b[NMAXCIRCLES];
double julia_x = 0.37468, julia_y = 0.21115;
double michelson_position = 0.0;
double oscillating_bc(int time, int j) { (void)time; (void)j; return 0.0; }
/* the program's normal deviates: Marsaglia's polar method on rand(), the
 * second deviate of each accepted pair kept for the next call */
double gaussian(void)
{
    static double v1, v2, s;
    static int second = 0;
    double x;
    if (!second) {
        do {
            v1 = 2 * ((double)rand() / RAND_MAX) - 1;
            v2 = 2 * ((double)rand() / RAND_MAX) - 1;
            s = v1 * v1 + v2 * v2;
        } while (s >= 1 || s == 0);
        x = v1 * sqrt(-2 * log(s) / s);
    } else {
        x = v2 * sqrt(-2 * log(s) / s);
    }
    second = !second;
    return x;
}
void xy_to_ij(double x, double y, int ij[2])
{
    ij[0] = (int)((x - XMIN) / (XMAX - XMIN) * NX);
    ij[1] = (int)((y - YMIN) / (YMAX - YMIN) * NY);
}

#define WAVEHIP_INIT
#ifdef LIST_INPUT
/* the rectangle lists a list-driven domain reads (global_pdes.c:534-545) */
#define NMAXPOLY 50000
typedef struct {
    double x1, y1, x2, y2;
    double posi1, posj1, posi2, posj2;
} t_rectangle;
typedef struct {
    double x1, y1, x2, y2;
    double width;
    double posi1, posj1, posi2, posj2, posi3, posj3, posi4, posj4;
} t_rect_rotated;
int npolyrect = NMAXPOLY, npolyrect_rot = NMAXPOLY;
t_rectangle polyrect[NMAXPOLY];
t_rect_rotated polyrectrot[NMAXPOLY];
#endif

#include "wavehip_compat.h"

#define NC ((size_t)NX * NY)

static void read_list(FILE *f, t_circle *c, int *n)
{
    if (fread(n, sizeof(int), 1, f) != 1 || *n < 0 || *n > NMAXCIRCLES) exit(2);
    double *a = malloc(sizeof(double) * 4 * (size_t)(*n > 0 ? *n : 1));
    if (!a || fread(a, sizeof(double), 4 * (size_t)*n, f) != 4 * (size_t)*n) exit(2);
    for (int k = 0; k < *n; k++) {
        c[k].xc = a[k];
        c[k].yc = a[*n + k];
        c[k].radius = a[2 * *n + k];
        c[k].active = a[3 * *n + k] != 0.0;
    }
    free(a);
}

#ifdef LIST_INPUT
/* a count, then rows runs of that many doubles */
static double *read_runs(FILE *f, int rows, int *n)
{
    if (fread(n, sizeof(int), 1, f) != 1 || *n < 0 || *n > NMAXPOLY) exit(2);
    double *a = malloc(sizeof(double) * rows * (size_t)(*n > 0 ? *n : 1));
    if (!a || fread(a, sizeof(double), rows * (size_t)*n, f) != rows * (size_t)*n) exit(2);
    return a;
}
#endif

int main(int argc, char **argv)
{
    if (argc == 2 && strcmp(argv[1], "--params") == 0) {
#ifdef XY_FROM_INPUT
        const int xyin = 1;
#else
        const int xyin = 0;
#endif
#ifdef LIST_INPUT
        const int lists = 1;
#else
        const int lists = 0;
#endif
        printf("nx=%d ny=%d twospeeds=%d domain=%d domain_b=%d comparison=%d variable_ior=%d ior=%d xy_from_input=%d "
               "lists=%d\n", NX, NY, TWOSPEEDS, B_DOMAIN, B_DOMAIN_B, COMPARISON, VARIABLE_IOR, IOR, xyin, lists);
        return 0;
    }
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    read_list(f, circles, &ncircles);
    read_list(f, circles_b, &ncircles_b);
#ifdef LIST_INPUT
    int n;
    double *a = read_runs(f, 4, &n);
    for (int k = 0; k < n; k++) {
        polyrect[k].x1 = a[k]; polyrect[k].y1 = a[n + k]; polyrect[k].x2 = a[2 * n + k]; polyrect[k].y2 = a[3 * n + k];
    }
    npolyrect = n;
    free(a);
    a = read_runs(f, 5, &n);
    for (int k = 0; k < n; k++) {
        polyrectrot[k].x1 = a[k]; polyrectrot[k].y1 = a[n + k]; polyrectrot[k].x2 = a[2 * n + k];
        polyrectrot[k].y2 = a[3 * n + k]; polyrectrot[k].width = a[4 * n + k];
    }
    npolyrect_rot = n;
    free(a);
#endif
#ifdef XY_FROM_INPUT
    static short xy_in_file[(size_t)NX * NY];
    int seed = 0;
    if (fread(xy_in_file, sizeof(short), NC, f) != NC || fread(&seed, sizeof(int), 1, f) != 1) return 2;
#endif
    fclose(f);
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 2;
#ifndef WAVEHIP_FLAT
    double *phi[NX], *psi[NX], *tcc[NX], *tc[NX], *tg[NX];
    short int *xy[NX];
    for (int i = 0; i < NX; i++) {
        phi[i] = malloc(sizeof(double) * NY); psi[i] = malloc(sizeof(double) * NY);
        tcc[i] = malloc(sizeof(double) * NY); tc[i] = malloc(sizeof(double) * NY);
        tg[i] = malloc(sizeof(double) * NY); xy[i] = malloc(sizeof(short) * NY);
        for (int j = 0; j < NY; j++) phi[i][j] = psi[i][j] = 1.0;  /* init sets them to 0 */
    }
    wavehip_init_wave_flat(phi, psi, xy);
#ifdef XY_FROM_INPUT
    /* the program's own mask (XYIN_INITIALISED-like) and rand() sequence */
    for (int i = 0; i < NX; i++) memcpy(xy[i], xy_in_file + (size_t)i * NY, sizeof(short) * NY);
    srand((unsigned)seed);
#endif
    wavehip_init_ior_2d(xy, tcc, tc, tg, 0.0);
    for (int i = 0; i < NX; i++) fwrite(xy[i], sizeof(short), NY, g);
    double **outs[5] = {phi, psi, tcc, tc, tg};
    for (int a = 0; a < 5; a++)
        for (int i = 0; i < NX; i++) fwrite(outs[a][i], sizeof(double), NY, g);
#else
    double *phi = malloc(sizeof(double) * NC), *psi = malloc(sizeof(double) * NC);
    short *xy = malloc(sizeof(short) * NC);
    if (!phi || !psi || !xy) return 2;
    for (size_t k = 0; k < NC; k++) phi[k] = psi[k] = 1.0;
    wavehip_init_wave_flat_mod(phi, psi, xy);
    fwrite(xy, sizeof(short), NC, g);
    fwrite(phi, sizeof(double), NC, g);
    fwrite(psi, sizeof(double), NC, g);
#endif
    fclose(g);
    wv_destroy(wavehip_ctx());
    return 0;
}
