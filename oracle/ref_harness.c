/*
 * ref_harness.c -- appended after the reference functions by
 * oracle/build_ref.sh.  Exposes them on flat k = i*NY + j arrays so the
 * tests can call them through ctypes.  FLAT_TWIN / TABLE_TWIN select which
 * reference stepper was compiled in (wave_billiard.c row-pointer, wave_3d.c
 * flat, or the wave_3d.c neighbour-table stepper).
 */
/* Neumann ghost points handed to the reference stepper (n_neumann_points of
 * global_pdes.c says how many are used; none until ref_set_neumann) */
#define REF_NEU_MAX 4096
static t_neumann_bc ref_neumann[REF_NEU_MAX];
__attribute__((constructor)) static void ref_init(void) { n_neumann_points = 0; }

int ref_nx(void) { return NX; }
int ref_ny(void) { return NY; }

int ref_set_neumann(int n, const int *pi, const int *pj, const int *bi1, const int *bj1,
                    const int *bi2, const int *bj2, const double *nx, const double *ny)
{
    if (n < 0 || n > REF_NEU_MAX) return -1;
    memset(ref_neumann, 0, sizeof ref_neumann);
    for (int k = 0; k < n; k++) {
        ref_neumann[k].i = pi[k]; ref_neumann[k].j = pj[k];
        ref_neumann[k].nbi1 = bi1[k]; ref_neumann[k].nbj1 = bj1[k];
        ref_neumann[k].nbi2 = bi2[k]; ref_neumann[k].nbj2 = bj2[k];
        ref_neumann[k].nx = nx[k]; ref_neumann[k].ny = ny[k];
        ref_neumann[k].active = 1;
    }
    n_neumann_points = n;
    return 0;
}

/* the left-edge drive, sub_wave.c:11965 */
double ref_oscillating_bc(int time, int j) { return oscillating_bc(time, j); }
int ref_beam_bc(void) { return BEAM_BC; }

#if defined(IOR_TWIN)
/* init_ior_2d (sub_wave.c:12164) on flat tables, after srand(seed) */
void ref_init_ior(short *xy_in, double *tcc, double *tc, double *tgamma, unsigned seed)
{
    double *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    srand(seed);
    init_ior_2d(m, t1, t2, t3, 0.0);
}

/* the same at a rotation angle (the REFRESH_IOR call, wave_billiard.c:943) */
void ref_init_ior_at(short *xy_in, double *tcc, double *tc, double *tgamma, unsigned seed, double angle)
{
    double *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    srand(seed);
    init_ior_2d(m, t1, t2, t3, angle);
}

int ref_nwells(void) { return NGRIDX * NGRIDY; }

/* the wells init_ior_2d places: IOR_PERIODIC_WELLS (sub_wave.c:12317-12328) or
 * IOR_RANDOM_WELLS (12352-12365, drawn with the reference's gaussian() from
 * the rand() sequence of srand(seed)); height 1 for the periodic wells */
double ref_wells(unsigned seed, double *xc, double *yc, double *h)
{
    const double dx = (XMAX - XMIN) / (double)NGRIDX, dy = (YMAX - YMIN) / (double)NGRIDY;
    srand(seed);
    for (int i = 0; i < NGRIDX; i++)
        for (int j = 0; j < NGRIDY; j++) {
            const int n = j * NGRIDX + i;
            if (IOR == IOR_RANDOM_WELLS) {
                xc[n] = XMIN + dx * ((double)i + 0.5 + 0.1 * gaussian());
                yc[n] = YMIN + dy * ((double)j + 0.5 + 0.1 * gaussian());
                h[n] = 0.5 + 0.5 * gaussian();
                if (h[n] > 1.0) h[n] = 1.0;
                if (h[n] < 0.0) h[n] = 0.0;
            } else {
                xc[n] = XMIN + dx * ((double)i + 0.5);
                yc[n] = YMIN + dy * ((double)j + 0.5);
                if (j % 2 == 1) yc[n] += 0.5 * dx;
                h[n] = 1.0;
            }
        }
    return 0.2 * dx * dx; /* sigma */
}
#elif defined(PULSE_TWIN)
/* init_circular_wave (wave_common.c:89) and add_circular_wave (337) on flat
 * arrays through row pointers; init writes the mask xy_in_billiard gives */
void ref_init_circular_wave(double x, double y, double *phi, double *psi, short *xy)
{
    double *a[NX], *b[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi + (size_t)i * NY; b[i] = psi + (size_t)i * NY; m[i] = xy + (size_t)i * NY;
    }
    init_circular_wave(x, y, a, b, m);
}
void ref_add_circular_wave(double factor, double x, double y, double *phi, double *psi, short *xy)
{
    double *a[NX], *b[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi + (size_t)i * NY; b[i] = psi + (size_t)i * NY; m[i] = xy + (size_t)i * NY;
    }
    add_circular_wave(factor, x, y, a, b, m);
}
#elif defined(MASK_TWIN)
/* xy_in_billiard_single_domain (sub_wave.c:5712) at the cell corners
 * ij_to_xy gives (sub_wave.c:469-476), flat out[i*NY + j], with the circle
 * list of the caller (D_CIRCLES) */
int ref_mask(int domain, int ncirc, const double *xc, const double *yc, const double *r, const int *active,
             short *out)
{
    if (ncirc < 0 || ncirc > NMAXCIRCLES) return -1;
    for (int k = 0; k < ncirc; k++) {
        circles[k].xc = xc[k];
        circles[k].yc = yc[k];
        circles[k].radius = r[k];
        circles[k].active = (short)active[k];
    }
    ncircles = ncirc;  /* the disc lattices read the global count */
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            double xy[2];
            ij_to_xy(i, j, xy);
            out[(size_t)i * NY + j] = (short)xy_in_billiard_single_domain(xy[0], xy[1], domain, ncirc, circles);
        }
    return 0;
}
/* the mirror position D_MICHELSON_MOVING reads (global_pdes.c:570) */
void ref_set_michelson(double position) { michelson_position = position; }
/* the lists the list-driven domains read (global_pdes.c:533-545), as the
 * program's init functions leave them, from flat arrays: rectangles x1, y1,
 * x2, y2 (polyrect); rotated rectangles x1, y1, x2, y2, width (polyrectrot);
 * polygons xc, yc, radius, angle, nsides, active (polygons); vertices x, y
 * (polyline); each field a run of n values */
int ref_set_lists(int nr, const double *rect, int nrr, const double *rrot, int np, const double *pg, int npl,
                  const double *pl)
{
    if (nr < 0 || nr > NMAXPOLY || nrr < 0 || nrr > NMAXPOLY || np < 0 || np > NMAXCIRCLES || npl < 0 ||
        npl > NMAXPOLY)
        return -1;
    for (int k = 0; k < nr; k++) {
        polyrect[k].x1 = rect[k]; polyrect[k].y1 = rect[nr + k];
        polyrect[k].x2 = rect[2 * nr + k]; polyrect[k].y2 = rect[3 * nr + k];
    }
    npolyrect = nr;
    for (int k = 0; k < nrr; k++) {
        polyrectrot[k].x1 = rrot[k]; polyrectrot[k].y1 = rrot[nrr + k];
        polyrectrot[k].x2 = rrot[2 * nrr + k]; polyrectrot[k].y2 = rrot[3 * nrr + k];
        polyrectrot[k].width = rrot[4 * nrr + k];
    }
    npolyrect_rot = nrr;
    for (int k = 0; k < np; k++) {
        polygons[k].xc = pg[k]; polygons[k].yc = pg[np + k]; polygons[k].radius = pg[2 * np + k];
        polygons[k].angle = pg[3 * np + k]; polygons[k].nsides = (int)pg[4 * np + k];
        polygons[k].active = (short)pg[5 * np + k];
    }
    for (int k = 0; k < npl; k++) {
        polyline[k].x = pl[k];
        polyline[k].y = pl[npl + k];
    }
    npolyline = npl;
    return 0;
}
/* xy_in_billiard itself (sub_wave.c:7593-7602: with COMPARISON, B_DOMAIN
 * and circles where y > 0, B_DOMAIN_B and circles_b where y <= 0) at the cell
 * corners, with the two circle lists of the caller */
int ref_mask_billiard(int nc, const double *xc, const double *yc, const double *r, const int *act, int nb,
                      const double *xb, const double *yb, const double *rb, const int *actb, short *out)
{
    if (nc < 0 || nc > NMAXCIRCLES || nb < 0 || nb > NMAXCIRCLES) return -1;
    for (int k = 0; k < nc; k++) {
        circles[k].xc = xc[k]; circles[k].yc = yc[k]; circles[k].radius = r[k]; circles[k].active = (short)act[k];
    }
    for (int k = 0; k < nb; k++) {
        circles_b[k].xc = xb[k]; circles_b[k].yc = yb[k]; circles_b[k].radius = rb[k];
        circles_b[k].active = (short)actb[k];
    }
    ncircles = nc;
    ncircles_b = nb;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            double xy[2];
            ij_to_xy(i, j, xy);
            out[(size_t)i * NY + j] = (short)xy_in_billiard(xy[0], xy[1]);
        }
    return 0;
}
#elif defined(COLOR_TWIN)
/* wave_value (wave_common.c:1053) at (i, j) on flat arrays; total_flux is
 * flat already (index 2*j*NX + 2*i) */
double ref_wave_value(double *phi, double *psi, double *te, double *ae, double *tf, double *cs, short *xy,
                      double scale, int time, int plot, int palette, int i, int j, double *rgb)
{
    double *a[NX], *b[NX], *e1[NX], *e2[NX], *c[NX];
    short int *m[NX];
    for (int r = 0; r < NX; r++) {
        a[r] = phi + (size_t)r * NY; b[r] = psi + (size_t)r * NY;
        e1[r] = te + (size_t)r * NY; e2[r] = ae + (size_t)r * NY; c[r] = cs + (size_t)r * NY;
        m[r] = xy + (size_t)r * NY;
    }
    return wave_value(i, j, a, b, e1, e2, tf, c, m, scale, time, plot, palette, rgb);
}

/* a colour table of colormaps.c / turbo_colormap.c, for the direct
 * cross-check of the table palettes (entries, or 0 if none) */
int ref_palette_table(int palette, float *out)
{
    const float(*t)[3] = NULL;
    int n = 256;
    switch (palette) {
    case COL_TURBO: case COL_TURBO_CYCLIC: t = turbo_srgb_floats; break;
    case COL_VIRIDIS: t = viridis_srgb_floats; break;
    case COL_MAGMA: t = magma_srgb_floats; break;
    case COL_INFERNO: t = inferno_srgb_floats; break;
    case COL_PLASMA: t = plasma_srgb_floats; break;
    case COL_CIVIDIS: t = cividis_srgb_floats; break;
    case COL_PARULA: t = parula_srgb_floats; break;
    case COL_TWILIGHT: case COL_TWILIGHT_SHIFTED: t = twilight_srgb_floats; n = 511; break;
    default: return 0;
    }
    memcpy(out, t, sizeof(float) * 3 * n);
    return n;
}
#elif defined(TABLE_TWIN)
#include <fcntl.h>
#include <unistd.h>
/* init_laplacian_coords (sub_wave.c:11555) on a zeroed table for the field at
 * `base`, its progress message discarded.  A table depends only on the
 * address it was built for, so three are kept by address (the three buffers
 * of a rotation). */
static double *ref_tbl_base[3];
static t_laplacian *ref_tbl[3];
static int ref_tbl_next;
static t_laplacian *ref_table_for(double *base)
{
    for (int q = 0; q < 3; q++)
        if (ref_tbl[q] && ref_tbl_base[q] == base) return ref_tbl[q];
    const int q = ref_tbl_next;
    ref_tbl_next = (ref_tbl_next + 1) % 3;
    free(ref_tbl[q]);
    ref_tbl[q] = (t_laplacian *)calloc((size_t)NX * NY, sizeof(t_laplacian));
    ref_tbl_base[q] = base;
    fflush(stdout);
    const int keep = dup(1), null = open("/dev/null", O_WRONLY);
    dup2(null, 1);
    init_laplacian_coords(ref_tbl[q], base);
    fflush(stdout);
    dup2(keep, 1);
    close(keep);
    close(null);
    return ref_tbl[q];
}

/* the table as counts and flat indices (-1 where the builder set nothing) */
void ref_laplacian_table(int *nn, int *nb)
{
    double *f = (double *)calloc((size_t)NX * NY, sizeof(double));
    t_laplacian *L = ref_table_for(f);
    for (long k = 0; k < (long)NX * NY; k++) {
        nn[k] = L[k].nneighb;
        for (int q = 0; q < 4; q++) nb[4 * k + q] = L[k].nghb[q] ? (int)(L[k].nghb[q] - f) : -1;
    }
    for (int q = 0; q < 3; q++)
        if (ref_tbl[q] == L) { free(L); ref_tbl[q] = NULL; ref_tbl_base[q] = NULL; }
    free(f);
}

/* the L-shape lines i1, j1, i2, j2, i3, j3 as init_laplacian_coords computes them */
void ref_lshape_lines(int *lsh)
{
    int ij[2];
    xy_to_ij(-LAMBDA, -1.0, ij);
    lsh[0] = ij[0] + 1; lsh[1] = ij[1] + 1;
    xy_to_ij(0.0, 0.0, ij);
    lsh[2] = ij[0] - 1; lsh[3] = ij[1] - 1;
    xy_to_ij(LAMBDA, 1.0, ij);
    lsh[4] = ij[0] - 1; lsh[5] = ij[1] - 1;
}

/* the corner drive of wave_3d.c:575-579 */
double ref_corner_drive(int time) { return AMPLITUDE * cos((double)time * OMEGA); }

/* evolve_wave_half_new, wave_3d.c:345, with the table built for phi_in (as
 * evolve_wave passes laplace / laplace1 / laplace2, wave_3d.c:1019-1021) */
void ref_evolve_wave_half_new(double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                              double *tc, double *tcc, double *tgamma)
{
    evolve_wave_half_new(phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma, ref_table_for(phi_in));
}
#elif !defined(FLAT_TWIN)
/* evolve_wave_half, wave_billiard.c:302 */
void ref_evolve_wave_half(double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                          double *tcc, double *tc, double *tgamma)
{
    double *a[NX], *b[NX], *c[NX], *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi_in + (size_t)i * NY; b[i] = psi_in + (size_t)i * NY; c[i] = phi_out + (size_t)i * NY;
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    evolve_wave_half(a, b, c, m, t1, t2, t3, ref_neumann);
}

/* evolve_wave, wave_billiard.c:613 */
void ref_evolve_wave(double *phi, double *psi, double *tmp, short *xy_in, double *tcc, double *tc,
                     double *tgamma)
{
    double *a[NX], *b[NX], *c[NX], *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi + (size_t)i * NY; b[i] = psi + (size_t)i * NY; c[i] = tmp + (size_t)i * NY;
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    evolve_wave(a, b, c, m, t1, t2, t3, ref_neumann);
}

/* compute_energy, wave_common.c:453 */
double ref_compute_energy(double *phi, double *psi, short *xy_in, int i, int j)
{
    double *a[NX], *b[NX];
    short int *m[NX];
    for (int r = 0; r < NX; r++) {
        a[r] = phi + (size_t)r * NY; b[r] = psi + (size_t)r * NY; m[r] = xy_in + (size_t)r * NY;
    }
    return compute_energy(a, b, m, i, j);
}
#else
/* evolve_wave_half, wave_3d.c:600 (table order tc, tcc, tgamma) */
void ref_evolve_wave_half_flat(double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                               double *tc, double *tcc, double *tgamma)
{
    evolve_wave_half(phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma, ref_neumann);
}
#endif
