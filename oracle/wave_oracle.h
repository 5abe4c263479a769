/*
 * wave_oracle.h -- CPU oracle for the 2D wave leapfrog stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use it to check and time against.
 *
 * It restates, in plain C + OpenMP, the reference's masked leapfrog update
 *   evolve_wave_half   wave_billiard.c:302-610  (row-pointer layout)
 *   evolve_wave        wave_billiard.c:613-631  (3-buffer rotation)
 *   evolve_wave_half   wave_3d.c:600-1007       (flat twin, adds BC_ABS_REFLECT)
 * with the reference's compile-time macros (NX, NY, B_COND, TWOSPEEDS, KAPPA,
 * KAPPA_SIDES, GAMMA_SIDES, KAPPA_TOPBOT, GAMMA_TOPBOT, OSCILLATE_TOPBOT) turned
 * into the runtime struct below.  Same loop order, same left-to-right
 * expression order, same edge / corner quirks, so results are bit-identical
 * to the reference compiled with its own flags (gcc -O3 -fopenmp, no FMA).
 *
 * Forms with no reference counterpart (9-point bulk Laplacian, soft point
 * sources, the FP64 energy sum order, the FP32 double-float "split" state)
 * are defined here once; the HIP kernels implement the same arithmetic.
 */
#ifndef WAVE_ORACLE_H
#define WAVE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* B_COND codes, global_pdes.c:270-275 */
#define ORC_BC_DIRICHLET  0
#define ORC_BC_PERIODIC   1
#define ORC_BC_ABSORBING  2
#define ORC_BC_VPER_HABS  3
#define ORC_BC_ABS_REFLECT 4

typedef struct {
    int nx, ny;               /* NX, NY: i (x) is the slow index, j (y) contiguous   */
    int bcond;                /* B_COND                                             */
    int twospeeds;            /* TWOSPEEDS                                          */
    int stencil;              /* 5 (reference) or 9 (build-defined bulk Laplacian)  */
    int oscillate_topbot;     /* OSCILLATE_TOPBOT                                   */
    int tb_shift;             /* tb_shift, wave_billiard.c:319                      */
    double kappa;             /* KAPPA                                              */
    double kappa_sides;       /* KAPPA_SIDES                                        */
    double gamma_sides;       /* GAMMA_SIDES                                        */
    double kappa_topbot;      /* KAPPA_TOPBOT                                       */
    double gamma_topbot;      /* GAMMA_TOPBOT                                       */
    int comparison;           /* COMPARISON (flat twin, wave_3d.c:736-990): two     */
                              /* halves, interface columns NY/2-1 and NY/2          */
} orc_params;

/* 1.0/6.0: weight of the build-defined 9-point Laplacian
 * lap9 = ((4*(E+W+N+S) + (NE+NW+SE+SW)) - 20*x) * ORC_W9 */
#define ORC_W9 (1.0 / 6.0)

/* OpenMP threads the parallel loops use */
int orc_num_threads(void);
/* OpenMP threads of later parallel regions; returns the previous count */
int orc_set_num_threads(int n);

/* One leapfrog step on row pointers (wave_billiard.c:302-610). */
void orc_evolve_wave_half(const orc_params *p,
                          double *const *phi_in, double *const *psi_in, double *const *phi_out,
                          short *const *xy_in, double *const *tcc, double *const *tc,
                          double *const *tgamma);

/* One leapfrog step on flat k=i*NY+j arrays; table order (tc, tcc, tgamma)
 * as in wave_3d.c:600. */
void orc_evolve_wave_half_flat(const orc_params *p,
                               double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                               double *tc, double *tcc, double *tgamma);

/* Three steps with the (phi,psi,tmp) rotation of wave_billiard.c:613-631,
 * flat arrays. */
void orc_evolve_wave_flat(const orc_params *p, double *phi, double *psi, double *tmp,
                          short *xy_in, double *tcc, double *tc, double *tgamma);

/* nsteps steps on three flat buffers b[0..2].  *state is the rotation index s:
 * phi = b[s], psi = b[(s+1)%3], tmp = b[(s+2)%3]; each step writes tmp and
 * sets s = (s+2)%3, exactly the rotation of evolve_wave.  After step number
 * g (g = step0 + m + 1, m = 0..nsteps-1) and when g-1 < amp_len, every source
 * k adds amp[k*amp_len + g-1] to the new phi at (si[k], sj[k]). */
void orc_run(const orc_params *p, double *b0, double *b1, double *b2, int *state,
             long nsteps, const short *xy_in, const double *tcc, const double *tc,
             const double *tgamma, int nsrc, const int *si, const int *sj,
             const double *amp, long amp_len, long step0);

/* Reference features outside the plain stepper, applied by orc_run_x in the
 * order of evolve_wave_half (wave_billiard.c:302-610):
 *   bulk (361-377) -> Neumann ghost points (379-389) -> edges (393-583,
 *   where a driven left-edge cell takes its drive value instead of the edge
 *   formula, 397) -> driven corners (585-590) -> FLOOR (592-609)
 * and then the build-defined point sources.
 * Neumann point k: phi_out[ni][nj] = nnx*phi_in[nbi1][nbj1] + nny*phi_in[nbi2][nbj2]
 * (two roundings, as phi_out = a; phi_out += b).  Driven cell k: after step
 * g (g = step0 + m + 1), while g-1 < dlen, phi_out[di][dj] = dval[k*dlen+g-1]
 * (the reference computes oscillating_bc(time, j) there, sub_wave.c:11965).
 * FLOOR (vmax > 0): every cell with xy_in != 0 is clamped to [-vmax, vmax]. */
typedef struct orc_extras {
    int n_neu;
    const int *ni, *nj, *nbi1, *nbj1, *nbi2, *nbj2;
    const double *nnx, *nny;
    int n_drv;
    const int *di, *dj;
    const double *dval;
    long dlen;
    double vmax;
} orc_extras;

void orc_run_x(const orc_params *p, double *b0, double *b1, double *b2, int *state,
               long nsteps, const short *xy_in, const double *tcc, const double *tc,
               const double *tgamma, int nsrc, const int *si, const int *sj,
               const double *amp, long amp_len, long step0, const orc_extras *x);

/* ---- neighbour-table path (PRECOMPUTE_BC) --------------------------------
 * The reference's second stepper, evolve_wave_half_new (wave_3d.c:345-597),
 * takes the Laplacian's neighbours from a table (t_laplacian,
 * global_pdes.c:493-497: a count and up to four pointers into the field the
 * table was built for).  Here an entry is nneighb[k] and nghb[4k+q], the flat
 * index i*NY+j of neighbour q in whichever buffer is phi_in.
 *
 * orc_laplacian_table restates init_laplacian_coords (sub_wave.c:11555-11892)
 * for B_COND 0-3 and BC_LSHAPE (genus-2 wrap of the L-shaped billiard, lines
 * i1, i2, i3 / j1, j2, j3 given in lsh[6], computed by the reference from
 * LAMBDA with xy_to_ij).  Entries it leaves unset (edges of B_COND 4 and
 * BC_LSHAPE) are nneighb 0, indices -1. */
#define ORC_BC_LSHAPE 10 /* global_pdes.c:276 */

void orc_laplacian_table(const orc_params *p, const int lsh[6], int *nneighb, int *nghb);

/* nsteps steps of evolve_wave_half_new with the rotation of orc_run (the
 * rotation of evolve_wave, wave_3d.c:1015-1021) and the same drive / FLOOR /
 * point-source conventions as orc_run_x (no Neumann points in this path).
 * In one step: Laplacian (compute_laplacian, sub_wave.c:11895-11963: bulk
 * (((E+W)+N)+S)-4x, or for BC_LSHAPE -4x plus the four table neighbours in
 * order; edge cells -n*x plus their n table neighbours in order), bulk cells
 * with TWOSPEEDS or xy_in != 0, left / right / top / bottom edge cells per
 * B_COND (no edge formula for BC_LSHAPE), driven cells, FLOOR, sources.
 * The reference leaves the Laplacian of cells with xy_in == 0 unset and reads
 * it when TWOSPEEDS is on; here it is computed like any other cell. */
void orc_run_table(const orc_params *p, double *b0, double *b1, double *b2, int *state,
                   long nsteps, const short *xy_in, const double *tcc, const double *tc,
                   const double *tgamma, const int *nneighb, const int *nghb, int nsrc,
                   const int *si, const int *sj, const double *amp, long amp_len, long step0,
                   const orc_extras *x);

/* ---- frame colours (the drawing path's colour computation) ---------------
 * orc_frame restates what draw_wave_highres_palette (wave_common.c:1748-1847)
 * computes before it hands colours to OpenGL: for every block cell (i, j),
 * i and j multiples of `size`, the plotted value and colour of wave_value
 * (wave_common.c:1053-1171: plot types P_AMPLITUDE .. P_LOG_AVERAGE_ENERGY,
 * with the running accumulators of the mean / average / flux plots) through
 * color_scheme_palette / color_scheme_asym_palette (colors_waves.c:161-290)
 * and the palettes of colormaps.c (hsl_to_rgb_jet 2077-2114,
 * hsl_to_rgb_palette 2343-2447, amp_to_rgb_palette 2449-2548); then the
 * SHADE_2D pseudo-3D shading, FADE_IN_OBSTACLE and fade, and the
 * (TWOSPEEDS || xy_in) draw test.
 * Table palettes (COL_TURBO .. COL_PARULA: 256 entries, COL_TWILIGHT and
 * COL_TWILIGHT_SHIFTED: 511, COL_TURBO_CYCLIC: the 256-entry Turbo table)
 * read the caller's table lut[3*n] (the program's own colour table).
 * Shading takes the gradient between block cells (i+size, j) and (i, j+size);
 * the reference reads cells (i+1, j) and (i, j+1), which it only fills when
 * size is 1.  Indices below 0 that amp_to_rgb_palette would form are
 * clamped to 0. */
typedef struct {
    int plot, scheme, palette;   /* PLOT, COLOR_SCHEME, palette argument       */
    int asym;                    /* COLOR_PALETTE >= COL_TURBO (wave_value)    */
    int size, time, shade, rescale_center, fade_in_obstacle, fade;
    long nsteps;                 /* NSTEPS (C_LUM hue drift)                   */
    double scale, fade_value;
    double slope, color_range, attenuation;        /* SLOPE, COLOR_RANGE, ATTENUATION */
    double vshift_amplitude, vscale_amplitude;     /* VSHIFT/VSCALE_AMPLITUDE   */
    double log_shift, log_scale, flux_scale, avrg_e_factor;
    double colorhue, colordrift, lummean, lumamp, huemean, hueamp;
    double shade_scale, light_x, light_y;          /* SHADE_SCALE_2D, light[2]  */
    double courant, courantb, e_scale;             /* compute_energy constants  */
} orc_frame_params;

/* Colours of one frame on flat k = i*NY + j arrays.  total_energy,
 * average_energy (NX*NY) and total_flux (2*NX*NY, index 2*j*NX + 2*i) are
 * the running accumulators (updated); color_scale and fade_table may be
 * NULL when unused.  Outputs at block cells: values[k], rgb[3k..3k+2]
 * (double, before the float conversion of glColor3f) and drawn[k]. */
void orc_frame(const orc_params *p, const orc_frame_params *f, const double *phi, const double *psi,
               const short *xy_in, double *total_energy, double *average_energy, double *total_flux,
               const double *color_scale, const double *fade_table, const float *lut, int lut_n,
               double *values, double *rgb, unsigned char *drawn);

/* the colour of one value: color_scheme_palette (asym = 0) or
 * color_scheme_asym_palette (asym = 1) */
void orc_color(const orc_frame_params *f, int scheme, int asym, double value, double scale, int time,
               const float *lut, int lut_n, double rgb[3]);

/* ---- media tables ----------------------------------------------------------
 * init_ior_2d's Gaussian-well models (sub_wave.c:12315-12555): at cell (i, j),
 * (x, y) from ij_to_xy over box = (XMIN, XMAX, YMIN, YMAX),
 *   s = sum_n exp(-r_n^2/sigma) * h[n]   (h = NULL: sum_n exp(-r_n^2/sigma)),
 *   s = tanh(s) if use_tanh, tcc = courant*s + courantb*(1 - s), tgamma = gamma,
 * then tc = sqrt(tcc) where xy_in != 0, courantb where TWOSPEEDS and
 * xy_in == 0, untouched elsewhere (sub_wave.c:12970-12981). */
/* init_circular_wave (wave_common.c:89-111) with mode 0: phi =
 * amp*exp(-d2/variance)*cos(-sqrt(d2)/wavelength) where xy != 0 or
 * TWOSPEEDS, 0 elsewhere, psi = 0; add_circular_wave (337-351) with mode 1:
 * phi += amp*factor*exp(...)*cos(...) on the same cells, psi untouched.
 * d2 from the cell corner of ij_to_xy over box to (x0, y0); glibc libm. */
void orc_circular_wave(const orc_params *p, int mode, double factor, double x0, double y0, double amp,
                       double variance, double wavelength, const double box[4], double *phi, double *psi,
                       const short *xy);
void orc_wells_tables(const orc_params *p, const short *xy, int nw, const double *xc, const double *yc,
                      const double *h, double sigma, int use_tanh, double courant, double courantb,
                      double gamma, const double box[4], double *tcc, double *tc, double *tg);

/* the billiard mask of init_wave_flat (D_SINAI 3, D_EXT_ELLIPSE 199,
 * D_CIRCLES 20), flat out[i*NY + j]; -1 for another domain */
int orc_mask(int NX, int NY, const double box[4], int domain, double lambda, double mu, int ncirc,
             const double *circ, short *out);

/* HSLuv to sRGB (hsluv2rgb of the hsluv-c library the reference vendors,
 * hsluv.c:396-410): h in degrees, s and l in 0..100, rgb in 0..1 */
void orc_hsluv_rgb(double h, double s, double l, double rgb[3]);

/* The macros of xy_in_billiard_single_domain (sub_wave.c:5712) the closed-form
 * domains read; same layout as the library's wv_domain_params. */
typedef struct {
    double lambda, mu, mu_b, apoly, wall_width, wall_width_b, julia_scale, julia_x, julia_y, mandellimit;
    int npoly, mdepth, mratio, mandellevel, ngridx, ngridy;
    double wall_width_asym, wall_width_asym_b, oscil_ymid, michelson_position;
} orc_domain_params;

/* xy_in_billiard_single_domain(x, y) at the cell corners of ij_to_xy over
 * box (sub_wave.c:469-476), flat out[i*NY + j], for the closed-form domains
 * (codes of global_pdes.c:14-236); -1 for another code. */
int orc_domain(int NX, int NY, const double box[4], int domain, const orc_domain_params *dp, int ncirc,
               const double *circ, short *out);

/* The lists of the list-driven domains, layout of wv_domain_lists: flat
 * runs of n values per field. */
typedef struct {
    int ncircles;  const double *circles;   /* xc, yc, radius, active        */
    int nrects;    const double *rects;     /* x1, y1, x2, y2 (polyrect)      */
    int nrectrot;  const double *rectrot;   /* x1, y1, x2, y2, width          */
    int npolygons; const double *polygons;  /* xc, yc, radius, angle, nsides, active */
    int npolyline; const double *polyline;  /* x, y                           */
} orc_domain_lists;

/* xy_in_billiard_single_domain for the domains that read those lists
 * (polyrect, polyrectrot, polygons, polyline, the global circle list);
 * -1 for another code or a list too short for the code */
int orc_list_domain(int NX, int NY, const double box[4], int domain, const orc_domain_params *dp,
                    const orc_domain_lists *L, short *out);

/* xy_in_billiard with COMPARISON (sub_wave.c:7593-7602): domain and circ
 * where y > 0, domain_b and circ_b where y <= 0; -1 for an unknown code */
int orc_comparison_mask(int NX, int NY, const double box[4], int domain, int domain_b, const orc_domain_params *dp,
                        int nc, const double *circ, int nb, const double *circ_b, short *out);

/* The macros of init_ior_2d (sub_wave.c:12164); layout of wv_ior_params. */
typedef struct {
    double courant, courantb, gamma, gammab, lambda, mu, mandel_ior_scale, mandellimit;
    int mandellevel, mdepth;
} orc_ior_params;

/* init_ior_2d's per-cell models (IOR codes of global_pdes.c:239-264; 0 is the
 * VARIABLE_IOR == 0 branch) on zeroed flat tables, then the tc rule
 * (sub_wave.c:12970-12981); cells the reference leaves unwritten stay 0.
 * Returns -1 for another code. */
int orc_ior_tables(const orc_params *p, const short *xy, int ior, const double box[4], const orc_ior_params *ip,
                   double *tcc, double *tc, double *tg);

/* the rotating-well centres (sub_wave.c:12390-12481) at ior_angle: n =
 * j*ngridx + i; returns sigma_stat */
double orc_rotating_wells(int ngridx, int ngridy, int large, double ior_angle, const double box[4], double *xc,
                          double *yc);

/* Per-cell energy, compute_energy wave_common.c:453-472, with the Courant
 * constants given explicitly (c2_in = COURANT^2 for xy_in != 0, c2_out =
 * COURANTB^2 used when TWOSPEEDS). */
double orc_energy_cell(const orc_params *p, const double *phi, const double *psi,
                       const short *xy_in, double courant, double courantb,
                       double e_scale, int i, int j);

/* Build-defined global energy: per-cell energy with c^2 taken from tcc
 * (cells with xy_in != 0, or all cells if TWOSPEEDS; 0 elsewhere; with
 * COMPARISON no gradient across the halves, wave_common.c:1965-1971), summed
 * in a fixed order: for each row, 64 lane-strided partials then a halving
 * tree (32,16,...,1); the NX row sums are reduced the same way. */
double orc_energy_total(const orc_params *p, const double *phi, const double *psi,
                        const short *xy_in, const double *tcc, double e_scale);
/* Same order over a precomputed per-cell energy array e[NX*NY]. */
double orc_fixed_order_sum(int nx, int ny, const double *e);

/* ---- FP32 double-float ("split") form, build-defined ----------------------
 * State per cell: u = hi + lo (two floats, |lo| <= ulp(hi)/2) and the
 * velocity v = u^n - u^{n-1} as one float.  Interior active cells:
 *   D(a) = (h_a - hc) + (l_a - lc)   (a difference of u = h + l, one
 *          rounding); D(W), D(S) are taken by the west / south neighbour
 *          towards the cell: D(W) = (hc - hW) + (lc - lW)
 *   lap  = (D(E) - D(W)) + (D(N) - D(S))
 *          (9-point: lap = 4*that + ((D(NE) - D(NW)) + (D(SE) - D(SW))),
 *          NE = (i+1, j+1), NW = (i-1, j+1), unscaled)
 *   (tcc_hi, tcc_lo) = binary32 pair of tcc*w, w = 1 (5-point) or ORC_W9
 *          (9-point), w applied in binary64
 *   vn   = ((vc + fmaf(tcc_hi, lap, tcc_lo*lap)) - kappa*hc) - gamma*vc
 *          (a term whose coefficient is zero is not evaluated)
 *   t = lc + vn;  hi' = hc + t;  lo' = t - (hi' - hc)   (Fast2Sum)
 * Edge cells use the FP64 reference formula on u = (double)hi + (double)lo,
 * u^{n-1} = u - (double)v.  Inactive cells are untouched. */
void orc_split_from_fp64(long n, const double *u_n, const double *u_nm1,
                         float *hi, float *lo, float *v);
void orc_split_to_fp64(long n, const float *hi, const float *lo, const float *v,
                       double *u_n, double *u_nm1);
void orc_split_run(const orc_params *p, float *hi, float *lo, float *v, long nsteps,
                   const short *xy_in, const double *tcc, const double *tc,
                   const double *tgamma, int nsrc, const int *si, const int *sj,
                   const double *amp, long amp_len, long step0);
double orc_split_energy_total(const orc_params *p, const float *hi, const float *lo,
                              const float *v, const short *xy_in, const double *tcc,
                              double e_scale);

#ifdef __cplusplus
}
#endif
/* ---- TIFF frames (the frame writer writetiff, sub_wave.c:62-123) --------
 * orc_lzw_strip: one strip coded as libtiff's LZW coder does (out needs
 * 3n/2 + 16 bytes); returns its length.  orc_tiff: the whole file that
 * writetiff(name, description, 0, 0, width, height, compression) writes for
 * the glReadPixels rows rgb (bottom row first, 3 bytes a pixel, no row
 * padding); compression 1 (none) or 5 (LZW); out holds orc_tiff_bound()
 * bytes; returns the file length or -1. */
long orc_lzw_strip(const unsigned char *in, long n, unsigned char *out);
long orc_tiff_bound(int width, int height, const char *description);
long orc_tiff(const unsigned char *rgb, int width, int height, const char *description, int compression,
              unsigned char *out, long cap);

#endif
