/*
 * wavehip.h -- C ABI of libwavehip.so, the MI355X (gfx950) stepper for the
 * masked leapfrog update of the 2D linear wave equation.
 *
 * It replaces the time-stepping core of nilsberglund-orleans/YouTube-simulations:
 *   evolve_wave_half  wave_billiard.c:302-610   (row pointers, double *phi[NX])
 *   evolve_wave       wave_billiard.c:613-631   (3 steps, (phi,psi,tmp) rotation)
 *   evolve_wave_half  wave_3d.c:600-1007        (flat k = i*NY + j)
 *   evolve_wave       wave_3d.c:1010-1031
 * The reference fixes NX, NY, B_COND, TWOSPEEDS, KAPPA ... as compile-time
 * macros and keeps the step counter in function statics; here they live in a
 * context.  include/wavehip_compat.h maps the reference signatures onto this
 * ABI for a host program that keeps its macros.
 *
 * Conventions: every array is flat, k = i*NY + j (i = x, slow; j = y,
 * contiguous), host memory, caller-owned.  Every int-returning call returns
 * WV_OK (0) or a negative WV_E* code and then wv_last_error() says why.
 * A context belongs to one host thread.  Device state stays resident between
 * calls; wv_step() is asynchronous on the context's HIP stream, wv_get_*,
 * wv_energy() and wv_sync() wait for it.
 */
#ifndef WAVEHIP_H
#define WAVEHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WV_OK            0
#define WV_EINVAL       -1   /* bad argument or parameter combination       */
#define WV_EHIP         -2   /* HIP runtime error (message in wv_last_error) */
#define WV_ESTATE       -3   /* call out of order (no tables / no state)     */
#define WV_EUNSUPPORTED -4   /* reference feature not provided (see text)    */
#define WV_ENOMEM       -5
#define WV_EIO          -6   /* file could not be written                    */

/* B_COND codes, global_pdes.c:270-274 */
#define WV_BC_DIRICHLET   0  /* zero-flux outer edges (3 neighbours - 3x)  */
#define WV_BC_PERIODIC    1
#define WV_BC_ABSORBING   2
#define WV_BC_VPER_HABS   3
#define WV_BC_ABS_REFLECT 4  /* wave_3d.c; equals ABSORBING without COMPARISON */
#define WV_BC_LSHAPE     10  /* global_pdes.c:276, neighbour-table stepper only */

/* storage / arithmetic form */
#define WV_FORM_FP64      0  /* binary64 leapfrog, bit-identical to the reference */
#define WV_FORM_F32SPLIT  1  /* binary32 double-float u = hi + lo plus velocity v */

typedef struct wv_ctx wv_ctx;

typedef struct wv_params {
    int nx, ny;            /* NX, NY (>= 3 each)                                   */
    int bcond;             /* B_COND, WV_BC_*                                      */
    int twospeeds;         /* TWOSPEEDS                                            */
    int stencil;           /* 5 = reference Laplacian; 9 = isotropic 9-point bulk   */
    int form;              /* WV_FORM_*                                            */
    int oscillate_topbot;  /* OSCILLATE_TOPBOT                                     */
    int tb_shift;          /* tb_shift of the reference (wave_billiard.c:319)      */
    double kappa;          /* KAPPA                                                */
    double kappa_sides;    /* KAPPA_SIDES                                          */
    double gamma_sides;    /* GAMMA_SIDES                                          */
    double kappa_topbot;   /* KAPPA_TOPBOT                                         */
    double gamma_topbot;   /* GAMMA_TOPBOT                                         */
    double e_scale;        /* E_SCALE of compute_energy (wave_common.c:469)        */
    int tb_depth;          /* steps fused per launch: 0 = tuned default, 1..8 fixed */
    int rows_per_tile;     /* rows per tile of a fused launch: 0 = default (boundary tiles may take fewer) */
    int device;            /* HIP device ordinal                                   */
    int flags;             /* WV_FLAG_*                                            */
} wv_params;

/* drop-in entry points upload the tables on the first call only */
#define WV_FLAG_TABLES_FIXED 1
/* FP64 form: evaluate -KAPPA*x - tgamma*(x-y) at bulk cells even when KAPPA and
 * every tgamma are 0.  Without the flag those zero terms are left out for states
 * with no -0, infinity or NaN (checked at wv_set_state*), which gives the same
 * bits for every finite field; the flag also reproduces the reference's NaN
 * where a field that overflows meets a zero coefficient. */
#define WV_FLAG_FULL_FORMULA 2
/* COMPARISON of the flat program (wave_3d.c:736-990): the grid is two halves,
 * j < NY/2 and j >= NY/2, each with its own top and bottom rows (the
 * interface columns NY/2-1 and NY/2 take the B_COND rules of wave_3d.c:873-982,
 * the periodic wraps of the outer rows stay inside their half).  FP64 form,
 * 5-point, NY >= 4; one step per launch. */
#define WV_FLAG_COMPARISON 4

/* Fill *p with the reference's defaults for an nx x ny grid (B_COND 0, 5-point,
 * FP64, all constants 0, e_scale 1, tb_shift nx, tuned tiling). */
void wv_params_default(wv_params *p, int nx, int ny);

int  wv_create(wv_ctx **out, const wv_params *p);
void wv_destroy(wv_ctx *c);
const char *wv_last_error(const wv_ctx *c);  /* never NULL; c may be NULL */

/* Tables of init_ior_2d (sub_wave.c:12164) and the domain mask of
 * init_wave_flat (wave_common.c:294).  Cells are updated as the reference
 * does: interior cells when TWOSPEEDS or xy_in > 0 (wave_billiard.c:365),
 * edge cells when TWOSPEEDS or xy_in != 0; other cells are never written. */
int wv_set_tables(wv_ctx *c, const short *xy_in, const double *tcc, const double *tc,
                  const double *tgamma);

/* Variable media of init_ior_2d's Gaussian-well models (IOR_PERIODIC_WELLS,
 * IOR_RANDOM_WELLS, IOR_POISSON_WELLS, IOR_PPP_WELLS, sub_wave.c:12315-12555),
 * the sum computed on the GPU: at cell (i, j), (x, y) of ij_to_xy
 * (sub_wave.c:469-476) over box = (XMIN, XMAX, YMIN, YMAX),
 *   s = sum_n exp(-r_n^2/sigma) * height[n]   (height NULL: without the factor),
 *   s = tanh(s) if use_tanh, tcc = courant*s + courantb*(1 - s), tgamma = gamma,
 * then tc = sqrt(tcc) where xy_in != 0, courantb where TWOSPEEDS (0 elsewhere),
 * sub_wave.c:12970-12981.  The well centres and heights are the program's
 * (drawn with its own generator).  The tables are installed as by
 * wv_set_tables and copied to tcc / tc / tgamma (NX*NY) when not NULL.
 * exp and tanh are the GPU's: tcc agrees with the reference to about 1e-16. */
int wv_set_wells_tables(wv_ctx *c, const short *xy_in, int nwells, const double *xc, const double *yc,
                        const double *height, double sigma, int use_tanh, double courant, double courantb,
                        double gamma, const double box[4], double *tcc, double *tc, double *tgamma);

/* The initial drop of init_circular_wave (wave_common.c:89-111) and the
 * pulses of add_circular_wave (wave_common.c:337-351), computed on the GPU in
 * the context's state, so a program that keeps its field on the GPU adds the
 * frame loop's pulses (wave_billiard.c:916-934) without a round trip.  At the
 * cell corner (x, y) of ij_to_xy (sub_wave.c:469-476) over box = (XMIN,
 * XMAX, YMIN, YMAX), d2 = (x - x0)^2 + (y - y0)^2, on the cells the drawing
 * rule reaches (xy_in != 0 of the installed mask, or TWOSPEEDS):
 *   init: phi = amp*exp(-d2/variance)*cos(-sqrt(d2)/wavelength), 0 on the
 *         other cells, psi = 0 (a state, as wv_set_state would install it);
 *   add:  phi += amp*factor*exp(-d2/variance)*cos(-sqrt(d2)/wavelength),
 *         psi unchanged (a kick of the velocity).
 * amp, variance, wavelength are INITIAL_AMP, INITIAL_VARIANCE,
 * INITIAL_WAVELENGTH.  The mask must be installed (wv_set_tables, unlike the
 * reference's init, which also computes xy_in: wv_build_mask does that).
 * exp, cos and sqrt are the GPU's (within an ulp or two of glibc's).  Split
 * form: applied to u = hi + lo in binary64 and to the velocity.  Asynchronous
 * on the context stream. */
int wv_init_circular_wave(wv_ctx *c, double x0, double y0, double amp, double variance, double wavelength,
                          const double box[4]);
int wv_add_circular_wave(wv_ctx *c, double factor, double x0, double y0, double amp, double variance,
                         double wavelength, const double box[4]);

/* Domain codes of xy_in_billiard_single_domain (global_pdes.c:18-50) that
 * wv_build_mask evaluates */
#define WV_D_SINAI        3
#define WV_D_EXT_ELLIPSE  199
#define WV_D_CIRCLES      20

/* The billiard mask of init_wave_flat (wave_common.c:294-319), computed on
 * the GPU for the context's (first) grid: xy_in[i*NY + j] =
 * xy_in_billiard_single_domain(x, y) (sub_wave.c:5712) at the cell corner
 * (x, y) of ij_to_xy (sub_wave.c:469-476) over box = (XMIN, XMAX, YMIN, YMAX):
 * D_SINAI 1 outside the disc of radius lambda (5774-5779), D_EXT_ELLIPSE 1
 * outside the ellipse of half-axes lambda, mu (5744-5749), D_CIRCLES 0 inside
 * any active circle of the list (6216-6227; circles = xc[n], yc[n],
 * radius[n], active[n] for n < ncircles, 4*ncircles doubles, active != 0).
 * The mask is returned, not installed: pass it on to wv_set_tables or
 * wv_set_wells_tables with the media tables. */
int wv_build_mask(wv_ctx *c, int domain, const double box[4], double lambda, double mu, int ncircles,
                  const double *circles, short *xy_in);

/* wv_build_mask with no device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight frame loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The circle list and the mask
 * pass through one fixed scratch array of 8 MiB per device, kept until the
 * process ends (at most 32,768 circles; rows of at most 2,097,152 cells; any
 * NX, in bands of rows); larger inputs are refused with WV_EINVAL.  Every
 * call on a device shares that array, so calls that overlap on one device
 * (several threads) race; a call waits for its own work before it returns.
 * Same arguments and results as wv_build_mask. */
int wv_build_mask_static(wv_ctx *c, int domain, const double box[4], double lambda, double mu, int ncircles,
                         const double *circles, short *xy_in);

/* The macros xy_in_billiard_single_domain (sub_wave.c:5712) reads besides
 * LAMBDA and MU, by their reference names (wave_billiard.c:73-109), and the
 * Julia parameter (julia_x, julia_y: global_pdes.c:567). */
typedef struct wv_domain_params {
    double lambda, mu, mu_b;          /* LAMBDA, MU, MU_B                       */
    double apoly;                     /* APOLY: turn, in units of pi/2          */
    double wall_width, wall_width_b;  /* WALL_WIDTH, WALL_WIDTH_B               */
    double julia_scale;               /* JULIA_SCALE                            */
    double julia_x, julia_y;          /* julia_x, julia_y                       */
    double mandellimit;               /* MANDELLIMIT                            */
    int npoly;                        /* NPOLY (1..4096)                        */
    int mdepth, mratio;               /* MDEPTH (0..30), MRATIO (>= 1)          */
    int mandellevel;                  /* MANDELLEVEL                            */
    int ngridx, ngridy;               /* NGRIDX, NGRIDY (disk grids, lattices)  */
    double wall_width_asym;           /* WALL_WIDTH_ASYM (non-isotropic lattices) */
    double wall_width_asym_b;         /* WALL_WIDTH_ASYM_B                      */
    double oscil_ymid;                /* OSCIL_YMID (D_TREE_BEAM)               */
    double michelson_position;        /* michelson_position (global_pdes.c:570), the
                                         moving mirror of D_MICHELSON_MOVING    */
} wv_domain_params;

/* The billiard mask of init_wave_flat (wave_common.c:294-319) on the GPU for
 * the context's (first) grid, as wv_build_mask, for every domain of
 * xy_in_billiard_single_domain (sub_wave.c:5712-7602) that is a function of
 * (x, y) and the parameters alone.  Codes are the reference's B_DOMAIN codes
 * (global_pdes.c:14-236):
 *   D_NOTHING 999, D_RECTANGLE 0, D_ELLIPSE 1, D_CIRCLE_LAYERS 196,
 *   D_EXT_ELLIPSE 199, D_EXT_ELLIPSE_CURVED 198, D_EXT_ELLIPSE_CURVED_BDRY 197,
 *   D_STADIUM 2, D_SINAI 3, D_DIAMOND 4, D_TRIANGLE 5, D_FLAT 6, D_ANNULUS 7,
 *   D_POLYGON 8, D_YOUNG 9, D_GRATING 10, D_EHRENFEST 11, D_DISK_GRID 12,
 *   D_DISK_HEX 13, D_MENGER 15, D_JULIA_INT 16, D_MENGER_ROTATED 17,
 *   D_PARABOLA 18, D_TWO_PARABOLAS 19, D_TWO_PARABOLAS_ASYM 191,
 *   D_TWO_PARABOLAS_CLOSED_SEMITRANS 195, D_TWO_PARABOLAS_CLOSED_LENS 1951,
 *   D_TWO_PARABOLAS_CLOSED_LENS_R 1952, D_CIRCLES 20, D_CIRCLES_IN_RECT 201,
 *   D_CIRCLES_NEUMANN 202, D_MANDELBROT 24, D_JULIA 25, D_MANDELBROT_CIRCLE 26,
 *   D_FOUR_PARABOLAS 31, D_POLY_PARABOLAS 32, D_PENROSE 33, D_HYPERBOLA 34,
 *   D_LSHAPE 50, D_WAVEGUIDE 51, D_WAVEGUIDES_COUPLED 522, D_WAVEGUIDE_S 523,
 *   D_WAVEGUIDE_S_SHORT 524, D_WAVEGUIDES_COUPLED_N 525, D_LENS 61,
 *   D_LENS_WALL 62, D_TWO_LENSES_WALL 63, D_TWO_LENSES_OBSTACLE 64,
 *   D_LENS_ROTATED 67, D_LENS_CONCAVE 68, D_LENS_CONVEX_CONCAVE 69, D_TREE 73,
 *   D_MICHELSON 74, D_GRADIENT_INDEX_LENS 76, D_TWOCIRCLES 79,
 *   D_POLYCIRCLES 791, D_TWO_ELLIPSES 795, D_TWO_RECTANGLES 796,
 *   D_CARDIOID 90, D_NEPHROID 91, D_TOKARSKY 35, D_FRESNEL 43,
 *   D_DOUBLE_FRESNEL 45, D_NOISEPANEL 44, D_NOISEPANEL_RECT 441, D_QRD 46,
 *   D_QRD_ASYM 461, D_CIRCLE_SEGMENT 47, D_GROOVE 48, D_FABRY_PEROT 49,
 *   D_WAVEGUIDE_W 52, D_WAVEGUIDES_W 521, D_WAVEGUIDE_BADSPLICE 526,
 *   D_CHESSBOARD 55, D_TRIANGLE_TILES 56, D_HEX_TILES 57, D_FUNNELS 58,
 *   D_ONE_FUNNEL 581, D_LENSES_RING 59, D_FRESNEL_ZONE_PLATE 65,
 *   D_FRESNEL_ZONE_PLATE_INV 66, D_TREE_BEAM 731,
 *   D_RITCHEY_CHRETIEN_SPHERICAL 75, D_RITCHEY_CHRETIEN_HYPERBOLIC 751,
 *   D_MAGNETRON 78, D_MAGNETRON_CATHODE 781, D_POLYCIRCLES_ANGLED 792,
 *   D_TWOCIRCLES_ELLIPSE 793, D_CIRCLES_ELLIPSE 794, D_TWO_POLYGONS 797,
 *   D_TWO_POLY_PARABOLAS 798, D_CIRCLE_LATTICE 93, D_CIRCLE_LATTICE_NONISO 934,
 *   D_CIRCLE_LATTICE_HALF 935, D_CIRCLE_LATTICE_HALF_V 936,
 *   D_CIRCLE_LATTICE_STRIP 937, D_CIRCLE_LATTICE_HEX 94,
 *   D_CIRCLE_LATTICE_HEX_NONISO 942, D_CIRCLE_LATTICE_HEX_STRIP 947,
 *   D_DISC_WAVEGUIDE 98, D_DISC_WAVEGUIDE_SHIFTED 981,
 *   D_DISC_WAVEGUIDE_ELLIPSE 982, D_WAVEGUIDE_ELLIPSE_OBLIQUE 983,
 *   D_CIRCLE_PAIRS 99, D_CIRCLE_POLYGON 991, D_ANNULUS_HEATED 21,
 *   D_MENGER_HEATED 22, D_MENGER_H_OPEN 23, D_MICHELSON_MOVING 741,
 *   D_IMAGE 77 (1 everywhere: the program's mask comes from its image).
 * Values are the reference's (0, 1 and, for some domains, 2; the heated
 * Menger gaskets 22 and 23 give k + 2 in a hole of depth k).  The one-time
 * constants the reference keeps in static variables (D_TRIANGLE_TILES,
 * D_HEX_TILES, D_LENSES_RING, D_RITCHEY_CHRETIEN_HYPERBOLIC) are those of a
 * program that uses this one domain.  The circle
 * list (4*ncircles doubles, as wv_build_mask) is read by the D_CIRCLES
 * family.  Cosines, sines and tangents of fixed angles (polygon sides, PI/6
 * of D_TREE) are taken on the host with the C library, as the reference
 * does (the zone plates' per-cell cosine enters by its sign only); the
 * per-cell arithmetic is IEEE binary64 without contraction, so masks equal
 * the reference's cell for cell.  Other codes return WV_EUNSUPPORTED (they
 * read lists that other reference functions fill: polygons, mazes, images,
 * random lattices). */
int wv_build_domain_mask(wv_ctx *c, int domain, const double box[4], const wv_domain_params *p, int ncircles,
                         const double *circles, short *xy_in);

/* xy_in_billiard with COMPARISON (sub_wave.c:7593-7602), the mask of the
 * flat program's two halves (WV_FLAG_COMPARISON): domain B_DOMAIN with the
 * circle list circles where y > 0, domain_b B_DOMAIN_B with circles_b where
 * y <= 0 (both from the codes and with the parameters of
 * wv_build_domain_mask), at the cell corners of ij_to_xy over box. */
int wv_build_comparison_mask(wv_ctx *c, int domain, int domain_b, const double box[4], const wv_domain_params *p,
                             int ncircles, const double *circles, int ncircles_b, const double *circles_b,
                             short *xy_in);

/* The lists the remaining domains of xy_in_billiard_single_domain read,
 * which the program's init functions fill (globals of global_pdes.c:533-545
 * and its circle list): each a count and a flat array holding, for every
 * field in turn, that many values; a count of 0 leaves the pointer unread. */
typedef struct wv_domain_lists {
    int ncircles;  const double *circles;   /* xc, yc, radius, active (circles, ncircles)        */
    int nrects;    const double *rects;     /* x1, y1, x2, y2 (polyrect, npolyrect)              */
    int nrectrot;  const double *rectrot;   /* x1, y1, x2, y2, width (polyrectrot, npolyrect_rot) */
    int npolygons; const double *polygons;  /* xc, yc, radius, angle, nsides, active (polygons)  */
    int npolyline; const double *polyline;  /* x, y (polyline, npolyline)                        */
} wv_domain_lists;

/* wv_build_domain_mask for the domains that read those lists, and for every
 * code wv_build_domain_mask takes (the circle list from lists):
 *   D_POLYGONS 40, D_VONKOCH 41, D_STAR 42, D_STAR_CHANNEL 421, D_MAZE 53,
 *   D_MAZE_SMALL 531, D_MAZE_CLOSED 54, D_MAZE_CHANNELS 541,
 *   D_TWO_PARABOLAS_ASYM_GUIDE 192, D_TWO_PARABOLAS_SEMITRANS 193,
 *   D_TWO_PARABOLAS_LENS 194, D_CIRCLE_LATTICE_RANDOM 931,
 *   D_CIRCLE_LATTICE_HONEY 95, D_CIRCLE_LATTICE_RHOMBUS 97,
 *   D_CIRCLE_LATTICE_POISSON 96, D_CIRCLE_LATTICE_SQUARE_CROSS 932,
 *   D_CIRCLE_LATTICE_MAZE 933, D_CIRCLE_LATTICE_HEX_MAZE 941,
 *   D_VONKOCH_HEATED 27 (sub_wave.c:5903-7590).  D_POLYGONS reads npolygons polygons, the disc
 * lattices 96, 932, 933 and 941 every circle of the list (radius MU, the
 * active flag unread, as the reference reads them).  D_STAR needs NPOLY + 1
 * vertices, D_STAR_CHANNEL NPOLY + 4, D_VONKOCH and D_VONKOCH_HEATED MDEPTH <= 15; polygons 1-4096
 * sides; at most 1,048,576 entries a list.  Sines and cosines of the polygon
 * sides are taken on the host, as for the regular polygons; masks equal the
 * reference's cell for cell.  Others return WV_EUNSUPPORTED. */
int wv_build_list_domain_mask(wv_ctx *c, int domain, const double box[4], const wv_domain_params *p,
                              const wv_domain_lists *lists, short *xy_in);

/* The macros init_ior_2d (sub_wave.c:12164-13004) reads, by reference name. */
typedef struct wv_ior_params {
    double courant, courantb;   /* COURANT, COURANTB                          */
    double gamma, gammab;       /* GAMMA, GAMMAB                              */
    double lambda, mu;          /* LAMBDA, MU                                 */
    double mandel_ior_scale;    /* MANDEL_IOR_SCALE                           */
    double mandellimit;         /* MANDELLIMIT                                */
    int mandellevel;            /* MANDELLEVEL                                */
    int mdepth;                 /* MDEPTH (IOR_LUNEBURG_LAYERS, >= 1)         */
} wv_ior_params;

/* Media codes of wv_set_ior_tables: the reference's IOR codes
 * (global_pdes.c:239-264), and 0 for VARIABLE_IOR == 0 */
#define WV_IOR_CONSTANT            0
#define WV_IOR_MANDELBROT          1
#define WV_IOR_MANDELBROT_LIN      100
#define WV_IOR_EARTH               2
#define WV_IOR_LENS_WALL           10
#define WV_IOR_LENS_OBSTACLE       11
#define WV_IOR_LENS_CONCAVE        12
#define WV_IOR_LENS_CONVEX_CONCAVE 13
#define WV_IOR_TREE                14
#define WV_IOR_WAVE_GUIDES_COUPLED   15
#define WV_IOR_WAVE_GUIDES_COUPLED_B 151
#define WV_IOR_WAVE_GUIDE_COATED   16
#define WV_IOR_MICHELSON           17
#define WV_IOR_GRADIENT_INDEX_LENS   18
#define WV_IOR_GRADIENT_INDEX_LENS_B 181
#define WV_IOR_LUNEBURG_LENS       182
#define WV_IOR_LUNEBURG_LAYERS     183
#define WV_IOR_LINEAR_X_A          19
#define WV_IOR_LINEAR_X_B          191

/* init_ior_2d (sub_wave.c:12164) computed on the GPU and installed as by
 * wv_set_tables: tcc and tgamma of every cell by the model's rule at the cell
 * corner (x, y) of ij_to_xy (sub_wave.c:469-476) over box, reading xy_in
 * where the model does (lens, tree, wave-guide, Michelson models); then tc =
 * sqrt(tcc) where xy_in != 0, COURANTB where TWOSPEEDS (sub_wave.c:12970-12981).
 * WV_IOR_CONSTANT is the VARIABLE_IOR == 0 branch (12984-13003: COURANT,
 * COURANT^2, GAMMA or GAMMAB by xy_in; COURANTB, COURANTB^2, GAMMAB on the
 * outside cells when TWOSPEEDS).  Cells the reference leaves unwritten get 0.
 * The tables are copied to tcc / tc / tgamma (NX*NY) when not NULL.  All
 * arithmetic is binary64 without contraction: tables equal the reference's
 * except IOR_MANDELBROT's log (the device library's, within an ulp or two of
 * glibc's; the escape count itself is exact).  Returns WV_EUNSUPPORTED for
 * the Gaussian-well models 4-9 (wv_set_wells_tables, with wv_rotating_wells
 * for 6 and 7), for IOR_EXPLO_LENSING (3: the reference indexes its row
 * pointer array xy_in[i*NY+j], sub_wave.c:12287, past its NX entries) and for
 * codes init_ior_2d does not handle. */
int wv_set_ior_tables(wv_ctx *c, const short *xy_in, int ior, const double box[4], const wv_ior_params *p,
                      double *tcc, double *tc, double *tgamma);

/* The well centres of IOR_PERIODIC_WELLS_ROTATING (large = 0) and
 * IOR_PERIODIC_WELLS_ROTATING_LARGE (large = 1), sub_wave.c:12390-12481, at
 * the rotation angle ior_angle: ngridx*ngridy centres into xc, yc (index
 * n = j*ngridx + i) and sigma_stat into *sigma.  Host code, no device; pass
 * them to wv_set_wells_tables (height NULL, use_tanh 0) to rebuild the media
 * every REFRESH_IOR frame (wave_billiard.c:939-945) on the GPU. */
int wv_rotating_wells(int ngridx, int ngridy, int large, double ior_angle, const double box[4], double *xc,
                      double *yc, double *sigma);

/* ---- TIFF frames: writetiff, sub_wave.c:62-123 ----------------------------
 * The file writetiff(filename, description, 0, 0, width, height, compression)
 * writes through libtiff (4.x): RGB, 8 bits a sample, one row per strip,
 * Orientation BOTLEFT, the rows of glReadPixels (bottom row first, 3 bytes a
 * pixel, GL_PACK_ALIGNMENT 1) stored as lines height-1 .. 0, in libtiff's
 * layout, byte for byte.  rgb is that host buffer.  Under WV_TIFF_LZW (what
 * every call site of the reference asks for) the strips are LZW-coded on the
 * device, one workgroup a strip; WV_TIFF_NONE stores the rows as they are.
 * wv_tiff_bound gives the output size to provide (0 for a frame or width out
 * of range: 1..WV_TIFF_MAX_WIDTH x 1..65535); *len receives the file size.
 * wv_tiff_write writes the file (WV_EIO when it cannot), where writetiff
 * returns 1. */
#define WV_TIFF_NONE      1
#define WV_TIFF_LZW       5
#define WV_TIFF_MAX_WIDTH 8192
size_t wv_tiff_bound(int width, int height, const char *description);
int wv_tiff_encode(wv_ctx *c, const unsigned char *rgb, int width, int height, const char *description,
                   int compression, unsigned char *out, size_t cap, size_t *len);
int wv_tiff_write(wv_ctx *c, const char *filename, const char *description, const unsigned char *rgb, int width,
                  int height, int compression);

/* u^n (phi) and u^{n-1} (psi); the third buffer of the rotation starts at 0. */
int wv_set_state(wv_ctx *c, const double *phi, const double *psi);
/* all three buffers of evolve_wave's rotation (phi, psi, tmp) */
int wv_set_state3(wv_ctx *c, const double *phi, const double *psi, const double *tmp);
int wv_get_state(wv_ctx *c, double *phi, double *psi);
int wv_get_state3(wv_ctx *c, double *phi, double *psi, double *tmp);

/* Soft point sources (no reference counterpart).  After step g (g = 1, 2, ...
 * counted from this call), source k adds amp[k*amp_len + g-1] to u at
 * (i[k], j[k]) while g <= amp_len.  Source cells must be active cells.
 * n = 0 removes all sources. */
int wv_set_sources(wv_ctx *c, int n, const int *i, const int *j, const double *amp,
                   long amp_len);

/* Reference features outside the plain stepper (FP64 form; set them before
 * wv_set_state -- changed later, the context steps one step per launch until
 * the next wv_set_state).  In one step the order is that of
 * evolve_wave_half: bulk cells, Neumann points, edge cells, driven cells,
 * FLOOR, then the point sources of wv_set_sources.
 *
 * Neumann ghost points, wave_billiard.c:379-389 (t_neumann_bc of
 * global_pdes.c:518-528 as plain arrays, active points only): point q at
 * (i[q], j[q]) takes nx[q]*u(nbi1,nbj1) + ny[q]*u(nbi2,nbj2) of the previous
 * time level; an edge cell that the edge formulas write keeps the edge value.
 * Both neighbours must lie in the grid within one cell of the point (what
 * init_neumann_bc, sub_wave.c:13720-13760, builds).  n = 0 removes them. */
int wv_set_neumann(wv_ctx *c, int n, const int *i, const int *j, const int *nbi1, const int *nbj1,
                   const int *nbi2, const int *nbj2, const double *nx, const double *ny);
/* Driven cells -- the OSCILLATE_LEFT edge and corners, wave_billiard.c:397 and
 * 585-590, with the schedule (oscillating_bc, sub_wave.c:11965) evaluated by
 * the caller: after step g (g = 1, 2, ... counted from this call) cell k is
 * set to val[k*len + g-1] while g <= len.  A later call with the same cells
 * only replaces the values.  n = 0 removes them. */
int wv_set_drive(wv_ctx *c, int n, const int *i, const int *j, const double *val, long len);
/* FLOOR, wave_billiard.c:592-609: after each step every cell with xy_in != 0
 * is clamped to [-vmax, vmax] (with fused steps the cells no step writes hold
 * +0); vmax = 0 turns it off. */
int wv_set_floor(wv_ctx *c, double vmax);

/* Probes -- the time series of wave_billiard.c:900-908 (SAVE_TIME_SERIES
 * writes phi at two cells after every evolve_wave call): after every
 * every-th step, counted from this call, the value of each cell (i[k],
 * j[k]) is recorded on the GPU, whatever the number of steps fused per
 * launch; up to `capacity` samples per probe (FP64 form).  n = 0 removes
 * them. */
int wv_set_probes(wv_ctx *c, int n, const int *i, const int *j, long every, long capacity);
/* the samples recorded since wv_set_probes or the last call, out[s*n + k]
 * for s < *count (at most max_samples), then counting restarts after them;
 * *lost (may be NULL): samples that did not fit the capacity */
int wv_get_probes(wv_ctx *c, double *out, long max_samples, long *count, long *lost);
/* The lines SAVE_TIME_SERIES writes for one probe (wave_billiard.c:902-905:
 * (long)(phi*1.0e16) printed with "%019ld\n"), for the samples v[s*stride],
 * s < count (stride n picks probe k of wv_get_probes' out + k).  Writes at
 * most cap bytes to out (no terminating zero) and the byte count to
 * *written; WV_EINVAL if the lines do not fit (at most 21 bytes each).  Host
 * code only: no context, no device. */
int wv_format_time_series(const double *v, long count, long stride, char *out, long cap, long *written);

/* Neighbour-table stepping, the reference's PRECOMPUTE_BC path:
 * evolve_wave_half_new (wave_3d.c:345-597) with the t_laplacian table built
 * by init_laplacian_coords (sub_wave.c:11555-11892) and read by
 * compute_laplacian (sub_wave.c:11895-11963).  nneighb[k] and nghb[4k+q]
 * (k = i*NY + j) are a table entry's count and the flat indices of its
 * neighbours in the field (the reference's pointers minus the field's base).
 * Where the stepper uses them, entries are summed as -n*x + neighbour 0 + ...
 * in table order: edge cells whose B_COND branch takes the bulk formula, and
 * for BC_LSHAPE (genus-2 wrap of the L-shaped billiard) every bulk cell (four
 * neighbours); other bulk cells keep (((E+W)+N)+S)-4x, and entries the
 * stepper never reads are not looked at.  The table stepper also updates
 * bulk cells with xy_in != 0 (not only > 0), has no edge formula for
 * BC_LSHAPE and no Neumann points; driven cells, FLOOR and point sources
 * apply as with the plain stepper.  FP64 form, 5-point, one step per launch.
 * nneighb = NULL returns to the plain stepper (not for BC_LSHAPE). */
int wv_set_laplacian(wv_ctx *c, const int *nneighb, const int *nghb);

/* Advance nsteps leapfrog steps: the 3-buffer rotation of evolve_wave
 * (wave_billiard.c:613-631) applied nsteps times, in fused launches of the
 * tiling's depth; asynchronous on the context stream. */
int wv_step(wv_ctx *c, long nsteps);
/* Do now the host-side work the first wv_step(c, nsteps) after a change of
 * tables, sources, point lists or tiling would do before its first launch:
 * build the tile tables of every fused depth the sequence uses and load
 * every stepping kernel it launches.  The state is unchanged; waits.  (No
 * counterpart in the reference, whose stepper has no set-up.) */
int wv_prepare(wv_ctx *c, long nsteps);
/* wv_step with a HIP event on the context stream before the first launch and
 * after every launch; waits; *ms is the device time from the first event to
 * the last and *launches the kernel launches in it */
int wv_step_timed(wv_ctx *c, long nsteps, double *ms, long *launches);
int wv_sync(wv_ctx *c);
long wv_steps_done(const wv_ctx *c);
/* Kernel time of the last wv_step_timed: launch q lasts from event q to event
 * q+1 (the gap between two launches counts to the later one); *avg_ms is the
 * average over every launch of the full fused depth and *timed_launches
 * their number.  A launch covers every tile for one fused block of steps. */
int wv_last_kernel_ms(const wv_ctx *c, double *avg_ms, double *timed_launches);

/* Total energy: compute_energy (wave_common.c:453-472) per cell with c^2 from
 * tcc, summed in FP64 in a fixed order (64 lane-strided partials per row, a
 * halving tree, then the same over rows): deterministic run to run. */
int wv_energy(wv_ctx *c, double *e);

/* Frame colours: what draw_wave_highres_palette (wave_common.c:1748-1847)
 * hands to OpenGL, computed on the GPU from the field in place.  For every
 * block cell (i, j) = (bi*size, bj*size): the value and colour of wave_value
 * (wave_common.c:1053-1171, plot types P_AMPLITUDE 0 .. P_LOG_AVERAGE_ENERGY 9
 * with their running sums), colour schemes C_LUM 0, C_HUE 1, C_ONEDIM 3,
 * C_ONEDIM_LINEAR 4, C_BASIC_LINEAR 5 (colors_waves.c:161-290), palettes
 * COL_JET 0, COL_HSLUV 1 (hsluv2rgb of hsluv.c:396-410, the HSLuv library the
 * reference vendors), COL_GRAY 2 and the table palettes COL_TURBO 10 .. COL_PARULA 16,
 * COL_TWILIGHT 17, COL_TWILIGHT_SHIFTED 18, COL_TURBO_CYCLIC 101
 * (colormaps.c:2077-2548; the program's own table, 256 or 511 entries, via
 * wv_frame_set_tables), then SHADE_2D shading (gradient towards blocks
 * (bi+1, bj) and (bi, bj+1)), FADE_IN_OBSTACLE, fade, and the draw test
 * TWOSPEEDS || xy_in.  rgba[4*(bi*nby + bj) ..] = (r, g, b, drawn) as the
 * floats glColor3f receives; nbx = ceil(NX/size), nby = ceil(NY/size).
 * Computed in FP64 like the reference; tanh, exp, log, atan are the GPU's
 * (within an ulp or two of glibc's). */
typedef struct {
    int plot;              /* PLOT (P_* code)                                     */
    int scheme;            /* COLOR_SCHEME                                        */
    int palette;           /* palette argument (COLOR_PALETTE or COLOR_PALETTE_B) */
    int asym;              /* COLOR_PALETTE >= COL_TURBO: asymmetric energy scheme */
    int size;              /* block size of draw_wave_highres_palette             */
    int time;              /* frame time argument                                 */
    int shade;             /* SHADE_2D                                            */
    int rescale_center;    /* RESCALE_COLOR_IN_CENTER (needs color_scale)         */
    int fade_in_obstacle;  /* FADE_IN_OBSTACLE (needs fade_table)                 */
    int fade;              /* fade argument, with fade_value                      */
    long nsteps;           /* NSTEPS (C_LUM hue drift)                            */
    double scale, fade_value;
    double slope, color_range, attenuation;        /* SLOPE, COLOR_RANGE, ATTENUATION */
    double vshift_amplitude, vscale_amplitude;     /* VSHIFT_AMPLITUDE, VSCALE_AMPLITUDE */
    double log_shift, log_scale, flux_scale, avrg_e_factor;
    double colorhue, colordrift, lummean, lumamp, huemean, hueamp;
    double shade_scale, light_x, light_y;          /* SHADE_SCALE_2D, light[0..1] */
    double courant, courantb;                      /* compute_energy's COURANT, COURANTB */
} wv_frame_params;

/* the values of the checked-in wave_billiard.c:200-241, 287 */
void wv_frame_params_default(wv_frame_params *f);
/* palette table (lut_n x 3 floats) or NULL; color_scale and fade_table
 * (NX*NY, k = i*NY + j) or NULL -- replaces all three */
int wv_frame_set_tables(wv_ctx *c, const float *lut, int lut_n, const double *color_scale,
                        const double *fade_table);
/* zero the running sums of the mean / average / flux plots */
int wv_frame_reset(wv_ctx *c);
/* one frame (updates the running sums of its plot type; a change of size
 * starts them from zero); rgba: 4*nbx*nby floats */
int wv_frame(wv_ctx *c, const wv_frame_params *f, float *rgba);
/* the same frame as an 8-bit colour buffer holds it: rgba[4*(bi*nby + bj) ..]
 * = (q(r), q(g), q(b), drawn ? 255 : 0), q(c) = (int)(min(max(c, 0), 1)*255 + 0.5)
 * in binary32 on the float colour; 4 bytes per block */
int wv_frame_rgba8(wv_ctx *c, const wv_frame_params *f, unsigned char *rgba);

/* What wv_step launches: fused steps per launch, rows per tile, tile columns */
int wv_get_tiling(const wv_ctx *c, int *tb_depth, int *rows_per_tile, int *tile_cols);
int wv_set_tiling(wv_ctx *c, int tb_depth, int rows_per_tile);
/* Tiles of one stepping launch of fused depth tb_depth (0: the current
 * one): wide tiles (split form: a workgroup of waves sharing their column
 * halos through LDS), one-wave tiles, workgroups launched, and the output
 * cells the tiles cover (a cell two tiles write counts twice).  No
 * counterpart in the reference (a tuning and test aid). */
int wv_get_tile_counts(wv_ctx *c, int tb_depth, int *wide, int *narrow, int *groups, long *cells);
/* Time candidate tilings on the device for the current state (state is
 * restored afterwards) and keep the fastest; *best_gcells is its throughput. */
int wv_autotune(wv_ctx *c, long steps_per_trial, double *best_gcells);

/* ---- drop-in forms of the reference functions (one context per program) ---
 * Each call uploads phi_in, psi_in and phi_out (so cells the step does not
 * write keep the caller's values), runs one step (or three) and writes the
 * result back into the caller's arrays, exactly as the reference does. */

/* evolve_wave_half, wave_billiard.c:302-304 */
int wv_evolve_wave_half_rows(wv_ctx *c, double *const *phi_in, double *const *psi_in,
                             double *const *phi_out, short *const *xy_in,
                             double *const *tcc, double *const *tc, double *const *tgamma);
/* evolve_wave, wave_billiard.c:613 */
int wv_evolve_wave_rows(wv_ctx *c, double *const *phi, double *const *psi, double *const *tmp,
                        short *const *xy_in, double *const *tcc_table,
                        double *const *tc_table, double *const *tgamma_table);
/* evolve_wave_half, wave_3d.c:600 (table order tc, tcc, tgamma) */
int wv_evolve_wave_half_flat(wv_ctx *c, double *phi_in, double *psi_in, double *phi_out,
                             short *xy_in, double *tc, double *tcc, double *tgamma);
/* evolve_wave, wave_3d.c:1010 (without the t_laplacian tables: PRECOMPUTE_BC 0) */
int wv_evolve_wave_flat(wv_ctx *c, double *phi, double *psi, double *tmp, short *xy_in,
                        double *tc, double *tcc, double *tgamma);

/* ---- batches of independent grids (no reference counterpart) ------------
 * n instances of any sizes, BCs and tables advance together: one launch per
 * fused block of steps covers the tiles of every instance, and the launch
 * sequence of a wv_batch_step is captured once into a HIP graph and replayed.
 * Every instance must use the same form and stencil. */
typedef struct wv_batch wv_batch;
int  wv_batch_create(wv_batch **out, int n, const wv_params *ps);
void wv_batch_destroy(wv_batch *b);
const char *wv_batch_last_error(const wv_batch *b);
int wv_batch_set_tables(wv_batch *b, int k, const short *xy_in, const double *tcc,
                        const double *tc, const double *tgamma);
int wv_batch_set_state(wv_batch *b, int k, const double *phi, const double *psi);
int wv_batch_get_state(wv_batch *b, int k, double *phi, double *psi);
int wv_batch_step(wv_batch *b, long nsteps);
/* wv_prepare for a batch; also captures the graph of an nsteps sequence */
int wv_batch_prepare(wv_batch *b, long nsteps);
int wv_batch_step_timed(wv_batch *b, long nsteps, double *ms, long *launches);
int wv_batch_last_kernel_ms(const wv_batch *b, double *avg_ms, double *timed_launches);
int wv_batch_energy(wv_batch *b, double *e /* n values */);
int wv_batch_set_tiling(wv_batch *b, int tb_depth, int rows_per_tile);
int wv_batch_get_tiling(const wv_batch *b, int *tb_depth, int *rows_per_tile, int *tile_cols);
/* wv_init_circular_wave / wv_add_circular_wave on grid k of a batch */
int wv_batch_init_circular_wave(wv_batch *b, int k, double x0, double y0, double amp, double variance,
                                double wavelength, const double box[4]);
int wv_batch_add_circular_wave(wv_batch *b, int k, double factor, double x0, double y0, double amp,
                               double variance, double wavelength, const double box[4]);
/* wv_set_wells_tables, wv_set_ior_tables and wv_build_domain_mask on grid k of
 * a batch (the media installed for that grid only; the mask is returned for
 * that grid's NX x NY) */
int wv_batch_set_wells_tables(wv_batch *b, int k, const short *xy_in, int nwells, const double *xc,
                              const double *yc, const double *height, double sigma, int use_tanh, double courant,
                              double courantb, double gamma, const double box[4], double *tcc, double *tc,
                              double *tgamma);
int wv_batch_set_ior_tables(wv_batch *b, int k, const short *xy_in, int ior, const double box[4],
                            const wv_ior_params *p, double *tcc, double *tc, double *tgamma);
int wv_batch_build_domain_mask(wv_batch *b, int k, int domain, const double box[4], const wv_domain_params *p,
                               int ncircles, const double *circles, short *xy_in);
int wv_batch_build_list_domain_mask(wv_batch *b, int k, int domain, const double box[4], const wv_domain_params *p,
                                    const wv_domain_lists *lists, short *xy_in);
/* wv_autotune over all grids of the batch (one tiling for every launch) */
int wv_batch_autotune(wv_batch *b, long steps_per_trial, double *best_gcells);

#ifdef __cplusplus
}
#endif
#endif
