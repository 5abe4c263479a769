// wv_internal.h -- device-side data layout shared by the kernels and the host
// code of libwavehip.so.
//
// Layout in HBM, per grid ("instance"):
//   state      FP64 form: four NX x ldy arrays of double (three rotate like
//              evolve_wave's phi/psi/tmp, the fourth lets a fused launch write
//              two time levels out of place).  Split form: two ping-pong sets
//              of three NX x ldy float arrays (hi, lo, v).  ldy = NY rounded
//              up to 64 elements, so every row starts 512-byte aligned.
//   act        1 bit per cell, uint64 words, ldw = ldy/64 words per row:
//              cell is written by a step (interior: TWOSPEEDS || xy_in > 0,
//              edge: TWOSPEEDS || xy_in != 0).
//   eact       1 bit per cell: cell counts in the energy (xy_in != 0 || TWOSPEEDS).
//   med        1 bit per cell: which of two (tcc, tgamma) pairs applies
//              (coefficient mode WV_COEF_TWO only).
//   tccf, tgf  per-cell coefficient tables (mode WV_COEF_FIELD only).
//   tc         per-cell tc (read at edge cells only).
//   sources    CSR by row: src_ptr[NX+1], src_col[], src_id[]; amplitudes
//              amp[id*amp_len + g].
//   Neumann points, driven cells, table Laplacians: CSR by row, columns
//              ascending within a row (binary search).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WV_COEF_UNIFORM 0
#define WV_COEF_TWO 1
#define WV_COEF_FIELD 2

// coefficient modes a stepping kernel is built for: one mode (the value of
// WV_COEF_*) or any mixture (batches); a kernel built for fewer modes holds
// fewer code paths and needs no more registers than those paths do
#define WV_CM_ALL 3
// any mixture, plus Neumann ghost points, driven cells and FLOOR (FP64 form)
#define WV_CM_ALLX 4
// WV_CM_ALLX plus the neighbour-table Laplacian and COMPARISON (FP64 form,
// one step per launch, 5-point)
#define WV_CM_TBL 5

// rows in flight in a stepping kernel's prefetch ring (3 or 6)
#ifndef WV_RING
#define WV_RING 3
#endif

// cells per lane along j in the stepping kernels (a wave covers 64*WV_C columns)
#ifndef WV_C
#define WV_C 1
#endif

struct WvInst {
    // FP64 form: u[0..3]; split form: hi[2], lo[2], v[2]
    double *u[4];
    float *hi[2], *lo[2], *v[2];
    const uint64_t *act, *eact, *med;
    const double *tccf, *tgf, *tc;       // per-cell tables (FP64)
    const float *tccf_hi, *tccf_lo, *tgf32;  // per-cell tables (split form)
    const int *src_ptr, *src_col, *src_id;
    const double *amp;
    long amp_len;
    int nx, ny, ldy, ldw;
    int bc, osc_tb, tb_shift, coef;
    int rows_per_tile, ncb;              // tiling of this instance
    int damped;                          // KAPPA or some tgamma is non-zero
    int fp64_full;                       // FP64 bulk cells evaluate -KAPPA*x - tgamma*(x-y)
                                         // even with zero coefficients (see wv_api.cpp)
    int tile0;                           // first tile index of this instance
    // FP64 form only: Neumann ghost points (CSR by row: column, packed
    // neighbour offsets, weights nx, ny), driven cells (CSR by row: column,
    // id; values drv_val[id*drv_len + g - drv_base]) and the FLOOR clamp
    // (vmax > 0; nz = xy_in != 0, 1 bit per cell)
    const int *neu_ptr, *neu_col, *neu_code;
    const double *neu_w;
    const int *drv_ptr, *drv_col, *drv_id;
    const double *drv_val;
    long drv_len, drv_base;
    const uint64_t *nz;
    double vmax;
    // neighbour-table stepping (evolve_wave_half_new): tbl = 0 off, 1 on,
    // 2 BC_LSHAPE.  Cells whose Laplacian is not the plain one, CSR by row:
    // column, neighbour count, four neighbour offsets i*ldy + j.
    int tbl;
    int cmp;                             // COMPARISON: two halves (one step per launch)
    const int *tbl_ptr, *tbl_col, *tbl_n;
    const long long *tbl_off;
    // probes (FP64 form): cells whose value is recorded after every
    // prb_every-th step counted from prb_base, CSR by row (column, id); sample
    // s of probe k at prb_buf[s*prb_n + k] while s < prb_cap
    const int *prb_ptr, *prb_col, *prb_id;
    double *prb_buf;
    int prb_n;
    long prb_every, prb_cap, prb_base;
    double tcc0, tcc1, tg0, tg1;         // coefficient pairs (modes UNIFORM / TWO)
    // split form: binary32 pairs of tcc0 and tcc1 times the stencil weight
    // (1 for 5-point, 1/6 for 9-point), and tgamma in binary32
    float tcch0, tccl0, tcch1, tccl1, tgf0, tgf1;
    double kappa, ks, gs, kt, gt, es2;
};

// one tile of a stepping launch: instance, column block, row strip, flags
#define WV_TILE_INTERIOR 1  // rows and columns the tile reads never reach the outer box
#define WV_TILE_ALLACT 2    // ... and every one of those cells is written by a step
#define WV_TILE_SRC 4       // a point source lies in the cells the tile computes
#define WV_TILE_NEU 8       // ... a Neumann ghost point (never an INTERIOR tile)
#define WV_TILE_DRV 16      // ... a driven cell (never an INTERIOR tile)
#define WV_TILE_PRB 32      // ... a probe (never an INTERIOR tile)
#define WV_TILE_EDGEJ 64    // not INTERIOR only because its columns reach column 0 or NY-1, or
                            // its rows row 0 or NX-1 of a grid not periodic along i
#define WV_TILE_WIDE 128    // an INTERIOR tile computed by the WV_WG waves of one workgroup
// One tile: output columns [jb, jb + width) of grid inst, rows [i0, i0 +
// rows), with width = (flags >> WV_TILE_WIDTH_SHIFT) & 255 and rows = flags
// >> WV_TILE_ROWS_SHIFT (the low bits hold WV_TILE_*).  A one-wave tile's
// lanes cover columns jb - H ... jb - H + 64C - 1 (H = wv_halo(K) halo
// columns per side); a WIDE tile's wave w covers jb - K + 62w ... jb - K +
// 62w + 63, its lanes 0 and 63 mirroring lane 62 of wave w-1 and lane 1 of
// wave w+1 (exchanged through LDS at every time level).
#define WV_TILE_WIDTH_SHIFT 8
#define WV_TILE_ROWS_SHIFT 16
struct WvTile {
    int inst, jb, i0, flags;
};

// waves per workgroup of the split-form stepping kernel: a WIDE tile, or up
// to WV_WG one-wave tiles side by side (independent of each other)
#define WV_WG 4

// Stepping kernels with the edge-column path (tiles flagged WV_TILE_EDGEJ
// evaluate their cells as interior cells and then their edge cells by the
// edge formulas; elsewhere such tiles take the boundary path): fused depth 4
// and more, the 5-point stencil, grids of one or two media (cm: the
// kernel's coefficient modes)
static constexpr bool wv_edgej_kernel(int k, int stencil, int cm)
{
    return k >= 4 && stencil == 5 && (cm == WV_COEF_UNIFORM || cm == WV_COEF_TWO || cm == WV_CM_ALL);
}

// number of output columns of one wave for a fused depth k
static inline int wv_halo(int k) { return (k + WV_C - 1) / WV_C * WV_C; }
static inline int wv_tile_cols(int k) { return 64 * WV_C - 2 * wv_halo(k); }
// number of output columns of a WIDE tile (WV_C == 1 only): 62 per wave,
// less K - 1 halo columns on either side
static inline int wv_wide_cols(int k) { return 62 * WV_WG - 2 * k + 2; }

// Launchers (wv_step_*.hip, dispatched in wv_energy.hip).  cm: WV_COEF_* of
// every grid of the launch, or WV_CM_ALL.  Ti/ni: interior
// tiles, Tb/nb: boundary tiles (two kernels, one after the other).  ia/ib:
// buffers holding u^n, u^{n-1}; oc/od: buffers receiving u^{n+k}, u^{n+k-1}
// (od unused when k == 1).  Split form: ia = input set, oc = output set.
// The amplitude index of the first step is *gstep + g0.
int wv_launch_fp64(int k, int stencil, int cm, const WvInst *d_inst, const WvTile *Ti, int ni, const WvTile *Tb,
                   int nb, const long *gstep, long g0, int ia, int ib, int oc, int od, hipStream_t s);
int wv_launch_split(int k, int stencil, int cm, const WvInst *d_inst, const WvTile *Ti, int ni, const WvTile *Tb,
                    int nb, const long *gstep, long g0, int ia, int oc, hipStream_t s);
// Persistent split-form launch (wv_step_split.hip): `epochs` launches of
// depth k over the ng workgroups of tile table T in one cooperative launch,
// epoch e reading set (ia0 + e) & 1; ctr: 8 * epochs * WV_PCTR_PAD ints of
// device memory (work counters, zeroed by the launch); *grid: workgroups
// launched
// ints per work counter (one counter per 128-byte line)
#define WV_PCTR_PAD 32
int wv_launch_split_persist(int k, int stencil, int cm, const WvInst *d_inst, const WvTile *T, int ng,
                            const long *gstep, long g0, int epochs, int ia0, int *ctr, hipStream_t s, int *grid);
bool wv_have_depth(int k);
// energy: rows table (inst,row) pairs; partial[nrows_total]; out[ninst]
int wv_launch_energy(int split, const WvInst *d_inst, const int2 *d_rows, int nrows,
                     double *d_partial, const int *d_row0, int ninst, double *d_out, int ia, int ib,
                     hipStream_t s);
int wv_launch_bump(long *gstep, long n, hipStream_t s);

// Frame colours (wv_frame.hip): one frame of the drawing path's colours for
// the block cells (i, j) = (bi*size, bj*size), bi < nbx, bj < nby.  Running
// sums are double-buffered: read from *0, written to *1 (compact block
// layout, tf two per block).  out[bi*nby + bj] = (r, g, b, drawn), or
// out8[...] = the same quantized to 8 bits.
struct WvFrame {
    int plot, scheme, palette, asym, size, time, shade, rescale, fadeobs, fade;
    long nsteps;
    double scale, fade_value, slope, color_range, attenuation, vshift, vscale, log_shift, log_scale;
    double flux_scale, avrg, colorhue, colordrift, lummean, lumamp, huemean, hueamp;
    double shade_scale, light_x, light_y, courant, courantb, es;
    int nx, ny, ldy, ldw, nbx, nby, twospeeds, lut_n;
    const double *phi, *psi;             // FP64 form
    const float *hi, *lo, *v;            // split form (phi = hi + lo, psi = phi - v)
    const uint64_t *nz;                  // xy_in != 0
    const double *cs, *ft;               // color_scale, fade_table (NX x ldy) or null
    const float *lut;                    // palette table, lut_n x 3
    const double *te0, *ae0, *tf0;
    double *te1, *ae1, *tf1;
    float4 *out;
    uchar4 *out8;                        // if set: 8-bit colours instead of out
};
int wv_launch_frame(const WvFrame &f, hipStream_t s);
// Gaussian-well media (wv_media.hip): tcc and tc of every cell of an NX x NY
// grid over (xmin, xmax, ymin, ymax), into arrays of row stride ldy; h may be
// null (unit heights); tcc_hi / tcc_lo (split form: the binary32 pair of
// tcc * split_weight) may be null
struct WvWells {
    int nx, ny, ldy, ldw, nw, use_tanh, twospeeds;
    double xmin, xmax, ymin, ymax, sigma, courant, courantb, split_weight;
    const double *xc, *yc, *h;
    const uint64_t *nz;
    double *tcc, *tc;
    float *tcc_hi, *tcc_lo;
};
int wv_launch_wells(const WvWells &a, hipStream_t s);
// Billiard masks (wv_media.hip): xy_in of every cell of an NX x NY grid over
// (xmin, xmax, ymin, ymax), flat out[i*NY + j]; kind 0 D_SINAI, 1
// D_EXT_ELLIPSE, 2 D_CIRCLES with circ = ncirc each of xc, yc, radius, active
// init_circular_wave / add_circular_wave (wv_media.hip) on a grid's device
// state: mode 0 sets phi = amp*exp(-d2/variance)*cos(-sqrt(d2)/wavelength)
// on drawn cells (xy_in != 0: nz plane, or TWOSPEEDS), 0 elsewhere, and psi
// = 0; mode 1 adds amp*factor*exp(...)*cos(...) to phi on drawn cells.
// FP64 form: phi, psi (row stride ldy); split form: both sets of hi, lo, v.
struct WvPulse {
    int nx, ny, ldy, ldw, mode, twospeeds;
    double xmin, xmax, ymin, ymax, x0, y0, amp, factor, variance, wavelength;
    const uint64_t *nz;
    double *phi, *psi;
    float *hi[2], *lo[2], *v[2];
};
int wv_launch_pulse(const WvPulse &a, hipStream_t s);
// Rows [i0, i0 + rows) only (the whole grid: i0 = 0, rows = nx), into
// out[(i - i0)*NY + j].
struct WvMask {
    int nx, ny, kind, ncirc, i0, rows;
    double xmin, xmax, ymin, ymax, lambda, mu;
    const double *circ;
    short *out;
};
// Every closed-form domain of xy_in_billiard_single_domain (wv_media.hip,
// k_domain): domain = the reference's code, the macros it reads, and the
// constants the host takes with the C library (cosines and sines of fixed
// angles, square roots of parameter expressions).  trig: ntrig (cos, sin)
// pairs of the domain's fixed angles (polygon sides), after the circle list
// in the same device buffer.  Rows [i0, i0 + rows) into out[(i - i0)*NY + j];
// with half set, only that half of the cells is written.
struct WvDomain {
    int nx, ny, domain, ncirc, i0, rows;
    int half;   // 0: every cell; 1: cells with y > 0 only; 2: y <= 0 only
    int npoly, mdepth, mratio, mandellevel, ngridx, ngridy;
    double xmin, xmax, ymin, ymax;
    double lambda, mu, mu_b, wall_width, wall_width_b, julia_scale, julia_x, julia_y, mandellimit;
    double wall_width_asym, wall_width_asym_b, oscil_ymid, michelson_position;
    double c0, c1, c2;   // per-domain host constants (see k_domain)
    // the list-driven domains: rectangles x1, y1, x2, y2; rotated ones x1, y1,
    // x2, y2, width; polygons xc, yc, radius, active, nsides, cos(omega/2) and
    // the offset of their sides' cosines and sines in trig; vertices x, y
    // (each field a run of n values)
    int nrect, nrectrot, npolyg, npline;
    const double *rect, *rectrot, *polyg, *pline;
    const double *circ, *trig;
    short *out;
};
int wv_launch_domain(const WvDomain &a, hipStream_t s);
// init_ior_2d's per-cell media models (wv_media.hip, k_ior): model = the
// reference's IOR code (0: VARIABLE_IOR == 0); tcc, tc, tgamma flat
// (k = i*NY + j); xy flat; a0 the model's host constant (lens wall abscissa)
struct WvIor {
    int nx, ny, model, mandellevel, mdepth, twospeeds;
    double xmin, xmax, ymin, ymax;
    double courant, courantb, gamma, gammab, lambda, mu, mandel_ior_scale, mandellimit, a0;
    const short *xy;
    double *tcc, *tc, *tg;
};
int wv_launch_ior(const WvIor &a, hipStream_t s);
int wv_launch_mask(const WvMask &a, hipStream_t s);
// The mask builder's fixed device scratch (wv_build_mask_static): its
// address on the current device (resolved once per device) and its size.
#define WV_MASK_SCRATCH_BYTES (8u << 20)
#define WV_MASK_SCRATCH_MAX_CIRCLES 32768
int wv_mask_scratch(int device, void **ptr);

// TIFF frames: LZW strips of n bytes (3*width), one per row, each into
// strips + r*stride (stride = wv_lzw_stride(n), a bound on one coded strip),
// lens[r] its length; then the strips packed at body + off[r]
long wv_lzw_stride(int n);
int wv_launch_lzw(const uint8_t *rgb, int n, int rows, uint8_t *strips, long stride, int *lens, hipStream_t s);
int wv_launch_tiff_pack(const uint8_t *strips, long stride, const int *lens, const long *off, int rows, uint8_t *body,
                        hipStream_t s);
