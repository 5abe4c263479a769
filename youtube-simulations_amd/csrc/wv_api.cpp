// wv_api.cpp -- host side of libwavehip.so: the C ABI of include/wavehip.h.
//
// One Engine drives n independent grids ("instances") that share a form
// (FP64 / FP32 split), a stencil and a tiling; a wv_ctx is an Engine with one
// instance plus the drop-in entry points, a wv_batch is an Engine with many.
// Each stepping launch covers the tiles of every instance (one tile = one
// wave64 streaming a strip of rows), so small grids of a batch fill the chip
// together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/wavehip.h"
#include "wv_internal.h"

namespace {

struct Grid {
    wv_params p{};
    int ldy = 0, ldw = 0;
    double *u[4] = {nullptr, nullptr, nullptr, nullptr};
    float *hi[2] = {nullptr, nullptr}, *lo[2] = {nullptr, nullptr}, *v[2] = {nullptr, nullptr};
    uint64_t *act = nullptr, *eact = nullptr, *med = nullptr;
    double *tccf = nullptr, *tgf = nullptr, *tc = nullptr;
    float *tccf_hi = nullptr, *tccf_lo = nullptr, *tgf32 = nullptr;
    int *src_ptr = nullptr, *src_col = nullptr, *src_id = nullptr;
    double *amp = nullptr;
    long amp_len = 0;
    int coef = WV_COEF_UNIFORM;
    double tcc0 = 0, tcc1 = 0, tg0 = 0, tg1 = 0;
    bool tables = false, state = false;
    bool exact = false;                 // masked cells not all +0: one step per launch
    // the state holds no -0, infinity or NaN, and the source amplitudes are
    // finite: undamped bulk cells may skip their zero damping terms
    bool clean_state = true, clean_amp = true;
    std::vector<uint64_t> h_act;        // host copy of the active plane
    std::vector<int> h_si, h_sj;        // source cells
    // FP64 form: Neumann ghost points, driven cells, FLOOR
    uint64_t *nz = nullptr;             // xy_in != 0, 1 bit per cell
    std::vector<uint64_t> h_nz;
    int *neu_ptr = nullptr, *neu_col = nullptr, *neu_code = nullptr;
    double *neu_w = nullptr;
    int *drv_ptr = nullptr, *drv_col = nullptr, *drv_id = nullptr;
    double *drv_val = nullptr;
    long drv_len = 0, drv_base = 0;
    double vmax = 0;
    std::vector<int> h_ni, h_nj, h_di, h_dj;
    // cells a step writes: h_act plus the Neumann points and driven cells
    std::vector<uint64_t> h_wr;
    // neighbour-table stepping (evolve_wave_half_new): 0 off, 1 on, 2 BC_LSHAPE;
    // the cells with a table Laplacian (CSR by row) and the active plane of
    // the plain stepper (h_act holds the one in use)
    int tbl = 0;
    int *tbl_ptr = nullptr, *tbl_col = nullptr, *tbl_n = nullptr;
    long long *tbl_off = nullptr;
    std::vector<uint64_t> h_act_plain;
    // frame colours: palette table, color_scale and fade_table (NX x ldy),
    // double-buffered running sums per block cell, output blocks
    float *fr_lut = nullptr;
    int fr_lut_n = 0;
    double *fr_cs = nullptr, *fr_ft = nullptr;
    double *fr_te[2] = {nullptr, nullptr}, *fr_ae[2] = {nullptr, nullptr}, *fr_tf[2] = {nullptr, nullptr};
    int fr_te_cur = 0, fr_ae_cur = 0, fr_tf_cur = 0, fr_size = 0;
    float4 *fr_out = nullptr;
    // probes: CSR by row (column, id), sample buffer, sampling period and
    // capacity, and where counting starts (device step counter; host steps)
    int *prb_ptr = nullptr, *prb_col = nullptr, *prb_id = nullptr;
    double *prb_buf = nullptr;
    std::vector<int> h_pi, h_pj;
    long prb_every = 0, prb_cap = 0, prb_base = 0, prb_host0 = 0;
};

struct Engine {
    int n = 0, device = 0, form = WV_FORM_FP64, stencil = 5;
    hipStream_t stream = nullptr;
    std::vector<Grid> g;
    int role[4] = {0, 1, 2, 3};  // FP64: phi, psi, free1 (tmp), free2
    int scur = 0;                // split: current set
    WvInst *d_inst = nullptr;
    // ni workgroups of tiles (WV_WG entries each in the split form, one in the FP64 form); nb unused (0)
    struct Tiles { WvTile *d = nullptr; int ni = 0, nb = 0; };
    std::map<int, Tiles> tiles;  // depth -> device tile tables
    int K = 0, L = 0;            // tiling (0 = untuned default)
    long steps_done = 0, last_launches = 0;
    long *d_gstep = nullptr;
    double *d_partial = nullptr, *d_eout = nullptr;
    int2 *d_erows = nullptr;
    int *d_row0 = nullptr;
    int nrows_total = 0;
    std::string err;
    // batch graph cache: (nsteps, rotation state) -> executable graph
    std::map<std::pair<long, long>, hipGraphExec_t> graphs;
    bool use_graph = false;
    // staging for the drop-in entry points
    std::vector<double> st[3], stt[3];
    std::vector<short> stm;
    bool dropin_tables = false;
    bool xcd_order = true;  // WAVEHIP_TILE_ORDER=linear turns the XCD interleave off
    bool slow_first = true;  // WAVEHIP_TILE_ORDER=inplace: slower tiles keep their place in the strip order
    // split form: the full-depth launches of a step sequence as one persistent
    // cooperative launch (WAVEHIP_PERSISTENT=1, one-wave tiles only; batches
    // then step without a graph)
    bool persistent = false;
    int *d_pctr = nullptr;  // its work counters (8 per epoch)
    long pctr_cap = 0;
    // split form: interior strips in wide tiles (WAVEHIP_WIDE=1); measured
    // slower than one-wave tiles on every workload, so off by default
    bool wide = false;
    // per-kernel timing of the last wv_step_timed: one event before the first
    // stepping launch and one after every launch (kev[q], kev[q+1] bracket
    // launch q; consecutive launches share an event).  These take timestamps
    // only (no system-scope fence between launches: that writes the caches
    // back and costs ~5 us a record at 4096^2); kev_end, recorded after the
    // last, is an ordinary event, the one the host waits for.
    bool time_kernels = false;
    std::vector<hipEvent_t> kev;
    hipEvent_t kev_end = nullptr;
    std::vector<int> kev_depth;  // fused depth of launch q
    std::vector<int> kev_epochs;  // launches of that depth it stands for (persistent: several)
    size_t kev_used = 0;         // events recorded (the last one follows the step counter's bump)
    bool kev_pending = false;    // per-launch durations of the last timed run not yet read
    double last_avg_ms = 0, last_timed = 0;
    // one workgroup of tiles of instance -1: a launch over it loads the kernel and returns
    WvTile *d_pad = nullptr;
    // TIFF frames: device rows, coded strips, their lengths and offsets, the
    // packed strips (grown to the largest frame seen)
    uint8_t *tf_rgb = nullptr, *tf_strips = nullptr, *tf_body = nullptr;
    int *tf_lens = nullptr;
    long *tf_off = nullptr;
    size_t tf_cap_rgb = 0, tf_cap_strips = 0, tf_cap_body = 0;
    int tf_cap_rows = 0;
};

int fail(Engine *e, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return code;
}

#define HIPCHK(e, call)                                                                          \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess)                                                                  \
            return fail(e, WV_EHIP, "%s failed: %s", #call, hipGetErrorString(err_));           \
    } while (0)

int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Fused steps per launch without tuning: 8 (measured best at 4096^2 for
// both forms); 6 on grids under 8 M cells and 4 under 2 M (at 1280x720 a
// launch lasts as long as its slowest boundary tile, which grows with K; a
// sweep of K and strip rows there finds the defaults within 4 % of the best)
int default_depth(const Engine *e)
{
    long cells = 0;
    for (const Grid &G : e->g) cells += (long)G.p.nx * G.p.ny;
    return cells < (2L << 20) ? 4 : (cells < (8L << 20) ? 6 : 8);
}
// Rows per strip without tuning: the measured best at 4096^2 (64 for both
// forms), shortened for smaller grids until the launch has enough tiles to
// fill the chip one and a quarter times over (256 CUs x 4 SIMDs x 4 waves)
int default_rows(const Engine *e)
{
    const int wout = wv_tile_cols(e->K > 0 ? e->K : default_depth(e));
    for (int L : {64, 48, 32, 24, 16, 12}) {
        long tiles = 0;
        for (const Grid &G : e->g) tiles += (long)((G.p.ny + wout - 1) / wout) * ((G.p.nx + L - 1) / L);
        if (tiles >= 5120) return L;
    }
    return 8;
}
int depth(const Engine *e) { return e->K > 0 ? e->K : default_depth(e); }
int rows(const Engine *e) { return e->L > 0 ? e->L : default_rows(e); }

int check_params(Engine *e, const wv_params *p)
{
    if (!p) return fail(e, WV_EINVAL, "params is NULL");
    if (p->nx < 3 || p->ny < 3) return fail(e, WV_EINVAL, "nx, ny must be >= 3 (got %d, %d)", p->nx, p->ny);
    if ((long)p->nx * p->ny > (1L << 31)) return fail(e, WV_EINVAL, "grid too large");
    if ((p->bcond < 0 || p->bcond > 4) && p->bcond != WV_BC_LSHAPE)
        return fail(e, WV_EINVAL, "bcond %d is not a B_COND code 0..4 or 10", p->bcond);
    if (p->bcond == WV_BC_LSHAPE && (p->form != WV_FORM_FP64 || p->stencil != 5))
        return fail(e, WV_EUNSUPPORTED, "BC_LSHAPE is provided for the FP64 form and the 5-point stencil");
    if ((p->flags & WV_FLAG_COMPARISON) && (p->form != WV_FORM_FP64 || p->stencil != 5 || p->ny < 4 ||
                                            p->bcond == WV_BC_LSHAPE))
        return fail(e, WV_EUNSUPPORTED, "COMPARISON is provided for the FP64 form, 5-point, NY >= 4");
    if (p->stencil != 5 && p->stencil != 9) return fail(e, WV_EINVAL, "stencil must be 5 or 9");
    if (p->form != WV_FORM_FP64 && p->form != WV_FORM_F32SPLIT) return fail(e, WV_EINVAL, "unknown form %d", p->form);
    if (p->tb_depth < 0 || p->tb_depth > 8 || (p->tb_depth && !wv_have_depth(p->tb_depth)))
        return fail(e, WV_EINVAL, "tb_depth must be 0 or one of 1,2,3,4,6,8");
    if (p->rows_per_tile < 0) return fail(e, WV_EINVAL, "rows_per_tile must be >= 0");
    return WV_OK;
}

int alloc_grid(Engine *e, Grid &G)
{
    const size_t cells = (size_t)G.p.nx * G.ldy;
    const size_t words = (size_t)G.p.nx * G.ldw;
    if (e->form == WV_FORM_FP64) {
        for (int k = 0; k < 4; k++) {
            HIPCHK(e, hipMalloc(&G.u[k], cells * sizeof(double)));
            HIPCHK(e, hipMemset(G.u[k], 0, cells * sizeof(double)));
        }
    } else {
        for (int k = 0; k < 2; k++) {
            HIPCHK(e, hipMalloc(&G.hi[k], cells * sizeof(float)));
            HIPCHK(e, hipMalloc(&G.lo[k], cells * sizeof(float)));
            HIPCHK(e, hipMalloc(&G.v[k], cells * sizeof(float)));
            HIPCHK(e, hipMemset(G.hi[k], 0, cells * sizeof(float)));
            HIPCHK(e, hipMemset(G.lo[k], 0, cells * sizeof(float)));
            HIPCHK(e, hipMemset(G.v[k], 0, cells * sizeof(float)));
        }
    }
    HIPCHK(e, hipMalloc(&G.act, words * 8));
    HIPCHK(e, hipMalloc(&G.eact, words * 8));
    HIPCHK(e, hipMalloc(&G.med, words * 8));
    HIPCHK(e, hipMalloc(&G.nz, words * 8));
    HIPCHK(e, hipMalloc(&G.tc, cells * sizeof(double)));
    HIPCHK(e, hipMemset(G.act, 0, words * 8));
    HIPCHK(e, hipMemset(G.eact, 0, words * 8));
    HIPCHK(e, hipMemset(G.med, 0, words * 8));
    HIPCHK(e, hipMemset(G.tc, 0, cells * sizeof(double)));
    return WV_OK;
}

void free_grid(Grid &G)
{
    void *ptrs[] = {G.u[0], G.u[1], G.u[2], G.u[3], G.hi[0], G.hi[1], G.lo[0], G.lo[1], G.v[0], G.v[1],
                    G.act, G.eact, G.med, G.tccf, G.tgf, G.tc, G.tccf_hi, G.tccf_lo, G.tgf32,
                    G.src_ptr, G.src_col, G.src_id, G.amp, G.nz, G.neu_ptr, G.neu_col, G.neu_code,
                    G.neu_w, G.drv_ptr, G.drv_col, G.drv_id, G.drv_val, G.tbl_ptr, G.tbl_col, G.tbl_n,
                    G.tbl_off, G.fr_lut, G.fr_cs, G.fr_ft, G.fr_te[0], G.fr_te[1], G.fr_ae[0], G.fr_ae[1],
                    G.fr_tf[0], G.fr_tf[1], G.fr_out, G.prb_ptr, G.prb_col, G.prb_id, G.prb_buf};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

// weight the split form folds into its coefficient pair: 1 for the 5-point
// Laplacian, the 9-point weight 1/6 (as the FP64 stencil's W9) otherwise
double split_weight(const Engine *e) { return e->stencil == 9 ? 1.0 / 6.0 : 1.0; }

WvInst make_inst(const Engine *e, const Grid &G)
{
    WvInst d{};
    for (int k = 0; k < 4; k++) d.u[k] = G.u[k];
    for (int k = 0; k < 2; k++) { d.hi[k] = G.hi[k]; d.lo[k] = G.lo[k]; d.v[k] = G.v[k]; }
    d.act = G.act; d.eact = G.eact; d.med = G.med;
    d.tccf = G.tccf; d.tgf = G.tgf; d.tc = G.tc;
    d.tccf_hi = G.tccf_hi; d.tccf_lo = G.tccf_lo; d.tgf32 = G.tgf32;
    d.src_ptr = G.src_ptr; d.src_col = G.src_col; d.src_id = G.src_id;
    d.amp = G.amp; d.amp_len = G.amp_len;
    d.neu_ptr = G.neu_ptr; d.neu_col = G.neu_col; d.neu_code = G.neu_code; d.neu_w = G.neu_w;
    d.drv_ptr = G.drv_ptr; d.drv_col = G.drv_col; d.drv_id = G.drv_id; d.drv_val = G.drv_val;
    d.drv_len = G.drv_len; d.drv_base = G.drv_base;
    d.nz = G.nz;
    d.vmax = G.vmax;
    d.prb_ptr = G.prb_ptr; d.prb_col = G.prb_col; d.prb_id = G.prb_id; d.prb_buf = G.prb_buf;
    d.prb_n = (int)G.h_pi.size(); d.prb_every = G.prb_every; d.prb_cap = G.prb_cap; d.prb_base = G.prb_base;
    d.tbl = G.tbl;
    d.cmp = (G.p.flags & WV_FLAG_COMPARISON) != 0;
    d.tbl_ptr = G.tbl_ptr; d.tbl_col = G.tbl_col; d.tbl_n = G.tbl_n; d.tbl_off = G.tbl_off;
    d.nx = G.p.nx; d.ny = G.p.ny; d.ldy = G.ldy; d.ldw = G.ldw;
    d.bc = G.p.bcond; d.osc_tb = G.p.oscillate_topbot; d.tb_shift = G.p.tb_shift; d.coef = G.coef;
    d.rows_per_tile = rows(e);
    d.tcc0 = G.tcc0; d.tcc1 = G.tcc1; d.tg0 = G.tg0; d.tg1 = G.tg1;
    // split form: the binary32 pair of tcc times the stencil's weight (tcc/6
    // for the 9-point stencil, applied to an unscaled sum of differences)
    const double sw = split_weight(e);
    const double c0 = G.tcc0 * sw, c1 = G.tcc1 * sw;
    d.tcch0 = (float)c0; d.tccl0 = (float)(c0 - (double)d.tcch0);
    d.tcch1 = (float)c1; d.tccl1 = (float)(c1 - (double)d.tcch1);
    d.tgf0 = (float)G.tg0; d.tgf1 = (float)G.tg1;
    d.kappa = G.p.kappa; d.ks = G.p.kappa_sides; d.gs = G.p.gamma_sides;
    d.kt = G.p.kappa_topbot; d.gt = G.p.gamma_topbot;
    d.es2 = G.p.e_scale * G.p.e_scale;
    d.damped = G.p.kappa != 0.0 || G.tg0 != 0.0 || G.tg1 != 0.0 || G.tgf != nullptr;
    // a - KAPPA*x - tgamma*(x-y) with zero coefficients equals a for every
    // value a step can produce from a clean state (leap_undamped, wv_device.h)
    d.fp64_full = d.damped || !G.clean_state || !G.clean_amp || (G.p.flags & WV_FLAG_FULL_FORMULA);
    return d;
}

int sync_inst(Engine *e)
{
    std::vector<WvInst> h(e->n);
    for (int k = 0; k < e->n; k++) h[k] = make_inst(e, e->g[k]);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipMemcpy(e->d_inst, h.data(), sizeof(WvInst) * e->n, hipMemcpyHostToDevice));
    return WV_OK;
}

void drop_graphs(Engine *e)
{
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    e->graphs.clear();
}

void drop_tiles(Engine *e)
{
    for (auto &kv : e->tiles) (void)hipFree(kv.second.d);
    e->tiles.clear();
    drop_graphs(e);
}

// Tile table of a launch of depth k: every instance, row strips (of rows(e)
// rows where a tile can be interior), column blocks of wv_tile_cols(k)
// columns for one-wave tiles; in the split form the inner columns of strips
// that can be interior go to wide tiles (a workgroup of WV_WG waves sharing
// their j-halos through LDS, wv_wide_cols(k) columns).  The table lists
// workgroups: WV_WG tile entries each in the split form (a wide tile
// repeated in every entry, or up to WV_WG one-wave tiles, padded with
// inst = -1), one in the FP64 form.  Workgroups are dealt to the 8 XCDs
// round-robin by workgroup id, so workgroup 8j + x of the table is the j-th
// one XCD x runs.  Each XCD gets one contiguous run of strips in row order:
// the K halo rows a tile reads above and below its strip were read shortly
// before by the neighbouring strip on the same XCD, so they come from that
// XCD's L2 instead of HBM, and so do the halo columns two neighbouring tiles
// of a strip share.  Workgroups of inst = -1 only pad the shorter runs.
int step_cm(const Engine *e);

int get_tiles(Engine *e, int k, Engine::Tiles *out)
{
    auto it = e->tiles.find(k);
    if (it != e->tiles.end()) {
        *out = it->second;
        return WV_OK;
    }
    // (a tile's row count lives in the upper bits of its flags)
    const int wout = wv_tile_cols(k), L = std::min(rows(e), 32767), H = wv_halo(k);
    const int G = e->form == WV_FORM_F32SPLIT ? WV_WG : 1;  // tile entries per workgroup
    const bool wide = e->form == WV_FORM_F32SPLIT && WV_C == 1 && e->wide;
    const int ww = wv_wide_cols(k);
    // Strips of L rows, except where a tile cannot be interior for its
    // position: the first and last rows of a grid and the column blocks that
    // reach its outer columns get strips of at most 8 rows.  Those tiles run
    // the edge formulas and take several times as long per row; short, they
    // finish within the launch instead of setting its length.
    //
    // Short strips cost redundant rows (2K per strip), so they are only as
    // short as keeps one boundary tile within a slot's share of the launch
    // (half a share for strips of the full length L): an estimate from the
    // geometry alone, counting a boundary row as 8 interior rows (measured
    // per tile at 4096^2, K = 8) and 4,096 resident waves.
#ifndef WV_LB_MAX
#define WV_LB_MAX 8  // rows of a short strip (4 and 16 measured slower at 4096^2)
#endif
    //
    // A launch with fewer tiles than resident waves lasts as long as its
    // slowest tile: there the short strips have 4 rows (at 1280x720, K = 4:
    // +30 % over 8), but at least K so that the next strip can be interior.
    int Lb = std::min(L, WV_LB_MAX);
    // the kernel variant of this launch has the edge-column path: tiles of
    // the outer columns take it, and so do the tiles of the first and last
    // rows of grids not periodic along i (grids of one or two media)
    bool ej_on = WV_C == 1 && wv_edgej_kernel(k, e->stencil, step_cm(e));
    for (int q = 0; q < e->n; q++) ej_on = ej_on && e->g[q].coef != WV_COEF_FIELD;
    double tot = 0;  // launch work in interior-row units
    long ntiles = 0;  // one-wave tiles at strips of L rows
    {
        // wide tiles cover an inner column with 64 WV_WG / ww lanes instead of 64 / wout
        const double fw = wide && ww > 0 ? (double)(64 * WV_WG) / ww / (64.0 / wout) : 1.0;
        for (int q = 0; q < e->n; q++) {
            const Grid &G = e->g[q];
            const int ncb = (G.p.ny + wout - 1) / wout, nsb = (G.p.nx + L - 1) / L;
            ntiles += (long)ncb * nsb;
            int ob = 0;
            for (int cb = 0; cb < ncb; cb++) {
                const int js = cb * wout - H;
                ob += !(js >= 1 && js + 64 * WV_C <= G.p.ny - 1);
            }
            const double nb = (double)ob * nsb + (double)(ncb - ob) * std::min(nsb, 2);
            const double ni = (double)(ncb - ob) * std::max(nsb - 2, 0);
            tot += (double)(L + 2 * k) * (ni * fw + 8.0 * nb);
        }
        if (ntiles < 4096) Lb = std::min(L, std::max(k, std::min(4, WV_LB_MAX)));
        // on grids large enough, longer boundary strips (fewer halo rows)
        // still finish within a slot's share of the launch (they start
        // first): at 8192^2 32 rows (+5 % over 8); at 4096^2 even 16 rows
        // would set the launch length
        for (int c : {16, 24, 32, 48, 64})
            if (c <= L && 8.0 * (c + 2 * k) <= tot / 4096.0) Lb = c;
        if (8.0 * (L + 2 * k) <= 0.5 * tot / 4096.0) Lb = L;
    }
    if (const char *v = getenv("WAVEHIP_BOUNDARY_ROWS")) Lb = std::max(k, std::min(L, atoi(v)));
    // Strips of the outer column blocks between their first and last rows.
    // Where the kernel has the edge-column path (K = 8, 5-point, uniform
    // medium: interior evaluation plus the edge formula of the outer columns)
    // such a row costs about 3.5 interior rows, and the strips are as long as
    // keeps one tile within a slot's share of the launch (at 4096^2: 24 rows,
    // +8 % for the split form over 8; 48 rows and more make the edge tiles the
    // critical path).  Otherwise they are boundary strips of Lb rows.
    // WAVEHIP_EDGE_ROWS fixes the length (tuning).
    int Lj = Lb;
    {
        if (ej_on && Lb < L)
            for (int c : {16, 24, 32, 48, 64, 96, 128})
                if (c <= L && 3.5 * (c + 2 * k) <= tot / 4096.0) Lj = c;
    }
    if (const char *v = getenv("WAVEHIP_EDGE_ROWS")) Lj = std::max(1, std::min(L, atoi(v)));
    // flags of a tile reading columns [jlo, jhi) (wrapped where the boundary
    // wraps) for output rows [ibeg, iend): INTERIOR when every row and column
    // the tile reads lies inside the outer box (no edge formula, no
    // wrapping); ALLACT when, in addition, every one of those cells is written
    // by a step
    auto flags_of = [&](int q, int jlo, int jhi, int ibeg, int iend) {
        const Grid &G = e->g[q];
        const int nx = G.p.nx, ny = G.p.ny;
        // (BC_LSHAPE: every bulk cell sums its Laplacian in table order)
        // (COMPARISON: neighbours across the grid, one step per launch)
        const bool interior = jlo >= 1 && jhi <= ny - 1 && ibeg >= k && iend <= nx - k && G.tbl != 2 &&
                              !(G.p.flags & WV_FLAG_COMPARISON);
        // boundary tiles: the point lists whose cells lie in the rows and
        // columns the tile computes, wrapped where the boundary wraps (rows
        // and columns clamped onto the edge are halo cells outside the
        // dependence cone of the tile's output, and rows outside the box are
        // skipped)
        const bool wrapi = G.p.bcond == 1, wrapj = G.p.bcond == 1 || G.p.bcond == 3;
        auto in_range = [](int p, int lo, int hi, int n, bool wrap) {
            if (!wrap) return p >= lo && p < hi;
            if (hi - lo >= n) return true;
            return ((p - lo) % n + n) % n < hi - lo;
        };
        auto any_hit = [&](const std::vector<int> &pi, const std::vector<int> &pj, int rlo, int rhi) {
            for (size_t q = 0; q < pi.size(); q++)
                if (in_range(pi[q], rlo, rhi, nx, wrapi) && in_range(pj[q], jlo, jhi, ny, wrapj)) return true;
            return false;
        };
        const int bflags = (any_hit(G.h_si, G.h_sj, ibeg - k, iend + k) ? WV_TILE_SRC : 0) |
                           (any_hit(G.h_ni, G.h_nj, ibeg - k, iend + k) ? WV_TILE_NEU : 0) |
                           (any_hit(G.h_di, G.h_dj, ibeg - k, iend + k) ? WV_TILE_DRV : 0) |
                           (any_hit(G.h_pi, G.h_pj, ibeg, iend) ? WV_TILE_PRB : 0);
        if (!interior) {
            // EDGEJ: rows inside the box (columns reaching column 0 or NY-1), or
            // rows reaching row 0 or NX-1 of a grid not periodic along i, on a
            // grid with no point lists and no FLOOR (the kernels with the edge-
            // column path evaluate such a tile as interior, its outer columns
            // and edge rows by the edge formulas)
            const bool rows_ok = (ibeg >= k && iend <= nx - k) || G.p.bcond != 1;
            const bool edgej = WV_C == 1 && rows_ok && G.tbl != 2 && !(G.p.flags & WV_FLAG_COMPARISON) &&
                               bflags == 0 && !(G.vmax > 0);
            return edgej ? (int)WV_TILE_EDGEJ : bflags;
        }
        int xflags = 0;
        for (size_t q = 0; q < G.h_ni.size() && !(xflags & WV_TILE_NEU); q++)
            if (G.h_ni[q] >= ibeg - k && G.h_ni[q] < iend + k && G.h_nj[q] >= jlo && G.h_nj[q] < jhi)
                xflags |= WV_TILE_NEU;
        for (size_t q = 0; q < G.h_di.size() && !(xflags & WV_TILE_DRV); q++)
            if (G.h_di[q] >= ibeg - k && G.h_di[q] < iend + k && G.h_dj[q] >= jlo && G.h_dj[q] < jhi)
                xflags |= WV_TILE_DRV;
        for (size_t q = 0; q < G.h_pi.size() && !(xflags & WV_TILE_PRB); q++)
            if (G.h_pi[q] >= ibeg && G.h_pi[q] < iend && G.h_pj[q] >= jlo && G.h_pj[q] < jhi)
                xflags |= WV_TILE_PRB;
        // FLOOR and the two lists run on the boundary path of the kernels
        if (xflags || G.vmax > 0) return xflags | (G.h_si.empty() ? 0 : WV_TILE_SRC);
        bool all = !G.h_wr.empty();
        for (int i = ibeg - k; all && i < iend + k; i++)
            for (int j = jlo; j < jhi; j++)
                if (!((G.h_wr[(size_t)i * G.ldw + (j >> 6)] >> (j & 63)) & 1u)) { all = false; break; }
        // a tile that adds a point source runs on the boundary path, so the
        // interior variants carry no source code at all
        bool src = false;
        for (size_t q = 0; q < G.h_si.size() && !src; q++)
            src = G.h_si[q] >= ibeg - k && G.h_si[q] < iend + k && G.h_sj[q] >= jlo && G.h_sj[q] < jhi;
        if (src) return (int)WV_TILE_SRC;
        return WV_TILE_INTERIOR | (all ? WV_TILE_ALLACT : 0);
    };
    // a one-wave tile with output columns [jb, jb + w) (w <= wout) reads
    // [jb - H, jb - H + 64C)
    auto narrow = [&](int q, int jb, int w, int i0, int i1) {
        return WvTile{q, jb, i0,
                      flags_of(q, jb - H, jb - H + 64 * WV_C, i0, i1) | (w << WV_TILE_WIDTH_SHIFT) |
                          ((i1 - i0) << WV_TILE_ROWS_SHIFT)};
    };
    // every tile in grid order: by grid, first row, columns
    std::vector<WvTile> tl;
    for (int q = 0; q < e->n; q++) {
        const Grid &Gq = e->g[q];
        const int nx = Gq.p.nx, ny = Gq.p.ny;
        const int ncb = (ny + wout - 1) / wout;
        // grids whose every tile is a boundary tile keep strips of L rows
        const bool allb = Gq.tbl == 2 || (Gq.p.flags & WV_FLAG_COMPARISON) || Gq.vmax > 0;
        const int lb = allb ? L : Lb;
        // row cuts of the inner column blocks: lb rows, strips of L, lb rows
        std::vector<int> cut{0};
        const int head = std::min(lb, nx), tail = std::max(head, nx - lb);
        if (head > 0) cut.push_back(head);
        {
            // Strip lengths of the inner rows.  A launch of fewer than three
            // tiles per resident wave (3,072 at three waves per SIMD) ends with
            // the last strips to start, so strips of 64 rows and more come in a
            // repeating mix of 2L, L/2, L/2 rows instead: the same redundant
            // halo rows per row (2K per strip), and the longest-first order
            // runs the long strips first and the short ones last (4096^2,
            // K = 8: split 142 -> 127 us, FP64 160 -> 151 us per launch).  With
            // more tiles the tail is short, and the mixed strips lose the L2
            // reuse of the halo rows two neighbouring strips share (8192^2
            // slower by a fifth).  WAVEHIP_STRIP_MIX sets the mix (a comma list
            // of row counts; one count for strips of one length).
            std::vector<int> mix{L};
            if (L >= 64 && ntiles < 3 * 3072) mix = {2 * L, L / 2, L / 2};
            if (const char *v = getenv("WAVEHIP_STRIP_MIX")) {
                mix.clear();
                for (const char *q = v; *q;) {
                    const int r = atoi(q);
                    if (r > 0) mix.push_back(std::min(r, 32767));
                    while (*q && *q != ',') q++;
                    if (*q == ',') q++;
                }
                if (mix.empty()) mix.push_back(L);
            }
            size_t m = 0;
            for (int i = head + mix[0]; i < tail; i += mix[++m % mix.size()]) cut.push_back(i);
        }
        if (tail > cut.back()) cut.push_back(tail);
        if (nx > cut.back()) cut.push_back(nx);
        // row cuts of the outer column blocks: lb rows at either end, strips
        // of lj rows between
        const int lj = allb ? L : Lj;
        std::vector<int> cutb{0};
        if (head > 0) cutb.push_back(head);
        for (int i = head + lj; i < tail; i += lj) cutb.push_back(i);
        if (tail > cutb.back()) cutb.push_back(tail);
        if (nx > cutb.back()) cutb.push_back(nx);
        auto outer = [&](int cb) {
            const int js = cb * wout - H;
            return !(js >= 1 && js + 64 * WV_C <= ny - 1);
        };
        // the inner column blocks form one run [cb_lo, cb_hi]: output columns [A, B)
        int cb_lo = ncb, cb_hi = -1;
        for (int cb = 0; cb < ncb; cb++)
            if (!outer(cb)) cb_lo = std::min(cb_lo, cb), cb_hi = cb;
        const int A = cb_lo * wout, B = std::min((cb_hi + 1) * wout, ny);
        // merge the two cut lists so tiles come out ordered by first row
        size_t a = 0, b = 0;
        while (a + 1 < cut.size() || b + 1 < cutb.size()) {
            const bool takeb = b + 1 < cutb.size() && (a + 1 >= cut.size() || cutb[b] < cut[a]);
            const int i0 = takeb ? cutb[b] : cut[a], i1 = takeb ? cutb[b + 1] : cut[a + 1];
            if (takeb || !wide || cb_hi < cb_lo || B - A < ww) {
                for (int cb = 0; cb < ncb; cb++)
                    if (outer(cb) == takeb) tl.push_back(narrow(q, cb * wout, wout, i0, i1));
            } else {
                // inner columns of an inner strip: wide tiles from A (each
                // reading [jb - K, jb + ww + K), inside the box since the
                // inner blocks are), the rest [A + m ww, B) in one-wave tiles
                // ending at B (the first may overlap the last wide tile:
                // those columns get the same values from both)
                const int m = (B - A) / ww, rest = B - A - m * ww;
                for (int t = 0; t < m; t++) {
                    const int jb = A + t * ww;
                    const int f = flags_of(q, jb - k, jb + ww + k, i0, i1);
                    if (f & WV_TILE_INTERIOR) {
                        tl.push_back(WvTile{q, jb, i0, f | WV_TILE_WIDE | (ww << WV_TILE_WIDTH_SHIFT) |
                                                           ((i1 - i0) << WV_TILE_ROWS_SHIFT)});
                    } else {
                        for (int j = jb; j < jb + ww; j += wout) tl.push_back(narrow(q, j, std::min(wout, jb + ww - j), i0, i1));
                    }
                }
                const int nr = (rest + wout - 1) / wout;
                for (int r = 0; r < nr; r++) tl.push_back(narrow(q, B - (nr - r) * wout, wout, i0, i1));
            }
            (takeb ? b : a)++;
        }
    }
    // Estimated run time of a tile (of each of its waves) in units of one
    // interior level-row (one level of one row of 64 lanes): K*rows +
    // K*(K-1) level-rows (the skewed pipeline fills and drains over 2K halo
    // rows; half of that at the first and last rows of a grid not periodic
    // along i, whose rows outside the box are skipped), times the cost of a
    // level-row on the tile's path, measured per tile at 4096^2 (K = 8, both
    // forms): interior 1 (1.15 with inactive cells), edge-column tiles 4.5,
    // tiles of the first or last rows 8 (measured 6, corners included, on the
    // edge-column path; counted as 8 so that they start before the interior
    // tiles: started last, they set the launch length), the corner tiles on
    // the boundary path 16.
    auto tile_cost = [&](const WvTile &t) {
        const Grid &Gq = e->g[t.inst];
        const int n = t.flags >> WV_TILE_ROWS_SHIFT, nx = Gq.p.nx;
        const bool wrapi = Gq.p.bcond == 1;
        const bool rowedge = t.i0 < k || t.i0 + n > nx - k;
        const double lr = (double)k * n + (rowedge && !wrapi ? 0.5 : 1.0) * k * (k - 1);
        const int js = t.jb - H;
        const bool coledge = !(js >= 1 && js + 64 * WV_C <= Gq.p.ny - 1);
        double f;
        if (t.flags & WV_TILE_INTERIOR) f = (t.flags & WV_TILE_ALLACT) ? 1.0 : 1.15;
        else if (ej_on && (t.flags & WV_TILE_EDGEJ)) f = rowedge ? 8.0 : 4.5;
        else f = (rowedge && coledge) ? 16.0 : 8.0;
        return lr * f;
    };
    auto slots = [&](const WvTile &t) { return (t.flags & WV_TILE_WIDE) ? WV_WG : 1; };
    // one list of workgroups in XCD-interleave order.  Each XCD takes one
    // contiguous run of the grid order holding an eighth of the estimated
    // work (wave-slots times run time), and starts with its longest tiles:
    // started last, they would run on after the rest had finished.
    // (Balancing rows instead of work gave the two XCDs holding the first and
    // last rows, with their slow boundary tiles, 12 % more work and a launch
    // 20 % longer than the other six needed.)  One-wave tiles of the split
    // form fill workgroups WV_WG at a time in that order.
    const WvTile padt{-1, 0, 0, 0};
    auto to_groups = [&](const std::vector<WvTile> &v, std::vector<WvTile> &dst) {
        std::vector<WvTile> pack;
        for (const WvTile &t : v) {
            if (t.flags & WV_TILE_WIDE) {
                for (int w = 0; w < G; w++) dst.push_back(t);
                continue;
            }
            pack.push_back(t);
            if ((int)pack.size() == G) {
                dst.insert(dst.end(), pack.begin(), pack.end());
                pack.clear();
            }
        }
        if (!pack.empty()) {
            while ((int)pack.size() < G) pack.push_back(padt);
            dst.insert(dst.end(), pack.begin(), pack.end());
        }
    };
    std::vector<WvTile> h;
    if (e->xcd_order) {
        const int X = 8;
        std::vector<std::vector<WvTile>> per(X);
        std::vector<double> cost(tl.size());
        double total = 0;
        for (size_t q = 0; q < tl.size(); q++) total += (cost[q] = tile_cost(tl[q])) * slots(tl[q]);
        double acc = 0;
        std::vector<std::vector<double>> pc(X);
        for (size_t q = 0; q < tl.size(); q++) {
            const double w = cost[q] * slots(tl[q]);
            const int x = std::min(X - 1, (int)((acc + 0.5 * w) * X / std::max(total, 1e-30)));
            per[x].push_back(tl[q]);
            pc[x].push_back(cost[q]);
            acc += w;
        }
        if (e->slow_first) {
            for (int x = 0; x < X; x++) {
                std::vector<size_t> idx(per[x].size());
                for (size_t q = 0; q < idx.size(); q++) idx[q] = q;
                // longest first; tiles of equal estimate keep their grid order
                std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return pc[x][a] > pc[x][b]; });
                std::vector<WvTile> v(idx.size());
                for (size_t q = 0; q < idx.size(); q++) v[q] = per[x][idx[q]];
                per[x].swap(v);
            }
        }
        std::vector<std::vector<WvTile>> grp(X);
        size_t len = 0;
        for (int x = 0; x < X; x++) {
            to_groups(per[x], grp[x]);
            len = std::max(len, grp[x].size() / G);
        }
        for (size_t j = 0; j < len; j++)
            for (int x = 0; x < X; x++)
                for (int w = 0; w < G; w++) h.push_back(j * G < grp[x].size() ? grp[x][j * G + w] : padt);
        while (h.size() >= (size_t)G) {
            bool pad = true;
            for (int w = 0; w < G; w++) pad = pad && h[h.size() - G + w].inst < 0;
            if (!pad) break;
            h.resize(h.size() - G);
        }
    } else {
        to_groups(tl, h);
    }
    Engine::Tiles T;
    T.ni = (int)(h.size() / G);
    T.nb = 0;
    if (!h.empty()) {
        HIPCHK(e, hipMalloc(&T.d, sizeof(WvTile) * h.size()));
        HIPCHK(e, hipMemcpy(T.d, h.data(), sizeof(WvTile) * h.size(), hipMemcpyHostToDevice));
    }
    e->tiles[k] = T;
    *out = T;
    return WV_OK;
}

int engine_create(Engine **out, int n, const wv_params *ps)
{
    Engine *e = new Engine();
    *out = e;
    if (n < 1) return fail(e, WV_EINVAL, "need at least one grid");
    for (int k = 0; k < n; k++) {
        int rc = check_params(e, &ps[k]);
        if (rc) return rc;
        if (ps[k].form != ps[0].form || ps[k].stencil != ps[0].stencil || ps[k].device != ps[0].device)
            return fail(e, WV_EINVAL, "all grids of a batch need the same form, stencil and device");
    }
    e->n = n;
    e->device = ps[0].device;
    e->form = ps[0].form;
    e->stencil = ps[0].stencil;
    e->K = ps[0].tb_depth;
    e->L = ps[0].rows_per_tile;
    if (const char *o = getenv("WAVEHIP_TILE_ORDER")) {
        e->xcd_order = strcmp(o, "linear") != 0;
        e->slow_first = strcmp(o, "inplace") != 0;
    }
    if (const char *o = getenv("WAVEHIP_WIDE")) e->wide = atoi(o) != 0;
    if (const char *o = getenv("WAVEHIP_PERSISTENT")) e->persistent = atoi(o) != 0;
    int ndev = 0;
    HIPCHK(e, hipGetDeviceCount(&ndev));
    if (e->device < 0 || e->device >= ndev) return fail(e, WV_EINVAL, "device %d not present (%d devices)", e->device, ndev);
    HIPCHK(e, hipSetDevice(e->device));
    // the tile order deals each XCD a run of the table, assuming 8 XCDs that
    // take workgroups round-robin (the whole chip, 256 CUs); a partition with
    // fewer CUs gets the grid order
    int cus = 0;
    HIPCHK(e, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
    if (cus < 256) e->xcd_order = false;
    HIPCHK(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->g.resize(n);
    int totrows = 0;
    for (int k = 0; k < n; k++) {
        Grid &G = e->g[k];
        G.p = ps[k];
        G.ldy = round_up(G.p.ny, 64);
        G.ldw = G.ldy / 64;
        int rc = alloc_grid(e, G);
        if (rc) return rc;
        totrows += G.p.nx;
    }
    HIPCHK(e, hipMalloc(&e->d_inst, sizeof(WvInst) * n));
    HIPCHK(e, hipMalloc(&e->d_gstep, sizeof(long)));
    HIPCHK(e, hipMemset(e->d_gstep, 0, sizeof(long)));
    // energy row table
    std::vector<int2> rowsv;
    std::vector<int> row0(n + 1, 0);
    for (int k = 0; k < n; k++) {
        row0[k] = (int)rowsv.size();
        for (int i = 0; i < e->g[k].p.nx; i++) rowsv.push_back(make_int2(k, i));
    }
    row0[n] = (int)rowsv.size();
    e->nrows_total = totrows;
    HIPCHK(e, hipMalloc(&e->d_erows, sizeof(int2) * rowsv.size()));
    HIPCHK(e, hipMemcpy(e->d_erows, rowsv.data(), sizeof(int2) * rowsv.size(), hipMemcpyHostToDevice));
    HIPCHK(e, hipMalloc(&e->d_row0, sizeof(int) * (n + 1)));
    HIPCHK(e, hipMemcpy(e->d_row0, row0.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(e, hipMalloc(&e->d_partial, sizeof(double) * rowsv.size()));
    HIPCHK(e, hipMalloc(&e->d_eout, sizeof(double) * n));
    const WvTile pad[WV_WG] = {{-1, 0, 0, 0}, {-1, 0, 0, 0}, {-1, 0, 0, 0}, {-1, 0, 0, 0}};
    static_assert(WV_WG == 4, "padding workgroup");
    HIPCHK(e, hipMalloc(&e->d_pad, sizeof(pad)));
    HIPCHK(e, hipMemcpy(e->d_pad, pad, sizeof(pad), hipMemcpyHostToDevice));
    return sync_inst(e);
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_tiles(e);
    for (hipEvent_t ev : e->kev) (void)hipEventDestroy(ev);
    if (e->kev_end) (void)hipEventDestroy(e->kev_end);
    for (auto &G : e->g) free_grid(G);
    void *ptrs[] = {e->d_inst, e->d_gstep, e->d_partial, e->d_eout, e->d_erows, e->d_row0,
                    e->tf_rgb, e->tf_strips, e->tf_body, e->tf_lens, e->tf_off, e->d_pad, e->d_pctr};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

inline bool bit(const std::vector<uint64_t> &pl, int ldw, int i, int j)
{
    return (pl[(size_t)i * ldw + (j >> 6)] >> (j & 63)) & 1u;
}

// the cells a step writes: the active plane plus Neumann points and driven cells
void rebuild_written(Grid &G)
{
    G.h_wr = G.h_act;
    if (G.h_wr.empty()) return;
    for (size_t q = 0; q < G.h_ni.size(); q++)
        G.h_wr[(size_t)G.h_ni[q] * G.ldw + (G.h_nj[q] >> 6)] |= 1ull << (G.h_nj[q] & 63);
    for (size_t q = 0; q < G.h_di.size(); q++)
        G.h_wr[(size_t)G.h_di[q] * G.ldw + (G.h_dj[q] >> 6)] |= 1ull << (G.h_dj[q] & 63);
}

inline bool bits_zero(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    return b == 0;
}

// do the masked (never written) cells of these arrays all hold +0?
// no -0, infinity or NaN among n values (NULL counts as clean)
bool clean_values(const double *a, size_t n)
{
    if (!a) return true;
    for (size_t c = 0; c < n; c++) {
        const double v = a[c];
        if (!std::isfinite(v) || (v == 0.0 && std::signbit(v))) return false;
    }
    return true;
}

bool masked_zero(const Grid &G, const double *a)
{
    if (!a) return true;
    const int nx = G.p.nx, ny = G.p.ny;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            if (!bit(G.h_wr, G.ldw, i, j) && !bits_zero(a[(size_t)i * ny + j])) return false;
    return true;
}

// Active plane of the neighbour-table stepper (wave_3d.c:390-572): bulk and
// edge cells with TWOSPEEDS or xy_in != 0 (the plain stepper's bulk rule is
// xy_in > 0); BC_LSHAPE has no edge formula, so of its edge cells only those
// of the OSCILLATE_TOPBOT branch are written.
std::vector<uint64_t> table_act(const Grid &G)
{
    const int nx = G.p.nx, ny = G.p.ny, ldw = G.ldw;
    std::vector<uint64_t> a((size_t)nx * ldw, 0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            bool on = G.p.twospeeds || bit(G.h_nz, ldw, i, j);
            const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
            if (on && edge && G.tbl == 2)
                on = (j == 0 || j == ny - 1) && G.p.oscillate_topbot && i < G.p.tb_shift && i < nx - 1 && i > 0;
            if (on) a[(size_t)i * ldw + (j >> 6)] |= 1ull << (j & 63);
        }
    return a;
}

int set_tables(Engine *e, int k, const short *xy, const double *tcc, const double *tc, const double *tg)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!xy || !tcc || !tc || !tg) return fail(e, WV_EINVAL, "table pointer is NULL");
    Grid &G = e->g[k];
    const int nx = G.p.nx, ny = G.p.ny, ldw = G.ldw, ldy = G.ldy;
    const bool ts = G.p.twospeeds != 0, cmp = (G.p.flags & WV_FLAG_COMPARISON) != 0;
    std::vector<uint64_t> act((size_t)nx * ldw, 0), eact((size_t)nx * ldw, 0), med((size_t)nx * ldw, 0);
    // distinct (tcc, tgamma) pairs over the cells a step or the energy reads
    std::vector<std::pair<uint64_t, uint64_t>> cls;
    bool field = false;
    auto key = [](double a) { uint64_t b; memcpy(&b, &a, 8); return b; };
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const size_t c = (size_t)i * ny + j;
            // (COMPARISON: the interface columns take the edge rule, wave_3d.c:875, 933)
            const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (cmp && (j == ny / 2 - 1 || j == ny / 2));
            const bool a = ts || (edge ? xy[c] != 0 : xy[c] > 0);
            const bool en = ts || xy[c] != 0;
            const size_t w = (size_t)i * ldw + (j >> 6);
            const uint64_t b = 1ull << (j & 63);
            if (a) act[w] |= b;
            if (en) eact[w] |= b;
            if (!(a || en) || field) continue;
            const auto pr = std::make_pair(key(tcc[c]), key(tg[c]));
            int idx = -1;
            for (size_t q = 0; q < cls.size(); q++)
                if (cls[q] == pr) idx = (int)q;
            if (idx < 0) {
                if (cls.size() == 2) { field = true; continue; }
                cls.push_back(pr);
                idx = (int)cls.size() - 1;
            }
            if (idx == 1) med[w] |= b;
        }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const size_t wb = (size_t)nx * ldw * 8;
    HIPCHK(e, hipMemcpy(G.act, act.data(), wb, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(G.eact, eact.data(), wb, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(G.med, med.data(), wb, hipMemcpyHostToDevice));
    std::vector<uint64_t> nzp((size_t)nx * ldw, 0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            if (xy[(size_t)i * ny + j] != 0) nzp[(size_t)i * ldw + (j >> 6)] |= 1ull << (j & 63);
    HIPCHK(e, hipMemcpy(G.nz, nzp.data(), wb, hipMemcpyHostToDevice));
    G.h_nz.swap(nzp);
    HIPCHK(e, hipMemcpy2D(G.tc, ldy * sizeof(double), tc, ny * sizeof(double), ny * sizeof(double), nx,
                          hipMemcpyHostToDevice));
    for (double **q : {&G.tccf, &G.tgf})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    for (float **q : {&G.tccf_hi, &G.tccf_lo, &G.tgf32})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    auto dbl = [](uint64_t b) { double a; memcpy(&a, &b, 8); return a; };
    if (!field) {
        G.coef = cls.size() == 2 ? WV_COEF_TWO : WV_COEF_UNIFORM;
        G.tcc0 = cls.empty() ? 0.0 : dbl(cls[0].first);
        G.tg0 = cls.empty() ? 0.0 : dbl(cls[0].second);
        G.tcc1 = cls.size() == 2 ? dbl(cls[1].first) : G.tcc0;
        G.tg1 = cls.size() == 2 ? dbl(cls[1].second) : G.tg0;
    } else {
        G.coef = WV_COEF_FIELD;
        const size_t cells = (size_t)nx * ldy;
        // tgamma: one value over the read cells, or a field
        bool tg_uniform = true;
        double tgu = 0;
        bool first = true;
        for (int i = 0; i < nx && tg_uniform; i++)
            for (int j = 0; j < ny; j++) {
                const size_t c = (size_t)i * ny + j;
                if (!(bit(act, ldw, i, j) || bit(eact, ldw, i, j))) continue;
                if (first) { tgu = tg[c]; first = false; }
                else if (key(tg[c]) != key(tgu)) { tg_uniform = false; break; }
            }
        G.tg0 = G.tg1 = tgu;
        HIPCHK(e, hipMalloc(&G.tccf, cells * sizeof(double)));
        HIPCHK(e, hipMemcpy2D(G.tccf, ldy * sizeof(double), tcc, ny * sizeof(double), ny * sizeof(double),
                              nx, hipMemcpyHostToDevice));
        if (!tg_uniform) {
            HIPCHK(e, hipMalloc(&G.tgf, cells * sizeof(double)));
            HIPCHK(e, hipMemcpy2D(G.tgf, ldy * sizeof(double), tg, ny * sizeof(double), ny * sizeof(double),
                                  nx, hipMemcpyHostToDevice));
        }
        if (e->form == WV_FORM_F32SPLIT) {
            std::vector<float> h((size_t)nx * ldy, 0.f), l((size_t)nx * ldy, 0.f);
            bool anylo = false;
            const double sw = split_weight(e);
            for (int i = 0; i < nx; i++)
                for (int j = 0; j < ny; j++) {
                    const double t = tcc[(size_t)i * ny + j] * sw;
                    const float th = (float)t;
                    const float tl = (float)(t - (double)th);
                    h[(size_t)i * ldy + j] = th;
                    l[(size_t)i * ldy + j] = tl;
                    if (tl != 0.0f) anylo = true;
                }
            HIPCHK(e, hipMalloc(&G.tccf_hi, cells * sizeof(float)));
            HIPCHK(e, hipMemcpy(G.tccf_hi, h.data(), cells * sizeof(float), hipMemcpyHostToDevice));
            if (anylo) {
                HIPCHK(e, hipMalloc(&G.tccf_lo, cells * sizeof(float)));
                HIPCHK(e, hipMemcpy(G.tccf_lo, l.data(), cells * sizeof(float), hipMemcpyHostToDevice));
            }
            if (!tg_uniform) {
                for (int i = 0; i < nx; i++)
                    for (int j = 0; j < ny; j++) h[(size_t)i * ldy + j] = (float)tg[(size_t)i * ny + j];
                HIPCHK(e, hipMalloc(&G.tgf32, cells * sizeof(float)));
                HIPCHK(e, hipMemcpy(G.tgf32, h.data(), cells * sizeof(float), hipMemcpyHostToDevice));
            }
        }
    }
    // a state set under another mask may hold non-zero values on cells that
    // are now never written: keep it, but step it one step per launch
    G.h_act_plain = act;
    if (G.tbl) act = table_act(G);
    if (G.state && G.h_act != act) G.exact = true;
    HIPCHK(e, hipMemcpy(G.act, act.data(), wb, hipMemcpyHostToDevice));
    G.h_act.swap(act);
    rebuild_written(G);
    G.tables = true;
    drop_tiles(e);  // tile flags depend on the mask
    return sync_inst(e);
}

int upload2d(Engine *e, void *dst, int ldy, const void *src, int nx, int ny, size_t es)
{
    HIPCHK(e, hipMemcpy2DAsync(dst, ldy * es, src, ny * es, ny * es, nx, hipMemcpyHostToDevice, e->stream));
    return WV_OK;
}

int download2d(Engine *e, void *dst, const void *src, int ldy, int nx, int ny, size_t es)
{
    HIPCHK(e, hipMemcpy2DAsync(dst, ny * es, src, ldy * es, ny * es, nx, hipMemcpyDeviceToHost, e->stream));
    return WV_OK;
}

int set_state(Engine *e, int k, const double *phi, const double *psi, const double *tmp)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!phi || !psi) return fail(e, WV_EINVAL, "state pointer is NULL");
    Grid &G = e->g[k];
    if (!G.tables) return fail(e, WV_ESTATE, "set the tables before the state");
    const int nx = G.p.nx, ny = G.p.ny;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    int rc;
    if (e->form == WV_FORM_FP64) {
        const size_t bytes = (size_t)nx * G.ldy * sizeof(double);
        if ((rc = upload2d(e, G.u[e->role[0]], G.ldy, phi, nx, ny, 8))) return rc;
        if ((rc = upload2d(e, G.u[e->role[1]], G.ldy, psi, nx, ny, 8))) return rc;
        if (tmp) {
            if ((rc = upload2d(e, G.u[e->role[2]], G.ldy, tmp, nx, ny, 8))) return rc;
        } else {
            HIPCHK(e, hipMemsetAsync(G.u[e->role[2]], 0, bytes, e->stream));
        }
        HIPCHK(e, hipMemsetAsync(G.u[e->role[3]], 0, bytes, e->stream));
        G.exact = !(masked_zero(G, phi) && masked_zero(G, psi) && masked_zero(G, tmp));

        const size_t n = (size_t)nx * ny;
        const bool clean = clean_values(phi, n) && clean_values(psi, n) && clean_values(tmp, n);
        if (clean != G.clean_state) {
            G.clean_state = clean;
            if ((rc = sync_inst(e))) return rc;
        }
    } else {
        const size_t n = (size_t)nx * ny;
        std::vector<float> h(n), l(n), v(n);
        for (size_t c = 0; c < n; c++) {
            const float hh = (float)phi[c];
            h[c] = hh;
            l[c] = (float)(phi[c] - (double)hh);
            v[c] = (float)(phi[c] - psi[c]);
        }
        for (int s = 0; s < 2; s++) {
            if ((rc = upload2d(e, G.hi[s], G.ldy, h.data(), nx, ny, 4))) return rc;
            if ((rc = upload2d(e, G.lo[s], G.ldy, l.data(), nx, ny, 4))) return rc;
            if ((rc = upload2d(e, G.v[s], G.ldy, v.data(), nx, ny, 4))) return rc;
        }
        HIPCHK(e, hipStreamSynchronize(e->stream));
        G.exact = false;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    G.state = true;
    return WV_OK;
}

int get_state(Engine *e, int k, double *phi, double *psi, double *tmp)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (!G.state) return fail(e, WV_ESTATE, "no state set");
    const int nx = G.p.nx, ny = G.p.ny;
    HIPCHK(e, hipSetDevice(e->device));
    int rc;
    if (e->form == WV_FORM_FP64) {
        if (phi && (rc = download2d(e, phi, G.u[e->role[0]], G.ldy, nx, ny, 8))) return rc;
        if (psi && (rc = download2d(e, psi, G.u[e->role[1]], G.ldy, nx, ny, 8))) return rc;
        if (tmp && (rc = download2d(e, tmp, G.u[e->role[2]], G.ldy, nx, ny, 8))) return rc;
        HIPCHK(e, hipStreamSynchronize(e->stream));
    } else {
        if (tmp) return fail(e, WV_EUNSUPPORTED, "the split form keeps one time level and a velocity; no third buffer");
        const size_t n = (size_t)nx * ny;
        std::vector<float> h(n), l(n), v(n);
        const int s = e->scur;
        if ((rc = download2d(e, h.data(), G.hi[s], G.ldy, nx, ny, 4))) return rc;
        if ((rc = download2d(e, l.data(), G.lo[s], G.ldy, nx, ny, 4))) return rc;
        if ((rc = download2d(e, v.data(), G.v[s], G.ldy, nx, ny, 4))) return rc;
        HIPCHK(e, hipStreamSynchronize(e->stream));
        for (size_t c = 0; c < n; c++) {
            const double u = (double)h[c] + (double)l[c];
            if (phi) phi[c] = u;
            if (psi) psi[c] = u - (double)v[c];
        }
    }
    return WV_OK;
}

int set_sources(Engine *e, int k, int n, const int *si, const int *sj, const double *amp, long amp_len)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (!G.tables) return fail(e, WV_ESTATE, "set the tables before the sources");
    // every argument is checked before anything changes: a rejected call
    // leaves the sources (and the step counter) as they were
    if (n < 0) return fail(e, WV_EINVAL, "negative source count");
    if (n > 0) {
        if (!si || !sj || !amp || amp_len <= 0) return fail(e, WV_EINVAL, "source arrays missing");
        for (int q = 0; q < n; q++) {
            if (si[q] < 0 || si[q] >= G.p.nx || sj[q] < 0 || sj[q] >= G.p.ny)
                return fail(e, WV_EINVAL, "source %d at (%d,%d) is outside the grid", q, si[q], sj[q]);
            if (!bit(G.h_wr, G.ldw, si[q], sj[q]))
                return fail(e, WV_EINVAL, "source %d at (%d,%d) is on a cell the step never writes", q, si[q], sj[q]);
        }
    }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (int **q : {&G.src_ptr, &G.src_col, &G.src_id})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (G.amp) { (void)hipFree(G.amp); G.amp = nullptr; }
    G.amp_len = 0;
    G.h_si.clear();
    G.h_sj.clear();
    G.clean_amp = true;
    drop_tiles(e);  // tile flags mark the tiles that hold a source
    if (n > 0) {
        std::vector<int> order(n);
        for (int q = 0; q < n; q++) order[q] = q;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return si[a] < si[b]; });
        std::vector<int> ptr(G.p.nx + 1, 0), col(n), id(n);
        for (int q = 0; q < n; q++) ptr[si[q] + 1]++;
        for (int i = 0; i < G.p.nx; i++) ptr[i + 1] += ptr[i];
        for (int q = 0; q < n; q++) { col[q] = sj[order[q]]; id[q] = order[q]; }
        hipError_t err = hipMalloc(&G.src_ptr, sizeof(int) * ptr.size());
        if (err == hipSuccess) err = hipMalloc(&G.src_col, sizeof(int) * n);
        if (err == hipSuccess) err = hipMalloc(&G.src_id, sizeof(int) * n);
        if (err == hipSuccess) err = hipMalloc(&G.amp, sizeof(double) * (size_t)n * amp_len);
        if (err == hipSuccess)
            err = hipMemcpy(G.src_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice);
        if (err == hipSuccess) err = hipMemcpy(G.src_col, col.data(), sizeof(int) * n, hipMemcpyHostToDevice);
        if (err == hipSuccess) err = hipMemcpy(G.src_id, id.data(), sizeof(int) * n, hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(G.amp, amp, sizeof(double) * (size_t)n * amp_len, hipMemcpyHostToDevice);
        if (err != hipSuccess) {
            // the grid is left with no sources, and the device table says so
            for (int **q : {&G.src_ptr, &G.src_col, &G.src_id})
                if (*q) { (void)hipFree(*q); *q = nullptr; }
            if (G.amp) { (void)hipFree(G.amp); G.amp = nullptr; }
            (void)sync_inst(e);
            return fail(e, WV_EHIP, "source upload failed: %s", hipGetErrorString(err));
        }
        G.amp_len = amp_len;
        G.h_si.assign(si, si + n);
        G.h_sj.assign(sj, sj + n);
        for (size_t q = 0; q < (size_t)n * amp_len && G.clean_amp; q++) G.clean_amp = std::isfinite(amp[q]);
    }
    // the source counter restarts at 0; driven cells keep counting from their own start
    long gnow = 0;
    HIPCHK(e, hipMemcpy(&gnow, e->d_gstep, sizeof(long), hipMemcpyDeviceToHost));
    for (auto &H : e->g) H.drv_base -= gnow;
    HIPCHK(e, hipMemset(e->d_gstep, 0, sizeof(long)));
    drop_graphs(e);
    return sync_inst(e);
}

// rows of a CSR list by row: ptr[nx+1], and the order of the entries (by
// row, then column, entries of one cell in list order: the kernels search a
// row's columns and take the last entry of a cell)
void csr_by_row(int nx, int n, const int *row, const int *col, std::vector<int> &ptr, std::vector<int> &order)
{
    order.resize(n);
    for (int q = 0; q < n; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return row[a] != row[b] ? row[a] < row[b] : col[a] < col[b]; });
    ptr.assign(nx + 1, 0);
    for (int q = 0; q < n; q++) ptr[row[q] + 1]++;
    for (int i = 0; i < nx; i++) ptr[i + 1] += ptr[i];
}

template <typename T>
int upload_vec(Engine *e, T **dst, const std::vector<T> &v)
{
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    if (v.empty()) return WV_OK;
    HIPCHK(e, hipMalloc(dst, sizeof(T) * v.size()));
    HIPCHK(e, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return WV_OK;
}

int set_neumann(Engine *e, int k, int n, const int *pi, const int *pj, const int *bi1, const int *bj1,
                const int *bi2, const int *bj2, const double *wx, const double *wy)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (G.tbl && n > 0)
        return fail(e, WV_EUNSUPPORTED, "the neighbour-table stepper has no Neumann points (wave_3d.c:345-597)");
    if (e->form != WV_FORM_FP64 || WV_C != 1)
        return fail(e, WV_EUNSUPPORTED, "Neumann ghost points are provided for the FP64 form");
    if (n < 0) return fail(e, WV_EINVAL, "negative point count");
    if (n > 0 && (!pi || !pj || !bi1 || !bj1 || !bi2 || !bj2 || !wx || !wy))
        return fail(e, WV_EINVAL, "Neumann arrays missing");
    const int nx = G.p.nx, ny = G.p.ny;
    std::vector<uint64_t> seen((size_t)nx * G.ldw, 0);
    for (int q = 0; q < n; q++) {
        const int i = pi[q], j = pj[q];
        if (i < 0 || i >= nx || j < 0 || j >= ny) return fail(e, WV_EINVAL, "Neumann point %d at (%d,%d) is outside the grid", q, i, j);
        const int b[4] = {bi1[q], bj1[q], bi2[q], bj2[q]};
        for (int m = 0; m < 4; m += 2)
            if (b[m] < 0 || b[m] >= nx || b[m + 1] < 0 || b[m + 1] >= ny || std::abs(b[m] - i) > 1 || std::abs(b[m + 1] - j) > 1)
                return fail(e, WV_EINVAL,
                            "Neumann point %d at (%d,%d): neighbour (%d,%d) must lie in the grid within one cell", q, i,
                            j, b[m], b[m + 1]);
        uint64_t &w = seen[(size_t)i * G.ldw + (j >> 6)];
        if ((w >> (j & 63)) & 1u) return fail(e, WV_EINVAL, "two Neumann points at (%d,%d)", i, j);
        w |= 1ull << (j & 63);
    }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::vector<int> ptr, order;
    csr_by_row(nx, n, pi, pj, ptr, order);
    std::vector<int> col(n), code(n);
    std::vector<double> wv(2 * (size_t)n);
    for (int q = 0; q < n; q++) {
        const int o = order[q], i = pi[o], j = pj[o];
        col[q] = j;
        // (row offset + 1) | (column offset + 1) << 2, first neighbour low
        code[q] = ((bi1[o] - i + 1) | ((bj1[o] - j + 1) << 2)) | (((bi2[o] - i + 1) | ((bj2[o] - j + 1) << 2)) << 4);
        wv[2 * (size_t)q] = wx[o];
        wv[2 * (size_t)q + 1] = wy[o];
    }
    int rc;
    if (n == 0) ptr.clear();
    if ((rc = upload_vec(e, &G.neu_ptr, ptr)) || (rc = upload_vec(e, &G.neu_col, col)) ||
        (rc = upload_vec(e, &G.neu_code, code)) || (rc = upload_vec(e, &G.neu_w, wv)))
        return rc;
    G.h_ni.assign(pi, pi + n);
    G.h_nj.assign(pj, pj + n);
    rebuild_written(G);
    if (G.state) G.exact = true;  // the set of written cells changed under the state
    drop_tiles(e);
    return sync_inst(e);
}

int set_drive(Engine *e, int k, int n, const int *pi, const int *pj, const double *val, long len)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (e->form != WV_FORM_FP64) return fail(e, WV_EUNSUPPORTED, "driven cells are provided for the FP64 form");
    if (n < 0 || len < 0) return fail(e, WV_EINVAL, "negative count");
    if (n > 0 && (!pi || !pj || !val || len < 1)) return fail(e, WV_EINVAL, "drive arrays missing");
    const int nx = G.p.nx, ny = G.p.ny;
    for (int q = 0; q < n; q++)
        if (pi[q] < 0 || pi[q] >= nx || pj[q] < 0 || pj[q] >= ny)
            return fail(e, WV_EINVAL, "driven cell %d at (%d,%d) is outside the grid", q, pi[q], pj[q]);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    // the drive counts steps from this call (device counter minus its value now)
    long gnow = 0;
    HIPCHK(e, hipMemcpy(&gnow, e->d_gstep, sizeof(long), hipMemcpyDeviceToHost));
    const bool same = G.h_di.size() == (size_t)n && std::equal(pi, pi + n, G.h_di.begin()) &&
                      std::equal(pj, pj + n, G.h_dj.begin());
    int rc;
    if (!same) {
        std::vector<int> ptr, order;
        csr_by_row(nx, n, pi, pj, ptr, order);
        std::vector<int> col(n), id(n);
        for (int q = 0; q < n; q++) { col[q] = pj[order[q]]; id[q] = order[q]; }
        if (n == 0) ptr.clear();
        if ((rc = upload_vec(e, &G.drv_ptr, ptr)) || (rc = upload_vec(e, &G.drv_col, col)) ||
            (rc = upload_vec(e, &G.drv_id, id)))
            return rc;
        G.h_di.assign(pi, pi + n);
        G.h_dj.assign(pj, pj + n);
        rebuild_written(G);
        if (G.state) G.exact = true;
        drop_tiles(e);
    }
    std::vector<double> v(val, val + (size_t)n * len);
    if ((rc = upload_vec(e, &G.drv_val, v))) return rc;
    G.drv_len = n ? len : 0;
    G.drv_base = gnow;
    drop_graphs(e);
    return sync_inst(e);
}

int set_floor(Engine *e, int k, double vmax)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (e->form != WV_FORM_FP64) return fail(e, WV_EUNSUPPORTED, "FLOOR is provided for the FP64 form");
    if (!(vmax >= 0)) return fail(e, WV_EINVAL, "vmax must be >= 0 (0 turns FLOOR off)");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const double old = G.vmax;
    G.vmax = vmax;
    if ((old > 0) != (vmax > 0)) drop_tiles(e);
    return sync_inst(e);
}

// Does the neighbour-table stepper use the Laplacian of edge cell (i, j)?
// wave_3d.c:404-572: the branches that take the bulk formula.
bool table_edge_uses_delta(const wv_params &p, int i, int j)
{
    const int nx = p.nx, ny = p.ny, bc = p.bcond;
    if (j == 0 || j == ny - 1) {
        if (p.oscillate_topbot && i < p.tb_shift && i < nx - 1 && i > 0) return true;
        return bc == WV_BC_DIRICHLET || bc == WV_BC_PERIODIC || (bc == WV_BC_VPER_HABS && i != 0);
    }
    return bc == WV_BC_DIRICHLET || bc == WV_BC_PERIODIC;
}

int set_laplacian(Engine *e, int k, const int *nn, const int *nb)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (e->form != WV_FORM_FP64 || e->stencil != 5 || WV_C != 1)
        return fail(e, WV_EUNSUPPORTED, "the neighbour-table stepper is provided for the FP64 form, 5-point");
    if (nn && !nb) return fail(e, WV_EINVAL, "neighbour index table is NULL");
    if (nn && !G.h_ni.empty())
        return fail(e, WV_EUNSUPPORTED, "the neighbour-table stepper has no Neumann points (wave_3d.c:345-597)");
    if (nn && (G.p.flags & WV_FLAG_COMPARISON))
        return fail(e, WV_EUNSUPPORTED, "the neighbour-table stepper has no COMPARISON (wave_3d.c:345-597)");
    const int nx = G.p.nx, ny = G.p.ny;
    const long ncell = (long)nx * ny;
    const bool lshape = G.p.bcond == WV_BC_LSHAPE;
    if (!nn && lshape) return fail(e, WV_EINVAL, "BC_LSHAPE steps with its neighbour table");
    // the cells whose Laplacian is not the plain 5-point one, in row-major
    // order (so already sorted by row, then column)
    std::vector<int> ptr, col, cnt;
    std::vector<long long> off;
    if (nn) {
        ptr.assign(nx + 1, 0);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                const long c = (long)i * ny + j;
                const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
                int n;
                if (edge) {
                    if (!table_edge_uses_delta(G.p, i, j)) continue;
                    n = nn[c];
                    if (n < 0 || n > 4)
                        return fail(e, WV_EINVAL, "cell (%d,%d): %d neighbours (0..4)", i, j, n);
                } else {
                    if (!lshape) continue;  // compute_laplacian's direct bulk formula
                    const long reg[4] = {c + ny, c - ny, c + 1, c - 1};
                    if (nb[4 * c] == reg[0] && nb[4 * c + 1] == reg[1] && nb[4 * c + 2] == reg[2] &&
                        nb[4 * c + 3] == reg[3])
                        continue;  // E, W, N, S: summed in the kernel from registers
                    n = 4;
                }
                for (int m = 0; m < 4; m++) {
                    long long o = 0;
                    if (m < n) {
                        const long q = nb[4 * c + m];
                        if (q < 0 || q >= ncell)
                            return fail(e, WV_EINVAL, "cell (%d,%d): neighbour %d index %ld outside the grid", i, j,
                                        m, q);
                        o = (long long)(q / ny) * G.ldy + q % ny;
                    }
                    off.push_back(o);
                }
                col.push_back(j);
                cnt.push_back(n);
                ptr[i + 1]++;
            }
        for (int i = 0; i < nx; i++) ptr[i + 1] += ptr[i];
    }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    int rc;
    if (col.empty()) ptr.clear();
    if ((rc = upload_vec(e, &G.tbl_ptr, ptr)) || (rc = upload_vec(e, &G.tbl_col, col)) ||
        (rc = upload_vec(e, &G.tbl_n, cnt)) || (rc = upload_vec(e, &G.tbl_off, off)))
        return rc;
    // no entries at all (e.g. absorbing edges): still the table stepper's rules
    if (nn && !G.tbl_ptr) {
        const std::vector<int> zero(nx + 1, 0);
        if ((rc = upload_vec(e, &G.tbl_ptr, zero))) return rc;
    }
    G.tbl = !nn ? 0 : (lshape ? 2 : 1);
    if (G.tables) {
        std::vector<uint64_t> act = G.tbl ? table_act(G) : G.h_act_plain;
        HIPCHK(e, hipMemcpy(G.act, act.data(), act.size() * 8, hipMemcpyHostToDevice));
        if (G.state && act != G.h_act) G.exact = true;
        G.h_act.swap(act);
        rebuild_written(G);
    }
    drop_tiles(e);
    return sync_inst(e);
}

// ------------------------------------------------------------------ probes

int set_probes(Engine *e, int k, int n, const int *pi, const int *pj, long every, long cap)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (e->form != WV_FORM_FP64 || WV_C != 1) return fail(e, WV_EUNSUPPORTED, "probes are provided for the FP64 form");
    if (n < 0) return fail(e, WV_EINVAL, "negative probe count");
    if (n > 0 && (!pi || !pj || every < 1 || cap < 1)) return fail(e, WV_EINVAL, "probe arrays, period or capacity");
    if (n > 0 && (double)n * (double)cap > 1e9) return fail(e, WV_EINVAL, "probe buffer above 8 GB");
    const int nx = G.p.nx, ny = G.p.ny;
    for (int q = 0; q < n; q++)
        if (pi[q] < 0 || pi[q] >= nx || pj[q] < 0 || pj[q] >= ny)
            return fail(e, WV_EINVAL, "probe %d at (%d,%d) is outside the grid", q, pi[q], pj[q]);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::vector<int> ptr, order, col(n), id(n);
    csr_by_row(nx, n, pi, pj, ptr, order);
    for (int q = 0; q < n; q++) { col[q] = pj[order[q]]; id[q] = order[q]; }
    if (n == 0) ptr.clear();
    int rc;
    if ((rc = upload_vec(e, &G.prb_ptr, ptr)) || (rc = upload_vec(e, &G.prb_col, col)) ||
        (rc = upload_vec(e, &G.prb_id, id)))
        return rc;
    if (G.prb_buf) { (void)hipFree(G.prb_buf); G.prb_buf = nullptr; }
    if (n > 0) {
        HIPCHK(e, hipMalloc(&G.prb_buf, sizeof(double) * (size_t)n * cap));
        HIPCHK(e, hipMemset(G.prb_buf, 0, sizeof(double) * (size_t)n * cap));
    }
    long gnow = 0;
    HIPCHK(e, hipMemcpy(&gnow, e->d_gstep, sizeof(long), hipMemcpyDeviceToHost));
    G.h_pi.assign(pi, pi + n);
    G.h_pj.assign(pj, pj + n);
    G.prb_every = n ? every : 0;
    G.prb_cap = n ? cap : 0;
    G.prb_base = gnow;
    G.prb_host0 = e->steps_done;
    drop_tiles(e);
    return sync_inst(e);
}

// copy the samples recorded since the probes were set or last read
// (out[s*n + k], s < *count) and restart the numbering after them; *lost:
// samples beyond the capacity, not recorded
int get_probes(Engine *e, int k, double *out, long max_samples, long *count, long *lost)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    const long n = (long)G.h_pi.size();
    if (!n) return fail(e, WV_ESTATE, "no probes set");
    if (!out || !count || max_samples < 0) return fail(e, WV_EINVAL, "output is NULL");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const long taken = (e->steps_done - G.prb_host0) / G.prb_every;
    const long have = std::min(taken, G.prb_cap);
    if (have > max_samples) return fail(e, WV_EINVAL, "%ld samples recorded, room for %ld", have, max_samples);
    if (have) HIPCHK(e, hipMemcpy(out, G.prb_buf, sizeof(double) * n * have, hipMemcpyDeviceToHost));
    *count = have;
    if (lost) *lost = taken - have;
    G.prb_base += taken * G.prb_every;
    G.prb_host0 += taken * G.prb_every;
    return sync_inst(e);
}

bool any_exact(const Engine *e)
{
    for (const auto &G : e->g)
        if (G.exact) return true;
    return false;
}

int ready(Engine *e)
{
    for (int k = 0; k < e->n; k++) {
        if (!e->g[k].tables) return fail(e, WV_ESTATE, "grid %d has no tables", k);
        if (!e->g[k].state) return fail(e, WV_ESTATE, "grid %d has no state", k);
        if (e->g[k].p.bcond == WV_BC_LSHAPE && !e->g[k].tbl)
            return fail(e, WV_ESTATE, "grid %d: BC_LSHAPE needs its neighbour table (wv_set_laplacian)", k);
    }
    return WV_OK;
}

// grids whose steps read neighbours anywhere in the field: the
// neighbour-table stepper and COMPARISON
bool any_table(const Engine *e)
{
    for (const auto &G : e->g)
        if (G.tbl || (G.p.flags & WV_FLAG_COMPARISON)) return true;
    return false;
}

// the neighbour-table Laplacian reads its neighbours from the step's input
// field, so it steps one step per launch (with the exact 3-buffer rotation)
int max_depth(const Engine *e, bool force_k1)
{
    return (force_k1 || (e->form == WV_FORM_FP64 && (any_exact(e) || any_table(e)))) ? 1 : depth(e);
}

// depth of the next launch when `left` steps remain
int next_depth(int kmax, long left)
{
    int k = (int)std::min<long>(kmax, left);
    while (!wv_have_depth(k)) k--;
    return k;
}

// host-side rotation after one launch of depth k
void advance_roles(Engine *e, int k)
{
    if (e->form == WV_FORM_FP64) {
        const int r0 = e->role[0], r1 = e->role[1], r2 = e->role[2], r3 = e->role[3];
        if (k == 1) { e->role[0] = r2; e->role[1] = r0; e->role[2] = r1; e->role[3] = r3; }
        else { e->role[0] = r2; e->role[1] = r3; e->role[2] = r0; e->role[3] = r1; }
    } else {
        e->scur = 1 - e->scur;
    }
}

// make every tile table a step sequence needs (no allocation during capture)
int prepare_tiles(Engine *e, long nsteps, int kmax)
{
    long done = 0;
    while (done < nsteps) {
        const int k = next_depth(kmax, nsteps - done);
        Engine::Tiles t;
        int rc;
        if ((rc = get_tiles(e, k, &t))) return rc;
        done += k;
        if (k == kmax) done += (nsteps - done) / kmax * kmax;  // the rest of the full blocks
    }
    return WV_OK;
}

// coefficient modes present in the engine: the kernel variant built for just those
int step_cm(const Engine *e)
{
    int cm = e->g[0].coef;
    bool extras = false;
    for (const Grid &G : e->g) {
        if (G.coef != cm) cm = WV_CM_ALL;
        extras = extras || !G.h_ni.empty() || !G.h_di.empty() || G.vmax > 0 || !G.h_pi.empty();
    }
    if (extras && e->form == WV_FORM_FP64) cm = WV_CM_ALLX;  // the variant with those paths
    if (e->form == WV_FORM_FP64 && any_table(e)) cm = WV_CM_TBL;
    return cm;
}

// one stepping launch of depth k over ni + nb tiles starting at T
int launch_depth(Engine *e, int k, int cm, const WvTile *T, int ni, int nb, long done)
{
    if (e->form == WV_FORM_FP64)
        return wv_launch_fp64(k, e->stencil, cm, e->d_inst, T, ni, T + ni, nb, e->d_gstep, done, e->role[0],
                              e->role[1], e->role[2], e->role[3], e->stream);
    return wv_launch_split(k, e->stencil, cm, e->d_inst, T, ni, T + ni, nb, e->d_gstep, done, e->scur,
                           1 - e->scur, e->stream);
}

// at least n timing events exist
int reserve_kev(Engine *e, size_t n)
{
    if (!e->kev_end) HIPCHK(e, hipEventCreate(&e->kev_end));
    while (e->kev.size() < n) {
        hipEvent_t ev;
        HIPCHK(e, hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
        e->kev.push_back(ev);
        e->kev_depth.push_back(0);
        e->kev_epochs.push_back(1);
    }
    return WV_OK;
}

// record timing event number q on the stepping stream (timed runs)
int record_kev(Engine *e, int depth, int epochs = 1)
{
    int rc;
    if ((rc = reserve_kev(e, e->kev_used + 1))) return rc;
    e->kev_depth[e->kev_used] = depth;
    e->kev_epochs[e->kev_used] = epochs;
    HIPCHK(e, hipEventRecord(e->kev[e->kev_used], e->stream));
    e->kev_used++;
    return WV_OK;
}

// the persistent launch's work counters for `epochs` epochs (grown, not
// shrunk; wv_prepare sizes them before a timed run)
int persist_counters(Engine *e, long epochs)
{
    const long n = 8 * epochs * WV_PCTR_PAD;
    if (n <= e->pctr_cap) return WV_OK;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (e->d_pctr) HIPCHK(e, hipFree(e->d_pctr));
    e->d_pctr = nullptr;
    e->pctr_cap = 0;
    HIPCHK(e, hipMalloc(&e->d_pctr, sizeof(int) * n));
    e->pctr_cap = n;
    return WV_OK;
}

// enqueue nsteps steps; force_k1 keeps one step per launch (drop-in forms)
int enqueue_steps(Engine *e, long nsteps, bool force_k1, long *launches)
{
    long done = 0, nl = 0;
    const int kmax = max_depth(e, force_k1);
    const int cm = step_cm(e);
    int rc;
    // timed runs: an event before the first launch and one after each launch
    if (e->time_kernels && (rc = record_kev(e, 0))) return rc;
    while (done < nsteps) {
        const int k = next_depth(kmax, nsteps - done);
        Engine::Tiles t;
        if ((rc = get_tiles(e, k, &t))) return rc;
        const long ep = (nsteps - done) / kmax;  // full-depth launches left
        if (e->persistent && !e->wide && e->form == WV_FORM_F32SPLIT && !force_k1 && k == kmax && ep >= 2 &&
            ep <= INT_MAX / 8) {
            if ((rc = persist_counters(e, ep))) return rc;
            int g = 0;
            const int lr = wv_launch_split_persist(k, e->stencil, cm, e->d_inst, t.d, t.ni, e->d_gstep, done, (int)ep,
                                                   e->scur, e->d_pctr, e->stream, &g);
            if (lr) return fail(e, WV_EHIP, "persistent launch failed: %s", hipGetErrorString((hipError_t)lr));
            for (long q = 0; q < ep; q++) advance_roles(e, k);
            if (e->time_kernels && (rc = record_kev(e, k, (int)ep))) return rc;
            done += ep * k;
            nl += ep;
            continue;
        }
        const int lr = launch_depth(e, k, cm, t.d, t.ni, t.nb, done);
        advance_roles(e, k);
        if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
        if (e->time_kernels && (rc = record_kev(e, k))) return rc;
        done += k;
        nl++;
    }
    int lr = wv_launch_bump(e->d_gstep, nsteps, e->stream);
    if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
    if (e->time_kernels && (rc = record_kev(e, -1))) return rc;  // end of the sequence
    e->steps_done += nsteps;
    if (launches) *launches = nl;
    return WV_OK;
}

long rot_key(const Engine *e)
{
    return e->form == WV_FORM_FP64 ? e->role[0] + 4 * e->role[1] + 16 * e->role[2] + 64 * e->role[3] : e->scur;
}

// the executable graph of an nsteps sequence from the current rotation
// state (batches), captured on first use
int step_graph(Engine *e, long nsteps, hipGraphExec_t *out)
{
    const int kmax = max_depth(e, false);
    int rc;
    if ((rc = prepare_tiles(e, nsteps, kmax))) return rc;
    const auto key = std::make_pair(nsteps, rot_key(e));
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        int role0[4];
        memcpy(role0, e->role, sizeof role0);
        const int s0 = e->scur;
        const long sd0 = e->steps_done;
        hipGraph_t graph;
        HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        long nl = 0;
        rc = enqueue_steps(e, nsteps, false, &nl);
        hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        // capture only recorded the work: rewind the bookkeeping
        memcpy(e->role, role0, sizeof role0);
        e->scur = s0;
        e->steps_done = sd0;
        if (rc) return rc;
        if (ce != hipSuccess) return fail(e, WV_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
        hipGraphExec_t ex;
        hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(e, WV_EHIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        it = e->graphs.emplace(key, ex).first;
    }
    *out = it->second;
    return WV_OK;
}

int engine_step(Engine *e, long nsteps, bool force_k1)
{
    if (nsteps < 0) return fail(e, WV_EINVAL, "negative step count");
    int rc = ready(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    if (nsteps == 0) return WV_OK;
    if (!e->use_graph || force_k1 || e->persistent) {
        rc = enqueue_steps(e, nsteps, force_k1, &e->last_launches);
        return rc;
    }
    // batch: capture the launch sequence once per (nsteps, rotation) and replay
    hipGraphExec_t ex;
    if ((rc = step_graph(e, nsteps, &ex))) return rc;
    HIPCHK(e, hipGraphLaunch(ex, e->stream));
    // advance the host-side rotation exactly as the captured sequence does
    const int kmax = max_depth(e, false);
    long done = 0, nl = 0;
    while (done < nsteps) {
        const int k = next_depth(kmax, nsteps - done);
        advance_roles(e, k);
        done += k;
        nl++;
    }
    e->steps_done += nsteps;
    e->last_launches = nl;
    return WV_OK;
}

// Everything a later engine_step(nsteps) would do on the host before its
// first launch, done now: the tile tables of every fused depth the sequence
// uses, the first load of every stepping kernel it launches (one launch over a
// padding tile, which returns at once), the energy kernel, and for batches the
// captured graph.  The state does not change.
int kernel_times(Engine *e);

int engine_prepare(Engine *e, long nsteps)
{
    if (nsteps < 0) return fail(e, WV_EINVAL, "negative step count");
    int rc = ready(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    if (nsteps == 0) return WV_OK;
    const int kmax = max_depth(e, false);
    if ((rc = prepare_tiles(e, nsteps, kmax))) return rc;
    const int cm = step_cm(e);
    bool seen[9] = {false};
    long done = 0;
    while (done < nsteps) {
        const int k = next_depth(kmax, nsteps - done);
        if (!seen[k]) {
            seen[k] = true;
            const int lr = launch_depth(e, k, cm, e->d_pad, 1, 0, 0);
            if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
            if (e->persistent && !e->wide && e->form == WV_FORM_F32SPLIT && k == kmax) {
                if ((rc = persist_counters(e, nsteps / kmax))) return rc;
                int g = 0;
                const int pr = wv_launch_split_persist(k, e->stencil, cm, e->d_inst, e->d_pad, 1, e->d_gstep, 0, 1,
                                                       e->scur, e->d_pctr, e->stream, &g);
                if (pr) return fail(e, WV_EHIP, "persistent launch failed: %s", hipGetErrorString((hipError_t)pr));
            }
        }
        done += k;
        if (k == kmax) done += (nsteps - done) / kmax * kmax;
    }
    int lr = wv_launch_bump(e->d_gstep, 0, e->stream);
    if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
    // the events a timed run of this sequence records (one per launch and two
    // more); a timed run's durations not read yet are read before they are
    // recorded over
    if ((rc = kernel_times(e))) return rc;
    long nl = 0;
    for (long d = 0; d < nsteps; nl++) d += next_depth(kmax, nsteps - d);
    if ((rc = reserve_kev(e, (size_t)nl + 2))) return rc;
    // record each once: the runtime sets an event up on its first record
    for (size_t q = 0; q < (size_t)nl + 2; q++) HIPCHK(e, hipEventRecord(e->kev[q], e->stream));
    HIPCHK(e, hipEventRecord(e->kev_end, e->stream));
    if (e->use_graph && !e->persistent) {
        hipGraphExec_t ex;
        if ((rc = step_graph(e, nsteps, &ex))) return rc;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return WV_OK;
}

// ------------------------------------------------------------ media tables

// init_ior_2d's Gaussian-well media, built in place: the mask planes on the
// host (as wv_set_tables), tcc and tc by the GPU into the engine's per-cell
// tables (field coefficient mode, one tgamma)
int set_wells_tables(Engine *e, int k, const short *xy, int nw, const double *xc, const double *yc, const double *h,
                     double sigma, int use_tanh, double courant, double courantb, double gamma, const double *box,
                     double *tcc_out, double *tc_out, double *tg_out)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!xy || !box || (nw > 0 && (!xc || !yc))) return fail(e, WV_EINVAL, "mask, box or well arrays missing");
    if (nw < 0) return fail(e, WV_EINVAL, "negative well count");
    if (!(sigma > 0) || !(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "sigma or box");
    Grid &G = e->g[k];
    const int nx = G.p.nx, ny = G.p.ny, ldw = G.ldw, ldy = G.ldy;
    const bool ts = G.p.twospeeds != 0, cmp = (G.p.flags & WV_FLAG_COMPARISON) != 0;
    // mask planes, the rules of set_tables
    std::vector<uint64_t> act((size_t)nx * ldw, 0), eact((size_t)nx * ldw, 0), nzp((size_t)nx * ldw, 0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const size_t c = (size_t)i * ny + j;
            const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (cmp && (j == ny / 2 - 1 || j == ny / 2));
            const size_t w = (size_t)i * ldw + (j >> 6);
            const uint64_t b = 1ull << (j & 63);
            if (ts || (edge ? xy[c] != 0 : xy[c] > 0)) act[w] |= b;
            if (ts || xy[c] != 0) eact[w] |= b;
            if (xy[c] != 0) nzp[w] |= b;
        }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const size_t wb = (size_t)nx * ldw * 8, cells = (size_t)nx * ldy;
    HIPCHK(e, hipMemcpy(G.eact, eact.data(), wb, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemset(G.med, 0, wb));
    HIPCHK(e, hipMemcpy(G.nz, nzp.data(), wb, hipMemcpyHostToDevice));
    G.h_nz.swap(nzp);
    for (double **q : {&G.tgf}) if (*q) { (void)hipFree(*q); *q = nullptr; }
    for (float **q : {&G.tgf32}) if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (!G.tccf) HIPCHK(e, hipMalloc(&G.tccf, cells * sizeof(double)));
    const bool split = e->form == WV_FORM_F32SPLIT;
    if (split) {
        if (!G.tccf_hi) HIPCHK(e, hipMalloc(&G.tccf_hi, cells * sizeof(float)));
        if (!G.tccf_lo) HIPCHK(e, hipMalloc(&G.tccf_lo, cells * sizeof(float)));
    }
    double *d_w = nullptr;
    const size_t nwv = 3 * (size_t)std::max(nw, 1);
    std::vector<double> w(nwv, 1.0);
    for (int n = 0; n < nw; n++) {
        w[n] = xc[n];
        w[nw + n] = yc[n];
        if (h) w[2 * nw + n] = h[n];
    }
    HIPCHK(e, hipMalloc(&d_w, sizeof(double) * nwv));
    hipError_t err = hipMemcpy(d_w, w.data(), sizeof(double) * nwv, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        WvWells a{};
        a.nx = nx; a.ny = ny; a.ldy = ldy; a.ldw = ldw; a.nw = nw; a.use_tanh = use_tanh != 0; a.twospeeds = ts;
        a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
        a.sigma = sigma; a.courant = courant; a.courantb = courantb;
        a.xc = d_w; a.yc = d_w + nw; a.h = h ? d_w + 2 * nw : nullptr;
        a.nz = G.nz; a.tcc = G.tccf; a.tc = G.tc;
        a.tcc_hi = split ? G.tccf_hi : nullptr; a.tcc_lo = split ? G.tccf_lo : nullptr;
        a.split_weight = split_weight(e);
        err = (hipError_t)wv_launch_wells(a, e->stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(d_w);
    if (err != hipSuccess) return fail(e, WV_EHIP, "wells media: %s", hipGetErrorString(err));
    if (tcc_out) HIPCHK(e, hipMemcpy2D(tcc_out, ny * sizeof(double), G.tccf, ldy * sizeof(double), ny * sizeof(double),
                                      nx, hipMemcpyDeviceToHost));
    if (tc_out) HIPCHK(e, hipMemcpy2D(tc_out, ny * sizeof(double), G.tc, ldy * sizeof(double), ny * sizeof(double), nx,
                                     hipMemcpyDeviceToHost));
    if (tg_out) std::fill(tg_out, tg_out + (size_t)nx * ny, gamma);
    G.coef = WV_COEF_FIELD;
    G.tcc0 = G.tcc1 = courantb;
    G.tg0 = G.tg1 = gamma;
    // as set_tables: a state set under another mask keeps its values, one step per launch
    G.h_act_plain = act;
    if (G.tbl) act = table_act(G);
    if (G.state && G.h_act != act) G.exact = true;
    HIPCHK(e, hipMemcpy(G.act, act.data(), wb, hipMemcpyHostToDevice));
    G.h_act.swap(act);
    rebuild_written(G);
    G.tables = true;
    drop_tiles(e);
    return sync_inst(e);
}

// --------------------------------------------------- initial drop, pulses

// is there a cell the drawing rule reaches (xy_in != 0 or TWOSPEEDS) that no
// step writes?  A drop or pulse that lands there leaves a non-zero value on
// a never-written cell, which the two-level fused launches would not keep
bool drawn_unwritten(const Grid &G)
{
    const int nx = G.p.nx, ny = G.p.ny, ldw = G.ldw;
    if (G.h_wr.empty()) return false;
    for (int i = 0; i < nx; i++)
        for (int w = 0; w < ldw; w++) {
            const int nb = std::min(64, ny - 64 * w);
            if (nb <= 0) break;
            const uint64_t valid = nb == 64 ? ~0ull : ((1ull << nb) - 1);
            const uint64_t drawn = G.p.twospeeds ? valid : (G.h_nz[(size_t)i * ldw + w] & valid);
            if (drawn & ~G.h_wr[(size_t)i * ldw + w]) return true;
        }
    return false;
}

// init_circular_wave (mode 0) / add_circular_wave (mode 1) on grid k
int circular_wave(Engine *e, int k, int mode, double factor, double x0, double y0, double amp, double variance,
                  double wavelength, const double *box)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!box || !(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "box missing or empty");
    if (!std::isfinite(factor) || !std::isfinite(x0) || !std::isfinite(y0) || !std::isfinite(amp) ||
        !(variance > 0) || !std::isfinite(variance) || !(wavelength != 0) || !std::isfinite(wavelength))
        return fail(e, WV_EINVAL, "drop parameters must be finite, with variance > 0 and wavelength != 0");
    Grid &G = e->g[k];
    if (!G.tables) return fail(e, WV_ESTATE, "set the tables (the mask) before a drop");
    if (mode == 1 && !G.state) return fail(e, WV_ESTATE, "no state to add a pulse to");
    HIPCHK(e, hipSetDevice(e->device));
    WvPulse a{};
    a.nx = G.p.nx; a.ny = G.p.ny; a.ldy = G.ldy; a.ldw = G.ldw; a.mode = mode; a.twospeeds = G.p.twospeeds != 0;
    a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
    a.x0 = x0; a.y0 = y0; a.amp = amp; a.factor = factor; a.variance = variance; a.wavelength = wavelength;
    a.nz = G.nz;
    const size_t bytes = (size_t)G.p.nx * G.ldy * sizeof(double);
    if (e->form == WV_FORM_FP64) {
        a.phi = G.u[e->role[0]];
        a.psi = G.u[e->role[1]];
        if (mode == 0) {
            // the other two buffers start at zero, as wv_set_state leaves them
            HIPCHK(e, hipMemsetAsync(G.u[e->role[2]], 0, bytes, e->stream));
            HIPCHK(e, hipMemsetAsync(G.u[e->role[3]], 0, bytes, e->stream));
        }
    } else {
        for (int s = 0; s < 2; s++) { a.hi[s] = G.hi[s]; a.lo[s] = G.lo[s]; a.v[s] = G.v[s]; }
    }
    const int lr = wv_launch_pulse(a, e->stream);
    if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
    if (e->form == WV_FORM_FP64) {
        // finite values, no -0 (x + t is -0 only if both are): the state stays
        // clean; values on never-written cells need one step per launch
        const bool ex = drawn_unwritten(G);
        G.exact = mode == 0 ? ex : (G.exact || ex);
        if (mode == 0 && !G.clean_state) {
            G.clean_state = true;
            int rc = sync_inst(e);
            if (rc) return rc;
        }
    }
    G.state = true;
    return WV_OK;
}

// ------------------------------------------------------------ frame colours

int frame_set_tables(Engine *e, int k, const float *lut, int lut_n, const double *cs, const double *ft)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    if (lut && lut_n != 256 && lut_n != 511) return fail(e, WV_EINVAL, "palette table of %d entries (256 or 511)", lut_n);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (void **q : {(void **)&G.fr_lut, (void **)&G.fr_cs, (void **)&G.fr_ft})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    G.fr_lut_n = 0;
    const int nx = G.p.nx, ny = G.p.ny;
    if (lut) {
        HIPCHK(e, hipMalloc(&G.fr_lut, sizeof(float) * 3 * lut_n));
        HIPCHK(e, hipMemcpy(G.fr_lut, lut, sizeof(float) * 3 * lut_n, hipMemcpyHostToDevice));
        G.fr_lut_n = lut_n;
    }
    for (auto pr : {std::make_pair(&G.fr_cs, cs), std::make_pair(&G.fr_ft, ft)}) {
        if (!pr.second) continue;
        HIPCHK(e, hipMalloc(pr.first, sizeof(double) * (size_t)nx * G.ldy));
        HIPCHK(e, hipMemcpy2D(*pr.first, G.ldy * sizeof(double), pr.second, ny * sizeof(double), ny * sizeof(double),
                              nx, hipMemcpyHostToDevice));
    }
    return WV_OK;
}

int frame_reset(Engine *e, int k)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    Grid &G = e->g[k];
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (double **q : {&G.fr_te[0], &G.fr_te[1], &G.fr_ae[0], &G.fr_ae[1], &G.fr_tf[0], &G.fr_tf[1]})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (G.fr_out) { (void)hipFree(G.fr_out); G.fr_out = nullptr; }
    G.fr_size = 0;
    G.fr_te_cur = G.fr_ae_cur = G.fr_tf_cur = 0;
    return WV_OK;
}

int frame(Engine *e, int k, const wv_frame_params *f, void *rgba, bool u8)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!f || !rgba) return fail(e, WV_EINVAL, "frame parameters or output is NULL");
    int rc = ready(e);
    if (rc) return rc;
    Grid &G = e->g[k];
    const int pal = f->palette, sch = f->scheme, plot = f->plot;
    if (plot < 0 || plot > 9) return fail(e, WV_EINVAL, "plot %d is not a wave plot type 0..9", plot);
    if (sch != 0 && sch != 1 && sch != 3 && sch != 4 && sch != 5)
        return fail(e, WV_EUNSUPPORTED, "colour scheme %d (0, 1, 3, 4, 5 are provided)", sch);
    const bool table = (pal >= 10 && pal <= 18) || pal == 101;
    if (pal != 0 && pal != 1 && pal != 2 && !table)
        return fail(e, WV_EUNSUPPORTED, "palette %d (COL_JET, COL_HSLUV, COL_GRAY and the table palettes are provided)",
                    pal);
    const int need = (pal == 17 || pal == 18) ? 511 : 256;
    if (table && G.fr_lut_n != need)
        return fail(e, WV_ESTATE, "palette %d needs its table of %d entries (wv_frame_set_tables)", pal, need);
    // the asymmetric schemes leave the colour unset for the linear schemes
    const bool asym_used = plot == 6 || (f->asym && (plot == 1 || plot == 3 || plot == 8));
    if (asym_used && (sch == 4 || sch == 5))
        return fail(e, WV_EUNSUPPORTED, "plot %d with scheme %d: color_scheme_asym_palette sets no colour", plot, sch);
    if (f->size < 1) return fail(e, WV_EINVAL, "block size must be >= 1");
    if (f->rescale_center && !G.fr_cs) return fail(e, WV_ESTATE, "RESCALE_COLOR_IN_CENTER needs color_scale");
    if (f->fade_in_obstacle && !G.fr_ft) return fail(e, WV_ESTATE, "FADE_IN_OBSTACLE needs fade_table");
    if (sch == 0 && f->nsteps == 0) return fail(e, WV_EINVAL, "C_LUM divides by nsteps");
    const int nx = G.p.nx, ny = G.p.ny, s = f->size;
    const int nbx = (nx + s - 1) / s, nby = (ny + s - 1) / s;
    if (nbx > 4 * 65535) return fail(e, WV_EINVAL, "more than 262140 block rows");
    HIPCHK(e, hipSetDevice(e->device));
    const size_t nb = (size_t)nbx * nby;
    if (G.fr_size != s) {
        if ((rc = frame_reset(e, k))) return rc;
        for (double **q : {&G.fr_te[0], &G.fr_te[1], &G.fr_ae[0], &G.fr_ae[1]}) {
            HIPCHK(e, hipMalloc(q, sizeof(double) * nb));
            HIPCHK(e, hipMemset(*q, 0, sizeof(double) * nb));
        }
        for (double **q : {&G.fr_tf[0], &G.fr_tf[1]}) {
            HIPCHK(e, hipMalloc(q, sizeof(double) * 2 * nb));
            HIPCHK(e, hipMemset(*q, 0, sizeof(double) * 2 * nb));
        }
        HIPCHK(e, hipMalloc(&G.fr_out, sizeof(float4) * nb));
        G.fr_size = s;
    }
    WvFrame a{};
    a.plot = plot; a.scheme = sch; a.palette = pal; a.asym = f->asym != 0; a.size = s; a.time = f->time;
    a.shade = f->shade != 0; a.rescale = f->rescale_center != 0; a.fadeobs = f->fade_in_obstacle != 0;
    a.fade = f->fade != 0; a.nsteps = f->nsteps;
    a.scale = f->scale; a.fade_value = f->fade_value; a.slope = f->slope; a.color_range = f->color_range;
    a.attenuation = f->attenuation; a.vshift = f->vshift_amplitude; a.vscale = f->vscale_amplitude;
    a.log_shift = f->log_shift; a.log_scale = f->log_scale; a.flux_scale = f->flux_scale;
    a.avrg = f->avrg_e_factor; a.colorhue = f->colorhue; a.colordrift = f->colordrift; a.lummean = f->lummean;
    a.lumamp = f->lumamp; a.huemean = f->huemean; a.hueamp = f->hueamp; a.shade_scale = f->shade_scale;
    a.light_x = f->light_x; a.light_y = f->light_y; a.courant = f->courant; a.courantb = f->courantb;
    a.es = G.p.e_scale;
    a.nx = nx; a.ny = ny; a.ldy = G.ldy; a.ldw = G.ldw; a.nbx = nbx; a.nby = nby; a.twospeeds = G.p.twospeeds != 0;
    a.lut_n = G.fr_lut_n;
    if (e->form == WV_FORM_FP64) {
        a.phi = G.u[e->role[0]];
        a.psi = G.u[e->role[1]];
    } else {
        a.hi = G.hi[e->scur]; a.lo = G.lo[e->scur]; a.v = G.v[e->scur];
    }
    a.nz = G.nz; a.cs = G.fr_cs; a.ft = G.fr_ft; a.lut = G.fr_lut;
    a.te0 = G.fr_te[G.fr_te_cur]; a.te1 = G.fr_te[1 - G.fr_te_cur];
    a.ae0 = G.fr_ae[G.fr_ae_cur]; a.ae1 = G.fr_ae[1 - G.fr_ae_cur];
    a.tf0 = G.fr_tf[G.fr_tf_cur]; a.tf1 = G.fr_tf[1 - G.fr_tf_cur];
    a.out = G.fr_out;
    a.out8 = u8 ? (uchar4 *)G.fr_out : nullptr;  // the float buffer holds the 8-bit colours too
    int lr = wv_launch_frame(a, e->stream);
    if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
    if (plot == 3 || plot == 5) G.fr_te_cur ^= 1;
    if (plot == 8 || plot == 9) G.fr_ae_cur ^= 1;
    if (plot == 7) G.fr_tf_cur ^= 1;
    HIPCHK(e, hipMemcpyAsync(rgba, G.fr_out, (u8 ? sizeof(uchar4) : sizeof(float4)) * nb, hipMemcpyDeviceToHost,
                             e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return WV_OK;
}

int engine_energy(Engine *e, double *out)
{
    int rc = ready(e);
    if (rc) return rc;
    if (!out) return fail(e, WV_EINVAL, "output pointer is NULL");
    HIPCHK(e, hipSetDevice(e->device));
    const bool split = e->form == WV_FORM_F32SPLIT;
    int lr = wv_launch_energy(split, e->d_inst, e->d_erows, e->nrows_total, e->d_partial, e->d_row0, e->n,
                              e->d_eout, split ? e->scur : e->role[0], split ? 0 : e->role[1], e->stream);
    if (lr) return fail(e, WV_EHIP, "energy launch failed: %s", hipGetErrorString((hipError_t)lr));
    HIPCHK(e, hipMemcpyAsync(out, e->d_eout, sizeof(double) * e->n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return WV_OK;
}

int engine_step_timed(Engine *e, long nsteps, double *ms, long *launches)
{
    int rc = ready(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    // per-kernel events need the launches themselves, not a replayed graph
    const bool graph = e->use_graph;
    e->use_graph = false;
    e->time_kernels = true;
    e->kev_used = 0;
    e->kev_pending = false;
    e->last_avg_ms = e->last_timed = 0;
    rc = engine_step(e, nsteps, false);
    e->time_kernels = false;
    e->use_graph = graph;
    if (rc) return rc;
    if (e->kev_used < 3) {  // no launch (nsteps == 0)
        if (ms) *ms = 0;
        if (launches) *launches = 0;
        return WV_OK;
    }
    // wait by polling the closing (fenced) event: a blocking wait can sleep
    // and wake up tens of microseconds after the work is done, which a caller
    // timing a short sequence by wall clock would count
    HIPCHK(e, hipEventRecord(e->kev_end, e->stream));
    hipError_t q;
    while ((q = hipEventQuery(e->kev_end)) == hipErrorNotReady) {
    }
    if (q != hipSuccess) return fail(e, WV_EHIP, "waiting for the timed steps: %s", hipGetErrorString(q));
    float t = 0;
    HIPCHK(e, hipEventElapsedTime(&t, e->kev[0], e->kev[e->kev_used - 1]));
    e->kev_pending = true;  // per-launch durations are read when asked for
    if (ms) *ms = t;
    if (launches) *launches = e->last_launches;
    return WV_OK;
}

// per-launch durations of the last timed run: launch q runs between events q
// and q+1; the average is over the launches of full depth (a shorter last one
// differs)
int kernel_times(Engine *e)
{
    if (!e->kev_pending) return WV_OK;
    e->kev_pending = false;
    const int kmax = max_depth(e, false);
    double ti = 0;
    int nt = 0;
    for (size_t q = 1; q + 1 < e->kev_used; q++) {
        if (e->kev_depth[q] != kmax) continue;
        float x = 0;
        HIPCHK(e, hipEventElapsedTime(&x, e->kev[q - 1], e->kev[q]));
        ti += x;
        nt += e->kev_epochs[q];
    }
    e->last_avg_ms = nt ? ti / nt : 0.0;
    e->last_timed = nt;
    return WV_OK;
}

int engine_set_tiling(Engine *e, int k, int L)
{
    if (k < 0 || k > 8 || (k && !wv_have_depth(k))) return fail(e, WV_EINVAL, "tb_depth must be 0 or one of 1,2,3,4,6,8");
    if (L < 0) return fail(e, WV_EINVAL, "rows_per_tile must be >= 0");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->K = k;
    e->L = L;
    drop_tiles(e);
    return sync_inst(e);
}

}  // namespace

struct wv_ctx {
    Engine *e;
};
struct wv_batch {
    Engine *e;
};

static const char *no_ctx = "no context";

extern "C" {

void wv_params_default(wv_params *p, int nx, int ny)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->nx = nx;
    p->ny = ny;
    p->stencil = 5;
    p->form = WV_FORM_FP64;
    p->tb_shift = nx;
    p->e_scale = 1.0;
}

int wv_create(wv_ctx **out, const wv_params *p)
{
    if (!out) return WV_EINVAL;
    *out = new wv_ctx{nullptr};
    return engine_create(&(*out)->e, 1, p);
}

void wv_destroy(wv_ctx *c)
{
    if (!c) return;
    engine_destroy(c->e);
    delete c;
}

const char *wv_last_error(const wv_ctx *c) { return c && c->e ? c->e->err.c_str() : no_ctx; }

int wv_set_tables(wv_ctx *c, const short *xy_in, const double *tcc, const double *tc, const double *tgamma)
{
    if (!c) return WV_EINVAL;
    return set_tables(c->e, 0, xy_in, tcc, tc, tgamma);
}

int wv_set_state(wv_ctx *c, const double *phi, const double *psi)
{
    if (!c) return WV_EINVAL;
    return set_state(c->e, 0, phi, psi, nullptr);
}

int wv_set_state3(wv_ctx *c, const double *phi, const double *psi, const double *tmp)
{
    if (!c) return WV_EINVAL;
    if (!tmp) return fail(c->e, WV_EINVAL, "tmp is NULL");
    if (c->e->form != WV_FORM_FP64) return fail(c->e, WV_EUNSUPPORTED, "three-buffer state needs the FP64 form");
    return set_state(c->e, 0, phi, psi, tmp);
}

int wv_get_state(wv_ctx *c, double *phi, double *psi)
{
    if (!c) return WV_EINVAL;
    return get_state(c->e, 0, phi, psi, nullptr);
}

int wv_get_state3(wv_ctx *c, double *phi, double *psi, double *tmp)
{
    if (!c) return WV_EINVAL;
    if (c->e->form != WV_FORM_FP64) return fail(c->e, WV_EUNSUPPORTED, "three-buffer state needs the FP64 form");
    return get_state(c->e, 0, phi, psi, tmp);
}

int wv_set_sources(wv_ctx *c, int n, const int *i, const int *j, const double *amp, long amp_len)
{
    if (!c) return WV_EINVAL;
    return set_sources(c->e, 0, n, i, j, amp, amp_len);
}

int wv_step(wv_ctx *c, long nsteps)
{
    if (!c) return WV_EINVAL;
    return engine_step(c->e, nsteps, false);
}

int wv_prepare(wv_ctx *c, long nsteps)
{
    if (!c) return WV_EINVAL;
    return engine_prepare(c->e, nsteps);
}
int wv_step_timed(wv_ctx *c, long nsteps, double *ms, long *launches)
{
    if (!c) return WV_EINVAL;
    return engine_step_timed(c->e, nsteps, ms, launches);
}

int wv_sync(wv_ctx *c)
{
    if (!c) return WV_EINVAL;
    Engine *e = c->e;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return WV_OK;
}

long wv_steps_done(const wv_ctx *c) { return c ? c->e->steps_done : -1; }

int wv_last_kernel_ms(const wv_ctx *c, double *avg_ms, double *timed_launches)
{
    if (!c) return WV_EINVAL;
    int rc = kernel_times(c->e);
    if (rc) return rc;
    if (avg_ms) *avg_ms = c->e->last_avg_ms;
    if (timed_launches) *timed_launches = c->e->last_timed;
    return WV_OK;
}

int wv_energy(wv_ctx *c, double *e)
{
    if (!c) return WV_EINVAL;
    return engine_energy(c->e, e);
}

void wv_frame_params_default(wv_frame_params *f)
{
    if (!f) return;
    memset(f, 0, sizeof *f);
    // the drawing macros of the checked-in wave_billiard.c:200-241, 287
    f->plot = 0;
    f->scheme = 3;
    f->palette = 0;
    f->size = 1;
    f->nsteps = 500;
    f->scale = 1.0;
    f->fade_value = 1.0;
    f->slope = 1.0;
    f->color_range = 0.8;
    f->vshift_amplitude = -0.2;
    f->vscale_amplitude = 1.0;
    f->log_shift = 0.75;
    f->log_scale = 0.75;
    f->flux_scale = 250.0;
    f->avrg_e_factor = 0.95;
    f->colorhue = 260.0;
    f->lummean = 0.5;
    f->lumamp = 0.3;
    f->huemean = 180.0;
    f->hueamp = -180.0;
    f->shade_scale = 0.5;
    f->light_x = 0.40824829;
    f->light_y = 0.816496581;
    f->courant = 0.1;
    f->courantb = 0.05;
}

int wv_frame_set_tables(wv_ctx *c, const float *lut, int lut_n, const double *color_scale, const double *fade_table)
{
    if (!c) return WV_EINVAL;
    return frame_set_tables(c->e, 0, lut, lut_n, color_scale, fade_table);
}

int wv_frame_reset(wv_ctx *c)
{
    if (!c) return WV_EINVAL;
    return frame_reset(c->e, 0);
}

int wv_frame(wv_ctx *c, const wv_frame_params *f, float *rgba)
{
    if (!c) return WV_EINVAL;
    return frame(c->e, 0, f, rgba, false);
}

int wv_frame_rgba8(wv_ctx *c, const wv_frame_params *f, unsigned char *rgba)
{
    if (!c) return WV_EINVAL;
    return frame(c->e, 0, f, rgba, true);
}

int wv_get_tiling(const wv_ctx *c, int *tb_depth, int *rows_per_tile, int *tile_cols)
{
    if (!c) return WV_EINVAL;
    const Engine *e = c->e;
    if (tb_depth) *tb_depth = depth(e);
    if (rows_per_tile) *rows_per_tile = rows(e);
    if (tile_cols) *tile_cols = wv_tile_cols(depth(e));
    return WV_OK;
}

// tiles of one launch of depth tb_depth (0: the current depth), by kind, and
// the output cells they cover (a cell written by two tiles counts twice)
static int tile_counts(Engine *e, int k, int *wide, int *narrow, int *groups, long *cells)
{
    if (k == 0) k = depth(e);
    if (!wv_have_depth(k)) return fail(e, WV_EINVAL, "no kernel for fused depth %d", k);
    int rc = ready(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    Engine::Tiles t;
    if ((rc = get_tiles(e, k, &t))) return rc;
    const int G = e->form == WV_FORM_F32SPLIT ? WV_WG : 1;
    std::vector<WvTile> h((size_t)t.ni * G);
    if (!h.empty()) HIPCHK(e, hipMemcpy(h.data(), t.d, sizeof(WvTile) * h.size(), hipMemcpyDeviceToHost));
    int nw = 0, nn = 0;
    long nc = 0;
    for (size_t q = 0; q < h.size(); q++) {
        const WvTile &u = h[q];
        if (u.inst < 0) continue;
        const bool w = u.flags & WV_TILE_WIDE;
        if (w && q % G) continue;  // a wide tile fills its workgroup's entries
        (w ? nw : nn)++;
        const Grid &Gq = e->g[u.inst];
        const int j1 = std::min(u.jb + ((u.flags >> WV_TILE_WIDTH_SHIFT) & 255), Gq.p.ny);
        const int i1 = std::min(u.i0 + (u.flags >> WV_TILE_ROWS_SHIFT), Gq.p.nx);
        nc += (long)std::max(0, j1 - std::max(u.jb, 0)) * std::max(0, i1 - u.i0);
    }
    if (wide) *wide = nw;
    if (narrow) *narrow = nn;
    if (groups) *groups = t.ni;
    if (cells) *cells = nc;
    return WV_OK;
}

int wv_get_tile_counts(wv_ctx *c, int tb_depth, int *wide, int *narrow, int *groups, long *cells)
{
    if (!c) return WV_EINVAL;
    return tile_counts(c->e, tb_depth, wide, narrow, groups, cells);
}

int wv_set_tiling(wv_ctx *c, int tb_depth, int rows_per_tile)
{
    if (!c) return WV_EINVAL;
    return engine_set_tiling(c->e, tb_depth, rows_per_tile);
}

int wv_autotune(wv_ctx *c, long steps_per_trial, double *best_gcells);

// ---- drop-in forms ----------------------------------------------------------

static int dropin_common(wv_ctx *c, int nsteps, double *const *a, double *const *b, double *const *t,
                         short *const *m, double *const *tcc, double *const *tc, double *const *tg,
                         bool rowptr, bool three)
{
    if (!c) return WV_EINVAL;
    Engine *e = c->e;
    if (e->form != WV_FORM_FP64) return fail(e, WV_EUNSUPPORTED, "the drop-in entry points use the FP64 form");
    const Grid &G = e->g[0];
    const int nx = G.p.nx, ny = G.p.ny;
    const size_t n = (size_t)nx * ny;
    if (!a || !b || !t || !m || !tcc || !tc || !tg) return fail(e, WV_EINVAL, "NULL argument");
    const double *pa, *pb, *pt;
    if (rowptr) {
        for (int k = 0; k < 3; k++) e->st[k].resize(n);
        for (int i = 0; i < nx; i++) {
            memcpy(&e->st[0][(size_t)i * ny], a[i], ny * sizeof(double));
            memcpy(&e->st[1][(size_t)i * ny], b[i], ny * sizeof(double));
            memcpy(&e->st[2][(size_t)i * ny], t[i], ny * sizeof(double));
        }
        pa = e->st[0].data(); pb = e->st[1].data(); pt = e->st[2].data();
    } else {
        pa = (const double *)a; pb = (const double *)b; pt = (const double *)t;
    }
    int rc;
    if (!e->dropin_tables || !(G.p.flags & WV_FLAG_TABLES_FIXED)) {
        if (rowptr) {
            e->stm.resize(n);
            for (int k = 0; k < 3; k++) e->stt[k].resize(n);
            for (int i = 0; i < nx; i++) {
                memcpy(&e->stm[(size_t)i * ny], m[i], ny * sizeof(short));
                memcpy(&e->stt[0][(size_t)i * ny], tcc[i], ny * sizeof(double));
                memcpy(&e->stt[1][(size_t)i * ny], tc[i], ny * sizeof(double));
                memcpy(&e->stt[2][(size_t)i * ny], tg[i], ny * sizeof(double));
            }
            rc = set_tables(e, 0, e->stm.data(), e->stt[0].data(), e->stt[1].data(), e->stt[2].data());
        } else {
            rc = set_tables(e, 0, (const short *)m, (const double *)tcc, (const double *)tc, (const double *)tg);
        }
        if (rc) return rc;
        e->dropin_tables = true;
    }
    // evolve_wave_half(phi_in, psi_in, phi_out): the three rotation buffers
    // are (phi_in, psi_in, phi_out); evolve_wave(phi, psi, tmp) likewise
    if ((rc = set_state(e, 0, pa, pb, pt))) return rc;
    if ((rc = engine_step(e, nsteps, true))) return rc;
    if (rowptr) {
        double *o[3] = {e->st[0].data(), e->st[1].data(), e->st[2].data()};
        if (three) {
            // after 3 steps the roles are back: phi, psi, tmp = u^{n+3}, u^{n+2}, u^{n+1}
            if ((rc = get_state(e, 0, o[0], o[1], o[2]))) return rc;
            for (int i = 0; i < nx; i++) {
                memcpy(a[i], &e->st[0][(size_t)i * ny], ny * sizeof(double));
                memcpy(b[i], &e->st[1][(size_t)i * ny], ny * sizeof(double));
                memcpy(t[i], &e->st[2][(size_t)i * ny], ny * sizeof(double));
            }
        } else {
            // one step: the new phi is the caller's phi_out
            if ((rc = get_state(e, 0, o[2], nullptr, nullptr))) return rc;
            for (int i = 0; i < nx; i++) memcpy(t[i], &e->st[2][(size_t)i * ny], ny * sizeof(double));
        }
    } else {
        if (three) rc = get_state(e, 0, (double *)a, (double *)b, (double *)t);
        else rc = get_state(e, 0, (double *)t, nullptr, nullptr);
        if (rc) return rc;
    }
    return WV_OK;
}

int wv_evolve_wave_half_rows(wv_ctx *c, double *const *phi_in, double *const *psi_in, double *const *phi_out,
                             short *const *xy_in, double *const *tcc, double *const *tc,
                             double *const *tgamma)
{
    return dropin_common(c, 1, phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma, true, false);
}

int wv_evolve_wave_rows(wv_ctx *c, double *const *phi, double *const *psi, double *const *tmp,
                        short *const *xy_in, double *const *tcc_table, double *const *tc_table,
                        double *const *tgamma_table)
{
    return dropin_common(c, 3, phi, psi, tmp, xy_in, tcc_table, tc_table, tgamma_table, true, true);
}

int wv_evolve_wave_half_flat(wv_ctx *c, double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                             double *tc, double *tcc, double *tgamma)
{
    return dropin_common(c, 1, (double *const *)phi_in, (double *const *)psi_in, (double *const *)phi_out,
                         (short *const *)xy_in, (double *const *)tcc, (double *const *)tc,
                         (double *const *)tgamma, false, false);
}

int wv_evolve_wave_flat(wv_ctx *c, double *phi, double *psi, double *tmp, short *xy_in, double *tc,
                        double *tcc, double *tgamma)
{
    return dropin_common(c, 3, (double *const *)phi, (double *const *)psi, (double *const *)tmp,
                         (short *const *)xy_in, (double *const *)tcc, (double *const *)tc,
                         (double *const *)tgamma, false, true);
}

// ---- batches ----------------------------------------------------------------

int wv_batch_create(wv_batch **out, int n, const wv_params *ps)
{
    if (!out) return WV_EINVAL;
    *out = new wv_batch{nullptr};
    int rc = engine_create(&(*out)->e, n, ps);
    if (!rc) (*out)->e->use_graph = true;
    return rc;
}

void wv_batch_destroy(wv_batch *b)
{
    if (!b) return;
    engine_destroy(b->e);
    delete b;
}

const char *wv_batch_last_error(const wv_batch *b) { return b && b->e ? b->e->err.c_str() : no_ctx; }

int wv_batch_set_tables(wv_batch *b, int k, const short *xy_in, const double *tcc, const double *tc,
                        const double *tgamma)
{
    if (!b) return WV_EINVAL;
    return set_tables(b->e, k, xy_in, tcc, tc, tgamma);
}

int wv_batch_set_state(wv_batch *b, int k, const double *phi, const double *psi)
{
    if (!b) return WV_EINVAL;
    return set_state(b->e, k, phi, psi, nullptr);
}

int wv_batch_get_state(wv_batch *b, int k, double *phi, double *psi)
{
    if (!b) return WV_EINVAL;
    return get_state(b->e, k, phi, psi, nullptr);
}

int wv_batch_step(wv_batch *b, long nsteps)
{
    if (!b) return WV_EINVAL;
    return engine_step(b->e, nsteps, false);
}

int wv_batch_prepare(wv_batch *b, long nsteps)
{
    if (!b) return WV_EINVAL;
    return engine_prepare(b->e, nsteps);
}
int wv_batch_step_timed(wv_batch *b, long nsteps, double *ms, long *launches)
{
    if (!b) return WV_EINVAL;
    return engine_step_timed(b->e, nsteps, ms, launches);
}

int wv_batch_last_kernel_ms(const wv_batch *b, double *avg_ms, double *timed_launches)
{
    if (!b) return WV_EINVAL;
    int rc = kernel_times(b->e);
    if (rc) return rc;
    if (avg_ms) *avg_ms = b->e->last_avg_ms;
    if (timed_launches) *timed_launches = b->e->last_timed;
    return WV_OK;
}

int wv_batch_energy(wv_batch *b, double *e)
{
    if (!b) return WV_EINVAL;
    return engine_energy(b->e, e);
}

int wv_batch_set_tiling(wv_batch *b, int tb_depth, int rows_per_tile)
{
    if (!b) return WV_EINVAL;
    return engine_set_tiling(b->e, tb_depth, rows_per_tile);
}

int wv_batch_get_tiling(const wv_batch *b, int *tb_depth, int *rows_per_tile, int *tile_cols)
{
    if (!b) return WV_EINVAL;
    const Engine *e = b->e;
    if (tb_depth) *tb_depth = depth(e);
    if (rows_per_tile) *rows_per_tile = rows(e);
    if (tile_cols) *tile_cols = wv_tile_cols(depth(e));
    return WV_OK;
}

}  // extern "C"

// ---- autotuner ----------------------------------------------------------------
// Times every (depth, rows per tile) candidate on the current state of every
// grid of the engine and keeps the fastest.  The states are saved to spare
// device buffers first and restored afterwards, so tuning does not advance
// the simulation.
static int engine_autotune(Engine *e, long steps_per_trial, double *best_gcells)
{
    int rc = ready(e);
    if (rc) return rc;
    if (steps_per_trial < 1) return fail(e, WV_EINVAL, "steps_per_trial must be >= 1");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    // save
    std::vector<std::pair<void *, size_t>> what;
    double cells_total = 0;
    for (const Grid &G : e->g) {
        const size_t cells = (size_t)G.p.nx * G.ldy;
        cells_total += (double)G.p.nx * G.p.ny;
        if (e->form == WV_FORM_FP64) {
            for (int k = 0; k < 4; k++) what.push_back({G.u[k], cells * 8});
        } else {
            for (int k = 0; k < 2; k++) {
                what.push_back({G.hi[k], cells * 4});
                what.push_back({G.lo[k], cells * 4});
                what.push_back({G.v[k], cells * 4});
            }
        }
    }
    std::vector<void *> save;
    auto release = [&]() {
        for (void *q : save) (void)hipFree(q);
        save.clear();
    };
    for (auto &w : what) {
        void *q = nullptr;
        if (hipMalloc(&q, w.second) != hipSuccess) {
            release();
            return fail(e, WV_ENOMEM, "no device memory to save the state while tuning");
        }
        save.push_back(q);
        if (hipMemcpy(q, w.first, w.second, hipMemcpyDeviceToDevice) != hipSuccess) {
            release();
            return fail(e, WV_EHIP, "saving the state failed");
        }
    }
    int role0[4];
    memcpy(role0, e->role, sizeof role0);
    const int s0 = e->scur;
    const long sd0 = e->steps_done;
    const int K0 = e->K, L0 = e->L;
    long gstep0 = 0;
    HIPCHK(e, hipMemcpy(&gstep0, e->d_gstep, sizeof(long), hipMemcpyDeviceToHost));
    const int depths[] = {1, 2, 3, 4, 6, 8};
    const int rowsc[] = {8, 12, 16, 24, 32, 64};
    double best = -1;
    int bk = depth(e), bl = rows(e);
    const bool exact = e->form == WV_FORM_FP64 && any_exact(e);
    for (int k : depths) {
        if (exact && k > 1) continue;
        for (int L : rowsc) {
            if ((rc = engine_set_tiling(e, k, L))) break;
            double ms = 0;
            long nl = 0;
            if ((rc = engine_step_timed(e, std::max<long>(1, steps_per_trial / 4), &ms, &nl))) break;  // warm
            if ((rc = engine_step_timed(e, steps_per_trial, &ms, &nl))) break;
            const double gc = cells_total * steps_per_trial / (ms * 1e6);
            if (gc > best) { best = gc; bk = k; bl = L; }
        }
        if (rc) break;
    }
    // restore
    for (size_t q = 0; q < what.size(); q++)
        (void)hipMemcpy(what[q].first, save[q], what[q].second, hipMemcpyDeviceToDevice);
    release();
    (void)hipMemcpy(e->d_gstep, &gstep0, sizeof(long), hipMemcpyHostToDevice);
    memcpy(e->role, role0, sizeof role0);
    e->scur = s0;
    e->steps_done = sd0;
    if (rc) {
        engine_set_tiling(e, K0, L0);
        return rc;
    }
    if ((rc = engine_set_tiling(e, bk, bl))) return rc;
    if (best_gcells) *best_gcells = best;
    return WV_OK;
}

extern "C" int wv_set_neumann(wv_ctx *c, int n, const int *i, const int *j, const int *nbi1, const int *nbj1,
                              const int *nbi2, const int *nbj2, const double *nx, const double *ny)
{
    if (!c) return WV_EINVAL;
    return set_neumann(c->e, 0, n, i, j, nbi1, nbj1, nbi2, nbj2, nx, ny);
}

extern "C" int wv_set_drive(wv_ctx *c, int n, const int *i, const int *j, const double *val, long len)
{
    if (!c) return WV_EINVAL;
    return set_drive(c->e, 0, n, i, j, val, len);
}

extern "C" int wv_set_floor(wv_ctx *c, double vmax)
{
    if (!c) return WV_EINVAL;
    return set_floor(c->e, 0, vmax);
}

extern "C" int wv_set_wells_tables(wv_ctx *c, const short *xy_in, int nwells, const double *xc, const double *yc,
                                   const double *height, double sigma, int use_tanh, double courant,
                                   double courantb, double gamma, const double box[4], double *tcc, double *tc,
                                   double *tgamma)
{
    if (!c) return WV_EINVAL;
    return set_wells_tables(c->e, 0, xy_in, nwells, xc, yc, height, sigma, use_tanh, courant, courantb, gamma, box,
                            tcc, tc, tgamma);
}

extern "C" int wv_batch_set_wells_tables(wv_batch *b, int k, const short *xy_in, int nwells, const double *xc,
                                         const double *yc, const double *height, double sigma, int use_tanh,
                                         double courant, double courantb, double gamma, const double box[4],
                                         double *tcc, double *tc, double *tgamma)
{
    if (!b) return WV_EINVAL;
    return set_wells_tables(b->e, k, xy_in, nwells, xc, yc, height, sigma, use_tanh, courant, courantb, gamma, box,
                            tcc, tc, tgamma);
}

// The billiard mask of init_wave_flat on the GPU (see wavehip.h)
static int build_mask(Engine *e, int domain, const double *box, double lambda, double mu, int ncirc, const double *circ,
                      short *xy)
{
    int kind;
    switch (domain) {
    case WV_D_SINAI: kind = 0; break;
    case WV_D_EXT_ELLIPSE: kind = 1; break;
    case WV_D_CIRCLES: kind = 2; break;
    default: return fail(e, WV_EUNSUPPORTED, "domain %d: D_SINAI, D_EXT_ELLIPSE and D_CIRCLES are provided", domain);
    }
    if (!box || !xy) return fail(e, WV_EINVAL, "box or output is NULL");
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "empty box");
    if (kind == 2 && (ncirc < 0 || (ncirc > 0 && !circ))) return fail(e, WV_EINVAL, "circle list");
    const Grid &G = e->g[0];
    const int nx = G.p.nx, ny = G.p.ny;
    const size_t cells = (size_t)nx * ny;
    HIPCHK(e, hipSetDevice(e->device));
    short *d_out = nullptr;
    double *d_c = nullptr;
    const int nc = kind == 2 ? ncirc : 0;
    hipError_t err = hipMalloc(&d_out, cells * sizeof(short));
    if (err == hipSuccess) err = hipMalloc(&d_c, sizeof(double) * 4 * (size_t)std::max(nc, 1));
    if (err == hipSuccess && nc > 0) err = hipMemcpy(d_c, circ, sizeof(double) * 4 * (size_t)nc, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        WvMask a{};
        a.nx = nx; a.ny = ny; a.kind = kind; a.ncirc = nc; a.i0 = 0; a.rows = nx;
        a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
        a.lambda = lambda; a.mu = mu; a.circ = d_c; a.out = d_out;
        err = (hipError_t)wv_launch_mask(a, e->stream);
    }
    if (err == hipSuccess) err = hipMemcpyAsync(xy, d_out, cells * sizeof(short), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(d_out);
    (void)hipFree(d_c);
    if (err != hipSuccess) return fail(e, WV_EHIP, "mask: %s", hipGetErrorString(err));
    return WV_OK;
}

extern "C" int wv_build_mask(wv_ctx *c, int domain, const double box[4], double lambda, double mu, int ncircles,
                             const double *circles, short *xy_in)
{
    if (!c) return WV_EINVAL;
    return build_mask(c->e, domain, box, lambda, mu, ncircles, circles, xy_in);
}

// build_mask without a device allocation: the circle list and the mask go
// through the library's fixed scratch array (WV_MASK_SCRATCH_BYTES on each
// device), the mask in bands of rows copied out one after the other on the
// context's stream.  Calls that overlap on one device race for the array.
static int build_mask_static(Engine *e, int domain, const double *box, double lambda, double mu, int ncirc,
                             const double *circ, short *xy)
{
    int kind;
    switch (domain) {
    case WV_D_SINAI: kind = 0; break;
    case WV_D_EXT_ELLIPSE: kind = 1; break;
    case WV_D_CIRCLES: kind = 2; break;
    default: return fail(e, WV_EUNSUPPORTED, "domain %d: D_SINAI, D_EXT_ELLIPSE and D_CIRCLES are provided", domain);
    }
    if (!box || !xy) return fail(e, WV_EINVAL, "box or output is NULL");
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "empty box");
    if (kind == 2 && (ncirc < 0 || (ncirc > 0 && !circ))) return fail(e, WV_EINVAL, "circle list");
    const int nc = kind == 2 ? ncirc : 0;
    if (nc > WV_MASK_SCRATCH_MAX_CIRCLES)
        return fail(e, WV_EINVAL, "%d circles: the allocation-free mask builder takes at most %d", nc,
                    WV_MASK_SCRATCH_MAX_CIRCLES);
    const Grid &G = e->g[0];
    const int nx = G.p.nx, ny = G.p.ny;
    // scratch layout: circle list (fixed room), then the row band
    const size_t circ_room = sizeof(double) * 4 * (size_t)WV_MASK_SCRATCH_MAX_CIRCLES;
    const size_t band_room = (size_t)WV_MASK_SCRATCH_BYTES - circ_room;
    const size_t row_bytes = sizeof(short) * (size_t)ny;
    if (row_bytes > band_room)
        return fail(e, WV_EINVAL, "NY = %d: the allocation-free mask builder takes rows of at most %zu cells", ny,
                    band_room / sizeof(short));
    const int band = (int)std::min<size_t>((size_t)nx, band_room / row_bytes);
    HIPCHK(e, hipSetDevice(e->device));
    void *base = nullptr;
    const int se = wv_mask_scratch(e->device, &base);
    if (se) return fail(e, WV_EHIP, "mask scratch: %s", hipGetErrorString((hipError_t)se));
    double *d_c = (double *)base;
    short *d_band = (short *)((unsigned char *)base + circ_room);
    hipError_t err = hipSuccess;
    if (nc > 0) err = hipMemcpyAsync(d_c, circ, sizeof(double) * 4 * (size_t)nc, hipMemcpyHostToDevice, e->stream);
    for (int i0 = 0; i0 < nx && err == hipSuccess; i0 += band) {
        WvMask a{};
        a.nx = nx; a.ny = ny; a.kind = kind; a.ncirc = nc; a.i0 = i0; a.rows = std::min(band, nx - i0);
        a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
        a.lambda = lambda; a.mu = mu; a.circ = d_c; a.out = d_band;
        err = (hipError_t)wv_launch_mask(a, e->stream);
        // (the next band's kernel follows this copy on the same stream)
        if (err == hipSuccess)
            err = hipMemcpyAsync(xy + (size_t)i0 * ny, d_band, row_bytes * a.rows, hipMemcpyDeviceToHost, e->stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return fail(e, WV_EHIP, "mask: %s", hipGetErrorString(err));
    return WV_OK;
}

extern "C" int wv_init_circular_wave(wv_ctx *c, double x, double y, double amp, double variance, double wavelength,
                                     const double box[4])
{
    if (!c) return WV_EINVAL;
    return circular_wave(c->e, 0, 0, 0.0, x, y, amp, variance, wavelength, box);
}

extern "C" int wv_add_circular_wave(wv_ctx *c, double factor, double x, double y, double amp, double variance,
                                    double wavelength, const double box[4])
{
    if (!c) return WV_EINVAL;
    return circular_wave(c->e, 0, 1, factor, x, y, amp, variance, wavelength, box);
}

extern "C" int wv_batch_init_circular_wave(wv_batch *b, int k, double x, double y, double amp, double variance,
                                           double wavelength, const double box[4])
{
    if (!b) return WV_EINVAL;
    return circular_wave(b->e, k, 0, 0.0, x, y, amp, variance, wavelength, box);
}

extern "C" int wv_batch_add_circular_wave(wv_batch *b, int k, double factor, double x, double y, double amp,
                                          double variance, double wavelength, const double box[4])
{
    if (!b) return WV_EINVAL;
    return circular_wave(b->e, k, 1, factor, x, y, amp, variance, wavelength, box);
}

extern "C" int wv_build_mask_static(wv_ctx *c, int domain, const double box[4], double lambda, double mu,
                                    int ncircles, const double *circles, short *xy_in)
{
    if (!c) return WV_EINVAL;
    return build_mask_static(c->e, domain, box, lambda, mu, ncircles, circles, xy_in);
}

// ---- every closed-form domain of xy_in_billiard_single_domain ----------------

// the reference's angle constants (global_pdes.c:8-10)
static const double REF_PI = 3.141592654, REF_DPI = 6.283185307, REF_PID = 1.570796327;

static bool domain_known(int d)
{
    static const int codes[] = {999, 0,   1,   196, 199, 198,  197,  2,   3,   4,   5,   6,   7,   8,   9,
                                10,  11,  12,  13,  15,  16,   17,   18,  19,  191, 195, 1951, 1952, 20,
                                201, 202, 24,  25,  26,  31,   32,   33,  34,  50,  51,  522, 523, 524, 525,
                                61,  62,  63,  64,  67,  68,   69,   73,  74,  76,  79,  791, 795, 796, 90, 91,
                                35,  43,  45,  44,  441, 46,   461,  47,  48,  49,  52,  521, 526, 55,  56,
                                57,  58,  581, 59,  65,  66,   731,  75,  751, 78,  781, 792, 793, 794, 797,
                                798, 93,  934, 935, 936, 937,  94,   942, 947, 98,  981, 982, 983, 99,  991,
                                21,  22,  23,  77,  741};
    for (int c : codes)
        if (c == d) return true;
    return false;
}

// the domains that read the lists other reference functions fill
static bool list_domain_known(int d)
{
    static const int codes[] = {40, 41, 42, 421, 53, 531, 54, 541, 192, 193, 194, 931, 95, 97, 96, 932, 933, 941, 27};
    for (int c : codes)
        if (c == d) return true;
    return false;
}

// the caller's lists (wv_domain_lists; the circle list alone for the
// closed-form entry points)
struct DomLists {
    int nc = 0;
    const double *c = nullptr;
    int nr = 0;
    const double *r = nullptr;
    int nrr = 0;
    const double *rr = nullptr;
    int np = 0;
    const double *pg = nullptr;
    int npl = 0;
    const double *pl = nullptr;
};

// one domain's launch parameters: checks, the host constants (the
// reference's expressions in its order), the angle table and, for
// D_POLYGONS, the polygons with their sides' table (polyg: xc, yc, radius,
// active, nsides, cos(omega/2), offset in trig; runs of n values)
static int domain_setup(Engine *e, const Grid &G, int domain, const double *box, const wv_domain_params *p,
                        const DomLists &Ls, bool with_lists, WvDomain &a, std::vector<double> &trig,
                        std::vector<double> &polyg)
{
    const bool listed = with_lists && list_domain_known(domain);
    if (!domain_known(domain) && !listed)
        return fail(e, WV_EUNSUPPORTED, "domain %d is not among the domains provided", domain);
    const bool circles = domain == 20 || domain == 201 || domain == 202 ||
                         (listed && (domain == 96 || domain == 932 || domain == 933 || domain == 941));
    const int ncirc = Ls.nc;
    if (circles && (ncirc < 0 || (ncirc > 0 && !Ls.c))) return fail(e, WV_EINVAL, "circle list");
    if (listed) {
        const auto bad = [](int n, const double *q) { return n < 0 || n > (1 << 20) || (n > 0 && !q); };
        if (bad(Ls.nc, Ls.c) || bad(Ls.nr, Ls.r) || bad(Ls.nrr, Ls.rr) || bad(Ls.np, Ls.pg) || bad(Ls.npl, Ls.pl))
            return fail(e, WV_EINVAL, "lists: counts 0..1048576 and an array for each nonzero count");
        if ((domain == 41 || domain == 27) && (Ls.npl < 1 || p->mdepth < 0 || p->mdepth > 15))
            return fail(e, WV_EINVAL, "D_VONKOCH: at least one vertex and MDEPTH 0..15");
        if (domain == 42 && (p->npoly < 1 || Ls.npl < p->npoly + 1))
            return fail(e, WV_EINVAL, "D_STAR: NPOLY >= 1 and NPOLY + 1 vertices");
        if (domain == 421 && (p->npoly < 0 || Ls.npl < p->npoly + 4))
            return fail(e, WV_EINVAL, "D_STAR_CHANNEL: NPOLY >= 0 and NPOLY + 4 vertices");
        if ((domain == 931 || domain == 95 || domain == 97) &&
            (p->ngridx < 1 || p->ngridy < 1 || (double)(p->ngridx + 1) * (p->ngridy + 1) > 1e6))
            return fail(e, WV_EINVAL, "NGRIDX %d, NGRIDY %d: both >= 1, at most 1e6 disks", p->ngridx, p->ngridy);
        if (domain == 40)
            for (int k = 0; k < Ls.np; k++) {
                const double ns = Ls.pg[4 * (size_t)Ls.np + k];
                if (!(ns >= 1.0 && ns <= 4096.0)) return fail(e, WV_EINVAL, "polygon %d: 1..4096 sides", k);
            }
    }
    const bool poly = domain == 8 || domain == 32 || domain == 791 || domain == 59 || domain == 792 ||
                      domain == 794 || domain == 797 || domain == 798 || domain == 991;
    if (poly && (p->npoly < 1 || p->npoly > 4096)) return fail(e, WV_EINVAL, "NPOLY %d: 1..4096", p->npoly);
    const bool lattice = domain == 93 || domain == 934 || domain == 935 || domain == 936 || domain == 937 ||
                         domain == 94 || domain == 942 || domain == 947;
    if (lattice && (p->ngridx < 1 || p->ngridy < 1 || (double)(p->ngridx + 2) * (p->ngridy + 1) > 1e6))
        return fail(e, WV_EINVAL, "NGRIDX %d, NGRIDY %d: both >= 1, at most 1e6 disks", p->ngridx, p->ngridy);
    if ((domain == 99 || domain == 991) && (p->ngridx < 1 || p->ngridx > 1000000))
        return fail(e, WV_EINVAL, "NGRIDX %d: 1..1e6", p->ngridx);
    // the loops and integer conversions of the periodic domains: bounded by
    // the box over LAMBDA (or LAMBDA over MU), and int-sized
    const double ext = std::max(std::max(std::fabs(box[0]), std::fabs(box[1])),
                                std::max(std::fabs(box[2]), std::fabs(box[3]))) + 2.0;
    if ((domain == 43 || domain == 45) && (!(p->mu > 0) || !(std::fabs(p->lambda) / p->mu <= 1e6)))
        return fail(e, WV_EINVAL, "D_FRESNEL: MU > 0 and |LAMBDA|/MU <= 1e6");
    if ((domain == 44 || domain == 441 || domain == 48 || domain == 55 || domain == 56 || domain == 57) &&
        (!(p->lambda > 0) || !(ext / p->lambda <= 1e6)))
        return fail(e, WV_EINVAL, "domain %d: LAMBDA > 0 and the box at most 1e6 LAMBDA", domain);
    if ((domain == 46 || domain == 461) && (!(p->lambda > 0) || !(ext / p->lambda < 46000.0)))
        return fail(e, WV_EINVAL, "D_QRD: LAMBDA > 0 and (box/LAMBDA)^2 within an int");
    if (domain == 15 || domain == 17 || domain == 22 || domain == 23) {
        // the digits (int)x1 % MRATIO of MRATIO^MDEPTH must fit an int
        if (p->mratio < 1 || p->mdepth < 0 || p->mdepth > 30 ||
            std::pow((double)p->mratio, (double)p->mdepth) >= 2147483647.0)
            return fail(e, WV_EINVAL, "MRATIO %d, MDEPTH %d: MRATIO >= 1 and MRATIO^MDEPTH < 2^31", p->mratio,
                        p->mdepth);
    }
    if ((domain == 16 || domain == 24 || domain == 25 || domain == 26) && (p->mandellevel < 0))
        return fail(e, WV_EINVAL, "MANDELLEVEL %d", p->mandellevel);
    if ((domain == 16 || domain == 25 || domain == 794) && p->julia_scale == 0.0)
        return fail(e, WV_EINVAL, "JULIA_SCALE 0");
    if (domain == 10) {
        // one term per line of the grating: bounded by the box over LAMBDA
        if (!(p->lambda > 0) || (std::fabs(box[2]) + std::fabs(box[3])) / p->lambda > 1e6)
            return fail(e, WV_EINVAL, "D_GRATING: LAMBDA > 0 and at most 1e6 lines");
    }
    if ((domain == 12 || domain == 13) &&
        (p->ngridx < 0 || p->ngridy < 1 || (double)p->ngridx * (p->ngridy + 1) > 1e6))
        return fail(e, WV_EINVAL, "NGRIDX %d, NGRIDY %d: NGRIDY >= 1, at most 1e6 disks", p->ngridx, p->ngridy);
    const double L = p->lambda, M = p->mu;
    a = WvDomain{};
    trig.clear();
    polyg.clear();
    if (domain == 40 && listed) {  // in_tpolygon, sub_wave.c:380-400, per polygon
        const size_t n = (size_t)Ls.np;
        polyg.assign(7 * n, 0.0);
        for (size_t k = 0; k < n; k++) {
            const int ns = (int)Ls.pg[4 * n + k];
            const double ang = Ls.pg[3 * n + k], omega = REF_DPI / ((double)ns);
            polyg[k] = Ls.pg[k];
            polyg[n + k] = Ls.pg[n + k];
            polyg[2 * n + k] = Ls.pg[2 * n + k];
            polyg[3 * n + k] = (short)Ls.pg[5 * n + k] != 0 ? 1.0 : 0.0;  // the short active flag
            polyg[4 * n + k] = (double)ns;
            polyg[5 * n + k] = std::cos(omega * 0.5);
            polyg[6 * n + k] = (double)trig.size();
            for (int q = 0; q < ns; q++) {
                const double angle = ang * REF_PID + ((double)q + 0.5) * omega;
                trig.push_back(std::cos(angle));
                trig.push_back(std::sin(angle));
            }
        }
    } else if (domain == 8 || domain == 32) {  // sub_wave.c:5812-5817, 6034-6041
        const double omega = REF_DPI / ((double)p->npoly);
        for (int k = 0; k < p->npoly; k++) {
            const double angle = p->apoly * REF_PID + ((double)k + 0.5) * omega;
            trig.push_back(std::cos(angle));
            trig.push_back(std::sin(angle));
        }
        a.c0 = L * std::cos(omega * 0.5);
    } else if (domain == 791) {  // sub_wave.c:6978-6981
        for (int k = 0; k < p->npoly; k++) {
            const double angle = p->apoly * REF_PID + (double)k * REF_DPI / (double)p->npoly;
            trig.push_back(std::cos(angle));
            trig.push_back(std::sin(angle));
        }
    } else if (domain == 792 || domain == 794) {  // sub_wave.c:6995-7000, 7035
        for (int k = 0; k < p->npoly; k++) {
            const double angle = p->apoly * REF_PID + (double)k * REF_DPI / (double)p->npoly;
            trig.push_back(std::cos(angle));
            trig.push_back(std::sin(angle));
        }
        a.c0 = std::cos(-0.5 * REF_PID);  // D_POLYCIRCLES_ANGLED's cb, sb (6995-6997)
        a.c1 = std::sin(-0.5 * REF_PID);
    } else if (domain == 797 || domain == 798 || domain == 991) {  // in_polygon 362-376, 7071-7078
        const double apoly = domain == 991 ? 1.0 - 2.0 / (double)p->npoly : p->apoly;
        const double omega = REF_DPI / ((double)p->npoly);
        for (int k = 0; k < p->npoly; k++) {
            const double angle = apoly * REF_PID + ((double)k + 0.5) * omega;
            trig.push_back(std::cos(angle));
            trig.push_back(std::sin(angle));
        }
        a.c0 = std::cos(omega * 0.5);
    } else if (domain == 59) {  // sub_wave.c:6685-6694
        const double salpha = REF_DPI / (double)p->npoly, h = L * std::tan(REF_PI / (double)p->npoly);
        a.c0 = h < M ? std::sqrt(M * M - h * h) : 0.0;
        for (int i = 0; i < p->npoly; i++) {
            trig.push_back(std::cos((double)i * salpha + p->apoly * REF_PID));
            trig.push_back(std::sin((double)i * salpha + p->apoly * REF_PID));
        }
    } else if (domain == 56) {  // sub_wave.c:6633-6640
        a.c0 = L / (2.0 * std::sqrt(3.0));
        a.c1 = a.c0 * 3.0;
        a.c2 = std::sin(REF_DPI / 3.0);
        trig.push_back(std::sin(-REF_DPI / 3.0));
    } else if (domain == 57) {  // sub_wave.c:6649-6650
        a.c0 = -1.0 / std::sqrt(3.0);
        a.c1 = -2.0 * a.c0;
    } else if (domain == 731) {  // sub_wave.c:6835
        a.c0 = std::cos(REF_PI / 6.0);
    } else if (domain == 78 || domain == 781) {  // sub_wave.c:6936-6941
        a.c0 = std::sqrt(2.0);
        for (int k = 0; k < 8; k++) {
            trig.push_back(std::cos((double)k * 0.5 * REF_PID));
            trig.push_back(std::sin((double)k * 0.5 * REF_PID));
        }
    } else if (domain == 983) {  // sub_wave.c:7400
        a.c0 = std::tan(p->apoly * REF_PID);
    } else if (domain == 751) {  // rc_hyp's one-time constants, sub_wave.c:5444-5472
        const double m = L, d = 2.5, b = 3.0;
        const double f = m * d + b, f1 = f / m, dprime = f1 - d, f2 = m * dprime / (m + 1);
        const double r1 = 2.0 * f / m, r2 = 2.0 * b / (m - 1.0);
        const double e1 = std::sqrt(1.0 + 2.0 * b / (m * m * m * d));
        const double e2 = std::sqrt(1.0 + 2.0 * (m * (2.0 * m - 1.0) + b / d) / ((m - 1) * (m - 1) * (m - 1)));
        const double a1 = f1 / e1, a2 = f2 / e2;
        for (double v : {a1, std::sqrt(a1 * r1), 0.5 * d + a1, a2, std::sqrt(a2 * r2), -0.5 * d + a2})
            trig.push_back(v);
    } else if (domain == 67) {  // sub_wave.c:6751-6752
        a.c0 = std::cos(p->apoly * REF_PID);
        a.c1 = std::sin(p->apoly * REF_PID);
    } else if (domain == 73) {  // sub_wave.c:6793
        a.c0 = std::cos(REF_PI / 6.0);
    } else if (domain == 74 || domain == 741) {  // sub_wave.c:6870, 6882
        a.c0 = std::sqrt(2.0);
    } else if (domain == 33) {  // sub_wave.c:6047
        a.c0 = std::sqrt(L * L - (1.0 - M) * (1.0 - M));
    } else if (domain == 1951 || domain == 1952) {  // sub_wave.c:5991-5996, 6011-6016
        const double y1 = domain == 1951 ? 0.2 : p->wall_width_b;
        const double x1 = 0.5 * L + 0.25 * M - y1 * y1 / M;
        const double x2 = x1 + p->wall_width;
        double r2 = domain == 1951 ? 1.5 * 0.25 * L : p->mu_b;
        r2 = r2 * r2;
        a.c0 = std::sqrt(r2 - y1 * y1) - x2;
        a.c1 = x1;
        a.c2 = r2;
    }
    a.nx = G.p.nx; a.ny = G.p.ny; a.domain = domain; a.ncirc = circles ? ncirc : 0; a.i0 = 0; a.rows = G.p.nx;
    if (listed) {
        a.nrect = Ls.nr; a.nrectrot = Ls.nrr; a.npolyg = domain == 40 ? Ls.np : 0; a.npline = Ls.npl;
    }
    a.npoly = p->npoly; a.mdepth = p->mdepth; a.mratio = p->mratio; a.mandellevel = p->mandellevel;
    a.ngridx = p->ngridx; a.ngridy = p->ngridy;
    a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
    a.lambda = L; a.mu = M; a.mu_b = p->mu_b; a.wall_width = p->wall_width; a.wall_width_b = p->wall_width_b;
    a.julia_scale = p->julia_scale; a.julia_x = p->julia_x; a.julia_y = p->julia_y;
    a.mandellimit = p->mandellimit;
    a.wall_width_asym = p->wall_width_asym; a.wall_width_asym_b = p->wall_width_asym_b;
    a.oscil_ymid = p->oscil_ymid;
    a.michelson_position = p->michelson_position;
    return WV_OK;
}

// xy_in_billiard over grid k: one domain everywhere (nd = 1), or with
// COMPARISON (nd = 2) domain[0] where y > 0 and domain[1] where y <= 0
// (sub_wave.c:7593-7602), each with its own circle list
static int build_domain_masks(Engine *e, int k, int nd, const int *domain, const double *box,
                              const wv_domain_params *p, const DomLists *Ls, bool with_lists, short *xy)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (!box || !p || !xy) return fail(e, WV_EINVAL, "box, parameters or output is NULL");
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "empty box");
    WvDomain a[2];
    // one device block per domain: circles | rects | rectrot | polygons |
    // polyline | angle table, each list only where the domain reads it
    std::vector<double> blob[2];
    int rc;
    for (int q = 0; q < nd; q++) {
        std::vector<double> trig, polyg;
        if ((rc = domain_setup(e, e->g[k], domain[q], box, p, Ls[q], with_lists, a[q], trig, polyg))) return rc;
        a[q].half = nd == 1 ? 0 : q + 1;
        std::vector<double> &b = blob[q];
        b.insert(b.end(), Ls[q].c, Ls[q].c + 4 * (size_t)a[q].ncirc);
        if (a[q].nrect) b.insert(b.end(), Ls[q].r, Ls[q].r + 4 * (size_t)a[q].nrect);
        if (a[q].nrectrot) b.insert(b.end(), Ls[q].rr, Ls[q].rr + 5 * (size_t)a[q].nrectrot);
        b.insert(b.end(), polyg.begin(), polyg.end());
        if (a[q].npline) b.insert(b.end(), Ls[q].pl, Ls[q].pl + 2 * (size_t)a[q].npline);
        b.insert(b.end(), trig.begin(), trig.end());
    }
    const Grid &G = e->g[k];
    const size_t cells = (size_t)G.p.nx * G.p.ny;
    HIPCHK(e, hipSetDevice(e->device));
    short *d_out = nullptr;
    double *d_c[2] = {nullptr, nullptr};
    hipError_t err = hipMalloc(&d_out, cells * sizeof(short));
    for (int q = 0; q < nd && err == hipSuccess; q++) {
        err = hipMalloc(&d_c[q], sizeof(double) * std::max<size_t>(blob[q].size(), 1));
        if (err == hipSuccess && !blob[q].empty())
            err = hipMemcpy(d_c[q], blob[q].data(), sizeof(double) * blob[q].size(), hipMemcpyHostToDevice);
        double *base = d_c[q];
        a[q].circ = base;
        base += 4 * (size_t)a[q].ncirc;
        a[q].rect = base;
        base += 4 * (size_t)a[q].nrect;
        a[q].rectrot = base;
        base += 5 * (size_t)a[q].nrectrot;
        a[q].polyg = base;
        base += 7 * (size_t)a[q].npolyg;
        a[q].pline = base;
        base += 2 * (size_t)a[q].npline;
        a[q].trig = base;
        a[q].out = d_out;
        if (err == hipSuccess) err = (hipError_t)wv_launch_domain(a[q], e->stream);
    }
    if (err == hipSuccess) err = hipMemcpyAsync(xy, d_out, cells * sizeof(short), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(d_out);
    for (double *d : d_c) (void)hipFree(d);
    if (err != hipSuccess) return fail(e, WV_EHIP, "domain mask: %s", hipGetErrorString(err));
    return WV_OK;
}

static DomLists circle_lists(int ncircles, const double *circles)
{
    DomLists L;
    L.nc = ncircles;
    L.c = circles;
    return L;
}

static DomLists from_lists(const wv_domain_lists *l)
{
    DomLists L;
    L.nc = l->ncircles; L.c = l->circles;
    L.nr = l->nrects; L.r = l->rects;
    L.nrr = l->nrectrot; L.rr = l->rectrot;
    L.np = l->npolygons; L.pg = l->polygons;
    L.npl = l->npolyline; L.pl = l->polyline;
    return L;
}

extern "C" int wv_build_domain_mask(wv_ctx *c, int domain, const double box[4], const wv_domain_params *p,
                                    int ncircles, const double *circles, short *xy_in)
{
    if (!c) return WV_EINVAL;
    const DomLists L = circle_lists(ncircles, circles);
    return build_domain_masks(c->e, 0, 1, &domain, box, p, &L, false, xy_in);
}

extern "C" int wv_batch_build_domain_mask(wv_batch *b, int k, int domain, const double box[4],
                                          const wv_domain_params *p, int ncircles, const double *circles,
                                          short *xy_in)
{
    if (!b) return WV_EINVAL;
    const DomLists L = circle_lists(ncircles, circles);
    return build_domain_masks(b->e, k, 1, &domain, box, p, &L, false, xy_in);
}

extern "C" int wv_build_list_domain_mask(wv_ctx *c, int domain, const double box[4], const wv_domain_params *p,
                                         const wv_domain_lists *lists, short *xy_in)
{
    if (!c) return WV_EINVAL;
    if (!lists) return fail(c->e, WV_EINVAL, "lists is NULL");
    const DomLists L = from_lists(lists);
    return build_domain_masks(c->e, 0, 1, &domain, box, p, &L, true, xy_in);
}

extern "C" int wv_batch_build_list_domain_mask(wv_batch *b, int k, int domain, const double box[4],
                                               const wv_domain_params *p, const wv_domain_lists *lists,
                                               short *xy_in)
{
    if (!b) return WV_EINVAL;
    if (!lists) return fail(b->e, WV_EINVAL, "lists is NULL");
    const DomLists L = from_lists(lists);
    return build_domain_masks(b->e, k, 1, &domain, box, p, &L, true, xy_in);
}

extern "C" int wv_build_comparison_mask(wv_ctx *c, int domain, int domain_b, const double box[4],
                                        const wv_domain_params *p, int ncircles, const double *circles,
                                        int ncircles_b, const double *circles_b, short *xy_in)
{
    if (!c) return WV_EINVAL;
    const int d[2] = {domain, domain_b};
    const DomLists L[2] = {circle_lists(ncircles, circles), circle_lists(ncircles_b, circles_b)};
    return build_domain_masks(c->e, 0, 2, d, box, p, L, false, xy_in);
}

// ---- init_ior_2d's per-cell media models -------------------------------------

static bool ior_known(int m)
{
    static const int codes[] = {0, 1, 100, 2, 10, 11, 12, 13, 14, 15, 151, 16, 17, 18, 181, 182, 183, 19, 191};
    for (int c : codes)
        if (c == m) return true;
    return false;
}

static int set_ior_tables(Engine *e, int k, const short *xy, int ior, const double *box, const wv_ior_params *p,
                          double *tcc_out, double *tc_out, double *tg_out)
{
    if (k < 0 || k >= e->n) return fail(e, WV_EINVAL, "grid index %d out of range", k);
    if (ior >= 4 && ior <= 9)
        return fail(e, WV_EUNSUPPORTED, "IOR %d is a Gaussian-well model: wv_set_wells_tables builds it", ior);
    if (ior == 3)
        return fail(e, WV_EUNSUPPORTED, "IOR_EXPLO_LENSING: the reference reads xy_in[i*NY+j] of its NX row pointers");
    if (!ior_known(ior)) return fail(e, WV_EUNSUPPORTED, "IOR %d is not a model of init_ior_2d provided here", ior);
    if (!xy || !box || !p) return fail(e, WV_EINVAL, "mask, box or parameters NULL");
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return fail(e, WV_EINVAL, "empty box");
    if ((ior == 1 || ior == 100) && p->mandellevel < (ior == 100 ? 1 : 0))
        return fail(e, WV_EINVAL, "MANDELLEVEL %d", p->mandellevel);
    if (ior == 183 && p->mdepth < 1) return fail(e, WV_EINVAL, "MDEPTH %d: >= 1", p->mdepth);
    Grid &G = e->g[k];
    const int nx = G.p.nx, ny = G.p.ny;
    const size_t cells = (size_t)nx * ny;
    WvIor a{};
    a.nx = nx; a.ny = ny; a.model = ior; a.mandellevel = p->mandellevel; a.mdepth = p->mdepth;
    a.twospeeds = G.p.twospeeds != 0;
    a.xmin = box[0]; a.xmax = box[1]; a.ymin = box[2]; a.ymax = box[3];
    a.courant = p->courant; a.courantb = p->courantb; a.gamma = p->gamma; a.gammab = p->gammab;
    a.lambda = p->lambda; a.mu = p->mu; a.mandel_ior_scale = p->mandel_ior_scale; a.mandellimit = p->mandellimit;
    // the lens-wall abscissa of IOR_LENS_CONCAVE / _CONVEX_CONCAVE, sub_wave.c:12625, 12655
    if (ior == 12 || ior == 13) {
        a.a0 = p->lambda + 0.1 - std::sqrt(p->lambda * p->lambda - p->mu * p->mu);
        if (ior == 13) a.a0 = a.a0 + 0.6;
    }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    short *d_xy = nullptr;
    double *d_t = nullptr;
    hipError_t err = hipMalloc(&d_xy, cells * sizeof(short));
    if (err == hipSuccess) err = hipMalloc(&d_t, 3 * cells * sizeof(double));
    if (err == hipSuccess) err = hipMemcpy(d_xy, xy, cells * sizeof(short), hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        a.xy = d_xy; a.tcc = d_t; a.tc = d_t + cells; a.tg = d_t + 2 * cells;
        err = (hipError_t)wv_launch_ior(a, e->stream);
    }
    // the tables come back straight into the caller's arrays when all three
    // are given, else into uninitialised staging
    const bool direct = tcc_out && tc_out && tg_out;
    std::unique_ptr<double[]> stage(direct ? nullptr : new (std::nothrow) double[3 * cells]);
    if (!direct && !stage) err = hipErrorOutOfMemory;
    double *h[3] = {direct ? tcc_out : stage.get(), direct ? tc_out : stage.get() + cells,
                    direct ? tg_out : stage.get() + 2 * cells};
    for (int q = 0; q < 3 && err == hipSuccess; q++)
        err = hipMemcpyAsync(h[q], d_t + q * cells, cells * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(d_xy);
    (void)hipFree(d_t);
    if (err != hipSuccess) return fail(e, WV_EHIP, "media tables: %s", hipGetErrorString(err));
    // installed as wv_set_tables installs tables (media classes, masks, tiles)
    const int rc = set_tables(e, k, xy, h[0], h[1], h[2]);
    if (rc) return rc;
    if (!direct) {
        double *o[3] = {tcc_out, tc_out, tg_out};
        for (int q = 0; q < 3; q++)
            if (o[q]) memcpy(o[q], h[q], cells * sizeof(double));
    }
    return WV_OK;
}

extern "C" int wv_set_ior_tables(wv_ctx *c, const short *xy_in, int ior, const double box[4], const wv_ior_params *p,
                                 double *tcc, double *tc, double *tgamma)
{
    if (!c) return WV_EINVAL;
    return set_ior_tables(c->e, 0, xy_in, ior, box, p, tcc, tc, tgamma);
}

extern "C" int wv_batch_set_ior_tables(wv_batch *b, int k, const short *xy_in, int ior, const double box[4],
                                       const wv_ior_params *p, double *tcc, double *tc, double *tgamma)
{
    if (!b) return WV_EINVAL;
    return set_ior_tables(b->e, k, xy_in, ior, box, p, tcc, tc, tgamma);
}

// The well centres of the rotating-well models, sub_wave.c:12390-12481:
// the static lattice (first call of the reference) turned by ior_angle.
extern "C" int wv_rotating_wells(int ngridx, int ngridy, int large, double ior_angle, const double box[4], double *xc,
                                 double *yc, double *sigma)
{
    if (ngridx < 1 || ngridy < 1 || !box || !xc || !yc || !sigma) return WV_EINVAL;
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return WV_EINVAL;
    const double xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3];
    // IOR_PERIODIC_WELLS_ROTATING 12394-12403; _LARGE 12442-12451
    const double dx = large ? (1.5 * xmax - 1.5 * xmin) / (double)ngridx : (xmax - xmin) / (double)ngridx;
    const double x0 = large ? 1.5 * xmin : xmin, y0 = large ? 1.5 * ymin : ymin;
    *sigma = 0.2 * dx * dx;
    const double ca = std::cos(ior_angle), sa = std::sin(ior_angle);
    for (int i = 0; i < ngridx; i++)
        for (int j = 0; j < ngridy; j++) {
            const int n = j * ngridx + i;
            const double dy = large ? (1.5 * ymax - 1.5 * ymin) / (double)ngridy : (ymax - ymin) / (double)ngridy;
            const double xs = x0 + dx * ((double)i + 0.5);
            double ys = y0 + dy * ((double)j + 0.5);
            if (j % 2 == 1) ys += 0.5 * dx;
            xc[n] = xs * ca + ys * sa;
            yc[n] = -xs * sa + ys * ca;
        }
    return WV_OK;
}

// ---- TIFF frames (writetiff, sub_wave.c:62-123) ------------------------------

static uint8_t *le16(uint8_t *p, unsigned v)
{
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    return p + 2;
}

static uint8_t *le32(uint8_t *p, uint32_t v)
{
    for (int b = 0; b < 4; b++) p[b] = (uint8_t)(v >> (8 * b));
    return p + 4;
}

static uint8_t *ifd_entry(uint8_t *p, unsigned tag, unsigned type, uint32_t count, uint32_t value)
{
    return le32(le32(le16(le16(p, tag), type), count), value);
}

extern "C" size_t wv_tiff_bound(int width, int height, const char *description)
{
    if (width < 1 || width > WV_TIFF_MAX_WIDTH || height < 1 || height > 65535 || !description) return 0;
    const size_t strip = (size_t)wv_lzw_stride(3 * width);
    return 8 + (size_t)height * strip + 1 + 2 + 12 * 12 + 4 + 6 + 8 * (size_t)height + strlen(description) + 1;
}

// grow a device buffer to at least want bytes
static int grow(Engine *e, void **q, size_t *cap, size_t want)
{
    if (*cap >= want) return WV_OK;
    if (*q) (void)hipFree(*q);
    *q = nullptr;
    *cap = 0;
    HIPCHK(e, hipMalloc(q, want));
    *cap = want;
    return WV_OK;
}

static int tiff_encode(Engine *e, const unsigned char *rgb, int width, int height, const char *description,
                       int compression, unsigned char *out, size_t cap, size_t *len)
{
    if (!rgb || !description || !out || !len) return fail(e, WV_EINVAL, "NULL argument");
    if (width < 1 || width > WV_TIFF_MAX_WIDTH || height < 1 || height > 65535)
        return fail(e, WV_EINVAL, "frame %dx%d: width 1..%d and height 1..65535 are provided", width, height,
                    WV_TIFF_MAX_WIDTH);
    if (compression != WV_TIFF_NONE && compression != WV_TIFF_LZW)
        return fail(e, WV_EUNSUPPORTED, "compression %d (1 = none and 5 = LZW are provided)", compression);
    const size_t need = wv_tiff_bound(width, height, description);
    if (cap < need) return fail(e, WV_EINVAL, "output holds %zu bytes, %zu needed (wv_tiff_bound)", cap, need);
    const int n = 3 * width;
    std::vector<uint32_t> soff(height), scnt(height);  // by TIFF line
    size_t body = 0;
    if (compression == WV_TIFF_NONE) {
        // rows are stored as they are: a copy, nothing to compute
        for (int r = 0; r < height; r++) {
            memcpy(out + 8 + body, rgb + (size_t)r * n, n);
            soff[height - 1 - r] = (uint32_t)(8 + body);
            scnt[height - 1 - r] = (uint32_t)n;
            body += n;
        }
    } else {
        HIPCHK(e, hipSetDevice(e->device));
        const long stride = wv_lzw_stride(n);
        int rc;
        size_t cap_rows = (size_t)e->tf_cap_rows;
        if ((rc = grow(e, (void **)&e->tf_rgb, &e->tf_cap_rgb, (size_t)height * n))) return rc;
        if ((rc = grow(e, (void **)&e->tf_strips, &e->tf_cap_strips, (size_t)height * stride))) return rc;
        if (cap_rows < (size_t)height) {
            size_t c1 = 0, c2 = 0;
            if (e->tf_lens) (void)hipFree(e->tf_lens);
            if (e->tf_off) (void)hipFree(e->tf_off);
            e->tf_lens = nullptr;
            e->tf_off = nullptr;
            e->tf_cap_rows = 0;
            if ((rc = grow(e, (void **)&e->tf_lens, &c1, sizeof(int) * height))) return rc;
            if ((rc = grow(e, (void **)&e->tf_off, &c2, sizeof(long) * height))) return rc;
            e->tf_cap_rows = height;
        }
        HIPCHK(e, hipMemcpyAsync(e->tf_rgb, rgb, (size_t)height * n, hipMemcpyHostToDevice, e->stream));
        int lr = wv_launch_lzw(e->tf_rgb, n, height, e->tf_strips, stride, e->tf_lens, e->stream);
        if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
        std::vector<int> lens(height);
        std::vector<long> off(height);
        HIPCHK(e, hipMemcpyAsync(lens.data(), e->tf_lens, sizeof(int) * height, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        // strips go to the file in the order they are written: row 0 (the
        // bottom row) first, as line height-1 (sub_wave.c:113-122)
        for (int r = 0; r < height; r++) {
            if (lens[r] < 3 || lens[r] > stride) return fail(e, WV_EHIP, "strip %d coded to %d bytes", r, lens[r]);
            off[r] = (long)body;
            soff[height - 1 - r] = (uint32_t)(8 + body);
            scnt[height - 1 - r] = (uint32_t)lens[r];
            body += (size_t)lens[r];
        }
        if ((rc = grow(e, (void **)&e->tf_body, &e->tf_cap_body, body))) return rc;
        HIPCHK(e, hipMemcpyAsync(e->tf_off, off.data(), sizeof(long) * height, hipMemcpyHostToDevice, e->stream));
        lr = wv_launch_tiff_pack(e->tf_strips, stride, e->tf_lens, e->tf_off, height, e->tf_body, e->stream);
        if (lr) return fail(e, WV_EHIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)lr));
        HIPCHK(e, hipMemcpyAsync(out + 8, e->tf_body, body, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    // header, then the IFD at the next even offset and the values that do not
    // fit in an entry, laid out as libtiff lays out this tag set
    memcpy(out, "II*\0", 4);
    size_t at = 8 + body;
    if (at & 1) out[at++] = 0;
    const uint32_t ifd = (uint32_t)at;
    le32(out + 4, ifd);
    const size_t dlen = strlen(description) + 1;
    const bool one = height == 1;
    // StripByteCounts are LONG for one strip or when the uncompressed strip
    // could code to 65535 bytes or more (LZW allowed ten times its input)
    const bool sbc_long = one || (compression == WV_TIFF_LZW ? n >= 65535 / 10 : n > 65535);
    const size_t sbc_bytes = (sbc_long ? 4 : 2) * (size_t)height;
    uint32_t data = ifd + 2 + 12 * 12 + 4;
    const uint32_t bps_at = data;
    data += 6;
    const uint32_t sbc_at = data;
    if (sbc_bytes > 4) data += (uint32_t)sbc_bytes;
    const uint32_t so_at = data;
    if (!one) data += 4 * (uint32_t)height;
    const uint32_t desc_at = data;
    uint32_t dval = desc_at;
    if (dlen <= 4) {
        uint8_t b[4] = {0, 0, 0, 0};
        memcpy(b, description, dlen);
        dval = b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
    }
    uint8_t *p = le16(out + at, 12);
    p = ifd_entry(p, 256, 3, 1, (uint32_t)width);         // ImageWidth
    p = ifd_entry(p, 257, 3, 1, (uint32_t)height);        // ImageLength
    p = ifd_entry(p, 258, 3, 3, bps_at);                  // BitsPerSample 8,8,8
    p = ifd_entry(p, 259, 3, 1, (uint32_t)compression);   // Compression
    p = ifd_entry(p, 262, 3, 1, 2);                       // Photometric RGB
    p = ifd_entry(p, 270, 2, (uint32_t)dlen, dval);       // ImageDescription
    p = ifd_entry(p, 273, 4, (uint32_t)height, one ? soff[0] : so_at);  // StripOffsets
    p = ifd_entry(p, 274, 3, 1, 4);                       // Orientation BOTLEFT
    p = ifd_entry(p, 277, 3, 1, 3);                       // SamplesPerPixel
    p = ifd_entry(p, 278, 3, 1, 1);                       // RowsPerStrip
    if (sbc_bytes <= 4)                                   // StripByteCounts
        p = ifd_entry(p, 279, sbc_long ? 4 : 3, (uint32_t)height, sbc_long ? scnt[0] : scnt[0] | scnt[1] << 16);
    else
        p = ifd_entry(p, 279, sbc_long ? 4 : 3, (uint32_t)height, sbc_at);
    p = ifd_entry(p, 284, 3, 1, 1);                       // PlanarConfig CONTIG
    p = le32(p, 0);
    p = le16(le16(le16(p, 8), 8), 8);
    if (sbc_bytes > 4)
        for (int l = 0; l < height; l++) p = sbc_long ? le32(p, scnt[l]) : le16(p, scnt[l]);
    if (!one)
        for (int l = 0; l < height; l++) p = le32(p, soff[l]);
    if (dlen > 4) {
        memcpy(p, description, dlen);
        p += dlen;
    }
    *len = (size_t)(p - out);
    return WV_OK;
}

extern "C" int wv_tiff_encode(wv_ctx *c, const unsigned char *rgb, int width, int height, const char *description,
                              int compression, unsigned char *out, size_t cap, size_t *len)
{
    if (!c) return WV_EINVAL;
    return tiff_encode(c->e, rgb, width, height, description, compression, out, cap, len);
}

extern "C" int wv_tiff_write(wv_ctx *c, const char *filename, const char *description, const unsigned char *rgb,
                             int width, int height, int compression)
{
    if (!c) return WV_EINVAL;
    if (!filename) return fail(c->e, WV_EINVAL, "NULL file name");
    const size_t cap = wv_tiff_bound(width, height, description);
    if (!cap) return fail(c->e, WV_EINVAL, "frame %dx%d or description not provided", width, height);
    std::vector<unsigned char> buf(cap);
    size_t len = 0;
    int rc = tiff_encode(c->e, rgb, width, height, description, compression, buf.data(), cap, &len);
    if (rc) return rc;
    FILE *f = fopen(filename, "wb");
    if (!f) return fail(c->e, WV_EIO, "cannot open %s for writing", filename);
    const bool ok = fwrite(buf.data(), 1, len, f) == len;
    if (fclose(f) != 0 || !ok) return fail(c->e, WV_EIO, "short write to %s", filename);
    return WV_OK;
}

extern "C" int wv_set_probes(wv_ctx *c, int n, const int *i, const int *j, long every, long capacity)
{
    if (!c) return WV_EINVAL;
    return set_probes(c->e, 0, n, i, j, every, capacity);
}

extern "C" int wv_get_probes(wv_ctx *c, double *out, long max_samples, long *count, long *lost)
{
    if (!c) return WV_EINVAL;
    return get_probes(c->e, 0, out, max_samples, count, lost);
}

// wave_billiard.c:902-905: the value scaled by 1e16 in binary64 and
// truncated to a long, then "%019ld\n" (19 characters, sign included, zero
// padded)
extern "C" int wv_format_time_series(const double *v, long count, long stride, char *out, long cap, long *written)
{
    if (written) *written = 0;
    if (count < 0 || (count > 0 && (!v || !out || stride < 1))) return WV_EINVAL;
    long pos = 0;
    char line[32];
    for (long s = 0; s < count; s++) {
        const long w = (long)(v[s * stride] * 1.0e16);
        const int m = snprintf(line, sizeof line, "%019ld\n", w);
        if (m <= 0 || pos + m > cap) return WV_EINVAL;
        memcpy(out + pos, line, (size_t)m);
        pos += m;
    }
    if (written) *written = pos;
    return WV_OK;
}

extern "C" int wv_set_laplacian(wv_ctx *c, const int *nneighb, const int *nghb)
{
    if (!c) return WV_EINVAL;
    return set_laplacian(c->e, 0, nneighb, nghb);
}

extern "C" int wv_autotune(wv_ctx *c, long steps_per_trial, double *best_gcells)
{
    if (!c) return WV_EINVAL;
    return engine_autotune(c->e, steps_per_trial, best_gcells);
}

extern "C" int wv_batch_autotune(wv_batch *b, long steps_per_trial, double *best_gcells)
{
    if (!b) return WV_EINVAL;
    return engine_autotune(b->e, steps_per_trial, best_gcells);
}
