// wv_step_fp64.hip -- FP64 stepping kernel, one fused depth K per build
// (compiled once per K with -DWV_K=k).
//
// A workgroup is one wave64 that owns a tile of 64*WV_C columns (halo
// included) and a strip of rows, and streams down the strip one row per
// iteration.  Skewed pipeline: at iteration s, level t (1..K) computes row
// s-2t from the level t-1 rows produced in iterations s-1 (row below: E),
// s-2 (centre) and s-3 (row above: W), so the K levels of one iteration are
// independent of each other and their arithmetic interleaves; new rows are
// committed after all levels have read.  Level-0 rows (u^n, u^{n-1} and the
// active / medium bits) arrive through a 3-deep prefetch ring: row s+3 is
// requested at the end of iteration s.  Window slots rotate with the 3-way
// unrolled loop: at phase P the rows of iterations s-1, s-2, s-3 sit in
// slots (P+2)%3, (P+1)%3 and P.  Level K at row s-2K is final and is written
// out with level K-1 of the same row (the next launch's u^{n-1}).
//
// The outermost columns and rows of a tile are recomputed by the neighbouring
// tiles (overlapped tiling): a tile reads K halo rows above and below its
// strip and H >= K halo columns on each side, and writes only its interior.
// With periodic edges the halo indices wrap, so a halo cell is an exact copy
// of the cell it aliases.  Every cell evaluates the reference expression in
// the reference order whatever K, strip height or tile width, so results are
// bit-identical to one step at a time and to the reference stepper.
#include "wv_device.h"

#ifndef WV_K
#error "compile with -DWV_K=<fused depth>"
#endif

using namespace wv;

namespace {

// Coefficient access of a kernel variant
constexpr int CF_UNIFORM = 0;  // one (tcc, tgamma) pair: values in registers
constexpr int CF_TWO = 1;      // two pairs, selected by the medium bit
constexpr int CF_FIELD = 2;    // per-cell tcc, one tgamma: loaded one iteration ahead
constexpr int CF_FIELDG = 3;   // per-cell tcc and tgamma: loaded one iteration ahead
constexpr int CF_ANY = 4;      // per-cell tables read in place (boundary tiles)

// INTERIOR: the tile's rows and columns never reach the outer box (bulk
// formula only, no index wrapping); otherwise every cell is classified.
// CF: coefficient access (above).  ALLACT (INTERIOR only): every cell the
// tile reads is written by a step, so no active-bit plane is read.  DAMP:
// bulk cells evaluate -KAPPA*x - tgamma*(x-y); without it they use
// leap_undamped(), which the host selects only where that is bit-identical.
// XT: Neumann ghost points, driven cells and FLOOR on boundary tiles.  TB
// (with XT, K = 1): the neighbour-table Laplacian of evolve_wave_half_new on
// boundary tiles of grids that use it.
//
// EJ (edge columns): columns reaching column 0 or NY-1, rows inside the box
// or reaching row 0 or NX-1 of a grid not periodic along i; bulk cells as on
// interior rows, outer-column cells of rows inside the box by the top/bottom-
// row edge formula with the row tests resolved (edge_tb_inside), rows 0 and
// NX-1 by the full edge formula, rows outside the box skipped.
//
// EI (with EJ): the tile's rows reach row 0 or NX-1 of a grid not periodic
// along i (a variant of its own, so that the edge-column loop of the other
// tiles keeps its size).
template <int K, int ST, bool INTERIOR, int CF, bool ALLACT, bool DAMP, bool XT = false, bool TB = false,
          bool EJ = false, bool EI = false>
__device__ __forceinline__ void run_fp64(const WvInst &d, const WvTile t, long gbase, int ia, int ib,
                                         int oc, int od)
{
    constexpr bool PREF = CF == CF_FIELD || CF == CF_FIELDG;
    constexpr int COEF = CF == CF_UNIFORM ? WV_COEF_UNIFORM : (CF == CF_TWO ? WV_COEF_TWO : WV_COEF_FIELD);
    constexpr bool ROWI = INTERIOR || EJ;  // rows inside the outer box
    static_assert(!(INTERIOR && EJ) && !(EJ && (XT || TB)), "edge-column tiles: plain variant only");
    static_assert(INTERIOR || !PREF, "prefetched coefficients are for interior tiles");
    static_assert(ROWI || DAMP, "boundary tiles evaluate the full formula");
    constexpr int C = WV_C;
    constexpr int H = Tile<K>::H;
    constexpr unsigned CMASK = (1u << C) - 1u;
    constexpr bool FAST = INTERIOR;
    static_assert(2 * K * C <= 32, "bit shift register holds 2K rows of C bits");
    static_assert(INTERIOR || !ALLACT, "ALLACT tiles are interior tiles");
    const Par p(d);
    const Src src(d);
    const int nx = p.nx, ny = p.ny, ldy = p.ldy, ldw = p.ldw;
    const bool wrapi = p.bc == 1, wrapj = (p.bc == 1 || p.bc == 3);
    const int jb = t.jb;
    const Cols<FAST> L(jb - H + (int)threadIdx.x * C, jb, (t.flags >> WV_TILE_WIDTH_SHIFT) & 255, ny, wrapj);
    const int ibeg = t.i0;
    const int iend = min(ibeg + (t.flags >> WV_TILE_ROWS_SHIFT), nx);
    const bool has_src = !ROWI && (t.flags & WV_TILE_SRC) && src.any();  // (source tiles are boundary tiles)
    // Neumann ghost points, driven cells and FLOOR (XT: built into the kernel
    // variant for grids that use them; boundary-path tiles only)
    const Xtra xq(d);
    const bool xneu = XT && !INTERIOR && WV_C == 1 && (t.flags & WV_TILE_NEU);
    const bool xdrv = XT && !INTERIOR && (t.flags & WV_TILE_DRV);
    const bool xfloor = XT && !INTERIOR && xq.vmax > 0.0;
    const bool xprb = XT && !INTERIOR && (t.flags & WV_TILE_PRB);
    static_assert(!TB || (XT && K == 1 && ST == 5), "table Laplacians: 5-point, one step per launch");
    const int tblm = TB && !INTERIOR ? d.tbl : 0;  // 1: table Laplacians, 2: BC_LSHAPE
    const bool cmpm = TB && !INTERIOR && d.cmp;    // COMPARISON
    const int jtop = ny / 2 - 1, jbot = ny / 2;
    const double *u0 = d.u[ia], *um1 = d.u[ib];
    const uint64_t *act = d.act, *med = d.med;
    const double *tcp = d.tc, *tccf = d.tccf, *tgf = d.tgf;
    gptr<double> out0 = gm(d.u[oc]), out1 = gm(d.u[od]);

    double U[K][3][C], Up[K][3][C];  // windows: u^t and u^{t-1} per row
    double QU[WV_RING][C], QP[WV_RING][C];  // prefetch ring: u^n, u^{n-1} rows
    unsigned QB[WV_RING], QM[WV_RING];      // prefetched active / medium bit words
    unsigned mreg = 0, medreg = 0;   // bits of rows s-1, s-2, ... (C bits each, newest lowest)
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < C; c++) U[a][b][c] = Up[a][b][c] = 0.0;

    const int s0 = ibeg - K;         // first level-0 row
    const int slast = iend + K - 1;  // last level-0 row the strip needs
    const int send = iend + 2 * K;   // iterations s0 .. send-1

    auto fetch = [&](int row, double(&qu)[C], double(&qp)[C], unsigned &qb, unsigned &qm) {
        const int rm = ROWI ? min(max(row, 0), nx - 1) : map_idx(row, nx, wrapi);
        load_row(u0, ldy, rm, L, qu);
        load_row(um1, ldy, rm, L, qp);
        // unconditional loads: a load under a runtime branch makes the compiler
        // drain every outstanding load (the prefetch ring) where the branch joins
        qb = ALLACT ? ~0u : lane_word(act, ldw, rm, L);
        qm = COEF == WV_COEF_TWO ? lane_word(med, ldw, rm, L) : 0u;
    };
#pragma unroll
    for (int q = 0; q < WV_RING; q++) {
        fetch(min(s0 + q, slast), QU[q], QP[q], QB[q], QM[q]);
        // issue the ring slots in order: the loop waits for slot q with a
        // count of the loads issued after it, and a preheader that issued them
        // in another order would make that count (the minimum over both ways
        // into the loop) drain the ring one slot early
        __builtin_amdgcn_sched_barrier(0);
    }

    // per-cell coefficients: level lv computes row s-2lv at iteration s and
    // row s+1-2lv at the next one, so each level loads its next row's values
    // one iteration ahead into registers of its own
    double TC[K][C], TG[K][C];
    auto fetch_coef = [&](int lv, int row) {
        const int rm = min(max(row, 0), nx - 1);
#pragma unroll
        for (int c = 0; c < C; c++) {
            const size_t k = (size_t)rm * ldy + L.jm[c];
            TC[lv - 1][c] = gc(tccf)[k];
            TG[lv - 1][c] = CF == CF_FIELDG ? gc(tgf)[k] : p.tg0;
        }
    };
    if (PREF) {
#pragma unroll
        for (int lv = 1; lv <= K; lv++) fetch_coef(lv, s0 - 2 * lv);
    }

    // boundary tiles: the tc value of each level's next row, loaded one
    // iteration ahead (a load inside the edge branch would make the level wait
    // for it, and for every load in flight, at the branch join)
    constexpr bool PTV = !INTERIOR;
    double TV[K][C];
    auto fetch_tv = [&](int lv, int row) {
        const int rm = map_idx(row, nx, wrapi);
#pragma unroll
        for (int c = 0; c < C; c++) TV[lv - 1][c] = gc(tcp)[(size_t)rm * ldy + L.jm[c]];
    };
    if (PTV) {
#pragma unroll
        for (int lv = 1; lv <= K; lv++) fetch_tv(lv, s0 - 2 * lv);
    }
    auto iter = [&](int s, auto phase) {
        constexpr int PR = decltype(phase)::value;  // ring slot of row s
        constexpr int P = PR % 3;
        constexpr int PE = (P + 2) % 3, PC = (P + 1) % 3, PW = P;
        // levels in decreasing order: level lv overwrites its own slot P (row
        // s-3 of level lv) after level lv+1 has read it, and reads level lv-1
        // before level lv-1 is advanced -- no temporaries, no copies
#pragma unroll
        for (int lv = K; lv >= 1; lv--) {
            // level lv is needed at rows [ibeg-K+lv, iend+K-lv) only (wave-uniform),
            // i.e. at iterations s in [ibeg-K+3lv, iend+K+lv): one unsigned compare
            const bool live = (unsigned)(s - s0 - 3 * lv) < (unsigned)(iend - ibeg + 2 * (K - lv));
            // rows outside the box (plain boundary tiles of grids not periodic
            // along i): no edge formula of rows 0 and NX-1 reads them
            const bool inbox = (ROWI && !EI) || XT || TB || wrapi || (unsigned)(s - 2 * lv) < (unsigned)nx;
            if (live && inbox) {
                const int r = s - 2 * lv;
                // rows outside the dependence cone of the strip's output are never
                // used: INTERIOR tiles only clamp them so table reads stay in bounds
                const int rm = ROWI ? min(max(r, 0), nx - 1) : map_idx(r, nx, wrapi);
                // (an edge-column tile may reach row 0 or NX-1 of a grid not
                // periodic along i: the row is wave-uniform)
                const bool rowint = (ROWI && !EI) || (rm >= 1 && rm <= nx - 2);
                const double(&Xe)[C] = U[lv - 1][PE];
                const double(&Xc)[C] = U[lv - 1][PC];
                const double(&Xw)[C] = U[lv - 1][PW];
                const double(&Yc)[C] = Up[lv - 1][PC];
                const int sh = (2 * lv - 1) * C;
                const unsigned ab = ALLACT ? CMASK : (mreg >> sh) & CMASK;
                const unsigned mb = (medreg >> sh) & CMASK;
                const double cn = from_next(Xc[0]), cp = from_prev(Xc[C - 1]);
                double en = 0, ep = 0, wn = 0, wp = 0;
                if (ST == 9 || xneu) {
                    en = from_next(Xe[0]); ep = from_prev(Xe[C - 1]);
                    wn = from_next(Xw[0]); wp = from_prev(Xw[C - 1]);
                }
                double nu[C], np[C];
                bool wr[C], nzc[C] = {};
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int cn1 = c < C - 1 ? c + 1 : 0, cp1 = c > 0 ? c - 1 : 0;
                    const double x = Xc[c], y = Yc[c], E = Xe[c], Wv = Xw[c];
                    const double N = (c < C - 1) ? Xc[cn1] : cn;
                    const double S = (c > 0) ? Xc[cp1] : cp;
                    double cc, g;
                    if (PREF) {
                        cc = TC[lv - 1][c];
                        g = TG[lv - 1][c];
                    } else {
                        coef_fp64<COEF>(p, tccf, tgf, rm, L.jm[c], (mb >> c) & 1u, cc, g);
                    }
                    double v;
                    const bool cmpcol = cmpm && (L.jm[c] == jtop || L.jm[c] == jbot);
                    const bool bulk = rowint && L.colint[c] && !cmpcol;
                    if (bulk) {
                        double delta;
                        if (ST == 9) {
                            const double NE = (c < C - 1) ? Xe[cn1] : en;
                            const double SE = (c > 0) ? Xe[cp1] : ep;
                            const double NW = (c < C - 1) ? Xw[cn1] : wn;
                            const double SW = (c > 0) ? Xw[cp1] : wp;
                            const double s4 = ((E + Wv) + N) + S;
                            const double sd = ((NE + NW) + SE) + SW;
                            delta = ((4.0 * s4 + sd) - 20.0 * x) * W9;
                        } else if (TB && tblm == 2) {
                            // BC_LSHAPE bulk, sub_wave.c:11901-11912: -4x, then the four
                            // table neighbours (E, W, N, S unless the cell lies on a seam)
                            if (!tbl_delta(xq, u0, rm, L.jm[c], x, 4, delta)) {
                                delta = -4.0 * x;
                                delta = delta + E;
                                delta = delta + Wv;
                                delta = delta + N;
                                delta = delta + S;
                            }
                        } else {
                            delta = (((E + Wv) + N) + S) - 4.0 * x;
                        }
                        v = DAMP ? leap(x, y, delta, cc, p.kappa, g) : leap_undamped(x, y, delta, cc);
                    } else {
                        const double tcv = PTV ? TV[lv - 1][c] : gc(tcp)[(size_t)rm * ldy + L.jm[c]];
                        if (TB && tblm) {
                            // an edge formula that uses no Laplacian has no table entry
                            double delta = 0.0;
                            (void)tbl_delta(xq, u0, rm, L.jm[c], x, 0, delta);
                            v = edge_table_fp64(p, rm, L.jm[c], x, y, E, Wv, N, S, delta, cc, g, tcv);
                        } else if (cmpm && (cmpcol || L.jm[c] == 0 || L.jm[c] == ny - 1)) {
                            v = cmp_fp64(p, u0, rm, L.jm[c], x, y, cc, g, tcv);
                        } else {
                            v = !EJ ? edge_fp64(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv)
                                    : (!EI || rowint ? edge_tb_inside(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv)
                                                     : edge_row_box(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv));
                        }
                    }
                    bool w = ALLACT || ((ab >> c) & 1u);
                    if (XT && !INTERIOR) {
                        // wave_billiard.c:379-389: a Neumann point takes the weighted sum of
                        // two neighbours unless it is an edge cell the edge loops write
                        int code = 0;
                        double w1 = 0, w2 = 0;
                        if (xneu && neu_find(xq, rm, L.jm[c], code, w1, w2) && !(w && !bulk)) {
                            double t = w1 * pick9(code & 15, wp, Wv, wn, S, x, N, ep, E, en);
                            t = t + w2 * pick9(code >> 4, wp, Wv, wn, S, x, N, ep, E, en);
                            v = t;
                            w = true;
                        }
                        // wave_billiard.c:397, 585-590: a driven cell takes its prescribed value
                        double dv = 0;
                        if (xdrv && drv_find(xq, rm, L.jm[c], gbase + lv - 1 - xq.dbase, dv)) {
                            v = dv;
                            w = true;
                        }
                        // wave_billiard.c:592-609: FLOOR clamps cells with xy_in != 0
                        nzc[c] = xfloor && ((gc(xq.nz)[(size_t)rm * xq.ldw + (L.jm[c] >> 6)] >> (L.jm[c] & 63)) & 1u);
                        if (nzc[c] && w) {
                            if (v > xq.vmax) v = xq.vmax;
                            else if (v < -xq.vmax) v = -xq.vmax;
                        }
                    }
                    if (has_src) {
                        const long g1 = gbase + lv - 1;
                        if (g1 < src.len) v = add_sources(src, rm, L.jm[c], g1, v);
                    }
                    wr[c] = w;
                    nu[c] = w ? v : x;  // cells the step never writes keep u
                    np[c] = x;
                    // probes: the tile that writes the cell records it (at every level)
                    if (xprb && r >= ibeg && r < iend && L.wcol[c]) {
                        double q = nu[c];
                        if (K == 1 && !w) {
                            // one step per launch (three buffers): the cell keeps what the
                            // output buffer holds, clamped by FLOOR where it applies
                            q = out0[(size_t)r * ldy + L.jm[c]];
                            if (nzc[c]) q = q > xq.vmax ? xq.vmax : (q < -xq.vmax ? -xq.vmax : q);
                        }
                        prb_record(xq, rm, L.jm[c], gbase + lv - 1 - xq.pbase, q);
                    }
                }
                if (lv == K) {
                    // level K at row s-2K is final (the live test puts it inside the strip)
                    gptr<double> o0 = out0 + (size_t)r * ldy + L.j0;
                    gptr<double> o1 = out1 + (size_t)r * ldy + L.j0;
                    if (C == 2 && FAST && (ALLACT || ab == CMASK) && L.wcol[0] && L.wcol[C - 1]) {
                        const dvec2 a = {nu[0], nu[C - 1]};
                        *(gptr<dvec2>)o0 = a;
                        if (K >= 2) {
                            const dvec2 b = {np[0], np[C - 1]};
                            *(gptr<dvec2>)o1 = b;
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < C; c++)
                            if (L.wcol[c] && wr[c]) {
                                o0[c] = nu[c];
                                if (K >= 2) o1[c] = np[c];
                            } else if (K == 1 && L.wcol[c] && nzc[c]) {
                                // FLOOR also clamps what a cell no step writes holds in
                                // the output buffer (one step per launch, three buffers;
                                // with fused launches such cells are +0)
                                double q = o0[c];
                                if (q > xq.vmax) q = xq.vmax;
                                else if (q < -xq.vmax) q = -xq.vmax;
                                o0[c] = q;
                            }
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        U[lv < K ? lv : 0][P][c] = nu[c];
                        Up[lv < K ? lv : 0][P][c] = np[c];
                    }
                }
            }
            // unconditional, like the ring (a load under the branch would
            // make the compiler wait for every load in flight at the join)
            if (PREF) fetch_coef(lv, s + 1 - 2 * lv);
            if (PTV) fetch_tv(lv, s + 1 - 2 * lv);
        }
        // level 0 from the prefetch ring, then request row s+3 into the freed slot
#pragma unroll
        for (int c = 0; c < C; c++) {
            U[0][P][c] = ring_take(QU[PR][c]);
            Up[0][P][c] = ring_take(QP[PR][c]);
        }
        mreg = (mreg << C) | bits_of(QB[PR], L);
        medreg = (medreg << C) | bits_of(QM[PR], L);
        // unconditional (past the last needed row it re-reads that cached row)
        fetch(min(s + WV_RING, slast), QU[PR], QP[PR], QB[PR], QM[PR]);
    };

    // one exit per trip (up to two trailing iterations past send-1 find every
    // level dead and re-read the last row): an exit between the phases would
    // make the compiler copy ring registers at the loop edges
    static_assert(WV_RING == 3 || WV_RING == 6, "ring of 3 or 6 rows");
    for (int s = s0; s < send; s += WV_RING) {
        iter(s, std::integral_constant<int, 0>{});
        iter(s + 1, std::integral_constant<int, 1>{});
        iter(s + 2, std::integral_constant<int, 2>{});
        if (WV_RING == 6) {
            iter(s + 3, std::integral_constant<int, 3 % WV_RING>{});
            iter(s + 4, std::integral_constant<int, 4 % WV_RING>{});
            iter(s + 5, std::integral_constant<int, 5 % WV_RING>{});
        }
    }
}

// Waves per SIMD the register allocation must allow: the rarely taken
// boundary path (per-lane column mapping, edge formulas) would otherwise set
// the register count of the whole kernel.  The uniform-medium kernel at K = 8
// fits three waves (168 registers, two spilled on the boundary path): +16 %
// at 4096^2 over two, which its own allocation (170) would give.
#ifndef WV_OCC_FP64_U8
#define WV_OCC_FP64_U8 3
#endif
// The edge-column variant is built into the uniform-medium 5-point kernel at
// K = 8 only, like the split kernel's.
template <int K, int ST, int CM>
struct EdgeJ {
    static constexpr bool on = wv_edgej_kernel(K, ST, CM);
};
template <int K, int CM>
struct Occ {
#ifdef WV_OCC_FP64
    static constexpr int waves = WV_OCC_FP64;
#else
    static constexpr int waves = K <= 4 ? 3 : (K == 8 && CM == WV_COEF_UNIFORM ? WV_OCC_FP64_U8 : 2);
#endif
};

template <int K, int ST, int CM>
__global__ void __launch_bounds__(64, (Occ<K, CM>::waves))
k_step_fp64(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, const long *gstep,
            long g0, int ia, int ib, int oc, int od)
{
    const WvTile t = tiles[blockIdx.x];
    if (t.inst < 0) return;  // padding of the XCD interleave
    const WvInst &d = I[t.inst];
    const long gbase = (gstep ? *gstep : 0) + g0;
    const bool interior = t.flags & WV_TILE_INTERIOR;
    const bool allact = t.flags & WV_TILE_ALLACT;
    const bool edgej = t.flags & WV_TILE_EDGEJ;
    const bool damp = d.fp64_full;
#define WV_RUN(CF, A, D) run_fp64<K, ST, true, CF, A, D>(d, t, gbase, ia, ib, oc, od)
#define WV_INTERIOR(CF)                            \
    if (allact) {                                  \
        if (damp) WV_RUN(CF, true, true);          \
        else WV_RUN(CF, true, false);              \
    } else {                                       \
        if (damp) WV_RUN(CF, false, true);         \
        else WV_RUN(CF, false, false);             \
    }
    // the modes this kernel is not built for are folded away
    constexpr bool TB = CM == WV_CM_TBL;
    constexpr bool XT = CM == WV_CM_ALLX || TB;
    const int coef = CM >= WV_CM_ALL ? d.coef : CM;
    switch (coef) {
    case WV_COEF_UNIFORM:
        if (interior) { WV_INTERIOR(CF_UNIFORM) }
        else if (EdgeJ<K, ST, CM>::on && edgej) {
            // edge-column tiles that reach row 0 or NX-1 (of a grid not periodic along i)
            const bool edgei = t.i0 < K || t.i0 + (t.flags >> WV_TILE_ROWS_SHIFT) > d.nx - K;
            if (edgei) {
                if (damp) run_fp64<K, ST, false, CF_UNIFORM, false, true, false, false, true, true>(d, t, gbase, ia, ib, oc, od);
                else run_fp64<K, ST, false, CF_UNIFORM, false, false, false, false, true, true>(d, t, gbase, ia, ib, oc, od);
            } else {
                if (damp) run_fp64<K, ST, false, CF_UNIFORM, false, true, false, false, true>(d, t, gbase, ia, ib, oc, od);
                else run_fp64<K, ST, false, CF_UNIFORM, false, false, false, false, true>(d, t, gbase, ia, ib, oc, od);
            }
        } else run_fp64<K, ST, false, CF_UNIFORM, false, true, XT, TB>(d, t, gbase, ia, ib, oc, od);
        break;
    case WV_COEF_TWO:
        if (interior) { WV_INTERIOR(CF_TWO) }
        else if (EdgeJ<K, ST, CM>::on && edgej) {
            const bool edgei = t.i0 < K || t.i0 + (t.flags >> WV_TILE_ROWS_SHIFT) > d.nx - K;
            if (edgei) {
                if (damp) run_fp64<K, ST, false, CF_TWO, false, true, false, false, true, true>(d, t, gbase, ia, ib, oc, od);
                else run_fp64<K, ST, false, CF_TWO, false, false, false, false, true, true>(d, t, gbase, ia, ib, oc, od);
            } else {
                if (damp) run_fp64<K, ST, false, CF_TWO, false, true, false, false, true>(d, t, gbase, ia, ib, oc, od);
                else run_fp64<K, ST, false, CF_TWO, false, false, false, false, true>(d, t, gbase, ia, ib, oc, od);
            }
        } else run_fp64<K, ST, false, CF_TWO, false, true, XT, TB>(d, t, gbase, ia, ib, oc, od);
        break;
    default:
        if (!interior) run_fp64<K, ST, false, CF_ANY, false, true, XT, TB>(d, t, gbase, ia, ib, oc, od);
        else if (d.tgf) {
            // a tgamma field implies damping
            if (allact) WV_RUN(CF_FIELDG, true, true);
            else WV_RUN(CF_FIELDG, false, true);
        } else WV_INTERIOR(CF_FIELD)
        break;
    }
#undef WV_INTERIOR
#undef WV_RUN
}

}  // namespace

#define WV_CAT2(a, b) a##b
#define WV_CAT(a, b) WV_CAT2(a, b)

int WV_CAT(wv_launch_fp64_k, WV_K)(int stencil, int cm, const WvInst *I, const WvTile *Ti, int ni, const WvTile *Tb,
                                    int nb, const long *gs, long g0, int ia, int ib, int oc, int od,
                                    hipStream_t s)
{
    // the two tile lists are contiguous in memory: one launch covers both
    const WvTile *T = ni ? Ti : Tb;
    const int n = ni + nb;
    if (!n) return 0;
    const dim3 grid(n), block(64);
#define WV_GO(ST, CM) hipLaunchKernelGGL((k_step_fp64<WV_K, ST, CM>), grid, block, 0, s, I, T, gs, g0, ia, ib, oc, od)
    // the table variant exists for one step per launch and the 5-point stencil only
#if WV_K == 1
#define WV_TBL(ST)                         \
    if (ST == 5) WV_GO(5, WV_CM_TBL);      \
    else return (int)hipErrorInvalidValue;
#else
#define WV_TBL(ST) return (int)hipErrorInvalidValue;
#endif
#define WV_ST(ST)                                 \
    switch (cm) {                                 \
    case WV_COEF_UNIFORM: WV_GO(ST, WV_COEF_UNIFORM); break; \
    case WV_COEF_TWO: WV_GO(ST, WV_COEF_TWO); break;         \
    case WV_COEF_FIELD: WV_GO(ST, WV_COEF_FIELD); break;     \
    case WV_CM_ALLX: WV_GO(ST, WV_CM_ALLX); break;           \
    case WV_CM_TBL: WV_TBL(ST); break;                       \
    default: WV_GO(ST, WV_CM_ALL); break;                     \
    }
    if (stencil == 9) {
        WV_ST(9)
    } else {
        WV_ST(5)
    }
#undef WV_ST
#undef WV_TBL
#undef WV_GO
    return (int)hipGetLastError();
}
