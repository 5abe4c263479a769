// wv_step_split.hip -- FP32 double-float ("split") stepping kernel, one
// fused depth K per build (compiled once per K with -DWV_K=k).
//
// State per cell: u = hi + lo (two binary32 numbers) and the velocity
// v = u^n - u^{n-1} (one binary32).  The leapfrog update is carried on v,
//   v' = v + tcc*lap(u) - KAPPA*u - tgamma*v,   u' = u + v'
// which is the reference's -y + 2x + tcc*delta - KAPPA*x - tgamma*(x-y)
// rearranged so that no 2x - y cancellation happens in binary32: lap(u) is
// built from neighbour differences of u, each the sum of the hi difference
// and the lo difference (differences of nearby values are exact), tcc is
// applied as a hi + lo pair (for the
// 9-point stencil the pair holds tcc/6, rounded once from binary64, and the
// sum 4*(edge differences) + (diagonal differences) is not scaled: a binary32
// 1/6 would change the wave speed by 3e-8 relative, a phase error that grows
// with every step), and u' is
// formed by adding v' to lo and that sum to hi with a Fast2Sum.  Edge cells
// (outer rows and columns) are evaluated in FP64 with the reference's own
// formulas.
//
// Same tile, skewed pipeline and prefetch ring as the FP64 kernel
// (wv_step_fp64.hip); a window entry is (hi, lo, v) of one row.
#include <hip/hip_cooperative_groups.h>

#include "wv_device.h"

#ifndef WV_K
#error "compile with -DWV_K=<fused depth>"
#endif

using namespace wv;

namespace {

// Coefficient access of a kernel variant
constexpr int CF_UNIFORM = 0;  // one (tcc, tgamma) pair: values in registers
constexpr int CF_TWO = 1;      // two pairs, selected by the medium bit
constexpr int CF_FIELD = 2;    // per-cell tcc_hi, one tgamma: loaded one iteration ahead
constexpr int CF_FIELDG = 3;   // per-cell tcc_hi and tgamma
constexpr int CF_FIELDL = 4;   // per-cell tcc_hi and tcc_lo, one tgamma
constexpr int CF_FIELDLG = 5;  // per-cell tcc_hi, tcc_lo and tgamma
constexpr int CF_ANY = 6;      // per-cell tables read in place (boundary tiles)

// INTERIOR: the tile's rows and columns never reach the outer box (bulk
// formula only, no index wrapping); otherwise every cell is classified.
// CF: coefficient access (above).  ALLACT (INTERIOR only): every cell the
// tile reads is written by a step.  DAMP: the grid has a non-zero KAPPA or
// tgamma (terms with a zero coefficient are not evaluated).
//
// Every neighbour difference is a difference of u = hi + lo, taken as (hi
// difference) + (lo difference): (east / north / diagonal neighbour) minus
// (cell); a cell's west, south and two western diagonal terms are minus the
// difference that neighbour takes towards it.  Interior tiles evaluate each
// difference once: along i it is kept for the next iteration of the same
// level, along j and the diagonals it comes from the adjacent lane.  Boundary
// tiles evaluate the same subtractions directly; the results are the same.
// (Adding the hi and lo differences before the second difference rounds each
// first difference once, at 2^-24 of its size; that error is a fraction
// 2*courant of the velocity's own rounding per step, and it saves two
// additions per cell and level against keeping the two sums apart.)
//
// EJ (edge columns): the tile's columns reach the outer box, and its rows
// either stay inside it or reach row 0 or NX-1 of a grid not periodic along i.
// On rows inside the box every cell takes the interior evaluation and the
// cells of the outer columns (j = 0, NY-1, and the halo lanes mapped onto
// them) then take the FP64 edge formula instead: a bulk cell's neighbours are
// its true neighbours whatever lane holds them, and an edge cell's own value
// is replaced, not blended.  Rows 0 and NX-1 take the edge formula on every
// lane (a branch the whole wave takes), and rows outside the box are skipped
// (no edge formula reads them).
//
// WIDE (interior tiles only): the WV_WG waves of the workgroup compute one
// tile side by side, wave w on columns jb - K + 62w + lane.  Lanes 1..62 of
// a wave are its own; lane 0 holds the column of lane 62 of wave w-1 and
// lane 63 that of lane 1 of wave w+1.  After every level lanes 1 and 62 put
// their new (hi, lo) into LDS, and before the next level reads that row
// lanes 0 and 63 take their neighbours' values from there: the j-halo is
// one column per wave instead of K, and only the workgroup's two outer
// waves carry the K-column dependence cone.  One barrier per row: a level's
// row is written in one iteration and read in the next, from one of three
// buffers (by iteration phase), so a buffer is rewritten two barriers after
// it was read.  The same cells get the same operations in the same order as
// in a one-wave tile.
//
// EI (with EJ): the tile's rows reach row 0 or NX-1 of a grid not periodic
// along i (a variant of its own, so that the edge-column loop of the other
// tiles keeps its size).
template <int K, int ST, bool INTERIOR, int CF, bool ALLACT, bool DAMP, bool EJ = false, bool WIDE = false,
          bool EI = false>
__device__ __forceinline__ void run_split(const WvInst &d, const WvTile t, long gbase, int ia, int oc, lptr2 gh)
{
    constexpr bool ROWI = INTERIOR || EJ;  // rows inside the outer box
    constexpr bool PREF = CF >= CF_FIELD && CF <= CF_FIELDLG;
    constexpr bool PG = CF == CF_FIELDG || CF == CF_FIELDLG, PL = CF == CF_FIELDL || CF == CF_FIELDLG;
    constexpr int COEF = CF == CF_UNIFORM ? WV_COEF_UNIFORM : (CF == CF_TWO ? WV_COEF_TWO : WV_COEF_FIELD);
    constexpr bool SHARE = ROWI && WV_C == 1;
    static_assert(!(INTERIOR && EJ), "edge-column tiles are not interior tiles");
    static_assert(!EJ || WV_C == 1, "edge-column tiles: one cell per lane");
    static_assert(ROWI || !PREF, "prefetched coefficients are for tiles with interior rows");
    static_assert(!WIDE || (INTERIOR && WV_C == 1), "wide tiles: interior, one cell per lane");
    static_assert(!EI || EJ, "edge rows: edge-column tiles only");
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
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int jb = t.jb;
    const Cols<FAST> L(WIDE ? jb - K + 62 * wave + lane : jb - H + lane * C, jb, (t.flags >> WV_TILE_WIDTH_SHIFT) & 255,
                       ny, wrapj, !WIDE || (lane >= 1 && lane <= 62));
    // LDS slots of the exchange: [phase][wave][level-1][side] (float2 hi,
    // lo), side 0 written by lane 1, side 1 by lane 62.  Lane 0 reads side 1
    // of wave w-1 and lane 63 side 0 of wave w+1 (not at the workgroup's outer
    // edges).  Only those lanes touch LDS (writing or reading whole rows costs
    // the wave an eighth of its time).
    constexpr int GL = K > 1 ? K - 1 : 1;
    const bool gwr = lane == 1 || lane == 62;
    const bool grd = (lane == 0 && wave > 0) || (lane == 63 && wave < WV_WG - 1);
    const int gw = wave * GL * 2 + (lane == 62);
    const int gr = grd ? (lane == 0 ? wave - 1 : wave + 1) * GL * 2 + (lane == 0) : 0;
    (void)gwr;
    (void)grd;
    (void)gw;
    (void)gr;
    const int ibeg = t.i0;
    const int iend = min(ibeg + (t.flags >> WV_TILE_ROWS_SHIFT), nx);
    const bool has_src = !ROWI && (t.flags & WV_TILE_SRC) && src.any();  // (source tiles are boundary tiles)
    const uint64_t *act = d.act, *med = d.med;
    const double *tcp = d.tc, *tccf = d.tccf, *tgf = d.tgf;
    const float *tccf_hi = d.tccf_hi, *tccf_lo = d.tccf_lo, *tgf32 = d.tgf32;
    const float kf = (float)p.kappa;
    const float *hin = d.hi[ia], *lin = d.lo[ia], *vin = d.v[ia];
    gptr<float> hout = gm(d.hi[oc]), lout = gm(d.lo[oc]), vout = gm(d.v[oc]);

    float Hw[K][3][C], Lw[K][3][C], Vw[K][3][C];
    float QH[WV_RING][C], QL[WV_RING][C], QV[WV_RING][C];
    unsigned QB[WV_RING], QM[WV_RING];
    unsigned mreg = 0, medreg = 0;
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < C; c++) Hw[a][b][c] = Lw[a][b][c] = Vw[a][b][c] = 0.0f;

    const int s0 = ibeg - K;
    const int slast = iend + K - 1;
    const int send = iend + 2 * K;

    auto fetch = [&](int row, float(&qh)[C], float(&ql)[C], float(&qv)[C], unsigned &qb, unsigned &qm) {
        const int rm = ROWI ? min(max(row, 0), nx - 1) : map_idx(row, nx, wrapi);
        load_row(hin, ldy, rm, L, qh);
        load_row(lin, ldy, rm, L, ql);
        load_row(vin, ldy, rm, L, qv);
        // unconditional loads: a load under a runtime branch makes the compiler
        // drain every outstanding load (the prefetch ring) where the branch joins
        qb = ALLACT ? ~0u : lane_word(act, ldw, rm, L);
        qm = COEF == WV_COEF_TWO ? lane_word(med, ldw, rm, L) : 0u;
    };
#pragma unroll
    for (int q = 0; q < WV_RING; q++) {
        fetch(min(s0 + q, slast), QH[q], QL[q], QV[q], QB[q], QM[q]);
        // issue the ring slots in order: the loop waits for slot q with a
        // count of the loads issued after it, and a preheader that issued them
        // in another order would make that count (the minimum over both ways
        // into the loop) drain the ring one slot early
        __builtin_amdgcn_sched_barrier(0);
    }

    // per-cell coefficients, one iteration ahead per level (see the FP64 kernel)
    float TH[K][C], TL[K][C], TG[K][C];
    auto fetch_coef = [&](int lv, int row) {
        const int rm = min(max(row, 0), nx - 1);
#pragma unroll
        for (int c = 0; c < C; c++) {
            const size_t k = (size_t)rm * ldy + L.jm[c];
            TH[lv - 1][c] = gc(tccf_hi)[k];
            TL[lv - 1][c] = PL ? gc(tccf_lo)[k] : 0.0f;
            TG[lv - 1][c] = PG ? gc(tgf32)[k] : p.tgf0;
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
    // differences kept from the previous iteration of each level (SHARE):
    // row r+1 minus row r, and the two diagonal ones (9-point)
    float DE[K], DNE[K], DSE[K];
#pragma unroll
    for (int a = 0; a < K; a++) DE[a] = DNE[a] = DSE[a] = 0.0f;

    auto iter = [&](int s, auto phase) {
        constexpr int PR = decltype(phase)::value;  // ring slot of row s
        constexpr int P = PR % 3;
        constexpr int PE = (P + 2) % 3, PC = (P + 1) % 3, PW = P;
        // levels in decreasing order: each level overwrites its own slot P in
        // place (see the FP64 kernel)
#pragma unroll
        for (int lv = K; lv >= 1; lv--) {
            // differences of this iteration's rows (east minus centre), also in
            // iterations where the level is not needed: the first needed one
            // takes its west difference from here
            float De = 0, Dne = 0, Dse = 0;
            if (SHARE) {
                const float h0 = Hw[lv - 1][PC][0], l0 = Lw[lv - 1][PC][0];
                const float he0 = Hw[lv - 1][PE][0], le0 = Lw[lv - 1][PE][0];
                De = (he0 - h0) + (le0 - l0);
                if (ST == 9) {
                    Dne = (from_next(he0) - h0) + (from_next(le0) - l0);
                    Dse = (from_prev(he0) - h0) + (from_prev(le0) - l0);
                }
            }
            // level lv is live at iterations s in [ibeg-K+3lv, iend+K+lv)
            const bool live = (unsigned)(s - s0 - 3 * lv) < (unsigned)(iend - ibeg + 2 * (K - lv));
            // rows outside the box (boundary tiles of grids not periodic along i):
            // no edge formula of rows 0 and NX-1 reads them, so they are skipped
            const bool inbox = (ROWI && !EI) || wrapi || (unsigned)(s - 2 * lv) < (unsigned)nx;
            if (live && inbox) {
                const int r = s - 2 * lv;
                const int rm = ROWI ? min(max(r, 0), nx - 1) : map_idx(r, nx, wrapi);
                // (wave-uniform; an edge-column tile may also reach row 0 or NX-1
                // of a grid not periodic along i, whose rows take the edge formula
                // on every lane)
                const bool rowint = (ROWI && !EI) || (rm >= 1 && rm <= nx - 2);
                const float(&he)[C] = Hw[lv - 1][PE];
                const float(&hc)[C] = Hw[lv - 1][PC];
                const float(&hw)[C] = Hw[lv - 1][PW];
                const float(&le)[C] = Lw[lv - 1][PE];
                const float(&lc)[C] = Lw[lv - 1][PC];
                const float(&lw)[C] = Lw[lv - 1][PW];
                const float(&vc)[C] = Vw[lv - 1][PC];
                const int sh = (2 * lv - 1) * C;
                const unsigned ab = ALLACT ? CMASK : (mreg >> sh) & CMASK;
                const unsigned mb = (medreg >> sh) & CMASK;
                const float hcn = from_next(hc[0]), hcp = from_prev(hc[C - 1]);
                const float lcn = from_next(lc[0]), lcp = from_prev(lc[C - 1]);
                float hen = 0, hep = 0, hwn = 0, hwp = 0, len_ = 0, lep = 0, lwn = 0, lwp = 0;
                if (ST == 9) {
                    hen = from_next(he[0]); hep = from_prev(he[C - 1]);
                    hwn = from_next(hw[0]); hwp = from_prev(hw[C - 1]);
                    len_ = from_next(le[0]); lep = from_prev(le[C - 1]);
                    lwn = from_next(lw[0]); lwp = from_prev(lw[C - 1]);
                }
                float nh[C], nl[C], nv[C];
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int cn = c < C - 1 ? c + 1 : 0, cp = c > 0 ? c - 1 : 0;
                    const float hN = (c < C - 1) ? hc[cn] : hcn, hS = (c > 0) ? hc[cp] : hcp;
                    const float lN = (c < C - 1) ? lc[cn] : lcn, lS = (c > 0) ? lc[cp] : lcp;
                    const unsigned mbit = (mb >> c) & 1u;
                    // edge cells in FP64 with the reference formulas
                    auto edge_cell = [&](float &oh, float &ol, float &ov) {
                        const double x = (double)hc[c] + (double)lc[c];
                        const double y = x - (double)vc[c];
                        const double E = (double)he[c] + (double)le[c];
                        const double Wv = (double)hw[c] + (double)lw[c];
                        const double N = (double)hN + (double)lN;
                        const double S = (double)hS + (double)lS;
                        double cc, g;
                        coef_fp64<COEF>(p, tccf, tgf, rm, L.jm[c], mbit, cc, g);
                        const double tcv = PTV ? TV[lv - 1][c] : gc(tcp)[(size_t)rm * ldy + L.jm[c]];
                        double out = !EJ ? edge_fp64(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv)
                                         : (!EI || rowint ? edge_tb_inside(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv)
                                                          : edge_row_box(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv));
                        if (has_src) {
                            const long g1 = gbase + lv - 1;
                            if (g1 < src.len) out = add_sources(src, rm, L.jm[c], g1, out);
                        }
                        oh = (float)out;
                        ol = (float)(out - (double)oh);
                        ov = (float)(out - x);
                    };
                    float oh, ol, ov;
                    if (SHARE) {
                        // h0 - hw is the east difference of row r-1 (previous iteration),
                        // h0 - hS the north difference of lane j-1
                        const float h0 = hc[c], l0 = lc[c];
                        const float Dn = (hN - h0) + (lN - l0);
                        const float dsum = (De - DE[lv - 1]) + (Dn - from_prev(Dn));
                        float lap;
                        if (ST == 9) {
                            // u0 - uNW is the SE difference of lane j+1 in row r-1,
                            // u0 - uSW the NE difference of lane j-1 in row r-1
                            const float dd = (Dne - from_next(DSE[lv - 1])) + (Dse - from_prev(DNE[lv - 1]));
                            lap = 4.0f * dsum + dd;
                        } else {
                            lap = dsum;
                        }
                        float ch, cl, gf;
                        if (PREF) {
                            ch = TH[lv - 1][c];
                            cl = TL[lv - 1][c];
                            gf = TG[lv - 1][c];
                        } else if (CF == CF_UNIFORM) {
                            ch = p.tcch0;
                            cl = p.tccl0;
                            gf = p.tgf0;
                        } else {
                            ch = mbit ? p.tcch1 : p.tcch0;
                            cl = mbit ? p.tccl1 : p.tccl0;
                            gf = mbit ? p.tgf1 : p.tgf0;
                        }
                        const float v0 = vc[c];
                        float vn = v0 + __fmaf_rn(ch, lap, cl * lap);
                        if (DAMP) {
                            if (kf != 0.0f) vn = vn - kf * h0;
                            vn = (gf != 0.0f) ? vn - gf * v0 : vn;
                        }
                        if (has_src) {
                            const long g1 = gbase + lv - 1;
                            if (g1 < src.len) vn = add_sources(src, rm, L.jm[c], g1, vn);
                        }
                        dd_add_f(h0, l0, vn, oh, ol);
                        ov = vn;
                        // (rows 0 and NX-1: every lane, the interior values discarded)
                        if (EJ && (!rowint || !L.colint[c])) edge_cell(oh, ol, ov);
                    } else if (rowint && L.colint[c]) {
                        const float h0 = hc[c], l0 = lc[c];
                        // differences of u = hi + lo: (hi difference) + (lo difference)
                        const float dE = (he[c] - h0) + (le[c] - l0), dW = (h0 - hw[c]) + (l0 - lw[c]);
                        const float dN = (hN - h0) + (lN - l0), dS = (h0 - hS) + (l0 - lS);
                        const float dsum = (dE - dW) + (dN - dS);
                        float lap;
                        if (ST == 9) {
                            const float hNE = (c < C - 1) ? he[cn] : hen, hSE = (c > 0) ? he[cp] : hep;
                            const float hNW = (c < C - 1) ? hw[cn] : hwn, hSW = (c > 0) ? hw[cp] : hwp;
                            const float lNE = (c < C - 1) ? le[cn] : len_, lSE = (c > 0) ? le[cp] : lep;
                            const float lNW = (c < C - 1) ? lw[cn] : lwn, lSW = (c > 0) ? lw[cp] : lwp;
                            const float dNE = (hNE - h0) + (lNE - l0), dNW = (h0 - hNW) + (l0 - lNW);
                            const float dSE = (hSE - h0) + (lSE - l0), dSW = (h0 - hSW) + (l0 - lSW);
                            lap = 4.0f * dsum + ((dNE - dNW) + (dSE - dSW));
                        } else {
                            lap = dsum;
                        }
                        float ch, cl, gf;
                        if (PREF) {
                            ch = TH[lv - 1][c];
                            cl = TL[lv - 1][c];
                            gf = TG[lv - 1][c];
                        } else if (COEF == WV_COEF_FIELD) {
                            const size_t k = (size_t)rm * ldy + L.jm[c];
                            ch = gc(tccf_hi)[k];
                            cl = tccf_lo ? gc(tccf_lo)[k] : 0.0f;
                            gf = tgf32 ? gc(tgf32)[k] : p.tgf0;
                        } else if (COEF == WV_COEF_UNIFORM) {
                            ch = p.tcch0;
                            cl = p.tccl0;
                            gf = p.tgf0;
                        } else {
                            ch = mbit ? p.tcch1 : p.tcch0;
                            cl = mbit ? p.tccl1 : p.tccl0;
                            gf = mbit ? p.tgf1 : p.tgf0;
                        }
                        const float v0 = vc[c];
                        float vn = v0 + __fmaf_rn(ch, lap, cl * lap);
                        if (DAMP) {  // a term whose coefficient is zero is not evaluated
                            if (kf != 0.0f) vn = vn - kf * h0;
                            vn = (gf != 0.0f) ? vn - gf * v0 : vn;
                        }
                        if (has_src) {
                            const long g1 = gbase + lv - 1;
                            if (g1 < src.len) vn = add_sources(src, rm, L.jm[c], g1, vn);
                        }
                        dd_add_f(h0, l0, vn, oh, ol);
                        ov = vn;
                    } else {
                        edge_cell(oh, ol, ov);
                    }
                    const bool live = ALLACT || ((ab >> c) & 1u);  // cells the step never writes keep their state
                    nh[c] = live ? oh : hc[c];
                    nl[c] = live ? ol : lc[c];
                    nv[c] = live ? ov : vc[c];
                }
                if (lv == K) {
                    const size_t ro = (size_t)r * ldy + L.j0;
                    if (C == 2 && FAST && (ALLACT || ab == CMASK) && L.wcol[0] && L.wcol[C - 1]) {
                        const fvec2 a = {nh[0], nh[C - 1]}, b = {nl[0], nl[C - 1]}, e = {nv[0], nv[C - 1]};
                        *(gptr<fvec2>)(hout + ro) = a;
                        *(gptr<fvec2>)(lout + ro) = b;
                        *(gptr<fvec2>)(vout + ro) = e;
                    } else {
#pragma unroll
                        for (int c = 0; c < C; c++)
                            if (L.wcol[c] && ((ab >> c) & 1u)) {
                                hout[ro + c] = nh[c];
                                lout[ro + c] = nl[c];
                                vout[ro + c] = nv[c];
                            }
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        Hw[lv < K ? lv : 0][P][c] = nh[c];
                        Lw[lv < K ? lv : 0][P][c] = nl[c];
                        Vw[lv < K ? lv : 0][P][c] = nv[c];
                    }
                }
            }
            if (SHARE) {
                DE[lv - 1] = De;
                if (ST == 9) {
                    DNE[lv - 1] = Dne;
                    DSE[lv - 1] = Dse;
                }
            }
            // unconditional, like the ring
            if (PREF) fetch_coef(lv, s + 1 - 2 * lv);
            if (PTV) fetch_tv(lv, s + 1 - 2 * lv);
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            Hw[0][P][c] = ring_take(QH[PR][c]);
            Lw[0][P][c] = ring_take(QL[PR][c]);
            Vw[0][P][c] = ring_take(QV[PR][c]);
        }
        mreg = (mreg << C) | bits_of(QB[PR], L);
        medreg = (medreg << C) | bits_of(QM[PR], L);
        // unconditional (past the last needed row it re-reads that cached row)
        fetch(min(s + WV_RING, slast), QH[PR], QL[PR], QV[PR], QB[PR], QM[PR]);
        if (WIDE) {
            // The rows of levels 1..K-1 this iteration computed (slot P): lanes
            // 1 and 62 publish theirs, and after the barrier lanes 0 and 63 take
            // their neighbours'.  A level's row is first read with its ghost
            // lanes in the next iteration (9-point diagonals, slot PE) or the one
            // after (slot PC), so the loads complete in the shadow of the next
            // row's work.  (All levels together, once per row: exec-masked
            // accesses next to the levels' liveness branches make the compiler
            // restructure the whole loop.  Every level writes its slot, live or
            // not: one that was not live left the slot as it was, and nobody
            // reads it where it is not needed.)
            if (gwr) {
#pragma unroll
                for (int lv = 1; lv < K; lv++) gh[gw + (P * WV_WG * GL + lv - 1) * 2] = fvec2{Hw[lv][P][0], Lw[lv][P][0]};
            }
            // no wait for the global loads in flight
            lds_barrier();
            if (grd) {
#pragma unroll
                for (int lv = 1; lv < K; lv++) {
                    const fvec2 g = gh[gr + (P * WV_WG * GL + lv - 1) * 2];
                    Hw[lv][P][0] = g.x;
                    Lw[lv][P][0] = g.y;
                }
            }
        }
    };

    // one exit per trip (up to two trailing iterations past send-1 find every
    // level dead and re-read the last row): an exit between the phases would
    // make the compiler copy ring registers at the loop edges
    static_assert(WV_RING == 3 || WV_RING == 6, "ring of 3 or 6 rows");
    static_assert(!WIDE || WV_RING == 3, "wide tiles: the exchange buffers follow the 3-row ring");
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
// the register count of the whole kernel.
// The kernels with the edge-column variant at K = 8 (5-point, one or two
// media) are held to three waves (168 registers): the variant alone would
// take the two-medium kernel to 177 and two waves.
// The 9-point kernels also run three waves at K = 8 (8192^2 smooth medium:
// +7 % over two, with some spills on paths that medium does not take).
template <int K, int ST, int CM>
struct EdgeJ {
    static constexpr bool on = wv_edgej_kernel(K, ST, CM);
};
template <int K, int ST, int CM>
struct Occ {
#ifdef WV_OCC_SPLIT
    static constexpr int waves = WV_OCC_SPLIT;
#else
    static constexpr int waves =
        K <= 4 ? 3 : (((K == 8 && ST == 5 && (CM == WV_COEF_UNIFORM || CM == WV_COEF_TWO)) || ST == 9) ? 3 : 2);
#endif
};

// LDS of a workgroup: the exchange rows of a wide tile (see run_split)
template <int K>
struct GhostLds {
    static constexpr int n = 3 * WV_WG * (K > 1 ? K - 1 : 1) * 2;
};

// Tile entry q of the table, computed by this wave (with the other waves of
// its workgroup if the tile is wide); gbase is the amplitude index of the
// first step.  q is wave-uniform.
template <int K, int ST, int CM>
__device__ __forceinline__ void step_tile(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, int q,
                                          long gbase, int ia, int oc, lptr2 lg)
{
    const WvTile t = tiles[q];
    if (t.inst < 0) return;  // padding (one-wave tiles only: a wide tile fills its workgroup)
    const WvInst &d = I[t.inst];
    const bool interior = t.flags & WV_TILE_INTERIOR;
    const bool allact = t.flags & WV_TILE_ALLACT;
    const bool edgej = t.flags & WV_TILE_EDGEJ;
    const bool wide = t.flags & WV_TILE_WIDE;
    const bool damp = d.damped;
#define WV_RUNW(CF, A, D, W) run_split<K, ST, true, CF, A, D, false, W>(d, t, gbase, ia, oc, lg)
#define WV_RUN(CF, A, D)                  \
    if (wide) WV_RUNW(CF, A, D, WV_C == 1);    \
    else WV_RUNW(CF, A, D, false);
    // edge-column tiles that reach row 0 or NX-1 (of a grid not periodic along i)
    const bool edgei = t.i0 < K || t.i0 + (t.flags >> WV_TILE_ROWS_SHIFT) > d.nx - K;
#define WV_EDGEJ1(CF, EI)                                                                      \
    if (damp) run_split<K, ST, false, CF, false, true, true, false, EI>(d, t, gbase, ia, oc, lg); \
    else run_split<K, ST, false, CF, false, false, true, false, EI>(d, t, gbase, ia, oc, lg);
#define WV_EDGEJ(CF)               \
    if (edgei) { WV_EDGEJ1(CF, true) } \
    else { WV_EDGEJ1(CF, false) }
#define WV_INTERIOR(CF)                            \
    if (allact) {                                  \
        if (damp) { WV_RUN(CF, true, true) }       \
        else { WV_RUN(CF, true, false) }           \
    } else {                                       \
        if (damp) { WV_RUN(CF, false, true) }      \
        else { WV_RUN(CF, false, false) }          \
    }
#define WV_DAMPED(CF)                              \
    if (allact) { WV_RUN(CF, true, true) }         \
    else { WV_RUN(CF, false, true) }
    // the modes this kernel is not built for are folded away
    const int coef = CM == WV_CM_ALL ? d.coef : CM;
    switch (coef) {
    case WV_COEF_UNIFORM:
        if (interior) { WV_INTERIOR(CF_UNIFORM) }
        else if (EdgeJ<K, ST, CM>::on && edgej) { WV_EDGEJ(CF_UNIFORM) }
        else run_split<K, ST, false, CF_UNIFORM, false, true>(d, t, gbase, ia, oc, lg);
        break;
    case WV_COEF_TWO:
        if (interior) { WV_INTERIOR(CF_TWO) }
        else if (EdgeJ<K, ST, CM>::on && edgej) { WV_EDGEJ(CF_TWO) }
        else run_split<K, ST, false, CF_TWO, false, true>(d, t, gbase, ia, oc, lg);
        break;
    default:
        if (!interior) run_split<K, ST, false, CF_ANY, false, true>(d, t, gbase, ia, oc, lg);
        else if (d.tccf_lo) {
            if (d.tgf32) { WV_DAMPED(CF_FIELDLG) }  // a tgamma field implies damping
            else WV_INTERIOR(CF_FIELDL)
        } else {
            if (d.tgf32) { WV_DAMPED(CF_FIELDG) }
            else WV_INTERIOR(CF_FIELD)
        }
        break;
    }
#undef WV_EDGEJ
#undef WV_EDGEJ1
#undef WV_DAMPED
#undef WV_INTERIOR
#undef WV_RUN
#undef WV_RUNW
}

template <int K, int ST, int CM>
__global__ void __launch_bounds__(64 * WV_WG, (Occ<K, ST, CM>::waves))
k_step_split(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, const long *gstep,
            long g0, int ia, int oc)
{
    __shared__ fvec2 gh[GhostLds<K>::n];
    // (the wave index through a scalar register: the tile and everything
    // derived from it stay wave-uniform)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    step_tile<K, ST, CM>(I, tiles, blockIdx.x * WV_WG + wave, (gstep ? *gstep : 0) + g0, ia, oc, (lptr2)gh);
}

// Persistent form: `epochs` launches of depth K over the same tile table (of
// one-wave tiles) in one cooperative launch of workgroups that all stay
// resident.  Epoch e reads set (ia0 + e) & 1 and writes the other.  The table
// lists XCD x's workgroups at positions 8j + x, longest first; a wave on XCD
// x (its hardware XCC id) takes their tile entries one at a time, in that
// order, from the counter ctr[8e + x] (zero before the launch), then helps
// with the lists of the other XCDs in turn (so every entry is taken whatever
// the partition mode).  Waves take tiles independently, as the hardware would
// start them: a workgroup moving from one table workgroup to the next as a
// whole would keep three waves waiting for the longest of four tiles.  A
// grid-wide barrier (release / acquire at device scope) separates the
// epochs.  Every wave runs every epoch to its end: each counter passes the
// length of its list once, and then every wave reads a value past it.
template <int K, int ST, int CM>
__global__ void __launch_bounds__(64 * WV_WG, (Occ<K, ST, CM>::waves))
k_step_split_persist(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, int ngroups,
                     const long *gstep, long g0, int epochs, int ia0, int *ctr)
{
    __shared__ fvec2 gh[GhostLds<K>::n];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const long gb = (gstep ? *gstep : 0) + g0;
    unsigned xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    const int x0 = (int)(xcc & 7);
    const bool lead = (threadIdx.x & 63) == 0;
    for (int e = 0; e < epochs; e++) {
        const int ia = (ia0 + e) & 1;
        for (int c = 0; c < 8; c++) {
            const int x = (x0 + c) & 7;
            const int ne = (ngroups - x + 7) / 8 * WV_WG;  // entries of table workgroups 8j + x
            for (;;) {
                int m = 0;
                if (lead) m = atomicAdd(&ctr[(8 * e + x) * WV_PCTR_PAD], 1);
                m = __builtin_amdgcn_readfirstlane(m);
                if (m >= ne) break;
                step_tile<K, ST, CM>(I, tiles, (8 * (m / WV_WG) + x) * WV_WG + m % WV_WG, gb + (long)e * K, ia, ia ^ 1,
                                     (lptr2)gh);
            }
        }
        grid.sync();
    }
}

__global__ void k_zero_int(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

#define WV_CAT2(a, b) a##b
#define WV_CAT(a, b) WV_CAT2(a, b)

int WV_CAT(wv_launch_split_k, WV_K)(int stencil, int cm, const WvInst *I, const WvTile *Ti, int ni, const WvTile *Tb,
                                     int nb, const long *gs, long g0, int ia, int oc, hipStream_t s)
{
    // the two tile lists are contiguous in memory: one launch covers both;
    // ni and nb count workgroups (WV_WG tile entries each)
    const WvTile *T = ni ? Ti : Tb;
    const int n = ni + nb;
    if (!n) return 0;
    const dim3 grid(n), block(64 * WV_WG);
#define WV_GO(ST, CM) hipLaunchKernelGGL((k_step_split<WV_K, ST, CM>), grid, block, 0, s, I, T, gs, g0, ia, oc)
#define WV_ST(ST)                                 \
    switch (cm) {                                 \
    case WV_COEF_UNIFORM: WV_GO(ST, WV_COEF_UNIFORM); break; \
    case WV_COEF_TWO: WV_GO(ST, WV_COEF_TWO); break;         \
    case WV_COEF_FIELD: WV_GO(ST, WV_COEF_FIELD); break;     \
    default: WV_GO(ST, WV_CM_ALL); break;                     \
    }
    if (stencil == 9) {
        WV_ST(9)
    } else {
        WV_ST(5)
    }
#undef WV_ST
#undef WV_GO
    return (int)hipGetLastError();
}

// Persistent launch of `epochs` launches of depth WV_K over the tile table T
// of ng workgroups: as many workgroups as stay resident together (a multiple
// of 8), launched cooperatively, after ctr[0 .. 8 epochs) is zeroed.
// *grid_out: the workgroups launched.
int WV_CAT(wv_launch_split_persist_k, WV_K)(int stencil, int cm, const WvInst *I, const WvTile *T, int ng,
                                             const long *gs, long g0, int epochs, int ia0, int *ctr,
                                             hipStream_t s, int *grid_out)
{
    if (!ng || epochs <= 0) return 0;
    const void *fn = nullptr;
#define WV_FN(ST, CM) fn = (const void *)k_step_split_persist<WV_K, ST, CM>
#define WV_ST(ST)                                 \
    switch (cm) {                                 \
    case WV_COEF_UNIFORM: WV_FN(ST, WV_COEF_UNIFORM); break; \
    case WV_COEF_TWO: WV_FN(ST, WV_COEF_TWO); break;         \
    case WV_COEF_FIELD: WV_FN(ST, WV_COEF_FIELD); break;     \
    default: WV_FN(ST, WV_CM_ALL); break;                     \
    }
    if (stencil == 9) {
        WV_ST(9)
    } else {
        WV_ST(5)
    }
#undef WV_ST
#undef WV_FN
    int dev = 0, cus = 0, per = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) || (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) ||
        (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * WV_WG, 0)))
        return (int)e;
    int g = per * cus;
    g -= g % 8;
    if (g < 8) return (int)hipErrorInvalidConfiguration;
    const int nc = 8 * epochs * WV_PCTR_PAD;
    hipLaunchKernelGGL(k_zero_int, dim3((nc + 255) / 256), dim3(256), 0, s, ctr, nc);
    if ((e = hipGetLastError())) return (int)e;
    void *args[] = {(void *)&I, (void *)&T, (void *)&ng, (void *)&gs, (void *)&g0, (void *)&epochs, (void *)&ia0,
                    (void *)&ctr};
    e = hipLaunchCooperativeKernel(fn, dim3(g), dim3(64 * WV_WG), args, 0, s);
    if (grid_out) *grid_out = g;
    return (int)e;
}
