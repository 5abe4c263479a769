// wv_kernels.hip -- CDNA4 (gfx950) kernels of the wave stepper.
//
// One kernel family per storage form, templated on the fused depth K (time
// steps per launch) and the stencil (5 or 9 point).  A workgroup is one
// wave64 that owns a tile of 64*WV_C columns (halo included) and a strip of
// rows, and streams down the strip one row per iteration:
//
//   iteration s loads row s of u^n (and u^{n-1}), then computes time level t
//   at row s-t for t = 1..K from a 3-row register window of level t-1.
//   Level K at row s-K is final and is written out, together with level K-1
//   of the same row (the next launch's u^{n-1}).
//
// Neighbours along i come from the register windows, neighbours along j from
// the adjacent lane (cross-lane shuffle) or the lane's own other column.  The
// outermost columns and rows of a tile are recomputed by the neighbouring
// tiles (overlapped tiling): a tile reads K halo rows above and below its
// strip and H >= K halo columns on each side, and writes only its interior.
// With periodic edges the halo indices wrap, so a halo cell is an exact copy
// of the cell it aliases.  Every cell evaluates the same expression in the
// same order whatever K, strip height or tile width, so FP64 results are
// bit-identical to one step at a time and to the reference.
//
// Compile with -ffp-contract=off: the reference's arithmetic has no FMA.
#include "wv_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double W9 = 1.0 / 6.0;  // 9-point weight: ((4*s4 + sd) - 20x) * W9

__device__ __forceinline__ int map_idx(int r, int n, bool wrap)
{
    if (wrap) {
        r %= n;
        return r < 0 ? r + n : r;
    }
    return r < 0 ? 0 : (r >= n ? n - 1 : r);
}

// value held by lane+1 / lane-1 of the wave (edge lanes get their own value)
__device__ __forceinline__ double from_next(double v) { return __shfl_down(v, 1, 64); }
__device__ __forceinline__ double from_prev(double v) { return __shfl_up(v, 1, 64); }
__device__ __forceinline__ float from_next(float v) { return __shfl_down(v, 1, 64); }
__device__ __forceinline__ float from_prev(float v) { return __shfl_up(v, 1, 64); }

// wave_billiard.c:373  -y + 2*x + tcc*delta - KAPPA*x - tgamma*(x-y)
__device__ __forceinline__ double leap(double x, double y, double delta, double cc, double kappa,
                                       double g)
{
    double a = -y + 2.0 * x;
    a = a + cc * delta;
    a = a - kappa * x;
    return a - g * (x - y);
}

// wave_billiard.c:418  x - tc*(x - inward) - K*x - G*(x-y)
__device__ __forceinline__ double absorb(double x, double y, double inward, double c, double k,
                                         double g)
{
    double a = x - c * (x - inward);
    a = a - k * x;
    return a - g * (x - y);
}

// Edge cell (rm, jm) of the outer box, every B_COND branch of
// wave_billiard.c:393-583 and wave_3d.c:649-871.  E/Wv: rows rm+1 / rm-1,
// N/S: columns jm+1 / jm-1 (already wrapped where the BC wraps).
__device__ double edge_fp64(const WvInst &d, int rm, int jm, double x, double y, double E,
                            double Wv, double N, double S, double cc, double g, double tcv)
{
    const int nx = d.nx, ny = d.ny, bc = d.bc;
    double delta;
    if (jm == ny - 1 || jm == 0) {
        // top (j = NY-1) and bottom (j = 0) rows, corners included
        const bool top = (jm == ny - 1);
        const double V = top ? S : N;  // inward neighbour along j
        const double O = top ? N : S;  // neighbour across the periodic seam
        if (d.osc_tb && rm < d.tb_shift && rm < nx - 1 && rm > 0) {
            delta = (E + Wv) + -2.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        }
        switch (bc) {
        case 0: {
            const double Ec = (rm == nx - 1) ? x : E;
            const double Wc = (rm == 0) ? x : Wv;
            delta = ((Ec + Wc) + V) - 3.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        }
        case 1:
            delta = (((E + Wv) + V) + O) - 4.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        case 3: {
            if (rm == 0) return absorb(x, y, E, tcv, d.ks, d.gs);
            const double Ec = (rm == nx - 1) ? x : E;
            delta = (((Ec + Wv) + V) + O) - 4.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        }
        default:
            return absorb(x, y, V, tcv, d.kt, d.gt);
        }
    }
    if (rm == 0) {  // left column, 1 <= j <= NY-2
        switch (bc) {
        case 0: delta = ((E + N) + S) - 3.0 * x; return leap(x, y, delta, cc, d.kappa, g);
        case 1: delta = (((E + Wv) + N) + S) - 4.0 * x; return leap(x, y, delta, cc, d.kappa, g);
        default: return absorb(x, y, E, tcv, d.ks, d.gs);
        }
    }
    // right column i = NX-1
    switch (bc) {
    case 0: delta = ((Wv + N) + S) - 3.0 * x; return leap(x, y, delta, cc, d.kappa, g);
    case 1: delta = (((Wv + E) + N) + S) - 4.0 * x; return leap(x, y, delta, cc, d.kappa, g);
    default: return absorb(x, y, Wv, tcv, d.ks, d.gs);
    }
}

// bits of the lane's WV_C cells in a bit plane (columns mapped per cell)
__device__ __forceinline__ unsigned lane_bits(const uint64_t *plane, const WvInst &d, int rm,
                                              int j0, const int (&jm)[WV_C], bool inside)
{
    const uint64_t *row = plane + (size_t)rm * d.ldw;
    if (inside) {  // j0 is a multiple of WV_C, so the WV_C bits share one word
        const uint64_t w = row[j0 >> 6];
        return (unsigned)(w >> (j0 & 63)) & ((1u << WV_C) - 1u);
    }
    unsigned b = 0;
#pragma unroll
    for (int c = 0; c < WV_C; c++) b |= (unsigned)((row[jm[c] >> 6] >> (jm[c] & 63)) & 1u) << c;
    return b;
}

__device__ __forceinline__ void load_row(const double *base, const WvInst &d, int rm, int j0,
                                         const int (&jm)[WV_C], bool inside, double (&out)[WV_C])
{
    const double *row = base + (size_t)rm * d.ldy;
    if (inside) {
        const double2 v = *reinterpret_cast<const double2 *>(row + j0);
        out[0] = v.x;
        out[1] = v.y;
    } else {
#pragma unroll
        for (int c = 0; c < WV_C; c++) out[c] = row[jm[c]];
    }
}

__device__ __forceinline__ void load_row(const float *base, const WvInst &d, int rm, int j0,
                                         const int (&jm)[WV_C], bool inside, float (&out)[WV_C])
{
    const float *row = base + (size_t)rm * d.ldy;
    if (inside) {
        const float2 v = *reinterpret_cast<const float2 *>(row + j0);
        out[0] = v.x;
        out[1] = v.y;
    } else {
#pragma unroll
        for (int c = 0; c < WV_C; c++) out[c] = row[jm[c]];
    }
}

// (tcc, tgamma) of a cell by coefficient mode
__device__ __forceinline__ void coef_fp64(const WvInst &d, int rm, int jm, unsigned medbit,
                                          double &cc, double &g)
{
    if (d.coef == WV_COEF_UNIFORM) {
        cc = d.tcc0; g = d.tg0;
    } else if (d.coef == WV_COEF_TWO) {
        cc = medbit ? d.tcc1 : d.tcc0;
        g = medbit ? d.tg1 : d.tg0;
    } else {
        const size_t k = (size_t)rm * d.ldy + jm;
        cc = d.tccf[k];
        g = d.tgf ? d.tgf[k] : d.tg0;
    }
}

// sum of the source amplitudes at (rm, jm) for amplitude index g (FP64 form
// adds them one by one, in source order, to the updated value)
template <typename T>
__device__ __forceinline__ T add_sources(const WvInst &d, int rm, int jm, long g, T v)
{
    const int k0 = d.src_ptr[rm], k1 = d.src_ptr[rm + 1];
    for (int k = k0; k < k1; k++)
        if (d.src_col[k] == jm) v = v + (T)d.amp[(size_t)d.src_id[k] * d.amp_len + g];
    return v;
}

// ---------------------------------------------------------------- FP64 ----

template <int K, int ST>
__global__ void __launch_bounds__(64)
k_step_fp64(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, const long *gstep,
            long g0, int ia, int ib, int oc, int od)
{
    constexpr int C = WV_C;
    constexpr int H = (K + C - 1) / C * C;
    constexpr int WOUT = 64 * C - 2 * H;
    const WvTile t = tiles[blockIdx.x];
    const WvInst &d = I[t.inst];
    const int lane = threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    const bool wrapi = d.bc == 1, wrapj = (d.bc == 1 || d.bc == 3);
    const int jb = t.cb * WOUT;
    const int j0 = jb - H + lane * C;
    const int ibeg = t.sb * d.rows_per_tile;
    const int iend = min(ibeg + d.rows_per_tile, nx);
    const long gbase = (gstep ? *gstep : 0) + g0;
    const bool has_src = d.src_ptr != nullptr && d.amp_len > 0;

    const double *__restrict__ u0 = d.u[ia];
    const double *__restrict__ um1 = d.u[ib];
    double *__restrict__ out0 = d.u[oc];
    double *__restrict__ out1 = d.u[od];

    int jm[C];
    bool colint[C], wcol[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int jj = j0 + c;
        jm[c] = map_idx(jj, ny, wrapj);
        colint[c] = jm[c] >= 1 && jm[c] <= ny - 2;
        wcol[c] = jj >= jb && jj < jb + WOUT && jj < ny;
    }
    const bool inside = j0 >= 0 && j0 + C <= ny;

    double Wn[K][3][C];  // register windows of levels 0..K-1
    double Y[C];         // u^{n-1} row s-1
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < C; c++) Wn[a][b][c] = 0.0;

    auto iter = [&](int s, auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int PM = (P + 2) % 3;  // slot of the previous row (centre)
        constexpr int PO = (P + 1) % 3;  // slot of the row before (oldest)
        load_row(u0, d, map_idx(s, nx, wrapi), j0, jm, inside, Wn[0][P]);
        load_row(um1, d, map_idx(s - 1, nx, wrapi), j0, jm, inside, Y);
#pragma unroll
        for (int lv = 1; lv <= K; lv++) {
            const int r = s - lv;
            const int rm = map_idx(r, nx, wrapi);
            const bool rowint = rm >= 1 && rm <= nx - 2;
            const double(&Xe)[C] = Wn[lv - 1][P];   // row r+1
            const double(&Xc)[C] = Wn[lv - 1][PM];  // row r
            const double(&Xw)[C] = Wn[lv - 1][PO];  // row r-1
            const double(&Yv)[C] = (lv == 1) ? Y : Wn[lv >= 2 ? lv - 2 : 0][PO];
            const double cn = from_next(Xc[0]), cp = from_prev(Xc[C - 1]);
            double en = 0, ep = 0, wn = 0, wp = 0;
            if (ST == 9) {
                en = from_next(Xe[0]); ep = from_prev(Xe[C - 1]);
                wn = from_next(Xw[0]); wp = from_prev(Xw[C - 1]);
            }
            const unsigned ab = lane_bits(d.act, d, rm, j0, jm, inside);
            const unsigned mb = d.coef == WV_COEF_TWO ? lane_bits(d.med, d, rm, j0, jm, inside) : 0u;
            double nv[C];
#pragma unroll
            for (int c = 0; c < C; c++) {
                const double x = Xc[c], y = Yv[c], E = Xe[c], Wv = Xw[c];
                const double N = (c < C - 1) ? Xc[c < C - 1 ? c + 1 : 0] : cn;
                const double S = (c > 0) ? Xc[c > 0 ? c - 1 : 0] : cp;
                double v = x;
                if ((ab >> c) & 1u) {
                    double cc, g;
                    coef_fp64(d, rm, jm[c], (mb >> c) & 1u, cc, g);
                    if (rowint && colint[c]) {
                        double delta;
                        if (ST == 9) {
                            const double NE = (c < C - 1) ? Xe[c < C - 1 ? c + 1 : 0] : en;
                            const double SE = (c > 0) ? Xe[c > 0 ? c - 1 : 0] : ep;
                            const double NW = (c < C - 1) ? Xw[c < C - 1 ? c + 1 : 0] : wn;
                            const double SW = (c > 0) ? Xw[c > 0 ? c - 1 : 0] : wp;
                            const double s4 = ((E + Wv) + N) + S;
                            const double sd = ((NE + NW) + SE) + SW;
                            delta = ((4.0 * s4 + sd) - 20.0 * x) * W9;
                        } else {
                            delta = (((E + Wv) + N) + S) - 4.0 * x;
                        }
                        v = leap(x, y, delta, cc, d.kappa, g);
                    } else {
                        const double tcv = d.tc[(size_t)rm * d.ldy + jm[c]];
                        v = edge_fp64(d, rm, jm[c], x, y, E, Wv, N, S, cc, g, tcv);
                    }
                    if (has_src) {
                        const long g1 = gbase + lv - 1;
                        if (g1 < d.amp_len) v = add_sources(d, rm, jm[c], g1, v);
                    }
                }
                nv[c] = v;
            }
            if (lv < K) {
#pragma unroll
                for (int c = 0; c < C; c++) Wn[lv < K ? lv : 0][P][c] = nv[c];
            } else if (r >= ibeg && r < iend) {
                double *o0 = out0 + (size_t)r * d.ldy;
                double *o1 = out1 + (size_t)r * d.ldy;
#pragma unroll
                for (int c = 0; c < C; c++)
                    if (wcol[c] && ((ab >> c) & 1u)) {
                        o0[j0 + c] = nv[c];
                        if (K >= 2) o1[j0 + c] = Xc[c];
                    }
            }
        }
    };

    const int send = iend + K;
    for (int s = ibeg - K; s < send; s += 3) {
        iter(s, std::integral_constant<int, 0>{});
        if (s + 1 >= send) break;
        iter(s + 1, std::integral_constant<int, 1>{});
        if (s + 2 >= send) break;
        iter(s + 2, std::integral_constant<int, 2>{});
    }
}

// ------------------------------------------------------------ FP32 split ----

__device__ __forceinline__ void dd_add_f(float h, float l, float w, float &ho, float &lo)
{
    const float s = h + w;
    const float bb = s - h;
    const float e = (h - (s - bb)) + (w - bb);
    const float t = l + e;
    const float hn = s + t;
    ho = hn;
    lo = t - (hn - s);
}

template <int K, int ST>
__global__ void __launch_bounds__(64)
k_step_split(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, const long *gstep,
             long g0, int ia, int oc)
{
    constexpr int C = WV_C;
    constexpr int H = (K + C - 1) / C * C;
    constexpr int WOUT = 64 * C - 2 * H;
    const WvTile t = tiles[blockIdx.x];
    const WvInst &d = I[t.inst];
    const int lane = threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    const bool wrapi = d.bc == 1, wrapj = (d.bc == 1 || d.bc == 3);
    const int jb = t.cb * WOUT;
    const int j0 = jb - H + lane * C;
    const int ibeg = t.sb * d.rows_per_tile;
    const int iend = min(ibeg + d.rows_per_tile, nx);
    const long gbase = (gstep ? *gstep : 0) + g0;
    const bool has_src = d.src_ptr != nullptr && d.amp_len > 0;
    const float kf = (float)d.kappa;
    const float w9 = (float)W9;

    const float *__restrict__ hin = d.hi[ia];
    const float *__restrict__ lin = d.lo[ia];
    const float *__restrict__ vin = d.v[ia];
    float *__restrict__ hout = d.hi[oc];
    float *__restrict__ lout = d.lo[oc];
    float *__restrict__ vout = d.v[oc];

    int jm[C];
    bool colint[C], wcol[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int jj = j0 + c;
        jm[c] = map_idx(jj, ny, wrapj);
        colint[c] = jm[c] >= 1 && jm[c] <= ny - 2;
        wcol[c] = jj >= jb && jj < jb + WOUT && jj < ny;
    }
    const bool inside = j0 >= 0 && j0 + C <= ny;

    float Hw[K][3][C], Lw[K][3][C], Vw[K][3][C];
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < C; c++) Hw[a][b][c] = Lw[a][b][c] = Vw[a][b][c] = 0.0f;

    auto iter = [&](int s, auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int PM = (P + 2) % 3;
        constexpr int PO = (P + 1) % 3;
        {
            const int rm = map_idx(s, nx, wrapi);
            load_row(hin, d, rm, j0, jm, inside, Hw[0][P]);
            load_row(lin, d, rm, j0, jm, inside, Lw[0][P]);
            load_row(vin, d, rm, j0, jm, inside, Vw[0][P]);
        }
#pragma unroll
        for (int lv = 1; lv <= K; lv++) {
            const int r = s - lv;
            const int rm = map_idx(r, nx, wrapi);
            const bool rowint = rm >= 1 && rm <= nx - 2;
            const float(&he)[C] = Hw[lv - 1][P];
            const float(&hc)[C] = Hw[lv - 1][PM];
            const float(&hw)[C] = Hw[lv - 1][PO];
            const float(&le)[C] = Lw[lv - 1][P];
            const float(&lc)[C] = Lw[lv - 1][PM];
            const float(&lw)[C] = Lw[lv - 1][PO];
            const float(&vc)[C] = Vw[lv - 1][PM];
            const float hcn = from_next(hc[0]), hcp = from_prev(hc[C - 1]);
            const float lcn = from_next(lc[0]), lcp = from_prev(lc[C - 1]);
            float hen = 0, hep = 0, hwn = 0, hwp = 0, len_ = 0, lep = 0, lwn = 0, lwp = 0;
            if (ST == 9) {
                hen = from_next(he[0]); hep = from_prev(he[C - 1]);
                hwn = from_next(hw[0]); hwp = from_prev(hw[C - 1]);
                len_ = from_next(le[0]); lep = from_prev(le[C - 1]);
                lwn = from_next(lw[0]); lwp = from_prev(lw[C - 1]);
            }
            const unsigned ab = lane_bits(d.act, d, rm, j0, jm, inside);
            const unsigned mb = d.coef == WV_COEF_TWO ? lane_bits(d.med, d, rm, j0, jm, inside) : 0u;
            float nh[C], nl[C], nvv[C];
#pragma unroll
            for (int c = 0; c < C; c++) {
                constexpr int dummy = 0;
                (void)dummy;
                const int cn = c < C - 1 ? c + 1 : 0, cp = c > 0 ? c - 1 : 0;
                const float hN = (c < C - 1) ? hc[cn] : hcn, hS = (c > 0) ? hc[cp] : hcp;
                const float lN = (c < C - 1) ? lc[cn] : lcn, lS = (c > 0) ? lc[cp] : lcp;
                float oh = hc[c], ol = lc[c], ov = vc[c];
                if ((ab >> c) & 1u) {
                    const unsigned mbit = (mb >> c) & 1u;
                    if (rowint && colint[c]) {
                        const float h0 = hc[c], l0 = lc[c];
                        const float dh = ((he[c] - h0) + (hw[c] - h0)) + ((hN - h0) + (hS - h0));
                        const float dl = ((le[c] - l0) + (lw[c] - l0)) + ((lN - l0) + (lS - l0));
                        float lap;
                        if (ST == 9) {
                            const float hNE = (c < C - 1) ? he[cn] : hen, hSE = (c > 0) ? he[cp] : hep;
                            const float hNW = (c < C - 1) ? hw[cn] : hwn, hSW = (c > 0) ? hw[cp] : hwp;
                            const float lNE = (c < C - 1) ? le[cn] : len_, lSE = (c > 0) ? le[cp] : lep;
                            const float lNW = (c < C - 1) ? lw[cn] : lwn, lSW = (c > 0) ? lw[cp] : lwp;
                            const float dhd = ((hNE - h0) + (hNW - h0)) + ((hSE - h0) + (hSW - h0));
                            const float dld = ((lNE - l0) + (lNW - l0)) + ((lSE - l0) + (lSW - l0));
                            lap = ((4.0f * dh + dhd) + (4.0f * dl + dld)) * w9;
                        } else {
                            lap = dh + dl;
                        }
                        float ch, cl, gf;
                        if (d.coef == WV_COEF_FIELD) {
                            const size_t k = (size_t)rm * d.ldy + jm[c];
                            ch = d.tccf_hi[k];
                            cl = d.tccf_lo ? d.tccf_lo[k] : 0.0f;
                            gf = d.tgf32 ? d.tgf32[k] : d.tgf0;
                        } else {
                            ch = mbit ? d.tcch1 : d.tcch0;
                            cl = mbit ? d.tccl1 : d.tccl0;
                            gf = mbit ? d.tgf1 : d.tgf0;
                        }
                        const float v0 = vc[c];
                        float vn = ((v0 + __fmaf_rn(ch, lap, cl * lap)) - kf * h0) - gf * v0;
                        if (has_src) {
                            const long g1 = gbase + lv - 1;
                            if (g1 < d.amp_len) vn = add_sources(d, rm, jm[c], g1, vn);
                        }
                        dd_add_f(h0, l0, vn, oh, ol);
                        ov = vn;
                    } else {
                        // edge cells in FP64 with the reference formulas
                        const double x = (double)hc[c] + (double)lc[c];
                        const double y = x - (double)vc[c];
                        const double E = (double)he[c] + (double)le[c];
                        const double Wv = (double)hw[c] + (double)lw[c];
                        const double N = (double)hN + (double)lN;
                        const double S = (double)hS + (double)lS;
                        double cc, g;
                        coef_fp64(d, rm, jm[c], mbit, cc, g);
                        const double tcv = d.tc[(size_t)rm * d.ldy + jm[c]];
                        double out = edge_fp64(d, rm, jm[c], x, y, E, Wv, N, S, cc, g, tcv);
                        if (has_src) {
                            const long g1 = gbase + lv - 1;
                            if (g1 < d.amp_len) out = add_sources(d, rm, jm[c], g1, out);
                        }
                        oh = (float)out;
                        ol = (float)(out - (double)oh);
                        ov = (float)(out - x);
                    }
                }
                nh[c] = oh; nl[c] = ol; nvv[c] = ov;
            }
            if (lv < K) {
#pragma unroll
                for (int c = 0; c < C; c++) {
                    Hw[lv < K ? lv : 0][P][c] = nh[c];
                    Lw[lv < K ? lv : 0][P][c] = nl[c];
                    Vw[lv < K ? lv : 0][P][c] = nvv[c];
                }
            } else if (r >= ibeg && r < iend) {
                const size_t ro = (size_t)r * d.ldy;
#pragma unroll
                for (int c = 0; c < C; c++)
                    if (wcol[c] && ((ab >> c) & 1u)) {
                        hout[ro + j0 + c] = nh[c];
                        lout[ro + j0 + c] = nl[c];
                        vout[ro + j0 + c] = nvv[c];
                    }
            }
        }
    };

    const int send = iend + K;
    for (int s = ibeg - K; s < send; s += 3) {
        iter(s, std::integral_constant<int, 0>{});
        if (s + 1 >= send) break;
        iter(s + 1, std::integral_constant<int, 1>{});
        if (s + 2 >= send) break;
        iter(s + 2, std::integral_constant<int, 2>{});
    }
}

// --------------------------------------------------------------- energy ----

// one wave per (instance, row): 64 lane-strided partial sums, then a halving
// tree; the same order as the oracle's fixed-order sum
template <int SPLIT>
__global__ void __launch_bounds__(64)
k_energy_rows(const WvInst *__restrict__ I, const int2 *__restrict__ rows, double *partial, int ia,
              int ib)
{
    const int2 rr = rows[blockIdx.x];
    const WvInst &d = I[rr.x];
    const int i = rr.y, lane = threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    const int ip = i + 1 == nx ? nx - 1 : i + 1, im = i == 0 ? 0 : i - 1;
    double acc = 0.0;
    for (int j = lane; j < ny; j += 64) {
        double e = 0.0;
        if ((d.eact[(size_t)i * d.ldw + (j >> 6)] >> (j & 63)) & 1u) {
            const int jp = j + 1 == ny ? ny - 1 : j + 1, jn = j == 0 ? 0 : j - 1;
            double c, v, uip, uim, ujp, ujm;
            if (SPLIT) {
                auto U = [&](int a, int b) {
                    const size_t k = (size_t)a * d.ldy + b;
                    return (double)d.hi[ia][k] + (double)d.lo[ia][k];
                };
                c = U(i, j);
                v = (double)d.v[ia][(size_t)i * d.ldy + j];
                uip = U(ip, j); uim = U(im, j); ujp = U(i, jp); ujm = U(i, jn);
            } else {
                const double *u = d.u[ia];
                c = u[(size_t)i * d.ldy + j];
                v = c - d.u[ib][(size_t)i * d.ldy + j];
                uip = u[(size_t)ip * d.ldy + j]; uim = u[(size_t)im * d.ldy + j];
                ujp = u[(size_t)i * d.ldy + jp]; ujm = u[(size_t)i * d.ldy + jn];
            }
            double cc, g;
            const unsigned mb = d.coef == WV_COEF_TWO
                                    ? (unsigned)((d.med[(size_t)i * d.ldw + (j >> 6)] >> (j & 63)) & 1u)
                                    : 0u;
            coef_fp64(d, i, j, mb, cc, g);
            const double a1 = uip - c, a2 = c - uim, b1 = ujp - c, b2 = c - ujm;
            const double gx2 = a1 * a1 + a2 * a2, gy2 = b1 * b1 + b2 * b2;
            e = d.es2 * (v * v + 0.5 * cc * (gx2 + gy2));
        }
        acc += e;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off, 64);
    if (lane == 0) partial[blockIdx.x] = acc;
}

// one wave per instance over its row sums (same order as above)
__global__ void __launch_bounds__(64)
k_energy_total(const double *partial, const int *row0, double *out)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const int a = row0[k], n = row0[k + 1] - a;
    double acc = 0.0;
    for (int i = lane; i < n; i += 64) acc += partial[a + i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off, 64);
    if (lane == 0) out[k] = acc;
}

__global__ void k_bump(long *g, long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *g += n;
}

template <int K>
int launch_fp64_k(int stencil, const WvInst *I, const WvTile *T, int n, const long *gs, long g0,
                  int ia, int ib, int oc, int od, hipStream_t s)
{
    if (stencil == 9)
        hipLaunchKernelGGL((k_step_fp64<K, 9>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, ib, oc, od);
    else
        hipLaunchKernelGGL((k_step_fp64<K, 5>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, ib, oc, od);
    return (int)hipGetLastError();
}

template <int K>
int launch_split_k(int stencil, const WvInst *I, const WvTile *T, int n, const long *gs, long g0,
                   int ia, int oc, hipStream_t s)
{
    if (stencil == 9)
        hipLaunchKernelGGL((k_step_split<K, 9>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, oc);
    else
        hipLaunchKernelGGL((k_step_split<K, 5>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, oc);
    return (int)hipGetLastError();
}

}  // namespace

bool wv_have_depth(int k) { return k == 1 || k == 2 || k == 3 || k == 4 || k == 6 || k == 8; }

int wv_launch_fp64(int k, int stencil, const WvInst *I, const WvTile *T, int n, const long *gs,
                   long g0, int ia, int ib, int oc, int od, hipStream_t s)
{
    switch (k) {
    case 1: return launch_fp64_k<1>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    case 2: return launch_fp64_k<2>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    case 3: return launch_fp64_k<3>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    case 4: return launch_fp64_k<4>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    case 6: return launch_fp64_k<6>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    case 8: return launch_fp64_k<8>(stencil, I, T, n, gs, g0, ia, ib, oc, od, s);
    }
    return -1;
}

int wv_launch_split(int k, int stencil, const WvInst *I, const WvTile *T, int n, const long *gs,
                    long g0, int ia, int oc, hipStream_t s)
{
    switch (k) {
    case 1: return launch_split_k<1>(stencil, I, T, n, gs, g0, ia, oc, s);
    case 2: return launch_split_k<2>(stencil, I, T, n, gs, g0, ia, oc, s);
    case 3: return launch_split_k<3>(stencil, I, T, n, gs, g0, ia, oc, s);
    case 4: return launch_split_k<4>(stencil, I, T, n, gs, g0, ia, oc, s);
    case 6: return launch_split_k<6>(stencil, I, T, n, gs, g0, ia, oc, s);
    case 8: return launch_split_k<8>(stencil, I, T, n, gs, g0, ia, oc, s);
    }
    return -1;
}

int wv_launch_energy(int split, const WvInst *I, const int2 *rows, int nrows, double *partial,
                     const int *row0, int ninst, double *out, int ia, int ib, hipStream_t s)
{
    if (split)
        hipLaunchKernelGGL((k_energy_rows<1>), dim3(nrows), dim3(64), 0, s, I, rows, partial, ia, ib);
    else
        hipLaunchKernelGGL((k_energy_rows<0>), dim3(nrows), dim3(64), 0, s, I, rows, partial, ia, ib);
    hipLaunchKernelGGL(k_energy_total, dim3(ninst), dim3(64), 0, s, partial, row0, out);
    return (int)hipGetLastError();
}

int wv_launch_bump(long *g, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, s, g, n);
    return (int)hipGetLastError();
}
