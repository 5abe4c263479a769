// wv_step_split.hip -- FP32 double-float ("split") stepping kernel, one
// fused depth K per build (compiled once per K with -DWV_K=k).
//
// State per cell: u = hi + lo (two binary32 numbers) and the velocity
// v = u^n - u^{n-1} (one binary32).  The leapfrog update is carried on v,
//   v' = v + tcc*lap(u) - KAPPA*u - tgamma*v,   u' = u + v'
// which is the reference's -y + 2x + tcc*delta - KAPPA*x - tgamma*(x-y)
// rearranged so that no 2x - y cancellation happens in binary32: lap(u) is
// summed from neighbour differences of hi and of lo separately (differences
// of nearby values are exact), tcc is applied as a hi + lo pair, and u' is
// formed with an exact two-sum.  Edge cells (outer rows and columns) are
// evaluated in FP64 with the reference's own formulas.
//
// Same tile, skewed pipeline and prefetch ring as the FP64 kernel
// (wv_step_fp64.hip); a window entry is (hi, lo, v) of one row.
#include "wv_device.h"

#ifndef WV_K
#error "compile with -DWV_K=<fused depth>"
#endif

using namespace wv;

namespace {

// INTERIOR: the tile's rows and columns never reach the outer box (bulk
// formula only, no index wrapping); otherwise every cell is classified.
template <int K, int ST, bool INTERIOR, bool FIELD, bool SRC>
__device__ __forceinline__ void run_split(const WvInst &d, const WvTile t, long gbase, int ia, int oc)
{
    constexpr int C = WV_C;
    constexpr int H = Tile<K>::H, WOUT = Tile<K>::WOUT;
    constexpr unsigned CMASK = (1u << C) - 1u;
    static_assert(2 * K * C <= 32, "bit shift register holds 2K rows of C bits");
    const Par p(d);
    const Src src(d);
    const int nx = p.nx, ny = p.ny, ldy = p.ldy, ldw = p.ldw;
    const bool wrapi = p.bc == 1, wrapj = (p.bc == 1 || p.bc == 3);
    const int jb = t.cb * WOUT;
    constexpr bool FAST = INTERIOR;
    const Cols<FAST> L(jb - H + (int)threadIdx.x * C, jb, WOUT, ny, wrapj);
    const int ibeg = t.sb * d.rows_per_tile;
    const int iend = min(ibeg + d.rows_per_tile, nx);
    const bool has_src = SRC && src.any();
    const uint64_t *act = d.act, *med = d.med;
    const double *tcp = d.tc, *tccf = d.tccf, *tgf = d.tgf;
    const float *tccf_hi = d.tccf_hi, *tccf_lo = d.tccf_lo, *tgf32 = d.tgf32;
    const float kf = (float)p.kappa;
    const float w9 = (float)W9;
    const float *hin = d.hi[ia], *lin = d.lo[ia], *vin = d.v[ia];
    gptr<float> hout = gm(d.hi[oc]), lout = gm(d.lo[oc]), vout = gm(d.v[oc]);

    float Hw[K][3][C], Lw[K][3][C], Vw[K][3][C];
    float QH[3][C], QL[3][C], QV[3][C];
    unsigned QB[3], QM[3];
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
        const int rm = INTERIOR ? min(max(row, 0), nx - 1) : map_idx(row, nx, wrapi);
        load_row(hin, ldy, rm, L, qh);
        load_row(lin, ldy, rm, L, ql);
        load_row(vin, ldy, rm, L, qv);
        qb = lane_word(act, ldw, rm, L);
        // unconditional: a load under a branch makes the compiler drain every
        // outstanding load (the prefetch ring) where the branch joins
        qm = lane_word(med, ldw, rm, L);
    };
    fetch(s0, QH[0], QL[0], QV[0], QB[0], QM[0]);
    fetch(s0 + 1, QH[1], QL[1], QV[1], QB[1], QM[1]);
    fetch(s0 + 2, QH[2], QL[2], QV[2], QB[2], QM[2]);

    auto iter = [&](int s, auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int PE = (P + 2) % 3, PC = (P + 1) % 3, PW = P;
        float NH[K + 1][C], NL[K + 1][C], NV[K + 1][C];
        unsigned abK = 0;
#pragma unroll
        for (int lv = 1; lv <= K; lv++) {
            const int r = s - 2 * lv;
            // rows outside the strip's dependence cone are computed but never used:
            // INTERIOR tiles only clamp them so that table reads stay in bounds
            const int rm = INTERIOR ? min(max(r, 0), nx - 1) : map_idx(r, nx, wrapi);
            const bool rowint = INTERIOR || (rm >= 1 && rm <= nx - 2);
            const float(&he)[C] = Hw[lv - 1][PE];
            const float(&hc)[C] = Hw[lv - 1][PC];
            const float(&hw)[C] = Hw[lv - 1][PW];
            const float(&le)[C] = Lw[lv - 1][PE];
            const float(&lc)[C] = Lw[lv - 1][PC];
            const float(&lw)[C] = Lw[lv - 1][PW];
            const float(&vc)[C] = Vw[lv - 1][PC];
            const int sh = (2 * lv - 1) * C;
            const unsigned ab = (mreg >> sh) & CMASK;
            const unsigned mb = (medreg >> sh) & CMASK;
            if (lv == K) abK = ab;
            const float hcn = from_next(hc[0]), hcp = from_prev(hc[C - 1]);
            const float lcn = from_next(lc[0]), lcp = from_prev(lc[C - 1]);
            float hen = 0, hep = 0, hwn = 0, hwp = 0, len_ = 0, lep = 0, lwn = 0, lwp = 0;
            if (ST == 9) {
                hen = from_next(he[0]); hep = from_prev(he[C - 1]);
                hwn = from_next(hw[0]); hwp = from_prev(hw[C - 1]);
                len_ = from_next(le[0]); lep = from_prev(le[C - 1]);
                lwn = from_next(lw[0]); lwp = from_prev(lw[C - 1]);
            }
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int cn = c < C - 1 ? c + 1 : 0, cp = c > 0 ? c - 1 : 0;
                const float hN = (c < C - 1) ? hc[cn] : hcn, hS = (c > 0) ? hc[cp] : hcp;
                const float lN = (c < C - 1) ? lc[cn] : lcn, lS = (c > 0) ? lc[cp] : lcp;
                const unsigned mbit = (mb >> c) & 1u;
                float oh, ol, ov;
                if (rowint && L.colint[c]) {
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
                    if (FIELD) {
                        const size_t k = (size_t)rm * ldy + L.jm[c];
                        ch = gc(tccf_hi)[k];
                        cl = tccf_lo ? gc(tccf_lo)[k] : 0.0f;
                        gf = tgf32 ? gc(tgf32)[k] : p.tgf0;
                    } else {
                        ch = mbit ? p.tcch1 : p.tcch0;
                        cl = mbit ? p.tccl1 : p.tccl0;
                        gf = mbit ? p.tgf1 : p.tgf0;
                    }
                    const float v0 = vc[c];
                    float vn = ((v0 + __fmaf_rn(ch, lap, cl * lap)) - kf * h0) - gf * v0;
                    if (has_src) {
                        const long g1 = gbase + lv - 1;
                        if (g1 < src.len) vn = add_sources(src, rm, L.jm[c], g1, vn);
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
                    coef_fp64<FIELD>(p, tccf, tgf, rm, L.jm[c], mbit, cc, g);
                    const double tcv = gc(tcp)[(size_t)rm * ldy + L.jm[c]];
                    double out = edge_fp64(p, rm, L.jm[c], x, y, E, Wv, N, S, cc, g, tcv);
                    if (has_src) {
                        const long g1 = gbase + lv - 1;
                        if (g1 < src.len) out = add_sources(src, rm, L.jm[c], g1, out);
                    }
                    oh = (float)out;
                    ol = (float)(out - (double)oh);
                    ov = (float)(out - x);
                }
                const bool act = (ab >> c) & 1u;  // cells the step never writes keep their state
                NH[lv][c] = act ? oh : hc[c];
                NL[lv][c] = act ? ol : lc[c];
                NV[lv][c] = act ? ov : vc[c];
            }
        }
        const int r = s - 2 * K;
        if (r >= ibeg && r < iend) {
            const size_t ro = (size_t)r * ldy + L.j0;
            if (FAST && abK == CMASK && L.wcol[0] && L.wcol[C - 1]) {
                const fvec2 a = {NH[K][0], NH[K][1]}, b = {NL[K][0], NL[K][1]}, e = {NV[K][0], NV[K][1]};
                *(gptr<fvec2>)(hout + ro) = a;
                *(gptr<fvec2>)(lout + ro) = b;
                *(gptr<fvec2>)(vout + ro) = e;
            } else {
#pragma unroll
                for (int c = 0; c < C; c++)
                    if (L.wcol[c] && ((abK >> c) & 1u)) {
                        hout[ro + c] = NH[K][c];
                        lout[ro + c] = NL[K][c];
                        vout[ro + c] = NV[K][c];
                    }
            }
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            Hw[0][P][c] = QH[P][c];
            Lw[0][P][c] = QL[P][c];
            Vw[0][P][c] = QV[P][c];
        }
        mreg = (mreg << C) | bits_of(QB[P], L);
        medreg = (medreg << C) | bits_of(QM[P], L);
#pragma unroll
        for (int lv = 1; lv < K; lv++)
#pragma unroll
            for (int c = 0; c < C; c++) {
                Hw[lv][P][c] = NH[lv][c];
                Lw[lv][P][c] = NL[lv][c];
                Vw[lv][P][c] = NV[lv][c];
            }
        // unconditional (past the last needed row it re-reads that cached row):
        // a conditional load would make the compiler drain the ring each iteration
        fetch(min(s + 3, slast), QH[P], QL[P], QV[P], QB[P], QM[P]);
    };

    for (int s = s0; s < send; s += 3) {
        iter(s, std::integral_constant<int, 0>{});
        if (s + 1 >= send) break;
        iter(s + 1, std::integral_constant<int, 1>{});
        if (s + 2 >= send) break;
        iter(s + 2, std::integral_constant<int, 2>{});
    }
}

template <int K, int ST>
__global__ void __launch_bounds__(64)
k_step_split(const WvInst *__restrict__ I, const WvTile *__restrict__ tiles, const long *gstep,
             long g0, int ia, int oc)
{
    const WvTile t = tiles[blockIdx.x];
    if (t.inst < 0) return;  // padding of the XCD interleave
    const WvInst &d = I[t.inst];
    const int js = t.cb * Tile<K>::WOUT - Tile<K>::H;
    const long gbase = (gstep ? *gstep : 0) + g0;
    const int L = d.rows_per_tile, ibeg = t.sb * L, iend = min(ibeg + L, d.nx);
    const bool interior = js >= 1 && js + 64 * WV_C <= d.ny - 1 && ibeg >= K && iend <= d.nx - K;
    const bool field = d.coef == WV_COEF_FIELD;
    const bool src = d.src_ptr != nullptr && d.amp_len > 0;
    if (interior) {
        if (field) {
            if (src) run_split<K, ST, true, true, true>(d, t, gbase, ia, oc);
            else run_split<K, ST, true, true, false>(d, t, gbase, ia, oc);
        } else {
            if (src) run_split<K, ST, true, false, true>(d, t, gbase, ia, oc);
            else run_split<K, ST, true, false, false>(d, t, gbase, ia, oc);
        }
    } else {
        if (field) run_split<K, ST, false, true, true>(d, t, gbase, ia, oc);
        else run_split<K, ST, false, false, true>(d, t, gbase, ia, oc);
    }
}

}  // namespace

#define WV_CAT2(a, b) a##b
#define WV_CAT(a, b) WV_CAT2(a, b)

int WV_CAT(wv_launch_split_k, WV_K)(int stencil, const WvInst *I, const WvTile *T, int n, const long *gs,
                                     long g0, int ia, int oc, hipStream_t s)
{
    if (stencil == 9)
        hipLaunchKernelGGL((k_step_split<WV_K, 9>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, oc);
    else
        hipLaunchKernelGGL((k_step_split<WV_K, 5>), dim3(n), dim3(64), 0, s, I, T, gs, g0, ia, oc);
    return (int)hipGetLastError();
}
