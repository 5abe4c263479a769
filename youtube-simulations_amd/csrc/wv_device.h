// wv_device.h -- device helpers shared by the stepping and energy kernels.
//
// Arithmetic follows the reference expression by expression (the files are
// compiled with -ffp-contract=off, so nothing is fused into an FMA that the
// reference does not have).
#pragma once
#include <type_traits>

#include "wv_internal.h"

#pragma clang fp contract(off)

namespace wv {

// global-address-space views of pointers read from the instance table (a
// plain pointer loaded from memory would be accessed with flat instructions)
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<const T> gc(const T *p) { return (gptr<const T>)p; }
template <typename T>
__device__ __forceinline__ gptr<T> gm(T *p) { return (gptr<T>)p; }

// native two-element vectors (usable in the global address space)
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
// LDS view of a workgroup array
typedef __attribute__((address_space(3))) fvec2 *lptr2;

// Workgroup barrier for LDS traffic only: this wave's LDS writes are done
// (lgkmcnt) and every wave has arrived.  Global loads in flight stay in
// flight (a fence would wait for them, draining a prefetch ring), and the
// memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

constexpr double W9 = 1.0 / 6.0;  // 9-point weight: ((4*s4 + sd) - 20x) * W9

__device__ __forceinline__ int map_idx(int r, int n, bool wrap)
{
    if (wrap) {
        r %= n;
        return r < 0 ? r + n : r;
    }
    return r < 0 ? 0 : (r >= n ? n - 1 : r);
}

// value held by lane+1 / lane-1 of the wave: DPP wave_shl:1 / wave_shr:1.
// Lanes 63 / 0 have no source lane and read 0 (bound_ctrl): they are halo
// columns of every tile, and a halo of H >= K columns keeps whatever they
// compute out of the K-step dependence cone of the written columns.  Zero
// fill lets the move write its destination directly (no copy of the old
// value first).
__device__ __forceinline__ int dpp_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ float from_next(float v) { return __int_as_float(dpp_next(__float_as_int(v))); }
__device__ __forceinline__ float from_prev(float v) { return __int_as_float(dpp_prev(__float_as_int(v))); }
__device__ __forceinline__ double from_next(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_next((int)b), hi = dpp_next((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_prev(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_prev((int)b), hi = dpp_prev((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Move a prefetched value out of its ring register with a copy the compiler
// keeps.  Without it the copy is folded away, the ring register goes on living
// as a window row, the next load into that ring slot gets another register,
// and the loop back-edge copies it back -- which waits for that load and so
// drains the prefetch ring once per loop trip.
__device__ __forceinline__ double ring_take(double x)
{
    double y;
    asm volatile("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
    return y;
}
__device__ __forceinline__ float ring_take(float x)
{
    float y;
    asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
    return y;
}

// wave_billiard.c:373  -y + 2*x + tcc*delta - KAPPA*x - tgamma*(x-y)
__device__ __forceinline__ double leap(double x, double y, double delta, double cc, double kappa,
                                       double g)
{
    double a = -y + 2.0 * x;
    a = a + cc * delta;
    a = a - kappa * x;
    return a - g * (x - y);
}

// leap() without its last two terms.  With KAPPA = 0 and tgamma = 0,
// a - KAPPA*x - tgamma*(x-y) equals a unless a is -0 or x, y are not finite,
// and a = (-y + 2x) + tcc*delta is -0 only when x is -0 (a sum is -0 only if
// both addends are), so for a state without -0, infinities or NaNs the steps
// never create one and the two forms agree bit for bit.  The host selects
// this form only for such states.
__device__ __forceinline__ double leap_undamped(double x, double y, double delta, double cc)
{
    double a = -y + 2.0 * x;
    return a + cc * delta;
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
// Scalar parameters of one grid, read once into registers before a kernel's
// main loop (reading them through the instance table inside the loop would
// re-load them after every store, since a store could alias the table).
struct Par {
    int nx, ny, ldy, ldw, bc, osc_tb, tb_shift, coef;
    double kappa, ks, gs, kt, gt;
    double tcc0, tcc1, tg0, tg1;
    float tcch0, tccl0, tcch1, tccl1, tgf0, tgf1;
    __device__ __forceinline__ explicit Par(const WvInst &d)
        : nx(d.nx), ny(d.ny), ldy(d.ldy), ldw(d.ldw), bc(d.bc), osc_tb(d.osc_tb),
          tb_shift(d.tb_shift), coef(d.coef), kappa(d.kappa), ks(d.ks), gs(d.gs), kt(d.kt),
          gt(d.gt), tcc0(d.tcc0), tcc1(d.tcc1), tg0(d.tg0), tg1(d.tg1), tcch0(d.tcch0),
          tccl0(d.tccl0), tcch1(d.tcch1), tccl1(d.tccl1), tgf0(d.tgf0), tgf1(d.tgf1)
    {
    }
};

__device__ inline double edge_fp64(const Par &d, int rm, int jm, double x, double y, double E,
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

// edge_fp64() for a cell of the top or bottom row (jm = 0 or NY-1) whose row
// lies inside the box (1 <= rm <= NX-2): the corner tests resolve, B_COND 1
// and 3 take the same four-neighbour sum, and what is left is one uniform
// switch on B_COND (same arithmetic, same order as edge_fp64)
__device__ __forceinline__ double edge_tb_inside(const Par &d, int rm, int jm, double x, double y, double E,
                                                 double Wv, double N, double S, double cc, double g,
                                                 double tcv)
{
    const bool top = (jm == d.ny - 1);
    const double V = top ? S : N;  // inward neighbour along j
    const double O = top ? N : S;  // neighbour across the periodic seam
    double delta;
    if (d.osc_tb && rm < d.tb_shift) {
        delta = (E + Wv) + -2.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    }
    switch (d.bc) {
    case 0:
        delta = ((E + Wv) + V) - 3.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    case 1:
    case 3:
        delta = (((E + Wv) + V) + O) - 4.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    default:
        return absorb(x, y, V, tcv, d.kt, d.gt);
    }
}

// edge_fp64() for a cell of row 0 or NX-1 (rm) of a grid not periodic along
// i (B_COND != 1): the corner cells (jm = 0 or NY-1) by the top/bottom-row
// formulas, the others by the left/right-column ones; OSCILLATE_TOPBOT does
// not reach rows 0 and NX-1 (same arithmetic, same order as edge_fp64)
__device__ __forceinline__ double edge_row_box(const Par &d, int rm, int jm, double x, double y, double E,
                                               double Wv, double N, double S, double cc, double g, double tcv)
{
    const bool first = rm == 0;
    double delta;
    if (jm == d.ny - 1 || jm == 0) {
        const bool top = (jm == d.ny - 1);
        const double V = top ? S : N;  // inward neighbour along j
        const double O = top ? N : S;  // neighbour across the periodic seam
        switch (d.bc) {
        case 0: {
            const double Ec = first ? E : x;
            const double Wc = first ? x : Wv;
            delta = ((Ec + Wc) + V) - 3.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        }
        case 3:
            if (first) return absorb(x, y, E, tcv, d.ks, d.gs);
            delta = (((x + Wv) + V) + O) - 4.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        default:
            return absorb(x, y, V, tcv, d.kt, d.gt);
        }
    }
    if (d.bc == 0) {
        delta = first ? ((E + N) + S) - 3.0 * x : ((Wv + N) + S) - 3.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    }
    return absorb(x, y, first ? E : Wv, tcv, d.ks, d.gs);
}

// COMPARISON, wave_3d.c:736-990: the flat grid in two halves, j < NY/2 and
// j >= NY/2.  Cells of the top and bottom rows (periodic wraps stay inside
// their half) and of the interface columns jtop = NY/2-1 and jbot = NY/2, with
// every neighbour read from the step's input field u0 (one step per launch).
// The reference's quirks are kept: the periodic jbot branch adds the cell
// itself, and the VPER_HABS corner (0, jbot) reads cell (1, 0).
__device__ inline double cmp_fp64(const Par &d, const double *u0, int rm, int jm, double x, double y, double cc,
                                  double g, double tcv)
{
    const int nx = d.nx, ny = d.ny, bc = d.bc, jtop = ny / 2 - 1, jbot = ny / 2;
    auto U = [&](int i, int j) { return gc(u0)[(size_t)i * d.ldy + j]; };
    const int ip = rm + 1 == nx ? nx - 1 : rm + 1, im = rm == 0 ? 0 : rm - 1;
    const int ipw = rm + 1 == nx ? 0 : rm + 1, imw = rm == 0 ? nx - 1 : rm - 1;
    double delta;
    if (jm == ny - 1 || jm == 0) {
        const bool top = jm == ny - 1;
        const int jin = top ? ny - 2 : 1, jseam = top ? jbot : jtop;
        if (d.osc_tb && rm < d.tb_shift && rm < nx - 1 && rm > 0) {
            delta = (U(rm + 1, jm) + U(rm - 1, jm)) + -2.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        }
        switch (bc) {
        case 0:
            delta = ((U(ip, jm) + U(im, jm)) + U(rm, jin)) - 3.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        case 1:
            delta = (((U(ipw, jm) + U(imw, jm)) + U(rm, jin)) + U(rm, jseam)) - 4.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        case 3:
            if (rm == 0) return absorb(x, y, U(1, jm), tcv, d.ks, d.gs);
            delta = (((U(ip, jm) + U(im, jm)) + U(rm, jin)) + U(rm, jseam)) - 4.0 * x;
            return leap(x, y, delta, cc, d.kappa, g);
        default:
            return absorb(x, y, U(rm, jin), tcv, d.kt, d.gt);
        }
    }
    // interface columns: the neighbour inside the half, the seam of the half
    const bool low = jm == jtop;
    const int jin = low ? jtop - 1 : jbot + 1;
    switch (bc) {
    case 1:
        delta = (((U(ipw, jm) + U(imw, jm)) + U(rm, jin)) + U(rm, low ? 0 : jbot)) - 4.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    case 2:
        return absorb(x, y, U(rm, jin), tcv, d.kt, d.gt);
    case 3:
        if (rm == 0) return absorb(x, y, low ? U(1, jtop) : U(1, 0), tcv, d.ks, d.gs);
        delta = (((U(ip, jm) + U(im, jm)) + U(rm, jin)) + U(rm, low ? 0 : ny - 1)) - 4.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    default:  // DIRICHLET, ABS_REFLECT
        delta = ((U(ip, jm) + U(im, jm)) + U(rm, jin)) - 3.0 * x;
        return leap(x, y, delta, cc, d.kappa, g);
    }
}

// Column geometry of one lane: FAST lanes lie in 1..ny-2 and read/write
// contiguously; other lanes map every column (wrap or clamp) one by one, on
// every lane, with no branch on the lane's position: a load under a
// per-lane branch would make the compiler wait for every load in flight (the
// prefetch ring) where the branch joins.
template <bool FAST>
struct Cols {
    int j0;
    int jm[WV_C];
    bool colint[WV_C], wcol[WV_C];
    // own: the lane may write its output columns (the mirrored lanes of a
    // wide tile do not)
    __device__ __forceinline__ Cols(int j0_, int jb, int wout, int ny, bool wrapj, bool own = true) : j0(j0_)
    {
#pragma unroll
        for (int c = 0; c < WV_C; c++) {
            const int jj = j0 + c;
            jm[c] = FAST ? jj : map_idx(jj, ny, wrapj);
            colint[c] = FAST ? true : (jm[c] >= 1 && jm[c] <= ny - 2);
            wcol[c] = own && jj >= jb && jj < jb + wout && jj < ny;
        }
    }
};

// The 32-bit word of a bit plane that holds the lane's WV_C bits, returned
// raw (bits at position j0 & 31) so that a prefetched word is not touched
// until it is consumed; bits_of() extracts them.  j0 is a multiple of WV_C
// (1 or 2), so the lane's bits never straddle a word.  Lanes with mapped
// columns assemble their bits at the same position.
template <bool FAST>
__device__ __forceinline__ unsigned lane_word(const uint64_t *plane, int ldw, int rm, const Cols<FAST> &L)
{
    if (FAST) {
        gptr<const uint32_t> row = (gptr<const uint32_t>)(gc(plane) + (size_t)rm * ldw);
        return row[L.j0 >> 5];
    }
    gptr<const uint64_t> row = gc(plane) + (size_t)rm * ldw;
    unsigned b = 0;
#pragma unroll
    for (int c = 0; c < WV_C; c++) b |= (unsigned)((row[L.jm[c] >> 6] >> (L.jm[c] & 63)) & 1u) << c;
    return b << (L.j0 & 31);
}

template <bool FAST>
__device__ __forceinline__ unsigned bits_of(unsigned w, const Cols<FAST> &L)
{
    return (w >> (L.j0 & 31)) & ((1u << WV_C) - 1u);
}

template <bool FAST>
__device__ __forceinline__ void load_row(const double *base, int ldy, int rm, const Cols<FAST> &L,
                                         double (&out)[WV_C])
{
    gptr<const double> row = gc(base) + (size_t)rm * ldy;
    if (FAST) {
        if (WV_C == 2) {
            const dvec2 v = *(gptr<const dvec2>)(row + L.j0);
            out[0] = v.x;
            out[WV_C - 1] = v.y;
        } else {
#pragma unroll
            for (int c = 0; c < WV_C; c++) out[c] = row[L.j0 + c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < WV_C; c++) out[c] = row[L.jm[c]];
    }
}

template <bool FAST>
__device__ __forceinline__ void load_row(const float *base, int ldy, int rm, const Cols<FAST> &L,
                                         float (&out)[WV_C])
{
    gptr<const float> row = gc(base) + (size_t)rm * ldy;
    if (FAST) {
        if (WV_C == 2) {
            const fvec2 v = *(gptr<const fvec2>)(row + L.j0);
            out[0] = v.x;
            out[WV_C - 1] = v.y;
        } else {
#pragma unroll
            for (int c = 0; c < WV_C; c++) out[c] = row[L.j0 + c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < WV_C; c++) out[c] = row[L.jm[c]];
    }
}

// (tcc, tgamma) of a cell by coefficient mode (values, not loads, for the
// two constant-media modes)
template <int COEF>
__device__ __forceinline__ void coef_fp64(const Par &d, const double *tccf, const double *tgf, int rm,
                                          int jm, unsigned medbit, double &cc, double &g)
{
    if (COEF == WV_COEF_UNIFORM) {
        cc = d.tcc0;
        g = d.tg0;
    } else if (COEF == WV_COEF_FIELD) {
        const size_t k = (size_t)rm * d.ldy + jm;
        cc = gc(tccf)[k];
        g = tgf ? gc(tgf)[k] : d.tg0;
    } else {
        cc = medbit ? d.tcc1 : d.tcc0;
        g = medbit ? d.tg1 : d.tg0;
    }
}

// Sources of one grid (CSR by row) and their amplitudes
struct Src {
    const int *ptr, *col, *id;
    const double *amp;
    long len;
    __device__ __forceinline__ explicit Src(const WvInst &d)
        : ptr(d.src_ptr), col(d.src_col), id(d.src_id), amp(d.amp), len(d.amp_len)
    {
    }
    __device__ __forceinline__ bool any() const { return ptr != nullptr && len > 0; }
};

// source amplitudes at (rm, jm) for amplitude index g, added one by one in
// source order to the updated value
template <typename T>
__device__ __forceinline__ T add_sources(const Src &q, int rm, int jm, long g, T v)
{
    const int k0 = gc(q.ptr)[rm], k1 = gc(q.ptr)[rm + 1];
    for (int k = k0; k < k1; k++)
        if (gc(q.col)[k] == jm) v = v + (T)gc(q.amp)[(size_t)gc(q.id)[k] * q.len + g];
    return v;
}

// Neumann ghost points, driven cells, the FLOOR clamp and the table
// Laplacians of one grid
struct Xtra {
    const int *nptr, *ncol, *ncode;
    const double *nw;
    const int *dptr, *dcol, *did;
    const double *dval;
    long dlen, dbase;
    const uint64_t *nz;
    int ldw;
    double vmax;
    const int *tptr, *tcol, *tn;
    const long long *toff;
    const int *pptr, *pcol, *pid;
    double *pbuf;
    int pn;
    long pevery, pcap, pbase;
    __device__ __forceinline__ explicit Xtra(const WvInst &d)
        : nptr(d.neu_ptr), ncol(d.neu_col), ncode(d.neu_code), nw(d.neu_w), dptr(d.drv_ptr),
          dcol(d.drv_col), did(d.drv_id), dval(d.drv_val), dlen(d.drv_len), dbase(d.drv_base), nz(d.nz),
          ldw(d.ldw), vmax(d.vmax), tptr(d.tbl_ptr), tcol(d.tbl_col), tn(d.tbl_n), toff(d.tbl_off),
          pptr(d.prb_ptr), pcol(d.prb_col), pid(d.prb_id), pbuf(d.prb_buf), pn(d.prb_n), pevery(d.prb_every),
          pcap(d.prb_cap), pbase(d.prb_base)
    {
    }
};

// last entry of cell (rm, jm) in a CSR-by-row list whose columns ascend
// within a row (entries of one cell in list order), or -1: binary search (a
// row along an edge can hold thousands of entries)
__device__ inline int row_find(const int *ptr, const int *col, int rm, int jm)
{
    const int beg = gc(ptr)[rm];
    int lo = beg, hi = gc(ptr)[rm + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gc(col)[mid] <= jm) lo = mid + 1;
        else hi = mid;
    }
    return (lo > beg && gc(col)[lo - 1] == jm) ? lo - 1 : -1;
}

// Neumann point at (rm, jm): its packed neighbour offsets and weights
__device__ inline bool neu_find(const Xtra &q, int rm, int jm, int &code, double &w1, double &w2)
{
    const int k = row_find(q.nptr, q.ncol, rm, jm);
    if (k < 0) return false;
    code = gc(q.ncode)[k];
    w1 = gc(q.nw)[2 * k];
    w2 = gc(q.nw)[2 * k + 1];
    return true;
}

// value of a driven cell (rm, jm) at amplitude index g, if it is driven then
__device__ inline bool drv_find(const Xtra &q, int rm, int jm, long g, double &val)
{
    if (g < 0 || g >= q.dlen) return false;
    const int k = row_find(q.dptr, q.dcol, rm, jm);
    if (k < 0) return false;
    val = gc(q.dval)[(size_t)gc(q.did)[k] * q.dlen + g];
    return true;
}

// Probe at (rm, jm): record value v of step g (0-based, counted from the
// probes' start) when g+1 is a multiple of the sampling period
__device__ inline void prb_record(const Xtra &q, int rm, int jm, long g, double v)
{
    const long n = g + 1;
    if (n <= 0 || n % q.pevery) return;
    const long s = n / q.pevery - 1;
    if (s >= q.pcap) return;
    const int k = row_find(q.pptr, q.pcol, rm, jm);
    if (k >= 0) gm(q.pbuf)[(size_t)s * q.pn + gc(q.pid)[k]] = v;
}

// Table Laplacian of cell (rm, jm), sub_wave.c:11908-11962: -n*x, then the
// n neighbours of its entry added in table order, read from the step's input
// field u0 (one step per launch).  n = nfix if non-zero (the bulk of
// BC_LSHAPE always adds four).  False if the cell has no entry.
__device__ inline bool tbl_delta(const Xtra &q, const double *u0, int rm, int jm, double x, int nfix,
                                 double &delta)
{
    const int k = row_find(q.tptr, q.tcol, rm, jm);
    if (k < 0) return false;
    const int n = nfix ? nfix : gc(q.tn)[k];
    double d = -(double)n * x;
    for (int m = 0; m < n; m++) d = d + gc(u0)[gc(q.toff)[4 * k + m]];
    delta = d;
    return true;
}

// Edge cell of the neighbour-table stepper, wave_3d.c:404-572: the formulas
// of each B_COND branch with the table Laplacian `delta` where a branch uses
// one (no edge formula for BC_LSHAPE: those cells are not written).
__device__ inline double edge_table_fp64(const Par &d, int rm, int jm, double x, double y, double E,
                                         double Wv, double N, double S, double delta, double cc,
                                         double g, double tcv)
{
    const int nx = d.nx, ny = d.ny, bc = d.bc;
    if (jm == ny - 1 || jm == 0) {
        // top (j = NY-1) and bottom (j = 0) rows, corners included
        if (d.osc_tb && rm < d.tb_shift && rm < nx - 1 && rm > 0) return leap(x, y, delta, cc, d.kappa, g);
        switch (bc) {
        case 0:
        case 1: return leap(x, y, delta, cc, d.kappa, g);
        case 3:
            if (rm == 0) return absorb(x, y, E, tcv, d.ks, d.gs);
            return leap(x, y, delta, cc, d.kappa, g);
        case 2:
        case 4: return absorb(x, y, jm == ny - 1 ? S : N, tcv, d.kt, d.gt);
        default: return x;
        }
    }
    // left (i = 0) and right (i = NX-1) columns, 1 <= j <= NY-2
    switch (bc) {
    case 0:
    case 1: return leap(x, y, delta, cc, d.kappa, g);
    case 2:
    case 3:
    case 4: return absorb(x, y, rm == 0 ? E : Wv, tcv, d.ks, d.gs);
    default: return x;
    }
}

// one of the 3x3 values around a cell: code = (row offset + 1) | (column offset + 1) << 2;
// wS, w, wN: row i-1 at columns j-1, j, j+1; cS, c, cN: row i; eS, e, eN: row i+1
__device__ __forceinline__ double pick9(int code, double wS, double w, double wN, double cS, double c,
                                        double cN, double eS, double e, double eN)
{
    const int a = code & 3, b = (code >> 2) & 3;
    const double r0 = b == 0 ? wS : (b == 1 ? w : wN);
    const double r1 = b == 0 ? cS : (b == 1 ? c : cN);
    const double r2 = b == 0 ? eS : (b == 1 ? e : eN);
    return a == 0 ? r0 : (a == 1 ? r1 : r2);
}

// u + w for u = h + l (double-float): the increment w takes the low word first
// (t = l + w, rounded at the increment's own binary32 precision), then one
// Fast2Sum with h, exact whenever |h| >= |t| -- the normal case for a per-step
// increment; near a zero crossing of u its error stays at the increment's
// rounding level
__device__ __forceinline__ void dd_add_f(float h, float l, float w, float &ho, float &lo)
{
    const float t = l + w;
    const float hn = h + t;
    ho = hn;
    lo = t - (hn - h);
}

// Tile geometry shared by the stepping kernels
template <int K>
struct Tile {
    static constexpr int C = WV_C;
    static constexpr int H = (K + C - 1) / C * C;  // halo columns per side (>= K)
    static constexpr int WOUT = 64 * C - 2 * H;    // output columns per tile
};

}  // namespace wv
