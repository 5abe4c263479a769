// wv_energy.hip -- FP64 energy reduction and the launch dispatch of the
// stepping kernels (one object per fused depth, see wv_step_*.hip).
#include "wv_device.h"

using namespace wv;

namespace {

// --------------------------------------------------------------- energy ----

// one wave per (instance, row): 64 lane-strided partial sums, then a halving
// tree; the same order as the oracle's fixed-order sum
template <int SPLIT>
__global__ void __launch_bounds__(64)
k_energy_rows(const WvInst *__restrict__ I, const int2 *__restrict__ rows, double *partial, int ia,
              int ib)
{
    // workgroups go round-robin to the 8 XCDs: each XCD takes a contiguous
    // band of the row list (a bijection), so the rows above and below a row
    // are read through the same L2 (-15 % FP64, -24 % split at 4096^2)
    const int nb = (int)gridDim.x, b = (int)blockIdx.x, x = b & 7;
    const int q = x * (nb >> 3) + min(x, nb & 7) + (b >> 3);
    const int2 rr = rows[q];
    const WvInst &d = I[rr.x];
    const int i = rr.y, lane = threadIdx.x;
    const int nx = d.nx, ny = d.ny;
    const int ip = i + 1 == nx ? nx - 1 : i + 1, im = i == 0 ? 0 : i - 1;
    double acc = 0.0;
    // every load unconditional (clamped indices stay in the grid) so that the
    // loads of several columns are in flight together; cells off the energy
    // plane add +0, as before, in the same order
#pragma unroll 4
    for (int j = lane; j < ny; j += 64) {
        const bool on = (gc(d.eact)[(size_t)i * d.ldw + (j >> 6)] >> (j & 63)) & 1u;
        double e;
        {
            int jp = j + 1 == ny ? ny - 1 : j + 1, jn = j == 0 ? 0 : j - 1;
            // COMPARISON: no gradient across the halves (compute_energy_mod,
            // wave_common.c:1965-1971)
            if (d.cmp && jp == ny / 2) jp = ny / 2 - 1;
            if (d.cmp && jn == ny / 2 - 1) jn = ny / 2;
            double c, v, uip, uim, ujp, ujm;
            if (SPLIT) {
                auto U = [&](int a, int b) {
                    const size_t k = (size_t)a * d.ldy + b;
                    return (double)gc(d.hi[ia])[k] + (double)gc(d.lo[ia])[k];
                };
                c = U(i, j);
                v = (double)gc(d.v[ia])[(size_t)i * d.ldy + j];
                uip = U(ip, j); uim = U(im, j); ujp = U(i, jp); ujm = U(i, jn);
            } else {
                gptr<const double> u = gc(d.u[ia]);
                c = u[(size_t)i * d.ldy + j];
                v = c - gc(d.u[ib])[(size_t)i * d.ldy + j];
                uip = u[(size_t)ip * d.ldy + j]; uim = u[(size_t)im * d.ldy + j];
                ujp = u[(size_t)i * d.ldy + jp]; ujm = u[(size_t)i * d.ldy + jn];
            }
            double cc, g;
            const unsigned mb = d.coef == WV_COEF_TWO
                                    ? (unsigned)((gc(d.med)[(size_t)i * d.ldw + (j >> 6)] >> (j & 63)) & 1u)
                                    : 0u;
            if (d.coef == WV_COEF_FIELD) coef_fp64<WV_COEF_FIELD>(Par(d), d.tccf, d.tgf, i, j, mb, cc, g);
            else if (d.coef == WV_COEF_TWO) coef_fp64<WV_COEF_TWO>(Par(d), d.tccf, d.tgf, i, j, mb, cc, g);
            else coef_fp64<WV_COEF_UNIFORM>(Par(d), d.tccf, d.tgf, i, j, mb, cc, g);
            const double a1 = uip - c, a2 = c - uim, b1 = ujp - c, b2 = c - ujm;
            const double gx2 = a1 * a1 + a2 * a2, gy2 = b1 * b1 + b2 * b2;
            e = d.es2 * (v * v + 0.5 * cc * (gx2 + gy2));
        }
        acc += on ? e : 0.0;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off, 64);
    if (lane == 0) partial[q] = acc;
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

}  // namespace

#define WV_DEPTHS(X) X(1) X(2) X(3) X(4) X(6) X(8)
#define WV_DECL(k)                                                                              \
    int wv_launch_fp64_k##k(int, int, const WvInst *, const WvTile *, int, const WvTile *, int,          \
                            const long *, long, int, int, int, int, hipStream_t);                 \
    int wv_launch_split_k##k(int, int, const WvInst *, const WvTile *, int, const WvTile *, int,         \
                             const long *, long, int, int, hipStream_t);                        \
    int wv_launch_split_persist_k##k(int, int, const WvInst *, const WvTile *, int, const long *, long, int, \
                                     int, int *, hipStream_t, int *);
WV_DEPTHS(WV_DECL)

bool wv_have_depth(int k) { return k == 1 || k == 2 || k == 3 || k == 4 || k == 6 || k == 8; }

int wv_launch_fp64(int k, int stencil, int cm, const WvInst *I, const WvTile *Ti, int ni, const WvTile *Tb,
                   int nb, const long *gs, long g0, int ia, int ib, int oc, int od, hipStream_t s)
{
    switch (k) {
#define WV_CASE(kk) \
    case kk: return wv_launch_fp64_k##kk(stencil, cm, I, Ti, ni, Tb, nb, gs, g0, ia, ib, oc, od, s);
        WV_DEPTHS(WV_CASE)
#undef WV_CASE
    }
    return -1;
}

int wv_launch_split(int k, int stencil, int cm, const WvInst *I, const WvTile *Ti, int ni, const WvTile *Tb,
                    int nb, const long *gs, long g0, int ia, int oc, hipStream_t s)
{
    switch (k) {
#define WV_CASE(kk) \
    case kk: return wv_launch_split_k##kk(stencil, cm, I, Ti, ni, Tb, nb, gs, g0, ia, oc, s);
        WV_DEPTHS(WV_CASE)
#undef WV_CASE
    }
    return -1;
}

int wv_launch_split_persist(int k, int stencil, int cm, const WvInst *I, const WvTile *T, int ng, const long *gs,
                            long g0, int epochs, int ia0, int *ctr, hipStream_t s, int *grid)
{
    switch (k) {
#define WV_CASE(kk) \
    case kk: return wv_launch_split_persist_k##kk(stencil, cm, I, T, ng, gs, g0, epochs, ia0, ctr, s, grid);
        WV_DEPTHS(WV_CASE)
#undef WV_CASE
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
