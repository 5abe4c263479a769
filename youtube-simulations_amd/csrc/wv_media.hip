// wv_media.hip -- media tables and billiard masks built on the GPU.
//
// init_ior_2d's Gaussian-well models (sub_wave.c:12315-12555:
// IOR_PERIODIC_WELLS, IOR_RANDOM_WELLS, IOR_POISSON_WELLS, IOR_PPP_WELLS) sum
// one Gaussian per well at every cell: O(NX*NY*wells) exponentials, done by
// the reference on one host thread before the first frame.  Here one thread
// per cell walks the well list (wave-uniform loads, summed in the
// reference's order) in FP64; exp and tanh are the device library's.
// init_ior_2d's other per-cell models (k_ior) and the closed-form domains of
// xy_in_billiard_single_domain (k_domain) are one thread per cell too, in
// binary64 without contraction and in the reference's order, so their
// results equal the reference's (IOR_MANDELBROT's log aside).
#include <mutex>

#include "wv_device.h"

using namespace wv;

namespace {

// tcc and tc of cell (i, j) into the engine's per-cell tables (row stride
// ldy); tc = sqrt(tcc) where xy_in != 0 (nz plane), courantb where TWOSPEEDS,
// 0 elsewhere (sub_wave.c:12970-12981); the split form's binary32 pair too
__global__ void __launch_bounds__(256) k_wells(WvWells a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= a.ny || i >= a.nx) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    double sum = 0.0;
    for (int n = 0; n < a.nw; n++) {
        const double dx = x - gc(a.xc)[n], dy = y - gc(a.yc)[n];
        const double r2 = dx * dx + dy * dy;
        const double e = exp(-r2 / a.sigma);
        sum = a.h ? sum + e * gc(a.h)[n] : sum + e;
    }
    if (a.use_tanh) sum = tanh(sum);
    const double tcc = a.courant * sum + a.courantb * (1.0 - sum);
    const size_t k = (size_t)i * a.ldy + j;
    gm(a.tcc)[k] = tcc;
    const bool nz = (gc(a.nz)[(size_t)i * a.ldw + (j >> 6)] >> (j & 63)) & 1u;
    gm(a.tc)[k] = nz ? sqrt(tcc) : (a.twospeeds ? a.courantb : 0.0);
    if (a.tcc_hi) {
        const double ts = tcc * a.split_weight;
        const float hi = (float)ts;
        gm(a.tcc_hi)[k] = hi;
        gm(a.tcc_lo)[k] = (float)(ts - (double)hi);
    }
}

// init_circular_wave (wave_common.c:89-111) and add_circular_wave (337-351)
// on the device state, one thread per cell, the reference's expressions in
// its order (exp, cos and sqrt are the device library's)
__global__ void __launch_bounds__(256) k_pulse(WvPulse a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= a.ny || i >= a.nx) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    const double dist2 = (x - a.x0) * (x - a.x0) + (y - a.y0) * (y - a.y0);
    const bool drawn = a.twospeeds || ((gc(a.nz)[(size_t)i * a.ldw + (j >> 6)] >> (j & 63)) & 1u);
    const size_t k = (size_t)i * a.ldy + j;
    double term = 0.0;
    if (drawn) {
        const double pre = a.mode == 0 ? a.amp : a.amp * a.factor;
        term = pre * exp(-dist2 / a.variance) * cos(-sqrt(dist2) / a.wavelength);
    }
    if (a.phi) {
        if (a.mode == 0) {
            gm(a.phi)[k] = term;   // 0 on cells not drawn
            gm(a.psi)[k] = 0.0;
        } else if (drawn) {
            gm(a.phi)[k] = gc(a.phi)[k] + term;
        }
        return;
    }
    // split form: u = hi + lo, v = u - psi; both sets
    for (int s = 0; s < 2; s++) {
        if (a.mode == 0) {
            const float h = (float)term;
            gm(a.hi[s])[k] = h;
            gm(a.lo[s])[k] = (float)(term - (double)h);
            gm(a.v[s])[k] = (float)term;
        } else if (drawn) {
            const double u = ((double)gc(a.hi[s])[k] + (double)gc(a.lo[s])[k]) + term;
            const float h = (float)u;
            gm(a.hi[s])[k] = h;
            gm(a.lo[s])[k] = (float)(u - (double)h);
            gm(a.v[s])[k] = (float)((double)gc(a.v[s])[k] + term);   // psi unchanged
        }
    }
}

// xy_in_billiard_single_domain (sub_wave.c:5712) at the cell corner of
// ij_to_xy (sub_wave.c:469-476), for the domains of the plain workloads:
// D_SINAI (5774-5779), D_EXT_ELLIPSE (5744-5749) and D_CIRCLES (6216-6227:
// 0 inside any active circle, tested in list order).  The circle list is
// read with wave-uniform loads.
__global__ void __launch_bounds__(256) k_mask(WvMask a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = a.i0 + (int)blockIdx.y;
    if (j >= a.ny || i >= a.nx || i >= a.i0 + a.rows) return;
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    short v = 1;
    if (a.kind == 0) {  // D_SINAI
        v = x * x + y * y > a.lambda * a.lambda;
    } else if (a.kind == 1) {  // D_EXT_ELLIPSE
        v = x * x / (a.lambda * a.lambda) + y * y / (a.mu * a.mu) > 1.0;
    } else {  // D_CIRCLES
        const int n = a.ncirc;
        for (int k = 0; k < n; k++) {
            if (gc(a.circ)[3 * n + k] == 0.0) continue;  // inactive
            const double x1 = gc(a.circ)[k], y1 = gc(a.circ)[n + k], r = gc(a.circ)[2 * n + k];
            const double r2 = r * r;
            if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < r2) {
                v = 0;
                break;
            }
        }
    }
    gm(a.out)[(size_t)(i - a.i0) * a.ny + j] = v;
}

// ---- every closed-form domain of xy_in_billiard_single_domain ---------------

constexpr double REF_DPI = 6.283185307;  // DPI, global_pdes.c:9

// vabs and module2, sub_wave.c:270-285
__device__ inline double vabs_(double x) { return x < 0.0 ? -x : x; }
__device__ inline double module2_(double x, double y) { return sqrt(x * x + y * y); }

// xy_in_triangle, sub_wave.c:5299-5311: all three wedge products >= 0
__device__ inline bool in_triangle(double x, double y, double ax, double ay, double bx, double by, double cx,
                                   double cy)
{
    const double v1 = (bx - ax) * (y - ay) - (by - ay) * (x - ax);
    const double v2 = (cx - bx) * (y - by) - (cy - by) * (x - bx);
    const double v3 = (ax - cx) * (y - cy) - (ay - cy) * (x - cx);
    return v1 >= 0.0 && v2 >= 0.0 && v3 >= 0.0;
}

// z -> z^2 + c from z0 until the level or the escape radius; |z|^2 at the end
__device__ inline double escape_norm2(double u, double v, double cx, double cy, int level, double limit)
{
    const double bound = 1000.0 * limit;
    int k = 0;
    while (k < level && u * u + v * v < bound) {
        const double u1 = u * u - v * v + cx;
        v = 2.0 * u * v + cy;
        u = u1;
        k++;
    }
    return u * u + v * v;
}

// the active circle the point falls in (D_CIRCLES family), in list order
__device__ inline bool in_circles(const WvDomain &a, double x, double y)
{
    const int n = a.ncirc;
    for (int k = 0; k < n; k++) {
        if (gc(a.circ)[3 * n + k] == 0.0) continue;
        const double x1 = gc(a.circ)[k], y1 = gc(a.circ)[n + k], r = gc(a.circ)[2 * n + k];
        if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < r * r) return true;
    }
    return false;
}

// The domains that read lists other reference functions fill (polyrect,
// polyrectrot, polygons, polyline, the global circle list): the lists are
// wave-uniform reads, tested in list order with the reference's expressions.
__device__ inline bool in_rect_open(const WvDomain &a, double x, double y, int k)   // inline tests, 5914, 6593
{
    const auto R = gc(a.rect);
    const int n = a.nrect;
    return x > R[k] && x < R[2 * n + k] && y > R[n + k] && y < R[3 * n + k];
}

__device__ inline bool in_rect_closed(const WvDomain &a, double x, double y, int k)  // xy_in_polyrect 5328
{
    const auto R = gc(a.rect);
    const int n = a.nrect;
    double x1 = R[k], y1 = R[n + k], x2 = R[2 * n + k], y2 = R[3 * n + k];
    if (!(x1 < x2)) { const double t = x1; x1 = x2; x2 = t; }
    if (!(y1 < y2)) { const double t = y1; y1 = y2; y2 = t; }
    return !(x < x1) && !(x > x2) && !(y < y1) && !(y > y2);
}

__device__ inline bool in_rectrot(const WvDomain &a, double x, double y, int k)      // xy_in_rectrotated 5393
{
    const auto R = gc(a.rectrot);
    const int n = a.nrectrot;
    const double X1 = R[k], Y1 = R[n + k], X2 = R[2 * n + k], Y2 = R[3 * n + k], W = R[4 * n + k];
    const double l = module2_(X2 - X1, Y2 - Y1);
    if (l == 0.0) return false;
    const double u1 = (X2 - X1) / l, u2 = (Y2 - Y1) / l, v1 = x - X1, v2 = y - Y1;
    const double pscal = u1 * v1 + u2 * v2;
    if (pscal < 0.0 || pscal > l) return false;
    const double h2 = v1 * v1 + v2 * v2 - pscal * pscal;
    return 4.0 * h2 <= W * W;
}

__device__ inline bool in_tri_v(const WvDomain &a, double x, double y, int p, int q, int r)  // 5313
{
    const auto P = gc(a.pline);
    const int n = a.npline;
    return in_triangle(x, y, P[p], P[n + p], P[q], P[n + q], P[r], P[n + r]);
}

__device__ short domain_value_list(const WvDomain &a, double x, double y)
{
    const double L = a.lambda, M = a.mu, WW = a.wall_width;
    switch (a.domain) {
    case 40: {                                                            // D_POLYGONS 6254, in_tpolygon 380
        const auto G = gc(a.polyg);
        const int n = a.npolyg;
        for (int i = 0; i < n; i++) {
            if (G[3 * n + i] == 0.0) continue;
            const double x1 = (x - G[i]) / G[2 * n + i], y1 = (y - G[n + i]) / G[2 * n + i];
            if (x1 * x1 + y1 * y1 >= 1.0) continue;
            const int ns = (int)G[4 * n + i], off = (int)G[6 * n + i];
            const double cw = G[5 * n + i];
            int cond = 1;
            for (int q = 0; q < ns; q++)
                cond = cond * (x1 * gc(a.trig)[off + 2 * q] + y1 * gc(a.trig)[off + 2 * q + 1] < cw);
            if (cond) return 0;
        }
        return 1;
    }
    case 41: case 27: {                                                   // D_VONKOCH 6260, D_VONKOCH_HEATED 7566
        if (a.domain == 27 && x * x + y * y > L * L) return 2;
        const int n = a.npline;
        int cond = in_tri_v(a, x, y, 0, n / 3, 2 * n / 3), m = 1, k = 1;
        for (int i = 0; i < a.mdepth; i++) {
            m = m * 4;
            for (int j = 0; j < n / m; j++) cond += in_tri_v(a, x, y, j * m + k, j * m + 2 * k, j * m + 3 * k);
            k = k * 4;
        }
        if (a.domain == 27) return cond > 0 ? 0 : 1;
        return cond >= 1;
    }
    case 42: {                                                            // D_STAR 6274
        const int N = a.npoly;
        int cond = in_tri_v(a, x, y, N, N - 1, 0);
        for (int i = 0; i < N - 1; i++) cond += in_tri_v(a, x, y, N, i, i + 1);
        return cond >= 1;
    }
    case 421: {                                                           // D_STAR_CHANNEL 6281
        const int N = a.npoly;
        int cond = x < 0.0 && vabs_(y) < 0.5 * WW;
        for (int i = 0; i < N + 3; i++) cond += in_tri_v(a, x, y, N + 3, i + 1, i);
        return cond >= 1;
    }
    case 53: case 531: case 54: case 541:                                 // D_MAZE 6590, _SMALL, _CLOSED, _CHANNELS
        for (int i = 0; i < a.nrect; i++)
            if (in_rect_open(a, x, y, i)) return 0;
        return 1;
    case 192: case 193: case 194: {                                       // D_TWO_PARABOLAS_ASYM_GUIDE 5903, _SEMITRANS 5921, _LENS 5942
        if (x > 0.0) {
            if (vabs_(y) > 0.5 * L) return 1;
            if (a.domain == 194 && !(vabs_(y) > 0.1 * L)) {
                const double xa = 0.24 * L, xb = xa + WW;
                double r2 = 1.5 * 0.25 * L;
                r2 = r2 * r2;
                const double x3 = sqrt(r2 - 0.01 * L * L) - xb;
                return ((x + x3) * (x + x3) + y * y < r2 && x > xa) ? 0 : 1;
            }
            const double x1 = -y * y / L + 0.25 * L;
            if (a.domain == 192) return x < x1 || x > x1 + WW;
            if (x < x1 || x > x1 + WW) return 1;
            if (a.domain == 193 && vabs_(y) < 0.1 * L) return 0;
            return 2;
        }
        for (int i = 0; i < a.nrect; i++)
            if (in_rect_open(a, x, y, i)) return a.domain == 192 ? 0 : 2;
        if (vabs_(y) > 0.5 * M) return 1;
        const double x1 = y * y / M - 0.25 * M;
        if (a.domain == 192) return x > x1 || x < x1 - WW;
        return (x > x1 || x < x1 - WW) ? 1 : 2;
    }
    case 931: {                                                           // D_CIRCLE_LATTICE_RANDOM 7128
        const double dx = (a.xmax - a.xmin) / (double)a.ngridx, dy = (a.ymax - a.ymin) / (double)a.ngridy;
        for (int i = 0; i < a.ngridx; i++) {
            const double x1 = a.xmin + ((double)i + 0.5) * dx;
            for (int j = 0; j < a.ngridy; j++)
                if (module2_(x - x1, y - (a.ymin + ((double)j + 0.5) * dy)) < M) return 1;
        }
        for (int i = 0; i < a.nrect; i++)
            if (in_rect_closed(a, x, y, i)) return 1;
        return 0;
    }
    case 95: case 97: {                                                   // D_CIRCLE_LATTICE_HONEY 7169, _RHOMBUS 7188
        const double dx = (a.xmax - a.xmin) / (double)a.ngridx, dy = (a.ymax - a.ymin) / (double)a.ngridy;
        for (int j = 0; j < a.ngridy + 1; j++) {
            const double y1 = a.ymin + ((double)j) * dy;
            for (int i = 0; i < a.ngridx + 1; i++) {
                if (a.domain == 95 && !(i % 3 != 2 * ((j + 1) % 2))) continue;
                double x1 = a.xmin + ((double)i) * dx;
                if (j % 2 == 1) x1 += 0.5 * dx;
                if (module2_(x - x1, y - y1) < M) return 1;
            }
        }
        for (int i = 0; i < a.nrectrot; i++)
            if (in_rectrot(a, x, y, i)) return 1;
        return 0;
    }
    case 96: case 932: case 933: case 941: {                              // D_CIRCLE_LATTICE_POISSON 7207, _SQUARE_CROSS 7215, _MAZE 7301, _HEX_MAZE 7311
        const int n = a.ncirc;
        for (int i = 0; i < n; i++)
            if (module2_(x - gc(a.circ)[i], y - gc(a.circ)[n + i]) < M) return 1;
        if (a.domain != 96)
            for (int i = 0; i < a.nrect; i++)
                if (in_rect_closed(a, x, y, i)) return 1;
        for (int i = 0; i < a.nrectrot; i++)
            if (in_rectrot(a, x, y, i)) return 1;
        return 0;
    }
    default: return 1;  // not reached: the launcher admits only the codes above
    }
}

// The cases of xy_in_billiard_single_domain beyond the families of
// domain_value, in the same manner.  The reference's one-time constants (its
// static `first` blocks, sub_wave.c:6631-6636, 6647-6652, 6685-6690, and
// rc_hyp's, 5444-5472) and every cosine, sine and tangent of a fixed angle
// come from the host in c0..c2 and the table.  The Fresnel zone plates
// compare a cosine of the cell's own argument with 0: any faithfully rounded
// cosine has the true value's sign there, the GPU's and the C library's alike.
__device__ short domain_value_more(const WvDomain &a, double x, double y)
{
    const double L = a.lambda, M = a.mu, MB = a.mu_b, WW = a.wall_width;
    const auto T = gc(a.trig);
    switch (a.domain) {
    case 35: {                                                            // D_TOKARSKY 6068
        const double x1 = 4.0 + x / (a.xmax - a.xmin) * 8.4, y1 = 2.0 + y / (a.xmax - a.xmin) * 8.4;
        if (x1 <= 0.0 || x1 >= 8.0) return 0;
        if (x1 < 1.0) return y1 > 2.0 && y1 < x1 + 2.0;
        if (x1 < 2.0) return y1 > 1.0 && y1 < 4.0;
        if (x1 < 3.0) return y1 > x1 - 2.0 && y1 < 3.0;
        if (x1 < 4.0) return y1 > 1.0 && y1 < 2.0;
        if (x1 < 5.0) return y1 > x1 - 4.0 && y1 < 2.0;
        if (x1 < 6.0) return y1 > 1.0 && y1 < 3.0;
        if (x1 < 7.0) return y1 > x1 - 6.0 && y1 < 10.0 - x1;
        return y1 > 2.0 && y1 < 3.0;
    }
    case 43: {                                                            // D_FRESNEL 6288
        if (vabs_(y) > 0.9 * vabs_(L)) return 1;
        if (vabs_(x) > M) return 1;
        double x1 = sqrt(L * L - y * y) - vabs_(L);
        while (x1 <= 0.0) x1 += M;
        if (L > 0.0) return x < x1 ? 0 : 1;
        return x > -x1 ? 0 : 1;
    }
    case 45: {                                                            // D_DOUBLE_FRESNEL 6307
        if (vabs_(y) > 0.9 * vabs_(L)) return 1;
        if (L > 0.0) {
            if (vabs_(x) > L + M) return 1;
            double x1 = sqrt(L * L - y * y) - L;
            while (x1 <= 0.0) x1 += M;
            x1 -= L;
            return vabs_(x) > -x1 ? 0 : 1;
        }
        if (vabs_(x) < -L - M) return 1;
        double x1 = sqrt(L * L - y * y) + L;
        while (x1 <= 0.0) x1 += M;
        x1 -= L;
        return vabs_(x) > x1 ? 1 : 0;
    }
    case 44: case 441: {                                                  // D_NOISEPANEL 6331, _RECT 6339
        double x1 = a.domain == 44 ? vabs_(x) : -x;
        if (a.domain == 441 && !(x1 > 0.02)) {                           // NPWIDTH, global_pdes.c:380
            if (x > 0.02) return vabs_(y) < a.ymax - 0.1 && x < a.xmax - 0.1;
            return 0;
        }
        while (x1 > 2.0 * L) x1 -= 2.0 * L;
        const double y1 = x1 <= L ? 0.1 + M * x1 / L : 0.1 + 2.0 * M - M * x1 / L;
        if (a.domain == 44) return y > a.ymin + y1 && y < a.ymax - y1;
        return y > a.ymin + y1 && y < a.ymax - y1 && x > a.xmin + 0.1;
    }
    case 46: case 461: {                                                  // D_QRD 6355, D_QRD_ASYM 6363
        const bool upper = a.domain == 46 || y > 0.0;
        const int p = upper ? 13 : 17;
        const double x1 = (upper ? vabs_(x) : vabs_(x + 1.0)) / L;
        const int k = (int)(x1 + 0.5), k1 = (k * k) % p;
        const double y1 = (M / (double)p) * ((double)(p + 1) - (double)k1);
        if (a.domain == 46) return y > a.ymin + y1 && y < a.ymax - y1;
        return upper ? y < a.ymax - y1 : y > a.ymin + y1;
    }
    case 47: {                                                            // D_CIRCLE_SEGMENT 6382
        if (vabs_(y) > 0.9 * vabs_(L)) return 1;
        const double y1 = 0.9 * L;
        double x1 = sqrt(L * L - y1 * y1) - vabs_(L) + M;
        if (L > 0.0 && x < x1) return 1;
        else if (L < 0.0 && x > -x1) return 1;
        x1 = sqrt(L * L - y * y) - vabs_(L) + M;
        if (L > 0.0) return x < x1 ? 0 : 1;
        return x > -x1 ? 0 : 1;
    }
    case 48: {                                                            // D_GROOVE 6403
        const double s = 0.85 * L, h = 0.5 * L;
        const double x1 = x - a.xmin - (double)((int)((x - a.xmin) / L)) * L;
        if (x1 < h) return y > a.ymin + L;
        return y > a.ymin + L + s;
    }
    case 49: return vabs_(x - y * L / a.ymax) > 0.5 * M;                  // D_FABRY_PEROT 6411
    case 52: {                                                            // D_WAVEGUIDE_W 6437
        const double x1 = vabs_(x), width = L - 2.0 * M, height = 0.5 * M;
        if (x1 > 2.0 * L - M) return 0;
        if (vabs_(y) > M + width + height) return 0;
        if (y >= height) {
            const double r = module2_(x1, y - height);
            if (r > M && r < M + width) return 1;
            return x1 > L + M;
        }
        if (y <= -height) {
            const double r = module2_(x1 - L, y + height);
            return r > M && r < M + width;
        }
        if (x1 > L + M) return 1;
        return x1 > M && x1 < M + width;
    }
    case 521: {                                                           // D_WAVEGUIDES_W 6461
        const double x1 = vabs_(x);
        double width = L - 2.0 * M, height = M + width, r;                // upper fibre
        if (y > M + width + height) return 0;
        if (y >= height) {
            r = module2_(x1, y - height);
            if (r > M && r < M + width) return 1;
            return y < height + M && x1 > L + M && x1 < L + M + width;
        }
        r = module2_(x1 - L, y - height);
        if (r > M && r < M + width) return 1;
        height = -M - width;                                              // lower fibre
        width = L - 2.0 * MB;
        if (y < -MB - width + height) return 0;
        if (y >= height) {
            r = module2_(x1, y - height);
            if (r > MB && r < MB + width) return 1;
            return y < height + MB && x1 > L + MB && x1 < L + MB + width;
        }
        r = module2_(x1 - L, y - height);
        return r > MB && r < MB + width;
    }
    case 526: {                                                           // D_WAVEGUIDE_BADSPLICE 6575
        const double y1 = x < 0.0 ? vabs_(vabs_(y) - 0.5) : (y > 0.0 ? vabs_(y - 0.5 - M) : vabs_(y + 0.5 - MB));
        if (y1 < 0.9 * L) return 1;
        return y1 < L ? 2 : 0;
    }
    case 55: {                                                            // D_CHESSBOARD 6622
        const int i = (int)(vabs_(x) / L + 0.5), j = (int)(vabs_(y) / L + 0.5);
        return (i + j) % 2 == 0;
    }
    case 56: {                                                            // D_TRIANGLE_TILES 6629
        // c0 = h, c1 = hh, c2 = sin(DPI/3), T[0] = sin(-DPI/3)
        const double h = a.c0, hh = a.c1;
        const int i = (int)((y + h) / hh + 10.0);
        double y1 = a.c2 * x - 0.5 * y;
        const int j = (int)((y1 + h) / hh + 10.0);
        y1 = T[0] * x - 0.5 * y;
        const int k = (int)((y1 + h) / hh + 10.0);
        return (i + j + k) % 2 == 0;
    }
    case 57: {                                                            // D_HEX_TILES 6645; c0 = ra, c1 = rb
        double x1 = (x + a.c0 * y) / L + 30.0, y1 = a.c1 * y / L + 30.0;
        x1 = x1 - (double)(3 * (int)(x1 / 3.0));
        y1 = y1 - (double)(3 * (int)(y1 / 3.0));
        if (x1 > 2.0 && y1 < 1.0) return 1;
        if (x1 < 1.0 && y1 > 2.0) return 1;
        if (x1 + y1 < 1.0) return 1;
        return x1 + y1 > 5.0;
    }
    case 58: {                                                            // D_FUNNELS 6665
        double y1 = y;
        if (y > 0.5 * a.ymax) y1 -= a.ymax;
        if (y < -0.5 * a.ymax) y1 += a.ymax;
        y1 = vabs_(y1 - M * x);
        y1 = vabs_(y1 - 0.5 * a.ymax) * 2.0 / a.ymax;
        return !(y1 > 0.25 * (1.0 + L + x * x));
    }
    case 581: return !(vabs_(y) > M + L * x * x);                        // D_ONE_FUNNEL 6675
    case 59: {                                                            // D_LENSES_RING 6681; c0 = ll
        for (int i = 0; i < a.npoly; i++) {
            const double ca = T[2 * i], sa = T[2 * i + 1];
            const double x1 = x * ca + y * sa, y1 = -x * sa + y * ca;
            if (module2_(x1 - L - a.c0, y1) < M && module2_(x1 - L + a.c0, y1) < M) return 0;
        }
        return 1;
    }
    case 65:                                                              // D_FRESNEL_ZONE_PLATE 6736
        if (vabs_(x + L) >= M) return 1;
        return cos(REF_DPI * (6.25 * y * y + 0.125)) < 0.0;
    case 66:                                                              // D_FRESNEL_ZONE_PLATE_INV 6743
        if (vabs_(x + L) >= M) return 1;
        return cos(REF_DPI * 6.25 * y * y) > 0.0;
    case 731: {                                                           // D_TREE_BEAM 6828; c0 = cos(PI/6)
        const double x1 = vabs_(x), lev[3] = {1.5, 0.25, -1.5}, rad[3] = {1.5, 2.0, 3.0};
        for (int k = 0; k < 3; k++) {
            const double r = rad[k] * L;
            if (in_triangle(x1, y - lev[k] * L, r * a.c0, -0.5 * r, 0.0, r, 0.0, -0.5 * r)) return 1;
        }
        const double ym = a.oscil_ymid;
        if (x < 0.0 && y > ym - 0.5 * WW && y < ym + 0.5 * WW) return 1;
        if (x < -((3.0 * L) * a.c0) + 0.5 * WW && y > ym - 0.75 * WW && y < ym + 0.75 * WW) return 2;
        return 0;
    }
    case 75: {                                                            // D_RITCHEY_CHRETIEN_SPHERICAL 6892
        const double d = 2.5, b = 3.5, r1 = 2.0 * (d + b / L), r2 = 6.0 / (L - 1.0);
        const double x1 = 0.5 * d - r1, x2 = -0.5 * d - r2, y1 = vabs_(y);
        if (x < 0.0) {
            if (y1 > 0.3) return 1;
            if (x < -0.5 * d - WW) return 1;
            return module2_(x - x2, y1) > r2;
        }
        if (y1 < M) return 1;
        if (x > 0.5 * d + WW) return 1;
        return module2_(x - x1, y1) < r1;
    }
    case 751: {                                                           // D_RITCHEY_CHRETIEN_HYPERBOLIC, rc_hyp 5437
        // T = a1, b1, x01, a2, b2, x02 (rc_hyp's one-time constants)
        const double d = 2.5, y1 = vabs_(y);
        if (x > 0.0) {
            if (y1 < M) return 1;
            if (x > 0.5 * d + WW) return 1;
            return (x - T[2]) * (x - T[2]) > T[0] * T[0] * (1.0 + y * y / (T[1] * T[1]));
        }
        if (y1 > 0.3) return 1;
        return x > T[5] || (x - T[5]) * (x - T[5]) < T[3] * T[3] * (1.0 + y * y / (T[4] * T[4]));
    }
    case 78: case 781: {                                                  // D_MAGNETRON 6929, _CATHODE 6947
        // c0 = sqrt(2); T = cos, sin of k*0.5*PID
        const double r = module2_(x, y);
        if (a.domain == 781 && r < WW) return 0;
        if (r > L) return 1;
        if (r < 0.33 * L) return 1;
        if (vabs_(y) < WW && x > -0.66 * L) return 1;
        if (vabs_(x) < WW && vabs_(y) < 0.66 * L) return 1;
        if (vabs_(x - y) < WW * a.c0 && r < 0.66 * L) return 1;
        if (vabs_(x + y) < WW * a.c0 && r < 0.66 * L) return 1;
        for (int k = 0; k < 8; k++)
            if (module2_(x - 0.66 * T[2 * k], y - 0.66 * T[2 * k + 1]) < M) return 1;
        return 0;
    }
    case 792: {                                                           // D_POLYCIRCLES_ANGLED 6992
        // c0 = cb, c1 = sb; T = cos, sin of the NPOLY angles
        if (module2_(x, y) < L) return 1;
        const double cb = a.c0, sb = a.c1;
        for (int k = 0; k < a.npoly; k++) {
            const double ca = T[2 * k], sa = T[2 * k + 1];
            const double x1 = x * ca + y * sa, y1 = -x * sa + y * ca;
            const double x2 = L + 2.0 * M * cb, y2 = 2.0 * M * sb;
            if (module2_(x1 - x2, y1 - y2) < M) return 1;
            const double c1 = -sb * L + cb * WW, c2 = -sb * L - cb * WW, b = -sb * x1 + cb * y1;
            if (x1 > 0.5 * L && x1 < L + M && b < c1 && b > c2) return 1;
        }
        return 0;
    }
    case 793: {                                                           // D_TWOCIRCLES_ELLIPSE 7015
        const double aa = 0.9 * a.xmax - 3.0 * M, b = aa / L;
        if (x * x / (aa * aa) + y * y / (b * b) < 1.0) return 1;
        const double x1 = aa + 2.0 * M;
        if ((x - x1) * (x - x1) + y * y < M * M) return 1;
        if ((x + x1) * (x + x1) + y * y < M * M) return 1;
        return vabs_(x) < aa + 2.0 * M && vabs_(y) < WW;
    }
    case 794: {                                                           // D_CIRCLES_ELLIPSE 7026; T = cos, sin
        const double aa = 0.9 * a.xmax - 3.0 * M, b = aa / L;
        x = x / a.julia_scale;
        y = y / a.julia_scale;
        if (x * x / (aa * aa) + y * y / (b * b) < 1.0) return 1;
        for (int k = 0; k < a.npoly; k++) {
            const double cp = T[2 * k], sp = T[2 * k + 1];
            const double x1 = aa * cp, y1 = b * sp;
            double nx = b * cp, ny = aa * sp;
            double r = 1.0 / module2_(nx, ny);
            nx *= r;
            ny *= r;
            const double x2 = x1 + 2.0 * M * nx, y2 = y1 + 2.0 * M * ny;
            if (module2_(x - x2, y - y2) < M) return 1;
            r = module2_(x - x1, y - y1);
            const double h = (x - x1) * nx + (y - y1) * ny;
            if (h > -M && h < 2.0 * M && r * r - h * h < WW * WW) return 1;
        }
        return 0;
    }
    case 797: case 798: {                                                 // D_TWO_POLYGONS 7062, D_TWO_POLY_PARABOLAS 7068
        // c0 = cos(omega/2); T = cos, sin of the side normals (in_polygon, sub_wave.c:362)
        int cond = 1;
        if (a.domain == 797) {
            x = vabs_(x) - L;
            for (int k = 0; k < a.npoly; k++) cond = cond * (x * T[2 * k] + y * T[2 * k + 1] < M * a.c0);
            return cond;
        }
        if (x > 0.0) x -= L;
        else x += L;
        for (int k = 0; k < a.npoly; k++) {
            const double ca = T[2 * k], sa = T[2 * k + 1];
            const double x1 = x * ca + y * sa, y1 = -x * sa + y * ca;
            cond = cond * (x1 < M - 0.25 * y1 * y1 / M);
        }
        return cond;
    }
    case 93: case 934: case 935: case 936: {                              // D_CIRCLE_LATTICE 7110, _NONISO 7225, _HALF 7243, _HALF_V 7262
        if (a.domain == 935 && y > 0.0) return 1;
        if (a.domain == 936 && x < 0.0) return 1;
        const double wa = a.domain == 93 ? WW : WW * a.wall_width_asym;
        const double dx = (a.xmax - a.xmin) / (double)a.ngridx, dy = (a.ymax - a.ymin) / (double)a.ngridy;
        for (int i = 0; i < a.ngridx; i++) {
            const double x1 = a.xmin + ((double)i + 0.5) * dx;
            if (vabs_(x - x1) < wa) return 1;
            for (int j = 0; j < a.ngridy; j++) {
                const double y1 = a.ymin + ((double)j + 0.5) * dy;
                if (module2_(x - x1, y - y1) < M) return 1;
                if (vabs_(y - y1) < WW) return 1;
            }
        }
        return 0;
    }
    case 937: {                                                           // D_CIRCLE_LATTICE_STRIP 7281
        if (x > L) return 1;
        if (x < -L) return 1;
        const double dx = 2.0 * L / (double)a.ngridx, dy = (a.ymax - a.ymin) / (double)a.ngridy;
        for (int i = 0; i < a.ngridx + 1; i++) {
            const double x1 = -L + ((double)i) * dx;
            if (vabs_(x - x1) < WW * a.wall_width_asym) return 1;
            for (int j = 0; j < a.ngridy; j++) {
                const double y1 = a.ymin + ((double)j + 0.5) * dy;
                if (module2_(x - x1, y - y1) < M) return 1;
                if (vabs_(y - y1) < WW) return 1;
            }
        }
        return 0;
    }
    case 94: case 942: case 947: {                                        // D_CIRCLE_LATTICE_HEX 7146, _HEX_NONISO 7321, _HEX_STRIP 7344
        const bool strip = a.domain == 947;
        if (strip && vabs_(x) > L) return 1;
        const double dx = strip ? 2.0 * L / (double)a.ngridx : (a.xmax - a.xmin) / (double)a.ngridx;
        const double dy = (a.ymax - a.ymin) / (double)a.ngridy, r = module2_(0.5 * dx, dy);
        const double nx = -dy / r, ny = dx / (2.0 * r), x0 = strip ? -L : a.xmin;
        const double wa = a.domain == 94 ? WW : WW * a.wall_width_asym;
        const double wb = a.domain == 94 ? WW : WW * a.wall_width_asym_b;
        for (int j = 0; j < a.ngridy + 1; j++) {
            const double y1 = a.ymin + ((double)j) * dy;
            if (vabs_(y - y1) < WW) return 1;
            for (int i = -1; i < a.ngridx; i++) {
                double x1 = x0 + ((double)i + 0.5) * dx;
                if (j % 2 == 0) x1 += 0.5 * dx;
                if (module2_(x - x1, y - y1) < M) return 1;
                if (vabs_(nx * (x - x1) + ny * (y - y1)) < wa) return 1;
                if (vabs_(-nx * (x - x1) + ny * (y - y1)) < wb) return 1;
            }
        }
        return 0;
    }
    case 98: {                                                            // D_DISC_WAVEGUIDE 7368
        if (x > 0.0) return x * x + y * y < L * L;
        const double y1 = y - 0.5 * WW, r = L - 0.5 * WW;
        if (x * x + y1 * y1 < r * r) return 1;
        if (y > -L && y < -L + WW) return 1;
        if (y > -L - 0.5 * WW && y < -L + 1.5 * WW) return 2;
        return 0;
    }
    case 981: case 982: case 983: {                                       // D_DISC_WAVEGUIDE_SHIFTED 7378, _ELLIPSE 7387, D_WAVEGUIDE_ELLIPSE_OBLIQUE 7396
        if (a.domain == 981 ? x * x + y * y < L * L : x * x / (L * L) + y * y < 1.0) return 1;
        if (x > 0.0) return 0;
        double y1;
        if (a.domain == 981) y1 = y + L - M - 0.5 * WW;
        else if (a.domain == 982) y1 = y + 1.0 - M - 0.5 * WW;
        else y1 = y + 1.0 - M - 0.5 * WW - (x - a.xmin) * a.c0;          // c0 = tan(APOLY*PID)
        if (vabs_(y1) < 0.5 * WW) return 1;
        return vabs_(y1) < 0.75 * WW ? 2 : 0;
    }
    case 99: case 991: {                                                  // D_CIRCLE_PAIRS 7405, D_CIRCLE_POLYGON 7418
        // 991: c0 = cos(omega/2); T = cos, sin of the side normals
        const double dx = (a.xmax - a.xmin) / (double)a.ngridx;
        for (int k = 0; k < a.ngridx; k++) {
            const double x1 = a.xmin + dx * ((double)k + 0.5), y1 = L;
            if (module2_(x - x1, y - y1) < M) return 1;
            if (a.domain == 99) {
                if (module2_(x - x1, y + y1) < MB) return 1;
            } else {
                int cond = 1;
                for (int q = 0; q < a.npoly; q++) cond = cond * ((x - x1) * T[2 * q] + (y + y1) * T[2 * q + 1] < MB * a.c0);
                if (cond) return 1;
            }
            if (vabs_(y) < L && vabs_(x - x1) < 0.5 * WW) return 1;
        }
        return 0;
    }
    case 77: return 1;                                                    // D_IMAGE 6924
    case 741: {                                                           // D_MICHELSON_MOVING 6879; c0 = sqrt(2)
        if (vabs_(x) > L && vabs_(y) > L) return 2;
        const double r = 0.5 * a.c0, h = a.ymax - M - 0.5 * WW, h1 = a.ymax - M - 0.5 * WW + a.michelson_position;
        const double x1 = (x + y) * r, y1 = (-x + y) * r;
        if (vabs_(y1) < 0.5 * WW && vabs_(x1) < L * a.c0) return 1;
        if (vabs_(x - h1) < 0.5 * WW && vabs_(y) < L) return 2;
        if (vabs_(x) < L && vabs_(y - h) < 0.5 * WW) return 2;
        return 0;
    }
    case 21: {                                                            // D_ANNULUS_HEATED 7475
        const double l2 = L * L, r2 = x * x + y * y, r2mu = (x - M) * (x - M) + y * y;
        if (r2mu > l2 && r2 < 1.0) return 1;
        return r2mu <= l2 ? 2 : 0;
    }
    case 22: case 23: {                                                   // D_MENGER_HEATED 7484, D_MENGER_H_OPEN 7500
        const bool inside = vabs_(x) < 1.0 && vabs_(y) < 1.0;
        if (a.domain == 22 && !inside) return 0;
        const int m = a.mratio;
        double x1 = 0.5 * (x + 1.0), y1 = 0.5 * (y + 1.0);
        for (int k = 0; k < a.mdepth; k++) {
            x1 = x1 * (double)m;
            y1 = y1 * (double)m;
            if (inside && ((int)x1 % m) == m / 2 && ((int)y1 % m) == m / 2) return (short)(k + 2);
        }
        return 1;
    }
    default: return domain_value_list(a, x, y);
    }
}

// the escape-time domains (D_JULIA_INT, D_MANDELBROT, D_MANDELBROT_CIRCLE,
// D_JULIA): up to MANDELLEVEL iterations a cell; they also run in a kernel of
// their own, whose loop the other cases do not crowd
__device__ short escape_domain_value(const WvDomain &a, double x, double y)
{
    const double L = a.lambda, M = a.mu;
    switch (a.domain) {
    case 16:                                                              // D_JULIA_INT 7443
        return escape_norm2(x / a.julia_scale, y / a.julia_scale, a.julia_x, a.julia_y, a.mandellevel,
                            a.mandellimit) < a.mandellimit;
    case 24: case 26:                                                     // D_MANDELBROT 7512, _CIRCLE 7531
        if (escape_norm2(0.0, 0.0, x, y, a.mandellevel, a.mandellimit) < a.mandellimit) return 0;
        if (a.domain == 24) return (x - 0.5) * (x - 0.5) / 3.0 + y * y / 1.0 > 1.2 ? 2 : 1;
        return (x - L) * (x - L) + (y - 0.5) * (y - 0.5) < M * M ? 2 : 1;
    default:                                                              // D_JULIA 7547
        if (escape_norm2(x / a.julia_scale, y / a.julia_scale, a.julia_x, a.julia_y, a.mandellevel,
                         a.mandellimit) < a.mandellimit)
            return 0;
        return x * x / 3.0 + y * y / 1.0 > 1.2 ? 2 : 1;
    }
}

static bool escape_domain(int d) { return d == 16 || d == 24 || d == 25 || d == 26; }

// xy_in_billiard_single_domain(x, y), sub_wave.c:5712-7602, for the
// closed-form domains; each case follows the reference's expressions in its
// order of evaluation (line numbers of sub_wave.c).  Host constants c0..c2
// and the trig table are set per domain by the launcher.
__device__ short domain_value(const WvDomain &a, double x, double y)
{
    const double L = a.lambda, M = a.mu, WW = a.wall_width;
    switch (a.domain) {
    case 999: return 1;                                                   // D_NOTHING 5721
    case 0: return vabs_(x) < L && vabs_(y) < 1.0;                        // D_RECTANGLE 5726
    case 1: return x * x / (L * L) + y * y < 1.0;                         // D_ELLIPSE 5732
    case 196: return x * x + y * y < L * L;                               // D_CIRCLE_LAYERS 5738
    case 199: return x * x / (L * L) + y * y / (M * M) > 1.0;             // D_EXT_ELLIPSE 5744
    case 197:                                                             // D_EXT_ELLIPSE_CURVED_BDRY 5757
        if (y > a.ymax - 0.05) return 0;
        if (y < a.ymin + 0.05) return 0;
        [[fallthrough]];
    case 198: {                                                           // D_EXT_ELLIPSE_CURVED 5750
        const double y1 = y + 0.4 * x * x;
        return x * x / (L * L) + y1 * y1 / (M * M) > 1.0;
    }
    case 2:                                                               // D_STADIUM 5766
        if (x > -0.5 * L && x < 0.5 * L && y > -1.0 && y < 1.0) return 1;
        if (module2_(x + 0.5 * L, y) < 1.0) return 1;
        return module2_(x - 0.5 * L, y) < 1.0;
    case 3: return x * x + y * y > L * L;                                 // D_SINAI 5774
    case 4: {                                                             // D_DIAMOND 5780
        const double l2 = L * L, r2 = l2 + (L - 1.0) * (L - 1.0);
        return x * x + y * y < 1.0 && (x - L) * (x - L) + (y - L) * (y - L) > r2 &&
               (x - L) * (x - L) + (y + L) * (y + L) > r2 && (x + L) * (x + L) + (y - L) * (y - L) > r2 &&
               (x + L) * (x + L) + (y + L) * (y + L) > r2;
    }
    case 5: return x > -L && y > -1.0 && L * y + x < 0.0;                 // D_TRIANGLE 5791
    case 6: return y > -L;                                                // D_FLAT 5797
    case 7: {                                                             // D_ANNULUS 5803
        const double l2 = L * L, r2 = x * x + y * y;
        return r2 > l2 && r2 < 1.0;
    }
    case 8: {                                                             // D_POLYGON 5810; c0 = LAMBDA*cos(omega/2)
        int cond = 1;
        for (int k = 0; k < a.npoly; k++)
            cond = cond * (x * gc(a.trig)[2 * k] + y * gc(a.trig)[2 * k + 1] < a.c0);
        return cond;
    }
    case 9:                                                               // D_YOUNG 5823
        if (x < -M || x > M) return 1;
        return vabs_(y - L) < M || vabs_(y + L) < M;
    case 10: {                                                            // D_GRATING 5829
        const int k1 = -(int)((-a.ymin) / L), k2 = (int)(a.ymax / L);
        int cond = 1;
        for (int i = k1; i <= k2; i++) {
            const double z = (double)i * L;
            cond = cond * (x * x + (y - z) * (y - z) > M * M);
        }
        return cond;
    }
    case 11:                                                              // D_EHRENFEST 5842
        return (x - 1.0) * (x - 1.0) + y * y < L * L || (x + 1.0) * (x + 1.0) + y * y < L * L ||
               (vabs_(x) < 1.0 && vabs_(y) < M);
    case 12: {                                                            // D_DISK_GRID 5846
        const double dy = (a.ymax - a.ymin) / ((double)a.ngridy);
        for (int i = -a.ngridx / 2; i < a.ngridx / 2; i++)
            for (int j = 0; j < a.ngridy; j++) {
                const double x1 = ((double)i + 0.5) * dy, y1 = a.ymin + ((double)j + 0.5) * dy;
                if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < M * M) return 0;
            }
        return 1;
    }
    case 13: {                                                            // D_DISK_HEX 5858
        const double dy = (a.ymax - a.ymin) / ((double)a.ngridy);
        for (int i = -a.ngridx / 2; i < a.ngridx / 2; i++)
            for (int j = -1; j < a.ngridy; j++) {
                const double x1 = ((double)i + 0.5) * dy;
                double y1 = a.ymin + ((double)j + 0.5) * dy;
                if ((i + a.ngridx) % 2 == 1) y1 += 0.5 * dy;
                if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < M * M) return 0;
            }
        return 1;
    }
    case 15: case 17: {                                                   // D_MENGER 7431, D_MENGER_ROTATED 7458
        double x1, y1;
        if (a.domain == 17) {
            const double x2 = 1.0 * (x + y), y2 = 1.0 * (x - y);
            if (!(vabs_(x2) < 1.0 && vabs_(y2) < 1.0)) return 1;
            x1 = 0.5 * (x2 + 1.0);
            y1 = 0.5 * (y2 + 1.0);
        } else {
            x1 = 0.5 * (x + 1.0);
            y1 = 0.5 * (y + 1.0);
        }
        // the test reads the unrotated point in both (7439, 7468)
        if (!(vabs_(x) < 1.0 && vabs_(y) < 1.0)) return 1;
        for (int k = 0; k < a.mdepth; k++) {
            x1 = x1 * (double)a.mratio;
            y1 = y1 * (double)a.mratio;
            if (((int)x1 % a.mratio) == a.mratio / 2 && ((int)y1 % a.mratio) == a.mratio / 2) return 0;
        }
        return 1;
    }
    case 16: return escape_domain_value(a, x, y);                        // D_JULIA_INT 7443
    case 18: return x > 0.25 * y * y / L - L;                             // D_PARABOLA 5872
    case 19: {                                                            // D_TWO_PARABOLAS 5876
        const double x1 = 0.25 * y * y / M - M - L, x2 = -x1;
        double width = 0.25 * M;
        if (width > 0.2) width = 0.2;
        if (vabs_(y) > 1.5 * M) return 1;
        if (x < x1 - width || x > x2 + width) return 1;
        return x > x1 && x < x2;
    }
    case 191:                                                             // D_TWO_PARABOLAS_ASYM 5887
        if (x > 0.0) {
            if (vabs_(y) > 0.5 * L) return 1;
            const double x1 = -y * y / L + 0.25 * L;
            return x < x1 || x > x1 + WW;
        } else {
            if (vabs_(y) > 0.5 * M) return 1;
            const double x1 = y * y / M - 0.25 * M;
            return x > x1 || x < x1 - WW;
        }
    case 195: case 1951: case 1952: {                                     // 5977, 5986, 6006
        if (vabs_(y) > 0.75 + WW) return 1;
        if (a.domain != 195) {
            const double ylens = a.domain == 1951 ? 0.2 : a.wall_width_b;
            if (x > 0.0 && vabs_(y) < ylens) {
                // c0 = x3, c1 = x1, c2 = r2 of the window lens
                if ((x + a.c0) * (x + a.c0) + y * y < a.c2 && x > a.c1) return 0;
                return 1;
            }
        }
        const double x1 = 0.5 * L + 0.25 * M - y * y / M;
        if (vabs_(y) > 0.75 && vabs_(x) < x1 + WW) return 2;
        if (vabs_(x) < x1 || vabs_(x) > x1 + WW) return 1;
        if (a.domain == 195 && x > 0.0 && vabs_(y) < 0.2) return 0;
        return 2;
    }
    case 20: case 202: return in_circles(a, x, y) ? 0 : 1;               // D_CIRCLES 6216, _NEUMANN 6228
    case 201:                                                             // D_CIRCLES_IN_RECT 6240
        if (in_circles(a, x, y)) return 0;
        return (vabs_(x) >= L || vabs_(y) >= a.ymax - a.xmax + L) ? 0 : 1;
    case 24: case 25: case 26: return escape_domain_value(a, x, y);      // D_MANDELBROT 7512, D_JULIA 7547, 7531
    case 31: {                                                            // D_FOUR_PARABOLAS 6026
        const double x1 = M + L - 0.25 * y * y / M, y1 = M + L - 0.25 * x * x / M;
        return vabs_(x) < x1 && vabs_(y) < y1;
    }
    case 32: {                                                            // D_POLY_PARABOLAS 6032
        int cond = 1;
        for (int k = 0; k < a.npoly; k++) {
            const double ca = gc(a.trig)[2 * k], sa = gc(a.trig)[2 * k + 1];
            const double x1 = x * ca + y * sa, y1 = -x * sa + y * ca;
            cond = cond * (x1 < L + M - 0.25 * y1 * y1 / M);
        }
        return cond;
    }
    case 33: {                                                            // D_PENROSE 6045; c0 = c
        const double c = a.c0, width = 0.1 * M, x1 = vabs_(x), y1 = vabs_(y);
        if (vabs_(x) >= L) return 0;
        if (vabs_(y) >= M && x * x / (L * L) + (y1 - M) * (y1 - M) / ((1.0 - M) * (1.0 - M)) >= 1.0) return 0;
        if (vabs_(x) <= c && 4.0 * (x1 - c) * (x1 - c) / (M * M) + y * y / (M * M) <= 1.0) return 0;
        if (vabs_(x) >= c && vabs_(y) <= width) return 0;
        return 1;
    }
    case 34: {                                                            // D_HYPERBOLA 6061
        const double b = M * sqrt(1.0 + x * x / (L * L - M * M));
        if (y > 1.02 * b) return 1;
        return y < 0.98 * b;
    }
    case 50:                                                              // D_LSHAPE 6415
        if (vabs_(x) > L || vabs_(y) > 1.0) return 0;
        return !(x > 0.0 && y > 0.0);
    case 51: {                                                            // D_WAVEGUIDE 6422
        const double x1 = a.xmin + M, x2 = a.xmax - 2.0 * M - 1.5 * L, y1 = 0.5 * L, y2 = 1.5 * L;
        if (x < x1) return 0;
        if (x > x2 + 1.5 * L) return 0;
        if (vabs_(y) > y2) return 0;
        if (x < x2) return vabs_(y) >= y1;
        const double r = module2_(x - x2, y);
        if (r < 0.5 * L) return 0;
        if (r > 1.5 * L) return 0;
        return 1;
    }
    case 522: case 525: {                                                 // D_WAVEGUIDES_COUPLED 6501, _N 6518
        const bool narrow = a.domain == 525;
        const double x1 = narrow ? vabs_(x) : vabs_(x) * 2.0 / a.xmax, y1 = vabs_(y);
        const double width = narrow ? 0.01 : 0.03;
        const double b = 1.0 * (1.0 - 2.0 * M - width), aa = M + 0.5 * width - b;
        double x6 = x1 * x1 * x1;
        const double x5 = x6 * x1 * x1;
        x6 = x5 * x1;
        if (!narrow) {
            const double f = aa + b * (1.0 + 2.0 * x6) / (1.0 + x6);
            const double fp = 6.0 * b * x5 / ((1.0 + x6) * (1.0 + x6));
            return vabs_(f - y1) < M * sqrt(1.0 + fp * fp);
        }
        const double c = 15.0;
        const double f = aa + b * (1.0 + 2.0 * c * x6) / (1.0 + c * x6);
        const double fp = 6.0 * b * c * x5 / ((1.0 + c * x6) * (1.0 + c * x6));
        if (vabs_(f - y1) < 0.9 * M * sqrt(1.0 + fp * fp)) return 1;
        if (vabs_(f - y1) < M * sqrt(1.0 + fp * fp)) return 2;
        return 0;
    }
    case 523: {                                                           // D_WAVEGUIDE_S 6538
        const double aa = 0.5 * (1.0 - M);
        if (x < 0.0) { x = -x; y = -y; }
        if (x < L) {
            if (vabs_(y - 2.0 * aa) < M) return 1;
            if (vabs_(y) < M) return 1;
            return vabs_(y + 2.0 * aa) < M;
        }
        const double r = module2_(x - L, y - aa);
        return vabs_(r - aa) < M;
    }
    case 524: {                                                           // D_WAVEGUIDE_S_SHORT 6556
        const double aa = -1.0 + M;
        if (y < 0.0) { x = -x; y = -y; }
        if (x < -L) {
            if (vabs_(y - 1.0 + M) > M) return 0;
            if (vabs_(y - 1.0 + M) > 0.8 * M) return 2;
            return 1;
        }
        const double r = module2_(x - aa, y);
        if (vabs_(r - 1.0 + M) < 0.8 * M) return 1;
        if (vabs_(r - 1.0 + M) < M) return 2;
        return 0;
    }
    case 61: case 62: case 67: {                                          // D_LENS 6701, _WALL 6708, _ROTATED 6749
        double x1 = x, y1 = y;
        if (a.domain == 67) {  // c0, c1 = cos, sin of APOLY*PID
            x1 = x * a.c0 + y * a.c1;
            y1 = -x * a.c1 + y * a.c0;
        }
        if (vabs_(y1) > M) return 1;
        const double aa = L - 0.25;
        if (x1 > 0) return (x1 + aa) * (x1 + aa) + y1 * y1 > L * L;
        return (x1 - aa) * (x1 - aa) + y1 * y1 > L * L;
    }
    case 63: case 64: {                                                   // D_TWO_LENSES_WALL 6715, _OBSTACLE 6726
        const double width = 0.2, aa = 1.9, b = 1.9 - 2.0 * L + 2.0 * width, x1 = vabs_(x);
        if (a.domain == 63 && x1 < WW && vabs_(y) > M) return 2;
        if (module2_(x1 - aa, y) > L) return 1;
        if (module2_(x1 - b, y) > L) return 1;
        return 0;
    }
    case 68: {                                                            // D_LENS_CONCAVE 6760
        const double width = 0.1, x1 = vabs_(x);
        if (vabs_(y) > M) return 1;
        return (x1 - L - width) * (x1 - L - width) + y * y < L * L;
    }
    case 69: {                                                            // D_LENS_CONVEX_CONCAVE 6767
        if (vabs_(y) > M) return 1;
        const double xshift = 0.6;
        if (x > 0.0) {
            const double x1 = x - xshift, aa = L - 0.25;
            if (x1 > 0) return (x1 + aa) * (x1 + aa) + y * y > L * L;
            return (x1 - aa) * (x1 - aa) + y * y > L * L;
        }
        const double width = 0.1, x1 = vabs_(x + xshift);
        return (x1 - L - width) * (x1 - L - width) + y * y < L * L;
    }
    case 73: {                                                            // D_TREE 6785; c0 = cos(PI/6)
        const double x1 = vabs_(x);
        const double h1 = 1.5 * L, r1 = 1.5 * L, y1 = y - h1;
        if (y1 > 0.95 * r1) return 1;
        const double h2 = 0.25 * L, r2 = 2.0 * L, y2 = y - h2;
        const double h3 = -1.5 * L, r3 = 3.0 * L, y3 = y - h3;
        const double rr[3] = {r1, r2, r3}, yy[3] = {y1, y2, y3};
        const double scale[3] = {1.0, 0.9, 0.8};
        const short val[3] = {0, 1, 2};
        for (int s = 0; s < 3; s++)
            for (int t = 0; t < 3; t++) {
                const double r = rr[t];
                const double px = s ? scale[s] * x1 : x1, py = s ? scale[s] * yy[t] : yy[t];
                if (in_triangle(px, py, r * a.c0, -0.5 * r, 0.0, r, 0.0, -0.5 * r)) return val[s];
            }
        return 1;
    }
    case 74: {                                                            // D_MICHELSON 6867; c0 = sqrt(2)
        if (vabs_(x) > L && vabs_(y) > L) return 2;
        const double r = 0.5 * a.c0, h = a.ymax - 0.5 * WW;
        const double x1 = (x + y) * r, y1 = (-x + y) * r;
        if (vabs_(y1) < 0.5 * WW && vabs_(x1) < L * a.c0) return 1;
        if (vabs_(x - h) < 0.5 * WW && vabs_(y) < L) return 2;
        if (vabs_(x) < L && vabs_(y - h) < 0.5 * WW) return 2;
        return 0;
    }
    case 76: return 1;                                                    // D_GRADIENT_INDEX_LENS 6919
    case 79: {                                                            // D_TWOCIRCLES 6966
        const double x1 = 0.25 * (L + 5.0 * M), x2 = -0.75 * (L + 2.0 * M);
        if (module2_(x - x1, y) < L) return 1;
        if (module2_(x - x2, y) < M) return 1;
        return x < x1 && x > x2 && vabs_(y) < WW;
    }
    case 791: {                                                           // D_POLYCIRCLES 6975
        if (module2_(x, y) < L) return 1;
        for (int k = 0; k < a.npoly; k++) {
            const double ca = gc(a.trig)[2 * k], sa = gc(a.trig)[2 * k + 1];
            double x1 = (L + 2.0 * M) * ca, y1 = (L + 2.0 * M) * sa;
            if (module2_(x - x1, y - y1) < M) return 1;
            x1 = x * ca + y * sa;
            y1 = -x * sa + y * ca;
            if (x1 > 0.0 && x1 < L + 2.0 * M && vabs_(y1) < WW) return 1;
        }
        return 0;
    }
    case 795: {                                                           // D_TWO_ELLIPSES 7052
        const double x1 = vabs_(x);
        return (x1 - L) * (x1 - L) / (M * M) + y * y < 1.0;
    }
    case 796: {                                                           // D_TWO_RECTANGLES 7057
        const double x1 = vabs_(x);
        return vabs_(x1 - L) < M && vabs_(y) < 1.0;
    }
    case 90: {                                                            // D_CARDIOID 7083
        const double y1 = y - M, r = x * x + y1 * y1;
        return r * r + 4.0 * L * y1 * r - 4.0 * L * L * x * x < 0.0;
    }
    case 91: {                                                            // D_NEPHROID 7089 (ipow, sub_wave.c:321)
        const double l4 = L * L * L * L, t = x * x + y * y - 4.0 * L * L;
        return 108 * l4 * x * x > t * t * t;
    }
    default: return domain_value_more(a, x, y);
    }
}

// 8 waves a SIMD (64 VGPRs): the escape loops of the fractal domains hold
// their wave for up to MANDELLEVEL iterations, so occupancy hides their
// latency; the cases of domain_value_more alone would ask for 67
template <bool ESCAPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_domain(WvDomain a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = a.i0 + (int)blockIdx.y;
    if (j >= a.ny || i >= a.nx || i >= a.i0 + a.rows) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    // COMPARISON: one domain for y > 0, the other for y <= 0 (sub_wave.c:7597-7600)
    if ((a.half == 1 && !(y > 0.0)) || (a.half == 2 && y > 0.0)) return;
    gm(a.out)[(size_t)(i - a.i0) * a.ny + j] = ESCAPE ? escape_domain_value(a, x, y) : domain_value(a, x, y);
}

// ---- init_ior_2d's per-cell media models -------------------------------------

// tcc and tgamma of one cell, sub_wave.c:12181-12968 (case lines below);
// false where the reference writes nothing (IOR_TREE off its four values)
__device__ bool ior_cell(const WvIor &a, double x, double y, short m, double &tcc, double &tg)
{
    const double c2 = a.courant * a.courant, cb2 = a.courantb * a.courantb;
    const double G = a.gamma, GB = a.gammab, L = a.lambda, M = a.mu;
    switch (a.model) {
    case 1: {                                                             // IOR_MANDELBROT 12183
        const double norm2 = escape_norm2(0.0, 0.0, x, y, a.mandellevel, a.mandellimit);
        if (norm2 < a.mandellimit) {
            tcc = c2;
        } else {
            double speed = 1.0 + a.mandel_ior_scale * log(1.0 + norm2 / a.mandellimit);
            if (speed < 0.01) speed = 0.01;
            tcc = c2 * speed;
        }
        tg = G;
        return true;
    }
    case 100: {                                                           // IOR_MANDELBROT_LIN 12220
        double u = 0.0, v = 0.0;
        int k = 0;
        const double bound = 1000.0 * a.mandellimit;
        while (k < a.mandellevel && u * u + v * v < bound) {
            const double u1 = u * u - v * v + x;
            v = 2.0 * u * v + y;
            u = u1;
            k++;
        }
        if (k >= a.mandellevel) {
            tcc = c2;
        } else {
            double speed = (double)k / (double)a.mandellevel;
            if (speed < 1.0e-10) speed = 1.0e-10;
            else if (speed > 10.0) speed = 10.0;
            tcc = c2 * speed;
        }
        tg = G;
        return true;
    }
    case 2: {                                                             // IOR_EARTH 12257
        const double r2 = x * x + y * y;
        double c;
        if (r2 > 1.0) c = 0.0;
        else if (r2 < 0.25 * 0.25) c = 0.8 * a.courant;
        else if (r2 < 0.58 * 0.58) c = a.courant * (0.68 - 0.55 * r2);
        else c = a.courant * (1.3 - 0.9 * r2);
        tcc = c;   // the reference stores c itself here
        tg = G;
        return true;
    }
    case 10: case 11: case 12: case 13: {                                 // IOR_LENS_* 12556-12682
        // 10: |x| < 0.05, |y| > MU; 11: |x| < 0.05, |y| < MU; 12, 13: |x| < a0, |y| > MU
        const double w = (a.model == 10 || a.model == 11) ? 0.05 : a.a0;
        const bool wall = vabs_(x) < w && (a.model == 11 ? vabs_(y) < M : vabs_(y) > M);
        if (wall) { tcc = 0.0; tg = G; }
        else if (m != 0) { tcc = c2; tg = G; }
        else { tcc = cb2; tg = GB; }
        return true;
    }
    case 14:                                                              // IOR_TREE 12683
        if (m == 0) { tcc = cb2; tg = GB; return true; }
        if (m == 1 || m == 3) { tcc = c2; tg = G; return true; }
        if (m == 2) { tcc = 0.0; tg = G; return true; }
        return false;
    case 15: case 151: {                                                  // IOR_WAVE_GUIDES_COUPLED 12720, _B 12746
        const bool outside = a.model == 15 ? vabs_(x) > L : (vabs_(x) > L && vabs_(y) > 0.7);
        if (m == 1) { tcc = c2; tg = G; }
        else if (outside) { tcc = 0.0; tg = 1.0; }
        else { tcc = cb2; tg = GB; }
        return true;
    }
    case 16: {                                                            // IOR_WAVE_GUIDE_COATED 12772
        const double cm = 0.8 * a.courant + 0.2 * a.courantb;
        if (m == 1) { tcc = c2; tg = G; }
        else if (m == 2) { tcc = cm * cm; tg = G; }
        else { tcc = cb2; tg = GB; }
        return true;
    }
    case 17:                                                              // IOR_MICHELSON 12800
        if (m == 0) { tcc = c2; tg = G; }
        else if (m == 2) { tcc = 0.0; tg = 1.0; }
        else { tcc = cb2; tg = GB; }
        return true;
    case 18: case 181: {                                                  // IOR_GRADIENT_INDEX_LENS 12823, _B 12851
        const double ax = vabs_(x), ay = vabs_(y);
        if (ax > L) { tcc = c2; tg = G; }
        else if (ay > 1.0) { tcc = 0.0; tg = 0.0; }
        else if (a.model == 18) { tcc = cb2 / (1.0 - M * ay * ay); tg = GB; }
        else {
            double speed = 1.0 / (1.0 - M * ay * ay);
            speed *= speed;
            tcc = cb2 * speed;
            tg = GB;
        }
        return true;
    }
    case 182: {                                                           // IOR_LUNEBURG_LENS 12881
        const double r = module2_(x, y);
        if (r > L) { tcc = c2; tg = G; }
        else { tcc = c2 / (2.0 - r / L); tg = GB; }
        return true;
    }
    case 183: {                                                           // IOR_LUNEBURG_LAYERS 12904
        // (the layer index is only used, and so only formed, inside the lens)
        const double r = module2_(x, y) / L;
        if (r > 1.0) { tcc = c2; tg = G; return true; }
        double r2 = r * (double)a.mdepth;
        r2 = (double)((int)r2) / (double)a.mdepth;
        tcc = c2 / (2.0 - r2);
        tg = GB;
        return true;
    }
    case 19: {                                                            // IOR_LINEAR_X_A 12929
        const double speed = a.courant * (1.0 - x) + a.courantb * x;
        tcc = speed * speed;
        tg = G;
        return true;
    }
    case 191: {                                                           // IOR_LINEAR_X_B 12945
        const double s = (x - a.xmin) / (a.xmax - a.xmin);
        const double speed = a.courant * (1.0 - s) + a.courantb * s;
        tcc = speed * speed;
        tg = G;
        return true;
    }
    default: return false;  // not reached: the launcher admits only the codes above
    }
}

__global__ void __launch_bounds__(256) k_ior(WvIor a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= a.ny || i >= a.nx) return;
    const size_t k = (size_t)i * a.ny + j;
    const short m = gc(a.xy)[k];
    double tcc = 0.0, tc = 0.0, tg = 0.0;
    if (a.model == 0) {
        // VARIABLE_IOR == 0, sub_wave.c:12984-13003
        if (m != 0) {
            tc = a.courant;
            tcc = a.courant * a.courant;
            tg = m == 1 ? a.gamma : a.gammab;
        } else if (a.twospeeds) {
            tc = a.courantb;
            tcc = a.courantb * a.courantb;
            tg = a.gammab;
        }
    } else {
        // ij_to_xy, sub_wave.c:469-476
        const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
        const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
        if (!ior_cell(a, x, y, m, tcc, tg)) tcc = tg = 0.0;
        // sub_wave.c:12970-12981
        if (m != 0) tc = sqrt(tcc);
        else if (a.twospeeds) tc = a.courantb;
    }
    gm(a.tcc)[k] = tcc;
    gm(a.tc)[k] = tc;
    gm(a.tg)[k] = tg;
}

// Fixed device scratch of the allocation-free mask builder: the circle list
// (at most WV_MASK_SCRATCH_MAX_CIRCLES x 4 doubles) and then bands of mask
// rows.  One copy per device, alive until the process ends; every call on a
// device shares it.
static __device__ __align__(256) unsigned char scratch[WV_MASK_SCRATCH_BYTES];

}  // namespace

// device address of the scratch array, looked up once per device with that
// device current (the array's name is a host-side stand-in, not an address)
int wv_mask_scratch(int device, void **ptr)
{
    constexpr int MAXDEV = 64;
    static void *addr[MAXDEV];
    static std::mutex mu;
    if (device < 0 || device >= MAXDEV) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t err = hipGetDevice(&cur);
        if (err == hipSuccess && cur != device) err = hipSetDevice(device);
        void *p = nullptr;
        if (err == hipSuccess) err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (err != hipSuccess) return (int)err;
        addr[device] = p;
    }
    *ptr = addr[device];
    return 0;
}

int wv_launch_pulse(const WvPulse &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.nx);
    hipLaunchKernelGGL(k_pulse, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_mask(const WvMask &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.rows);
    hipLaunchKernelGGL(k_mask, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_domain(const WvDomain &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.rows);
    if (escape_domain(a.domain)) hipLaunchKernelGGL(k_domain<true>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(k_domain<false>, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_ior(const WvIor &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.nx);
    hipLaunchKernelGGL(k_ior, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_wells(const WvWells &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.nx);
    hipLaunchKernelGGL(k_wells, grid, block, 0, s, a);
    return (int)hipGetLastError();
}
