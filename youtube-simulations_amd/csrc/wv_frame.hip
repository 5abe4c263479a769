// wv_frame.hip -- the colours of one frame of the drawing path on the GPU.
//
// draw_wave_highres_palette (wave_common.c:1748-1847) computes, for every
// block cell (i, j) of a size x size grid, the plotted value and colour of
// wave_value (wave_common.c:1053-1171), shades the colours with the value
// gradient (SHADE_2D), fades them and hands them to OpenGL.  Here one thread
// owns one block cell and writes (r, g, b, drawn) as floats, the values
// glColor3f receives, or as the 8-bit colours a frame buffer holds: a frame
// leaves the GPU as 16 or 4 bytes per block instead of the 16 bytes per cell
// of phi and psi.
//
// The running sums of the mean / average / flux plots are read from one
// buffer and written to the other, so a tile can evaluate the blocks of its
// halo (from their old sums and this frame's field) in the same pass without
// racing with their owners.
//
// Arithmetic follows the reference expression by expression (no FMA
// contraction); tanh, exp, log, atan, sqrt and fmod are the device library's,
// which may differ from glibc's in the last bit.
#include "wv_device.h"

using namespace wv;

namespace {

// global_pdes.c:318-327, 354-359, 364-378 and the constants of global_pdes.c:8-10
constexpr int P_AMPLITUDE = 0, P_ENERGY = 1, P_MIXED = 2, P_MEAN_ENERGY = 3, P_LOG_ENERGY = 4,
              P_LOG_MEAN_ENERGY = 5, P_ENERGY_FLUX = 6, P_TOTAL_ENERGY_FLUX = 7, P_AVERAGE_ENERGY = 8;
constexpr int C_LUM = 0, C_HUE = 1, C_ONEDIM = 3, C_ONEDIM_LINEAR = 4, C_BASIC_LINEAR = 5;
constexpr int COL_JET = 0, COL_HSLUV = 1, COL_GRAY = 2, COL_TURBO = 10, COL_PARULA = 16, COL_TWILIGHT = 17,
              COL_TWILIGHT_SHIFTED = 18, COL_TURBO_CYCLIC = 101;
constexpr double REF_PI = 3.141592654, REF_DPI = 6.283185307, REF_PID = 1.570796327;

struct Rgb {
    double c[3];
};

__device__ inline double lut_at(const WvFrame &f, int idx, int c)
{
    idx = idx < 0 ? 0 : (idx > f.lut_n - 1 ? f.lut_n - 1 : idx);
    return (double)gc(f.lut)[3 * idx + c];
}

// colormaps.c:2077-2114
__device__ inline Rgb hsl_jet(double h, double s, double l)
{
    const double c = (1.0 - fabs(2.0 * l - 1.0)) * s;
    const double m = 1.0 * (l - 0.5 * c);
    const double x = c * (1.0 - fabs(fmod(h / 60.0, 2.0) - 1.0));
    if (h >= 0.0 && h < 60.0) return {{c + m, x + m, m}};
    if (h < 120.0) return {{x + m, c + m, m}};
    if (h < 180.0) return {{m, c + m, x + m}};
    if (h < 240.0) return {{m, x + m, c + m}};
    if (h < 300.0) return {{x + m, m, c + m}};
    if (h < 360.0) return {{c + m, m, x + m}};
    return {{m, m, m}};
}

// HSLuv to sRGB (the published HSLuv algorithm, hsluv.org, which the
// reference vendors as hsluv-c: hsluv2rgb, hsluv.c:396-410).  The chroma of
// (L, H) is the HSLuv saturation times the largest chroma inside the sRGB
// gamut there: the nearest of six lines (each channel at 0 and at 1) along
// the hue's ray.  Then CIELUV -> CIEXYZ (D65) -> linear RGB -> sRGB gamma.
// The sRGB matrix, the white point's (u', v') and CIE's kappa and epsilon are
// the library's values; the library's order of operations is kept (sin, cos
// and pow are the device's).
__constant__ const double HS_M[9] = {
    3.24096994190452134377,  -1.53738317757009345794, -0.49861076029300328366,
    -0.96924363628087982613, 1.87596750150772066772,  0.04155505740717561247,
    0.05563007969699360846,  -0.20397695888897656435, 1.05697151424287856072};
constexpr double HS_REF_U = 0.19783000664283680764, HS_REF_V = 0.46831999493879100370;
constexpr double HS_KAPPA = 903.29629629629629629630, HS_EPS = 0.00885645167903563082;
constexpr double HS_DEG = 0.01745329251994329577;

__device__ inline double hs_gamma(double c)
{
    return c <= 0.0031308 ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055;
}

__device__ Rgb hsluv(double h, double s, double l)
{
    double c = 0.0;
    if (!(l > 99.9999999 || l < 0.00000001)) {
        const double tl = l + 16.0, sub1 = (tl * tl * tl) / 1560896.0;
        const double sub2 = sub1 > HS_EPS ? sub1 : (l / HS_KAPPA);
        const double hr = h * HS_DEG, sh = sin(hr), ch = cos(hr);
        double best = 1.7976931348623157e308;
        for (int q = 0; q < 6; q++) {
            const double m1 = HS_M[3 * (q >> 1)], m2 = HS_M[3 * (q >> 1) + 1], m3 = HS_M[3 * (q >> 1) + 2];
            const double t = (double)(q & 1);
            const double top1 = (284517.0 * m1 - 94839.0 * m3) * sub2;
            const double top2 = (838422.0 * m3 + 769860.0 * m2 + 731718.0 * m1) * l * sub2 - 769860.0 * t * l;
            const double bottom = (632260.0 * m3 - 126452.0 * m2) * sub2 + 126452.0 * t;
            const double len = (top2 / bottom) / (sh - (top1 / bottom) * ch);
            if (len >= 0 && len < best) best = len;
        }
        c = best / 100.0 * s;
    }
    if (s < 0.00000001) h = 0.0;
    const double hr = h * HS_DEG, u = cos(hr) * c, v = sin(hr) * c;
    double X = 0.0, Y = 0.0, Z = 0.0;
    if (l > 0.00000001) {
        const double vu = u / (13.0 * l) + HS_REF_U, vv = v / (13.0 * l) + HS_REF_V;
        if (l <= 8.0) {
            Y = l / HS_KAPPA;
        } else {
            const double w = (l + 16.0) / 116.0;
            Y = w * w * w;
        }
        X = -(9.0 * Y * vu) / ((vu - 4.0) * vv - vu * vv);
        Z = (9.0 * Y - (15.0 * vv * Y) - (vv * X)) / (3.0 * vv);
    }
    Rgb r;
    for (int k = 0; k < 3; k++) r.c[k] = hs_gamma(HS_M[3 * k] * X + HS_M[3 * k + 1] * Y + HS_M[3 * k + 2] * Z);
    return r;
}

__device__ inline Rgb lut_rgb(const WvFrame &f, int idx, double mul, bool scaled)
{
    Rgb r;
    for (int c = 0; c < 3; c++) r.c[c] = scaled ? mul * lut_at(f, idx, c) : lut_at(f, idx, c);
    return r;
}

__device__ inline Rgb turbo_blend(const WvFrame &f, double t)
{
    Rgb r;
    for (int c = 0; c < 3; c++) r.c[c] = (1.0 - t) * lut_at(f, 255, c) + t * lut_at(f, 0, c);
    return r;
}

// hsl_to_rgb_palette, colormaps.c:2343-2447 (hue in degrees)
__device__ inline Rgb hsl_palette(const WvFrame &f, double h, double s, double l)
{
    const double ratio = 0.711111111;
    const int pal = f.palette;
    if (pal == COL_JET) return hsl_jet(h, s, l);
    if (pal == COL_HSLUV) {  // colormaps.c:2355-2362
        Rgb r = hsluv(h, 100.0 * s, l);
        for (int c = 0; c < 3; c++) r.c[c] = 2.0 * l * r.c[c] * 100.0;
        return r;
    }
    if (pal == COL_GRAY) {
        const int ch = (int)(h / 360.0);
        return {{(double)ch, (double)ch, (double)ch}};
    }
    const double two_l = 2.0 * l;
    if (pal >= COL_TURBO && pal <= COL_PARULA) return lut_rgb(f, 255 - (int)(ratio * h), two_l, true);
    if (pal == COL_TWILIGHT) {
        int ch = 510 - (int)(2.0 * ratio * h);
        ch = ch > 510 ? 510 : (ch < 0 ? 0 : ch);
        return lut_rgb(f, ch, two_l, true);
    }
    if (pal == COL_TWILIGHT_SHIFTED) {
        int ch = (int)(2.0 * ratio * h) - 255;
        if (ch > 510) ch -= 510;
        else if (ch < 0) ch += 510;
        return lut_rgb(f, ch, two_l, true);
    }
    // COL_TURBO_CYCLIC
    if (h < 330.0) return lut_rgb(f, 255 - (int)(ratio * h * 12.0 / 11.0), two_l, true);
    if (h > 360.0) h = 360.0;
    return turbo_blend(f, (h - 330.0) / 30.0);
}

// amp_to_rgb_palette, colormaps.c:2449-2548 (amplitude in [0, 1])
__device__ inline Rgb amp_palette(const WvFrame &f, double h)
{
    const int pal = f.palette;
    int ch = (int)(256.0 * h);
    if (ch > 255) ch = 255;
    if (pal == COL_JET) return hsl_jet(360.0 * (1.0 - h), 0.9, 0.5);
    if (pal == COL_HSLUV) {  // colormaps.c:2463-2470
        Rgb r = hsluv(360.0 * (1.0 - h), 90.0, 0.5);
        for (int c = 0; c < 3; c++) r.c[c] = r.c[c] * 100.0;
        return r;
    }
    if (pal == COL_GRAY) return {{h, h, h}};
    if (pal >= COL_TURBO && pal <= COL_PARULA) return lut_rgb(f, ch, 1.0, false);
    if (pal == COL_TWILIGHT) {
        ch = (int)(511.0 * h);
        ch = ch > 510 ? 510 : (ch < 0 ? 0 : ch);
        return lut_rgb(f, ch, 1.0, false);
    }
    if (pal == COL_TWILIGHT_SHIFTED) {
        ch = 255 - (int)(511.0 * h);
        if (ch > 510) ch -= 510;
        else if (ch < 0) ch += 510;
        return lut_rgb(f, ch, 1.0, false);
    }
    if (h < 0.9) return lut_rgb(f, (int)(256.0 * h / 0.9), 1.0, false);
    if (h > 1.0) h = 1.0;
    return turbo_blend(f, (h - 0.9) / 0.1);
}

// color_amplitude, colors_waves.c:3-16
__device__ inline double amplitude(const WvFrame &f, double value, double scale)
{
    const double a = tanh(f.slope * value / scale);
    if (f.attenuation == 0.0) return f.color_range == 1.0 ? a : f.color_range * a;
    const double d = exp(-((double)f.time * f.attenuation));
    return f.color_range == 1.0 ? a * d : f.color_range * a * d;
}

// color_scheme_palette (colors_waves.c:161-218), color_scheme_asym_palette (220-290)
__device__ inline Rgb colour(const WvFrame &f, int scheme, bool asym, double value, double scale, int time)
{
    double hue, y, amp;
    switch (scheme) {
    case C_LUM:
        hue = f.colorhue + (double)time * f.colordrift / (double)f.nsteps;
        if (hue < 0.0) hue += 360.0;
        if (hue >= 360.0) hue -= 360.0;
        amp = amplitude(f, value, scale);
        y = f.lummean + amp * f.lumamp;
        y -= (double)(int)y;
        return hsl_palette(f, hue, 0.9, y);
    case C_HUE:
        if (asym) {
            amp = 2.0 * tanh(f.slope * value / scale) * exp(-((double)time * f.attenuation)) - 1.0;
            hue = f.huemean + 0.8 * amp * f.hueamp;
        } else {
            amp = amplitude(f, value, scale);
            hue = f.huemean + amp * f.hueamp;
        }
        if (hue < 0.0) hue += 360.0;
        if (hue >= 360.0) hue -= 360.0;
        return hsl_palette(f, hue, 0.9, 0.5);
    case C_ONEDIM:
        amp = amplitude(f, value, scale);
        return amp_palette(f, asym ? amp : 0.5 * (1.0 + amp));
    case C_ONEDIM_LINEAR:
        amp = f.slope * value / scale;
        if (amp > 1.0) amp -= 1.0;
        else if (amp < 0.0) amp += 1.0;
        return amp_palette(f, amp);
    default:  // C_BASIC_LINEAR
        if (value > 1.0) value -= 1.0;
        else if (value < 0.0) value += 1.0;
        return amp_palette(f, value);
    }
}

// phi and psi of a cell in either form
__device__ inline double phi_at(const WvFrame &f, int i, int j)
{
    const size_t k = (size_t)i * f.ldy + j;
    if (f.phi) return gc(f.phi)[k];
    return (double)gc(f.hi)[k] + (double)gc(f.lo)[k];
}
__device__ inline double psi_at(const WvFrame &f, int i, int j, double phi)
{
    const size_t k = (size_t)i * f.ldy + j;
    if (f.psi) return gc(f.psi)[k];
    return phi - (double)gc(f.v)[k];
}
__device__ inline bool nz_at(const WvFrame &f, int i, int j)
{
    return (gc(f.nz)[(size_t)i * f.ldw + (j >> 6)] >> (j & 63)) & 1u;
}

// compute_energy, wave_common.c:453-472
__device__ inline double energy(const WvFrame &f, int i, int j, double x, double y)
{
    const int ip = i + 1 == f.nx ? f.nx - 1 : i + 1, im = i == 0 ? 0 : i - 1;
    const int jp = j + 1 == f.ny ? f.ny - 1 : j + 1, jm = j == 0 ? 0 : j - 1;
    const double e = phi_at(f, ip, j), w = phi_at(f, im, j), n = phi_at(f, i, jp), s = phi_at(f, i, jm);
    const double v = x - y;
    const double gx2 = (e - x) * (e - x) + (x - w) * (x - w);
    const double gy2 = (n - x) * (n - x) + (x - s) * (x - s);
    if (nz_at(f, i, j)) return f.es * f.es * (v * v + 0.5 * f.courant * f.courant * (gx2 + gy2));
    if (f.twospeeds) return f.es * f.es * (v * v + 0.5 * f.courantb * f.courantb * (gx2 + gy2));
    return 0.0;
}

// argument, sub_wave.c:287-304
__device__ inline double argument(double x, double y)
{
    double a;
    if (x != 0.0) {
        a = atan(y / x);
        if (x < 0.0) a += REF_PI;
    } else {
        a = REF_PID;
        if (y < 0.0) a = REF_PI * 1.5;
    }
    return a;
}

// The value and colour of block (bi, bj), wave_value (wave_common.c:1053-1171),
// from the old running sums; the new sums go to te / ae / tf.
template <int PLOT>
__device__ __forceinline__ double wave_value(const WvFrame &f, int bi, int bj, Rgb &rgb, double &te, double &ae,
                                             double &tfx, double &tfy)
{
    const int i = bi * f.size, j = bj * f.size;
    const size_t b = (size_t)bi * f.nby + bj;
    const double x = phi_at(f, i, j), y = psi_at(f, i, j, x);
    const double sc = f.scale;
    const int t = f.time;
    double value, e;
    switch (PLOT) {
    case P_AMPLITUDE:
        value = x;
        rgb = colour(f, f.scheme, false, f.vshift + f.vscale * value, sc, t);
        break;
    case P_ENERGY:
        value = energy(f, i, j, x, y);
        rgb = colour(f, f.scheme, f.asym, value, sc, t);
        break;
    case P_MIXED:
        value = j > f.ny / 2 ? x : energy(f, i, j, x, y);
        rgb = colour(f, f.scheme, false, value, sc, t);
        break;
    case P_MEAN_ENERGY:
        te = gc(f.te0)[b] + energy(f, i, j, x, y);
        value = te / (double)(t + 1);
        rgb = colour(f, f.scheme, f.asym, value, sc, t);
        break;
    case P_LOG_ENERGY:
        e = energy(f, i, j, x, y);
        if (e < 1.0e-20) e = 1.0e-20;
        value = f.log_shift + f.log_scale * log(e);
        rgb = colour(f, f.scheme, false, value, sc, t);
        break;
    case P_LOG_MEAN_ENERGY:
        e = energy(f, i, j, x, y);
        if (e == 0.0) e = 1.0e-20;
        te = gc(f.te0)[b] + e;
        value = f.log_shift + f.log_scale * log(te / (double)(t + 1));
        rgb = colour(f, f.scheme, false, value, sc, t);
        break;
    case P_ENERGY_FLUX:
    case P_TOTAL_ENERGY_FLUX: {
        // compute_energy_flux, wave_common.c:475-512
        const double vel = (x - y) < 0.0 ? -(x - y) : (x - y);
        const int ip = i + 1 == f.nx ? f.nx - 1 : i + 1, im = i == 0 ? 0 : i - 1;
        const int jp = j + 1 == f.ny ? f.ny - 1 : j + 1, jm = j == 0 ? 0 : j - 1;
        const double gx = phi_at(f, ip, j) - phi_at(f, im, j);
        const double gy = phi_at(f, i, jp) - phi_at(f, i, jm);
        const bool on = nz_at(f, i, j) || f.twospeeds;
        const double mod = on ? vel * sqrt(gx * gx + gy * gy) : 0.0;
        if (PLOT == P_ENERGY_FLUX) {
            value = mod * f.flux_scale;
            rgb = colour(f, f.scheme, true, value, sc, t);
            break;
        }
        tfx = gc(f.tf0)[2 * b] * 0.99;
        tfy = gc(f.tf0)[2 * b + 1] * 0.99;
        tfx = tfx + (on ? vel * gx : 0.0);
        tfy = tfy + (on ? vel * gy : 0.0);
        const double m = sqrt(tfx * tfx + tfy * tfy);
        double arg = argument(tfx, tfy);
        if (arg < 0.0) arg += REF_DPI;
        value = arg / REF_DPI;
        rgb = colour(f, C_ONEDIM_LINEAR, false, value, 1.0, 1);
        const double ff = tanh(m * f.flux_scale);
        for (int c = 0; c < 3; c++) rgb.c[c] *= ff;
        break;
    }
    default:  // P_AVERAGE_ENERGY, P_LOG_AVERAGE_ENERGY
        e = energy(f, i, j, x, y);
        ae = f.avrg * gc(f.ae0)[b] + (1.0 - f.avrg) * e;
        if (PLOT == P_AVERAGE_ENERGY) {
            value = ae;
            rgb = colour(f, f.scheme, f.asym, value, sc, t);
        } else {
            value = ae == 0.0 ? f.log_shift - 10.0 * f.log_scale : f.log_shift + f.log_scale * log(ae);
            rgb = colour(f, f.scheme, false, value, sc, t);
        }
        break;
    }
    if (f.rescale) {
        const double c = gc(f.cs)[(size_t)i * f.ldy + j];
        value *= c;
        for (int q = 0; q < 3; q++) rgb.c[q] = 1.0 + c * (rgb.c[q] - 1.0);
    }
    return value;
}

// A workgroup owns a tile of TY x (TX*NC) block cells, each thread NC
// consecutive cells of one block row: the thread computes all of its values
// and colours before it stores anything, so their loads are in flight
// together (one cell per thread leaves too few bytes in flight to stream at
// HBM rate).  Values go to LDS; the shading gradient reads the next block
// row / column from there, and only the tile's halo (the row below, the
// column to the right) is evaluated a second time, by the tile that needs
// it, from the old running sums.
constexpr int TX = 64, TY = 4, NC = 4, TW = TX * NC;

template <int PLOT>
__global__ void __launch_bounds__(TX * TY) k_frame(const WvFrame f)
{
    __shared__ double sv[TY + 1][TW + 1];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int bi = blockIdx.y * TY + ty, bj0 = blockIdx.x * TW + tx * NC;
    Rgb rgb[NC];
    double val[NC], te[NC], ae[NC], tfx[NC], tfy[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        rgb[c] = Rgb{};
        val[c] = te[c] = ae[c] = tfx[c] = tfy[c] = 0.0;
        if (bi < f.nbx && bj0 + c < f.nby) val[c] = wave_value<PLOT>(f, bi, bj0 + c, rgb[c], te[c], ae[c], tfx[c], tfy[c]);
    }
    double hcol = 0.0;
    if (f.shade) {
        // halo: row TY of the tile (threads of row 0), column TW (thread 0 of each row)
        Rgb dummy;
        double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
        const int hbi = blockIdx.y * TY + TY, hbj = blockIdx.x * TW + TW;
        if (ty == 0 && hbi < f.nbx)
            for (int c = 0; c < NC; c++)
                if (bj0 + c < f.nby) sv[TY][tx * NC + c] = wave_value<PLOT>(f, hbi, bj0 + c, dummy, d0, d1, d2, d3);
        if (tx == 0 && bi < f.nbx && hbj < f.nby) hcol = wave_value<PLOT>(f, bi, hbj, dummy, d0, d1, d2, d3);
    }
#pragma unroll
    for (int c = 0; c < NC; c++) sv[ty][tx * NC + c] = val[c];
    if (tx == 0) sv[ty][TW] = hcol;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int bj = bj0 + c;
        if (bi >= f.nbx || bj >= f.nby) continue;
        const size_t b = (size_t)bi * f.nby + bj;
        if (PLOT == P_MEAN_ENERGY || PLOT == P_LOG_MEAN_ENERGY) gm(f.te1)[b] = te[c];
        if (PLOT >= P_AVERAGE_ENERGY) gm(f.ae1)[b] = ae[c];
        if (PLOT == P_TOTAL_ENERGY_FLUX) {
            gm(f.tf1)[2 * b] = tfx[c];
            gm(f.tf1)[2 * b + 1] = tfy[c];
        }
        const int i = bi * f.size, j = bj * f.size;
        // SHADE_2D, wave_common.c:1772-1791: gradient towards the next block cells
        if (f.shade && i < f.nx - f.size && j < f.ny - f.size) {
            const double gx = sv[ty + 1][tx * NC + c] - val[c], gy = sv[ty][tx * NC + c + 1] - val[c];
            const double norm = sqrt(f.shade_scale * f.shade_scale + gx * gx + gy * gy);
            const double pscal = -gx * f.light_x - gy * f.light_y + f.shade_scale;
            double ca = pscal / norm;
            ca = (ca + 1.0) * 0.4 + 0.2;
            for (int q = 0; q < 3; q++) {
                if (f.rescale)
                    rgb[c].c[q] =
                        1.0 + (ca * rgb[c].c[q] - 1.0) * (gc(f.cs)[(size_t)i * f.ldy + j] + 0.01) * (1.0 / 1.01);
                else
                    rgb[c].c[q] *= ca;
            }
        }
        // drawing, wave_common.c:1797-1820
        const bool drawn = f.twospeeds || nz_at(f, i, j);
        if (drawn) {
            if (f.fadeobs) {
                const double fd = gc(f.ft)[(size_t)i * f.ldy + j];
                for (int q = 0; q < 3; q++) rgb[c].c[q] *= fd;
            }
            if (f.fade)
                for (int q = 0; q < 3; q++) rgb[c].c[q] *= f.fade_value;
        }
        const float r = (float)rgb[c].c[0], g = (float)rgb[c].c[1], bl = (float)rgb[c].c[2];
        if (f.out8) {
            // an 8-bit colour buffer: clamp to [0, 1], scale by 255, round to nearest
            auto q = [](float x) { return (unsigned char)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f + 0.5f); };
            f.out8[b] = make_uchar4(q(r), q(g), q(bl), drawn ? 255 : 0);
        } else {
            f.out[b] = make_float4(r, g, bl, drawn ? 1.0f : 0.0f);
        }
    }
}

}  // namespace

int wv_launch_frame(const WvFrame &f, hipStream_t s)
{
    const dim3 block(TX, TY), grid((f.nby + TW - 1) / TW, (f.nbx + TY - 1) / TY);
    switch (f.plot) {
#define WV_PLOT(P) case P: hipLaunchKernelGGL(k_frame<P>, grid, block, 0, s, f); break;
    WV_PLOT(0) WV_PLOT(1) WV_PLOT(2) WV_PLOT(3) WV_PLOT(4) WV_PLOT(5) WV_PLOT(6) WV_PLOT(7) WV_PLOT(8)
    WV_PLOT(9)
#undef WV_PLOT
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}
