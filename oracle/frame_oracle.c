/*
 * frame_oracle.c -- CPU oracle of the frame colours (test infrastructure, see
 * wave_oracle.h).  Same flags as wave_oracle.c (no FMA contraction), so
 * every operation below is one binary64 operation in source order; libm
 * (tanh, exp, log, atan, sqrt, fmod) is glibc's, as for the reference.
 */
#include <math.h>
#include <stdlib.h>

#include "wave_oracle.h"

/* plot types, global_pdes.c:318-327 */
enum { P_AMPLITUDE, P_ENERGY, P_MIXED, P_MEAN_ENERGY, P_LOG_ENERGY, P_LOG_MEAN_ENERGY, P_ENERGY_FLUX,
       P_TOTAL_ENERGY_FLUX, P_AVERAGE_ENERGY, P_LOG_AVERAGE_ENERGY };
/* colour schemes, global_pdes.c:354-359 */
enum { C_LUM = 0, C_HUE = 1, C_ONEDIM = 3, C_ONEDIM_LINEAR = 4, C_BASIC_LINEAR = 5 };
/* palettes, global_pdes.c:364-378 */
enum { COL_JET = 0, COL_HSLUV = 1, COL_GRAY = 2, COL_TURBO = 10, COL_PARULA = 16, COL_TWILIGHT = 17,
       COL_TWILIGHT_SHIFTED = 18, COL_TURBO_CYCLIC = 101 };

/* the reference's own constants, global_pdes.c:8-10 */
#define REF_PI 3.141592654
#define REF_DPI 6.283185307
#define REF_PID 1.570796327

static double lut_at(const float *lut, int n, int idx, int c)
{
    if (idx < 0) idx = 0;
    if (idx > n - 1) idx = n - 1;
    return (double)lut[3 * idx + c];
}

/* colormaps.c:2077-2114 */
static void hsl_jet(double h, double s, double l, double rgb[3])
{
    double c = (1.0 - fabs(2.0 * l - 1.0)) * s;
    double m = 1.0 * (l - 0.5 * c);
    double x = c * (1.0 - fabs(fmod(h / 60.0, 2) - 1.0));
    if (h >= 0.0 && h < 60.0) { rgb[0] = c + m; rgb[1] = x + m; rgb[2] = m; }
    else if (h < 120.0) { rgb[0] = x + m; rgb[1] = c + m; rgb[2] = m; }
    else if (h < 180.0) { rgb[0] = m; rgb[1] = c + m; rgb[2] = x + m; }
    else if (h < 240.0) { rgb[0] = m; rgb[1] = x + m; rgb[2] = c + m; }
    else if (h < 300.0) { rgb[0] = x + m; rgb[1] = m; rgb[2] = c + m; }
    else if (h < 360.0) { rgb[0] = c + m; rgb[1] = m; rgb[2] = x + m; }
    else { rgb[0] = m; rgb[1] = m; rgb[2] = m; }
}

/* HSLuv to sRGB, the published HSLuv algorithm (hsluv.org) that the
 * reference vendors as hsluv-c (hsluv.c:396-410): HSLuv -> LCh with the
 * chroma scaled to the largest one inside the sRGB gamut at that lightness
 * and hue -> CIELUV -> CIEXYZ (D65) -> sRGB.  The sRGB matrix, the white
 * point's (u', v') and CIE's kappa and epsilon are the library's values
 * (hsluv.c:43-58); every operation is in the library's order, so results
 * are the library's bit for bit. */
static const double HS_M[3][3] = {
    {3.24096994190452134377, -1.53738317757009345794, -0.49861076029300328366},
    {-0.96924363628087982613, 1.87596750150772066772, 0.04155505740717561247},
    {0.05563007969699360846, -0.20397695888897656435, 1.05697151424287856072}};
#define HS_REF_U 0.19783000664283680764
#define HS_REF_V 0.46831999493879100370
#define HS_KAPPA 903.29629629629629629630
#define HS_EPS 0.00885645167903563082
#define HS_DEG 0.01745329251994329577

/* the shortest non-negative distance from the L axis along hue h to the six
 * gamut lines (one per channel at 0 and at 1) of lightness l */
static double hs_max_chroma(double l, double h)
{
    const double tl = l + 16.0, sub1 = (tl * tl * tl) / 1560896.0;
    const double sub2 = sub1 > HS_EPS ? sub1 : (l / HS_KAPPA);
    const double hrad = h * HS_DEG, sh = sin(hrad), chh = cos(hrad);
    double best = 1.7976931348623157e308;
    for (int ch = 0; ch < 3; ch++)
        for (int t = 0; t < 2; t++) {
            const double m1 = HS_M[ch][0], m2 = HS_M[ch][1], m3 = HS_M[ch][2];
            const double top1 = (284517.0 * m1 - 94839.0 * m3) * sub2;
            const double top2 = (838422.0 * m3 + 769860.0 * m2 + 731718.0 * m1) * l * sub2 - 769860.0 * t * l;
            const double bottom = (632260.0 * m3 - 126452.0 * m2) * sub2 + 126452.0 * t;
            const double slope = top1 / bottom, icpt = top2 / bottom;
            const double len = icpt / (sh - slope * chh);
            if (len >= 0 && len < best) best = len;
        }
    return best;
}

static double hs_gamma(double c)
{
    return c <= 0.0031308 ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055;
}

static void hsluv_rgb(double h, double s, double l, double rgb[3])
{
    /* HSLuv -> LCh */
    const double c = (l > 99.9999999 || l < 0.00000001) ? 0.0 : hs_max_chroma(l, h) / 100.0 * s;
    if (s < 0.00000001) h = 0.0;
    /* LCh -> Luv */
    const double hrad = h * HS_DEG, u = cos(hrad) * c, v = sin(hrad) * c;
    /* Luv -> XYZ */
    double X = 0.0, Y = 0.0, Z = 0.0;
    if (l > 0.00000001) {
        const double vu = u / (13.0 * l) + HS_REF_U, vv = v / (13.0 * l) + HS_REF_V;
        if (l <= 8.0) Y = l / HS_KAPPA;
        else {
            const double q = (l + 16.0) / 116.0;
            Y = q * q * q;
        }
        X = -(9.0 * Y * vu) / ((vu - 4.0) * vv - vu * vv);
        Z = (9.0 * Y - (15.0 * vv * Y) - (vv * X)) / (3.0 * vv);
    }
    /* XYZ -> sRGB */
    for (int k = 0; k < 3; k++) rgb[k] = hs_gamma(HS_M[k][0] * X + HS_M[k][1] * Y + HS_M[k][2] * Z);
}

/* colormaps.c:2343-2447 (hue in degrees) */
static void hsl_palette(double h, double s, double l, double rgb[3], int palette, const float *lut, int n)
{
    const double ratio = 0.711111111;
    int ch;
    if (palette == COL_JET) {
        hsl_jet(h, s, l, rgb);
    } else if (palette == COL_HSLUV) {  /* colormaps.c:2355-2362 */
        double q[3];
        hsluv_rgb(h, 100.0 * s, l, q);
        for (int c = 0; c < 3; c++) rgb[c] = 2.0 * l * q[c] * 100.0;
    } else if (palette == COL_GRAY) {
        ch = h / 360.0;
        rgb[0] = rgb[1] = rgb[2] = ch;
    } else if (palette >= COL_TURBO && palette <= COL_PARULA) {
        ch = 255 - (int)(ratio * h);
        for (int c = 0; c < 3; c++) rgb[c] = 2.0 * l * lut_at(lut, n, ch, c);
    } else if (palette == COL_TWILIGHT) {
        ch = 510 - (int)(2.0 * ratio * h);
        if (ch > 510) ch = 510;
        else if (ch < 0) ch = 0;
        for (int c = 0; c < 3; c++) rgb[c] = 2.0 * l * lut_at(lut, n, ch, c);
    } else if (palette == COL_TWILIGHT_SHIFTED) {
        ch = (int)(2.0 * ratio * h) - 255;
        if (ch > 510) ch -= 510;
        else if (ch < 0) ch += 510;
        for (int c = 0; c < 3; c++) rgb[c] = 2.0 * l * lut_at(lut, n, ch, c);
    } else if (palette == COL_TURBO_CYCLIC) {
        if (h < 330.0) {
            ch = 255 - (int)(ratio * h * 12.0 / 11.0);
            for (int c = 0; c < 3; c++) rgb[c] = 2.0 * l * lut_at(lut, n, ch, c);
        } else {
            if (h > 360.0) h = 360.0;
            const double t = (h - 330.0) / 30.0;
            for (int c = 0; c < 3; c++) rgb[c] = (1.0 - t) * lut_at(lut, n, 255, c) + t * lut_at(lut, n, 0, c);
        }
    }
}

/* colormaps.c:2449-2548 (amplitude in [0, 1]) */
static void amp_palette(double h, double rgb[3], int palette, const float *lut, int n)
{
    int ch = (int)(256.0 * h);
    if (ch > 255) ch = 255;
    if (palette == COL_JET) {
        hsl_jet(360.0 * (1.0 - h), 0.9, 0.5, rgb);
    } else if (palette == COL_HSLUV) {  /* colormaps.c:2463-2470 */
        double q[3];
        hsluv_rgb(360.0 * (1.0 - h), 90.0, 0.5, q);
        for (int c = 0; c < 3; c++) rgb[c] = q[c] * 100.0;
    } else if (palette == COL_GRAY) {
        rgb[0] = rgb[1] = rgb[2] = h;
    } else if (palette >= COL_TURBO && palette <= COL_PARULA) {
        for (int c = 0; c < 3; c++) rgb[c] = lut_at(lut, n, ch, c);
    } else if (palette == COL_TWILIGHT) {
        ch = (int)(511.0 * h);
        if (ch > 510) ch = 510;
        else if (ch < 0) ch = 0;
        for (int c = 0; c < 3; c++) rgb[c] = lut_at(lut, n, ch, c);
    } else if (palette == COL_TWILIGHT_SHIFTED) {
        ch = 255 - (int)(511.0 * h);
        if (ch > 510) ch -= 510;
        else if (ch < 0) ch += 510;
        for (int c = 0; c < 3; c++) rgb[c] = lut_at(lut, n, ch, c);
    } else if (palette == COL_TURBO_CYCLIC) {
        if (h < 0.9) {
            ch = (int)(256.0 * h / 0.9);
            for (int c = 0; c < 3; c++) rgb[c] = lut_at(lut, n, ch, c);
        } else {
            if (h > 1.0) h = 1.0;
            const double t = (h - 0.9) / 0.1;
            for (int c = 0; c < 3; c++) rgb[c] = (1.0 - t) * lut_at(lut, n, 255, c) + t * lut_at(lut, n, 0, c);
        }
    }
}

/* colors_waves.c:3-27 */
static double amplitude(const orc_frame_params *f, double value, double scale, int time)
{
    if (f->attenuation == 0.0) {
        if (f->color_range == 1.0) return tanh(f->slope * value / scale);
        return f->color_range * tanh(f->slope * value / scale);
    }
    if (f->color_range == 1.0) return tanh(f->slope * value / scale) * exp(-((double)time * f->attenuation));
    return f->color_range * tanh(f->slope * value / scale) * exp(-((double)time * f->attenuation));
}

static double amplitude_asym(const orc_frame_params *f, double value, double scale, int time)
{
    return 2.0 * tanh(f->slope * value / scale) * exp(-((double)time * f->attenuation)) - 1.0;
}

/* color_scheme_palette (colors_waves.c:161-218) and color_scheme_asym_palette (220-290) */
void orc_color(const orc_frame_params *f, int scheme, int asym, double value, double scale, int time,
               const float *lut, int n, double rgb[3])
{
    const int pal = f->palette;
    double hue, y, amp;
    switch (scheme) {
    case C_LUM:
        hue = f->colorhue + (double)time * f->colordrift / (double)f->nsteps;
        if (hue < 0.0) hue += 360.0;
        if (hue >= 360.0) hue -= 360.0;
        amp = amplitude(f, value, scale, time);
        y = f->lummean + amp * f->lumamp;
        y -= (double)(int)y;
        hsl_palette(hue, 0.9, y, rgb, pal, lut, n);
        break;
    case C_HUE:
        if (asym) {
            amp = amplitude_asym(f, value, scale, time);
            hue = f->huemean + 0.8 * amp * f->hueamp;
        } else {
            amp = amplitude(f, value, scale, time);
            hue = f->huemean + amp * f->hueamp;
        }
        if (hue < 0.0) hue += 360.0;
        if (hue >= 360.0) hue -= 360.0;
        hsl_palette(hue, 0.9, 0.5, rgb, pal, lut, n);
        break;
    case C_ONEDIM:
        amp = amplitude(f, value, scale, time);
        amp_palette(asym ? amp : 0.5 * (1.0 + amp), rgb, pal, lut, n);
        break;
    case C_ONEDIM_LINEAR:
        amp = f->slope * value / scale;
        if (amp > 1.0) amp -= 1.0;
        else if (amp < 0.0) amp += 1.0;
        amp_palette(amp, rgb, pal, lut, n);
        break;
    case C_BASIC_LINEAR:
        if (value > 1.0) value -= 1.0;
        else if (value < 0.0) value += 1.0;
        amp_palette(value, rgb, pal, lut, n);
        break;
    }
}

/* sub_wave.c:270-305 */
static double ref_vabs(double x) { return x < 0.0 ? -x : x; }
static double module2(double x, double y) { return sqrt(x * x + y * y); }
static double argument(double x, double y)
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

/* compute_energy_flux, wave_common.c:475-512 */
static void energy_flux(const orc_params *p, const double *phi, const double *psi, const short *xy, int i, int j,
                        double *gx, double *gy, double *arg, double *mod)
{
    const int NX = p->nx, NY = p->ny;
    const double v = ref_vabs(phi[(size_t)i * NY + j] - psi[(size_t)i * NY + j]);
    const int ip = i + 1 == NX ? NX - 1 : i + 1, im = i == 0 ? 0 : i - 1;
    const int jp = j + 1 == NY ? NY - 1 : j + 1, jm = j == 0 ? 0 : j - 1;
    const double ggx = phi[(size_t)ip * NY + j] - phi[(size_t)im * NY + j];
    const double ggy = phi[(size_t)i * NY + jp] - phi[(size_t)i * NY + jm];
    *arg = argument(ggx, ggy);
    if (*arg < 0.0) *arg += REF_DPI;
    if (*arg > REF_DPI) *arg -= REF_DPI;
    if (xy[(size_t)i * NY + j] || p->twospeeds) {
        *mod = v * module2(ggx, ggy);
        *gx = v * ggx;
        *gy = v * ggy;
    } else {
        *mod = *gx = *gy = 0.0;
    }
}

/* wave_value, wave_common.c:1053-1171, at cell (i, j) */
static double wave_value(const orc_params *p, const orc_frame_params *f, const double *phi, const double *psi,
                         const short *xy, double *te, double *ae, double *tf, const double *cs,
                         const float *lut, int n, int i, int j, double rgb[3])
{
    const int NX = p->nx, NY = p->ny;
    const size_t k = (size_t)i * NY + j;
    const double sc = f->scale;
    const int t = f->time;
    double value, e, gx, gy, arg, mod;
#define ENERGY() orc_energy_cell(p, phi, psi, xy, f->courant, f->courantb, f->e_scale, i, j)
    switch (f->plot) {
    case P_AMPLITUDE:
        value = phi[k];
        orc_color(f, f->scheme, 0, f->vshift_amplitude + f->vscale_amplitude * value, sc, t, lut, n, rgb);
        break;
    case P_ENERGY:
        value = ENERGY();
        orc_color(f, f->scheme, f->asym, value, sc, t, lut, n, rgb);
        break;
    case P_MIXED:
        value = j > NY / 2 ? phi[k] : ENERGY();
        orc_color(f, f->scheme, 0, value, sc, t, lut, n, rgb);
        break;
    case P_MEAN_ENERGY:
        te[k] += ENERGY();
        value = te[k] / (double)(t + 1);
        orc_color(f, f->scheme, f->asym, value, sc, t, lut, n, rgb);
        break;
    case P_LOG_ENERGY:
        e = ENERGY();
        if (e < 1.0e-20) e = 1.0e-20;
        value = f->log_shift + f->log_scale * log(e);
        orc_color(f, f->scheme, 0, value, sc, t, lut, n, rgb);
        break;
    case P_LOG_MEAN_ENERGY:
        e = ENERGY();
        if (e == 0.0) e = 1.0e-20;
        te[k] += e;
        value = f->log_shift + f->log_scale * log(te[k] / (double)(t + 1));
        orc_color(f, f->scheme, 0, value, sc, t, lut, n, rgb);
        break;
    case P_ENERGY_FLUX:
        energy_flux(p, phi, psi, xy, i, j, &gx, &gy, &arg, &mod);
        value = mod * f->flux_scale;
        orc_color(f, f->scheme, 1, value, sc, t, lut, n, rgb);
        break;
    case P_TOTAL_ENERGY_FLUX: {
        energy_flux(p, phi, psi, xy, i, j, &gx, &gy, &arg, &mod);
        const size_t q = 2 * (size_t)j * NX + 2 * (size_t)i;
        tf[q] *= 0.99;
        tf[q + 1] *= 0.99;
        tf[q] += gx;
        tf[q + 1] += gy;
        mod = module2(tf[q], tf[q + 1]);
        arg = argument(tf[q], tf[q + 1]);
        if (arg < 0.0) arg += REF_DPI;
        value = arg / REF_DPI;
        orc_color(f, C_ONEDIM_LINEAR, 0, value, 1.0, 1, lut, n, rgb);
        const double ff = tanh(mod * f->flux_scale);
        for (int c = 0; c < 3; c++) rgb[c] *= ff;
        break;
    }
    case P_AVERAGE_ENERGY:
        e = ENERGY();
        ae[k] = f->avrg_e_factor * ae[k] + (1.0 - f->avrg_e_factor) * e;
        value = ae[k];
        orc_color(f, f->scheme, f->asym, value, sc, t, lut, n, rgb);
        break;
    default: /* P_LOG_AVERAGE_ENERGY */
        e = ENERGY();
        ae[k] = f->avrg_e_factor * ae[k] + (1.0 - f->avrg_e_factor) * e;
        if (ae[k] == 0.0) value = f->log_shift - 10.0 * f->log_scale;
        else value = f->log_shift + f->log_scale * log(ae[k]);
        orc_color(f, f->scheme, 0, value, sc, t, lut, n, rgb);
        break;
    }
#undef ENERGY
    if (f->rescale_center) {
        value *= cs[k];
        for (int c = 0; c < 3; c++) rgb[c] = 1.0 + cs[k] * (rgb[c] - 1.0);
    }
    return value;
}

void orc_frame(const orc_params *p, const orc_frame_params *f, const double *phi, const double *psi,
               const short *xy, double *te, double *ae, double *tf, const double *cs, const double *fade_table,
               const float *lut, int n, double *values, double *rgb, unsigned char *drawn)
{
    const int NX = p->nx, NY = p->ny, s = f->size;
    const size_t NC = (size_t)NX * NY;
    double *rv = (double *)malloc(sizeof(double) * 3 * NC);
    /* values and colours, wave_common.c:1762-1770 */
    for (int i = 0; i < NX; i += s)
        for (int j = 0; j < NY; j += s) {
            double c[3] = {0.0, 0.0, 0.0};
            const size_t k = (size_t)i * NY + j;
            values[k] = wave_value(p, f, phi, psi, xy, te, ae, tf, cs, lut, n, i, j, c);
            for (int q = 0; q < 3; q++) rv[q * NC + k] = c[q];
        }
    /* SHADE_2D, wave_common.c:1772-1791 (gradient between block cells) */
    if (f->shade) {
        const double vscale2 = f->shade_scale * f->shade_scale, col_ratio = 1.0 / 1.01;
        for (int i = 0; i < NX - s; i += s)
            for (int j = 0; j < NY - s; j += s) {
                const size_t k = (size_t)i * NY + j;
                const double gx = values[k + (size_t)s * NY] - values[k];
                const double gy = values[k + s] - values[k];
                const double norm = sqrt(vscale2 + gx * gx + gy * gy);
                const double pscal = -gx * f->light_x - gy * f->light_y + f->shade_scale;
                double ca = pscal / norm;
                ca = (ca + 1.0) * 0.4 + 0.2;
                for (int q = 0; q < 3; q++) {
                    if (f->rescale_center)
                        rv[q * NC + k] = 1.0 + (ca * rv[q * NC + k] - 1.0) * (cs[k] + 0.01) * col_ratio;
                    else
                        rv[q * NC + k] *= ca;
                }
            }
    }
    /* drawing, wave_common.c:1797-1820 */
    for (int i = 0; i < NX; i += s)
        for (int j = 0; j < NY; j += s) {
            const size_t k = (size_t)i * NY + j;
            drawn[k] = p->twospeeds || xy[k];
            if (drawn[k]) {
                if (f->fade_in_obstacle)
                    for (int q = 0; q < 3; q++) rv[q * NC + k] *= fade_table[k];
                if (f->fade)
                    for (int q = 0; q < 3; q++) rv[q * NC + k] *= f->fade_value;
            }
            for (int q = 0; q < 3; q++) rgb[3 * k + q] = rv[q * NC + k];
        }
    free(rv);
}

void orc_hsluv_rgb(double h, double s, double l, double rgb[3]) { hsluv_rgb(h, s, l, rgb); }
