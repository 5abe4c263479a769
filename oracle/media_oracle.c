/*
 * media_oracle.c -- CPU oracle of init_ior_2d's media tables (Gaussian
 * wells, rotating wells, the per-cell models) and of the billiard masks of
 * xy_in_billiard_single_domain (test infrastructure, see wave_oracle.h).
 * Same flags as wave_oracle.c; exp, tanh, log, sqrt, cos and sin are
 * glibc's, as for the reference.
 */
#include <math.h>
#include <stddef.h>

#include "wave_oracle.h"

/* init_ior_2d, sub_wave.c:12315-12389 (IOR_PERIODIC_WELLS, IOR_RANDOM_WELLS;
 * IOR_POISSON_WELLS and IOR_PPP_WELLS 12482-12555 sum the same way over
 * their circles), then the tc rule of sub_wave.c:12970-12981 */
void orc_wells_tables(const orc_params *p, const short *xy, int nw, const double *xc, const double *yc,
                      const double *h, double sigma, int use_tanh, double courant, double courantb,
                      double gamma, const double box[4], double *tcc, double *tc, double *tg)
{
    const int NX = p->nx, NY = p->ny;
    const double xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3];
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            /* ij_to_xy, sub_wave.c:469-476 */
            const double x = xmin + ((double)i) * (xmax - xmin) / ((double)NX);
            const double y = ymin + ((double)j) * (ymax - ymin) / ((double)NY);
            double sum = 0.0;
            for (int n = 0; n < nw; n++) {
                const double r2 = (x - xc[n]) * (x - xc[n]) + (y - yc[n]) * (y - yc[n]);
                if (h) sum += exp(-r2 / (sigma)) * h[n];
                else sum += exp(-r2 / (sigma));
            }
            if (use_tanh) sum = tanh(sum);
            const size_t k = (size_t)i * NY + j;
            tcc[k] = courant * sum + courantb * (1.0 - sum);
            tg[k] = gamma;
            if (xy[k] != 0) tc[k] = sqrt(tcc[k]);
            else if (p->twospeeds) tc[k] = courantb;
        }
}

/* init_wave_flat's mask, wave_common.c:294-319: xy_in[i][j] =
 * xy_in_billiard_single_domain(ij_to_xy(i, j)), sub_wave.c:5712 with
 * ij_to_xy 469-476, for D_SINAI (5774-5779), D_EXT_ELLIPSE (5744-5749) and
 * D_CIRCLES (6216-6227; circ = xc[n], yc[n], radius[n], active[n]).
 * Returns -1 for another domain. */
int orc_mask(int NX, int NY, const double box[4], int domain, double lambda, double mu, int ncirc,
             const double *circ, short *out)
{
    if (domain != 3 && domain != 199 && domain != 20) return -1;
    const double xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3];
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            const double x = xmin + ((double)i) * (xmax - xmin) / ((double)NX);
            const double y = ymin + ((double)j) * (ymax - ymin) / ((double)NY);
            short v = 1;
            if (domain == 3) {
                v = x * x + y * y > lambda * lambda;
            } else if (domain == 199) {
                v = x * x / (lambda * lambda) + y * y / (mu * mu) > 1.0;
            } else {
                for (int k = 0; k < ncirc; k++)
                    if (circ[3 * ncirc + k] != 0.0) {
                        const double x1 = circ[k], y1 = circ[ncirc + k];
                        const double r2 = circ[2 * ncirc + k] * circ[2 * ncirc + k];
                        if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < r2) {
                            v = 0;
                            break;
                        }
                    }
            }
            out[(size_t)i * NY + j] = v;
        }
    return 0;
}

/* init_circular_wave (wave_common.c:89-111; mode 0: phi set, psi = 0, with
 * the mask xy given instead of computed) and add_circular_wave (337-351;
 * mode 1: phi += ..., psi untouched), at the cell corners of ij_to_xy
 * (sub_wave.c:469-476) over box, on cells with xy != 0 or TWOSPEEDS */
void orc_circular_wave(const orc_params *p, int mode, double factor, double x0, double y0, double amp,
                       double variance, double wavelength, const double box[4], double *phi, double *psi,
                       const short *xy)
{
    const int NX = p->nx, NY = p->ny;
    const double xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3];
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            const double x = xmin + ((double)i) * (xmax - xmin) / ((double)NX);
            const double y = ymin + ((double)j) * (ymax - ymin) / ((double)NY);
            const double dist2 = (x - x0) * (x - x0) + (y - y0) * (y - y0);
            const size_t k = (size_t)i * NY + j;
            const int drawn = xy[k] != 0 || p->twospeeds;
            if (mode == 0) {
                phi[k] = drawn ? amp * exp(-dist2 / variance) * cos(-sqrt(dist2) / wavelength) : 0.0;
                psi[k] = 0.0;
            } else if (drawn) {
                phi[k] += amp * factor * exp(-dist2 / variance) * cos(-sqrt(dist2) / wavelength);
            }
        }
}

/* ---- closed-form domains of xy_in_billiard_single_domain ------------------
 * One function per family, each in the reference's shape (same tests in the
 * same order, cosines and sines of the polygon angles taken per point as the
 * reference takes them); sub_wave.c line numbers at each case. */

#define O_PI 3.141592654  /* global_pdes.c:8-10 */
#define O_DPI 6.283185307
#define O_PID 1.570796327

static double o_abs(double x) { return x < 0.0 ? -x : x; }            /* vabs 270 */
static double o_mod(double x, double y) { return sqrt(x * x + y * y); } /* module2 279 */

static int o_tri(double x, double y, const double z1[2], const double z2[2], const double z3[2]) /* 5299 */
{
    double v1 = (z2[0] - z1[0]) * (y - z1[1]) - (z2[1] - z1[1]) * (x - z1[0]);
    double v2 = (z3[0] - z2[0]) * (y - z2[1]) - (z3[1] - z2[1]) * (x - z2[0]);
    double v3 = (z1[0] - z3[0]) * (y - z3[1]) - (z1[1] - z3[1]) * (x - z3[0]);
    return (v1 >= 0.0) && (v2 >= 0.0) && (v3 >= 0.0);
}

static int o_escapes(double u, double v, double cx, double cy, const orc_domain_params *d)
{
    int i = 0;
    while ((i < d->mandellevel) && (u * u + v * v < 1000.0 * d->mandellimit)) {
        double u1 = u * u - v * v + cx;
        v = 2.0 * u * v + cy;
        u = u1;
        i++;
    }
    return !(u * u + v * v < d->mandellimit);
}

/* in_polygon, sub_wave.c:362-376: inside the regular npoly-gon of
 * circumradius r turned by apoly*PID */
static int o_inpoly(double x, double y, double r, int npoly, double apoly)
{
    int condition = 1, k;
    const double omega = O_DPI / ((double)npoly), cw = cos(omega * 0.5);
    for (k = 0; k < npoly; k++) {
        const double angle = apoly * O_PID + ((double)k + 0.5) * omega;
        condition = condition * (x * cos(angle) + y * sin(angle) < r * cw);
    }
    return condition;
}

/* the closed-form cases of xy_in_billiard_single_domain beyond the families
 * of o_domain_at; the reference's one-time constants (its static `first`
 * blocks, 6631-6636, 6647-6652, 6685-6690, rc_hyp 5444-5472) are computed
 * per call, as a program that uses one domain computes them once */
static int o_domain_more(double x, double y, int dom, const orc_domain_params *d, const double box[4])
{
    const double LAMBDA = d->lambda, MU = d->mu, MU_B = d->mu_b, WALL_WIDTH = d->wall_width;
    const double XMIN = box[0], XMAX = box[1], YMIN = box[2], YMAX = box[3];
    const int NPOLY = d->npoly, NGRIDX = d->ngridx, NGRIDY = d->ngridy, MRATIO = d->mratio;
    double x1, y1, x2, y2, r, a, b, h, hh, dx, dy, nx, ny, width, height, angle, omega, ca, sa, wa, wb;
    int i, j, k, k1, condition;
    switch (dom) {
    case 35:                                                                         /* 6068 */
        x1 = 4.0 + x / (XMAX - XMIN) * 8.4;
        y1 = 2.0 + y / (XMAX - XMIN) * 8.4;
        if ((x1 <= 0.0) || (x1 >= 8.0)) return 0;
        if (x1 < 1.0) return (y1 > 2.0) && (y1 < x1 + 2.0);
        if (x1 < 2.0) return (y1 > 1.0) && (y1 < 4.0);
        if (x1 < 3.0) return (y1 > x1 - 2.0) && (y1 < 3.0);
        if (x1 < 4.0) return (y1 > 1.0) && (y1 < 2.0);
        if (x1 < 5.0) return (y1 > x1 - 4.0) && (y1 < 2.0);
        if (x1 < 6.0) return (y1 > 1.0) && (y1 < 3.0);
        if (x1 < 7.0) return (y1 > x1 - 6.0) && (y1 < 10.0 - x1);
        return (y1 > 2.0) && (y1 < 3.0);
    case 43:                                                                         /* 6288 */
        if (o_abs(y) > 0.9 * o_abs(LAMBDA)) return 1;
        if (o_abs(x) > MU) return 1;
        x1 = sqrt(LAMBDA * LAMBDA - y * y) - o_abs(LAMBDA);
        while (x1 <= 0.0) x1 += MU;
        if (LAMBDA > 0.0) return (x < x1) ? 0 : 1;
        return (x > -x1) ? 0 : 1;
    case 45:                                                                         /* 6307 */
        if (o_abs(y) > 0.9 * o_abs(LAMBDA)) return 1;
        if (LAMBDA > 0.0) {
            if (o_abs(x) > LAMBDA + MU) return 1;
            x1 = sqrt(LAMBDA * LAMBDA - y * y) - LAMBDA;
            while (x1 <= 0.0) x1 += MU;
            x1 -= LAMBDA;
            return (o_abs(x) > -x1) ? 0 : 1;
        }
        if (o_abs(x) < -LAMBDA - MU) return 1;
        x1 = sqrt(LAMBDA * LAMBDA - y * y) + LAMBDA;
        while (x1 <= 0.0) x1 += MU;
        x1 -= LAMBDA;
        return (o_abs(x) > x1) ? 1 : 0;
    case 44: case 441:                                                               /* 6331, 6339 */
        x1 = (dom == 44) ? o_abs(x) : -x;
        if (dom == 441 && !(x1 > 0.02)) {                                            /* NPWIDTH 0.02 */
            if (x > 0.02) return (o_abs(y) < YMAX - 0.1) && (x < XMAX - 0.1);
            return 0;
        }
        while (x1 > 2.0 * LAMBDA) x1 -= 2.0 * LAMBDA;
        if (x1 <= LAMBDA) y1 = 0.1 + MU * x1 / LAMBDA;
        else y1 = 0.1 + 2.0 * MU - MU * x1 / LAMBDA;
        if (dom == 44) return (y > YMIN + y1) && (y < YMAX - y1);
        return (y > YMIN + y1) && (y < YMAX - y1) && (x > XMIN + 0.1);
    case 46:                                                                         /* 6355 */
        x1 = o_abs(x) / LAMBDA;
        k = (int)(x1 + 0.5);
        k1 = (k * k) % 13;
        y1 = (MU / 13.0) * (14.0 - (double)k1);
        return (y > YMIN + y1) && (y < YMAX - y1);
    case 461:                                                                        /* 6363 */
        if (y > 0.0) {
            x1 = o_abs(x) / LAMBDA;
            k = (int)(x1 + 0.5);
            k1 = (k * k) % 13;
            y1 = (MU / 13.0) * (14.0 - (double)k1);
            return y < YMAX - y1;
        }
        x1 = o_abs(x + 1.0) / LAMBDA;
        k = (int)(x1 + 0.5);
        k1 = (k * k) % 17;
        y1 = (MU / 17.0) * (18.0 - (double)k1);
        return y > YMIN + y1;
    case 47:                                                                         /* 6382 */
        if (o_abs(y) > 0.9 * o_abs(LAMBDA)) return 1;
        y1 = 0.9 * LAMBDA;
        x1 = sqrt(LAMBDA * LAMBDA - y1 * y1) - o_abs(LAMBDA) + MU;
        if ((LAMBDA > 0.0) && (x < x1)) return 1;
        else if ((LAMBDA < 0.0) && (x > -x1)) return 1;
        x1 = sqrt(LAMBDA * LAMBDA - y * y) - o_abs(LAMBDA) + MU;
        if (LAMBDA > 0.0) return (x < x1) ? 0 : 1;
        return (x > -x1) ? 0 : 1;
    case 48:                                                                         /* 6403 */
        b = 0.85 * LAMBDA;
        a = 0.5 * LAMBDA;
        x1 = x - XMIN - (double)((int)((x - XMIN) / LAMBDA)) * LAMBDA;
        if (x1 < a) return y > YMIN + LAMBDA;
        return y > YMIN + LAMBDA + b;
    case 49: return o_abs(x - y * LAMBDA / YMAX) > 0.5 * MU;                         /* 6411 */
    case 52:                                                                         /* 6437 */
        x1 = o_abs(x);
        width = LAMBDA - 2.0 * MU;
        height = 0.5 * MU;
        if (x1 > 2.0 * LAMBDA - MU) return 0;
        if (o_abs(y) > MU + width + height) return 0;
        if (y >= height) {
            r = o_mod(x1, y - height);
            if ((r > MU) && (r < MU + width)) return 1;
            if (x1 > LAMBDA + MU) return 1;
            return 0;
        }
        if (y <= -height) {
            r = o_mod(x1 - LAMBDA, y + height);
            return (r > MU) && (r < MU + width);
        }
        if (x1 > LAMBDA + MU) return 1;
        return (x1 > MU) && (x1 < MU + width);
    case 521:                                                                        /* 6461 */
        x1 = o_abs(x);
        width = LAMBDA - 2.0 * MU;                                                   /* upper fibre */
        height = MU + width;
        if (y > MU + width + height) return 0;
        if (y >= height) {
            r = o_mod(x1, y - height);
            if ((r > MU) && (r < MU + width)) return 1;
            return (y < height + MU) && (x1 > LAMBDA + MU) && (x1 < LAMBDA + MU + width);
        }
        r = o_mod(x1 - LAMBDA, y - height);
        if ((r > MU) && (r < MU + width)) return 1;
        height = -MU - width;                                                        /* lower fibre */
        width = LAMBDA - 2.0 * MU_B;
        if (y < -MU_B - width + height) return 0;
        if (y >= height) {
            r = o_mod(x1, y - height);
            if ((r > MU_B) && (r < MU_B + width)) return 1;
            return (y < height + MU_B) && (x1 > LAMBDA + MU_B) && (x1 < LAMBDA + MU_B + width);
        }
        r = o_mod(x1 - LAMBDA, y - height);
        return (r > MU_B) && (r < MU_B + width);
    case 526:                                                                        /* 6575 */
        if (x < 0.0) y1 = o_abs(o_abs(y) - 0.5);
        else if (y > 0.0) y1 = o_abs(y - 0.5 - MU);
        else y1 = o_abs(y + 0.5 - MU_B);
        if (y1 < 0.9 * LAMBDA) return 1;
        if (y1 < LAMBDA) return 2;
        return 0;
    case 55:                                                                         /* 6622 */
        i = (int)(o_abs(x) / LAMBDA + 0.5);
        j = (int)(o_abs(y) / LAMBDA + 0.5);
        return (i + j) % 2 == 0;
    case 56:                                                                         /* 6629 */
        h = LAMBDA / (2.0 * sqrt(3.0));
        hh = h * 3.0;
        i = (int)((y + h) / hh + 10.0);
        y1 = sin(O_DPI / 3.0) * x - 0.5 * y;
        j = (int)((y1 + h) / hh + 10.0);
        y1 = sin(-O_DPI / 3.0) * x - 0.5 * y;
        k = (int)((y1 + h) / hh + 10.0);
        return (i + j + k) % 2 == 0;
    case 57:                                                                         /* 6645 */
        a = -1.0 / sqrt(3.0);
        b = -2.0 * a;
        x1 = (x + a * y) / LAMBDA + 30.0;
        y1 = b * y / LAMBDA + 30.0;
        x1 = x1 - (double)(3 * (int)(x1 / 3.0));
        y1 = y1 - (double)(3 * (int)(y1 / 3.0));
        if ((x1 > 2.0) && (y1 < 1.0)) return 1;
        if ((x1 < 1.0) && (y1 > 2.0)) return 1;
        if (x1 + y1 < 1.0) return 1;
        return x1 + y1 > 5.0;
    case 58:                                                                         /* 6665 */
        y1 = y;
        if (y > 0.5 * YMAX) y1 -= YMAX;
        if (y < -0.5 * YMAX) y1 += YMAX;
        y1 = o_abs(y1 - MU * x);
        y1 = o_abs(y1 - 0.5 * YMAX) * 2.0 / YMAX;
        return !(y1 > 0.25 * (1.0 + LAMBDA + x * x));
    case 581: return !(o_abs(y) > MU + LAMBDA * x * x);                              /* 6675 */
    case 59:                                                                         /* 6681 */
        a = O_DPI / (double)NPOLY;
        h = LAMBDA * tan(O_PI / (double)NPOLY);
        b = (h < MU) ? sqrt(MU * MU - h * h) : 0.0;
        for (i = 0; i < NPOLY; i++) {
            ca = cos((double)i * a + d->apoly * O_PID);
            sa = sin((double)i * a + d->apoly * O_PID);
            x1 = x * ca + y * sa;
            y1 = -x * sa + y * ca;
            if ((o_mod(x1 - LAMBDA - b, y1) < MU) && (o_mod(x1 - LAMBDA + b, y1) < MU)) return 0;
        }
        return 1;
    case 65:                                                                         /* 6736 */
        if (o_abs(x + LAMBDA) >= MU) return 1;
        return cos(O_DPI * (6.25 * y * y + 0.125)) < 0.0;
    case 66:                                                                         /* 6743 */
        if (o_abs(x + LAMBDA) >= MU) return 1;
        return cos(O_DPI * 6.25 * y * y) > 0.0;
    case 731: {                                                                      /* 6828 */
        const double c6 = cos(O_PI / 6.0), lev[3] = {1.5, 0.25, -1.5}, rad[3] = {1.5, 2.0, 3.0};
        x1 = o_abs(x);
        for (k = 0; k < 3; k++) {
            r = rad[k] * LAMBDA;
            const double z1[2] = {r * c6, -0.5 * r}, z2[2] = {0.0, r}, z3[2] = {0.0, -0.5 * r};
            if (o_tri(x1, y - lev[k] * LAMBDA, z1, z2, z3)) return 1;
        }
        if ((x < 0.0) && (y > d->oscil_ymid - 0.5 * WALL_WIDTH) && (y < d->oscil_ymid + 0.5 * WALL_WIDTH))
            return 1;
        if ((x < -(3.0 * LAMBDA) * c6 + 0.5 * WALL_WIDTH) && (y > d->oscil_ymid - 0.75 * WALL_WIDTH) &&
            (y < d->oscil_ymid + 0.75 * WALL_WIDTH))
            return 2;
        return 0;
    }
    case 75:                                                                         /* 6892 */
        a = 2.5;
        b = 3.5;
        x2 = 2.0 * (a + b / LAMBDA);                                                 /* r1 */
        y2 = 6.0 / (LAMBDA - 1.0);                                                   /* r2 */
        x1 = 0.5 * a - x2;
        h = -0.5 * a - y2;
        y1 = o_abs(y);
        if (x < 0.0) {
            if (y1 > 0.3) return 1;
            if (x < -0.5 * a - WALL_WIDTH) return 1;
            return o_mod(x - h, y1) > y2;
        }
        if (y1 < MU) return 1;
        if (x > 0.5 * a + WALL_WIDTH) return 1;
        return o_mod(x - x1, y1) < x2;
    case 751: {                                                                      /* rc_hyp 5437 */
        const double m = LAMBDA, dd = 2.5, bb = 3.0;
        const double f = m * dd + bb, f1 = f / m, dprime = f1 - dd, f2 = m * dprime / (m + 1);
        const double r1 = 2.0 * f / m, r2 = 2.0 * bb / (m - 1.0);
        const double e1 = sqrt(1.0 + 2.0 * bb / (m * m * m * dd));
        const double e2 = sqrt(1.0 + 2.0 * (m * (2.0 * m - 1.0) + bb / dd) / ((m - 1) * (m - 1) * (m - 1)));
        const double a1 = f1 / e1, b1 = sqrt(a1 * r1), x01 = 0.5 * dd + a1;
        const double a2 = f2 / e2, b2 = sqrt(a2 * r2), x02 = -0.5 * dd + a2;
        y1 = o_abs(y);
        if (x > 0.0) {
            if (y1 < MU) return 1;
            if (x > 0.5 * dd + WALL_WIDTH) return 1;
            return (x - x01) * (x - x01) > a1 * a1 * (1.0 + y * y / (b1 * b1));
        }
        if (y1 > 0.3) return 1;
        return (x > x02) || ((x - x02) * (x - x02) < a2 * a2 * (1.0 + y * y / (b2 * b2)));
    }
    case 78: case 781:                                                               /* 6929, 6947 */
        r = o_mod(x, y);
        if (dom == 781 && r < WALL_WIDTH) return 0;
        if (r > LAMBDA) return 1;
        if (r < 0.33 * LAMBDA) return 1;
        if ((o_abs(y) < WALL_WIDTH) && (x > -0.66 * LAMBDA)) return 1;
        if ((o_abs(x) < WALL_WIDTH) && (o_abs(y) < 0.66 * LAMBDA)) return 1;
        if ((o_abs(x - y) < WALL_WIDTH * sqrt(2.0)) && (r < 0.66 * LAMBDA)) return 1;
        if ((o_abs(x + y) < WALL_WIDTH * sqrt(2.0)) && (r < 0.66 * LAMBDA)) return 1;
        for (k = 0; k < 8; k++) {
            x1 = 0.66 * cos((double)k * 0.5 * O_PID);
            y1 = 0.66 * sin((double)k * 0.5 * O_PID);
            if (o_mod(x - x1, y - y1) < MU) return 1;
        }
        return 0;
    case 792: {                                                                      /* 6992 */
        if (o_mod(x, y) < LAMBDA) return 1;
        const double cb = cos(-0.5 * O_PID), sb = sin(-0.5 * O_PID);
        for (k = 0; k < NPOLY; k++) {
            angle = d->apoly * O_PID + (double)k * O_DPI / (double)NPOLY;
            ca = cos(angle);
            sa = sin(angle);
            x1 = x * ca + y * sa;
            y1 = -x * sa + y * ca;
            x2 = LAMBDA + 2.0 * MU * cb;
            y2 = 2.0 * MU * sb;
            if (o_mod(x1 - x2, y1 - y2) < MU) return 1;
            const double c1 = -sb * LAMBDA + cb * WALL_WIDTH, c2 = -sb * LAMBDA - cb * WALL_WIDTH;
            b = -sb * x1 + cb * y1;
            if ((x1 > 0.5 * LAMBDA) && (x1 < LAMBDA + MU) && (b < c1) && (b > c2)) return 1;
        }
        return 0;
    }
    case 793:                                                                        /* 7015 */
        a = 0.9 * XMAX - 3.0 * MU;
        b = a / LAMBDA;
        if (x * x / (a * a) + y * y / (b * b) < 1.0) return 1;
        x1 = a + 2.0 * MU;
        if ((x - x1) * (x - x1) + y * y < MU * MU) return 1;
        if ((x + x1) * (x + x1) + y * y < MU * MU) return 1;
        return (o_abs(x) < a + 2.0 * MU) && (o_abs(y) < WALL_WIDTH);
    case 794:                                                                        /* 7026 */
        a = 0.9 * XMAX - 3.0 * MU;
        b = a / LAMBDA;
        x = x / d->julia_scale;
        y = y / d->julia_scale;
        if (x * x / (a * a) + y * y / (b * b) < 1.0) return 1;
        for (k = 0; k < NPOLY; k++) {
            angle = d->apoly * O_PID + (double)k * O_DPI / (double)NPOLY;
            x1 = a * cos(angle);
            y1 = b * sin(angle);
            nx = b * cos(angle);
            ny = a * sin(angle);
            r = 1.0 / o_mod(nx, ny);
            nx *= r;
            ny *= r;
            x2 = x1 + 2.0 * MU * nx;
            y2 = y1 + 2.0 * MU * ny;
            if (o_mod(x - x2, y - y2) < MU) return 1;
            r = o_mod(x - x1, y - y1);
            h = (x - x1) * nx + (y - y1) * ny;
            if ((h > -MU) && (h < 2.0 * MU) && (r * r - h * h < WALL_WIDTH * WALL_WIDTH)) return 1;
        }
        return 0;
    case 797: return o_inpoly(o_abs(x) - LAMBDA, y, MU, NPOLY, d->apoly);            /* 7062 */
    case 798:                                                                        /* 7068 */
        condition = 1;
        omega = O_DPI / ((double)NPOLY);
        if (x > 0.0) x -= LAMBDA;
        else x += LAMBDA;
        for (k = 0; k < NPOLY; k++) {
            angle = d->apoly * O_PID + ((double)k + 0.5) * omega;
            x1 = x * cos(angle) + y * sin(angle);
            y1 = -x * sin(angle) + y * cos(angle);
            condition = condition * (x1 < MU - 0.25 * y1 * y1 / MU);
        }
        return condition;
    case 93: case 934: case 935: case 936:                                           /* 7110, 7225, 7243, 7262 */
        if (dom == 935 && y > 0.0) return 1;
        if (dom == 936 && x < 0.0) return 1;
        wa = (dom == 93) ? WALL_WIDTH : WALL_WIDTH * d->wall_width_asym;
        dx = (XMAX - XMIN) / (double)NGRIDX;
        dy = (YMAX - YMIN) / (double)NGRIDY;
        for (i = 0; i < NGRIDX; i++) {
            x1 = XMIN + ((double)i + 0.5) * dx;
            if (o_abs(x - x1) < wa) return 1;
            for (j = 0; j < NGRIDY; j++) {
                y1 = YMIN + ((double)j + 0.5) * dy;
                if (o_mod(x - x1, y - y1) < MU) return 1;
                if (o_abs(y - y1) < WALL_WIDTH) return 1;
            }
        }
        return 0;
    case 937:                                                                        /* 7281 */
        if (x > LAMBDA) return 1;
        if (x < -LAMBDA) return 1;
        dx = 2.0 * LAMBDA / (double)NGRIDX;
        dy = (YMAX - YMIN) / (double)NGRIDY;
        for (i = 0; i < NGRIDX + 1; i++) {
            x1 = -LAMBDA + ((double)i) * dx;
            if (o_abs(x - x1) < WALL_WIDTH * d->wall_width_asym) return 1;
            for (j = 0; j < NGRIDY; j++) {
                y1 = YMIN + ((double)j + 0.5) * dy;
                if (o_mod(x - x1, y - y1) < MU) return 1;
                if (o_abs(y - y1) < WALL_WIDTH) return 1;
            }
        }
        return 0;
    case 94: case 942: case 947:                                                     /* 7146, 7321, 7344 */
        if (dom == 947 && o_abs(x) > LAMBDA) return 1;
        dx = (dom == 947) ? 2.0 * LAMBDA / (double)NGRIDX : (XMAX - XMIN) / (double)NGRIDX;
        dy = (YMAX - YMIN) / (double)NGRIDY;
        r = o_mod(0.5 * dx, dy);
        nx = -dy / r;
        ny = dx / (2.0 * r);
        wa = (dom == 94) ? WALL_WIDTH : WALL_WIDTH * d->wall_width_asym;
        wb = (dom == 94) ? WALL_WIDTH : WALL_WIDTH * d->wall_width_asym_b;
        for (j = 0; j < NGRIDY + 1; j++) {
            y1 = YMIN + ((double)j) * dy;
            if (o_abs(y - y1) < WALL_WIDTH) return 1;
            for (i = -1; i < NGRIDX; i++) {
                x1 = ((dom == 947) ? -LAMBDA : XMIN) + ((double)i + 0.5) * dx;
                if (j % 2 == 0) x1 += 0.5 * dx;
                if (o_mod(x - x1, y - y1) < MU) return 1;
                if (o_abs(nx * (x - x1) + ny * (y - y1)) < wa) return 1;
                if (o_abs(-nx * (x - x1) + ny * (y - y1)) < wb) return 1;
            }
        }
        return 0;
    case 98:                                                                         /* 7368 */
        if (x > 0.0) return x * x + y * y < LAMBDA * LAMBDA;
        y1 = y - 0.5 * WALL_WIDTH;
        r = LAMBDA - 0.5 * WALL_WIDTH;
        if (x * x + y1 * y1 < r * r) return 1;
        if ((y > -LAMBDA) && (y < -LAMBDA + WALL_WIDTH)) return 1;
        if ((y > -LAMBDA - 0.5 * WALL_WIDTH) && (y < -LAMBDA + 1.5 * WALL_WIDTH)) return 2;
        return 0;
    case 981: case 982: case 983:                                                    /* 7378, 7387, 7396 */
        if (dom == 981 ? (x * x + y * y < LAMBDA * LAMBDA) : (x * x / (LAMBDA * LAMBDA) + y * y < 1.0)) return 1;
        if (x > 0.0) return 0;
        if (dom == 981) y1 = y + LAMBDA - MU - 0.5 * WALL_WIDTH;
        else if (dom == 982) y1 = y + 1.0 - MU - 0.5 * WALL_WIDTH;
        else y1 = y + 1.0 - MU - 0.5 * WALL_WIDTH - (x - XMIN) * tan(d->apoly * O_PID);
        if (o_abs(y1) < 0.5 * WALL_WIDTH) return 1;
        if (o_abs(y1) < 0.75 * WALL_WIDTH) return 2;
        return 0;
    case 99: case 991:                                                               /* 7405, 7418 */
        dx = (XMAX - XMIN) / (double)NGRIDX;
        for (k = 0; k < NGRIDX; k++) {
            x1 = XMIN + dx * ((double)k + 0.5);
            y1 = LAMBDA;
            if (o_mod(x - x1, y - y1) < MU) return 1;
            if (dom == 99 ? (o_mod(x - x1, y + y1) < MU_B)
                          : o_inpoly(x - x1, y + y1, MU_B, NPOLY, 1.0 - 2.0 / (double)NPOLY))
                return 1;
            if ((o_abs(y) < LAMBDA) && (o_abs(x - x1) < 0.5 * WALL_WIDTH)) return 1;
        }
        return 0;
    case 77: return 1;                                                               /* 6924 */
    case 741:                                                                        /* 6879 */
        if ((o_abs(x) > LAMBDA) && (o_abs(y) > LAMBDA)) return 2;
        r = 0.5 * sqrt(2.0);
        h = YMAX - MU - 0.5 * WALL_WIDTH;
        hh = YMAX - MU - 0.5 * WALL_WIDTH + d->michelson_position;
        x1 = (x + y) * r;
        y1 = (-x + y) * r;
        if ((o_abs(y1) < 0.5 * WALL_WIDTH) && (o_abs(x1) < LAMBDA * sqrt(2.0))) return 1;
        if ((o_abs(x - hh) < 0.5 * WALL_WIDTH) && (o_abs(y) < LAMBDA)) return 2;
        if ((o_abs(x) < LAMBDA) && (o_abs(y - h) < 0.5 * WALL_WIDTH)) return 2;
        return 0;
    case 21: {                                                                       /* 7475 */
        const double l2 = LAMBDA * LAMBDA, r2 = x * x + y * y, r2mu = (x - MU) * (x - MU) + y * y;
        if ((r2mu > l2) && (r2 < 1.0)) return 1;
        else if (r2mu <= l2) return 2;
        return 0;
    }
    case 22: case 23:                                                                /* 7484, 7500 */
        if (dom == 22 && ((o_abs(x) >= 1.0) || (o_abs(y) >= 1.0))) return 0;
        x1 = 0.5 * (x + 1.0);
        y1 = 0.5 * (y + 1.0);
        for (k = 0; k < d->mdepth; k++) {
            x1 = x1 * (double)MRATIO;
            y1 = y1 * (double)MRATIO;
            if ((dom == 22 || ((o_abs(x) < 1.0) && (o_abs(y) < 1.0))) && (((int)x1 % MRATIO) == MRATIO / 2) &&
                (((int)y1 % MRATIO) == MRATIO / 2))
                return k + 2;
        }
        return 1;
    }
    return -1;
}

static int o_domain_at(double x, double y, int dom, const orc_domain_params *d, const double box[4], int nc,
                       const double *circ)
{
    const double LAMBDA = d->lambda, MU = d->mu, WALL_WIDTH = d->wall_width;
    const double XMIN = box[0], XMAX = box[1], YMIN = box[2], YMAX = box[3];
    double x1, y1, x2, y2, r, r2, a, b, c, width, angle, omega;
    int i, j, k, condition;
    switch (dom) {
    case 999: return 1;
    case 0: return (o_abs(x) < LAMBDA) && (o_abs(y) < 1.0);                          /* 5726 */
    case 1: return x * x / (LAMBDA * LAMBDA) + y * y < 1.0;                          /* 5732 */
    case 196: return x * x + y * y < LAMBDA * LAMBDA;                                /* 5738 */
    case 199: return x * x / (LAMBDA * LAMBDA) + y * y / (MU * MU) > 1.0;            /* 5744 */
    case 198: case 197:                                                              /* 5750, 5757 */
        if (dom == 197 && y > YMAX - 0.05) return 0;
        if (dom == 197 && y < YMIN + 0.05) return 0;
        y1 = y + 0.4 * x * x;
        return x * x / (LAMBDA * LAMBDA) + y1 * y1 / (MU * MU) > 1.0;
    case 2:                                                                          /* 5766 */
        if ((x > -0.5 * LAMBDA) && (x < 0.5 * LAMBDA) && (y > -1.0) && (y < 1.0)) return 1;
        else if (o_mod(x + 0.5 * LAMBDA, y) < 1.0) return 1;
        else if (o_mod(x - 0.5 * LAMBDA, y) < 1.0) return 1;
        return 0;
    case 3: return x * x + y * y > LAMBDA * LAMBDA;                                  /* 5774 */
    case 4: {                                                                        /* 5780 */
        double l2 = LAMBDA * LAMBDA;
        r2 = l2 + (LAMBDA - 1.0) * (LAMBDA - 1.0);
        return (x * x + y * y < 1.0) && ((x - LAMBDA) * (x - LAMBDA) + (y - LAMBDA) * (y - LAMBDA) > r2) &&
               ((x - LAMBDA) * (x - LAMBDA) + (y + LAMBDA) * (y + LAMBDA) > r2) &&
               ((x + LAMBDA) * (x + LAMBDA) + (y - LAMBDA) * (y - LAMBDA) > r2) &&
               ((x + LAMBDA) * (x + LAMBDA) + (y + LAMBDA) * (y + LAMBDA) > r2);
    }
    case 5: return (x > -LAMBDA) && (y > -1.0) && (LAMBDA * y + x < 0.0);            /* 5791 */
    case 6: return y > -LAMBDA;                                                      /* 5797 */
    case 7: {                                                                        /* 5803 */
        double l2 = LAMBDA * LAMBDA;
        r2 = x * x + y * y;
        return (r2 > l2) && (r2 < 1.0);
    }
    case 8:                                                                          /* 5810 */
        condition = 1;
        omega = O_DPI / ((double)d->npoly);
        c = LAMBDA * cos(omega * 0.5);
        for (k = 0; k < d->npoly; k++) {
            angle = d->apoly * O_PID + (k + 0.5) * omega;
            condition = condition * (x * cos(angle) + y * sin(angle) < c);
        }
        return condition;
    case 9:                                                                          /* 5823 */
        if ((x < -MU) || (x > MU)) return 1;
        return (o_abs(y - LAMBDA) < MU) || (o_abs(y + LAMBDA) < MU);
    case 10: {                                                                       /* 5829 */
        int k1 = -(int)((-YMIN) / LAMBDA), k2 = (int)(YMAX / LAMBDA);
        condition = 1;
        for (i = k1; i <= k2; i++) {
            double z = (double)i * LAMBDA;
            condition = condition * (x * x + (y - z) * (y - z) > MU * MU);
        }
        return condition;
    }
    case 11:                                                                         /* 5842 */
        return ((x - 1.0) * (x - 1.0) + y * y < LAMBDA * LAMBDA) || ((x + 1.0) * (x + 1.0) + y * y < LAMBDA * LAMBDA) ||
               ((o_abs(x) < 1.0) && (o_abs(y) < MU));
    case 12: case 13: {                                                              /* 5846, 5858 */
        double dy = (YMAX - YMIN) / ((double)d->ngridy);
        for (i = -d->ngridx / 2; i < d->ngridx / 2; i++)
            for (j = (dom == 13 ? -1 : 0); j < d->ngridy; j++) {
                x1 = ((double)i + 0.5) * dy;
                y1 = YMIN + ((double)j + 0.5) * dy;
                if (dom == 13 && (i + d->ngridx) % 2 == 1) y1 += 0.5 * dy;
                if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < MU * MU) return 0;
            }
        return 1;
    }
    case 15:                                                                         /* 7431 */
        x1 = 0.5 * (x + 1.0);
        y1 = 0.5 * (y + 1.0);
        for (k = 0; k < d->mdepth; k++) {
            x1 = x1 * (double)d->mratio;
            y1 = y1 * (double)d->mratio;
            if ((o_abs(x) < 1.0) && (o_abs(y) < 1.0) && (((int)x1 % d->mratio) == d->mratio / 2) &&
                (((int)y1 % d->mratio) == d->mratio / 2))
                return 0;
        }
        return 1;
    case 17:                                                                         /* 7458 */
        x2 = 1.0 * (x + y);
        y2 = 1.0 * (x - y);
        if ((o_abs(x2) < 1.0) && (o_abs(y2) < 1.0)) {
            x1 = 0.5 * (x2 + 1.0);
            y1 = 0.5 * (y2 + 1.0);
            for (k = 0; k < d->mdepth; k++) {
                x1 = x1 * (double)d->mratio;
                y1 = y1 * (double)d->mratio;
                if ((o_abs(x) < 1.0) && (o_abs(y) < 1.0) && (((int)x1 % d->mratio) == d->mratio / 2) &&
                    (((int)y1 % d->mratio) == d->mratio / 2))
                    return 0;
            }
        }
        return 1;
    case 16: return !o_escapes(x / d->julia_scale, y / d->julia_scale, d->julia_x, d->julia_y, d);  /* 7443 */
    case 18: return x > 0.25 * y * y / LAMBDA - LAMBDA;                              /* 5872 */
    case 19:                                                                         /* 5876 */
        x1 = 0.25 * y * y / MU - MU - LAMBDA;
        x2 = -x1;
        width = 0.25 * MU;
        if (width > 0.2) width = 0.2;
        if (o_abs(y) > 1.5 * MU) return 1;
        else if ((x < x1 - width) || (x > x2 + width)) return 1;
        else if ((x > x1) && (x < x2)) return 1;
        return 0;
    case 191:                                                                        /* 5887 */
        if (x > 0.0) {
            if (o_abs(y) > 0.5 * LAMBDA) return 1;
            x1 = -y * y / LAMBDA + 0.25 * LAMBDA;
            return (x < x1) || (x > x1 + WALL_WIDTH);
        }
        if (o_abs(y) > 0.5 * MU) return 1;
        x1 = y * y / MU - 0.25 * MU;
        return (x > x1) || (x < x1 - WALL_WIDTH);
    case 195:                                                                        /* 5977 */
        if (o_abs(y) > 0.75 + WALL_WIDTH) return 1;
        x1 = 0.5 * LAMBDA + 0.25 * MU - y * y / MU;
        if ((o_abs(y) > 0.75) && (o_abs(x) < x1 + WALL_WIDTH)) return 2;
        if ((o_abs(x) < x1) || (o_abs(x) > x1 + WALL_WIDTH)) return 1;
        if ((x > 0.0) && (o_abs(y) < 0.2)) return 0;
        return 2;
    case 1951: case 1952: {                                                          /* 5986, 6006 */
        double ylens = dom == 1951 ? 0.2 : d->wall_width_b;
        if (o_abs(y) > 0.75 + WALL_WIDTH) return 1;
        if ((x > 0.0) && (o_abs(y) < ylens)) {
            double x3;
            y1 = ylens;
            x1 = 0.5 * LAMBDA + 0.25 * MU - y1 * y1 / MU;
            x2 = x1 + WALL_WIDTH;
            r2 = dom == 1951 ? 1.5 * 0.25 * LAMBDA : d->mu_b;
            r2 = r2 * r2;
            x3 = sqrt(r2 - y1 * y1) - x2;
            if (((x + x3) * (x + x3) + y * y < r2) && (x > x1)) return 0;
            return 1;
        }
        x1 = 0.5 * LAMBDA + 0.25 * MU - y * y / MU;
        if ((o_abs(y) > 0.75) && (o_abs(x) < x1 + WALL_WIDTH)) return 2;
        if ((o_abs(x) < x1) || (o_abs(x) > x1 + WALL_WIDTH)) return 1;
        return 2;
    }
    case 20: case 201: case 202:                                                     /* 6216, 6240, 6228 */
        for (i = 0; i < nc; i++)
            if (circ[3 * nc + i] != 0.0) {
                x1 = circ[i];
                y1 = circ[nc + i];
                r2 = circ[2 * nc + i] * circ[2 * nc + i];
                if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < r2) return 0;
            }
        if (dom == 201) return !((o_abs(x) >= LAMBDA) || (o_abs(y) >= YMAX - XMAX + LAMBDA));
        return 1;
    case 24: case 26:                                                                /* 7512, 7531 */
        if (!o_escapes(0.0, 0.0, x, y, d)) return 0;
        if (dom == 24) return ((x - 0.5) * (x - 0.5) / 3.0 + y * y / 1.0 > 1.2) ? 2 : 1;
        return ((x - LAMBDA) * (x - LAMBDA) + (y - 0.5) * (y - 0.5) < MU * MU) ? 2 : 1;
    case 25:                                                                         /* 7547 */
        if (!o_escapes(x / d->julia_scale, y / d->julia_scale, d->julia_x, d->julia_y, d)) return 0;
        return (x * x / 3.0 + y * y / 1.0 > 1.2) ? 2 : 1;
    case 31:                                                                         /* 6026 */
        x1 = MU + LAMBDA - 0.25 * y * y / MU;
        y1 = MU + LAMBDA - 0.25 * x * x / MU;
        return (o_abs(x) < x1) && (o_abs(y) < y1);
    case 32:                                                                         /* 6032 */
        condition = 1;
        omega = O_DPI / ((double)d->npoly);
        for (k = 0; k < d->npoly; k++) {
            angle = d->apoly * O_PID + ((double)k + 0.5) * omega;
            x1 = x * cos(angle) + y * sin(angle);
            y1 = -x * sin(angle) + y * cos(angle);
            condition = condition * (x1 < LAMBDA + MU - 0.25 * y1 * y1 / MU);
        }
        return condition;
    case 33:                                                                         /* 6045 */
        c = sqrt(LAMBDA * LAMBDA - (1.0 - MU) * (1.0 - MU));
        width = 0.1 * MU;
        x1 = o_abs(x);
        y1 = o_abs(y);
        if (o_abs(x) >= LAMBDA) return 0;
        else if ((o_abs(y) >= MU) &&
                 (x * x / (LAMBDA * LAMBDA) + (y1 - MU) * (y1 - MU) / ((1.0 - MU) * (1.0 - MU)) >= 1.0))
            return 0;
        else if ((o_abs(x) <= c) && (4.0 * (x1 - c) * (x1 - c) / (MU * MU) + y * y / (MU * MU) <= 1.0)) return 0;
        else if ((o_abs(x) >= c) && (o_abs(y) <= width)) return 0;
        return 1;
    case 34:                                                                         /* 6061 */
        b = MU * sqrt(1.0 + x * x / (LAMBDA * LAMBDA - MU * MU));
        if (y > 1.02 * b) return 1;
        else if (y < 0.98 * b) return 1;
        return 0;
    case 50:                                                                         /* 6415 */
        if (o_abs(x) > LAMBDA) return 0;
        else if (o_abs(y) > 1.0) return 0;
        else if ((x > 0.0) && (y > 0.0)) return 0;
        return 1;
    case 51:                                                                         /* 6422 */
        x1 = XMIN + MU;
        x2 = XMAX - 2.0 * MU - 1.5 * LAMBDA;
        y1 = 0.5 * LAMBDA;
        y2 = 1.5 * LAMBDA;
        if (x < x1) return 0;
        if (x > x2 + 1.5 * LAMBDA) return 0;
        if (o_abs(y) > y2) return 0;
        if (x < x2) return o_abs(y) >= y1;
        r = o_mod(x - x2, y);
        if (r < 0.5 * LAMBDA) return 0;
        if (r > 1.5 * LAMBDA) return 0;
        return 1;
    case 522: case 525: {                                                            /* 6501, 6518 */
        double x5, x6, f, fp, cc = 15.0;
        x1 = dom == 522 ? o_abs(x) * 2.0 / XMAX : o_abs(x);
        y1 = o_abs(y);
        width = dom == 522 ? 0.03 : 0.01;
        b = 1.0 * (1.0 - 2.0 * MU - width);
        a = MU + 0.5 * width - b;
        x6 = x1 * x1 * x1;
        x5 = x6 * x1 * x1;
        x6 = x5 * x1;
        if (dom == 522) {
            f = a + b * (1.0 + 2.0 * x6) / (1.0 + x6);
            fp = 6.0 * b * x5 / ((1.0 + x6) * (1.0 + x6));
            return o_abs(f - y1) < MU * sqrt(1.0 + fp * fp);
        }
        f = a + b * (1.0 + 2.0 * cc * x6) / (1.0 + cc * x6);
        fp = 6.0 * b * cc * x5 / ((1.0 + cc * x6) * (1.0 + cc * x6));
        if (o_abs(f - y1) < 0.9 * MU * sqrt(1.0 + fp * fp)) return 1;
        if (o_abs(f - y1) < MU * sqrt(1.0 + fp * fp)) return 2;
        return 0;
    }
    case 523:                                                                        /* 6538 */
        a = 0.5 * (1.0 - MU);
        if (x < 0.0) {
            x = -x;
            y = -y;
        }
        if (x < LAMBDA) {
            if (o_abs(y - 2.0 * a) < MU) return 1;
            if (o_abs(y) < MU) return 1;
            if (o_abs(y + 2.0 * a) < MU) return 1;
            return 0;
        }
        r = o_mod(x - LAMBDA, y - a);
        return o_abs(r - a) < MU;
    case 524:                                                                        /* 6556 */
        a = -1.0 + MU;
        if (y < 0.0) {
            x = -x;
            y = -y;
        }
        if (x < -LAMBDA) {
            if (o_abs(y - 1.0 + MU) > MU) return 0;
            if (o_abs(y - 1.0 + MU) > 0.8 * MU) return 2;
            return 1;
        }
        r = o_mod(x - a, y);
        if (o_abs(r - 1.0 + MU) < 0.8 * MU) return 1;
        if (o_abs(r - 1.0 + MU) < MU) return 2;
        return 0;
    case 61: case 62:                                                                /* 6701, 6708 */
        if (o_abs(y) > MU) return 1;
        a = LAMBDA - 0.25;
        if (x > 0) return (x + a) * (x + a) + y * y > LAMBDA * LAMBDA;
        return (x - a) * (x - a) + y * y > LAMBDA * LAMBDA;
    case 63: case 64:                                                                /* 6715, 6726 */
        width = 0.2;
        a = 1.9;
        b = 1.9 - 2.0 * LAMBDA + 2.0 * width;
        x1 = o_abs(x);
        if (dom == 63 && (x1 < WALL_WIDTH) && (o_abs(y) > MU)) return 2;
        if (o_mod(x1 - a, y) > LAMBDA) return 1;
        if (o_mod(x1 - b, y) > LAMBDA) return 1;
        return 0;
    case 67: {                                                                       /* 6749 */
        double ca = cos(d->apoly * O_PID), sa = sin(d->apoly * O_PID);
        x1 = x * ca + y * sa;
        y1 = -x * sa + y * ca;
        if (o_abs(y1) > MU) return 1;
        a = LAMBDA - 0.25;
        if (x1 > 0) return (x1 + a) * (x1 + a) + y1 * y1 > LAMBDA * LAMBDA;
        return (x1 - a) * (x1 - a) + y1 * y1 > LAMBDA * LAMBDA;
    }
    case 68:                                                                         /* 6760 */
        width = 0.1;
        x1 = o_abs(x);
        if (o_abs(y) > MU) return 1;
        return (x1 - LAMBDA - width) * (x1 - LAMBDA - width) + y * y < LAMBDA * LAMBDA;
    case 69:                                                                         /* 6767 */
        if (o_abs(y) > MU) return 1;
        if (x > 0.0) {
            x1 = x - 0.6;
            a = LAMBDA - 0.25;
            if (x1 > 0) return (x1 + a) * (x1 + a) + y * y > LAMBDA * LAMBDA;
            return (x1 - a) * (x1 - a) + y * y > LAMBDA * LAMBDA;
        }
        width = 0.1;
        x1 = o_abs(x + 0.6);
        return (x1 - LAMBDA - width) * (x1 - LAMBDA - width) + y * y < LAMBDA * LAMBDA;
    case 73: {                                                                       /* 6785 */
        double zz[9][2], h, y3;
        h = 1.5 * LAMBDA;
        r = 1.5 * LAMBDA;
        x1 = o_abs(x);
        y1 = y - h;
        zz[0][0] = r * cos(O_PI / 6.0); zz[0][1] = -0.5 * r;
        zz[1][0] = 0.0;                 zz[1][1] = r;
        zz[2][0] = 0.0;                 zz[2][1] = -0.5 * r;
        if (y1 > 0.95 * zz[1][1]) return 1;
        h = 0.25 * LAMBDA;
        r = 2.0 * LAMBDA;
        y2 = y - h;
        zz[3][0] = r * cos(O_PI / 6.0); zz[3][1] = -0.5 * r;
        zz[4][0] = 0.0;                 zz[4][1] = r;
        zz[5][0] = 0.0;                 zz[5][1] = -0.5 * r;
        h = -1.5 * LAMBDA;
        r = 3.0 * LAMBDA;
        y3 = y - h;
        zz[6][0] = r * cos(O_PI / 6.0); zz[6][1] = -0.5 * r;
        zz[7][0] = 0.0;                 zz[7][1] = r;
        zz[8][0] = 0.0;                 zz[8][1] = -0.5 * r;
        if (o_tri(x1, y1, zz[0], zz[1], zz[2])) return 0;
        if (o_tri(x1, y2, zz[3], zz[4], zz[5])) return 0;
        if (o_tri(x1, y3, zz[6], zz[7], zz[8])) return 0;
        if (o_tri(0.9 * x1, 0.9 * y1, zz[0], zz[1], zz[2])) return 1;
        if (o_tri(0.9 * x1, 0.9 * y2, zz[3], zz[4], zz[5])) return 1;
        if (o_tri(0.9 * x1, 0.9 * y3, zz[6], zz[7], zz[8])) return 1;
        if (o_tri(0.8 * x1, 0.8 * y1, zz[0], zz[1], zz[2])) return 2;
        if (o_tri(0.8 * x1, 0.8 * y2, zz[3], zz[4], zz[5])) return 2;
        if (o_tri(0.8 * x1, 0.8 * y3, zz[6], zz[7], zz[8])) return 2;
        return 1;
    }
    case 74: {                                                                       /* 6867 */
        double h;
        if ((o_abs(x) > LAMBDA) && (o_abs(y) > LAMBDA)) return 2;
        r = 0.5 * sqrt(2.0);
        h = YMAX - 0.5 * WALL_WIDTH;
        x1 = (x + y) * r;
        y1 = (-x + y) * r;
        if ((o_abs(y1) < 0.5 * WALL_WIDTH) && (o_abs(x1) < LAMBDA * sqrt(2.0))) return 1;
        if ((o_abs(x - h) < 0.5 * WALL_WIDTH) && (o_abs(y) < LAMBDA)) return 2;
        if ((o_abs(x) < LAMBDA) && (o_abs(y - h) < 0.5 * WALL_WIDTH)) return 2;
        return 0;
    }
    case 76: return 1;                                                               /* 6919 */
    case 79:                                                                         /* 6966 */
        x1 = 0.25 * (LAMBDA + 5.0 * MU);
        x2 = -0.75 * (LAMBDA + 2.0 * MU);
        if (o_mod(x - x1, y) < LAMBDA) return 1;
        if (o_mod(x - x2, y) < MU) return 1;
        if ((x < x1) && (x > x2) && (o_abs(y) < WALL_WIDTH)) return 1;
        return 0;
    case 791:                                                                        /* 6975 */
        if (o_mod(x, y) < LAMBDA) return 1;
        for (k = 0; k < d->npoly; k++) {
            double ca, sa;
            angle = d->apoly * O_PID + (double)k * O_DPI / (double)d->npoly;
            ca = cos(angle);
            sa = sin(angle);
            x1 = (LAMBDA + 2.0 * MU) * ca;
            y1 = (LAMBDA + 2.0 * MU) * sa;
            if (o_mod(x - x1, y - y1) < MU) return 1;
            x1 = x * ca + y * sa;
            y1 = -x * sa + y * ca;
            if ((x1 > 0.0) && (x1 < LAMBDA + 2.0 * MU) && (o_abs(y1) < WALL_WIDTH)) return 1;
        }
        return 0;
    case 795:                                                                        /* 7052 */
        x1 = o_abs(x);
        return (x1 - LAMBDA) * (x1 - LAMBDA) / (MU * MU) + y * y < 1.0;
    case 796:                                                                        /* 7057 */
        x1 = o_abs(x);
        return (o_abs(x1 - LAMBDA) < MU) && (o_abs(y) < 1.0);
    case 90:                                                                         /* 7083 */
        y1 = y - MU;
        r = x * x + y1 * y1;
        return r * r + 4.0 * LAMBDA * y1 * r - 4.0 * LAMBDA * LAMBDA * x * x < 0.0;
    case 91: {                                                                       /* 7089, ipow 321 */
        double l4 = LAMBDA, t = x * x + y * y - 4.0 * LAMBDA * LAMBDA, t3 = t;
        for (k = 1; k < 4; k++) l4 *= LAMBDA;
        for (k = 1; k < 3; k++) t3 *= t;
        return 108 * l4 * x * x > t3;
    }
    }
    return o_domain_more(x, y, dom, d, box);
}

/* ---- list-driven domains of xy_in_billiard_single_domain ------------------ */

/* xy_in_polyrect, sub_wave.c:5328-5358: the closed rectangle, corners in
 * either order */
static int o_in_rect(double x, double y, const double *R, int n, int k)
{
    double x1 = R[k], y1 = R[n + k], x2 = R[2 * n + k], y2 = R[3 * n + k], t;
    if (!(x1 < x2)) { t = x1; x1 = x2; x2 = t; }
    if (!(y1 < y2)) { t = y1; y1 = y2; y2 = t; }
    if (x < x1) return 0;
    if (x > x2) return 0;
    if (y < y1) return 0;
    if (y > y2) return 0;
    return 1;
}

/* the open rectangle the maze and guide domains test inline (6593, 5914) */
static int o_in_rect_open(double x, double y, const double *R, int n, int k)
{
    return (x > R[k]) && (x < R[2 * n + k]) && (y > R[n + k]) && (y < R[3 * n + k]);
}

/* xy_in_rectrotated, sub_wave.c:5393-5416 */
static int o_in_rectrot(double x, double y, const double *R, int n, int k)
{
    const double X1 = R[k], Y1 = R[n + k], X2 = R[2 * n + k], Y2 = R[3 * n + k], W = R[4 * n + k];
    const double l = o_mod(X2 - X1, Y2 - Y1);
    if (l == 0.0) return 0;
    const double u1 = (X2 - X1) / l, u2 = (Y2 - Y1) / l, v1 = x - X1, v2 = y - Y1;
    const double pscal = u1 * v1 + u2 * v2;
    if (pscal < 0.0) return 0;
    if (pscal > l) return 0;
    const double h2 = v1 * v1 + v2 * v2 - pscal * pscal;
    return 4.0 * h2 <= W * W;
}

/* in_tpolygon, sub_wave.c:380-400 */
static int o_in_tpolygon(double x, double y, const double *P, int n, int k)
{
    const double xc = P[k], yc = P[n + k], radius = P[2 * n + k], ang = P[3 * n + k];
    const int nsides = (int)P[4 * n + k];
    int condition = 1, q;
    const double x1 = (x - xc) / radius, y1 = (y - yc) / radius;
    if (x1 * x1 + y1 * y1 >= 1.0) return 0;
    const double omega = O_DPI / ((double)nsides), cw = cos(omega * 0.5);
    for (q = 0; q < nsides; q++) {
        const double angle = ang * O_PID + ((double)q + 0.5) * omega;
        condition = condition * (x1 * cos(angle) + y1 * sin(angle) < cw);
    }
    return condition;
}

/* xy_in_triangle_tvertex, sub_wave.c:5313-5325, on polyline vertices */
static int o_tri_v(double x, double y, const orc_domain_lists *L, int a, int b, int c)
{
    const int n = L->npolyline;
    const double z1[2] = {L->polyline[a], L->polyline[n + a]}, z2[2] = {L->polyline[b], L->polyline[n + b]};
    const double z3[2] = {L->polyline[c], L->polyline[n + c]};
    return o_tri(x, y, z1, z2, z3);
}

/* the lattice discs of D_CIRCLE_LATTICE_HONEY (7169) and _RHOMBUS (7188) */
static int o_lattice_rows(double x, double y, int honey, const orc_domain_params *d, const double box[4])
{
    const double dx = (box[1] - box[0]) / (double)d->ngridx, dy = (box[3] - box[2]) / (double)d->ngridy;
    int i, j;
    for (j = 0; j < d->ngridy + 1; j++) {
        const double y1 = box[2] + ((double)j) * dy;
        for (i = 0; i < d->ngridx + 1; i++) {
            if (honey && !(i % 3 != 2 * ((j + 1) % 2))) continue;
            double x1 = box[0] + ((double)i) * dx;
            if (j % 2 == 1) x1 += 0.5 * dx;
            if (o_mod(x - x1, y - y1) < d->mu) return 1;
        }
    }
    return 0;
}

static int o_list_at(double x, double y, int dom, const orc_domain_params *d, const double box[4],
                     const orc_domain_lists *L)
{
    const double LAMBDA = d->lambda, MU = d->mu, WALL_WIDTH = d->wall_width;
    const int nr = L->nrects, nrr = L->nrectrot, nc = L->ncircles, npl = L->npolyline, NPOLY = d->npoly;
    double x1;
    int i, j, k, m, condition;
    switch (dom) {
    case 40:                                                                         /* 6254 */
        for (i = 0; i < L->npolygons; i++)
            if ((L->polygons[5 * L->npolygons + i] != 0.0) && o_in_tpolygon(x, y, L->polygons, L->npolygons, i))
                return 0;
        return 1;
    case 41: case 27:                                                                /* 6260, 7566 */
        if (dom == 27 && x * x + y * y > LAMBDA * LAMBDA) return 2;
        condition = o_tri_v(x, y, L, 0, npl / 3, 2 * npl / 3);
        m = 1;
        k = 1;
        for (i = 0; i < d->mdepth; i++) {
            m = m * 4;
            for (j = 0; j < npl / m; j++) condition += o_tri_v(x, y, L, j * m + k, j * m + 2 * k, j * m + 3 * k);
            k = k * 4;
        }
        if (dom == 27) return (condition > 0) ? 0 : 1;
        return condition >= 1;
    case 42:                                                                         /* 6274 */
        condition = o_tri_v(x, y, L, NPOLY, NPOLY - 1, 0);
        for (i = 0; i < NPOLY - 1; i++) condition += o_tri_v(x, y, L, NPOLY, i, i + 1);
        return condition >= 1;
    case 421:                                                                        /* 6281 */
        condition = (x < 0.0) && (o_abs(y) < 0.5 * WALL_WIDTH);
        for (i = 0; i < NPOLY + 3; i++) condition += o_tri_v(x, y, L, NPOLY + 3, i + 1, i);
        return condition >= 1;
    case 53: case 531: case 54: case 541:                                            /* 6590-6611 */
        for (i = 0; i < nr; i++)
            if (o_in_rect_open(x, y, L->rects, nr, i)) return 0;
        return 1;
    case 192: case 193: case 194:                                                    /* 5903, 5921, 5942 */
        if (x > 0.0) {
            if (o_abs(y) > 0.5 * LAMBDA) return 1;
            if (dom == 194 && !(o_abs(y) > 0.1 * LAMBDA)) {
                const double xa = 0.24 * LAMBDA, xb = xa + WALL_WIDTH;
                double r2 = 1.5 * 0.25 * LAMBDA;
                r2 = r2 * r2;
                const double x3 = sqrt(r2 - 0.01 * LAMBDA * LAMBDA) - xb;
                if (((x + x3) * (x + x3) + y * y < r2) && (x > xa)) return 0;
                return 1;
            }
            x1 = -y * y / LAMBDA + 0.25 * LAMBDA;
            if (dom == 192) return (x < x1) || (x > x1 + WALL_WIDTH);
            if ((x < x1) || (x > x1 + WALL_WIDTH)) return 1;
            if (dom == 193 && o_abs(y) < 0.1 * LAMBDA) return 0;
            return 2;
        }
        for (i = 0; i < nr; i++)
            if (o_in_rect_open(x, y, L->rects, nr, i)) return (dom == 192) ? 0 : 2;
        if (o_abs(y) > 0.5 * MU) return 1;
        x1 = y * y / MU - 0.25 * MU;
        if (dom == 192) return (x > x1) || (x < x1 - WALL_WIDTH);
        if ((x > x1) || (x < x1 - WALL_WIDTH)) return 1;
        return 2;
    case 931: {                                                                      /* 7128 */
        const double dx = (box[1] - box[0]) / (double)d->ngridx, dy = (box[3] - box[2]) / (double)d->ngridy;
        for (i = 0; i < d->ngridx; i++) {
            x1 = box[0] + ((double)i + 0.5) * dx;
            for (j = 0; j < d->ngridy; j++)
                if (o_mod(x - x1, y - (box[2] + ((double)j + 0.5) * dy)) < MU) return 1;
        }
        for (i = 0; i < nr; i++)
            if (o_in_rect(x, y, L->rects, nr, i)) return 1;
        return 0;
    }
    case 95: case 97:                                                                /* 7169, 7188 */
        if (o_lattice_rows(x, y, dom == 95, d, box)) return 1;
        for (i = 0; i < nrr; i++)
            if (o_in_rectrot(x, y, L->rectrot, nrr, i)) return 1;
        return 0;
    case 96: case 932: case 933: case 941:                                           /* 7207, 7215, 7301, 7311 */
        for (i = 0; i < nc; i++)
            if (o_mod(x - L->circles[i], y - L->circles[nc + i]) < MU) return 1;
        if (dom != 96)
            for (i = 0; i < nr; i++)
                if (o_in_rect(x, y, L->rects, nr, i)) return 1;
        for (i = 0; i < nrr; i++)
            if (o_in_rectrot(x, y, L->rectrot, nrr, i)) return 1;
        return 0;
    }
    return -1;
}

int orc_list_domain(int NX, int NY, const double box[4], int domain, const orc_domain_params *dp,
                    const orc_domain_lists *L, short *out)
{
    /* the vertices each polyline domain reads */
    if ((domain == 41 || domain == 27) && (L->npolyline < 1 || dp->mdepth < 0 || dp->mdepth > 15)) return -1;
    if (domain == 42 && (dp->npoly < 1 || L->npolyline < dp->npoly + 1)) return -1;
    if (domain == 421 && (dp->npoly < 0 || L->npolyline < dp->npoly + 4)) return -1;
    if (o_list_at(0.0, 0.0, domain, dp, box, L) < 0) return -1;
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            const double x = box[0] + ((double)i) * (box[1] - box[0]) / ((double)NX);
            const double y = box[2] + ((double)j) * (box[3] - box[2]) / ((double)NY);
            out[(size_t)i * NY + j] = (short)o_list_at(x, y, domain, dp, box, L);
        }
    return 0;
}

int orc_domain(int NX, int NY, const double box[4], int domain, const orc_domain_params *dp, int ncirc,
               const double *circ, short *out)
{
    if (o_domain_at(0.0, 0.0, domain, dp, box, 0, circ) < 0) return -1;
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            /* ij_to_xy, sub_wave.c:469-476 */
            const double x = box[0] + ((double)i) * (box[1] - box[0]) / ((double)NX);
            const double y = box[2] + ((double)j) * (box[3] - box[2]) / ((double)NY);
            out[(size_t)i * NY + j] = (short)o_domain_at(x, y, domain, dp, box, ncirc, circ);
        }
    return 0;
}

/* ---- init_ior_2d's per-cell models, sub_wave.c:12181-12968 ------------------
 * Each model as a loop over the cells in the reference's shape; 0 is the
 * VARIABLE_IOR == 0 branch (12984-13003). */
int orc_ior_tables(const orc_params *p, const short *xy, int ior, const double box[4], const orc_ior_params *q,
                   double *tcc, double *tc, double *tg)
{
    static const int codes[] = {0, 1, 100, 2, 10, 11, 12, 13, 14, 15, 151, 16, 17, 18, 181, 182, 183, 19, 191};
    int known = 0;
    for (size_t n = 0; n < sizeof codes / sizeof codes[0]; n++) known |= codes[n] == ior;
    if (!known) return -1;
    const int NX = p->nx, NY = p->ny;
    const double COURANT = q->courant, COURANTB = q->courantb, GAMMA = q->gamma, GAMMAB = q->gammab;
    const double LAMBDA = q->lambda, MU = q->mu, XMIN = box[0], XMAX = box[1];
    const double courant2 = COURANT * COURANT, courantb2 = COURANTB * COURANTB;
    int i, j;
    for (size_t k = 0; k < (size_t)NX * NY; k++) tcc[k] = tc[k] = tg[k] = 0.0;
    if (ior == 0) {
        for (i = 0; i < NX; i++)
            for (j = 0; j < NY; j++) {
                const size_t k = (size_t)i * NY + j;
                if (xy[k] != 0) {
                    tc[k] = COURANT;
                    tcc[k] = courant2;
                    tg[k] = xy[k] == 1 ? GAMMA : GAMMAB;
                } else if (p->twospeeds) {
                    tc[k] = COURANTB;
                    tcc[k] = courantb2;
                    tg[k] = GAMMAB;
                }
            }
        return 0;
    }
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            const size_t k = (size_t)i * NY + j;
            const double x = box[0] + ((double)i) * (box[1] - box[0]) / ((double)NX);
            const double y = box[2] + ((double)j) * (box[3] - box[2]) / ((double)NY);
            double u, v, u1, r, r2, c, speed, a;
            int n;
            switch (ior) {
            case 1: case 100:                                          /* 12183, 12220 */
                u = 0.0;
                v = 0.0;
                n = 0;
                while ((n < q->mandellevel) && (u * u + v * v < 1000.0 * q->mandellimit)) {
                    u1 = u * u - v * v + x;
                    v = 2.0 * u * v + y;
                    u = u1;
                    n++;
                }
                if (ior == 1) {
                    double norm2 = u * u + v * v;
                    if (norm2 < q->mandellimit) tcc[k] = courant2;
                    else {
                        speed = 1.0 + q->mandel_ior_scale * log(1.0 + norm2 / q->mandellimit);
                        if (speed < 0.01) speed = 0.01;
                        tcc[k] = courant2 * speed;
                    }
                } else {
                    if (n >= q->mandellevel) tcc[k] = courant2;
                    else {
                        speed = (double)n / (double)q->mandellevel;
                        if (speed < 1.0e-10) speed = 1.0e-10;
                        else if (speed > 10.0) speed = 10.0;
                        tcc[k] = courant2 * speed;
                    }
                }
                tg[k] = GAMMA;
                break;
            case 2:                                                    /* 12257 */
                r2 = x * x + y * y;
                if (r2 > 1.0) c = 0.0;
                else if (r2 < 0.25 * 0.25) c = 0.8 * COURANT;
                else if (r2 < 0.58 * 0.58) c = COURANT * (0.68 - 0.55 * r2);
                else c = COURANT * (1.3 - 0.9 * r2);
                tcc[k] = c;
                tg[k] = GAMMA;
                break;
            case 10: case 11: case 12: case 13:                        /* 12556, 12593, 12623, 12653 */
                if (ior == 10 || ior == 11) a = 0.05;
                else a = LAMBDA + 0.1 - sqrt(LAMBDA * LAMBDA - MU * MU) + (ior == 13 ? 0.6 : 0.0);
                if ((o_abs(x) < a) && (ior == 11 ? (o_abs(y) < MU) : (o_abs(y) > MU))) {
                    tcc[k] = 0.0;
                    tg[k] = GAMMA;
                } else if (xy[k] != 0) {
                    tcc[k] = courant2;
                    tg[k] = GAMMA;
                } else {
                    tcc[k] = courantb2;
                    tg[k] = GAMMAB;
                }
                break;
            case 14:                                                   /* 12683 */
                switch (xy[k]) {
                case 0: tcc[k] = courantb2; tg[k] = GAMMAB; break;
                case 1: tcc[k] = courant2; tg[k] = GAMMA; break;
                case 2: tcc[k] = 0.0; tg[k] = GAMMA; break;
                case 3: tcc[k] = courant2; tg[k] = GAMMA; break;
                }
                break;
            case 15: case 151:                                         /* 12720, 12746 */
                if (xy[k] == 1) {
                    tcc[k] = courant2;
                    tg[k] = GAMMA;
                } else if (ior == 15 ? (o_abs(x) > LAMBDA) : ((o_abs(x) > LAMBDA) && (o_abs(y) > 0.7))) {
                    tcc[k] = 0.0;
                    tg[k] = 1.0;
                } else {
                    tcc[k] = courantb2;
                    tg[k] = GAMMAB;
                }
                break;
            case 16: {                                                 /* 12772 */
                double cm = 0.8 * COURANT + 0.2 * COURANTB;
                if (xy[k] == 1) { tcc[k] = courant2; tg[k] = GAMMA; }
                else if (xy[k] == 2) { tcc[k] = cm * cm; tg[k] = GAMMA; }
                else { tcc[k] = courantb2; tg[k] = GAMMAB; }
                break;
            }
            case 17:                                                   /* 12800 */
                if (xy[k] == 0) { tcc[k] = courant2; tg[k] = GAMMA; }
                else if (xy[k] == 2) { tcc[k] = 0.0; tg[k] = 1.0; }
                else { tcc[k] = courantb2; tg[k] = GAMMAB; }
                break;
            case 18: case 181: {                                       /* 12823, 12851 */
                double ax = o_abs(x), ay = o_abs(y);
                if (ax > LAMBDA) { tcc[k] = courant2; tg[k] = GAMMA; }
                else if (ay > 1.0) { tcc[k] = 0.0; tg[k] = 0.0; }
                else if (ior == 18) { tcc[k] = courantb2 / (1.0 - MU * ay * ay); tg[k] = GAMMAB; }
                else {
                    speed = 1.0 / (1.0 - MU * ay * ay);
                    speed *= speed;
                    tcc[k] = courantb2 * speed;
                    tg[k] = GAMMAB;
                }
                break;
            }
            case 182:                                                  /* 12881 */
                r = o_mod(x, y);
                if (r > LAMBDA) { tcc[k] = courant2; tg[k] = GAMMA; }
                else { tcc[k] = courant2 / (2.0 - r / LAMBDA); tg[k] = GAMMAB; }
                break;
            case 183:                                                  /* 12904 */
                r = o_mod(x, y) / LAMBDA;
                if (r > 1.0) { tcc[k] = courant2; tg[k] = GAMMA; }
                else {
                    r2 = r * (double)q->mdepth;
                    r2 = (double)((int)r2) / (double)q->mdepth;
                    tcc[k] = courant2 / (2.0 - r2);
                    tg[k] = GAMMAB;
                }
                break;
            case 19:                                                   /* 12929 */
                speed = COURANT * (1.0 - x) + COURANTB * x;
                tcc[k] = speed * speed;
                tg[k] = GAMMA;
                break;
            case 191:                                                  /* 12945 */
                a = (x - XMIN) / (XMAX - XMIN);
                speed = COURANT * (1.0 - a) + COURANTB * a;
                tcc[k] = speed * speed;
                tg[k] = GAMMA;
                break;
            }
            /* 12970-12981 */
            if (xy[k] != 0) tc[k] = sqrt(tcc[k]);
            else if (p->twospeeds) tc[k] = COURANTB;
        }
    return 0;
}

double orc_rotating_wells(int ngridx, int ngridy, int large, double ior_angle, const double box[4], double *xc,
                          double *yc)
{
    const double s = large ? 1.5 : 1.0;
    const double XMIN = box[0], XMAX = box[1], YMIN = box[2], YMAX = box[3];
    const double dx = large ? (s * XMAX - s * XMIN) / (double)ngridx : (XMAX - XMIN) / (double)ngridx;
    const double dy = large ? (s * YMAX - s * YMIN) / (double)ngridy : (YMAX - YMIN) / (double)ngridy;
    const double ca = cos(ior_angle), sa = sin(ior_angle);
    for (int i = 0; i < ngridx; i++)
        for (int j = 0; j < ngridy; j++) {
            const int n = j * ngridx + i;
            double xs = (large ? s * XMIN : XMIN) + dx * ((double)i + 0.5);
            double ys = (large ? s * YMIN : YMIN) + dy * ((double)j + 0.5);
            if (j % 2 == 1) ys += 0.5 * dx;
            xc[n] = xs * ca + ys * sa;
            yc[n] = -xs * sa + ys * ca;
        }
    return 0.2 * dx * dx;
}

/* xy_in_billiard with COMPARISON, sub_wave.c:7593-7602: domain with circ
 * where y > 0, domain_b with circ_b where y <= 0 */
int orc_comparison_mask(int NX, int NY, const double box[4], int domain, int domain_b, const orc_domain_params *dp,
                        int nc, const double *circ, int nb, const double *circ_b, short *out)
{
    if (o_domain_at(0.0, 0.0, domain, dp, box, 0, circ) < 0 || o_domain_at(0.0, 0.0, domain_b, dp, box, 0, circ_b) < 0)
        return -1;
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++) {
            const double x = box[0] + ((double)i) * (box[1] - box[0]) / ((double)NX);
            const double y = box[2] + ((double)j) * (box[3] - box[2]) / ((double)NY);
            out[(size_t)i * NY + j] = (short)(y > 0.0 ? o_domain_at(x, y, domain, dp, box, nc, circ)
                                                      : o_domain_at(x, y, domain_b, dp, box, nb, circ_b));
        }
    return 0;
}
