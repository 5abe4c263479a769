/*
 * media_oracle.c -- CPU oracle of the Gaussian-well media tables and of the
 * billiard masks (test infrastructure, see wave_oracle.h).  Same flags as
 * wave_oracle.c; exp, tanh and sqrt are glibc's, as for the reference.
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
