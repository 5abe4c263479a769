/*
 * wave_oracle.c -- CPU oracle (test infrastructure, see wave_oracle.h).
 *
 * Built with the reference's own flags (makefile:1-4: gcc -O3 -fopenmp, no
 * -march, no -ffast-math) plus -ffp-contract=off, so every + - * below is one
 * IEEE binary64 (or binary32 for the split form) operation in source order.
 */
#include "wave_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

int orc_num_threads(void) { return omp_get_max_threads(); }
/* OpenMP threads of later parallel regions; returns the previous count */
int orc_set_num_threads(int n)
{
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

#define ACTIVE_BULK(p, v)  ((p)->twospeeds || (v) > 0)     /* wave_billiard.c:365 */
#define ACTIVE_EDGE(p, v)  ((p)->twospeeds || (v) != 0)    /* wave_billiard.c:395,436,467,525 */

/* leapfrog update, wave_billiard.c:373: -y + 2*x + tcc*delta - KAPPA*x - tgamma*(x-y) */
static inline double leap(double x, double y, double delta, double cc, double kappa, double g)
{
    return -y + 2 * x + cc * delta - kappa * x - g * (x - y);
}

/* first-order absorbing edge, wave_billiard.c:418: x - tc*(x - inward) - K*x - G*(x-y) */
static inline double absorb(double x, double y, double inward, double c, double k, double g)
{
    return x - c * (x - inward) - k * x - g * (x - y);
}

/* Bulk 1<=i<=NX-2, 1<=j<=NY-2 -- wave_billiard.c:361-377 (5-point); the
 * 9-point variant has no reference counterpart. */
static void bulk_step(const orc_params *p, double *const *phi_in, double *const *psi_in,
                      double *const *phi_out, short *const *xy_in, double *const *tcc,
                      double *const *tgamma)
{
    const int NX = p->nx, NY = p->ny;
    const double KAPPA = p->kappa;
    int i, j;
    double delta, x, y;
    if (p->stencil == 9) {
#pragma omp parallel for private(i, j, delta, x, y)
        for (i = 1; i < NX - 1; i++)
            for (j = 1; j < NY - 1; j++)
                if (ACTIVE_BULK(p, xy_in[i][j])) {
                    x = phi_in[i][j];
                    y = psi_in[i][j];
                    double s4 = phi_in[i + 1][j] + phi_in[i - 1][j] + phi_in[i][j + 1] + phi_in[i][j - 1];
                    double sd = phi_in[i + 1][j + 1] + phi_in[i - 1][j + 1] + phi_in[i + 1][j - 1]
                                + phi_in[i - 1][j - 1];
                    delta = (4.0 * s4 + sd - 20.0 * x) * ORC_W9;
                    phi_out[i][j] = leap(x, y, delta, tcc[i][j], KAPPA, tgamma[i][j]);
                }
        return;
    }
#pragma omp parallel for private(i, j, delta, x, y)
    for (i = 1; i < NX - 1; i++)
        for (j = 1; j < NY - 1; j++)
            if (ACTIVE_BULK(p, xy_in[i][j])) {
                x = phi_in[i][j];
                y = psi_in[i][j];
                delta = phi_in[i + 1][j] + phi_in[i - 1][j] + phi_in[i][j + 1] + phi_in[i][j - 1] - 4.0 * x;
                phi_out[i][j] = leap(x, y, delta, tcc[i][j], KAPPA, tgamma[i][j]);
            }
}

/* Edge rows and columns -- wave_billiard.c:393-583 (row-pointer) and
 * wave_3d.c:649-871 (flat; BC_ABS_REFLECT). Corners belong to the top
 * (j=NY-1) and bottom (j=0) loops, as in the reference. */
static void edge_step(const orc_params *p, double *const *phi_in, double *const *psi_in,
                      double *const *phi_out, short *const *xy_in, double *const *tcc,
                      double *const *tc, double *const *tgamma)
{
    const int NX = p->nx, NY = p->ny, B = p->bcond;
    const double KAPPA = p->kappa, KS = p->kappa_sides, GS = p->gamma_sides;
    const double KT = p->kappa_topbot, GT = p->gamma_topbot;
    int i, j, iplus, iminus;
    double delta, x, y;
    /* COMPARISON (wave_3d.c:737-738, 806-807): the periodic wraps of the top and
     * bottom rows stay inside their half */
    const int jbot = p->comparison ? NY / 2 : 0, jtop = p->comparison ? NY / 2 - 1 : NY - 1;

    /* left boundary, wave_billiard.c:393-430 (OSCILLATE_LEFT not supported) */
    for (j = 1; j < NY - 1; j++)
        if (ACTIVE_EDGE(p, xy_in[0][j])) {
            x = phi_in[0][j];
            y = psi_in[0][j];
            switch (B) {
            case ORC_BC_DIRICHLET:
                delta = phi_in[1][j] + phi_in[0][j + 1] + phi_in[0][j - 1] - 3.0 * x;
                phi_out[0][j] = leap(x, y, delta, tcc[0][j], KAPPA, tgamma[0][j]);
                break;
            case ORC_BC_PERIODIC:
                delta = phi_in[1][j] + phi_in[NX - 1][j] + phi_in[0][j + 1] + phi_in[0][j - 1] - 4.0 * x;
                phi_out[0][j] = leap(x, y, delta, tcc[0][j], KAPPA, tgamma[0][j]);
                break;
            default: /* ABSORBING, VPER_HABS, ABS_REFLECT */
                phi_out[0][j] = absorb(x, y, phi_in[1][j], tc[0][j], KS, GS);
                break;
            }
        }

    /* right boundary, wave_billiard.c:433-467 */
    for (j = 1; j < NY - 1; j++)
        if (ACTIVE_EDGE(p, xy_in[NX - 1][j])) {
            x = phi_in[NX - 1][j];
            y = psi_in[NX - 1][j];
            switch (B) {
            case ORC_BC_DIRICHLET:
                delta = phi_in[NX - 2][j] + phi_in[NX - 1][j + 1] + phi_in[NX - 1][j - 1] - 3.0 * x;
                phi_out[NX - 1][j] = leap(x, y, delta, tcc[NX - 1][j], KAPPA, tgamma[NX - 1][j]);
                break;
            case ORC_BC_PERIODIC:
                delta = phi_in[NX - 2][j] + phi_in[0][j] + phi_in[NX - 1][j + 1] + phi_in[NX - 1][j - 1] - 4.0 * x;
                phi_out[NX - 1][j] = leap(x, y, delta, tcc[NX - 1][j], KAPPA, tgamma[NX - 1][j]);
                break;
            default:
                phi_out[NX - 1][j] = absorb(x, y, phi_in[NX - 2][j], tc[NX - 1][j], KS, GS);
                break;
            }
        }

    /* top boundary j = NY-1, wave_billiard.c:470-525 */
    for (i = 0; i < NX; i++)
        if (ACTIVE_EDGE(p, xy_in[i][NY - 1])) {
            x = phi_in[i][NY - 1];
            y = psi_in[i][NY - 1];
            if (p->oscillate_topbot && i < p->tb_shift && i < NX - 1 && i > 0) {
                delta = phi_in[i + 1][NY - 1] + phi_in[i - 1][NY - 1] + -2.0 * x;
                phi_out[i][NY - 1] = leap(x, y, delta, tcc[i][NY - 1], KAPPA, tgamma[i][NY - 1]);
            } else
                switch (B) {
                case ORC_BC_DIRICHLET:
                    iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                    iminus = i - 1; if (iminus == -1) iminus = 0;
                    delta = phi_in[iplus][NY - 1] + phi_in[iminus][NY - 1] + phi_in[i][NY - 2] - 3.0 * x;
                    phi_out[i][NY - 1] = leap(x, y, delta, tcc[i][NY - 1], KAPPA, tgamma[i][NY - 1]);
                    break;
                case ORC_BC_PERIODIC:
                    iplus = (i + 1) % NX;
                    iminus = (i - 1) % NX; if (iminus < 0) iminus += NX;
                    delta = phi_in[iplus][NY - 1] + phi_in[iminus][NY - 1] + phi_in[i][NY - 2] + phi_in[i][jbot] - 4.0 * x;
                    phi_out[i][NY - 1] = leap(x, y, delta, tcc[i][NY - 1], KAPPA, tgamma[i][NY - 1]);
                    break;
                case ORC_BC_VPER_HABS:
                    iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                    iminus = i - 1; if (iminus == -1) iminus = 0;
                    delta = phi_in[iplus][NY - 1] + phi_in[iminus][NY - 1] + phi_in[i][NY - 2] + phi_in[i][jbot] - 4.0 * x;
                    if (i == 0)
                        phi_out[0][NY - 1] = absorb(x, y, phi_in[1][NY - 1], tc[0][NY - 1], KS, GS);
                    else
                        phi_out[i][NY - 1] = leap(x, y, delta, tcc[i][NY - 1], KAPPA, tgamma[i][NY - 1]);
                    break;
                default: /* ABSORBING, ABS_REFLECT */
                    phi_out[i][NY - 1] = absorb(x, y, phi_in[i][NY - 2], tc[i][NY - 1], KT, GT);
                    break;
                }
        }

    /* bottom boundary j = 0, wave_billiard.c:528-583 */
    for (i = 0; i < NX; i++)
        if (ACTIVE_EDGE(p, xy_in[i][0])) {
            x = phi_in[i][0];
            y = psi_in[i][0];
            if (p->oscillate_topbot && i < p->tb_shift && i < NX - 1 && i > 0) {
                delta = phi_in[i + 1][0] + phi_in[i - 1][0] + -2.0 * x;
                phi_out[i][0] = leap(x, y, delta, tcc[i][0], KAPPA, tgamma[i][0]);
            } else
                switch (B) {
                case ORC_BC_DIRICHLET:
                    iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                    iminus = i - 1; if (iminus == -1) iminus = 0;
                    delta = phi_in[iplus][0] + phi_in[iminus][0] + phi_in[i][1] - 3.0 * x;
                    phi_out[i][0] = leap(x, y, delta, tcc[i][0], KAPPA, tgamma[i][0]);
                    break;
                case ORC_BC_PERIODIC:
                    iplus = (i + 1) % NX;
                    iminus = (i - 1) % NX; if (iminus < 0) iminus += NX;
                    delta = phi_in[iplus][0] + phi_in[iminus][0] + phi_in[i][1] + phi_in[i][jtop] - 4.0 * x;
                    phi_out[i][0] = leap(x, y, delta, tcc[i][0], KAPPA, tgamma[i][0]);
                    break;
                case ORC_BC_VPER_HABS:
                    iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                    iminus = i - 1; if (iminus == -1) iminus = 0;
                    delta = phi_in[iplus][0] + phi_in[iminus][0] + phi_in[i][1] + phi_in[i][jtop] - 4.0 * x;
                    if (i == 0)
                        phi_out[0][0] = absorb(x, y, phi_in[1][0], tc[0][0], KS, GS);
                    else
                        phi_out[i][0] = leap(x, y, delta, tcc[i][0], KAPPA, tgamma[i][0]);
                    break;
                default:
                    phi_out[i][0] = absorb(x, y, phi_in[i][1], tc[i][0], KT, GT);
                    break;
                }
        }
    if (!p->comparison) return;

    /* COMPARISON, top of the bottom half j = jtop, wave_3d.c:873-929 */
    for (i = 0; i < NX; i++)
        if (ACTIVE_EDGE(p, xy_in[i][jtop])) {
            x = phi_in[i][jtop];
            y = psi_in[i][jtop];
            switch (B) {
            case ORC_BC_PERIODIC:
                iplus = (i + 1) % NX;
                iminus = (i - 1) % NX; if (iminus < 0) iminus += NX;
                delta = phi_in[iplus][jtop] + phi_in[iminus][jtop] + phi_in[i][jtop - 1] + phi_in[i][0] - 4.0 * x;
                phi_out[i][jtop] = leap(x, y, delta, tcc[i][jtop], KAPPA, tgamma[i][jtop]);
                break;
            case ORC_BC_ABSORBING:
                phi_out[i][jtop] = absorb(x, y, phi_in[i][jtop - 1], tc[i][jtop], KT, GT);
                break;
            case ORC_BC_VPER_HABS:
                iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                iminus = i - 1; if (iminus == -1) iminus = 0;
                delta = phi_in[iplus][jtop] + phi_in[iminus][jtop] + phi_in[i][jtop - 1] + phi_in[i][0] - 4.0 * x;
                if (i == 0) phi_out[0][jtop] = absorb(x, y, phi_in[1][jtop], tc[0][jtop], KS, GS);
                else phi_out[i][jtop] = leap(x, y, delta, tcc[i][jtop], KAPPA, tgamma[i][jtop]);
                break;
            default: /* DIRICHLET, ABS_REFLECT */
                iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                iminus = i - 1; if (iminus == -1) iminus = 0;
                delta = phi_in[iplus][jtop] + phi_in[iminus][jtop] + phi_in[i][jtop - 1] - 3.0 * x;
                phi_out[i][jtop] = leap(x, y, delta, tcc[i][jtop], KAPPA, tgamma[i][jtop]);
                break;
            }
        }

    /* COMPARISON, bottom of the top half j = jbot, wave_3d.c:931-982; the
     * periodic branch wraps to the cell itself and the VPER_HABS corner reads
     * phi_in[NY] (cell (1, 0)), as the reference does */
    for (i = 0; i < NX; i++)
        if (ACTIVE_EDGE(p, xy_in[i][jbot])) {
            x = phi_in[i][jbot];
            y = psi_in[i][jbot];
            switch (B) {
            case ORC_BC_PERIODIC:
                iplus = (i + 1) % NX;
                iminus = (i - 1) % NX; if (iminus < 0) iminus += NX;
                delta = phi_in[iplus][jbot] + phi_in[iminus][jbot] + phi_in[i][jbot + 1] + phi_in[i][jbot] - 4.0 * x;
                phi_out[i][jbot] = leap(x, y, delta, tcc[i][jbot], KAPPA, tgamma[i][jbot]);
                break;
            case ORC_BC_ABSORBING:
                phi_out[i][jbot] = absorb(x, y, phi_in[i][jbot + 1], tc[i][jbot], KT, GT);
                break;
            case ORC_BC_VPER_HABS:
                iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                iminus = i - 1; if (iminus == -1) iminus = 0;
                delta = phi_in[iplus][jbot] + phi_in[iminus][jbot] + phi_in[i][jbot + 1] + phi_in[i][NY - 1] - 4.0 * x;
                if (i == 0) phi_out[0][jbot] = absorb(x, y, phi_in[1][0], tc[0][jbot], KS, GS);
                else phi_out[i][jbot] = leap(x, y, delta, tcc[i][jbot], KAPPA, tgamma[i][jbot]);
                break;
            default: /* DIRICHLET, ABS_REFLECT */
                iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
                iminus = i - 1; if (iminus == -1) iminus = 0;
                delta = phi_in[iplus][jbot] + phi_in[iminus][jbot] + phi_in[i][jbot + 1] - 3.0 * x;
                phi_out[i][jbot] = leap(x, y, delta, tcc[i][jbot], KAPPA, tgamma[i][jbot]);
                break;
            }
        }
}

void orc_evolve_wave_half(const orc_params *p, double *const *phi_in, double *const *psi_in,
                          double *const *phi_out, short *const *xy_in, double *const *tcc,
                          double *const *tc, double *const *tgamma)
{
    bulk_step(p, phi_in, psi_in, phi_out, xy_in, tcc, tgamma);
    edge_step(p, phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma);
}

/* row-pointer view of a flat k = i*NY + j array */
#define ROWS(name, base, type)                                   \
    type **name = (type **)malloc(sizeof(type *) * (size_t)p->nx); \
    for (int r_ = 0; r_ < p->nx; r_++) name[r_] = (type *)(base) + (size_t)r_ * p->ny

void orc_evolve_wave_half_flat(const orc_params *p, double *phi_in, double *psi_in,
                               double *phi_out, short *xy_in, double *tc, double *tcc,
                               double *tgamma)
{
    ROWS(a, phi_in, double); ROWS(b, psi_in, double); ROWS(c, phi_out, double);
    ROWS(m, xy_in, short); ROWS(t1, tcc, double); ROWS(t2, tc, double); ROWS(t3, tgamma, double);
    orc_evolve_wave_half(p, a, b, c, m, t1, t2, t3);
    free(a); free(b); free(c); free(m); free(t1); free(t2); free(t3);
}

void orc_evolve_wave_flat(const orc_params *p, double *phi, double *psi, double *tmp,
                          short *xy_in, double *tcc, double *tc, double *tgamma)
{
    /* wave_billiard.c:623-627 */
    orc_evolve_wave_half_flat(p, phi, psi, tmp, xy_in, tc, tcc, tgamma);
    orc_evolve_wave_half_flat(p, tmp, phi, psi, xy_in, tc, tcc, tgamma);
    orc_evolve_wave_half_flat(p, psi, tmp, phi, xy_in, tc, tcc, tgamma);
}

void orc_run(const orc_params *p, double *b0, double *b1, double *b2, int *state,
             long nsteps, const short *xy_in, const double *tcc, const double *tc,
             const double *tgamma, int nsrc, const int *si, const int *sj,
             const double *amp, long amp_len, long step0)
{
    double **R[3];
    double *base[3] = {b0, b1, b2};
    for (int k = 0; k < 3; k++) {
        R[k] = (double **)malloc(sizeof(double *) * (size_t)p->nx);
        for (int r = 0; r < p->nx; r++) R[k][r] = base[k] + (size_t)r * p->ny;
    }
    ROWS(m, xy_in, short); ROWS(t1, tcc, double); ROWS(t2, tc, double); ROWS(t3, tgamma, double);
    int s = *state;
    for (long n = 0; n < nsteps; n++) {
        double **phi = R[s], **psi = R[(s + 1) % 3], **tmp = R[(s + 2) % 3];
        orc_evolve_wave_half(p, phi, psi, tmp, m, t1, t2, t3);
        long g = step0 + n; /* amplitude index of this step */
        if (g < amp_len)
            for (int k = 0; k < nsrc; k++) tmp[si[k]][sj[k]] += amp[(size_t)k * amp_len + g];
        s = (s + 2) % 3;
    }
    *state = s;
    for (int k = 0; k < 3; k++) free(R[k]);
    free(m); free(t1); free(t2); free(t3);
}

/* ---------------------------------------------------------------- energy -- */

double orc_energy_cell(const orc_params *p, const double *phi, const double *psi,
                       const short *xy_in, double courant, double courantb,
                       double e_scale, int i, int j)
{
    /* wave_common.c:453-472 on flat arrays */
    const int NX = p->nx, NY = p->ny;
#define PH(a, b) phi[(size_t)(a) * NY + (b)]
    double velocity = PH(i, j) - psi[(size_t)i * NY + j];
    int iplus = i + 1;  if (iplus == NX) iplus = NX - 1;
    int iminus = i - 1; if (iminus == -1) iminus = 0;
    int jplus = j + 1;  if (jplus == NY) jplus = NY - 1;
    int jminus = j - 1; if (jminus == -1) jminus = 0;
    double gx2 = (PH(iplus, j) - PH(i, j)) * (PH(iplus, j) - PH(i, j))
                 + (PH(i, j) - PH(iminus, j)) * (PH(i, j) - PH(iminus, j));
    double gy2 = (PH(i, jplus) - PH(i, j)) * (PH(i, jplus) - PH(i, j))
                 + (PH(i, j) - PH(i, jminus)) * (PH(i, j) - PH(i, jminus));
#undef PH
    if (xy_in[(size_t)i * NY + j])
        return e_scale * e_scale * (velocity * velocity + 0.5 * courant * courant * (gx2 + gy2));
    else if (p->twospeeds)
        return e_scale * e_scale * (velocity * velocity + 0.5 * courantb * courantb * (gx2 + gy2));
    return 0.0;
}

static double tree64(const double *e, long n, long stride)
{
    double acc[64];
    for (int l = 0; l < 64; l++) {
        acc[l] = 0.0;
        for (long k = l; k < n; k += 64) acc[l] += e[k * stride];
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) acc[l] = acc[l] + acc[l + off];
    return acc[0];
}

double orc_fixed_order_sum(int nx, int ny, const double *e)
{
    double *rows = (double *)malloc(sizeof(double) * (size_t)nx);
    for (int i = 0; i < nx; i++) rows[i] = tree64(e + (size_t)i * ny, ny, 1);
    double t = tree64(rows, nx, 1);
    free(rows);
    return t;
}

/* per-cell energy with c^2 from the tcc table (build-defined total energy) */
static void energy_field(const orc_params *p, const double *phi, const double *psi,
                         const short *xy_in, const double *tcc, double e_scale, double *e)
{
    const int NX = p->nx, NY = p->ny;
    const double es2 = e_scale * e_scale;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            size_t k = (size_t)i * NY + j;
            if (!(xy_in[k] != 0 || p->twospeeds)) { e[k] = 0.0; continue; }
            int ip = i + 1 == NX ? NX - 1 : i + 1, im = i == 0 ? 0 : i - 1;
            int jp = j + 1 == NY ? NY - 1 : j + 1, jm = j == 0 ? 0 : j - 1;
            /* COMPARISON: no gradient across the halves, compute_energy_mod
             * wave_common.c:1965-1971 */
            if (p->comparison && jp == NY / 2) jp = NY / 2 - 1;
            if (p->comparison && jm == NY / 2 - 1) jm = NY / 2;
            double c = phi[k], v = c - psi[k];
            double a1 = phi[(size_t)ip * NY + j] - c, a2 = c - phi[(size_t)im * NY + j];
            double b1 = phi[(size_t)i * NY + jp] - c, b2 = c - phi[(size_t)i * NY + jm];
            double gx2 = a1 * a1 + a2 * a2, gy2 = b1 * b1 + b2 * b2;
            e[k] = es2 * (v * v + 0.5 * tcc[k] * (gx2 + gy2));
        }
}

double orc_energy_total(const orc_params *p, const double *phi, const double *psi,
                        const short *xy_in, const double *tcc, double e_scale)
{
    size_t n = (size_t)p->nx * p->ny;
    double *e = (double *)malloc(sizeof(double) * n);
    energy_field(p, phi, psi, xy_in, tcc, e_scale, e);
    double t = orc_fixed_order_sum(p->nx, p->ny, e);
    free(e);
    return t;
}

void orc_run_x(const orc_params *p, double *b0, double *b1, double *b2, int *state,
               long nsteps, const short *xy_in, const double *tcc, const double *tc,
               const double *tgamma, int nsrc, const int *si, const int *sj,
               const double *amp, long amp_len, long step0, const orc_extras *x)
{
    double **R[3];
    double *base[3] = {b0, b1, b2};
    for (int k = 0; k < 3; k++) {
        R[k] = (double **)malloc(sizeof(double *) * (size_t)p->nx);
        for (int r = 0; r < p->nx; r++) R[k][r] = base[k] + (size_t)r * p->ny;
    }
    ROWS(m, xy_in, short); ROWS(t1, tcc, double); ROWS(t2, tc, double); ROWS(t3, tgamma, double);
    int s = *state;
    for (long n = 0; n < nsteps; n++) {
        double **phi = R[s], **psi = R[(s + 1) % 3], **tmp = R[(s + 2) % 3];
        const long g = step0 + n; /* amplitude index of this step */
        bulk_step(p, phi, psi, tmp, m, t1, t3);
        /* wave_billiard.c:379-389 */
        for (int k = 0; k < x->n_neu; k++) {
            double v = x->nnx[k] * phi[x->nbi1[k]][x->nbj1[k]];
            v += x->nny[k] * phi[x->nbi2[k]][x->nbj2[k]];
            tmp[x->ni[k]][x->nj[k]] = v;
        }
        edge_step(p, phi, psi, tmp, m, t1, t2, t3);
        /* wave_billiard.c:397, 585-590 */
        if (g < x->dlen)
            for (int k = 0; k < x->n_drv; k++) tmp[x->di[k]][x->dj[k]] = x->dval[(size_t)k * x->dlen + g];
        /* wave_billiard.c:592-609 */
        if (x->vmax > 0.0)
            for (int i = 0; i < p->nx; i++)
                for (int j = 0; j < p->ny; j++)
                    if (m[i][j] != 0) {
                        if (tmp[i][j] > x->vmax) tmp[i][j] = x->vmax;
                        else if (tmp[i][j] < -x->vmax) tmp[i][j] = -x->vmax;
                    }
        if (g < amp_len)
            for (int k = 0; k < nsrc; k++) tmp[si[k]][sj[k]] += amp[(size_t)k * amp_len + g];
        s = (s + 2) % 3;
    }
    *state = s;
    for (int k = 0; k < 3; k++) free(R[k]);
    free(m); free(t1); free(t2); free(t3);
}

/* ------------------------------------------------- neighbour-table path -- */

/* one table entry: n neighbours (a, b, c, d), flat indices i*NY + j */
static void tbl_set(const orc_params *p, int *nn, int *nb, int i, int j, int n, int ia, int ja, int ib,
                    int jb, int ic, int jc, int id, int jd)
{
    const int NY = p->ny;
    const long k = (long)i * NY + j;
    nn[k] = n;
    nb[4 * k] = ia * NY + ja;
    nb[4 * k + 1] = ib * NY + jb;
    nb[4 * k + 2] = ic * NY + jc;
    nb[4 * k + 3] = n == 4 ? id * NY + jd : -1;
}

void orc_laplacian_table(const orc_params *p, const int lsh[6], int *nn, int *nb)
{
    const int NX = p->nx, NY = p->ny;
    for (long k = 0; k < (long)NX * NY; k++) {
        nn[k] = 0;
        nb[4 * k] = nb[4 * k + 1] = nb[4 * k + 2] = nb[4 * k + 3] = -1;
    }
    /* bulk, sub_wave.c:11563-11572: (i+1, j), (i-1, j), (i, j+1), (i, j-1) */
    for (int i = 1; i < NX - 1; i++)
        for (int j = 1; j < NY - 1; j++) tbl_set(p, nn, nb, i, j, 4, i + 1, j, i - 1, j, i, j + 1, i, j - 1);
    const int bc = p->bcond;
    if (bc >= ORC_BC_DIRICHLET && bc <= ORC_BC_VPER_HABS) {
        const int per = bc == ORC_BC_PERIODIC;
        /* left and right edges, 1 <= j <= NY-2 (11577-11594, 11625-11646,
         * 11680-11697, 11729-11746): three neighbours, periodic adds the wrap */
        for (int j = 1; j < NY - 1; j++) {
            if (per) {
                tbl_set(p, nn, nb, 0, j, 4, 1, j, NX - 1, j, 0, j + 1, 0, j - 1);
                tbl_set(p, nn, nb, NX - 1, j, 4, NX - 2, j, 0, j, NX - 1, j + 1, NX - 1, j - 1);
            } else {
                tbl_set(p, nn, nb, 0, j, 3, 1, j, 0, j + 1, 0, j - 1, 0, 0);
                tbl_set(p, nn, nb, NX - 1, j, 3, NX - 2, j, NX - 1, j + 1, NX - 1, j - 1, 0, 0);
            }
        }
        /* top (j = NY-1) and bottom (j = 0) rows, corners included: i+-1
         * clamped, or wrapped for periodic; periodic and VPER_HABS add the
         * neighbour across the y seam (11596-11620, 11648-11675, 11699-11724,
         * 11748-11775) */
        const int four = per || bc == ORC_BC_VPER_HABS;
        for (int i = 0; i < NX; i++) {
            int ip, im;
            if (per) {
                ip = (i + 1) % NX;
                im = (i - 1 + NX) % NX;
            } else {
                ip = i + 1 == NX ? NX - 1 : i + 1;
                im = i == 0 ? 0 : i - 1;
            }
            tbl_set(p, nn, nb, i, NY - 1, four ? 4 : 3, ip, NY - 1, im, NY - 1, i, NY - 2, i, 0);
            tbl_set(p, nn, nb, i, 0, four ? 4 : 3, ip, 0, im, 0, i, 1, i, NY - 1);
        }
    } else if (bc == ORC_BC_LSHAPE) {
        /* sub_wave.c:11778-11889, in the reference's order (later entries win) */
        const int i1 = lsh[0], j1 = lsh[1], i2 = lsh[2], j2 = lsh[3], i3 = lsh[4], j3 = lsh[5];
        for (int j = j1 + 1; j < j2; j++) tbl_set(p, nn, nb, i1, j, 4, i1 + 1, j, i3, j, i1, j + 1, i1, j - 1);
        for (int j = j2; j < j3 - 1; j++) tbl_set(p, nn, nb, i1, j, 4, i1 + 1, j, i2 - 1, j, i1, j + 1, i1, j - 1);
        for (int j = j1 + 1; j < j2; j++) tbl_set(p, nn, nb, i3, j, 4, i1, j, i3 - 1, j, i3, j + 1, i3, j - 1);
        for (int j = j2; j < j3 - 1; j++) tbl_set(p, nn, nb, i2, j, 4, i1, j, i2 - 1, j, i2, j + 1, i2, j - 1);
        for (int i = i1; i < i2; i++) tbl_set(p, nn, nb, i, j3, 4, i + 1, j3, i - 1, j3, i, j1, i, j3 - 1);
        for (int i = i2; i < i3; i++) tbl_set(p, nn, nb, i, j2, 4, i + 1, j2, i - 1, j2, i, j1, i, j2 - 1);
        for (int i = i1; i < i2; i++) tbl_set(p, nn, nb, i, j1, 4, i + 1, j1, i - 1, j1, i, j1 + 1, i, j3);
        for (int i = i2; i < i3; i++) tbl_set(p, nn, nb, i, j1, 4, i + 1, j1, i - 1, j1, i, j1 + 1, i, j2);
        tbl_set(p, nn, nb, i1, j1, 4, i1 + 1, j1, i3 - 1, j1, i1, j1 + 1, i1, j3 - 1);
        tbl_set(p, nn, nb, i3, j1, 4, i1, j1, i3 - 1, j1, i3, j1 + 1, i3, j3);
        tbl_set(p, nn, nb, i1, j3, 4, i1 + 1, j3, i3, j3, i1, j1, i1, j3 - 1);
    }
}

/* -n*x plus the n table neighbours in order (sub_wave.c:11927-11962) */
static inline double tbl_delta(const double *phi, const int *nn, const int *nb, long k, int nfix)
{
    const int n = nfix ? nfix : nn[k];
    double d = -(double)n * phi[k];
    for (int q = 0; q < n; q++) d += phi[nb[4 * k + q]];
    return d;
}

/* one step of evolve_wave_half_new, wave_3d.c:345-597 (Laplacian of
 * compute_laplacian, sub_wave.c:11895-11963) */
static void table_step(const orc_params *p, const double *phi_in, const double *psi_in, double *phi_out,
                       const short *xy, const double *tc, const double *tcc, const double *tg,
                       const int *nn, const int *nb)
{
    const int NX = p->nx, NY = p->ny, bc = p->bcond;
    const double KAPPA = p->kappa, KS = p->kappa_sides, GS = p->gamma_sides;
    const double KT = p->kappa_topbot, GT = p->gamma_topbot;
    const int lshape = bc == ORC_BC_LSHAPE;
    double *delta = (double *)malloc(sizeof(double) * (size_t)NX * NY);
    int i, j;
#pragma omp parallel for private(i, j)
    for (i = 1; i < NX - 1; i++)
        for (j = 1; j < NY - 1; j++) {
            const long k = (long)i * NY + j;
            delta[k] = lshape ? tbl_delta(phi_in, nn, nb, k, 4)
                              : phi_in[k + NY] + phi_in[k - NY] + phi_in[k + 1] + phi_in[k - 1] - 4.0 * phi_in[k];
        }
    for (i = 0; i < NX; i++) {
        delta[(long)i * NY] = tbl_delta(phi_in, nn, nb, (long)i * NY, 0);
        delta[(long)i * NY + NY - 1] = tbl_delta(phi_in, nn, nb, (long)i * NY + NY - 1, 0);
    }
    for (j = 1; j < NY - 1; j++) {
        delta[j] = tbl_delta(phi_in, nn, nb, j, 0);
        delta[(long)(NX - 1) * NY + j] = tbl_delta(phi_in, nn, nb, (long)(NX - 1) * NY + j, 0);
    }
    /* bulk, wave_3d.c:390-402 */
#pragma omp parallel for private(i, j)
    for (i = 1; i < NX - 1; i++)
        for (j = 1; j < NY - 1; j++) {
            const long k = (long)i * NY + j;
            if (p->twospeeds || xy[k] != 0) phi_out[k] = leap(phi_in[k], psi_in[k], delta[k], tcc[k], KAPPA, tg[k]);
        }
    /* left and right edges, wave_3d.c:404-482 */
    for (int side = 0; side < 2; side++) {
        const int ie = side ? NX - 1 : 0, in = side ? NX - 2 : 1;
        for (j = 1; j < NY - 1; j++) {
            const long k = (long)ie * NY + j;
            if (!(p->twospeeds || xy[k] != 0)) continue;
            const double x = phi_in[k], y = psi_in[k];
            if (bc == ORC_BC_DIRICHLET || bc == ORC_BC_PERIODIC)
                phi_out[k] = leap(x, y, delta[k], tcc[k], KAPPA, tg[k]);
            else if (bc >= ORC_BC_ABSORBING && bc <= ORC_BC_ABS_REFLECT)
                phi_out[k] = absorb(x, y, phi_in[(long)in * NY + j], tc[k], KS, GS);
        }
    }
    /* top (j = NY-1) and bottom (j = 0) rows, wave_3d.c:484-572 */
    for (int side = 0; side < 2; side++) {
        const int je = side ? 0 : NY - 1, jn = side ? 1 : NY - 2;
        for (i = 0; i < NX; i++) {
            const long k = (long)i * NY + je;
            if (!(p->twospeeds || xy[k] != 0)) continue;
            const double x = phi_in[k], y = psi_in[k];
            if (p->oscillate_topbot && i < p->tb_shift && i < NX - 1 && i > 0) {
                phi_out[k] = leap(x, y, delta[k], tcc[k], KAPPA, tg[k]);
                continue;
            }
            switch (bc) {
            case ORC_BC_DIRICHLET:
            case ORC_BC_PERIODIC:
                phi_out[k] = leap(x, y, delta[k], tcc[k], KAPPA, tg[k]);
                break;
            case ORC_BC_ABSORBING:
            case ORC_BC_ABS_REFLECT:
                phi_out[k] = absorb(x, y, phi_in[(long)i * NY + jn], tc[k], KT, GT);
                break;
            case ORC_BC_VPER_HABS:
                if (i == 0) phi_out[k] = absorb(x, y, phi_in[(long)NY + je], tc[k], KS, GS);
                else phi_out[k] = leap(x, y, delta[k], tcc[k], KAPPA, tg[k]);
                break;
            default: /* BC_LSHAPE: no edge formula */
                break;
            }
        }
    }
    free(delta);
}

void orc_run_table(const orc_params *p, double *b0, double *b1, double *b2, int *state,
                   long nsteps, const short *xy_in, const double *tcc, const double *tc,
                   const double *tgamma, const int *nneighb, const int *nghb, int nsrc,
                   const int *si, const int *sj, const double *amp, long amp_len, long step0,
                   const orc_extras *x)
{
    double *R[3] = {b0, b1, b2};
    const long NY = p->ny, n = (long)p->nx * p->ny;
    int s = *state;
    for (long m = 0; m < nsteps; m++) {
        double *phi = R[s], *psi = R[(s + 1) % 3], *tmp = R[(s + 2) % 3];
        const long g = step0 + m;
        table_step(p, phi, psi, tmp, xy_in, tc, tcc, tgamma, nneighb, nghb);
        /* driven cells: the left edge (wave_3d.c:413) and the corners (574-579) */
        if (x && g < x->dlen)
            for (int k = 0; k < x->n_drv; k++) tmp[x->di[k] * NY + x->dj[k]] = x->dval[(size_t)k * x->dlen + g];
        /* FLOOR, wave_3d.c:581-596 */
        if (x && x->vmax > 0.0)
            for (long k = 0; k < n; k++)
                if (xy_in[k] != 0) {
                    if (tmp[k] > x->vmax) tmp[k] = x->vmax;
                    if (tmp[k] < -x->vmax) tmp[k] = -x->vmax;
                }
        if (g < amp_len)
            for (int k = 0; k < nsrc; k++) tmp[si[k] * NY + sj[k]] += amp[(size_t)k * amp_len + g];
        s = (s + 2) % 3;
    }
    *state = s;
}

/* ------------------------------------------------------------ split form -- */

void orc_split_from_fp64(long n, const double *u_n, const double *u_nm1, float *hi, float *lo,
                         float *v)
{
    for (long k = 0; k < n; k++) {
        float h = (float)u_n[k];
        hi[k] = h;
        lo[k] = (float)(u_n[k] - (double)h);
        v[k] = (float)(u_n[k] - u_nm1[k]);
    }
}

void orc_split_to_fp64(long n, const float *hi, const float *lo, const float *v, double *u_n,
                       double *u_nm1)
{
    for (long k = 0; k < n; k++) {
        double u = (double)hi[k] + (double)lo[k];
        u_n[k] = u;
        u_nm1[k] = u - (double)v[k];
    }
}

/* u + w for u = h + l: w takes the low word first (t = l + w, rounded at the
 * increment's own precision), then one Fast2Sum with h */
static inline void dd_add_f(float h, float l, float w, float *ho, float *lo)
{
    float t = l + w;
    float hn = h + t;
    *ho = hn;
    *lo = t - (hn - h);
}

void orc_split_run(const orc_params *p, float *hi, float *lo, float *v, long nsteps,
                   const short *xy_in, const double *tcc, const double *tc,
                   const double *tgamma, int nsrc, const int *si, const int *sj,
                   const double *amp, long amp_len, long step0)
{
    const int NX = p->nx, NY = p->ny;
    const size_t n = (size_t)NX * NY;
    float *h2 = (float *)malloc(sizeof(float) * n), *l2 = (float *)malloc(sizeof(float) * n),
          *v2 = (float *)malloc(sizeof(float) * n);
    double *U = (double *)malloc(sizeof(double) * n), *Y = (double *)malloc(sizeof(double) * n),
           *O = (double *)malloc(sizeof(double) * n);
    const float kf = (float)p->kappa;
    /* the coefficient pair holds tcc times the stencil weight (1/6 for the
     * 9-point Laplacian, folded in binary64); the sum of differences is not
     * scaled */
    const double sw = p->stencil == 9 ? ORC_W9 : 1.0;
    ROWS(Ur, U, double); ROWS(Yr, Y, double); ROWS(Or, O, double);
    ROWS(m, xy_in, short); ROWS(t1, tcc, double); ROWS(t2, tc, double); ROWS(t3, tgamma, double);

    for (long step = 0; step < nsteps; step++) {
        const long g = step0 + step;
        memcpy(h2, hi, sizeof(float) * n);
        memcpy(l2, lo, sizeof(float) * n);
        memcpy(v2, v, sizeof(float) * n);
        /* interior cells in FP32 double-float */
#pragma omp parallel for
        for (int i = 1; i < NX - 1; i++)
            for (int j = 1; j < NY - 1; j++) {
                size_t k = (size_t)i * NY + j;
                if (!ACTIVE_BULK(p, xy_in[k])) continue;
                const float hc = hi[k], lc = lo[k], vc = v[k];
                /* every neighbour difference is a difference of u = hi + lo,
                 * (hi[a] - hi[b]) + (lo[a] - lo[b]) with a the east / north /
                 * diagonal neighbour and b the cell, evaluated once per pair of
                 * cells: the west and south ones enter as minus the difference
                 * their neighbour takes */
#define DIF(a, b) ((hi[a] - hi[b]) + (lo[a] - lo[b]))
                const float d = (DIF(k + NY, k) - DIF(k, k - NY)) + (DIF(k + 1, k) - DIF(k, k - 1));
                float lap;
                if (p->stencil == 9) {
                    const float dd = (DIF(k + NY + 1, k) - DIF(k, k - NY + 1)) + (DIF(k + NY - 1, k) - DIF(k, k - NY - 1));
                    lap = 4.0f * d + dd;
                } else {
                    lap = d;
                }
#undef DIF
                const double cw = tcc[k] * sw;
                const float ch = (float)cw;
                const float cl = (float)(cw - (double)ch);
                const float gf = (float)tgamma[k];
                /* a term whose coefficient is zero is not evaluated */
                float vn = vc + fmaf(ch, lap, cl * lap);
                if (kf != 0.0f) vn = vn - kf * hc;
                if (gf != 0.0f) vn = vn - gf * vc;
                for (int s = 0; s < nsrc; s++)
                    if (g < amp_len && si[s] == i && sj[s] == j) vn = vn + (float)amp[(size_t)s * amp_len + g];
                dd_add_f(hc, lc, vn, &h2[k], &l2[k]);
                v2[k] = vn;
            }
        /* edge cells in FP64 with the reference formulas */
        for (size_t k = 0; k < n; k++) {
            U[k] = (double)hi[k] + (double)lo[k];
            Y[k] = U[k] - (double)v[k];
        }
        edge_step(p, Ur, Yr, Or, m, t1, t2, t3);
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++) {
                if (i != 0 && i != NX - 1 && j != 0 && j != NY - 1) continue;
                size_t k = (size_t)i * NY + j;
                if (!ACTIVE_EDGE(p, xy_in[k])) continue;
                double out = O[k];
                for (int s = 0; s < nsrc; s++)
                    if (g < amp_len && si[s] == i && sj[s] == j) out += amp[(size_t)s * amp_len + g];
                float h = (float)out;
                h2[k] = h;
                l2[k] = (float)(out - (double)h);
                v2[k] = (float)(out - U[k]);
            }
        memcpy(hi, h2, sizeof(float) * n);
        memcpy(lo, l2, sizeof(float) * n);
        memcpy(v, v2, sizeof(float) * n);
    }
    free(h2); free(l2); free(v2); free(U); free(Y); free(O);
    free(Ur); free(Yr); free(Or); free(m); free(t1); free(t2); free(t3);
}

double orc_split_energy_total(const orc_params *p, const float *hi, const float *lo,
                              const float *v, const short *xy_in, const double *tcc,
                              double e_scale)
{
    size_t n = (size_t)p->nx * p->ny;
    double *U = (double *)malloc(sizeof(double) * n);
    for (size_t k = 0; k < n; k++) U[k] = (double)hi[k] + (double)lo[k];
    /* the split form carries the velocity v itself; gradients use u = hi + lo */
    double *e = (double *)malloc(sizeof(double) * n);
    const int NX = p->nx, NY = p->ny;
    const double es2 = e_scale * e_scale;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            size_t k = (size_t)i * NY + j;
            if (!(xy_in[k] != 0 || p->twospeeds)) { e[k] = 0.0; continue; }
            int ip = i + 1 == NX ? NX - 1 : i + 1, im = i == 0 ? 0 : i - 1;
            int jp = j + 1 == NY ? NY - 1 : j + 1, jm = j == 0 ? 0 : j - 1;
            double c = U[k], vv = (double)v[k];
            double a1 = U[(size_t)ip * NY + j] - c, a2 = c - U[(size_t)im * NY + j];
            double b1 = U[(size_t)i * NY + jp] - c, b2 = c - U[(size_t)i * NY + jm];
            double gx2 = a1 * a1 + a2 * a2, gy2 = b1 * b1 + b2 * b2;
            e[k] = es2 * (vv * vv + 0.5 * tcc[k] * (gx2 + gy2));
        }
    double t = orc_fixed_order_sum(NX, NY, e);
    free(U); free(e);
    return t;
}
