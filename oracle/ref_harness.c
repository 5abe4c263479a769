/*
 * ref_harness.c -- appended after the reference functions by
 * oracle/build_ref.sh.  Exposes them on flat k = i*NY + j arrays so the
 * tests can call them through ctypes.  FLAT_TWIN selects which reference
 * stepper was compiled in (wave_billiard.c row-pointer or wave_3d.c flat).
 */
static t_neumann_bc ref_no_neumann[1];

int ref_nx(void) { return NX; }
int ref_ny(void) { return NY; }

#ifndef FLAT_TWIN
/* evolve_wave_half, wave_billiard.c:302 */
void ref_evolve_wave_half(double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                          double *tcc, double *tc, double *tgamma)
{
    double *a[NX], *b[NX], *c[NX], *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi_in + (size_t)i * NY; b[i] = psi_in + (size_t)i * NY; c[i] = phi_out + (size_t)i * NY;
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    evolve_wave_half(a, b, c, m, t1, t2, t3, ref_no_neumann);
}

/* evolve_wave, wave_billiard.c:613 */
void ref_evolve_wave(double *phi, double *psi, double *tmp, short *xy_in, double *tcc, double *tc,
                     double *tgamma)
{
    double *a[NX], *b[NX], *c[NX], *t1[NX], *t2[NX], *t3[NX];
    short int *m[NX];
    for (int i = 0; i < NX; i++) {
        a[i] = phi + (size_t)i * NY; b[i] = psi + (size_t)i * NY; c[i] = tmp + (size_t)i * NY;
        m[i] = xy_in + (size_t)i * NY;
        t1[i] = tcc + (size_t)i * NY; t2[i] = tc + (size_t)i * NY; t3[i] = tgamma + (size_t)i * NY;
    }
    evolve_wave(a, b, c, m, t1, t2, t3, ref_no_neumann);
}

/* compute_energy, wave_common.c:453 */
double ref_compute_energy(double *phi, double *psi, short *xy_in, int i, int j)
{
    double *a[NX], *b[NX];
    short int *m[NX];
    for (int r = 0; r < NX; r++) {
        a[r] = phi + (size_t)r * NY; b[r] = psi + (size_t)r * NY; m[r] = xy_in + (size_t)r * NY;
    }
    return compute_energy(a, b, m, i, j);
}
#else
/* evolve_wave_half, wave_3d.c:600 (table order tc, tcc, tgamma) */
void ref_evolve_wave_half_flat(double *phi_in, double *psi_in, double *phi_out, short *xy_in,
                               double *tc, double *tcc, double *tgamma)
{
    evolve_wave_half(phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma, ref_no_neumann);
}
#endif
