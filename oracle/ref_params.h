/*
 * ref_params.h -- parameter block for compiling the reference stepper
 * (oracle/build_ref.sh).  The reference is configured by editing the
 * #define block at the top of each program (wave_billiard.c:46-285,
 * wave_3d.c:46-290); this file is that block for the fixture builds.  Each
 * value can be overridden with -D on the command line.  Only the macros
 * read by the compiled functions are listed.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef NX
#define NX 24
#endif
#ifndef NY
#define NY 17
#endif
#ifndef B_COND
#define B_COND 0
#endif
#ifndef TWOSPEEDS
#define TWOSPEEDS 0
#endif
#ifndef OSCILLATE_TOPBOT
#define OSCILLATE_TOPBOT 0
#endif
#ifndef KAPPA
#define KAPPA 0.0
#endif
#ifndef KAPPA_SIDES
#define KAPPA_SIDES 0.0
#endif
#ifndef GAMMA_SIDES
#define GAMMA_SIDES 0.0
#endif
#ifndef KAPPA_TOPBOT
#define KAPPA_TOPBOT 0.0
#endif
#ifndef GAMMA_TOPBOT
#define GAMMA_TOPBOT 0.0
#endif
#ifndef COURANT
#define COURANT 0.1
#endif
#ifndef COURANTB
#define COURANTB 0.05
#endif
#ifndef E_SCALE
#define E_SCALE 300.0
#endif
#define XMIN -2.0
#define XMAX 2.0
#define YMIN -1.125
#define YMAX 1.125
#ifndef OSCILLATE_LEFT
#define OSCILLATE_LEFT 0
#endif
#ifndef OSCILLATION_SCHEDULE
#define OSCILLATION_SCHEDULE 0
#endif
#define OSCIL_YMAX 1.2
#define OSCIL_YMID 0.0
#define OSCIL_LEFT_YSHIFT 0.0
#define INITIAL_SHIFT 20.0
#define WAVE_PACKET_SHIFT 200.0
#ifndef OMEGA
#define OMEGA 0.005
#endif
#define OMEGA_B 0.01
#define AMPLITUDE 1.0
#ifndef LAMBDA
#define LAMBDA 1.5      /* D_LSHAPE / BC_LSHAPE half-width, sub_wave.c:6417, 11780 */
#endif
#define ACHIRP 0.25
#define DAMPING 0.0
#define INITIAL_VARIANCE 0.0003
#ifndef BC_NEUMANN
#define BC_NEUMANN 0
#endif
#define COMPARISON 0
#ifndef FLOOR
#define FLOOR 0
#endif
#ifndef VMAX
#define VMAX 2.5
#endif
#define NSTEPS 10
#define BEAM_BC ((OSCILLATION_SCHEDULE == OSC_BEAM_SINE)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_DECREASING)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_TWOPERIODS)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_CHIRP))
