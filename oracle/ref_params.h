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
#ifndef OSCIL_YMID
#define OSCIL_YMID 0.0
#endif
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
#ifndef INITIAL_VARIANCE
#define INITIAL_VARIANCE 0.0003
#endif
#ifndef INITIAL_AMP
#define INITIAL_AMP 1.0
#endif
#ifndef INITIAL_WAVELENGTH
#define INITIAL_WAVELENGTH 0.01
#endif
#ifndef BC_NEUMANN
#define BC_NEUMANN 0
#endif
#ifndef COMPARISON
#define COMPARISON 0
#endif
#ifndef FLOOR
#define FLOOR 0
#endif
#ifndef VMAX
#define VMAX 2.5
#endif
#ifndef NSTEPS
#define NSTEPS 10
#endif
#define BEAM_BC ((OSCILLATION_SCHEDULE == OSC_BEAM_SINE)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_DECREASING)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_TWOPERIODS)||(OSCILLATION_SCHEDULE == OSC_BEAM_SINE_CHIRP))

/* colour macros of the drawing path (wave_billiard.c:200-241), for the
 * colour builds */
#ifndef COLOR_SCHEME
#define COLOR_SCHEME 3
#endif
#ifndef COLOR_PALETTE
#define COLOR_PALETTE 0
#endif
#ifndef SLOPE
#define SLOPE 1.0
#endif
#ifndef COLOR_RANGE
#define COLOR_RANGE 0.8
#endif
#ifndef ATTENUATION
#define ATTENUATION 0.0
#endif
#ifndef VSHIFT_AMPLITUDE
#define VSHIFT_AMPLITUDE -0.2
#endif
#ifndef VSCALE_AMPLITUDE
#define VSCALE_AMPLITUDE 1.0
#endif
#ifndef LOG_SCALE
#define LOG_SCALE 0.75
#endif
#ifndef LOG_SHIFT
#define LOG_SHIFT 0.75
#endif
#ifndef FLUX_SCALE
#define FLUX_SCALE 250.0
#endif
#ifndef AVRG_E_FACTOR
#define AVRG_E_FACTOR 0.95
#endif
#ifndef COLORHUE
#define COLORHUE 260
#endif
#ifndef COLORDRIFT
#define COLORDRIFT 0.0
#endif
#ifndef LUMMEAN
#define LUMMEAN 0.5
#endif
#ifndef LUMAMP
#define LUMAMP 0.3
#endif
#ifndef HUEMEAN
#define HUEMEAN 180.0
#endif
#ifndef HUEAMP
#define HUEAMP -180.0
#endif
#ifndef RESCALE_COLOR_IN_CENTER
#define RESCALE_COLOR_IN_CENTER 0
#endif

/* media models of init_ior_2d (sub_wave.c:12164-13004), for the media builds;
 * the macros its other branches and init_circle_config_pattern read are set
 * to the checked-in wave_billiard.c values or to any valid value (those
 * branches do not run) */
#ifndef VARIABLE_IOR
#define VARIABLE_IOR 1
#endif
#ifndef IOR
#define IOR 4
#endif
#ifndef NGRIDX
#define NGRIDX 10
#endif
#ifndef NGRIDY
#define NGRIDY 10
#endif
#ifndef GAMMA
#define GAMMA 1.0e-4
#endif
#ifndef GAMMAB
#define GAMMAB 2.0e-4
#endif
#ifndef MANDELLEVEL
#define MANDELLEVEL 100
#endif
#ifndef MANDELLIMIT
#define MANDELLIMIT 10.0
#endif
#ifndef MANDEL_IOR_SCALE
#define MANDEL_IOR_SCALE -0.05
#endif
#define NMAXCIRCLES 10000
#define HRES 1
#define INITIAL_TIME 0
#ifndef WALL_WIDTH
#define WALL_WIDTH 0.1
#endif
#ifndef MU
#define MU 0.1
#endif
/* domain parameters xy_in_billiard_single_domain reads (sub_wave.c:5712),
 * at the values of the checked-in wave_billiard.c:73-109 */
#ifndef B_DOMAIN
#define B_DOMAIN 20
#endif
#ifndef B_DOMAIN_B
#define B_DOMAIN_B 20
#endif
#ifndef MU_B
#define MU_B 0.08
#endif
#ifndef JULIA_SCALE
#define JULIA_SCALE 0.8
#endif
#ifndef MRATIO
#define MRATIO 3
#endif
#ifndef WALL_WIDTH_B
#define WALL_WIDTH_B 0.1
#endif
#ifndef WALL_WIDTH_ASYM
#define WALL_WIDTH_ASYM 0.75
#endif
#ifndef WALL_WIDTH_ASYM_B
#define WALL_WIDTH_ASYM_B 0.75
#endif
#define PDISC_FACTOR 3.0
#define P_PERCOL 0.25
#define NPOISSON 300
#define X_SHOOTER -0.2
#define X_TARGET 0.4
#define Y_SHOOTER -0.6
#define Y_TARGET 0.0
#define RADIUS_FACTOR 0.3
#ifndef NPOLY
#define NPOLY 6
#endif
#ifndef APOLY
#define APOLY 0.0
#endif
#ifndef MDEPTH
#define MDEPTH 4
#endif

