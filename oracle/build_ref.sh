#!/usr/bin/env bash
# Compile the reference's own stepper into oracle/_ref/ for the golden-fixture
# generator and the oracle cross-check.  The translation unit is assembled on
# stdin from the reference files where they lie (nothing is copied into the
# repository): the parameter block oracle/ref_params.h, global_pdes.c whole,
# then xy_to_ij (sub_wave.c:412-422), oscillating_bc (sub_wave.c:11965-12162),
# compute_energy (wave_common.c:453-472) and the stepper (wave_billiard.c:302-631
# or wave_3d.c:600-1007; for the neighbour-table stepper sub_wave.c:11555-11963
# and wave_3d.c:345-597; for the frame colours sub_wave.c:270-304, colors_waves.c
# with the colour tables it includes, wave_common.c:453-512 and 1053-1171; for
# the media tables sub_wave.c:270-304, 332-357, 469-476, 721-1387,
# 12164-13004; for the masks sub_wave.c:270-304, 321-330, 332-357, 362-411,
# 469-476, 721-1387, 5299-5486, 5682-7602; for the initial drop and pulses the
# same plus wave_common.c:89-111 and 337-351), then oracle/ref_harness.c.  Flags follow makefile:1-4
# (gcc -O3 -fopenmp) plus -ffp-contract=off.  Needs the reference sources at
# /root/reference; machines without them use the oracle alone.
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
[ -f "$REF/wave_billiard.c" ] || { echo "reference not found at $REF" >&2; exit 1; }
mkdir -p "$OUT"

build_one() {  # build_one <tag> <0 row pointers | 1 flat | 2 neighbour table | 3 colours | 4 media | 5 masks | 6 pulses> <gcc -D flags...>
    local tag=$1 flat=$2; shift 2
    local so="$OUT/libref_${tag}.so"
    {
        cat "$HERE/ref_params.h"
        echo "#include \"$REF/global_pdes.c\""
        sed -n '412,422p' "$REF/sub_wave.c"
        sed -n '11965,12162p' "$REF/sub_wave.c"
        if [ "$flat" = 6 ]; then
            echo "#define PULSE_TWIN 1"
            for r in 270,304 321,330 332,357 362,411 469,476 721,1387 5299,5486 5682,7602; do sed -n "${r}p" "$REF/sub_wave.c"; done
            sed -n '89,111p' "$REF/wave_common.c"
            sed -n '337,351p' "$REF/wave_common.c"
        elif [ "$flat" = 5 ]; then
            echo "#define MASK_TWIN 1"
            for r in 270,304 321,330 332,357 362,411 469,476 721,1387 5299,5486 5682,7602; do sed -n "${r}p" "$REF/sub_wave.c"; done
        elif [ "$flat" = 4 ]; then
            echo "#define IOR_TWIN 1"
            for r in 270,304 332,357 469,476 721,1387 12164,13004; do sed -n "${r}p" "$REF/sub_wave.c"; done
        elif [ "$flat" = 3 ]; then
            echo "#define COLOR_TWIN 1"
            sed -n '270,304p' "$REF/sub_wave.c"
            echo "#include \"$REF/colors_waves.c\""
            sed -n '453,512p' "$REF/wave_common.c"
            sed -n '1053,1171p' "$REF/wave_common.c"
        elif [ "$flat" = 2 ]; then
            echo "#define TABLE_TWIN 1"
            sed -n '11555,11963p' "$REF/sub_wave.c"
            sed -n '345,597p' "$REF/wave_3d.c"
        elif [ "$flat" = 1 ]; then
            echo "#define FLAT_TWIN 1"
            sed -n '600,1007p' "$REF/wave_3d.c"
        else
            sed -n '453,472p' "$REF/wave_common.c"
            sed -n '302,631p' "$REF/wave_billiard.c"
        fi
        cat "$HERE/ref_harness.c"
    } | gcc -x c - -O3 -fopenmp -fPIC -shared -ffp-contract=off -w "$@" -o "$so" -lm
}

KS="-DKAPPA=1.0e-3 -DKAPPA_SIDES=5.0e-4 -DGAMMA_SIDES=1.0e-4 -DKAPPA_TOPBOT=2.0e-4 -DGAMMA_TOPBOT=1.0e-7"
pids=()
for g in "24 17" "16 29"; do
    set -- $g; nx=$1; ny=$2
    for bc in 0 1 2 3; do
        for ts in 0 1; do
            build_one "row_${nx}x${ny}_bc${bc}_ts${ts}" 0 -DNX=$nx -DNY=$ny -DB_COND=$bc -DTWOSPEEDS=$ts $KS &
            pids+=($!)
        done
    done
    for bc in 0 1 2 3 4; do
        for ts in 0 1; do
            build_one "flat_${nx}x${ny}_bc${bc}_ts${ts}" 1 -DNX=$nx -DNY=$ny -DB_COND=$bc -DTWOSPEEDS=$ts $KS &
            pids+=($!)
        done
    done
    build_one "row_${nx}x${ny}_bc0_ts0_tb" 0 -DNX=$nx -DNY=$ny -DB_COND=0 -DTWOSPEEDS=0 -DOSCILLATE_TOPBOT=1 $KS &
    pids+=($!)
    build_one "flat_${nx}x${ny}_bc2_ts1_tb" 1 -DNX=$nx -DNY=$ny -DB_COND=2 -DTWOSPEEDS=1 -DOSCILLATE_TOPBOT=1 $KS &
    pids+=($!)
    # wait in batches of 8 to keep the container within its 8 CPUs
    for p in "${pids[@]}"; do wait "$p"; done; pids=()
done
# the CPU-runnable plain workload: 1280x720, 5-point, B_COND 0, courant 0.1
build_one "row_1280x720_bc0_ts0" 0 -DNX=1280 -DNY=720 -DB_COND=0 -DTWOSPEEDS=0
# the features of the checked-in wave_billiard.c the plain builds leave off:
# Neumann ghost points, the oscillating left edge (periodic schedule) and the
# FLOOR clamp; 16x29 so that the corner write at phi_out[0][NX-1]
# (wave_billiard.c:589) stays inside the row
for bc in 0 1 2; do
    build_one "row_16x29_bc${bc}_ts0_ext" 0 -DNX=16 -DNY=29 -DB_COND=$bc -DTWOSPEEDS=0 -DBC_NEUMANN=1 \
        -DOSCILLATE_LEFT=1 -DOSCILLATION_SCHEDULE=0 -DFLOOR=1 -DVMAX=0.9 $KS &
    pids+=($!)
done
build_one "row_16x29_bc2_ts1_ext" 0 -DNX=16 -DNY=29 -DB_COND=2 -DTWOSPEEDS=1 -DBC_NEUMANN=1 \
    -DOSCILLATE_LEFT=1 -DOSCILLATION_SCHEDULE=0 -DFLOOR=1 -DVMAX=0.9 $KS &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# COMPARISON: the flat twin split into two halves (wave_3d.c:736-990)
for g in "24 17" "16 29"; do
    set -- $g; nx=$1; ny=$2
    for bc in 0 1 2 3 4; do
        build_one "flat_${nx}x${ny}_bc${bc}_cmp" 1 -DNX=$nx -DNY=$ny -DB_COND=$bc -DTWOSPEEDS=$(( bc % 2 )) \
            -DCOMPARISON=1 $KS &
        pids+=($!)
    done
    for p in "${pids[@]}"; do wait "$p"; done; pids=()
done
# the neighbour-table stepper (PRECOMPUTE_BC): evolve_wave_half_new
# (wave_3d.c:345-597) with init_laplacian_coords / compute_laplacian
# (sub_wave.c:11555-11963), B_COND 0-4 and BC_LSHAPE (10)
for g in "24 17" "16 29"; do
    set -- $g; nx=$1; ny=$2
    for bc in 0 1 2 3 4 10; do
        build_one "tbl_${nx}x${ny}_bc${bc}_ts0" 2 -DNX=$nx -DNY=$ny -DB_COND=$bc -DTWOSPEEDS=0 $KS &
        pids+=($!)
    done
    for p in "${pids[@]}"; do wait "$p"; done; pids=()
done
build_one "tbl_24x17_bc1_ts1" 2 -DNX=24 -DNY=17 -DB_COND=1 -DTWOSPEEDS=1 $KS &
pids+=($!)
build_one "tbl_24x17_bc0_ts0_tb" 2 -DNX=24 -DNY=17 -DB_COND=0 -DTWOSPEEDS=0 -DOSCILLATE_TOPBOT=1 $KS &
pids+=($!)
build_one "tbl_16x29_bc3_ts0_tb" 2 -DNX=16 -DNY=29 -DB_COND=3 -DTWOSPEEDS=0 -DOSCILLATE_TOPBOT=1 $KS &
pids+=($!)
build_one "tbl_16x29_bc2_ts0_ext" 2 -DNX=16 -DNY=29 -DB_COND=2 -DTWOSPEEDS=0 -DOSCILLATE_LEFT=1 \
    -DOSCILLATION_SCHEDULE=0 -DFLOOR=1 -DVMAX=0.9 $KS &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# the colour path of the drawing functions: wave_value (wave_common.c:1053)
# through color_scheme_palette / color_scheme_asym_palette and the palettes
# of colormaps.c, one build per COLOR_SCHEME (COLOR_PALETTE >= COL_TURBO
# switches the energy plots to the asymmetric scheme)
COL="-DNX=24 -DNY=17 -DE_SCALE=50.0 -DCOURANT=0.1 -DCOURANTB=0.05 -DNSTEPS=500 -DCOLORDRIFT=30.0"
for sc in 0 1 3 4 5; do
    build_one "col_s${sc}" 3 $COL -DCOLOR_SCHEME=$sc -DTWOSPEEDS=0 &
    pids+=($!)
done
build_one "col_s1_asym" 3 $COL -DCOLOR_SCHEME=1 -DCOLOR_PALETTE=11 -DTWOSPEEDS=1 &
pids+=($!)
build_one "col_s3_asym" 3 $COL -DCOLOR_SCHEME=3 -DCOLOR_PALETTE=11 -DTWOSPEEDS=0 &
pids+=($!)
build_one "col_s3_att" 3 $COL -DCOLOR_SCHEME=3 -DCOLOR_RANGE=1.0 -DATTENUATION=0.01 \
    -DRESCALE_COLOR_IN_CENTER=1 -DTWOSPEEDS=1 &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# the Gaussian-well media of init_ior_2d (sub_wave.c:12315-12389)
build_one "ior_wells_p" 4 -DNX=40 -DNY=30 -DIOR=4 -DNGRIDX=6 -DNGRIDY=5 -DTWOSPEEDS=0 -DCOURANTB=0.05 &
pids+=($!)
build_one "ior_wells_r" 4 -DNX=40 -DNY=30 -DIOR=5 -DNGRIDX=6 -DNGRIDY=5 -DTWOSPEEDS=1 -DCOURANTB=0.05 &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# the per-cell media models of init_ior_2d (sub_wave.c:12181-12968), one
# build per IOR code (VARIABLE_IOR 0: the constant branch 12984-13003), and
# the rotating wells (12390-12481) for the REFRESH_IOR rebuild
IORP="-DNX=97 -DNY=61 -DLAMBDA=0.8 -DMU=0.3 -DCOURANT=0.1 -DCOURANTB=0.05 -DGAMMA=1.0e-4 -DGAMMAB=2.0e-4"
for m in 1 100 2 10 11 12 13 14 15 151 16 17 18 181 182 183 19 191; do
    build_one "ior_m${m}" 4 $IORP -DIOR=$m -DTWOSPEEDS=$(( m % 2 )) -DMDEPTH=6 &
    pids+=($!)
    [ ${#pids[@]} -ge 8 ] && { for p in "${pids[@]}"; do wait "$p"; done; pids=(); }
done
build_one "ior_const_ts0" 4 $IORP -DVARIABLE_IOR=0 -DTWOSPEEDS=0 &
pids+=($!)
build_one "ior_const_ts1" 4 $IORP -DVARIABLE_IOR=0 -DTWOSPEEDS=1 &
pids+=($!)
build_one "ior_rot" 4 -DNX=40 -DNY=30 -DIOR=6 -DNGRIDX=6 -DNGRIDY=5 -DTWOSPEEDS=0 -DCOURANTB=0.05 &
pids+=($!)
build_one "ior_rot_large" 4 -DNX=40 -DNY=30 -DIOR=7 -DNGRIDX=6 -DNGRIDY=5 -DTWOSPEEDS=1 -DCOURANTB=0.05 &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# every closed-form domain of xy_in_billiard_single_domain, two parameter sets
build_one "dom_a" 5 -DNX=97 -DNY=61 -DLAMBDA=0.6 -DMU=0.3 -DNPOLY=6 -DAPOLY=0.5 &
pids+=($!)
build_one "dom_b" 5 -DNX=160 -DNY=90 -DLAMBDA=1.2 -DMU=0.15 -DNPOLY=9 -DAPOLY=1.0 -DMDEPTH=3 -DJULIA_SCALE=1.0 \
    -DMANDELLEVEL=200 -DWALL_WIDTH=0.075 -DMU_B=0.3 -DWALL_WIDTH_B=0.15 -DNGRIDX=8 -DNGRIDY=6 \
    -DWALL_WIDTH_ASYM=1.25 -DWALL_WIDTH_ASYM_B=0.5 -DOSCIL_YMID=-0.3 &
pids+=($!)
build_one "dom_c" 5 -DNX=121 -DNY=77 -DLAMBDA=0.45 -DMU=0.05 -DNPOLY=12 -DAPOLY=0.25 -DMDEPTH=5 -DMRATIO=5 \
    -DJULIA_SCALE=1.3 -DMANDELLEVEL=500 -DMANDELLIMIT=4.0 -DWALL_WIDTH=0.04 -DMU_B=0.15 -DWALL_WIDTH_B=0.05 \
    -DNGRIDX=12 -DNGRIDY=8 -DWALL_WIDTH_ASYM=0.6 -DWALL_WIDTH_ASYM_B=1.4 -DOSCIL_YMID=0.45 &
pids+=($!)
# xy_in_billiard with COMPARISON (two domains, two circle lists)
build_one "dom_cmp_a" 5 -DNX=97 -DNY=61 -DLAMBDA=0.6 -DMU=0.3 -DNPOLY=6 -DAPOLY=0.5 -DCOMPARISON=1 -DB_DOMAIN=20 \
    -DB_DOMAIN_B=20 &
pids+=($!)
build_one "dom_cmp_b" 5 -DNX=97 -DNY=61 -DLAMBDA=0.6 -DMU=0.3 -DNPOLY=6 -DAPOLY=0.5 -DCOMPARISON=1 -DB_DOMAIN=73 \
    -DB_DOMAIN_B=8 &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# the billiard masks of init_wave_flat (xy_in_billiard_single_domain): the
# Sinai and exterior-ellipse domains of the plain workloads and D_CIRCLES
build_one "mask_97x61_l03" 5 -DNX=97 -DNY=61 -DLAMBDA=0.3 -DMU=0.1 &
pids+=($!)
build_one "mask_97x61_l06_m03" 5 -DNX=97 -DNY=61 -DLAMBDA=0.6 -DMU=0.3 &
pids+=($!)
build_one "mask_1280x720_l03" 5 -DNX=1280 -DNY=720 -DLAMBDA=0.3 -DMU=0.1 &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done; pids=()
# the precision workload's grid (4096 x 4096, absorbing edges, courant 0.1, no
# damping): the reference stepper timed as the CPU baseline of the benchmark
build_one "row_4096x4096_bc2_ts0" 0 -DNX=4096 -DNY=4096 -DB_COND=2 -DTWOSPEEDS=0 &
pids+=($!)
# the initial drop and the added pulses of init_circular_wave / add_circular_wave
# (wave_common.c:89-111, 337-351) on a Sinai billiard, one speed and two
build_one "pulse_97x61_sinai_ts0" 6 -DNX=97 -DNY=61 -DB_DOMAIN=3 -DLAMBDA=0.3 -DINITIAL_VARIANCE=0.05 \
    -DINITIAL_WAVELENGTH=0.1 -DINITIAL_AMP=0.75 &
pids+=($!)
build_one "pulse_97x61_sinai_ts1" 6 -DNX=97 -DNY=61 -DB_DOMAIN=3 -DLAMBDA=0.3 -DINITIAL_VARIANCE=0.05 \
    -DINITIAL_WAVELENGTH=0.1 -DINITIAL_AMP=0.75 -DTWOSPEEDS=1 &
pids+=($!)
# a lens medium with damping and absorbing edges (energy fixture, two speeds)
build_one "row_64x48_bc2_ts1" 0 -DNX=64 -DNY=48 -DB_COND=2 -DTWOSPEEDS=1 -DCOURANTB=0.0666666666666666667 $KS
for p in "${pids[@]}"; do wait "$p"; done; pids=()
ls "$OUT" | wc -l
