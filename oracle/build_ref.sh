#!/bin/sh
# Builds oracle/_ref/libstz_ref.so: the SZ3-free functions of the reference's
# canonical variant, compiled from the reference source where it lies.
#
#   F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
#   F5:1-17      its standard-library #includes
#   F5:24-1253   globals, split/merge, predictors, pre/de-preprocess, quantizer
#   F5:1577-1839 ROI / slice planners
#
# oracle/_ref/libstz_ref_stz2.so: preprocess_block of src/stz2.cpp (its lines
# 1-8 and 82-306, no SZ3 symbol) with ref2_harness.cpp.
# oracle/_ref/libstz_ref_dep.so: the staged predictors of src/dependency.cpp
# (its lines 1-9, 11-29 and 90-432: #includes, globals, interp_axis ..
# depreprocess_block; line 10 is the SZ3 #include) with ref3_harness.cpp.
#
# F5:18-22 (#include of the SZ3-ncp headers) and every function that calls
# SZ3-ncp (F5:1255-1576, F5:1841-) are left out: SZ3-ncp is not in this image,
# so the entropy-coding half of the reference cannot be built here.  The lines
# are read through a pipe; no reference source is written to disk.
# ref_harness.cpp (appended) only adds extern "C" wrappers that call them.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REFROOT=${STZ_REFERENCE_ROOT:-/root/reference}
F5="$REFROOT/src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp"
if [ ! -f "$F5" ]; then
    echo "reference source not present; oracle/_ref not built" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
# one build per predictor option set of F5 (its own compile-time switches,
# F5:177-203 PRED_WPOW2, F5:579-607 PRED_PLANE_AXIS, F5:609-686 PRED_CUBE_AXIS)
build() {  # $1 = output name, the rest = -D flags
    out="$HERE/_ref/$1"; shift
    {
        echo "#line 1 \"F5\""; sed -n '1,17p' "$F5"
        echo "#line 24 \"F5\""; sed -n '24,1253p' "$F5"
        echo "#line 1577 \"F5\""; sed -n '1577,1839p' "$F5"
        echo "#line 1 \"ref_harness.cpp\""; cat "$HERE/ref_harness.cpp"
    } | g++ -x c++ -std=c++17 -O3 -fopenmp -fPIC -shared -w "$@" -o "$out" -
    echo "built $out"
}
build libstz_ref.so & p0=$!
build libstz_ref_wpow2.so -DPRED_WPOW2 & p1=$!
build libstz_ref_plane.so -DPRED_PLANE_AXIS & p2=$!
build libstz_ref_cube.so -DPRED_CUBE_AXIS & p3=$!
build libstz_ref_all.so -DPRED_WPOW2 -DPRED_PLANE_AXIS -DPRED_CUBE_AXIS & p4=$!
# stz2.cpp's static plane / cube predictors (preprocess_block, stz2.cpp:82-306)
STZ2="$REFROOT/src/stz2.cpp"
{
    echo "#line 1 \"stz2.cpp\""; sed -n '1,8p' "$STZ2"
    echo "#line 82 \"stz2.cpp\""; sed -n '82,306p' "$STZ2"
    echo "#line 1 \"ref2_harness.cpp\""; cat "$HERE/ref2_harness.cpp"
} | g++ -x c++ -std=c++17 -O3 -fopenmp -fPIC -shared -w -o "$HERE/_ref/libstz_ref_stz2.so" - & p5=$!
# dependency.cpp's staged inverse-variance predictors (dependency.cpp:90-432)
DEP="$REFROOT/src/dependency.cpp"
{
    echo "#line 1 \"dependency.cpp\""; sed -n '1,9p' "$DEP"
    echo "#line 11 \"dependency.cpp\""; sed -n '11,29p' "$DEP"
    echo "#line 90 \"dependency.cpp\""; sed -n '90,432p' "$DEP"
    echo "#line 1 \"ref3_harness.cpp\""; cat "$HERE/ref3_harness.cpp"
} | g++ -x c++ -std=c++17 -O3 -fopenmp -fPIC -shared -w -o "$HERE/_ref/libstz_ref_dep.so" - & p6=$!
for p in $p0 $p1 $p2 $p3 $p4 $p5 $p6; do wait $p || exit 1; done
echo "built $HERE/_ref/libstz_ref_stz2.so"
echo "built $HERE/_ref/libstz_ref_dep.so"
