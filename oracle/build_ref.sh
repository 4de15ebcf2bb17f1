#!/bin/bash
# TEST INFRASTRUCTURE ONLY. Compiles the reference's own sources where they lie under
# /root/reference (lib/, planet/, tinyxml2/ -- the files its Makefile links into libaturan.a, with
# the .inc files it already carries -- plus cli/DefaultStream.cpp, the output stream its
# command-line build links) together with ref_harness.cpp into oracle/_ref/ref_harness.
# Nothing is copied out of the reference and its own build system is not run. Outputs go only to
# oracle/_ref/, which git ignores.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
mkdir -p "$OUT/obj"
CXX=${CXX:-g++}
# -O2 without fast-math on baseline x86-64 (no FMA): every double operation rounds as at -O0.
FLAGS="-O2 -fno-fast-math -ffp-contract=off -std=c++11 -fopenmp -w -I$REF/lib -I$REF/planet -I$REF/tinyxml2"
SRCS="lib/Array.cpp lib/Array_2D.cpp lib/Array_1D.cpp lib/Config.cpp lib/Utils.cpp lib/FFT.cpp
planet/cUranusModel.cpp planet/PrintMsg_Uran.cpp planet/RungeKutta_Uran_Turb.cpp planet/RHS_Uran_Turb.cpp
planet/ParaView_Uran.cpp planet/Thermo_Uran.cpp planet/InitVariables_Uran.cpp planet/InitVelocity_Uran.cpp
planet/FileIO_Uran.cpp planet/Pressure_Uran.cpp planet/SaturationAdjustmentUran.cpp planet/Delegations_Uran.cpp
planet/BC_Uran.cpp tinyxml2/tinyxml2.cpp cli/DefaultStream.cpp"
pids=()
objs=()
for s in $SRCS; do
    o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
    objs+=("$o")
    if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
        $CXX $FLAGS -c "$REF/$s" -o "$o" &
        pids+=($!)
    fi
done
for p in "${pids[@]}"; do wait "$p"; done
$CXX $FLAGS -c "$HERE/ref_harness.cpp" -o "$OUT/obj/ref_harness.o"
$CXX $FLAGS -o "$OUT/ref_harness" "$OUT/obj/ref_harness.o" "${objs[@]}" -lm
echo "built $OUT/ref_harness"
