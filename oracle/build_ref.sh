#!/usr/bin/env bash
# Build the reference SayakaAmr solve path from its own sources (read in place under
# the reference checkout) together with oracle/ref_harness.cpp into
#   oracle/_ref/libsayaka_ref2.so   (SAYAKA_SPACEDIM=2)
#   oracle/_ref/libsayaka_ref3.so   (SAYAKA_SPACEDIM=3)
# TEST INFRASTRUCTURE ONLY: the libraries pin the oracle and give the CPU baseline.
#
# Compile flags, and why each one is needed on Linux/g++ (the reference targets MSVC):
#   -D'__declspec(x)='      MSVC attribute on two deprecated constructors
#                           (SayakaFillPatch.h:75, SayakaBoundaryPatch.h:202)
#   -include cstring/tuple/functional
#                           standard headers MSVC pulls in transitively
#   -D_USE_MATH_DEFINES, -Dstricmp=strcasecmp   MSVC spellings
#   -DNDEBUG                the reference's Release configuration
#   2D only, SayakaBoundaryPatch.cpp: -D'static_assert(...)=' removes one
#                           static_assert(NDIM==3) inside a z-face BC routine that the
#                           solve never calls (SayakaBoundaryPatch.cpp:390)
# SayakaWriterHdf5.cpp (needs HDF5), the writers, config and the demo drivers are not
# part of the solve path and are not compiled.
set -euo pipefail
REF=${SAYAKA_REF_SRC:-/root/reference/src}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
    echo "build_ref.sh: reference sources not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
SRCS="SayakaBox SayakaChildIndex SayakaFaceIndex SayakaSurroundingIndex SayakaTree
      SayakaRegrid SayakaTreeData SayakaBlockData SayakaBoundaryPatch SayakaFillPatch
      SayakaInterpPatch SayakaTower SayakaMacSolver SayakaMacOperator SayakaMacUtil
      SayakaMacCoef SayakaMacIB log"
CXXFLAGS="-std=c++14 -O2 -fPIC -DNDEBUG -D__declspec(x)= -include cstring -include tuple
          -include functional -D_USE_MATH_DEFINES -Dstricmp=strcasecmp -w -I$REF"
JOBS=()
for D in 2 3; do
    for f in $SRCS; do
        EXTRA=""
        if [ "$D" = 2 ] && [ "$f" = SayakaBoundaryPatch ]; then EXTRA="-Dstatic_assert(...)="; fi
        o="$OUT/obj/${f}_$D.o"
        if [ ! -f "$o" ] || [ "$REF/$f.cpp" -nt "$o" ]; then
            g++ $CXXFLAGS -DSAYAKA_SPACEDIM=$D $EXTRA -c "$REF/$f.cpp" -o "$o" &
            JOBS+=($!)
        fi
    done
done
for j in "${JOBS[@]:-}"; do [ -n "$j" ] && wait "$j"; done
for D in 2 3; do
    g++ $CXXFLAGS -DSAYAKA_SPACEDIM=$D -c "$HERE/ref_harness.cpp" -o "$OUT/obj/ref_harness_$D.o"
    OBJS=""
    for f in $SRCS; do OBJS="$OBJS $OUT/obj/${f}_$D.o"; done
    g++ -shared -Wl,-Bsymbolic -o "$OUT/libsayaka_ref$D.so.tmp" $OBJS "$OUT/obj/ref_harness_$D.o"
    mv "$OUT/libsayaka_ref$D.so.tmp" "$OUT/libsayaka_ref$D.so"
done
# The projection driver (written against the reference's call surface; the product build
# compiles the same file against include/sayaka) linked with the reference's 3D objects.
DRV="$HERE/../sayakaamr_amd/csrc/drivers/projection_step.cpp"
OBJS3=""
for f in $SRCS; do OBJS3="$OBJS3 $OUT/obj/${f}_3.o"; done
g++ $CXXFLAGS -DSAYAKA_SPACEDIM=3 "$DRV" $OBJS3 -o "$OUT/ref_projection_step.tmp"
mv "$OUT/ref_projection_step.tmp" "$OUT/ref_projection_step"
echo "build_ref.sh: built $OUT/libsayaka_ref2.so $OUT/libsayaka_ref3.so $OUT/ref_projection_step"
