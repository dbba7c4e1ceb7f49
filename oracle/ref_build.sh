#!/bin/bash
# Builds oracle/_ref/libref_islands.so from the reference's own Boost-free sources,
# where they lie under /root/reference (ALGLIB 3.15.0 interpolation and its
# dependencies, libtricubic 1.0, src/field.cpp and src/analyticFields.cpp with
# the vendored exprtk), plus oracle/ref_wrap.cpp and oracle/ref_wrap_fields.cpp.
# Test infrastructure only; outputs go to oracle/_ref/ (git-ignored).
set -e
REF=${PT_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
[ -d "$REF/alglib-3.15.0/cpp/src" ] || { echo "ref_build: no reference sources, skipped"; exit 0; }
ALG=$REF/alglib-3.15.0/cpp/src
TRI=$REF/tricubic-1.0/src/libtricubic
OUT=$HERE/_ref
mkdir -p "$OUT/obj"
FLAGS="-O1 -fPIC -std=c++11 -ffp-contract=off -w -I$ALG -I$TRI"
pids=()
for f in ap alglibinternal alglibmisc linalg solvers optimization specialfunctions integration interpolation; do
    if [ ! -f "$OUT/obj/$f.o" ] || [ "$ALG/$f.cpp" -nt "$OUT/obj/$f.o" ]; then
        g++ $FLAGS -c "$ALG/$f.cpp" -o "$OUT/obj/$f.o" & pids+=($!)
    fi
done
g++ $FLAGS -c "$TRI/libtricubic.cpp" -o "$OUT/obj/libtricubic.o" & pids+=($!)
g++ $FLAGS -c "$TRI/tricubic_utils.cpp" -o "$OUT/obj/tricubic_utils.o" & pids+=($!)
g++ $FLAGS -c "$HERE/ref_wrap.cpp" -o "$OUT/obj/ref_wrap.o" & pids+=($!)
# field classes: the reference's include/ and exprtk/ on the include path only here
FFLAGS="-O1 -fPIC -std=c++11 -ffp-contract=off -w -I$REF/include -I$REF/exprtk"
for f in field analyticFields; do
    if [ ! -f "$OUT/obj/$f.o" ] || [ "$REF/src/$f.cpp" -nt "$OUT/obj/$f.o" ]; then
        g++ $FFLAGS -c "$REF/src/$f.cpp" -o "$OUT/obj/$f.o" & pids+=($!)
    fi
done
g++ $FFLAGS -c "$HERE/ref_wrap_fields.cpp" -o "$OUT/obj/ref_wrap_fields.o" & pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done
g++ -shared -o "$OUT/libref_islands.so" "$OUT"/obj/*.o
echo "ref_build: $OUT/libref_islands.so"
