#!/bin/bash
# Build the reference's own CPU code into oracle/_ref/ (test infrastructure
# only; outputs are git-ignored).  Needs the reference tree, so it runs in the
# development container, never where the reference sources are absent (the prebuilt files are used there).
#   libkirkland_ref.so : ImageSimulation/multislice/slicelib.cpp + floatTIFF.cpp
#                        + ref_harness.cpp
#   libfftw2f_ref.so   : ImageSimulation/fftw-2.1.5/fftw/*.c (float mode)
#                        + ref_fftw2_harness.c
# The reference's mulslice.cpp/cfpix.cpp are NOT built: they need the FFTW 3
# float library, which this image lacks.
set -euo pipefail
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MS="$REF/ImageSimulation/multislice"
FW="$REF/ImageSimulation/fftw-2.1.5/fftw"
[ -f "$MS/slicelib.cpp" ] || { echo "reference not found at $REF" >&2; exit 1; }
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
g++ -O2 -fPIC -shared -w -I"$MS" "$MS/slicelib.cpp" "$MS/floatTIFF.cpp" "$HERE/ref_harness.cpp" \
    -o "$OUT/libkirkland_ref.so"
objs=()
for f in "$FW"/*.c; do
  case "$(basename "$f")" in fftwf77.c) continue;; esac
  o="$TMP/$(basename "$f" .c).o"
  gcc -O2 -fPIC -w -DFFTW_ENABLE_FLOAT -I"$FW" -c "$f" -o "$o"
  objs+=("$o")
done
gcc -O2 -fPIC -w -fopenmp -DFFTW_ENABLE_FLOAT -I"$FW" -c "$HERE/ref_fftw2_harness.c" -o "$TMP/h.o"
gcc -shared -fopenmp "${objs[@]}" "$TMP/h.o" -lm -o "$OUT/libfftw2f_ref.so"
echo "built $OUT/libkirkland_ref.so $OUT/libfftw2f_ref.so"
