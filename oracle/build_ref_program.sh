#!/bin/bash
# Builds the reference's whole program, as its sources stand, into
# oracle/_ref/bgk32/BGK: default REAL(4) (the reference as built), AMD flang,
# the image's MPICH (/opt/conda: mpif.h, libmpifort, libmpi).  Test
# infrastructure only: tests/golden/make_reference_runs.py runs it at one
# MPI rank to record the reference's own RESIDUAL.RES / RESULTS*.RES for
# small decks, which the oracle is then checked against.
#
# Sources are compiled where they lie (nothing is copied or changed); every
# output goes under oracle/_ref/.  Files are compiled in module-dependency
# order, found by repeated passes over the USE statements; getavr.f90 is
# left out as the reference's own makefile leaves it out (nothing calls it).
set -euo pipefail
REF=${DGBGK_REFERENCE:-/root/reference}/src
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref/bgk32
MPI=${MPI_PREFIX:-/opt/conda}
FC=${FC:-amdflang}
[ -d "$REF" ] || { echo "reference sources not found at $REF; skipping the program"; exit 0; }
[ -f "$MPI/include/mpif.h" ] || { echo "mpif.h not found under $MPI/include; skipping"; exit 0; }
[ -f "$MPI/lib/libmpifort.so" ] || { echo "libmpifort not found under $MPI/lib; skipping"; exit 0; }
command -v "$FC" >/dev/null || { echo "$FC not found; skipping the program"; exit 0; }

FLAGS="-O2 -g -w -I$MPI/include"
rm -rf "$OUT"; mkdir -p "$OUT"
cd "$OUT"
pending=""
for f in "$REF"/*.f90; do
  b=$(basename "$f" .f90)
  [ "$b" = getavr ] && continue
  pending="$pending $b"
done
objs=""
while [ -n "${pending// /}" ]; do
  next=""; built=""
  for b in $pending; do
    ready=1
    for m in $(grep -ioE '^[[:space:]]*use[[:space:]]+[a-z0-9_]+' "$REF/$b.f90" |
               awk '{print tolower($2)}' | sort -u); do
      [ -f "$m.mod" ] || { ready=0; break; }
    done
    if [ $ready = 1 ]; then built="$built $b"; else next="$next $b"; fi
  done
  [ -n "${built// /}" ] || { echo "cannot order:$next"; exit 1; }
  for b in $built; do
    $FC -c $FLAGS "$REF/$b.f90" -o "$b.o" &
  done
  wait
  for b in $built; do [ -f "$b.o" ] || { echo "compile failed: $b"; exit 1; }; objs="$objs $b.o"; done
  pending=$next
done
$FC -o BGK $objs -L"$MPI/lib" -Wl,-rpath,"$MPI/lib" -lmpifort -lmpi
rm -f ./*.o ./*.mod
echo "built $OUT/BGK"
