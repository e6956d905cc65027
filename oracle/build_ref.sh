#!/bin/bash
# Builds oracle/_ref/libref64.so and libref32.so from the reference's own
# MPI-free hot-path routines, compiled where they lie under the reference
# tree, plus ref_wrap.f90 (this project's C entry points onto them).
#
# Test infrastructure only.  The routines that call MPI (ADVNCE, FLUCON2,
# COMPMU, MACROS, EDGFLXOPT, GETRES, GETMAC, INICON2) need an external MPI
# library and are not built.  getbou/getnor/getfor INCLUDE 'mpif.h' without
# calling MPI; the image's installed header satisfies that include.
set -euo pipefail
REF=${DGBGK_REFERENCE:-/root/reference}/src
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
MPIINC=${MPIF_INCLUDE:-/opt/conda/include}
FC=${FC:-amdflang}
[ -d "$REF" ] || { echo "reference sources not found at $REF; skipping _ref"; exit 0; }
[ -f "$MPIINC/mpif.h" ] || { echo "mpif.h not found under $MPIINC; skipping _ref"; exit 0; }
command -v "$FC" >/dev/null || { echo "$FC not found; skipping _ref"; exit 0; }
SRCS="rfillv rfillm rfilla getsid normal getgeo getlen getmat conmat getnor vspace gteqnf
      getflu getfla getelc getbou getinc cmminc adthem inflow alotim alotim2 getfor"
mkdir -p "$OUT"
for prec in 64 32; do
  flags="-O2 -fPIC -I$MPIINC"
  [ "$prec" = 64 ] && flags="$flags -fdefault-real-8 -fdefault-double-8"
  bdir=$OUT/build$prec
  rm -rf "$bdir"; mkdir -p "$bdir"
  objs=""
  for f in $SRCS; do
    (cd "$bdir" && $FC -c $flags "$REF/$f.f90" -o "$f.o")
    objs="$objs $bdir/$f.o"
  done
  (cd "$bdir" && $FC -c $flags "$HERE/ref_wrap.f90" -o ref_wrap.o)
  $FC -shared -o "$OUT/libref$prec.so" $objs "$bdir/ref_wrap.o"
  rm -rf "$bdir"
done
echo "built $OUT/libref64.so $OUT/libref32.so"
