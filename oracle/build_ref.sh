#!/bin/sh
# Builds oracle/_ref/libref.so from the reference's OWN sources, compiled where
# they lie (read-only) with amdflang: FFTPACK (cft_buff.f, rft_buff.f), the
# global-state module (declaration.f90) and the tridiagonal solvers
# (tridLU_3D.f90; it includes the image's MPICH mpif.h for declarations only
# and calls no MPI routine).  ref_shim.f90 is this repository's own bind(C)
# loader of the module state.  Outputs go to oracle/_ref/ only.
# The rest of the reference (FOU3D.f90, start.f90, littleharsh.f90) needs the
# MPI library at link time and is not built.
# Test infrastructure only.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${LH_REFERENCE:-/root/reference}
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "build_ref: $REF not present, skipping"
  exit 0
fi
FC=${FC:-amdflang}
FFLAGS="-O0 -fPIC"
mkdir -p "$OUT"
$FC $FFLAGS -c "$REF/cft_buff.f" -o "$OUT/cft_buff.o" 2>"$OUT/build.log"
$FC $FFLAGS -c "$REF/rft_buff.f" -o "$OUT/rft_buff.o" 2>>"$OUT/build.log"
$FC $FFLAGS -module-dir "$OUT" -c "$REF/declaration.f90" -o "$OUT/declaration.o" 2>>"$OUT/build.log"
$FC $FFLAGS -module-dir "$OUT" -I/opt/conda/include -c "$REF/tridLU_3D.f90" -o "$OUT/tridLU_3D.o" 2>>"$OUT/build.log"
$FC $FFLAGS -module-dir "$OUT" -I"$OUT" -c "$HERE/ref_shim.f90" -o "$OUT/ref_shim.o" 2>>"$OUT/build.log"
$FC -shared -o "$OUT/libref.so" "$OUT/cft_buff.o" "$OUT/rft_buff.o" "$OUT/declaration.o" \
    "$OUT/tridLU_3D.o" "$OUT/ref_shim.o" 2>>"$OUT/build.log"
echo "build_ref: built $OUT/libref.so"
