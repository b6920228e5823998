#!/bin/bash
# Builds the reference Fortran (MLFMAxwell) with flang into oracle/_ref/,
# together with ref_driver.f90, so that tests can check the C oracle against
# the reference itself.  TEST INFRASTRUCTURE ONLY: nothing in lywell_amd/
# links or runs what this produces.
#
# The reference sources are read where they lie (default /root/reference, or
# $LYWELL_REFERENCE).  flang rejects three ifort extensions in them, so the
# eight module files are copied to a temporary directory outside the
# repository, patched there by the three edits below, compiled, and the
# temporary copies are deleted.  No reference source is written into the
# repository; only the compiled driver lands in oracle/_ref/.
#   1. OPEN(..., CARRIAGECONTROL='FORTRAN') in the progress printer
#      (modGlobalMetod.f90:98,107): the specifier is dropped (console output
#      only).
#   2. LOGICAL compared with == (modIterativo.f90:242,258,407,423): .eqv.
#   3. REAL(8) values used as subscripts (modMLFMA.f90:1361,1386,1660):
#      wrapped in int(), which is what ifort does implicitly.
set -euo pipefail
REF="${LYWELL_REFERENCE:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/lib/llvm/bin/flang}"
if [ ! -f "$REF/modMLFMA.f90" ]; then
  echo "build_ref.sh: reference sources not found under $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
TMP="$(mktemp -d /tmp/lywell_ref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
MODS="modGlobalParam modGlobalMetod modReadStl modRWG modMoM modIterativo modRCS modMLFMA"
for m in $MODS; do cp "$REF/$m.f90" "$TMP/$m.f90"; done
sed -i "s/,CARRIAGECONTROL='FORTRAN'//" "$TMP/modGlobalMetod.f90"
sed -i 's/if (prec == \.false\.) then/if (prec .eqv. .false.) then/' "$TMP/modIterativo.f90"
sed -i 's/Datos(Mint(2,cont),:)/Datos(int(Mint(2,cont)),:)/' "$TMP/modMLFMA.f90"
sed -i 's/Vant(Mint(2,cont),:)/Vant(int(Mint(2,cont)),:)/g' "$TMP/modMLFMA.f90"
sed -i 's/W(i,octree(niv-1)%mat_interp(2,cont))/W(i,int(octree(niv-1)%mat_interp(2,cont)))/' "$TMP/modMLFMA.f90"
cd "$TMP"
objs=""
for m in $MODS; do
  "$FC" -O2 -c "$m.f90" -o "$m.o"
  objs="$objs $TMP/$m.o"
done
"$FC" -O2 -I"$TMP" -c "$HERE/ref_driver.f90" -o "$TMP/ref_driver.o"
"$FC" -O2 -o "$OUT/ref_driver" "$TMP/ref_driver.o" $objs
"$FC" -O2 -I"$TMP" -c "$HERE/ref_stl.f90" -o "$TMP/ref_stl.o"
"$FC" -O2 -o "$OUT/ref_stl" "$TMP/ref_stl.o" $objs
"$FC" -O2 -I"$TMP" -c "$HERE/ref_dat.f90" -o "$TMP/ref_dat.o"
"$FC" -O2 -o "$OUT/ref_dat" "$TMP/ref_dat.o" $objs
echo "built $OUT/ref_driver $OUT/ref_stl $OUT/ref_dat"
