#!/usr/bin/env bash
# Build the upstream Ixchel2D CPU program (serial flang build) for one mesh size.
#
# Test infrastructure only: the binary it writes is the reference that the
# golden fixtures under tests/golden/ and the "reference" CPU baseline come
# from.  Ixchel2D fixes its mesh extents at compile time
# (source/mod_malla.f90:15, `mi`, `nj`), so one binary is built per size.
#
# The upstream sources are compiled where they lie; only the mesh-extent line
# is changed, in a throw-away copy under a temporary directory that is removed
# afterwards.  OpenACC directives stay comments (flang rejects -fopenacc for
# AMD targets), so this is the reference's serial CPU build.  The only output
# is   oracle/_ref/ixchel2d_ref_m<MI>_n<NJ>
#
# With BODY set (trian | cuadr | chime) the driver's call
# definir_cuerpo(..., 'sincu') (source/main.f90:327) passes that option
# instead, in the same throw-away copy, and the binary is named
# ixchel2d_ref_m<MI>_n<NJ>_<BODY>: upstream's immersed-body shapes
# (source/mod_front_inmersa.f90:54-179) with nothing else changed.
#
# usage: [BODY=cuadr] oracle/build_ref.sh MI NJ [REFERENCE_ROOT]
set -euo pipefail
MI=${1:?mi}
NJ=${2:?nj}
REF=${3:-/root/reference}
SRC="$REF/source"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
FLANG=${FLANG:-/opt/rocm/lib/llvm/bin/flang}
BODY=${BODY:-}
mkdir -p "$OUT"
BIN="$OUT/ixchel2d_ref_m${MI}_n${NJ}${BODY:+_$BODY}"
if [[ -x "$BIN" && "${FORCE:-0}" != 1 ]]; then
  echo "$BIN"
  exit 0
fi
TMP="$(mktemp -d /tmp/ixchel2d_ref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
# the one configuration change upstream expects per mesh: mi, nj
sed -e "s/mi=128, nj=128/mi=${MI}, nj=${NJ}/" "$SRC/mod_malla.f90" > "$TMP/mod_malla.f90"
grep -q "mi=${MI}, nj=${NJ}" "$TMP/mod_malla.f90"
cd "$TMP"
# module dependency order of source/CMakeLists.txt:43-52
FFLAGS="-O2 -fopenmp -I/opt/rocm/lib/llvm/include/flang"
"$FLANG" $FFLAGS -c "$TMP/mod_malla.f90"
MAIN="$SRC/main.f90"
if [[ -n "$BODY" ]]; then
  case "$BODY" in trian|cuadr|chime) ;; *) echo "unknown body $BODY" >&2; exit 2 ;; esac
  sed -e "s/gamma_energ, 'sincu')/gamma_energ, '${BODY}')/" "$SRC/main.f90" > "$TMP/main.f90"
  grep -q "gamma_energ, '${BODY}')" "$TMP/main.f90"
  MAIN="$TMP/main.f90"
fi
for f in mod_cond_frontera mod_solucionador mod_residuos mod_postproceso \
         mod_ec_momento mod_ec_energia mod_front_inmersa; do
  "$FLANG" $FFLAGS -c "$SRC/$f.f90"
done
"$FLANG" $FFLAGS -c "$MAIN"
"$FLANG" $FFLAGS -o "$BIN" mod_malla.o mod_cond_frontera.o mod_solucionador.o \
  mod_residuos.o mod_postproceso.o mod_ec_momento.o mod_ec_energia.o \
  mod_front_inmersa.o main.o
echo "$BIN"
