#!/bin/bash
# Builds, from the reference's own sources where they lie under
# /root/reference (nothing is copied into this repository), into
# oracle/_ref/ only:
#
# 1. libhacapk_ref.so:
#      m_ppohBEM_user_func.f90        face_integral (the entry kernel)
#      m_ppohBEM_matrix_element_ij.f90 centroid collocation
#      m_HACApK_calc_entry_ij.f90     HACApK_entry_ij wrapper module
#      HACApK_lib.f90                 dotp / scaled norm / pivot helpers / med3
#    plus oracle/ref_bind.f90, a bind(C) binding that lets the tests call the
#    assumed-shape routines of HACApK_lib.f90 from C.  These four files
#    compile unmodified with flang; the rest of the reference needs MPI or
#    MAGMA and is not built, except:
#      ACA+: HACApK_calc_vec and HACApK_acaplus (m_HACApK_base.f90:677-699,
#      783-975), which use no MPI, are cut out of the preprocessed module
#      together with its header and interface block into a temporary
#      module ref_aca_mod (text unmodified; only the module name differs),
#      and oracle/ref_aca_bind.f90 forwards a C call to HACApK_acaplus.
#
# 2. ref_layout_magma.txt, ref_layout_plain.txt: the byte layout flang gives
#    the reference's declarations of st_HACApK_leafmtx / st_HACApK_leafmtxp
#    (m_HACApK_base.f90:54-235), preprocessed with and without -DHAVE_MAGMA
#    -DHAVE_MAGMA_BATCH -DISO_C_BINDING; printed by oracle/ref_layout.f90.
#
# 3. ref_driver: oracle/ref_driver.f90 compiled against the same
#    declarations (HAVE_MAGMA build) and linked with libhmv.so, so that a
#    test can run the reference's Fortran calling sequence on the device.
#    Needs hacapk-magma_amd/lib/libhmv.so built first.
#
# The type declarations are cut out of the preprocessed module into a
# temporary directory that is removed afterwards.
set -euo pipefail
REF=/root/reference
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
LIBDIR="$HERE/../hacapk-magma_amd/lib"
FLANG="${FLANG:-/opt/rocm/lib/llvm/bin/flang}"
mkdir -p "$OUT/mod"
TMP="$(mktemp -d)"
trap 'rm -rf "$TMP"' EXIT
: >"$OUT/build.log"

# ACA+ routines of m_HACApK_base.f90 as module ref_aca_mod: the module
# header (use, implicit), its interface block, HACApK_calc_vec and
# HACApK_acaplus, from the preprocessed source (plain build).
"$FLANG" -E -cpp "$REF/m_HACApK_base.f90" 2>>"$OUT/build.log" |
  awk '
    /^[ \t]*module[ \t]+m_HACApK_base[ \t]*$/ { print "module ref_aca_mod"; hdr = 1; next }
    hdr && /^[ \t]*type[ \t]*::/ { hdr = 0 }
    hdr && !/^#/ { print; next }
    !done_if && /^[ \t]*interface[ \t]*$/ { inif = 1 }
    inif && !/^#/ { print }
    inif && /^[ \t]*end[ \t]*interface[ \t]*$/ { inif = 0; done_if = 1; print "contains" }
    /^[ \t]*subroutine[ \t]+HACApK_calc_vec[ \t]*\(/ { cv = 1 }
    cv && !/^#/ { print }
    cv && /^[ \t]*end[ \t]*subroutine/ { cv = 0 }
    /^[ \t]*integer[ \t]+function[ \t]+HACApK_acaplus[ \t]*\(/ { ap = 1 }
    ap && !/^#/ { print }
    ap && /^[ \t]*end[ \t]*function/ { ap = 0 }
    END { print "end module ref_aca_mod" }
  ' >"$TMP/ref_aca_mod.f90"
grep -q "function HACApK_acaplus" "$TMP/ref_aca_mod.f90"
grep -q "subroutine HACApK_calc_vec" "$TMP/ref_aca_mod.f90"

"$FLANG" -O2 -fPIC -shared -J "$OUT/mod" -o "$OUT/libhacapk_ref.so.tmp" \
  "$REF/m_ppohBEM_user_func.f90" \
  "$REF/m_ppohBEM_matrix_element_ij.f90" \
  "$REF/m_HACApK_calc_entry_ij.f90" \
  "$REF/HACApK_lib.f90" \
  "$HERE/ref_bind.f90" \
  "$TMP/ref_aca_mod.f90" \
  "$HERE/ref_aca_bind.f90" 2>>"$OUT/build.log"
mv "$OUT/libhacapk_ref.so.tmp" "$OUT/libhacapk_ref.so"
# the three derived types of m_HACApK_base.f90 as the given flags declare them
extract_types() {  # $1 = output file, rest = preprocessor flags
  local out="$1"
  shift
  "$FLANG" -E -cpp "$@" "$REF/m_HACApK_base.f90" 2>>"$OUT/build.log" |
    awk '
      /^[ \t]*type[ \t]*::[ \t]*st_HACApK_(leafmtx|leafmtxp|lcontrol)[ \t]*$/ { on = 1 }
      on && !/^#/ { print }
      /^[ \t]*end[ \t]+type[ \t]+st_HACApK_(leafmtx|leafmtxp|lcontrol)[ \t]*$/ { on = 0 }
    ' >"$out"
  grep -q "end type st_HACApK_leafmtxp" "$out"
}
mkdir -p "$TMP/magma" "$TMP/plain"
extract_types "$TMP/magma/ref_types.inc" -DHAVE_MAGMA -DHAVE_MAGMA_BATCH -DISO_C_BINDING
extract_types "$TMP/plain/ref_types.inc"
"$FLANG" -O0 -cpp -DREF_MAGMA -I "$TMP/magma" -J "$TMP/magma" "$HERE/ref_layout.f90" \
  -o "$TMP/layout_magma" 2>>"$OUT/build.log"
"$FLANG" -O0 -cpp -I "$TMP/plain" -J "$TMP/plain" "$HERE/ref_layout.f90" \
  -o "$TMP/layout_plain" 2>>"$OUT/build.log"
"$TMP/layout_magma" >"$OUT/ref_layout_magma.txt"
"$TMP/layout_plain" >"$OUT/ref_layout_plain.txt"

if [ -f "$LIBDIR/libhmv.so" ]; then
  "$FLANG" -O2 -cpp -I "$TMP/magma" -J "$TMP/magma" "$HERE/ref_driver.f90" \
    -o "$OUT/ref_driver.tmp" -L "$LIBDIR" -lhmv \
    -Wl,-rpath,'$ORIGIN/../../hacapk-magma_amd/lib' -Wl,-rpath,/opt/rocm/lib 2>>"$OUT/build.log"
  mv "$OUT/ref_driver.tmp" "$OUT/ref_driver"
fi
