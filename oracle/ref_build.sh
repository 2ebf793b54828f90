#!/bin/bash
# Builds oracle/_ref/libhacapk_ref.so from the reference's own sources, in
# place under /root/reference (nothing is copied into this repository):
#   m_ppohBEM_user_func.f90        face_integral (the entry kernel)
#   m_ppohBEM_matrix_element_ij.f90 centroid collocation
#   m_HACApK_calc_entry_ij.f90     HACApK_entry_ij wrapper module
#   HACApK_lib.f90                 dotp / scaled norm / pivot helpers / med3
# plus oracle/ref_bind.f90, a bind(C) binding that lets the tests call the
# assumed-shape routines of HACApK_lib.f90 from C.  These four files compile
# unmodified with flang; the rest of the reference needs MPI or MAGMA and is
# not built.  Output goes only to oracle/_ref/.
set -euo pipefail
REF=/root/reference
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
FLANG="${FLANG:-/opt/rocm/lib/llvm/bin/flang}"
mkdir -p "$OUT/mod"
"$FLANG" -O2 -fPIC -shared -J "$OUT/mod" -o "$OUT/libhacapk_ref.so.tmp" \
  "$REF/m_ppohBEM_user_func.f90" \
  "$REF/m_ppohBEM_matrix_element_ij.f90" \
  "$REF/m_HACApK_calc_entry_ij.f90" \
  "$REF/HACApK_lib.f90" \
  "$HERE/ref_bind.f90" 2>"$OUT/build.log"
mv "$OUT/libhacapk_ref.so.tmp" "$OUT/libhacapk_ref.so"
