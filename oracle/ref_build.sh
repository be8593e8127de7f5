#!/bin/bash
# Builds oracle/_ref/libref_chlld.so from the reference's own sources where
# they lie under /root/reference (nothing is copied into the repository).
#   Storage_module.f90            compiled unmodified (it provides cpi4)
#   Solvers.f90:108-150           signum, minmod, linear
#   Solvers.f90:1143-1859         chlld
# Solvers.f90 as a whole cannot compile here: it needs the CUDA Fortran
# module cudafor.  The two routine ranges need nothing but STORAGE, so the
# recipe cuts them out (sed, unmodified text) into a module source in _ref/.
# The '!@cuf attributes(host, device)' lines are comments for a compiler
# without CUDA Fortran.  The cut-out module source and the compiler's
# module and object files are deleted once the library is linked, so
# oracle/_ref/ (git-ignored) ends up holding only libref_chlld.so.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${SET_REFERENCE:-/root/reference}/SET_Fortran_Cuda_MHD"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/llvm/bin/flang}"
if [ ! -f "$REF/Solvers.f90" ]; then
  echo "reference sources not present at $REF; skipping _ref build" >&2
  exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
{
  echo "module ref_chlld_mod"
  echo "use STORAGE"
  echo "contains"
  sed -n '108,150p' "$REF/Solvers.f90"
  sed -n '1143,1859p' "$REF/Solvers.f90"
  echo "end module ref_chlld_mod"
} > ref_chlld_mod.f90
FFLAGS="-O2 -fPIC -ffp-contract=off"
"$FC" $FFLAGS -c "$REF/Storage_module.f90" -o storage.o
"$FC" $FFLAGS -c ref_chlld_mod.f90 -o ref_chlld_mod.o
"$FC" $FFLAGS -c "$HERE/ref_bind.f90" -o ref_bind.o
"$FC" -shared -fPIC -o libref_chlld.so storage.o ref_chlld_mod.o ref_bind.o
rm -f ref_chlld_mod.f90 storage.mod ref_chlld_mod.mod ref_bind.mod storage.o ref_chlld_mod.o ref_bind.o
echo "built $OUT/libref_chlld.so"

# SET_ENA's rate functions (SET_ENA/SET_ENA.cpp): the file as a whole needs
# CGAL, Eigen and Boost, none of which is here; these routines need only
# <cmath>.  Cut, unmodified: the macros (:41-66), const_pi (:87),
# Velosity_1 and maxwell_distribution (:341-370), int_1_f1/f2/f3 and int_1
# (:426-542).
ENA="${SET_REFERENCE:-/root/reference}/SET_ENA/SET_ENA.cpp"
if [ -f "$ENA" ]; then
  { sed -n '41,66p' "$ENA"; sed -n '87p' "$ENA"; sed -n '341,370p' "$ENA"; sed -n '426,542p' "$ENA"; } \
    > "$OUT/ref_ena_cut.inc"
  g++ -O2 -fPIC -ffp-contract=off -shared -I "$OUT" -o "$OUT/libref_ena.so" "$HERE/ref_ena_bind.cpp"
  rm -f "$OUT/ref_ena_cut.inc"
  echo "built $OUT/libref_ena.so"
fi
