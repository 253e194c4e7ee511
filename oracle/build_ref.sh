#!/usr/bin/env bash
# Test infrastructure only.  Compiles the unmodified ADECO reference CPU path where it lies
# (/root/reference/src) together with oracle/ref_harness.cpp into oracle/_ref/libadeco_ref.so.
# The reference program's main.cpp is compiled with its main renamed (-Dmain=adeco_ref_cli_main)
# so the harness can run the whole program in-process (ref_cli_main).
# Flags follow the reference's tested CPU build: -O3 -DNDEBUG, no -march (SSE2 only, so the
# compiler cannot contract float multiply-adds), C++11 with GNU extensions, MPICH from /opt/conda
# for the one-rank MPI calls inside generate_tree.  Outputs go to oracle/_ref/ only.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${ADECO_REFERENCE_SRC:-/root/reference/src}"
OUT="$HERE/_ref"
if [ ! -f "$REF/AdDeposit2Grid.cpp" ]; then
    echo "build_ref.sh: reference sources not found at $REF; skipping" >&2
    exit 0
fi
MPI_INC=/opt/conda/include
MPI_LIB=/opt/conda/lib
mkdir -p "$OUT"
CXXFLAGS="-std=gnu++11 -O3 -DNDEBUG -DNO_HDF5 -fPIC -w -I$REF -I$MPI_INC"
g++ $CXXFLAGS -c "$REF/AdDeposit2Grid.cpp" -o "$OUT/AdDeposit2Grid.o"
g++ $CXXFLAGS -Dmain=adeco_ref_cli_main -c "$REF/main.cpp" -o "$OUT/main.o"
g++ $CXXFLAGS -c "$HERE/ref_harness.cpp" -o "$OUT/ref_harness.o"
g++ -shared -o "$OUT/libadeco_ref.so" "$OUT/AdDeposit2Grid.o" "$OUT/main.o" "$OUT/ref_harness.o" \
    -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpi -lpthread -fopenmp
rm -f "$OUT/AdDeposit2Grid.o" "$OUT/main.o" "$OUT/ref_harness.o"
echo "built $OUT/libadeco_ref.so"
