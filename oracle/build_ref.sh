#!/usr/bin/env bash
# Test infrastructure only.  Compiles the unmodified ADECO reference CPU path where it lies
# (/root/reference/src) together with oracle/ref_harness.cpp into oracle/_ref/libadeco_ref.so.
# The reference program's main.cpp is compiled with its main renamed (-Dmain=adeco_ref_cli_main)
# so the harness can run the whole program in-process (ref_cli_main).
# Flags follow the reference's tested CPU build: -O3 -DNDEBUG, no -march (SSE2 only, so the
# compiler cannot contract float multiply-adds), C++11 with GNU extensions, MPICH from /opt/conda
# for the one-rank MPI calls inside generate_tree.  Outputs go to oracle/_ref/ only.
#
# When the product library adeco_amd/libadeco_amd.so is built, a second harness,
# oracle/_ref/libadeco_ref_dropin.so, is linked the way a maintainer would drop the library into
# the reference: the same main.cpp and harness, with adeco_amd/cpp/AdDeposit2Grid_adeco.cpp (the
# reference's DepositWrapper::deposit_tets forwarding to adeco_deposit_patches) compiled in place
# of AdDeposit2Grid.cpp.  The tests run the reference program through it on the GPU.
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
rm -f "$OUT/AdDeposit2Grid.o"
echo "built $OUT/libadeco_ref.so"
PRODUCT="$HERE/../adeco_amd"
if [ -f "$PRODUCT/libadeco_amd.so" ]; then
    g++ $CXXFLAGS -I"$HERE/../include" -c "$PRODUCT/cpp/AdDeposit2Grid_adeco.cpp" -o "$OUT/AdDeposit2Grid_adeco.o"
    g++ -shared -o "$OUT/libadeco_ref_dropin.so" "$OUT/AdDeposit2Grid_adeco.o" "$OUT/main.o" "$OUT/ref_harness.o" \
        -L"$MPI_LIB" -Wl,-rpath,/usr/lib/x86_64-linux-gnu -Wl,-rpath,"$MPI_LIB" -L"$PRODUCT" \
        -Wl,-rpath,'$ORIGIN/../../adeco_amd' -ladeco_amd \
        -lmpi -lpthread -fopenmp
    rm -f "$OUT/AdDeposit2Grid_adeco.o"
    echo "built $OUT/libadeco_ref_dropin.so"
fi
rm -f "$OUT/main.o" "$OUT/ref_harness.o"
