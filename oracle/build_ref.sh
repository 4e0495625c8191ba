#!/bin/sh
# Builds the reference solver (its own sources, where they lie under the
# reference checkout) together with oracle/ref_harness.cpp into oracle/_ref/.
# Two variants:
#   ref_native  the reference makefile's flags (-O3 -march=native; GCC contracts
#               a*b+c into FMA by default on an FMA-capable host)
#   ref_nofma   -O3 -ffp-contract=off, baseline x86-64 (no FMA instructions):
#               the IEEE operation order the oracle restates bit for bit.
# MPI comes from the image's MPICH (/opt/conda); the conda wrapper's default
# compiler is absent, so MPICH_CXX selects the system g++; conda's older
# libstdc++ sits first on the wrapper's -L path, so the C++ runtime is linked
# statically from the system toolchain.
set -e
REF=${WENOFD_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MPICXX=${MPICXX:-/opt/conda/bin/mpicxx}
if [ ! -f "$REF/src/WENOFD.cpp" ]; then
    echo "reference sources not found at $REF; skipping oracle/_ref" >&2
    exit 0
fi
if [ ! -x "$MPICXX" ]; then
    echo "mpicxx not found at $MPICXX; skipping oracle/_ref" >&2
    exit 0
fi
mkdir -p "$OUT"
export MPICH_CXX=${MPICH_CXX:-g++}
SRCS="$REF/src/WENOFD.cpp $HERE/ref_harness.cpp"
INC="-I$REF/include -I$REF/utils"
$MPICXX -O3 -march=native $INC $SRCS -static-libstdc++ -static-libgcc -o "$OUT/ref_native"
$MPICXX -O3 -ffp-contract=off $INC $SRCS -static-libstdc++ -static-libgcc -o "$OUT/ref_nofma"
echo "built $OUT/ref_native $OUT/ref_nofma"
