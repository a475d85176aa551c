#!/usr/bin/env bash
# Builds the reference solver (DNS-solver/src, read where it lies under
# /root/reference) for one case into oracle/_ref/<name>/run-dns.
#
# The reference takes its case as compile-time macros in definitions.h
# (definitions.h:30-59) and includes that header by quote from its own
# directory.  This recipe writes a case copy of definitions.h with the
# requested macro values into oracle/_ref/<name>/ (a build product, never
# committed) and compiles the untouched sources with `-I<casedir>
# -I<srcdir> -I-`, which makes the quote-include search skip the source
# directory.  No source file is modified or copied.
#
# usage: build_ref.sh NAME XDV YDV ZDV REBS UB CFL TSTR TEND ITER [PSLIP]
#   ITER = number passed to `#define ITER` (the loop runs ITER+1 steps,
#          dns-const-q-xycut.c:235-237,485)
#   PSLIP = literal word PSLIP to enable slip posts (slip_posts_init.c)
# REF_XDIM (environment, default 1): MPI ranks along x (definitions.h:58,
# the rank count the binary must be started with; XDV a multiple of it)
# REF_SHIM=1 (environment): also link <name>/run-dns-gpu, the same sources
# with dns-solvers_amd/driver/xcomputations_gpu.c in place of
# xcomputations.c, against the built dns-solvers_amd/lib/libdnsgpu.so: the
# reference's own main on the GPU step (one rank)
set -euo pipefail
REF=${DNS_REFERENCE_SRC:-/root/reference/DNS-solver/src}
HERE=$(cd "$(dirname "$0")" && pwd)
NAME=$1; XDV=$2; YDV=$3; ZDV=$4; REBS=$5; UB=$6; CFL=$7; TSTR=$8; TEND=$9; ITER=${10}
SLIP=${11:-}
XDIM=${REF_XDIM:-1}
OUT=$HERE/_ref/$NAME
if [ ! -d "$REF" ]; then echo "reference sources not found at $REF" >&2; exit 3; fi
mkdir -p "$OUT"
sed -e "s/^#define Rebs .*/#define Rebs $REBS/" \
    -e "s/^#define ub .*/#define ub $UB/" \
    -e "s/^#define zdv .*/#define zdv $ZDV/" \
    -e "s/^#define ydv .*/#define ydv $YDV/" \
    -e "s/^#define xdv .*/#define xdv $XDV/" \
    -e "s/^#define TSTR .*/#define TSTR $TSTR/" \
    -e "s/^#define TEND .*/#define TEND $TEND/" \
    -e "s/^#define CFL .*/#define CFL $CFL/" \
    -e "s/^\/\/#define ITER .*/#define ITER $ITER/" \
    -e "s/^#define XDIM .*/#define XDIM $XDIM/" \
    "$REF/definitions.h" > "$OUT/definitions.h.tmp"
if [ "$SLIP" = "PSLIP" ]; then
  sed -i -e "s/^\/\/#define PSLIP/#define PSLIP/" "$OUT/definitions.h.tmp"
fi
mv "$OUT/definitions.h.tmp" "$OUT/definitions.h"
MPICC=${MPICC:-/opt/conda/bin/mpicc}
# -fcommon: definitions.h:123 has a tentative definition in a header.
# No -march: plain x86-64 SSE2 double arithmetic, no FMA contraction.
MPICH_CC=gcc "$MPICC" -O3 -fcommon -w -I"$OUT" -I"$REF" -I- "$REF"/*.c -lm -o "$OUT/run-dns.tmp" 2> "$OUT/build.log" \
  || { cat "$OUT/build.log" >&2; exit 1; }
mv "$OUT/run-dns.tmp" "$OUT/run-dns"
echo "$OUT/run-dns"
if [ "${REF_SHIM:-0}" = "1" ]; then
  REPO=$(cd "$HERE/.." && pwd)
  SRCS=$(ls "$REF"/*.c | grep -v '/xcomputations\.c$')
  # gcc with the MPI wrapper's flags spelled out: the HIP runtime needs the
  # system C++ runtime, which the wrapper's library directory (an older one)
  # would shadow, at link time and through the run path
  MPIDIR=$(dirname "$(dirname "$MPICC")")
  SYSLIB=/usr/lib/x86_64-linux-gnu
  gcc -O3 -fcommon -w -I"$OUT" -I"$REPO/include" -I"$REF" -I- -I"$MPIDIR/include" $SRCS \
    "$REPO/dns-solvers_amd/driver/xcomputations_gpu.c" -L"$REPO/dns-solvers_amd/lib" -ldnsgpu \
    -Wl,-rpath-link,"$SYSLIB" -L"$MPIDIR/lib" -lmpi -lm -Wl,--disable-new-dtags \
    -Wl,-rpath,"$SYSLIB:$MPIDIR/lib" -Wl,-rpath,'$ORIGIN/../../../dns-solvers_amd/lib' \
    -o "$OUT/run-dns-gpu.tmp" 2> "$OUT/build-gpu.log" || { cat "$OUT/build-gpu.log" >&2; exit 1; }
  mv "$OUT/run-dns-gpu.tmp" "$OUT/run-dns-gpu"
  echo "$OUT/run-dns-gpu"
fi
