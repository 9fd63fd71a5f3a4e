#!/bin/bash
# Builds the reference Fortran program with flang into oracle/_ref/ (TEST INFRASTRUCTURE).
#
# Usage: oracle/build_ref.sh <lord> <strat> [itraj] [iprc] [relexit]
#   lord    : GCR order written into `parameter (lordx=..,lord=..)` (reference default 9)
#   strat   : stratification N^2/g written over the hard-coded `strat=1.e-05`
#   itraj   : advector scheme written over the hard-coded `itraj=0` (1: velprd)
#   iprc    : gcrk's preconditioner switch written over its hard-coded `iprc=1` (0: none,
#             precon's own iprc=0 branch p = rhs)
#   relexit : 1 adds a relative exit test to gcrk: eps = eps0 * max|b|, b the residual of
#             a zero pressure (prfr + div of the incoming velocities), in place of eps0/dt
#
# The source is read where it lies (REF_SRC, default under /root/reference) and a patched
# copy is written to oracle/_ref/ only (git-ignored, never committed). Patches:
#   1. program nftincv10 -> subroutine nftmain, so ref_harness.f90 can be the main program
#   2. neighbour lists widened from 10 to 32 entries (extruded tet meshes have up to ~22
#      neighbours; 10 overflows the array)
#   3. GCR order `lord`
#   4. stratification constant (and, when given, the advector scheme itraj)
#   5. restart-file formats f14.6 -> es26.17e3 so the final state is written at full
#      precision
#   6. (optional) gcrk's iprc and the relative exit test above
# The build is FP64 (-fdefault-real-8 -fdefault-double-8), as the reference's REAL*4
# default would otherwise apply.
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
REF_SRC=${REF_SRC:-/root/reference/lippsLFroude_implgwopks3D_unstr3.f}
LORD=${1:-3}
STRAT=${2:-1.e-05}
ITRAJ=${3:-0}
IPRC=${4:-1}
RELEXIT=${5:-0}
FC=${FC:-flang}
OUT="$HERE/_ref"
TAG="l${LORD}_s${STRAT}"
if [ "$ITRAJ" != "0" ]; then TAG="${TAG}_t${ITRAJ}"; fi
if [ "$IPRC" != "1" ]; then TAG="${TAG}_p${IPRC}"; fi
if [ "$RELEXIT" != "0" ]; then TAG="${TAG}_r"; fi
mkdir -p "$OUT"
if [ ! -f "$REF_SRC" ]; then
  echo "reference source not found: $REF_SRC" >&2
  exit 2
fi
# relative exit test, inserted after gcrk's `iprc=` line (nml is set by then; r, qr and
# u2, v2, w2 are overwritten by the reference's own start of the solve right after)
REL="$OUT/relexit_${TAG}.inc"
if [ "$RELEXIT" != "0" ]; then
  cat > "$REL" <<'EOR'
c     relative exit: eps = eps0 * max|b|, b = residual of a zero pressure
      do k=1,nml
        qr(k)=0.
      enddo
      call prfr(qr,u1,v1,w1,u2,v2,w2,impfl)
      call div(u2,v2,w2,r,-1)
      bmxrel=0.
      do k=1,nml
        bmxrel=amax1(bmxrel,abs(r(k)))
      enddo
      eps=eps0*bmxrel
EOR
else
  : > "$REL"
fi
sed -e '1s/^      program nftincv10/      subroutine nftmain/' \
    -e 's/(npoinx,10)/(npoinx,32)/g' \
    -e "s/parameter (lordx=9,nn=npoinx,lord=9)/parameter (lordx=${LORD},nn=npoinx,lord=${LORD})/" \
    -e "s/^      strat=1.e-05/      strat=${STRAT}/" \
    -e "s/^      itraj=0\(\r\?\)\$/      itraj=${ITRAJ}\1/" \
    -e 's/format(3f14.6)/format(3es26.17e3)/' \
    -e 's/format(4f14.6)/format(4es26.17e3)/' \
    -e "s/^      iprc=1\(\r\?\)\$/      iprc=${IPRC}\1/" \
    -e "/^      iprc=[01]\r\?\$/r $REL" \
    "$REF_SRC" > "$OUT/ref_${TAG}.f"
FFLAGS="-O3 -fdefault-real-8 -fdefault-double-8"
$FC $FFLAGS -c "$OUT/ref_${TAG}.f" -o "$OUT/ref_${TAG}.o"
$FC $FFLAGS -c "$HERE/ref_harness.f90" -o "$OUT/ref_harness.o" -J "$OUT"
$FC $FFLAGS "$OUT/ref_harness.o" "$OUT/ref_${TAG}.o" -o "$OUT/nft_ref_${TAG}"
echo "$OUT/nft_ref_${TAG}"
