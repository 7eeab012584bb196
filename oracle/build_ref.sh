#!/bin/bash
# Builds the reference CPU solver (VAC 4.52) from its Fortran sources under
# /root/reference/src into oracle/_ref/, one executable per configuration.
#
#   oracle/build_ref.sh NAME NDIMNDIR PHYS G1[,G2[,G3]] [cpu|gpu] [poisson]
#   e.g. oracle/build_ref.sh vac_mhd22_132 22 mhd 132,132
#
# The user module is VAC's nul one (setvac -u=nul) unless VAC_USR names
# another of the reference's (e.g. VAC_USR=gravity: vacusr.t.gravity,
# vacusrpar.t.gravity and the vacusr.gravity.t they include).
#
# With "poisson" the Poisson module is switched on (what setvac -on=poisson
# sets: the $if_poisson switch of vacpp.pl), so process() can call projectb
# and vacpoisson.t is compiled in.
#
# With the fifth argument "gpu" the same VAC host is linked with the GPU
# drop-in instead of its own advect (vac_amd/fortran/vacgpu_advect.t, which
# calls libvacgpu through vac_amd/fortran/vacgpu_mod.f90): the CPU advect of
# vac.t is renamed advect_cpu and left unused, VAC's savefile is renamed
# savefile_cpu, which the drop-in's savefile calls after fetching the state,
# and the main loop's wold=w goes through the drop-in's vacgpu_store_wold.  This integration executable
# is what the drop-in test runs against the plain reference build.
#
# The recipe is this script, not the reference's setvac/Makefile:
#  * the needed .t sources are copied to a throw-away directory under /tmp,
#  * the physics module is chosen by copying vacpar.t.PHYS etc. to vacpar.t,
#  * the compile-time grid (ixGhi) line of vacdef.t is rewritten,
#  * two portability edits are applied (Perl >= 5.22 rejects defined(%hash)
#    in vacpp.pl:322; vacio.t's flushunit uses the NAG-only F90_UNIX_IO module),
#  * every .t file is expanded by the reference's own preprocessor
#    (perl vacpp.pl -d=NDIMNDIR), whose module switches already default to
#    cd,tvdlf,tvd on and mc,fct,impl,poisson,ct,gencoord,resist,rk,mpi off,
#  * the .f files are compiled with flang, FMA contraction off, so the
#    arithmetic is the plain IEEE sequence the source spells out.
# Only the final executable is written into oracle/_ref/; the scratch copy is
# deleted. Nothing from /root/reference is copied into the repository.
set -euo pipefail
NAME=$1; D=$2; PHYS=$3; G=$4; MODE=${5:-cpu}; POISSON=${6:-}
REF=${VAC_REF_SRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS=${FFLAGS:--O2 -ffp-contract=off}
[ -d "$REF" ] || { echo "reference sources not found at $REF" >&2; exit 2; }
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/vacref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
for f in vac.t vacio.t vacgrid.t vacgrid.setnozzle.t vacdef.t vaccd.t vactvdlf.t vactvd.t \
         vacusr.t.nul vacusrpar.t.nul vacnul.specialbound.t vacnul.specialini.t \
         vacnul.specialsource.t vacnul.specialio.t vacnul.addsource.t vacnul.process.t \
         vacpar.t.$PHYS vacphys0.t.$PHYS vacphys.t.$PHYS vacpp.pl; do
  cp "$REF/$f" .
done
for f in "$REF"/vacphys.*.t "$REF"/vacproc.*.t; do cp "$f" .; done
USR=${VAC_USR:-nul}
if [ "$USR" != nul ]; then
  cp "$REF/vacusr.t.$USR" "$REF/vacusrpar.t.$USR" "$REF/vacusr.$USR.t" .
fi
chmod u+w ./*
cp vacpar.t.$PHYS vacpar.t; cp vacphys0.t.$PHYS vacphys0.t; cp vacphys.t.$PHYS vacphys.t
cp vacusr.t.$USR vacusr.t; cp vacusrpar.t.$USR vacusrpar.t
# portability edit 1: Perl >= 5.22
sed -i 's/defined(%shift0)/%shift0/' vacpp.pl
# portability edit 2: F2003 flush statement instead of the NAG module
perl -0pi -e 's/use F90_UNIX_IO,only : flush[^\n]*\n//; s/call flush\(qunit,ierror\)[^\n]*/flush(qunit)/' vacio.t
if [ "$POISSON" = poisson ]; then
  cp "$REF/vacpoisson.t" .; chmod u+w vacpoisson.t
  perl -pi -e 's/\$if_poisson=0;/\$if_poisson=1;/' vacpp.pl
  grep -q '$if_poisson=1;' vacpp.pl || { echo "could not switch the Poisson module on" >&2; exit 1; }
fi
# compile-time grid
IFS=, read -r -a GA <<< "$G"
line="INTEGER,PARAMETER:: "; gmin=${GA[0]}; gmax=${GA[0]}
for i in "${!GA[@]}"; do
  line+="ixGhi$((i+1))=${GA[$i]},"
  [ "${GA[$i]}" -lt "$gmin" ] && gmin=${GA[$i]}
  [ "${GA[$i]}" -gt "$gmax" ] && gmax=${GA[$i]}
done
line+="ixGhimin=$gmin,ixGhimax=$gmax"
perl -pi -e "s/^INTEGER,PARAMETER:: ixGhi1=.*\$/$line/" vacdef.t
MODS="vac vacio vacgrid vacphys0 vacphys vacusr vaccd vactvdlf vactvd"
[ "$POISSON" = poisson ] && MODS="$MODS vacpoisson"
LINK=""
if [ "$MODE" = gpu ]; then
  REPO=$(cd "$HERE/.." && pwd)
  cp "$REPO/vac_amd/fortran/vacgpu_advect.t" "$REPO/vac_amd/fortran/vacgpu_mod.f90" .
  perl -pi -e 's/^subroutine advect\(method/subroutine advect_cpu(method/' vac.t
  # the drop-in's savefile fetches the device state, then calls VAC's
  perl -pi -e 's/^subroutine savefile\(ifile,w\)/subroutine savefile_cpu(ifile,w)/' vacio.t
  grep -q '^subroutine savefile_cpu(ifile,w)' vacio.t || { echo "savefile not found in vacio.t" >&2; exit 1; }
  # the main loop's copy of w into wold (src/vac.t:85) goes through the
  # drop-in, which skips it while the host's w lags the device state
  perl -pi -e 's/^(\s*)wold\(ixG\^S,1:nw\)=w\(ixG\^S,1:nw\)\s*$/$1call vacgpu_store_wold(w)\n/' vac.t
  grep -q 'call vacgpu_store_wold(w)' vac.t || { echo "wold copy not found in vac.t" >&2; exit 1; }
  $FC $FFLAGS -c vacgpu_mod.f90 -o vacgpu_mod.o
  MODS="$MODS vacgpu_advect"
  if [ "$POISSON" = poisson ]; then
    # VAC's projectb is renamed and unused: every projection runs on the device
    cp "$REPO/vac_amd/fortran/vacgpu_projectb.t" .
    perl -pi -e 's/^subroutine projectb\(w\)/subroutine projectb_cpu(w)/' vacproc.projectb.t
    MODS="$MODS vacgpu_projectb"
  fi
  LINK="vacgpu_mod.o -L$REPO/vac_amd -lvacgpu -Wl,-rpath,\$ORIGIN/../../vac_amd -L/opt/rocm/lib -Wl,-rpath,/opt/rocm/lib"
fi
for m in vacdef vacpar vacusrpar $MODS; do
  perl vacpp.pl -d=$D $m.t > $m.f
done
objs=""
for m in $MODS; do
  $FC $FFLAGS -ffree-form -c $m.f -o $m.o 2> $m.log || { cat $m.log >&2; exit 1; }
  objs+=" $m.o"
done
$FC $FFLAGS -o "$OUT/$NAME" $objs $LINK
echo "built $OUT/$NAME"
