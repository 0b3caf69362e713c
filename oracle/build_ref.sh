#!/usr/bin/env bash
# Build the reference FIM code itself, from its sources where they lie under
# /root/reference, into oracle/_ref/.  Test infrastructure only: nothing in the
# product links or runs what this script builds.
#
#   oracle/_ref/bin/grid, ginfo, getlvl   reference grid prep chain (default real 4,
#                                          as PREP_FFLAGS in macros.make.linuxpcgnu:25)
#   oracle/_ref/bin/lsfitref              getlvl's least-squares pseudo-inverse
#                                          (LSfitting + sgesvd) on given point sets
#   oracle/_ref/bin/dynref_r8             reference dynamics hot path, -fdefault-real-8
#   oracle/_ref/bin/dynref_r4             reference dynamics hot path, default real 4
#   oracle/_ref/bin/dynref_r8omp          the r8 build with -fopenmp (timing calibration)
#   oracle/_ref/bin/restref_r4            reference restart writer/reader of the
#                                          dynamics section, default real 4
#   oracle/_ref/bin/hycomref              reference HYCOM ocean continuity equation
#                                          (hycom_cnuity.F90), -fdefault-real-8
#   oracle/_ref/bin/hyctrref              reference HYCOM tracer transport
#   oracle/_ref/bin/hychybref             reference HYCOM hybrid grid generator (hybgn1)
#                                          (hycom_transp3d.F90 with hycom_fct3d.F90
#                                          and hycom_dffusn.F90), -fdefault-real-8
#   oracle/_ref/bin/hycmomref             reference HYCOM momentum equation
#                                          (hycom_momtum.F90 with hycom_dissip.F90),
#                                          -fdefault-real-8
#   oracle/_ref/bin/hycedgref             reference HYCOM edge interpolation
#                                          (hycom_edgvar.F90 bclin_edgvar), then
#                                          optionally cnuity, -fdefault-real-8
#
# The two dynref programs link the reference's own modules (cnuity, mass_adv
# and its flux limiter, trcadv, the class-B transport transp3d/fct3d, edgvar1,
# edgvar3, momtum, hybgen, hystat, abstart, globsum, dffusn and the
# control/constant modules they use) to the
# driver oracle/ref/dynref.F90.
# No reference source is copied; no stand-in modules are written.
set -euo pipefail
REF=${FIM_REFERENCE:-/root/reference}
S=$REF/FIMsrc
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-amdflang}
if [ ! -d "$S" ]; then
  echo "build_ref: $S not present; reference build skipped" >&2
  exit 0
fi
if ! command -v "$FC" >/dev/null 2>&1; then
  echo "build_ref: $FC not found; reference build skipped" >&2
  exit 0
fi
mkdir -p "$OUT/bin"

CNTL_SRC="$S/cntl/units.F90 $S/utils/module_initial_chem_namelists.F90 $S/cntl/fimnamelist.F90
          $S/cntl/module_wrf_control.F90 $S/cntl/module_control.F90"

compile() {  # compile <objdir> <flags> <files...>
  local od=$1; shift; local fl=$1; shift
  mkdir -p "$od"
  for f in "$@"; do
    local b; b=$(basename "${f%.*}")
    if [ ! -f "$od/$b.o" ] || [ "$f" -nt "$od/$b.o" ]; then
      (cd "$od" && $FC -c $fl -J"$od" -I"$od" "$f" -o "$od/$b.o")
    fi
  done
}

# ---- grid prep chain (default real 4) ------------------------------------
G=$S/prep/grid
GOBJ=$OUT/obj_grid
GFL="-cpp -DSERIAL -O2 -ffp-contract=off -I$S/dummygptl"
compile "$GOBJ" "$GFL" $CNTL_SRC $S/utils/headers.F90 $S/sys_share/sys_share.F90 \
  $G/datastru.F90 $G/avg.F90 $G/middle.F90 $G/third.F90 $G/rotate.F90 \
  $G/top_gridpoints.F90 $G/top_triangles.F90 $G/trisect.F90 $G/trisect_triangle.F90 \
  $G/bisect_triangle.F90 $G/icos.F90 \
  $G/GetRegions.F90 $G/hilbert.F90 $G/ijblock.F90 $G/IJblockLayout.F90 $G/mpi_stubs.F90 \
  $G/perm.F90 $G/SquareDecomp.F90 $G/SquareLayout.F90 $G/GridGen.F90 \
  $S/prep/lsfit/LSfitting.F90 $G/getlvl.F90
compile "$GOBJ" "-O2 -ffp-contract=off" $S/prep/lsfit/svd.f $S/prep/lsfit/blas.f $S/prep/lsfit/la_aux.f
C="$GOBJ/headers.o $GOBJ/module_initial_chem_namelists.o $GOBJ/fimnamelist.o $GOBJ/module_control.o
   $GOBJ/module_wrf_control.o $GOBJ/units.o $GOBJ/sys_share.o"
$FC -o "$OUT/bin/grid" $GOBJ/{avg,bisect_triangle,datastru,icos,middle,rotate,third,top_gridpoints,top_triangles,trisect,trisect_triangle}.o $C
$FC -o "$OUT/bin/ginfo" $GOBJ/{datastru,GetRegions,GridGen,hilbert,ijblock,IJblockLayout,mpi_stubs,perm,SquareDecomp,SquareLayout}.o $C
$FC -o "$OUT/bin/getlvl" $GOBJ/{getlvl,LSfitting,svd,blas,la_aux}.o $C
# the least-squares pseudo-inverse alone, on given point sets
compile "$GOBJ" "$GFL" "$HERE/ref/lsfitref.F90"
$FC -o "$OUT/bin/lsfitref" $GOBJ/{lsfitref,LSfitting,svd,blas,la_aux}.o $C

# ---- dynamics hot path, two precisions -----------------------------------
H=$S/fim/horizontal
DYN_SRC="$CNTL_SRC $S/cntl/module_constants.F90 $S/cntl/module_variables.F90
  $H/global_bounds.F90 $S/dummygptl/gptl.F90 $H/findmxmn.F90 $H/stencilprint.F90 $H/stenedgprint.F90
  $H/dffusn.F90 $H/cnuity_compute.F90 $H/cnuity.F90 $H/trcadv_compute.F90 $H/trcadv.F90
  $H/flux_limiter.F90 $H/massadv_compute.F90 $H/massadv.F90 $H/fct3d.F90 $H/transp3d.F90
  $H/edgvar1.F90 $H/edgvar3.F90 $H/momtum_compute.F90 $H/momtum.F90 $H/hybgen.F90
  $H/hystat.F90 $H/abstart.F90 $H/globsum.F90 $H/GetIpnGlobal.F90 $S/prep/ss2icos/lay2lay.F90 $S/prep/ss2icos/lin2stp.F90"
# r8omp: the FP64 build with the reference's own OpenMP directives on
# (-fopenmp), the CPU path the restatement's timing is calibrated against
for prec in r8 r4 r8omp; do
  case $prec in
    r8) PF="-fdefault-real-8 -fdefault-double-8" ;;
    r4) PF="" ;;
    r8omp) PF="-fdefault-real-8 -fdefault-double-8 -fopenmp" ;;
  esac
  OD=$OUT/obj_$prec
  FL="-cpp -DSERIAL -O2 -ffp-contract=off $PF -I$H -I$S/dummygptl"
  compile "$OD" "$FL" $DYN_SRC "$HERE/ref/dynref.F90"
  objs=""
  for f in $DYN_SRC "$HERE/ref/dynref.F90"; do b=$(basename "${f%.*}"); objs="$objs $OD/$b.o"; done
  LF=""
  [ $prec = r8omp ] && LF="-fopenmp"
  $FC $LF -o "$OUT/bin/dynref_$prec" $objs
done

# ---- HYCOM ocean continuity equation (hycom_cnuity.F90), FP64 --------------
HYC_SRC="$CNTL_SRC $S/cntl/module_constants.F90 $H/global_bounds.F90 $S/dummygptl/gptl.F90 $H/findmxmn.F90
  $H/stencilprint.F90 $H/stenedgprint.F90 $H/hycom_control.F90 $H/hycom_constants.F90 $H/hycom_variables.F90
  $H/hycom_sverbal.F90 $H/hycom_diag.F90 $H/hycom_cnuity.F90 $H/hycom_edgvar.F90 $H/hycom_dissip.F90
  $H/hycom_momtum.F90 $H/hycom_fct3d.F90 $H/hycom_dffusn.F90 $H/hycom_transp3d.F90 $H/hycom_sigetc.F90
  $H/hycom_hybgn1.F90 $H/GetIpnGlobal.F90 $H/its2time.F90"
OD=$OUT/obj_hycom
FL="-cpp -DSERIAL -O2 -ffp-contract=off -fdefault-real-8 -fdefault-double-8 -I$H -I$S/dummygptl"
compile "$OD" "$FL" $HYC_SRC "$HERE/ref/hycomref.F90" "$HERE/ref/hycedgref.F90" "$HERE/ref/hycmomref.F90" \
  "$HERE/ref/hyctrref.F90" "$HERE/ref/hychybref.F90"
for drv in hycomref hycedgref hycmomref hyctrref hychybref; do
  objs=""
  for f in $HYC_SRC "$HERE/ref/$drv.F90"; do b=$(basename "${f%.*}"); objs="$objs $OD/$b.o"; done
  $FC -o "$OUT/bin/$drv" $objs
done

# ---- restart writer/reader of the dynamics section (default real 4) -------
RST_SRC="$S/cntl/module_sfc_variables.F90 $H/restart_dyn.F90 $H/writearr32_r.F90 $H/readarr32_r.F90"
OD=$OUT/obj_r4
FL="-cpp -DSERIAL -O2 -ffp-contract=off -I$H -I$S/dummygptl"
compile "$OD" "$FL" $RST_SRC "$HERE/ref/restref.F90"
objs=""
for f in $DYN_SRC $RST_SRC "$HERE/ref/restref.F90"; do b=$(basename "${f%.*}"); objs="$objs $OD/$b.o"; done
$FC -o "$OUT/bin/restref_r4" $objs
echo "build_ref: reference programs in $OUT/bin"
