#!/bin/bash
# Build the reference's own compute modules (project-fluxo/fluxo Fortran, compiled where they
# lie under /root/reference/src) together with oracle/ref/fluxo_ref_harness.f90 into
# oracle/_ref/.  Test infrastructure only: the binaries produce golden values and the
# reported CPU baseline; the product never links or calls them.
#
# The reference's CMake build is not used.  The defines below are the ones its CMakeLists.txt
# sets for the named configurations (LGL nodes, split form, no MPI/AMR; shock capturing only in
# the NFVSE variant).  The
# two-point volume flux is fixed at compile time (FLUXO_EQN_VOLFLUX, CMakeLists.txt:242-245):
# flang's preprocessor rejects the run-time default's "#if PP_VolFlux==-1".  Translation units
# that need HDF5, the HOPR mesh reader, restart or the time-disc driver do not compile here and are not needed by the harness; the
# loop compiles in dependency order until no further file compiles; object files already
# built are kept, so a rebuild only recompiles the harness and relinks.
set -e
REF=${REF:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
FC=${FC:-flang}
[ -d "$REF" ] || { echo "reference sources not found at $REF; nothing built"; exit 0; }
mkdir -p "$OUT"
COMMON_NOOPT="-cpp -fdefault-real-8 -fdefault-double-8 -DPP_N=N -DPP_NodeType=2 -DPP_DiscType=2 -DMPI=0 -DUSE_AMR=0 -DPOSITIVITYPRES=0 -I$REF"
# golden-value builds: -O2; the CPU-baseline build (OPT set per variant below) optimises further
OPT_DEFAULT="-O2"

# VARIANTS="a b ..." builds only the named variants (all when unset)
build_variant() {
  local name=$1; shift
  case " ${VARIANTS:-$name} " in *" $name "*) ;; *) return 0 ;; esac
  local defs="$*"
  local COMMON="$COMMON_NOOPT ${OPT:-$OPT_DEFAULT}"
  local obj="$OUT/obj_$name"
  mkdir -p "$obj"
  # the lifting directory not selected by PP_Lifting is left out, as the reference's CMake does
  local eqn=navierstokes
  case "$defs" in *-Dmhd*) eqn=mhd ;; esac
  local others='mhd|maxwell|linearscalaradvection'
  [ "$eqn" = mhd ] && others='navierstokes|maxwell|linearscalaradvection'
  local lift=br2
  case "$defs" in *-DPP_Lifting=2*) lift=br1 ;; esac
  # FLUXO_TESTCASE (CMakeLists.txt:101-103): the testcase directory TC (default: default)
  local tc=${TC:-default}
  local tcs
  tcs=$(printf '%s\n' default taylorgreenvortex mhd_equilibrium ns_angularmomentum | grep -v -x "$tc" | paste -sd'|')
  local exclude="/amr/|/equation/($others)/|/testcase/($tcs)/|fluxo_post|/fluxo\.f90|/dg/$lift/"
  find "$REF" -name "*.f90" | grep -v -E "$exclude" | sort > "$obj/todo.txt"
  local pass n f b extra
  for pass in $(seq 1 15); do
    : > "$obj/next.txt"; n=0
    while read -r f; do
      b=$(basename "$f" .f90)
      # taylorgreenvortex/testcase_analyze.f90 tests '#ifdef MPI' where the rest of the code
      # tests '#if MPI'; with the serial build's MPI=0 it would take its MPI branch, so it is
      # compiled with MPI undefined, which selects its serial branch
      extra=""
      case "$f" in */taylorgreenvortex/testcase_analyze.f90) extra="-UMPI" ;; esac
      if [ -f "$obj/$b.o" ] || $FC $COMMON $defs $extra -module-dir "$obj" -c "$f" -o "$obj/$b.o" 2> "$obj/$b.err"; then
        n=$((n+1))
      else
        echo "$f" >> "$obj/next.txt"
      fi
    done < "$obj/todo.txt"
    cp "$obj/next.txt" "$obj/todo.txt"
    [ $n -eq 0 ] && break
  done
  $FC $COMMON $defs -module-dir "$obj" -I "$obj" -c "$HERE/fluxo_ref_harness.f90" -o "$obj/harness.o"
  $FC ${OPT:-$OPT_DEFAULT} -o "$OUT/fluxo_ref_$name" "$obj/harness.o" $(ls "$obj"/*.o | grep -v harness.o)
  echo "built $OUT/fluxo_ref_$name; reference files not compiled:"
  sed 's|^|    |' "$obj/todo.txt"
}

NS="-DPP_nVar=5 -Dnavierstokes -DPP_entropy_vars_exist -DPP_u_aux_exist"
NOSC="-DSHOCKCAPTURE=0 -DSHOCK_NFVSE=0"
build_variant euler_vf32 $NS $NOSC -DPP_VolFlux=32 -DPARABOLIC=0
build_variant euler_vf5 $NS $NOSC -DPP_VolFlux=5 -DPARABOLIC=0
# Ismail-Roe two-point flux (VolumeFlux=10)
build_variant euler_vf10 $NS $NOSC -DPP_VolFlux=10 -DPARABOLIC=0
# the other split-form volume fluxes (VolumeFlux 1-4 and 6-9, equation.f90:414-455)
for vf in 1 2 3 4 6 7 8 9; do
  build_variant euler_vf$vf $NS $NOSC -DPP_VolFlux=$vf -DPARABOLIC=0
done
build_variant ns_br1_vf10 $NS $NOSC -DPP_VolFlux=10 -DPARABOLIC=1 -DPP_Lifting=1 -DPP_Lifting_Var=3 -DPP_VISC=0
build_variant ns_br1_vf32 $NS $NOSC -DPP_VolFlux=32 -DPARABOLIC=1 -DPP_Lifting=1 -DPP_Lifting_Var=3 -DPP_VISC=0
# BR2 lifting (FLUXO_PARABOLIC_LIFTING=br2, CMakeLists.txt:278-283)
build_variant ns_br2_vf32 $NS $NOSC -DPP_VolFlux=32 -DPARABOLIC=1 -DPP_Lifting=2 -DPP_Lifting_Var=3 -DPP_VISC=0
# ideal GLM-MHD with Powell non-conservative terms (FLUXO_EQNSYSNAME=mhd, FLUXO_EQN_GLM,
# FLUXO_EQN_NONCONS=Powell; CMakeLists.txt:174-209), Derigs flux (VolumeFlux=10)
MHD="-DPP_nVar=9 -Dmhd -DPP_entropy_vars_exist -DPP_u_aux_exist -DPP_GLM=T -DNONCONS=1"
build_variant mhd_vf10 $MHD $NOSC -DPP_VolFlux=10 -DPARABOLIC=0
# resistive GLM-MHD: viscosity, resistivity and heat conduction through BR1 lifting of the
# entropy variables (FLUXO_PARABOLIC, CMakeLists.txt:253-283)
build_variant mhd_br1_vf10 $MHD $NOSC -DPP_VolFlux=10 -DPARABOLIC=1 -DPP_Lifting=1 -DPP_Lifting_Var=3
build_variant mhd_br2_vf10 $MHD $NOSC -DPP_VolFlux=10 -DPARABOLIC=1 -DPP_Lifting=2 -DPP_Lifting_Var=3
# the same with the Taylor-Green vortex testcase (FLUXO_TESTCASE=taylorgreenvortex), whose
# AnalyzeTestcase integrates kinetic and magnetic energy, enstrophy and dissipation rates
TC=taylorgreenvortex build_variant mhd_br1_vf10_tgv $MHD $NOSC -DPP_VolFlux=10 -DPARABOLIC=1 -DPP_Lifting=1 \
  -DPP_Lifting_Var=3
# NFVSE shock capturing (FLUXO_SHOCKCAPTURE + FLUXO_SHOCKCAP_NFVSE, CMakeLists.txt:316-331)
build_variant euler_nfvse_vf32 $NS -DSHOCKCAPTURE=1 -DSHOCK_NFVSE=1 -DPP_VolFlux=32 -DPARABOLIC=0
# Navier-Stokes (BR1) with NFVSE: the blended advective volume integral plus VolInt_visc (dg.f90:368-398)
build_variant ns_br1_nfvse_vf32 $NS -DSHOCKCAPTURE=1 -DSHOCK_NFVSE=1 -DPP_VolFlux=32 -DPARABOLIC=1 -DPP_Lifting=1 \
  -DPP_Lifting_Var=3 -DPP_VISC=0
# the CPU baseline: Navier-Stokes BR1 with the Ranocha flux, -O3 with AVX2/FMA (x86-64-v3, which
# the GPU hosts' CPUs support; -march=native would target this container's CPU instead)
OPT="-O3 -march=x86-64-v3" build_variant ns_br1_vf32_fast $NS $NOSC -DPP_VolFlux=32 -DPARABOLIC=1 -DPP_Lifting=1 \
  -DPP_Lifting_Var=3 -DPP_VISC=0
