!==================================================================================================
!> Harness that runs the reference's own compute modules (project-fluxo/fluxo, built from
!> /root/reference/src by oracle/ref/build_ref.sh) on a mesh and state given in a binary file.
!>
!> TEST INFRASTRUCTURE ONLY: it produces golden values for the oracle and the engine.
!> It replaces what the reference's InitMesh (src/mesh/mesh.f90:100-300) does after reading a
!> HOPR file -- side counts and ranges, ElemToSide/SideToElem, buildMappings, metric array
!> allocation and CalcMetrics -- and the TimeStepByLSERKW2 loop (src/timedisc/timedisc.f90:
!> 424-480, with CALCTIMESTEP before every step); everything else is the reference's code.
!> The reference's Analyze (src/analyze/analyze.f90:370-540: CalcErrorNorms, CalcBulk,
!> AnalyzeEquation, AnalyzeTestcase) runs once on the input state after its right-hand side;
!> the values it collects for AnalyzeToFile are written as 'name value' lines to analysis.txt
!> in the working directory.
!>
!> After the analysis it writes the face fluxes Flux_master/Flux_slave of that right-hand side
!> (FillFlux, src/dg/fillflux.f90:41-110) and, with shock capturing, VolInt_NFVSE's blended
!> volume integral (src/shockcapturing/nfvse.f90:475-561) of the same state.
!>
!> A mesh with non-conforming interfaces brings MortarType / MortarInfo in the input (the reference's
!> fillMeshInfo arrays, prepare_mesh.f90:700-718); InitMortarBase, CalcMetrics' mortar branch and
!> U_Mortar / Flux_Mortar in DGTimeDerivative are then the reference's.
!>
!> usage: fluxo_ref_harness parameter.ini input.bin output.bin
!==================================================================================================
#include "defines.h"
program fluxo_ref_harness
  use MOD_Globals
  use MOD_PreProc
  use MOD_MPI,                 only: InitMPI
  use MOD_Output,              only: DefineParametersOutput
  use MOD_ReadInTools,         only: prms
  use MOD_Interpolation,       only: DefineParametersInterpolation, InitInterpolation, GetVandermonde
  use MOD_Interpolation_Vars,  only: NodeType, NodeTypeGL, NodeTypeEQ
  use MOD_Mortar,              only: DefineParametersMortar, InitMortarBase, InitMortar
  use MOD_Equation,            only: DefineParametersEquation, InitEquation
  use MOD_Testcase,            only: DefineParametersTestcase, InitTestcase
  use MOD_GetBoundaryFlux,     only: InitBC
  use MOD_DG,                  only: InitDG, DGTimeDerivative
  use MOD_DG_Vars,             only: U, Ut, nTotalU, Flux_master, Flux_slave
  use MOD_VolInt,              only: VolInt_adv_SplitForm
  use MOD_Restart_Vars,        only: DoRestart, RestartInitIsDone
  use MOD_Mappings,            only: buildMappings
  use MOD_Metrics,             only: CalcMetrics
  use MOD_Mesh_Vars
  use MOD_TimeDisc_Vars,       only: CFLScale, CFLScaleAlpha, RKA, RKb, RKc, nRKStages, dtElem, SetTimeDiscCoefs
#if PARABOLIC
  use MOD_Lifting,             only: DefineParametersLifting, InitLifting
  use MOD_Lifting_Vars,        only: gradPx, gradPy, gradPz
  use MOD_TimeDisc_Vars,       only: DFLScale, DFLScaleAlpha, RelativeDFL
#endif
  use MOD_CalcTimeStep,        only: CALCTIMESTEP
  use MOD_Analyze,             only: DefineParametersAnalyze, InitAnalyze, Analyze
  use MOD_Analyze_Vars,        only: A2F_nVars, A2F_VarNames, A2F_Data
  use MOD_Vector,              only: VCopy, VAXPBY
#if SHOCK_NFVSE
  use MOD_ShockCapturing,      only: DefineParametersShockCapturing, InitShockCapturing
  use MOD_NFVSE_Vars,          only: alpha
  use MOD_NFVSE,               only: CalcBlendingCoefficient, VolInt_NFVSE
#endif
#ifdef PP_GLM
  use MOD_Equation_Vars,       only: GLM_ch, GLM_dtch1
#endif
  implicit none
  character(len=255) :: fpar
  character(len=512) :: fin, fout
  integer :: hdr(8), nSteps, iStep, iStage, errType, u_in, u_out, u_an, iv
  real :: rhdr(4), cflIn, dflIn, t, dt, tStage
  real, allocatable :: Utmp(:, :, :, :, :), b_dt(:)

  call get_command_argument(1, fpar)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  call InitMPI()
  call DefineParametersOutput()
  call DefineParametersInterpolation()
  call DefineParametersMortar()
  call DefineParametersEquation()
  call DefineParametersTestcase()
  call DefineParametersAnalyze()
#if PARABOLIC
  call DefineParametersLifting()
#endif
#if SHOCK_NFVSE
  call DefineParametersShockCapturing()
#endif
  call prms%read_options(fpar)

  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) hdr
  read(u_in) rhdr
  NGeo = hdr(2); nElems = hdr(3); nSides = hdr(4); nSteps = hdr(5)
  cflIn = rhdr(1); dflIn = rhdr(2); t = rhdr(3)

  call InitInterpolation()
  if (PP_N .ne. hdr(1)) stop 'N of parameter file and input differ'
  call InitMortarBase()

  ! ---- the part of InitMesh that follows readMesh: periodic mesh; hdr(7) big mortar sides come
  ! first (sides 1..hdr(7)), then the inner sides (setLocalSideIDs, prepare_mesh.f90:186-189)
  nGlobalElems = nElems; offsetElem = 0
  nMortarInnerSides = hdr(7); nMortarSides = hdr(7)
  nBCSides = 0; nInnerSides = nSides - nMortarInnerSides; nMPISides_MINE = 0; nMPISides_YOUR = 0
  nMPISides = 0; nMortarMPISides = 0; nBCs = 0; nUserBCs = 0; nAnalyzeSides = 0
  firstBCSide = 1
  firstMortarInnerSide = firstBCSide + nBCSides
  firstInnerSide = firstMortarInnerSide + nMortarInnerSides
  firstMPISide_MINE = firstInnerSide + nInnerSides
  firstMPISide_YOUR = firstMPISide_MINE + nMPISides_MINE
  firstMortarMPISide = firstMPISide_YOUR + nMPISides_YOUR
  lastBCSide = firstMortarInnerSide - 1
  lastMortarInnerSide = firstInnerSide - 1
  lastInnerSide = firstMPISide_MINE - 1
  lastMPISide_MINE = firstMPISide_YOUR - 1
  lastMPISide_YOUR = firstMortarMPISide - 1
  lastMortarMPISide = nSides
  firstSlaveSide = firstInnerSide
  lastSlaveSide = lastMPISide_YOUR
  nSidesMaster = nSides
  nSidesSlave = lastSlaveSide - firstSlaveSide + 1
  allocate(ElemToSide(2, 6, nElems), SideToElem(5, nSides), BC(1:nBCSides), AnalyzeSide(1:nSides))
  allocate(MortarType(2, 1:nSides), MortarInfo(MI_FLIP, 0:4, nMortarSides))
  allocate(BoundaryType(nBCs, 3))
  MortarType = -1; MortarInfo = -1; AnalyzeSide = 0
  allocate(NodeCoords(3, 0:NGeo, 0:NGeo, 0:NGeo, nElems))
  read(u_in) NodeCoords
  read(u_in) ElemToSide
  read(u_in) SideToElem
  if (nMortarSides > 0) then
    read(u_in) MortarType
    read(u_in) MortarInfo
  end if
  call buildMappings(PP_N, V2S=V2S, V2S2=V2S2, S2V=S2V, S2V2=S2V2, S2V3=S2V3, CS2V2=CS2V2, FS2M=FS2M)
  allocate(Vdm_GLN_N(0:PP_N, 0:PP_N), Vdm_N_GLN(0:PP_N, 0:PP_N))
  call GetVandermonde(PP_N, NodeTypeGL, PP_N, NodeType, Vdm_GLN_N, Vdm_N_GLN, modal=.false.)
  allocate(Elem_xGP(3, 0:PP_N, 0:PP_N, 0:PP_N, nElems))
  allocate(Metrics_fTilde(3, 0:PP_N, 0:PP_N, 0:PP_N, nElems), Metrics_gTilde(3, 0:PP_N, 0:PP_N, 0:PP_N, nElems))
  allocate(Metrics_hTilde(3, 0:PP_N, 0:PP_N, 0:PP_N, nElems), dXGL_N(3, 3, 0:PP_N, 0:PP_N, 0:PP_N, nElems))
  allocate(sJ(0:PP_N, 0:PP_N, 0:PP_N, nElems))
  NGeoRef = 3*NGeo
  allocate(DetJac_Ref(1, 0:NGeoRef, 0:NGeoRef, 0:NGeoRef, nElems), Elem_inCyl(nElems))
  allocate(Face_xGP(3, 0:PP_N, 0:PP_N, 1:nSides), NormVec(3, 0:PP_N, 0:PP_N, 1:nSides))
  allocate(TangVec1(3, 0:PP_N, 0:PP_N, 1:nSides), TangVec2(3, 0:PP_N, 0:PP_N, 1:nSides))
  allocate(SurfElem(0:PP_N, 0:PP_N, 1:nSides))
  crossProductMetrics = .false.
  NodeTypeMesh = NodeTypeEQ
  call CalcMetrics()
  Elem_inCyl = .false.
  MeshInitIsDone = .true.
  ! ---- end of the InitMesh part

  call InitMortar()
  call InitEquation()
  call InitBC()
  DoRestart = .true.; RestartInitIsDone = .true.   ! U comes from the input file, not FillIni
  call InitDG()
#if PARABOLIC
  call InitLifting()
#endif
  call InitTestcase()
  call InitAnalyze()
#if SHOCK_NFVSE
  call InitShockCapturing()
#endif
  read(u_in) U
  close(u_in)

  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) Metrics_fTilde, Metrics_gTilde, Metrics_hTilde, sJ, NormVec, TangVec1, TangVec2, SurfElem
  ! volume split-form contribution alone
  Ut = 0.
  call VolInt_adv_SplitForm(Ut)
  write(u_out) Ut
  ! full right-hand side
  call DGTimeDerivative(t)
  write(u_out) Ut
#if PARABOLIC
  write(u_out) gradPx, gradPy, gradPz
#endif
#if SHOCK_NFVSE
  write(u_out) alpha
#endif
  ! error norms, bulk, entropy (prints ' Entropy      :' and ' dSdU*Ut         :' on stdout;
  ! CalcEntropy is module-private) and the testcase analysis of this state
  call Analyze(t, 0_8)
  open(newunit=u_an, file='analysis.txt', status='replace')
  do iv = 5, A2F_nVars
    write(u_an, '(A,1X,ES25.17)') trim(A2F_VarNames(iv)), A2F_Data(iv)
  end do
  close(u_an)
  ! the face fluxes of FillFlux and Flux_Mortar from that right-hand side (module arrays of
  ! MOD_DG_Vars; Flux_slave spans the inner sides firstSlaveSide:lastSlaveSide)
  write(u_out) Flux_master, Flux_slave
#if SHOCK_NFVSE
  ! VolInt followed by VolInt_NFVSE alone: blending coefficient, volume split form, blend
  call CalcBlendingCoefficient(U)
  Ut = 0.
  call VolInt_adv_SplitForm(Ut)
  call VolInt_NFVSE(Ut)
  write(u_out) Ut
#endif
  ! CFL/DFL scaling of FillCFL_DFL (timedisc.f90:697-741) with the RK4-5 tables
  call SetTimeDiscCoefs('carpenterrk4-5')
  CFLScale = cflIn*CFLScaleAlpha(MIN(15, PP_N))/(2.*PP_N + 1.)
#if PARABOLIC
  DFLScale = dflIn*DFLScaleAlpha(MIN(10, PP_N))*RelativeDFL/(2.*PP_N + 1.)**2
#endif
  allocate(dtElem(nElems))
  dt = CALCTIMESTEP(errType)
  write(u_out) dt, real(errType)
#ifdef PP_GLM
  write(u_out) GLM_ch, GLM_dtch1
#endif
  ! nSteps of TimeStepByLSERKW2 with CALCTIMESTEP before each step
  allocate(Utmp, mold=U)
  allocate(b_dt(nRKStages))
  do iStep = 1, nSteps
    dt = CALCTIMESTEP(errType)
    b_dt = RKb*dt
    call DGTimeDerivative(t)
    call VCopy(nTotalU, Utmp, Ut)
    call VAXPBY(nTotalU, U, Ut, ConstIn=b_dt(1))
    do iStage = 2, nRKStages
      tStage = t + dt*RKc(iStage)
      call DGTimeDerivative(tStage)
      call VAXPBY(nTotalU, Utmp, Ut, ConstOut=-RKA(iStage))
      call VAXPBY(nTotalU, U, Utmp, ConstIn=b_dt(iStage))
    end do
    t = t + dt
  end do
  write(u_out) U, t, dt
  close(u_out)
end program fluxo_ref_harness
