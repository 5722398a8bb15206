!==================================================================================================
!> FLUXO-shaped host interface over the MI355X engine.
!>
!> The routine names and argument meanings are those of the reference so that a FLUXO-style
!> driver can call them unchanged:
!>   DGTimeDerivative_weakForm(t)    src/dg/dg.f90:282-452      (drop-in: U up, Ut down)
!>   VolInt_SplitForm(Ut)            src/dg/volint.f90:173-218
!>   Lifting(t)                      src/dg/br1/br1.f90:154-364
!>   CALCTIMESTEP(errType)           src/equation/navierstokes/calctimestep.f90:39-170
!>   TimeStepByLSERKW2(t)            src/timedisc/timedisc.f90:424-480 (device-resident U)
!>   CalcEntropy(S,dSdU_Ut)          src/equation/navierstokes/analyze_equation.f90:549-599
!>   ExactFunc(2|7)                  src/equation/navierstokes/equation.f90:632-651, 802-811
!>   CalcErrorNorms(t,L2,L2c,Linf)   src/analyze/analyze.f90:557-628 (on the device, fx_analyze)
!>   FillFlux(Fm,Fs,doMPISides)      src/dg/fillflux.f90:41-110
!>   SurfInt(Fm,Fs,Ut,doMPISides)    src/dg/surfint.f90:48-264
!>   VolInt_NFVSE(Ut)                src/shockcapturing/nfvse.f90:475-561
!>   TimeDiscRunTo(...)              src/timedisc/timedisc.f90:274-413 (end time as the reference)
!> Module variables play the role of MOD_DG_Vars / MOD_Mesh_Vars / MOD_TimeDisc_Vars.  PP_nVar is
!> 5 (Euler / Navier-Stokes) or 9 (GLM-MHD, ExactFunc of src/equation/mhd/equation.f90).
!==================================================================================================
module fxh_dg
  use, intrinsic :: iso_c_binding
  use fxh_basis, only: rk
  use fluxo_gpu_bind
  implicit none
  private
  public :: InitDG_GPU, FinalizeDG_GPU, DGTimeDerivative_weakForm, VolInt_SplitForm, Lifting
  public :: CALCTIMESTEP, TimeStepByLSERKW2, TimeDiscRun, TimeDiscRunTo, CalcEntropy, ExactFunc, FillIni
  public :: StateToHost, StateToDevice, CalcErrorNorms, FillFlux, SurfInt, VolInt_NFVSE, SetRefState
  public :: U, Ut, gradPx, gradPy, gradPz, Elem_xGP, sJ, PP_N, PP_nVar, nElems, nSides, dt, CFLScale, DFLScale
  public :: Kappa, AdvVel, IniAmplitude, IniExactFunc, cfg, RefStateCons, mu_0

  integer :: PP_N = 3, PP_nVar = 5, nElems = 0, nSides = 0, IniExactFunc = 2
  real(rk), allocatable, target :: U(:, :, :, :, :), Ut(:, :, :, :, :)
  real(rk), allocatable, target :: gradPx(:, :, :, :, :), gradPy(:, :, :, :, :), gradPz(:, :, :, :, :)
  real(rk), allocatable, target :: Elem_xGP(:, :, :, :, :), sJ(:, :, :, :)
  real(rk) :: dt = 0.0_rk, CFLScale = 0.9_rk, DFLScale = 0.9_rk
  real(rk) :: Kappa = 1.4_rk, AdvVel(3) = 1.0_rk, IniAmplitude = 0.1_rk
  real(rk) :: RefStateCons(9) = [1.0_rk, 0.0_rk, 0.0_rk, 0.0_rk, 2.5_rk, 0.0_rk, 0.0_rk, 0.0_rk, 0.0_rk]  ! ExactFunc 1
  real(rk) :: mu_0 = 1.0_rk   ! GLM-MHD magnetic permeability
  type(fx_config) :: cfg
  type(c_ptr) :: ctx = c_null_ptr

contains

  subroutine chk(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= FX_OK) then
      write(*, '(A,I0,A)') ' ERROR in '//where//' (code ', rc, '): '//trim(fx_error_text())
      stop 2
    end if
  end subroutine chk

  !> InitDG + mesh upload: cfgIn carries the parameter.ini options, the mesh arrays are in
  !> the reference layout.
  subroutine InitDG_GPU(cfgIn, e2s, s2e, Mf, Mg, Mh, sJin, NormVec, TangVec1, TangVec2, SurfElem, xGP, &
                        nMortarSides, MortarType, MortarInfo)
    type(fx_config), intent(in) :: cfgIn
    integer(c_int), intent(in), target :: e2s(:, :, :), s2e(:, :)
    ! non-conforming interfaces (fxio_mesh_mortars): the first nMortarSides sides are big mortar sides
    integer, intent(in), optional :: nMortarSides
    integer(c_int), intent(in), target, optional :: MortarType(:, :), MortarInfo(:, :, :)
    real(rk), intent(in), target :: Mf(:, :, :, :, :), Mg(:, :, :, :, :), Mh(:, :, :, :, :)
    real(rk), intent(in), target :: sJin(:, :, :, :), NormVec(:, :, :, :), TangVec1(:, :, :, :)
    real(rk), intent(in), target :: TangVec2(:, :, :, :), SurfElem(:, :, :), xGP(:, :, :, :, :)
    type(fx_mesh) :: m
    cfg = cfgIn
    PP_N = cfg%N; PP_nVar = cfg%nVar; nElems = cfg%nElems; nSides = cfg%nSides; Kappa = cfg%kappa
    mu_0 = cfg%mu_0
    call chk(fx_ctx_create(cfg, ctx), 'fx_ctx_create')
    m%ElemToSide = c_loc(e2s); m%SideToElem = c_loc(s2e)
    m%Metrics_fTilde = c_loc(Mf); m%Metrics_gTilde = c_loc(Mg); m%Metrics_hTilde = c_loc(Mh)
    m%sJ = c_loc(sJin); m%NormVec = c_loc(NormVec); m%TangVec1 = c_loc(TangVec1)
    m%TangVec2 = c_loc(TangVec2); m%SurfElem = c_loc(SurfElem)
    if (present(nMortarSides)) then
      if (nMortarSides > 0) then
        m%nMortarSides = nMortarSides
        m%MortarType = c_loc(MortarType); m%MortarInfo = c_loc(MortarInfo)
      end if
    end if
    call chk(fx_mesh_upload(ctx, m), 'fx_mesh_upload')
    allocate(U(PP_nVar, 0:PP_N, 0:PP_N, 0:PP_N, nElems), Ut(PP_nVar, 0:PP_N, 0:PP_N, 0:PP_N, nElems))
    allocate(Elem_xGP(3, 0:PP_N, 0:PP_N, 0:PP_N, nElems), sJ(0:PP_N, 0:PP_N, 0:PP_N, nElems))
    Elem_xGP = xGP; sJ = sJin
    call chk(fx_elem_xgp_upload(ctx, Elem_xGP), 'fx_elem_xgp_upload')
    U = 0.0_rk; Ut = 0.0_rk
  end subroutine InitDG_GPU

  subroutine FinalizeDG_GPU()
    if (c_associated(ctx)) call fx_ctx_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(U)) deallocate(U, Ut, Elem_xGP, sJ)
    if (allocated(gradPx)) deallocate(gradPx, gradPy, gradPz)
  end subroutine FinalizeDG_GPU

  subroutine StateToDevice()
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
  end subroutine StateToDevice

  subroutine StateToHost()
    call chk(fx_state_download(ctx, U), 'fx_state_download')
  end subroutine StateToHost

  !> Drop-in form: U is taken from the host array, Ut is returned in the host array.
  subroutine DGTimeDerivative_weakForm(tIn)
    real(rk), intent(in) :: tIn
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
    call chk(fx_rhs(ctx, tIn, Ut), 'fx_rhs')
  end subroutine DGTimeDerivative_weakForm

  !> Volume split-form contribution alone (before -sJ), on the current host U.
  subroutine VolInt_SplitForm(UtOut)
    real(rk), intent(out) :: UtOut(:, :, :, :, :)
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
    call chk(fx_volint(ctx, UtOut), 'fx_volint')
  end subroutine VolInt_SplitForm

  !> BR1 lifting of the current host U into gradPx/y/z (primitive-variable gradients).
  subroutine Lifting(tIn)
    real(rk), intent(in) :: tIn
    if (.not. allocated(gradPx)) then
      allocate(gradPx, mold=U); allocate(gradPy, mold=U); allocate(gradPz, mold=U)
    end if
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
    call chk(fx_lifting(ctx, gradPx, gradPy, gradPz), 'fx_lifting')
  end subroutine Lifting

  !> CFL/DFL time step of the device state, CFLScale/DFLScale as read from parameter.ini.
  real(rk) function CALCTIMESTEP(errType)
    integer, intent(out) :: errType
    integer(c_int) :: e
    real(c_double) :: d
    call chk(fx_calc_dt(ctx, CFLScale, DFLScale, d, e), 'fx_calc_dt')
    errType = e
    CALCTIMESTEP = d
  end function CALCTIMESTEP

  !> One Carpenter RK4-5 step of the device-resident state with the module's dt.
  subroutine TimeStepByLSERKW2(t)
    real(rk), intent(in) :: t
    call chk(fx_rk_step(ctx, t, dt), 'fx_rk_step')
  end subroutine TimeStepByLSERKW2

  !> nSteps of the TimeDisc loop (CALCTIMESTEP + TimeStepByLSERKW2) on the device.
  subroutine TimeDiscRun(nSteps, t, errType)
    integer, intent(in) :: nSteps
    real(rk), intent(inout) :: t
    integer, intent(out) :: errType
    integer(c_int) :: e
    real(c_double) :: tt, d
    tt = t
    call chk(fx_run(ctx, nSteps, CFLScale, DFLScale, tt, d, e), 'fx_run')
    t = tt; dt = d; errType = e
  end subroutine TimeDiscRun

  !> The TimeDisc loop up to tEnd on the device: CFL steps, the step that reaches tEnd within 1e-4
  !> of dt shortened to end there (timedisc.f90:313-317), at most maxIter steps (maxIter < 0: no
  !> limit).  t is the start time on entry, the final time on return; nIter the steps taken.
  subroutine TimeDiscRunTo(maxIter, tEnd, t, nIter, errType)
    integer, intent(in) :: maxIter
    real(rk), intent(in) :: tEnd
    real(rk), intent(inout) :: t
    integer, intent(out) :: nIter, errType
    integer(c_int) :: e, n
    real(c_double) :: tt, d
    tt = t
    call chk(fx_run_to(ctx, maxIter, CFLScale, DFLScale, tEnd, tt, d, n, e), 'fx_run_to')
    t = tt; dt = d; nIter = n; errType = e
  end subroutine TimeDiscRunTo

  !> FillFlux of the current host U: Flux_master / Flux_slave (PP_nVar,0:N,0:N,nSides).  All
  !> sides are inner sides, so the MPI pass (doMPISides = .TRUE.) has nothing to fill.
  subroutine FillFlux(Flux_master, Flux_slave, doMPISides)
    real(rk), intent(out) :: Flux_master(:, :, :, :), Flux_slave(:, :, :, :)
    logical, intent(in) :: doMPISides
    if (doMPISides) return
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
    call chk(fx_fillflux(ctx, Flux_master, Flux_slave), 'fx_fillflux')
  end subroutine FillFlux

  !> SurfInt: Ut += the surface integral of Flux_master / Flux_slave (no MPI sides here).
  subroutine SurfInt(Flux_master, Flux_slave, UtInOut, doMPISides)
    real(rk), intent(in) :: Flux_master(:, :, :, :), Flux_slave(:, :, :, :)
    real(rk), intent(inout) :: UtInOut(:, :, :, :, :)
    logical, intent(in) :: doMPISides
    if (doMPISides) return
    call chk(fx_surfint(ctx, Flux_master, Flux_slave, UtInOut), 'fx_surfint')
  end subroutine SurfInt

  !> VolInt_NFVSE of the current host U: on entry Ut holds VolInt_SplitForm of U (the call order
  !> of DGTimeDerivative_weakForm), on return the blended volume integral (1-alpha) VolInt +
  !> alpha F_FV with alpha from CalcBlendingCoefficient; the engine forms both on the device.
  subroutine VolInt_NFVSE(UtInOut)
    real(rk), intent(inout) :: UtInOut(:, :, :, :, :)
    call chk(fx_state_upload(ctx, U), 'fx_state_upload')
    call chk(fx_volint_nfvse(ctx, UtInOut), 'fx_volint_nfvse')
  end subroutine VolInt_NFVSE

  !> Entropy integral and dS/dU.Ut of the device state (Ut of the last RHS evaluation).
  subroutine CalcEntropy(S, dSdU_Ut)
    real(rk), intent(out) :: S, dSdU_Ut
    real(c_double) :: a, b
    call chk(fx_rhs(ctx, 0.0_rk, Ut), 'fx_rhs')
    call chk(fx_entropy(ctx, a, b), 'fx_entropy')
    S = a; dSdU_Ut = b
  end subroutine CalcEntropy

  !> RefState (primitive: rho, v, p, and B for GLM-MHD) as ExactFunc 1's conservative state
  !> (equation.f90:236-250; mhd/equation.f90:233-241 with psi = 0).
  subroutine SetRefState(prim)
    real(rk), intent(in) :: prim(:)
    real(rk) :: p(9)
    p = 0.0_rk
    p(1:size(prim)) = prim
    if (PP_nVar == 9) then
      call PrimToConsMHD(p, RefStateCons)
    else
      RefStateCons = 0.0_rk
      RefStateCons(1) = p(1); RefStateCons(2:4) = p(1)*p(2:4)
      RefStateCons(5) = p(5)/(Kappa - 1.0_rk) + 0.5_rk*p(1)*sum(p(2:4)**2)
    end if
  end subroutine SetRefState

  !> PrimToCons of GLM-MHD (mhd/equation_vars.f90:258-278)
  subroutine PrimToConsMHD(prim, cons)
    real(rk), intent(in) :: prim(9)
    real(rk), intent(out) :: cons(9)
    real(rk) :: sKappaM1, s2mu_0
    sKappaM1 = 1.0_rk/(Kappa - 1.0_rk); s2mu_0 = 0.5_rk/mu_0
    cons(1) = prim(1)
    cons(2:4) = prim(2:4)*prim(1)
    cons(5) = sKappaM1*prim(5) + 0.5_rk*sum(cons(2:4)*prim(2:4)) + s2mu_0*(sum(prim(6:9)*prim(6:9)))
    cons(6:9) = prim(6:9)
  end subroutine PrimToConsMHD

  !> ExactFunc: Euler / Navier-Stokes 1 (RefState), 2 (density wave), 7 (Taylor-Green vortex)
  !> (navierstokes/equation.f90:632-651, 802-811); GLM-MHD 1 (RefState) and 312 (3D Orszag-Tang
  !> vortex, mhd/equation.f90:1257-1266).
  subroutine ExactFunc(ExactFunction, tIn, x, Resu)
    integer, intent(in) :: ExactFunction
    real(rk), intent(in) :: tIn, x(3)
    real(rk), intent(out) :: Resu(:)
    real(rk) :: prim(9), cent(3), pi, A, Ms, skm1
    pi = acos(-1.0_rk)
    skm1 = 1.0_rk/(Kappa - 1.0_rk)
    if (PP_nVar == 9) then
      select case (ExactFunction)
      case (1)
        Resu = RefStateCons(1:9)
      case (312)
        prim = 0.0_rk
        prim(1) = 1.0_rk
        prim(2) = -(1.0_rk + 0.2_rk*sin(2.0_rk*pi*x(3)))*sin(2.0_rk*pi*x(2))
        prim(3) = (1.0_rk + 0.2_rk*sin(2.0_rk*pi*x(3)))*sin(2.0_rk*pi*x(1))
        prim(4) = 0.2_rk*sin(2.0_rk*pi*x(3))
        prim(5) = 1.0_rk/Kappa
        prim(6) = -sin(2.0_rk*pi*x(2))/Kappa
        prim(7) = sin(4.0_rk*pi*x(1))/Kappa
        call PrimToConsMHD(prim, Resu)
      case default
        write(*, '(A,I0,A)') ' ERROR: GLM-MHD ExactFunc ', ExactFunction, ' not supported (1, 312)'
        stop 2
      end select
      return
    end if
    select case (ExactFunction)
    case (1)  ! constant RefState
      Resu = RefStateCons(1:5)
    case (7)  ! Taylor-Green vortex
      A = 1.0_rk; Ms = 0.1_rk
      prim(1) = 1.0_rk
      prim(2) = A*sin(x(1))*cos(x(2))*cos(x(3))
      prim(3) = -A*cos(x(1))*sin(x(2))*cos(x(3))
      prim(4) = 0.0_rk
      prim(5) = (A/Ms*A/Ms/Kappa*prim(1))
      prim(5) = prim(5) + 1.0_rk/16.0_rk*A*A*prim(1)*(cos(2*x(1))*cos(2.0_rk*x(3)) + 2.0_rk*cos(2.0_rk*x(2)) &
                + 2.0_rk*cos(2.0_rk*x(1)) + cos(2*x(2))*cos(2.0_rk*x(3)))
      Resu(1) = prim(1)
      Resu(2:4) = prim(2:4)*prim(1)
      Resu(5) = skm1*prim(5) + 0.5_rk*sum(Resu(2:4)*prim(2:4))
    case default  ! 2: density wave
      prim(1) = 1.0_rk; prim(2:4) = AdvVel; prim(5) = 1.0_rk
      cent = x - AdvVel*tIn
      prim(1) = prim(1)*(1.0_rk + IniAmplitude*sin(pi*(1.0_rk*cent(1) + 1.0_rk*cent(2) + 1.0_rk*cent(3))))
      Resu(1) = prim(1)
      Resu(2:4) = prim(1)*prim(2:4)
      Resu(5) = prim(5)*skm1 + 0.5_rk*prim(1)*sum(prim(2:4)*prim(2:4))
    end select
  end subroutine ExactFunc

  subroutine FillIni()
    integer :: i, j, k, e
    do e = 1, nElems
      do k = 0, PP_N
        do j = 0, PP_N
          do i = 0, PP_N
            call ExactFunc(IniExactFunc, 0.0_rk, Elem_xGP(:, i, j, k, e), U(:, i, j, k, e))
          end do
        end do
      end do
    end do
  end subroutine FillIni

  !> CalcErrorNorms of the device state against ExactFunc IniExactFunc (1 the constant
  !> RefState, 2 the density wave) at time t: L2 on NAnalyze = 2(N+1) Gauss-Lobatto points, L2
  !> on the collocation points, Linf; computed on the device (fx_analyze), only the norms come
  !> back.
  subroutine CalcErrorNorms(t, L2, L2colloc, Linf)
    real(rk), intent(in) :: t
    real(rk), intent(out) :: L2(:), L2colloc(:), Linf(:)
    type(fx_analyze_config) :: ac
    type(fx_analysis) :: res
    ac%exactFunc = IniExactFunc
    ac%refState = RefStateCons
    ac%advVel = AdvVel
    ac%amplitude = IniAmplitude
    call chk(fx_analyze(ctx, t, ac, res), 'fx_analyze')
    L2 = res%L2(1:PP_nVar); L2colloc = res%L2colloc(1:PP_nVar); Linf = res%Linf(1:PP_nVar)
  end subroutine CalcErrorNorms

end module fxh_dg
