!==================================================================================================
!> ISO_C_BINDING interface to the engine's C ABI (include/fluxo_gpu.h).  The derived types are
!> interoperable with fx_config / fx_mesh; every function returns 0 or a negative FX_E* code.
!==================================================================================================
module fluxo_gpu_bind
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: FX_OK = 0, FX_PREC_FP64 = 0, FX_PREC_FP32_SPLIT = 1

  type, bind(C) :: fx_config
    integer(c_int) :: N = 3, nVar = 5, nElems = 0, nSides = 0
    integer(c_int) :: volumeFlux = 0, riemann = 1, parabolic = 0, lifting = 1
    integer(c_int) :: precision = 0, whichMortar = 0
    real(c_double) :: kappa = 1.4d0, mu0 = 0.0d0, Pr = 0.72d0, etaBR2 = 2.0d0
    integer(c_int) :: nfvse = 0, computeAlpha = 1, spacePropSweeps = 1, reserved3 = 0
    real(c_double) :: alpha_max = 0.5d0, alpha_min = 0.01d0, spacePropFactor = 0.5d0, shockBlendCoef = 0.0d0
    real(c_double) :: mu_0 = 1.0d0, GLM_scale = 0.5d0, eta = 0.0d0, reserved4 = 0.0d0
  end type fx_config

  type, bind(C) :: fx_mesh
    type(c_ptr) :: ElemToSide = c_null_ptr, SideToElem = c_null_ptr
    type(c_ptr) :: Metrics_fTilde = c_null_ptr, Metrics_gTilde = c_null_ptr, Metrics_hTilde = c_null_ptr
    type(c_ptr) :: sJ = c_null_ptr, NormVec = c_null_ptr, TangVec1 = c_null_ptr, TangVec2 = c_null_ptr
    type(c_ptr) :: SurfElem = c_null_ptr
    type(c_ptr) :: MortarType = c_null_ptr, MortarInfo = c_null_ptr
    integer(c_int) :: nMortarSides = 0, reserved_m = 0
  end type fx_mesh

  type, bind(C) :: fx_analyze_config
    integer(c_int) :: exactFunc = 0, nAnalyze = 0
    real(c_double) :: refState(9) = 0.0d0, advVel(3) = 1.0d0, waveNumber(3) = 1.0d0
    real(c_double) :: amplitude = 0.1d0, R = 287.058d0, dissFactor = 1.0d0
  end type fx_analyze_config

  type, bind(C) :: fx_analysis
    real(c_double) :: vol
    real(c_double) :: L2(9), L2colloc(9), Linf(9), bulk(9), bulk_t(9), maxabs_Ut(9)
    real(c_double) :: Ekin, Ekin_comp, Emag_comp, Enstrophy_comp, Mag_Enstrophy, Entropy_comp
    real(c_double) :: DR_u, DR_S, DR_Sd, DR_p, max_Vorticity, max_Current, mean_temperature, uprime
  end type fx_analysis

  ! libfluxo_io (include/fluxo_io.h): HOPR mesh files and state files
  interface
    integer(c_int) function fxio_mesh_info(file, nElems, nSides, Ngeo, nodeTypeGL) bind(C, name="fxio_mesh_info")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: file(*)
      integer(c_int), intent(out) :: nElems, nSides, Ngeo, nodeTypeGL
    end function fxio_mesh_info
    integer(c_int) function fxio_mesh_read(file, e2s, s2e, NodeCoords) bind(C, name="fxio_mesh_read")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: file(*)
      integer(c_int), intent(out) :: e2s(*), s2e(*)
      real(c_double), intent(out) :: NodeCoords(*)
    end function fxio_mesh_read
    integer(c_int) function fxio_mesh_mortars(file, nMortarSides, MortarType, MortarInfo) &
        bind(C, name="fxio_mesh_mortars")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: file(*)
      integer(c_int), intent(out) :: nMortarSides
      type(c_ptr), value :: MortarType, MortarInfo
    end function fxio_mesh_mortars
    integer(c_int) function fxio_state_write(file, project, meshFile, N, nVar, nElems, time, varNames, U) &
        bind(C, name="fxio_state_write")
      import :: c_int, c_char, c_double, c_ptr
      character(kind=c_char), intent(in) :: file(*), project(*), meshFile(*)
      integer(c_int), value :: N, nVar, nElems
      real(c_double), value :: time
      type(c_ptr), intent(in) :: varNames(*)
      real(c_double), intent(in) :: U(*)
    end function fxio_state_write
    integer(c_int) function fxio_state_info(file, N, nVar, nElems, time) bind(C, name="fxio_state_info")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: file(*)
      integer(c_int), intent(out) :: N, nVar, nElems
      real(c_double), intent(out) :: time
    end function fxio_state_info
    integer(c_int) function fxio_state_read(file, N, nVar, nElems, U) bind(C, name="fxio_state_read")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: file(*)
      integer(c_int), value :: N, nVar, nElems
      real(c_double), intent(out) :: U(*)
    end function fxio_state_read
    type(c_ptr) function fxio_last_error() bind(C, name="fxio_last_error")
      import :: c_ptr
    end function fxio_last_error
  end interface

  interface
    integer(c_int) function fx_elem_xgp_upload(ctx, xgp) bind(C, name="fx_elem_xgp_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xgp(*)
    end function fx_elem_xgp_upload
    integer(c_int) function fx_analyze(ctx, t, cfg, res) bind(C, name="fx_analyze")
      import :: c_int, c_ptr, c_double, fx_analyze_config, fx_analysis
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      type(fx_analyze_config), intent(in) :: cfg
      type(fx_analysis), intent(out) :: res
    end function fx_analyze
    integer(c_int) function fx_ctx_create(cfg, ctx) bind(C, name="fx_ctx_create")
      import :: c_int, c_ptr, fx_config
      type(fx_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function fx_ctx_create
    subroutine fx_ctx_destroy(ctx) bind(C, name="fx_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine fx_ctx_destroy
    type(c_ptr) function fx_last_error() bind(C, name="fx_last_error")
      import :: c_ptr
    end function fx_last_error
    integer(c_int) function fx_mesh_upload(ctx, mesh) bind(C, name="fx_mesh_upload")
      import :: c_int, c_ptr, fx_mesh
      type(c_ptr), value :: ctx
      type(fx_mesh), intent(in) :: mesh
    end function fx_mesh_upload
    integer(c_int) function fx_state_upload(ctx, U) bind(C, name="fx_state_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: U(*)
    end function fx_state_upload
    integer(c_int) function fx_state_download(ctx, U) bind(C, name="fx_state_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: U(*)
    end function fx_state_download
    integer(c_int) function fx_ut_download(ctx, Ut) bind(C, name="fx_ut_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: Ut(*)
    end function fx_ut_download
    integer(c_int) function fx_rhs(ctx, t, Ut) bind(C, name="fx_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      real(c_double), intent(out) :: Ut(*)
    end function fx_rhs
    integer(c_int) function fx_volint(ctx, Ut) bind(C, name="fx_volint")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: Ut(*)
    end function fx_volint
    integer(c_int) function fx_lifting(ctx, gx, gy, gz) bind(C, name="fx_lifting")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: gx(*), gy(*), gz(*)
    end function fx_lifting
    integer(c_int) function fx_blending(ctx, alpha) bind(C, name="fx_blending")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: alpha(*)
    end function fx_blending
    integer(c_int) function fx_calc_dt(ctx, cfl, dfl, dt, errType) bind(C, name="fx_calc_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl, dfl
      real(c_double), intent(out) :: dt
      integer(c_int), intent(out) :: errType
    end function fx_calc_dt
    integer(c_int) function fx_rk_step(ctx, t, dt) bind(C, name="fx_rk_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t, dt
    end function fx_rk_step
    integer(c_int) function fx_run(ctx, nSteps, cfl, dfl, t, dt, errType) bind(C, name="fx_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nSteps
      real(c_double), value :: cfl, dfl
      real(c_double), intent(inout) :: t
      real(c_double), intent(out) :: dt
      integer(c_int), intent(out) :: errType
    end function fx_run
    integer(c_int) function fx_run_to(ctx, maxSteps, cfl, dfl, tEnd, t, dt, steps, errType) bind(C, name="fx_run_to")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: maxSteps
      real(c_double), value :: cfl, dfl, tEnd
      real(c_double), intent(inout) :: t
      real(c_double), intent(out) :: dt
      integer(c_int), intent(out) :: steps, errType
    end function fx_run_to
    integer(c_int) function fx_fillflux(ctx, Flux_master, Flux_slave) bind(C, name="fx_fillflux")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: Flux_master(*), Flux_slave(*)
    end function fx_fillflux
    integer(c_int) function fx_surfint(ctx, Flux_master, Flux_slave, Ut) bind(C, name="fx_surfint")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: Flux_master(*), Flux_slave(*)
      real(c_double), intent(inout) :: Ut(*)
    end function fx_surfint
    integer(c_int) function fx_volint_nfvse(ctx, Ut) bind(C, name="fx_volint_nfvse")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: Ut(*)
    end function fx_volint_nfvse
    integer(c_int) function fx_entropy(ctx, S, dSdU_Ut) bind(C, name="fx_entropy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: S, dSdU_Ut
    end function fx_entropy
    integer(c_int) function fx_glm(ctx, glm_ch, glm_dtch1) bind(C, name="fx_glm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: glm_ch, glm_dtch1
    end function fx_glm
    integer(c_int) function fx_synchronize(ctx) bind(C, name="fx_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fx_synchronize
    integer(c_long_long) function fx_ctx_bytes(cfg) bind(C, name="fx_ctx_bytes")
      import :: c_long_long, fx_config
      type(fx_config), intent(in) :: cfg
    end function fx_ctx_bytes
  end interface

contains

  !> Text of the library's last error as a Fortran string.
  function fx_error_text() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(fx_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function fx_error_text

  function fxio_error_text() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(fxio_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function fxio_error_text

end module fluxo_gpu_bind
