!==================================================================================================
!> fluxo_gpu: parameter.ini-driven time loop on the MI355X engine (the reference's PROGRAM Fluxo,
!> src/fluxo.f90, restricted to periodic box meshes, with TimeDisc timedisc.f90:144-415).
!>
!> Keys (reference meaning unless noted): N, Riemann, VolumeFlux, IniExactFunc (1, 2, 7; GLM-MHD
!> 1, 312), RefState, AdvVel, IniAmplitude, kappa, mu0, Pr, CFLscale, DFLscale, TEnd (required),
!> maxIter (default -1: no limit), Analyze_dt.  The run stops at TEnd as the reference's TimeDisc
!> does (timedisc.f90:313-317): the step that reaches TEnd is shortened to end there.
!> EquationSystem (this driver's run-time form of the reference's compile-time FLUXO_EQNSYSNAME):
!> navierstokes (default) or mhd; GLM-MHD keys of mhd/equation.f90:174-208: mu_0, GLM_scale, eta,
!> mu (viscosity), Pr, kappa (default 5/3); GLM_scr other than 0 is rejected (not built).
!> ShockCapturing (logical, the run-time form of FLUXO_SHOCKCAPTURE + FLUXO_SHOCKCAP_NFVSE) with
!> the keys of InitNFVSE (nfvse.f90:63-146): ComputeAlpha, ShockBlendCoef, alpha_max, alpha_min,
!> SpacePropFactor, SpacePropSweeps; SubFVMethod (1), ModalThreshold (1), TimeRelFactor (0) and
!> ShockIndicatorQuantity (DensityTimesPressure) accept only the built values.
!> Mesh keys (this driver's periodic box generator replaces MeshFile): BoxElems (3 ints),
!> BoxCorner, BoxLength (3 reals each), Ngeo, meshdeform (logical), meshdeformAmp.
!> Parabolic (logical) selects Navier-Stokes at run time, with BR1 lifting or, with Lifting = 2
!> (the reference's compile-time PP_Lifting), BR2 with penalty etaBR2 (br2.f90:116, default 2).
!> DeviceResident (logical, default T): T keeps U on the device and runs whole steps there;
!> F calls DGTimeDerivative_weakForm once per stage from the host RK loop (drop-in form).
!> MeshFile (optional): a HOPR mesh file (periodic boundaries; libfluxo_io) instead of the box
!> generator; non-conforming meshes (mortars; Euler, Navier-Stokes with BR1) with whichMortar (0
!> projection, 1 collocation).  RestartFile (optional): initial state from a state file (Restart,
!> restart.f90:147-232; same number of variables and elements) instead of IniExactFunc, its Time
!> as the start time; a file of another N is interpolated (lower N) or projected conservatively
!> through J U (higher N) with the reference's modal Vandermonde matrices.  ProjectName (optional): WriteState (hdf5_output.f90:74-165) of the final state to
!> <ProjectName>_State_<time>.h5.  StateFile (optional): the final U(nVar,0:N,0:N,0:N,nElems) as
!> raw FP64 in Fortran order.
!==================================================================================================
program fluxo_gpu
  use, intrinsic :: iso_c_binding
  use fxh_basis
  use fxh_mesh
  use fxh_readini
  use fluxo_gpu_bind, only: fx_config, fxio_mesh_info, fxio_mesh_read, fxio_mesh_mortars, fxio_state_write, &
                            fxio_state_info, &
                            fxio_state_read, fxio_error_text
  use fxh_dg
  implicit none
  character(len=512) :: fname
  character(len=256) :: sfile, meshFile, project
  character(len=32) :: tstamp, eqsys, str
  character(len=40), target :: vnames(9)
  type(c_ptr) :: vptr(9)
  integer :: u_out, nodeGL, rN, rVar, rE, iv, nvar
  type(fx_config) :: c
  integer :: KB(3), Ngeo, E, S, ierr, errType, iter, maxIter, n
  real(rk) :: x0(3), L(3), amp, t, tEnd, Analyze_dt, tAnalyze, Ssum, dSdU, L2(9), L2c(9), Linf(9)
  real(rk) :: xg(0:15), wg(0:15), t0, t1, rate, A(5), b(5), RKc(5), prim(8)
  logical :: finished, lastStep
  integer(c_int), allocatable :: e2s(:, :, :), s2e(:, :)
  integer(c_int), allocatable, target :: MType(:, :), MInfo(:, :, :)
  integer(c_int) :: nMortar = 0
  real(rk), allocatable :: NC(:, :, :, :, :), Mf(:, :, :, :, :), Mg(:, :, :, :, :), Mh(:, :, :, :, :)
  real(rk), allocatable :: sJm(:, :, :, :), nv(:, :, :, :), t1v(:, :, :, :), t2v(:, :, :, :), se(:, :, :)
  real(rk), allocatable :: xGP(:, :, :, :, :), Uttmp(:, :, :, :, :), DetJacRef(:, :, :, :, :)
  logical :: warp, resident
  integer(8) :: c0, c1, cr

  if (command_argument_count() < 1) then
    write(*, '(A)') ' usage: fluxo_gpu parameter.ini'
    stop 2
  end if
  call get_command_argument(1, fname)
  call read_ini(trim(fname))

  c%N = GETINT('N')
  eqsys = GETSTR('EquationSystem', 'navierstokes')
  select case (trim(eqsys))
  case ('navierstokes')
    c%nVar = 5
  case ('mhd')
    c%nVar = 9
  case default
    write(*, '(A)') ' ERROR: EquationSystem must be navierstokes or mhd, not '//trim(eqsys)
    stop 2
  end select
  nvar = c%nVar
  c%riemann = GETINT('Riemann', '1')
  c%volumeFlux = GETINT('VolumeFlux', merge('10', '0 ', nvar == 9))
  c%parabolic = merge(1, 0, GETLOGICAL('Parabolic', 'F'))
  c%Pr = GETREAL('Pr', '0.72')
  c%lifting = GETINT('Lifting', '1')
  c%etaBR2 = GETREAL('etaBR2', '2.')
  if (nvar == 9) then  ! InitEquation of mhd/equation.f90:174-208
    c%kappa = GETREAL('kappa', '1.6666666666666667')
    c%mu_0 = GETREAL('mu_0', '1.')
    c%GLM_scale = GETREAL('GLM_scale', '0.5')
    c%eta = GETREAL('eta', '0.')
    c%mu0 = GETREAL('mu', '0.')
    if (GETREAL('GLM_scr', '0.') /= 0.0_rk) then
      write(*, '(A)') ' ERROR: GLM_scr (GLM damping source) is not built; only GLM_scr = 0'
      stop 2
    end if
  else
    c%kappa = GETREAL('kappa', '1.4')
    c%mu0 = GETREAL('mu0', '0.')
  end if
  if (GETLOGICAL('ShockCapturing', 'F')) then  ! InitNFVSE (nfvse.f90:63-146)
    c%nfvse = 1
    c%computeAlpha = GETINT('ComputeAlpha', '1')
    c%shockBlendCoef = GETREAL('ShockBlendCoef', '0.0')
    c%alpha_max = GETREAL('alpha_max', '0.5')
    c%alpha_min = GETREAL('alpha_min', '0.01')
    c%spacePropFactor = GETREAL('SpacePropFactor', '0.5')
    c%spacePropSweeps = GETINT('SpacePropSweeps', '1')
    if (GETINT('SubFVMethod', '1') /= 1 .or. GETINT('ModalThreshold', '1') /= 1 .or. &
        GETREAL('TimeRelFactor', '0.0') /= 0.0_rk) then
      write(*, '(A)') ' ERROR: shock capturing is built for SubFVMethod=1, ModalThreshold=1, TimeRelFactor=0'
      stop 2
    end if
    str = GETSTR('ShockIndicatorQuantity', 'DensityTimesPressure')
    if (trim(str) /= 'DensityTimesPressure') then
      write(*, '(A)') ' ERROR: ShockIndicatorQuantity '//trim(str)//' not built (DensityTimesPressure)'
      stop 2
    end if
  end if
  Kappa = c%kappa
  mu_0 = c%mu_0
  PP_nVar = nvar
  IniExactFunc = GETINT('IniExactFunc')
  AdvVel = GETREALARRAY('AdvVel', 3, '1.,1.,1.')
  IniAmplitude = GETREAL('IniAmplitude', '0.1')
  if (has_key('RefState')) then  ! primitive -> RefStateCons (equation.f90:236-250)
    prim = 0.0_rk
    prim(1:merge(8, 5, nvar == 9)) = GETREALARRAY('RefState', merge(8, 5, nvar == 9))
    call SetRefState(prim(1:merge(8, 5, nvar == 9)))
  end if
  CFLScale = GETREAL('CFLscale')
  DFLScale = GETREAL('DFLscale', '0.9')
  tEnd = GETREAL('TEnd')
  maxIter = GETINT('maxIter', '-1')
  Analyze_dt = GETREAL('Analyze_dt', '1.0e30')
  resident = GETLOGICAL('DeviceResident', 'T')
  n = c%N
  nodeGL = 0
  if (has_key('MeshFile')) then
    ! ReadMesh / setLocalSideIDs / fillMeshInfo of a HOPR file (libfluxo_io)
    meshFile = GETSTR('MeshFile')
    call io_chk(fxio_mesh_info(trim(meshFile)//c_null_char, E, S, Ngeo, nodeGL), 'fxio_mesh_info')
    allocate(e2s(2, 6, E), s2e(5, S), NC(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, E))
    call io_chk(fxio_mesh_read(trim(meshFile)//c_null_char, e2s, s2e, NC), 'fxio_mesh_read')
    ! non-conforming interfaces (mortars) and 'whichMortar' (mortar.f90:88-104)
    call io_chk(fxio_mesh_mortars(trim(meshFile)//c_null_char, nMortar, c_null_ptr, c_null_ptr), &
                'fxio_mesh_mortars')
    allocate(MType(2, S), MInfo(2, 0:4, max(1, nMortar)))
    call io_chk(fxio_mesh_mortars(trim(meshFile)//c_null_char, nMortar, c_loc(MType), c_loc(MInfo)), &
                'fxio_mesh_mortars')
    c%whichMortar = GETINT('whichMortar', '0')
  else
    meshFile = 'box'
    KB = GETINTARRAY('BoxElems', 3)
    x0 = GETREALARRAY('BoxCorner', 3, '0.,0.,0.')
    L = GETREALARRAY('BoxLength', 3, '2.,2.,2.')
    Ngeo = GETINT('Ngeo', '1')
    warp = GETLOGICAL('meshdeform', 'F')
    amp = GETREAL('meshdeformAmp', '0.1')
    E = KB(1)*KB(2)*KB(3); S = 3*E
    allocate(e2s(2, 6, E), s2e(5, S), NC(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, E))
    call box_connectivity(KB, e2s, s2e)
    call box_nodecoords(KB, x0, L, Ngeo, warp, amp, NC)
  end if
  c%nElems = E; c%nSides = S
  allocate(Mf(3, 0:n, 0:n, 0:n, E), Mg(3, 0:n, 0:n, 0:n, E), Mh(3, 0:n, 0:n, 0:n, E), xGP(3, 0:n, 0:n, 0:n, E))
  allocate(sJm(0:n, 0:n, 0:n, E), nv(3, 0:n, 0:n, S), t1v(3, 0:n, 0:n, S), t2v(3, 0:n, 0:n, S), se(0:n, 0:n, S))
  allocate(DetJacRef(1, 0:3*Ngeo, 0:3*Ngeo, 0:3*Ngeo, E))
  if (nMortar > 0) then
    call calc_metrics(n, Ngeo, E, S, NC, e2s, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP, ierr, nodeGL /= 0, DetJacRef, &
                      MType, MInfo)
  else
    call calc_metrics(n, Ngeo, E, S, NC, e2s, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP, ierr, nodeGL /= 0, DetJacRef)
  end if
  if (ierr /= 0) then
    write(*, '(A)') ' ERROR: negative or badly scaled Jacobian in the generated mesh'
    stop 2
  end if
  if (nMortar > 0) then
    call InitDG_GPU(c, e2s, s2e, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP, nMortar, MType, MInfo)
  else
    call InitDG_GPU(c, e2s, s2e, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP)
  end if
  call lgl_nodes(n, xg(0:n), wg(0:n))
  t = 0.0_rk
  if (has_key('RestartFile')) then
    sfile = GETSTR('RestartFile')
    call io_chk(fxio_state_info(trim(sfile)//c_null_char, rN, rVar, rE, t), 'fxio_state_info')
    if (rVar /= nvar .or. rE /= E) then
      write(*, '(A)') ' ERROR: restart file nVar / nElems differ from the run'
      stop 2
    end if
    if (rN == n) then
      call io_chk(fxio_state_read(trim(sfile)//c_null_char, n, nvar, E, U), 'fxio_state_read')
    else
      call restart_change_n(rN)
    end if
    write(*, '(A,A,A,ES14.6)') ' restart from ', trim(sfile), ' at t=', t
  else
    call FillIni()
  end if
  call StateToDevice()
  write(*, '(A,I0,A,I0,A,I0,A,I0)') ' fluxo_gpu: nElems=', E, ' N=', n, ' VolumeFlux=', c%volumeFlux, &
      ' Riemann=', c%riemann
  if (nvar == 5) then
    call CalcEntropy(Ssum, dSdU)
    write(*, '(A,ES22.14,A,ES22.14)') ' t=0 Entropy=', Ssum, ' dSdU*Ut=', dSdU
  end if
  tAnalyze = Analyze_dt
  iter = 0
  call system_clock(c0, cr)
  if (resident) then
    ! the TimeDisc loop on the device up to TEnd (or maxIter steps)
    call TimeDiscRunTo(maxIter, tEnd, t, iter, errType)
    if (errType /= 0) then
      write(*, '(A,I0)') ' ERROR: CALCTIMESTEP errType=', errType
      stop 3
    end if
  else
    ! host RK loop calling the drop-in RHS once per stage (timedisc.f90:424-480)
    A = [0.0_rk, 567301805773.0_rk/1357537059087.0_rk, 2404267990393.0_rk/2016746695238.0_rk, &
         3550918686646.0_rk/2091501179385.0_rk, 1275806237668.0_rk/842570457699.0_rk]
    b = [1432997174477.0_rk/9575080441755.0_rk, 5161836677717.0_rk/13612068292357.0_rk, &
         1720146321549.0_rk/2090206949498.0_rk, 3134564353537.0_rk/4481467310338.0_rk, &
         2277821191437.0_rk/14882151754819.0_rk]
    RKc = [0.0_rk, 1432997174477.0_rk/9575080441755.0_rk, 2526269341429.0_rk/6820363962896.0_rk, &
           2006345519317.0_rk/3224310063776.0_rk, 2802321613138.0_rk/2924317926251.0_rk]
    allocate(Uttmp, mold=U)
    finished = .not. (t < tEnd) .or. maxIter == 0   ! timedisc.f90:236
    do while (.not. finished)
      call StateToDevice()
      dt = CALCTIMESTEP(errType)
      if (errType /= 0) then
        write(*, '(A,I0)') ' ERROR: CALCTIMESTEP errType=', errType
        stop 3
      end if
      lastStep = tEnd - t <= dt*(1.0_rk + 1.0e-4_rk)   ! timedisc.f90:313-317
      if (lastStep) dt = tEnd - t
      call DGTimeDerivative_weakForm(t)
      Uttmp = Ut
      U = U + Ut*(b(1)*dt)
      do n = 2, 5
        call DGTimeDerivative_weakForm(t + dt*RKc(n))
        Uttmp = Uttmp*(-A(n)) + Ut
        U = U + Uttmp*(b(n)*dt)
      end do
      n = c%N
      t = t + dt
      iter = iter + 1
      finished = lastStep .or. iter == maxIter
    end do
    call StateToDevice()
  end if
  call system_clock(c1)
  t1 = real(c1 - c0, rk)/real(cr, rk)
  t0 = 0.0_rk
  rate = real(E, rk)*real((c%N+1)**3, rk)*5.0_rk*real(iter, rk)/max(t1 - t0, 1.0e-30_rk)
  call StateToHost()
  write(*, '(A,I0,A,ES22.14,A,ES14.6)') ' iter=', iter, ' t=', t, ' dt=', dt
  if (nvar == 5) then
    call CalcEntropy(Ssum, dSdU)
    write(*, '(A,ES22.14,A,ES22.14)') ' Entropy=', Ssum, ' dSdU*Ut=', dSdU
  end if
  if (IniExactFunc == 1 .or. (IniExactFunc == 2 .and. nvar == 5)) then
    call CalcErrorNorms(t, L2(1:nvar), L2c(1:nvar), Linf(1:nvar))
    write(*, '(A14,9ES21.12)') ' L_2        : ', L2(1:nvar)
    write(*, '(A14,9ES21.12)') ' L_2_colloc : ', L2c(1:nvar)
    write(*, '(A14,9ES21.12)') ' L_inf      : ', Linf(1:nvar)
  end if
  write(*, '(A,ES14.6,A,ES14.6)') ' DOF-updates/s=', rate, '  PID [s/(DOF*timestep)]=', 5.0_rk/rate
  if (has_key('ProjectName')) then
    ! WriteState: <ProjectName>_State_<TIMESTAMP>.h5 (hdf5_output.f90:74-165, globals.f90:285-303)
    project = GETSTR('ProjectName')
    write(tstamp, '(F17.9)') t
    do iv = 1, len_trim(tstamp)
      if (tstamp(iv:iv) == ' ') tstamp(iv:iv) = '0'
    end do
    sfile = trim(project)//'_State_'//trim(tstamp)//'.h5'
    if (nvar == 9) then  ! StrVarNames of mhd/equation_vars.f90:78-91
      vnames = [character(len=40) :: 'Density', 'MomentumX', 'MomentumY', 'MomentumZ', 'TotalEnergy', &
                'MagneticFieldX', 'MagneticFieldY', 'MagneticFieldZ', 'Psi']
    else
      vnames(1:5) = [character(len=40) :: 'Density', 'MomentumX', 'MomentumY', 'MomentumZ', &
                     'EnergyStagnationDensity']
    end if
    do iv = 1, nvar
      vnames(iv) = trim(vnames(iv))//c_null_char
      vptr(iv) = c_loc(vnames(iv))
    end do
    call io_chk(fxio_state_write(trim(sfile)//c_null_char, trim(project)//c_null_char, &
                                 trim(meshFile)//c_null_char, n, nvar, E, t, vptr, U), 'fxio_state_write')
    write(*, '(A,A)') ' WRITE STATE TO HDF5 FILE: ', trim(sfile)
  end if
  if (has_key('StateFile')) then
    sfile = GETSTR('StateFile')
    open(newunit=u_out, file=trim(sfile), access='stream', form='unformatted', status='replace')
    write(u_out) U
    close(u_out)
  end if
  write(*, '(A)') ' FLUXO_GPU FINISHED'
  call FinalizeDG_GPU()

contains

  !> Restart from a state of degree Nr (restart.f90:183-232, InterpolateSolution): the modal
  !> Vandermonde matrix Nr -> N (interpolation for Nr < N; for Nr > N the Legendre modes above N
  !> are dropped, applied to J U with J interpolated from the 3 Ngeo nodes to Nr, and the
  !> result divided by J again, which keeps each element's integral of U).
  subroutine restart_change_n(Nr)
    integer, intent(in) :: Nr
    real(rk), allocatable :: Ur(:, :, :, :, :), Vdm(:, :), Vdm3(:, :), JNR(:, :, :, :)
    integer :: iel, i, j, k
    allocate(Ur(nvar, 0:Nr, 0:Nr, 0:Nr, E), Vdm(0:n, 0:Nr), Vdm3(0:Nr, 0:3*Ngeo), JNR(1, 0:Nr, 0:Nr, 0:Nr))
    call io_chk(fxio_state_read(trim(sfile)//c_null_char, Nr, nvar, E, Ur), 'fxio_state_read')
    write(*, '(A,I0,A,I0)') ' interpolating the restart state from N=', Nr, ' to N=', n
    call vandermonde(Nr, NODE_GL, n, NODE_GL, .true., Vdm)
    call vandermonde(3*Ngeo, NODE_GL, Nr, NODE_GL, .true., Vdm3)
    do iel = 1, E
      if (Nr > n) then
        call change_basis_3d(1, 3*Ngeo, Nr, Vdm3, DetJacRef(:, :, :, :, iel), JNR)
        do k = 0, Nr
          do j = 0, Nr
            do i = 0, Nr
              Ur(:, i, j, k, iel) = Ur(:, i, j, k, iel)*JNR(1, i, j, k)
            end do
          end do
        end do
      end if
      call change_basis_3d(nvar, Nr, n, Vdm, Ur(:, :, :, :, iel), U(:, :, :, :, iel))
      if (Nr > n) then
        do k = 0, n
          do j = 0, n
            do i = 0, n
              U(:, i, j, k, iel) = U(:, i, j, k, iel)*sJ(i, j, k, iel)
            end do
          end do
        end do
      end if
    end do
  end subroutine restart_change_n

  subroutine io_chk(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(A,I0,A)') ' ERROR in '//where//' (code ', rc, '): '//trim(fxio_error_text())
      stop 2
    end if
  end subroutine io_chk
end program fluxo_gpu
