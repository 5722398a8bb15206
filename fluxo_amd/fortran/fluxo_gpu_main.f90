!==================================================================================================
!> fluxo_gpu: parameter.ini-driven time loop on the MI355X engine (the reference's PROGRAM Fluxo,
!> src/fluxo.f90, restricted to periodic box meshes, with TimeDisc timedisc.f90:144-415).
!>
!> Keys (reference meaning unless noted): N, Riemann, VolumeFlux, IniExactFunc, AdvVel,
!> IniAmplitude, kappa, mu0, Pr, CFLscale, DFLscale, tEnd, maxIter, Analyze_dt.
!> Mesh keys (this driver's periodic box generator replaces MeshFile): BoxElems (3 ints),
!> BoxCorner, BoxLength (3 reals each), Ngeo, meshdeform (logical), meshdeformAmp.
!> Parabolic (logical) selects Navier-Stokes at run time, with BR1 lifting or, with Lifting = 2
!> (the reference's compile-time PP_Lifting), BR2 with penalty etaBR2 (br2.f90:116, default 2).
!> DeviceResident (logical, default T): T keeps U on the device and runs whole steps there;
!> F calls DGTimeDerivative_weakForm once per stage from the host RK loop (drop-in form).
!> StateFile (optional): U(5,0:N,0:N,0:N,nElems) at the end, raw FP64 in Fortran order (the
!> place of the reference's WriteState, hdf5_output.f90:74-165).
!==================================================================================================
program fluxo_gpu
  use, intrinsic :: iso_c_binding
  use fxh_basis
  use fxh_mesh
  use fxh_readini
  use fluxo_gpu_bind, only: fx_config
  use fxh_dg
  implicit none
  character(len=512) :: fname
  character(len=256) :: sfile
  integer :: u_out
  type(fx_config) :: c
  integer :: KB(3), Ngeo, E, S, ierr, errType, iter, maxIter, nChunk, n
  real(rk) :: x0(3), L(3), amp, t, tEnd, Analyze_dt, tAnalyze, Ssum, dSdU, L2(5), L2c(5), Linf(5)
  real(rk) :: xg(0:15), wg(0:15), t0, t1, rate, A(5), b(5), RKc(5)
  integer(c_int), allocatable :: e2s(:, :, :), s2e(:, :)
  real(rk), allocatable :: NC(:, :, :, :, :), Mf(:, :, :, :, :), Mg(:, :, :, :, :), Mh(:, :, :, :, :)
  real(rk), allocatable :: sJm(:, :, :, :), nv(:, :, :, :), t1v(:, :, :, :), t2v(:, :, :, :), se(:, :, :)
  real(rk), allocatable :: xGP(:, :, :, :, :), Uttmp(:, :, :, :, :)
  logical :: warp, resident
  integer(8) :: c0, c1, cr

  if (command_argument_count() < 1) then
    write(*, '(A)') ' usage: fluxo_gpu parameter.ini'
    stop 2
  end if
  call get_command_argument(1, fname)
  call read_ini(trim(fname))

  c%N = GETINT('N')
  c%riemann = GETINT('Riemann', '1')
  c%volumeFlux = GETINT('VolumeFlux', '0')
  c%parabolic = merge(1, 0, GETLOGICAL('Parabolic', 'F'))
  c%kappa = GETREAL('kappa', '1.4')
  c%mu0 = GETREAL('mu0', '0.')
  c%Pr = GETREAL('Pr', '0.72')
  c%lifting = GETINT('Lifting', '1')
  c%etaBR2 = GETREAL('etaBR2', '2.')
  IniExactFunc = GETINT('IniExactFunc')
  AdvVel = GETREALARRAY('AdvVel', 3, '1.,1.,1.')
  IniAmplitude = GETREAL('IniAmplitude', '0.1')
  CFLScale = GETREAL('CFLscale')
  DFLScale = GETREAL('DFLscale', '0.9')
  tEnd = GETREAL('tEnd', '1.0e30')
  maxIter = GETINT('maxIter', '100')
  Analyze_dt = GETREAL('Analyze_dt', '1.0e30')
  resident = GETLOGICAL('DeviceResident', 'T')
  KB = GETINTARRAY('BoxElems', 3)
  x0 = GETREALARRAY('BoxCorner', 3, '0.,0.,0.')
  L = GETREALARRAY('BoxLength', 3, '2.,2.,2.')
  Ngeo = GETINT('Ngeo', '1')
  warp = GETLOGICAL('meshdeform', 'F')
  amp = GETREAL('meshdeformAmp', '0.1')

  E = KB(1)*KB(2)*KB(3); S = 3*E; n = c%N
  c%nElems = E; c%nSides = S
  allocate(e2s(2, 6, E), s2e(5, S), NC(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, E))
  allocate(Mf(3, 0:n, 0:n, 0:n, E), Mg(3, 0:n, 0:n, 0:n, E), Mh(3, 0:n, 0:n, 0:n, E), xGP(3, 0:n, 0:n, 0:n, E))
  allocate(sJm(0:n, 0:n, 0:n, E), nv(3, 0:n, 0:n, S), t1v(3, 0:n, 0:n, S), t2v(3, 0:n, 0:n, S), se(0:n, 0:n, S))
  call box_connectivity(KB, e2s, s2e)
  call box_nodecoords(KB, x0, L, Ngeo, warp, amp, NC)
  call calc_metrics(n, Ngeo, E, S, NC, e2s, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP, ierr)
  if (ierr /= 0) then
    write(*, '(A)') ' ERROR: negative or badly scaled Jacobian in the generated mesh'
    stop 2
  end if
  call InitDG_GPU(c, e2s, s2e, Mf, Mg, Mh, sJm, nv, t1v, t2v, se, xGP)
  call lgl_nodes(n, xg(0:n), wg(0:n))
  call FillIni()
  call StateToDevice()
  write(*, '(A,I0,A,I0,A,I0,A,I0)') ' fluxo_gpu: nElems=', E, ' N=', n, ' VolumeFlux=', c%volumeFlux, &
      ' Riemann=', c%riemann
  t = 0.0_rk
  call CalcEntropy(Ssum, dSdU)
  write(*, '(A,ES22.14,A,ES22.14)') ' t=0 Entropy=', Ssum, ' dSdU*Ut=', dSdU
  tAnalyze = Analyze_dt
  iter = 0
  call system_clock(c0, cr)
  if (resident) then
    do while (iter < maxIter .and. t < tEnd)
      nChunk = maxIter - iter
      call TimeDiscRun(nChunk, t, errType)
      iter = iter + nChunk
      if (errType /= 0) then
        write(*, '(A,I0)') ' ERROR: CALCTIMESTEP errType=', errType
        stop 3
      end if
    end do
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
    do while (iter < maxIter .and. t < tEnd)
      call StateToDevice()
      dt = CALCTIMESTEP(errType)
      if (errType /= 0) then
        write(*, '(A,I0)') ' ERROR: CALCTIMESTEP errType=', errType
        stop 3
      end if
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
    end do
    call StateToDevice()
  end if
  call system_clock(c1)
  t1 = real(c1 - c0, rk)/real(cr, rk)
  t0 = 0.0_rk
  rate = real(E, rk)*real((c%N+1)**3, rk)*5.0_rk*real(iter, rk)/max(t1 - t0, 1.0e-30_rk)
  call StateToHost()
  call CalcEntropy(Ssum, dSdU)
  write(*, '(A,I0,A,ES22.14,A,ES14.6)') ' iter=', iter, ' t=', t, ' dt=', dt
  write(*, '(A,ES22.14,A,ES22.14)') ' Entropy=', Ssum, ' dSdU*Ut=', dSdU
  if (IniExactFunc == 2) then
    call CalcErrorNorms(t, L2, L2c, Linf)
    write(*, '(A14,5ES21.12)') ' L_2        : ', L2
    write(*, '(A14,5ES21.12)') ' L_2_colloc : ', L2c
    write(*, '(A14,5ES21.12)') ' L_inf      : ', Linf
  end if
  write(*, '(A,ES14.6,A,ES14.6)') ' DOF-updates/s=', rate, '  PID [s/(DOF*timestep)]=', 5.0_rk/rate
  if (has_key('StateFile')) then
    sfile = GETSTR('StateFile')
    open(newunit=u_out, file=trim(sfile), access='stream', form='unformatted', status='replace')
    write(u_out) U
    close(u_out)
  end if
  write(*, '(A)') ' FLUXO_GPU FINISHED'
  call FinalizeDG_GPU()
end program fluxo_gpu
