!==================================================================================================
!> Host mesh for periodic hexahedral boxes and the curvilinear metric terms the engine consumes.
!>
!> Connectivity follows the reference's conventions: local side ids (src/defines.h:68-73),
!> ElemToSide(2,6,nElems) / SideToElem(5,nSides) (src/mesh/prepare_mesh.f90:679-710) and the
!> CGNS side/volume index mappings with flips (src/mesh/mappings.f90:262-515).  Element
!> e=(ix,iy,iz) is ix+Kx*(iy+Ky*iz)+1; side 3(e-1)+d is the +d face of e (master, flip 0) and
!> the -d face of its periodic neighbour (slave).  Geometry nodes are equidistant of degree
!> Ngeo; the optional periodic warp x += amp*prod_k sin(2 pi (x_k-x0_k)/L_k) equals the
!> reference's meshdeform (src/mesh/mesh.f90:238-245) on [0,2]^3 with amp=0.1.
!>
!> The metrics restate CalcMetrics (src/mesh/metrics.f90:78-323: curl-form metric terms,
!> Jacobian on NgeoRef=3*Ngeo projected modally to N) and CalcSurfMetrics/SurfMetricsFromJa
!> (metrics.f90:334-483) for Legendre-Gauss-Lobatto volume nodes, with the small sides of
!> non-conforming interfaces (Mortar_CalcSurfMetrics, src/mortar/mortar_metrics.f90:60-119).
!==================================================================================================
module fxh_mesh
  use, intrinsic :: iso_c_binding
  use fxh_basis
  implicit none
  private
  public :: box_connectivity, box_nodecoords, calc_metrics, fxh_box_mesh, fxh_metrics, fxh_metrics_nc

  integer, parameter :: ZETA_MINUS = 1, ETA_MINUS = 2, XI_PLUS = 3, ETA_PLUS = 4, XI_MINUS = 5, ZETA_PLUS = 6
  integer, parameter :: plus_side(3) = [XI_PLUS, ETA_PLUS, ZETA_PLUS]
  integer, parameter :: minus_side(3) = [XI_MINUS, ETA_MINUS, ZETA_MINUS]
  integer, parameter :: normal_dirs(6) = [3, 2, 1, 2, 1, 3]
  integer, parameter :: tang_dirs(6) = [1, 3, 2, 3, 2, 1]
  real(rk), parameter :: normal_signs(6) = [-1.0_rk, -1.0_rk, 1.0_rk, 1.0_rk, -1.0_rk, 1.0_rk]

contains

  subroutine flip_s2m(N, p, q, flip, pq)
    integer, intent(in) :: N, p, q, flip
    integer, intent(out) :: pq(2)
    select case (flip)
    case (0); pq = [p, q]
    case (1); pq = [q, p]
    case (2); pq = [N-p, q]
    case (3); pq = [N-q, N-p]
    case default; pq = [p, N-q]
    end select
  end subroutine flip_s2m

  subroutine side_to_vol(N, l, p, q, flip, loc, ijk)
    integer, intent(in) :: N, l, p, q, flip, loc
    integer, intent(out) :: ijk(3)
    integer :: pq(2)
    call flip_s2m(N, p, q, flip, pq)
    select case (loc)
    case (XI_MINUS);   ijk = [l, pq(2), pq(1)]
    case (XI_PLUS);    ijk = [N-l, pq(1), pq(2)]
    case (ETA_MINUS);  ijk = [pq(1), l, pq(2)]
    case (ETA_PLUS);   ijk = [N-pq(1), N-l, pq(2)]
    case (ZETA_MINUS); ijk = [pq(2), pq(1), l]
    case default;      ijk = [pq(1), pq(2), N-l]
    end select
  end subroutine side_to_vol

  subroutine side_to_vol2(N, p, q, loc, pq)
    integer, intent(in) :: N, p, q, loc
    integer, intent(out) :: pq(2)
    select case (loc)
    case (XI_MINUS, ZETA_MINUS); pq = [q, p]
    case (ETA_PLUS); pq = [N-p, q]
    case default; pq = [p, q]
    end select
  end subroutine side_to_vol2

  !> flip of the -d slave face so master point (p,q) and slave point hit the same node
  integer function find_flip(d)
    integer, intent(in) :: d
    integer :: f, p, q, m(3), s(3)
    logical :: ok
    find_flip = -1
    do f = 1, 4
      ok = .true.
      do q = 0, 2
        do p = 0, 2
          call side_to_vol(2, 0, p, q, 0, plus_side(d), m)
          call side_to_vol(2, 0, p, q, f, minus_side(d), s)
          m(d) = 0
          if (any(m /= s)) ok = .false.
        end do
      end do
      if (ok) then
        find_flip = f; return
      end if
    end do
  end function find_flip

  subroutine box_connectivity(KB, e2s, s2e)
    integer, intent(in) :: KB(3)
    integer, intent(out) :: e2s(2, 6, KB(1)*KB(2)*KB(3)), s2e(5, 3*KB(1)*KB(2)*KB(3))
    integer :: ix, iy, iz, d, e, en, s, c(3), cn(3), flips(3)
    do d = 1, 3
      flips(d) = find_flip(d)
    end do
    do iz = 0, KB(3)-1
      do iy = 0, KB(2)-1
        do ix = 0, KB(1)-1
          c = [ix, iy, iz]
          e = ix + KB(1)*(iy + KB(2)*iz) + 1
          do d = 1, 3
            cn = c
            cn(d) = mod(cn(d) + 1, KB(d))
            en = cn(1) + KB(1)*(cn(2) + KB(2)*cn(3)) + 1
            s = 3*(e-1) + d
            s2e(:, s) = [e, en, plus_side(d), minus_side(d), flips(d)]
            e2s(:, plus_side(d), e) = [s, 0]
            e2s(:, minus_side(d), en) = [s, flips(d)]
          end do
        end do
      end do
    end do
  end subroutine box_connectivity

  subroutine box_nodecoords(KB, x0, length, Ngeo, warp, amp, X)
    integer, intent(in) :: KB(3), Ngeo
    real(rk), intent(in) :: x0(3), length(3), amp
    logical, intent(in) :: warp
    real(rk), intent(out) :: X(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, KB(1)*KB(2)*KB(3))
    integer :: ix, iy, iz, i, j, k, d, e, c(3), loc(3)
    real(rk) :: xx(3), t, s, pi
    pi = acos(-1.0_rk)
    do iz = 0, KB(3)-1
      do iy = 0, KB(2)-1
        do ix = 0, KB(1)-1
          e = ix + KB(1)*(iy + KB(2)*iz) + 1
          c = [ix, iy, iz]
          do k = 0, Ngeo
            do j = 0, Ngeo
              do i = 0, Ngeo
                loc = [i, j, k]
                do d = 1, 3
                  t = real(c(d)*Ngeo + loc(d), rk)/real(KB(d)*Ngeo, rk)
                  xx(d) = x0(d) + t*length(d)
                end do
                if (warp) then
                  s = amp
                  do d = 1, 3
                    s = s*sin(2.0_rk*pi*(xx(d) - x0(d))/length(d))
                  end do
                  xx = xx + s
                end if
                X(:, i, j, k, e) = xx
              end do
            end do
          end do
        end do
      end do
    end do
  end subroutine box_nodecoords

  !> CalcMetrics + CalcSurfMetrics; ierr=1 for non-positive or badly scaled Jacobians.
  subroutine calc_metrics(N, Ngeo, nElems, nSides, NodeCoords, e2s, Mf, Mg, Mh, sJ, NormVec, &
                          TangVec1, TangVec2, SurfElem, Elem_xGP, ierr, nodesGL, DetJacRef, MortarType, &
                          MortarInfo)
    integer, intent(in) :: N, Ngeo, nElems, nSides
    logical, intent(in), optional :: nodesGL   ! NodeCoords on Gauss-Lobatto instead of equidistant nodes
    ! the Jacobian on the 3 Ngeo Gauss-Lobatto nodes (DetJac_Ref of metrics.f90), for Restart
    real(rk), intent(out), optional :: DetJacRef(1, 0:3*Ngeo, 0:3*Ngeo, 0:3*Ngeo, nElems)
    ! non-conforming interfaces (fxio_mesh_mortars): the small sides of a big mortar side get the
    ! big side's metric terms interpolated to them and scaled by the area ratio
    ! (Mortar_CalcSurfMetrics, mortar_metrics.f90:60-119)
    integer, intent(in), optional :: MortarType(2, nSides), MortarInfo(2, 0:4, *)
    real(rk), intent(in) :: NodeCoords(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, nElems)
    integer, intent(in) :: e2s(2, 6, nElems)
    real(rk), intent(out), dimension(3, 0:N, 0:N, 0:N, nElems) :: Mf, Mg, Mh, Elem_xGP
    real(rk), intent(out) :: sJ(0:N, 0:N, 0:N, nElems)
    real(rk), intent(out), dimension(3, 0:N, 0:N, nSides) :: NormVec, TangVec1, TangVec2
    real(rk), intent(out) :: SurfElem(0:N, 0:N, nSides)
    integer, intent(out) :: ierr
    integer :: NgeoRef, e, i, j, k, l, q, nn, loc, p, SideID, flip, nd, td, pq(2), ijk(3)
    real(rk) :: V_eq_gl(0:Ngeo, 0:Ngeo), DGLgeo(0:Ngeo, 0:Ngeo), V_gl_glN(0:N, 0:Ngeo), DGLN(0:N, 0:N)
    real(rk), allocatable :: V_gl_ref(:, :), V_ref_N(:, :), dXref(:, :, :, :, :), detRef(:, :, :, :)
    real(rk) :: xg(0:Ngeo), xN(0:N)
    real(rk) :: XGLgeo(3, 0:Ngeo, 0:Ngeo, 0:Ngeo), XGLN(3, 0:N, 0:N, 0:N)
    real(rk) :: dXgeo(3, 3, 0:Ngeo, 0:Ngeo, 0:Ngeo), dXN(3, 3, 0:N, 0:N, 0:N)
    real(rk) :: R(3, 3, 0:N, 0:N, 0:N), Ja(3, 3, 0:N, 0:N, 0:N), detN(1, 0:N, 0:N, 0:N)
    real(rk) :: sl_geo(3, 0:Ngeo, 0:Ngeo, 0:Ngeo), sl_N(3, 0:N, 0:N, 0:N)
    real(rk) :: jmin, jmax, ns
    real(rk), allocatable :: sl_ref(:, :, :, :)
    real(rk) :: Jaf(3, 3, 0:N, 0:N), Jas(3, 3, 0:N, 0:N, 4), MInt(0:N, 0:N, 2), MProj(0:N, 0:N, 2)
    integer :: im, nm, s2
    ierr = 0
    if (present(MortarType)) call mortar_ops(N, 0, MInt, MProj)
    NgeoRef = 3*Ngeo
    allocate(V_gl_ref(0:NgeoRef, 0:Ngeo), V_ref_N(0:N, 0:NgeoRef))
    allocate(dXref(3, 3, 0:NgeoRef, 0:NgeoRef, 0:NgeoRef), detRef(1, 0:NgeoRef, 0:NgeoRef, 0:NgeoRef))
    allocate(sl_ref(3, 0:NgeoRef, 0:NgeoRef, 0:NgeoRef))
    call vandermonde(Ngeo, NODE_EQUI, Ngeo, NODE_GL, .false., V_eq_gl)
    if (present(nodesGL)) then
      if (nodesGL) call vandermonde(Ngeo, NODE_GL, Ngeo, NODE_GL, .false., V_eq_gl)
    end if
    call nodes_of(Ngeo, NODE_GL, xg)
    call derivative_matrix(Ngeo, xg, DGLgeo)
    call vandermonde(Ngeo, NODE_GL, NgeoRef, NODE_GL, .false., V_gl_ref)
    call vandermonde(NgeoRef, NODE_GL, N, NODE_GL, .true., V_ref_N)
    call vandermonde(Ngeo, NODE_GL, N, NODE_GL, .false., V_gl_glN)
    call nodes_of(N, NODE_GL, xN)
    call derivative_matrix(N, xN, DGLN)
    do e = 1, nElems
      call change_basis_3d(3, Ngeo, Ngeo, V_eq_gl, NodeCoords(:, :, :, :, e), XGLgeo)
      call change_basis_3d(3, Ngeo, N, V_gl_glN, XGLgeo, XGLN)
      Elem_xGP(:, :, :, :, e) = XGLN
      dXgeo = 0.0_rk
      do k = 0, Ngeo
        do j = 0, Ngeo
          do i = 0, Ngeo
            do l = 0, Ngeo
              dXgeo(1, :, i, j, k) = dXgeo(1, :, i, j, k) + DGLgeo(i, l)*XGLgeo(:, l, j, k)
              dXgeo(2, :, i, j, k) = dXgeo(2, :, i, j, k) + DGLgeo(j, l)*XGLgeo(:, i, l, k)
              dXgeo(3, :, i, j, k) = dXgeo(3, :, i, j, k) + DGLgeo(k, l)*XGLgeo(:, i, j, l)
            end do
          end do
        end do
      end do
      do nn = 1, 3
        sl_geo = dXgeo(:, nn, :, :, :)
        call change_basis_3d(3, Ngeo, NgeoRef, V_gl_ref, sl_geo, sl_ref)
        dXref(:, nn, :, :, :) = sl_ref
      end do
      do k = 0, NgeoRef
        do j = 0, NgeoRef
          do i = 0, NgeoRef
            associate (a => dXref(:, :, i, j, k))
              detRef(1, i, j, k) = 0.0_rk + a(1,1)*(a(2,2)*a(3,3) - a(3,2)*a(2,3)) &
                                          + a(2,1)*(a(3,2)*a(1,3) - a(1,2)*a(3,3)) &
                                          + a(3,1)*(a(1,2)*a(2,3) - a(2,2)*a(1,3))
            end associate
          end do
        end do
      end do
      if (present(DetJacRef)) DetJacRef(:, :, :, :, e) = detRef
      call change_basis_3d(1, NgeoRef, N, V_ref_N, detRef, detN)
      sJ(:, :, :, e) = 1.0_rk/detN(1, :, :, :)
      jmin = minval(detN); jmax = maxval(detN)
      if (jmin <= 0.0_rk .or. jmin/jmax < 0.001_rk) ierr = 1
      if (N >= Ngeo) then
        do nn = 1, 3
          sl_geo = dXgeo(:, nn, :, :, :)
          call change_basis_3d(3, Ngeo, N, V_gl_glN, sl_geo, sl_N)
          dXN(:, nn, :, :, :) = sl_N
        end do
      else
        dXN = 0.0_rk
        do k = 0, N
          do j = 0, N
            do i = 0, N
              do l = 0, N
                dXN(1, :, i, j, k) = dXN(1, :, i, j, k) + DGLN(i, l)*XGLN(:, l, j, k)
                dXN(2, :, i, j, k) = dXN(2, :, i, j, k) + DGLN(j, l)*XGLN(:, i, l, k)
                dXN(3, :, i, j, k) = dXN(3, :, i, j, k) + DGLN(k, l)*XGLN(:, i, j, l)
              end do
            end do
          end do
        end do
      end if
      ! curl form: R(:,nn) = 1/2 (X x dX/dxi_dd) components; Ja = -curl R
      do k = 0, N
        do j = 0, N
          do i = 0, N
            associate (dx => dXN(:, :, i, j, k), X => XGLN(:, i, j, k))
              R(:, 1, i, j, k) = 0.5_rk*(X(3)*dx(:, 2) - X(2)*dx(:, 3))
              R(:, 2, i, j, k) = 0.5_rk*(X(1)*dx(:, 3) - X(3)*dx(:, 1))
              R(:, 3, i, j, k) = 0.5_rk*(X(2)*dx(:, 1) - X(1)*dx(:, 2))
            end associate
          end do
        end do
      end do
      Ja = 0.0_rk
      do k = 0, N
        do j = 0, N
          do i = 0, N
            do q = 0, N
              Ja(1, :, i, j, k) = Ja(1, :, i, j, k) - DGLN(j, q)*R(3, :, i, q, k)
              Ja(2, :, i, j, k) = Ja(2, :, i, j, k) - DGLN(k, q)*R(1, :, i, j, q)
              Ja(3, :, i, j, k) = Ja(3, :, i, j, k) - DGLN(i, q)*R(2, :, q, j, k)
            end do
            do q = 0, N
              Ja(1, :, i, j, k) = Ja(1, :, i, j, k) + DGLN(k, q)*R(2, :, i, j, q)
              Ja(2, :, i, j, k) = Ja(2, :, i, j, k) + DGLN(i, q)*R(3, :, q, j, k)
              Ja(3, :, i, j, k) = Ja(3, :, i, j, k) + DGLN(j, q)*R(1, :, i, q, k)
            end do
          end do
        end do
      end do
      Mf(:, :, :, :, e) = Ja(1, :, :, :, :)
      Mg(:, :, :, :, e) = Ja(2, :, :, :, :)
      Mh(:, :, :, :, e) = Ja(3, :, :, :, :)
      ! surface metrics on master sides
      do loc = 1, 6
        SideID = e2s(1, loc, e); flip = e2s(2, loc, e)
        if (flip /= 0) cycle
        nd = normal_dirs(loc); td = tang_dirs(loc); ns = normal_signs(loc)
        do q = 0, N
          do p = 0, N
            call side_to_vol2(N, p, q, loc, pq)
            select case (loc)
            case (XI_MINUS);   ijk = [0, pq(1), pq(2)]
            case (XI_PLUS);    ijk = [N, pq(1), pq(2)]
            case (ETA_MINUS);  ijk = [pq(1), 0, pq(2)]
            case (ETA_PLUS);   ijk = [pq(1), N, pq(2)]
            case (ZETA_MINUS); ijk = [pq(1), pq(2), 0]
            case default;      ijk = [pq(1), pq(2), N]
            end select
            Jaf(:, :, p, q) = Ja(:, :, ijk(1), ijk(2), ijk(3))
            call surf_from_ja(Jaf(nd, :, p, q), Jaf(td, :, p, q), ns, NormVec(:, p, q, SideID), &
                              TangVec1(:, p, q, SideID), TangVec2(:, p, q, SideID), SurfElem(p, q, SideID))
          end do
        end do
        if (.not. present(MortarType)) cycle
        if (MortarType(1, SideID) <= 0) cycle
        call interp_big_to_small(9, N, MortarType(1, SideID), MInt, Jaf, Jas)
        nm = merge(4, 2, MortarType(1, SideID) == 1)
        do im = 1, nm
          s2 = MortarInfo(1, im, MortarType(2, SideID))
          if (MortarInfo(2, im, MortarType(2, SideID)) > 0) cycle
          Jas(:, :, :, :, im) = merge(0.25_rk, 0.5_rk, nm == 4)*Jas(:, :, :, :, im)
          do q = 0, N
            do p = 0, N
              call surf_from_ja(Jas(nd, :, p, q, im), Jas(td, :, p, q, im), ns, NormVec(:, p, q, s2), &
                                TangVec1(:, p, q, s2), TangVec2(:, p, q, s2), SurfElem(p, q, s2))
            end do
          end do
        end do
      end do
    end do
  end subroutine calc_metrics

  !> Normal, tangents and surface element of one face point from the face's metric terms in the
  !> normal (jn) and first tangential (jt) direction (SurfMetricsFromJa, metrics.f90:460-483).
  subroutine surf_from_ja(jn, jt, ns, nv, t1, t2, se)
    real(rk), intent(in) :: jn(3), jt(3), ns
    real(rk), intent(out) :: nv(3), t1(3), t2(3), se
    real(rk) :: dot, tn
    se = sqrt(jn(1)*jn(1) + jn(2)*jn(2) + jn(3)*jn(3))
    nv = ns*jn/se
    dot = jt(1)*nv(1) + jt(2)*nv(2) + jt(3)*nv(3)
    t1 = jt - dot*nv
    tn = sqrt(t1(1)*t1(1) + t1(2)*t1(2) + t1(3)*t1(3))
    t1 = t1/tn
    t2 = [nv(2)*t1(3) - nv(3)*t1(2), nv(3)*t1(1) - nv(1)*t1(3), nv(1)*t1(2) - nv(2)*t1(1)]
  end subroutine surf_from_ja

  !> C entry for callers that build a periodic box mesh through the host (Python bench and
  !> tests): connectivity, node coordinates and all metric arrays in the reference layout.
  subroutine fxh_box_mesh(Kx, Ky, Kz, N, Ngeo, x0, length, warp, amp, e2s, s2e, NodeCoords, Mf, Mg, &
                          Mh, sJ, NormVec, TangVec1, TangVec2, SurfElem, Elem_xGP, ierr) &
                          bind(C, name="fxh_box_mesh")
    integer(c_int), value :: Kx, Ky, Kz, N, Ngeo, warp
    real(c_double), intent(in) :: x0(3), length(3)
    real(c_double), value :: amp
    integer(c_int), intent(out) :: e2s(2, 6, Kx*Ky*Kz), s2e(5, 3*Kx*Ky*Kz)
    real(c_double), intent(out) :: NodeCoords(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, Kx*Ky*Kz)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, 0:N, Kx*Ky*Kz) :: Mf, Mg, Mh, Elem_xGP
    real(c_double), intent(out) :: sJ(0:N, 0:N, 0:N, Kx*Ky*Kz)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, 3*Kx*Ky*Kz) :: NormVec, TangVec1, TangVec2
    real(c_double), intent(out) :: SurfElem(0:N, 0:N, 3*Kx*Ky*Kz)
    integer(c_int), intent(out) :: ierr
    integer :: KB(3), E, ie
    KB = [Kx, Ky, Kz]
    E = Kx*Ky*Kz
    if (any(KB < 1) .or. N < 1 .or. Ngeo < 1) then
      ierr = -1; return
    end if
    call box_connectivity(KB, e2s, s2e)
    call box_nodecoords(KB, x0, length, Ngeo, warp /= 0, amp, NodeCoords)
    call calc_metrics(N, Ngeo, E, 3*E, NodeCoords, e2s, Mf, Mg, Mh, sJ, NormVec, TangVec1, TangVec2, &
                      SurfElem, Elem_xGP, ie)
    ierr = ie
  end subroutine fxh_box_mesh

  !> C entry for a mesh read from a file (HOPR reader, include/fluxo_io.h): the metric arrays
  !> of CalcMetrics / CalcSurfMetrics for given connectivity and node coordinates.
  subroutine fxh_metrics(N, Ngeo, nElems, nodesGL, e2s, NodeCoords, Mf, Mg, Mh, sJ, NormVec, TangVec1, &
                         TangVec2, SurfElem, Elem_xGP, ierr) bind(C, name="fxh_metrics")
    integer(c_int), value :: N, Ngeo, nElems, nodesGL
    integer(c_int), intent(in) :: e2s(2, 6, nElems)
    real(c_double), intent(in) :: NodeCoords(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, nElems)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, 0:N, nElems) :: Mf, Mg, Mh, Elem_xGP
    real(c_double), intent(out) :: sJ(0:N, 0:N, 0:N, nElems)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, 3*nElems) :: NormVec, TangVec1, TangVec2
    real(c_double), intent(out) :: SurfElem(0:N, 0:N, 3*nElems)
    integer(c_int), intent(out) :: ierr
    integer :: ie
    if (N < 1 .or. Ngeo < 1 .or. nElems < 1) then
      ierr = -1; return
    end if
    call calc_metrics(N, Ngeo, nElems, 3*nElems, NodeCoords, e2s, Mf, Mg, Mh, sJ, NormVec, TangVec1, &
                      TangVec2, SurfElem, Elem_xGP, ie, nodesGL /= 0)
    ierr = ie
  end subroutine fxh_metrics

  !> fxh_metrics for a mesh with non-conforming interfaces: nSides sides of which the first
  !> nMortarSides are big mortar sides, MortarType(2,nSides) / MortarInfo(2,0:4,nMortarSides)
  !> as fxio_mesh_mortars gives them.
  subroutine fxh_metrics_nc(N, Ngeo, nElems, nSides, nMortarSides, nodesGL, e2s, MortarType, MortarInfo, &
                            NodeCoords, Mf, Mg, Mh, sJ, NormVec, TangVec1, TangVec2, SurfElem, Elem_xGP, ierr) &
                            bind(C, name="fxh_metrics_nc")
    integer(c_int), value :: N, Ngeo, nElems, nSides, nMortarSides, nodesGL
    integer(c_int), intent(in) :: e2s(2, 6, nElems), MortarType(2, nSides), MortarInfo(2, 0:4, max(1, nMortarSides))
    real(c_double), intent(in) :: NodeCoords(3, 0:Ngeo, 0:Ngeo, 0:Ngeo, nElems)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, 0:N, nElems) :: Mf, Mg, Mh, Elem_xGP
    real(c_double), intent(out) :: sJ(0:N, 0:N, 0:N, nElems)
    real(c_double), intent(out), dimension(3, 0:N, 0:N, nSides) :: NormVec, TangVec1, TangVec2
    real(c_double), intent(out) :: SurfElem(0:N, 0:N, nSides)
    integer(c_int), intent(out) :: ierr
    integer :: ie
    if (N < 1 .or. Ngeo < 1 .or. nElems < 1 .or. nSides < 1 .or. nMortarSides < 0) then
      ierr = -1; return
    end if
    call calc_metrics(N, Ngeo, nElems, nSides, NodeCoords, e2s, Mf, Mg, Mh, sJ, NormVec, TangVec1, &
                      TangVec2, SurfElem, Elem_xGP, ie, nodesGL /= 0, MortarType=MortarType, MortarInfo=MortarInfo)
    ierr = ie
  end subroutine fxh_metrics_nc

end module fxh_mesh
