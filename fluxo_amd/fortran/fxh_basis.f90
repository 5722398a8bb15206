!==================================================================================================
!> Host-side 1D polynomial bases on [-1,1]: Legendre-Gauss and Legendre-Gauss-Lobatto nodes and
!> weights, barycentric weights, Lagrange interpolation, derivative matrices, nodal and modal
!> Vandermonde matrices and the tensor-product change of basis.  Follows the reference's
!> src/interpolation/basis.f90 (LegendrePolynomialAndDerivative :286-323,
!> LegendreGaussNodesAndWeights :430-497, LegGaussLobNodesAndWeights :544-608,
!> BarycentricWeights :615-634, PolynomialDerivativeMatrix :641-663,
!> LagrangeInterpolationPolys :732-767, BuildLegendreVdm :196-252),
!> interpolation.f90 (GetVandermonde :271-339), changeBasis.f90 (ChangeBasis3D :225-275) and the
!> mortar operators of src/mortar/mortar_vars.f90 (MortarBasis_BigToSmall :70-96,
!> MortarBasis_SmallToBig_Collocation :107-146, MortarBasis_SmallToBig_Projection :153-196,
!> InterpolateBigToSmall :203-266).
!==================================================================================================
module fxh_basis
  use, intrinsic :: iso_c_binding, only: c_double
  implicit none
  private
  integer, parameter, public :: rk = c_double
  integer, parameter, public :: NODE_EQUI = 0, NODE_GL = 1, NODE_GAUSS = 2
  public :: legendre, gauss_nodes, lgl_nodes, nodes_of, barycentric, derivative_matrix
  public :: lagrange_polys, legendre_vdm, vandermonde, change_basis_3d, mortar_ops, interp_big_to_small

contains

  subroutine legendre(N, x, L, Lder)
    integer, intent(in) :: N
    real(rk), intent(in) :: x
    real(rk), intent(out) :: L, Lder
    real(rk) :: Lm1, Lm2, Ldm1, Ldm2
    integer :: i
    if (N == 0) then
      L = 1.0_rk; Lder = 0.0_rk
    else if (N == 1) then
      L = x; Lder = 1.0_rk
    else
      Lm2 = 1.0_rk; Lm1 = x; Ldm2 = 0.0_rk; Ldm1 = 1.0_rk
      do i = 2, N
        L = (real(2*i-1, rk)*x*Lm1 - real(i-1, rk)*Lm2)/real(i, rk)
        Lder = Ldm2 + real(2*i-1, rk)*Lm1
        Lm2 = Lm1; Lm1 = L; Ldm2 = Ldm1; Ldm1 = Lder
      end do
    end if
    L = L*sqrt(real(N, rk) + 0.5_rk)
    Lder = Lder*sqrt(real(N, rk) + 0.5_rk)
  end subroutine legendre

  subroutine gauss_nodes(N, x, w)
    integer, intent(in) :: N
    real(rk), intent(out) :: x(0:N), w(0:N)
    real(rk) :: cheb, xi, L, Ld, dx
    integer :: i, it
    if (N == 0) then
      x = 0.0_rk; w = 2.0_rk; return
    else if (N == 1) then
      x(0) = -sqrt(1.0_rk/3.0_rk); x(1) = -x(0); w = 1.0_rk; return
    end if
    cheb = 2.0_rk*atan(1.0_rk)/real(N+1, rk)
    do i = 0, (N+1)/2 - 1
      xi = -cos(cheb*real(2*i+1, rk))
      do it = 0, 10
        call legendre(N+1, xi, L, Ld)
        dx = -L/Ld
        xi = xi + dx
        if (abs(dx) < 1.0e-15_rk*abs(xi)) exit
      end do
      call legendre(N+1, xi, L, Ld)
      x(i) = xi; x(N-i) = -xi
      w(i) = (2.0_rk*N + 3)/((1.0_rk - xi*xi)*Ld*Ld); w(N-i) = w(i)
    end do
    if (mod(N, 2) == 0) then
      x(N/2) = 0.0_rk
      call legendre(N+1, 0.0_rk, L, Ld)
      w(N/2) = (2.0_rk*N + 3)/(Ld*Ld)
    end if
  end subroutine gauss_nodes

  subroutine q_and_l(N, x, q, qder, L)
    integer, intent(in) :: N
    real(rk), intent(in) :: x
    real(rk), intent(out) :: q, qder, L
    real(rk) :: Lm1, Lm2, Ld, Ldm1, Ldm2
    integer :: i
    Lm2 = 1.0_rk; Lm1 = x; Ldm2 = 0.0_rk; Ldm1 = 1.0_rk; L = 0.0_rk
    do i = 2, N
      L = (real(2*i-1, rk)*x*Lm1 - real(i-1, rk)*Lm2)/real(i, rk)
      Ld = Ldm2 + real(2*i-1, rk)*Lm1
      Lm2 = Lm1; Lm1 = L; Ldm2 = Ldm1; Ldm1 = Ld
    end do
    q = real(2*N+1, rk)/real(N+1, rk)*(x*L - Lm2)
    qder = real(2*N+1, rk)*L
  end subroutine q_and_l

  subroutine lgl_nodes(N, x, w)
    integer, intent(in) :: N
    real(rk), intent(out) :: x(0:N), w(0:N)
    real(rk) :: pi, c1, c2, xi, q, qd, L, dx
    integer :: i, it
    pi = acos(-1.0_rk)
    x(0) = -1.0_rk; x(N) = 1.0_rk
    w(0) = 2.0_rk/real(N*(N+1), rk); w(N) = w(0)
    if (N > 1) then
      c1 = pi/real(N, rk); c2 = 3.0_rk/(real(8*N, rk)*pi)
      do i = 1, (N+1)/2 - 1
        xi = -cos(c1*(real(i, rk) + 0.25_rk) - c2/(real(i, rk) + 0.25_rk))
        do it = 0, 10
          call q_and_l(N, xi, q, qd, L)
          dx = -q/qd
          xi = xi + dx
          if (abs(dx) < 1.0e-15_rk*abs(xi)) exit
        end do
        call q_and_l(N, xi, q, qd, L)
        x(i) = xi; x(N-i) = -xi
        w(i) = w(0)/(L*L); w(N-i) = w(i)
      end do
    end if
    if (mod(N, 2) == 0) then
      x(N/2) = 0.0_rk
      call q_and_l(N, 0.0_rk, q, qd, L)
      w(N/2) = w(0)/(L*L)
    end if
  end subroutine lgl_nodes

  subroutine nodes_of(N, ntype, x)
    integer, intent(in) :: N, ntype
    real(rk), intent(out) :: x(0:N)
    real(rk) :: w(0:N)
    integer :: i
    select case (ntype)
    case (NODE_GAUSS)
      call gauss_nodes(N, x, w)
    case (NODE_GL)
      call lgl_nodes(N, x, w)
    case default
      do i = 0, N
        x(i) = 2.0_rk*real(i, rk)/real(N, rk) - 1.0_rk
      end do
    end select
  end subroutine nodes_of

  subroutine barycentric(N, x, wb)
    integer, intent(in) :: N
    real(rk), intent(in) :: x(0:N)
    real(rk), intent(out) :: wb(0:N)
    integer :: i, j
    wb = 1.0_rk
    do i = 1, N
      do j = 0, i-1
        wb(j) = wb(j)*(x(j) - x(i))
        wb(i) = wb(i)*(x(i) - x(j))
      end do
    end do
    wb = 1.0_rk/wb
  end subroutine barycentric

  subroutine derivative_matrix(N, x, D)
    integer, intent(in) :: N
    real(rk), intent(in) :: x(0:N)
    real(rk), intent(out) :: D(0:N, 0:N)
    real(rk) :: wb(0:N)
    integer :: i, l
    call barycentric(N, x, wb)
    D = 0.0_rk
    do l = 0, N
      do i = 0, N
        if (l /= i) then
          D(i, l) = wb(l)/(wb(i)*(x(i) - x(l)))
          D(i, i) = D(i, i) - D(i, l)
        end if
      end do
    end do
  end subroutine derivative_matrix

  logical function almost_equal(a, b)
    real(rk), intent(in) :: a, b
    real(rk) :: tol
    tol = epsilon(1.0_rk)
    if (a == 0.0_rk .or. b == 0.0_rk) then
      almost_equal = abs(a - b) <= 2.0_rk*tol
    else
      almost_equal = (abs(a - b) <= tol*abs(a)) .and. (abs(a - b) <= tol*abs(b))
    end if
  end function almost_equal

  subroutine lagrange_polys(x, N, xGP, wb, L)
    real(rk), intent(in) :: x
    integer, intent(in) :: N
    real(rk), intent(in) :: xGP(0:N), wb(0:N)
    real(rk), intent(out) :: L(0:N)
    real(rk) :: s
    logical :: hit
    integer :: i
    hit = .false.
    do i = 0, N
      L(i) = 0.0_rk
      if (almost_equal(x, xGP(i))) then
        L(i) = 1.0_rk; hit = .true.
      end if
    end do
    if (hit) return
    s = 0.0_rk
    do i = 0, N
      L(i) = wb(i)/(x - xGP(i))
      s = s + L(i)
    end do
    do i = 0, N
      L(i) = L(i)/s
    end do
  end subroutine lagrange_polys

  subroutine nodal_vdm(Nin, Nout, wb, xin, xout, Vdm)
    integer, intent(in) :: Nin, Nout
    real(rk), intent(in) :: wb(0:Nin), xin(0:Nin), xout(0:Nout)
    real(rk), intent(out) :: Vdm(0:Nout, 0:Nin)
    integer :: r
    do r = 0, Nout
      call lagrange_polys(xout(r), Nin, xin, wb, Vdm(r, :))
    end do
  end subroutine nodal_vdm

  subroutine legendre_vdm(N, xi, Vdm_Leg, sVdm_Leg)
    integer, intent(in) :: N
    real(rk), intent(in) :: xi(0:N)
    real(rk), intent(out) :: Vdm_Leg(0:N, 0:N), sVdm_Leg(0:N, 0:N)
    integer :: Ng, i, j, r, c, k
    real(rk) :: wb(0:N), dummy, s
    real(rk), allocatable :: xg(:), wg(:), VlT(:, :), Vgp(:, :)
    Ng = 3*N/2 + 1
    allocate(xg(0:Ng), wg(0:Ng), VlT(0:N, 0:Ng), Vgp(0:Ng, 0:N))
    call barycentric(N, xi, wb)
    call gauss_nodes(Ng, xg, wg)
    do i = 0, Ng
      do j = 0, N
        call legendre(j, xg(i), VlT(j, i), dummy)
      end do
    end do
    do i = 0, Ng
      VlT(:, i) = VlT(:, i)*wg(i)
    end do
    call nodal_vdm(N, Ng, wb, xi, xg, Vgp)
    do r = 0, N
      do c = 0, N
        s = 0.0_rk
        do k = 0, Ng
          s = s + VlT(r, k)*Vgp(k, c)
        end do
        sVdm_Leg(r, c) = s
      end do
    end do
    do i = 0, N
      do j = 0, N
        call legendre(j, xi(i), Vdm_Leg(i, j), dummy)
      end do
    end do
  end subroutine legendre_vdm

  !> Vdm(0:Nout,0:Nin) from nodes of type typeIn to typeOut; modal projection when Nout<Nin.
  subroutine vandermonde(Nin, typeIn, Nout, typeOut, modal, Vdm)
    integer, intent(in) :: Nin, typeIn, Nout, typeOut
    logical, intent(in) :: modal
    real(rk), intent(out) :: Vdm(0:Nout, 0:Nin)
    real(rk) :: xin(0:Nin), xout(0:Nout), wbin(0:Nin), s
    real(rk) :: VLi(0:Nin, 0:Nin), sVLi(0:Nin, 0:Nin), VLo(0:Nout, 0:Nout), sVLo(0:Nout, 0:Nout)
    integer :: i, r, c, k
    if (typeIn == typeOut .and. Nin == Nout) then
      Vdm = 0.0_rk
      do i = 0, Nin
        Vdm(i, i) = 1.0_rk
      end do
      return
    end if
    call nodes_of(Nin, typeIn, xin)
    call barycentric(Nin, xin, wbin)
    call nodes_of(Nout, typeOut, xout)
    if (modal .and. Nout < Nin) then
      call legendre_vdm(Nin, xin, VLi, sVLi)
      call legendre_vdm(Nout, xout, VLo, sVLo)
      do r = 0, Nout
        do c = 0, Nin
          s = 0.0_rk
          do k = 0, Nout
            s = s + VLo(r, k)*sVLi(k, c)
          end do
          Vdm(r, c) = s
        end do
      end do
    else
      call nodal_vdm(Nin, Nout, wbin, xin, xout, Vdm)
    end if
  end subroutine vandermonde

  !> Xout(d,:,:,:) = (Vdm x Vdm x Vdm) Xin(d,:,:,:), one direction at a time.
  subroutine change_basis_3d(dim1, Nin, Nout, Vdm, Xin, Xout)
    integer, intent(in) :: dim1, Nin, Nout
    real(rk), intent(in) :: Vdm(0:Nout, 0:Nin), Xin(dim1, 0:Nin, 0:Nin, 0:Nin)
    real(rk), intent(out) :: Xout(dim1, 0:Nout, 0:Nout, 0:Nout)
    real(rk) :: b1(dim1, 0:Nout, 0:Nin, 0:Nin), b2(dim1, 0:Nout, 0:Nout, 0:Nin)
    integer :: i, j, k, io, jo, ko
    b1 = 0.0_rk
    do k = 0, Nin
      do j = 0, Nin
        do i = 0, Nin
          do io = 0, Nout
            b1(:, io, j, k) = b1(:, io, j, k) + Vdm(io, i)*Xin(:, i, j, k)
          end do
        end do
      end do
    end do
    b2 = 0.0_rk
    do k = 0, Nin
      do j = 0, Nin
        do jo = 0, Nout
          do io = 0, Nout
            b2(:, io, jo, k) = b2(:, io, jo, k) + Vdm(jo, j)*b1(:, io, j, k)
          end do
        end do
      end do
    end do
    Xout = 0.0_rk
    do k = 0, Nin
      do ko = 0, Nout
        do jo = 0, Nout
          do io = 0, Nout
            Xout(:, io, jo, ko) = Xout(:, io, jo, ko) + Vdm(ko, k)*b2(:, io, jo, k)
          end do
        end do
      end do
    end do
  end subroutine change_basis_3d

  !> Operators of a non-conforming interface between a big side and its two halves in one
  !> direction (interval 1 = [-1,0], 2 = [0,1]), on Legendre-Gauss-Lobatto nodes, in the
  !> reference's transposed layout (InitMortarBase, mortar.f90:63-110):
  !>   MInt(l,p,k)  = l_l(xi^k_p): big-side Lagrange polynomial l at node p of interval k;
  !>   MProj(l,p,k): weight of the interval-k node l in big-side node p of the flux projection,
  !>   whichMortar 0 = L2 projection through the Legendre modes with Gauss quadrature,
  !>   1 = collocation, M(i,j) = w_j / w_i l_i(xi^k_j).
  subroutine mortar_ops(N, whichMortar, MInt, MProj)
    integer, intent(in) :: N, whichMortar
    real(rk), intent(out) :: MInt(0:N, 0:N, 2), MProj(0:N, 0:N, 2)
    real(rk) :: x(0:N), w(0:N), wb(0:N), xs(0:N), V(0:N, 0:N, 2)
    real(rk) :: xg(0:N), wg(0:N), VGP(0:N, 0:N), VL(0:N, 0:N), Vphi(0:N, 0:N), dummy
    integer :: k, i, j
    call lgl_nodes(N, x, w)
    call barycentric(N, x, wb)
    do k = 1, 2
      if (k == 1) then
        xs = 0.5_rk*(x - 1.0_rk)
      else
        xs = 0.5_rk*(x + 1.0_rk)
      end if
      call nodal_vdm(N, N, wb, x, xs, V(:, :, k))   ! V(p,l) = l_l(xs_p)
      MInt(:, :, k) = transpose(V(:, :, k))
    end do
    if (whichMortar == 1) then
      do k = 1, 2
        do i = 0, N
          do j = 0, N
            MProj(j, i, k) = w(j)/w(i)*V(j, i, k)
          end do
        end do
      end do
    else
      call gauss_nodes(N, xg, wg)
      call vandermonde(N, NODE_GL, N, NODE_GAUSS, .false., VGP)   ! VGP(g,j) = l_j(xg_g)
      do i = 0, N
        VGP(i, :) = wg(i)*VGP(i, :)
      end do
      do i = 0, N
        do j = 0, N
          call legendre(j, x(i), VL(i, j), dummy)
        end do
      end do
      do k = 1, 2
        do i = 0, N
          do j = 0, N
            if (k == 1) then
              call legendre(j, 0.5_rk*(xg(i) - 1.0_rk), Vphi(i, j), dummy)
            else
              call legendre(j, 0.5_rk*(xg(i) + 1.0_rk), Vphi(i, j), dummy)
            end if
          end do
        end do
        MProj(:, :, k) = transpose(matmul(VL, matmul(transpose(Vphi), VGP)))
      end do
    end if
  end subroutine mortar_ops

  !> Big side data -> the 4 (type 1) or 2 (type 2: halves in eta, 3: halves in xi) small sides,
  !> in the reference's order of summation (InterpolateBigToSmall, mortar_vars.f90:203-266).
  subroutine interp_big_to_small(nd, N, mtype, MInt, Big, Small)
    integer, intent(in) :: nd, N, mtype
    real(rk), intent(in) :: MInt(0:N, 0:N, 2), Big(nd, 0:N, 0:N)
    real(rk), intent(out) :: Small(nd, 0:N, 0:N, 4)
    real(rk) :: tmp(nd, 0:N, 0:N)
    integer :: p, q, l, iNb, jNb, im
    Small = 0.0_rk
    select case (mtype)
    case (1)
      do jNb = 1, 2
        tmp = 0.0_rk
        do q = 0, N
          do p = 0, N
            do l = 0, N
              tmp(:, p, q) = tmp(:, p, q) + MInt(l, q, jNb)*Big(:, p, l)
            end do
          end do
        end do
        do iNb = 1, 2
          im = iNb + 2*(jNb - 1)
          do q = 0, N
            do p = 0, N
              do l = 0, N
                Small(:, p, q, im) = Small(:, p, q, im) + MInt(l, p, iNb)*tmp(:, l, q)
              end do
            end do
          end do
        end do
      end do
    case (2)
      do jNb = 1, 2
        do q = 0, N
          do p = 0, N
            do l = 0, N
              Small(:, p, q, jNb) = Small(:, p, q, jNb) + MInt(l, q, jNb)*Big(:, p, l)
            end do
          end do
        end do
      end do
    case (3)
      do iNb = 1, 2
        do q = 0, N
          do p = 0, N
            do l = 0, N
              Small(:, p, q, iNb) = Small(:, p, q, iNb) + MInt(l, p, iNb)*Big(:, l, q)
            end do
          end do
        end do
      end do
    end select
  end subroutine interp_big_to_small

end module fxh_basis
