!==================================================================================================
!> Minimal parameter.ini reader with the reference's accessor names (GETINT, GETREAL, GETSTR,
!> GETLOGICAL, GETREALARRAY, GETINTARRAY of src/readintools/readintools.f90:1002-1099).
!> Syntax: "key = value", keys case-insensitive, '!' and '#' start comments, a default is
!> used when the key is absent, and a missing key without default stops the run.
!==================================================================================================
module fxh_readini
  use fxh_basis, only: rk
  implicit none
  private
  public :: read_ini, GETINT, GETREAL, GETSTR, GETLOGICAL, GETREALARRAY, GETINTARRAY, has_key

  integer, parameter :: MAXKEYS = 256
  character(len=64) :: keys(MAXKEYS)
  character(len=256) :: vals(MAXKEYS)
  integer :: nkeys = 0

contains

  function lower(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: i
    r = s
    do i = 1, len(s)
      if (r(i:i) >= 'A' .and. r(i:i) <= 'Z') r(i:i) = achar(iachar(r(i:i)) + 32)
    end do
  end function lower

  subroutine read_ini(fname)
    character(len=*), intent(in) :: fname
    character(len=512) :: line
    integer :: u, ios, c, eq
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(A)') ' ERROR: cannot open parameter file '//trim(fname)
      stop 2
    end if
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      c = scan(line, '!#')
      if (c > 0) line = line(1:c-1)
      eq = index(line, '=')
      if (eq == 0) cycle
      if (nkeys >= MAXKEYS) exit
      nkeys = nkeys + 1
      keys(nkeys) = lower(adjustl(line(1:eq-1)))
      vals(nkeys) = adjustl(line(eq+1:))
    end do
    close(u)
  end subroutine read_ini

  logical function has_key(key)
    character(len=*), intent(in) :: key
    integer :: i
    has_key = .false.
    do i = 1, nkeys
      if (trim(keys(i)) == lower(trim(key))) has_key = .true.
    end do
  end function has_key

  function lookup(key, default) result(v)
    character(len=*), intent(in) :: key
    character(len=*), intent(in), optional :: default
    character(len=256) :: v
    integer :: i
    do i = nkeys, 1, -1
      if (trim(keys(i)) == lower(trim(key))) then
        v = vals(i); return
      end if
    end do
    if (present(default)) then
      v = default
    else
      write(*, '(A)') ' ERROR: parameter "'//trim(key)//'" not set and has no default'
      stop 2
    end if
  end function lookup

  integer function GETINT(key, default)
    character(len=*), intent(in) :: key
    character(len=*), intent(in), optional :: default
    character(len=256) :: v
    v = lookup(key, default)
    read(v, *) GETINT
  end function GETINT

  real(rk) function GETREAL(key, default)
    character(len=*), intent(in) :: key
    character(len=*), intent(in), optional :: default
    character(len=256) :: v
    v = lookup(key, default)
    read(v, *) GETREAL
  end function GETREAL

  function GETSTR(key, default) result(s)
    character(len=*), intent(in) :: key
    character(len=*), intent(in), optional :: default
    character(len=256) :: s
    s = lookup(key, default)
  end function GETSTR

  logical function GETLOGICAL(key, default)
    character(len=*), intent(in) :: key
    character(len=*), intent(in), optional :: default
    character(len=256) :: v
    v = lower(adjustl(lookup(key, default)))
    GETLOGICAL = (v(1:1) == 't' .or. v(1:2) == '.t' .or. v(1:1) == '1')
  end function GETLOGICAL

  function GETREALARRAY(key, n, default) result(a)
    character(len=*), intent(in) :: key
    integer, intent(in) :: n
    character(len=*), intent(in), optional :: default
    real(rk) :: a(n)
    character(len=256) :: v
    integer :: i
    v = lookup(key, default)
    do i = 1, len_trim(v)
      if (v(i:i) == '(' .or. v(i:i) == ')' .or. v(i:i) == '/') v(i:i) = ' '
    end do
    read(v, *) a
  end function GETREALARRAY

  function GETINTARRAY(key, n, default) result(a)
    character(len=*), intent(in) :: key
    integer, intent(in) :: n
    character(len=*), intent(in), optional :: default
    integer :: a(n)
    character(len=256) :: v
    integer :: i
    v = lookup(key, default)
    do i = 1, len_trim(v)
      if (v(i:i) == '(' .or. v(i:i) == ')' .or. v(i:i) == '/') v(i:i) = ' '
    end do
    read(v, *) a
  end function GETINTARRAY

end module fxh_readini
