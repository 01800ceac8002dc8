! xtfem_hip.f90 -- Fortran binding of libxtfem_hip.so for the XTFEM driver.
!
! Module xtfem_hip declares the C ABI of include/xtfem.h with BIND(C) interfaces
! (scalars by value, arrays by reference).  The external subroutines below keep the
! reference's names and argument lists (kronecker.f90:19-365, ilut.f:3-44,
! 152-495, 2383-2432), so a driver that calls them through implicit interfaces
! (main3d.f90) links against this file and the library in place of kronecker.f90 and
! ilut.f.  Library failures stop the program, as the reference's fatal paths do;
! the reference's own ierr conventions are passed through unchanged.  kronpgmres
! writes the residual history to unit iout in the reference's format
! (kronecker.f90:258, 311, 354).
module xtfem_hip
  use iso_c_binding
  implicit none

  interface
    function xt_last_error() bind(C, name="xt_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function

    integer(c_int) function xt_kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, x, y) &
        bind(C, name="xt_kronspmv")
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), b(*), kv(*), mv(*), x(*)
      integer(c_int), value :: q, p
      integer(c_int), intent(in) :: ki(*), kj(*), mi(*), mj(*)
      real(c_double), intent(inout) :: y(*)
    end function

    integer(c_int) function xt_kronspgemv(a, b, alpha, beta, q, ki, kj, kv, mi, mj, mv, &
        p, x, y) bind(C, name="xt_kronspgemv")
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), b(*), kv(*), mv(*), x(*)
      real(c_double), value :: alpha, beta
      integer(c_int), value :: q, p
      integer(c_int), intent(in) :: ki(*), kj(*), mi(*), mj(*)
      real(c_double), intent(inout) :: y(*)
    end function

    integer(c_int) function xt_amux(n, x, y, a, ja, ia) bind(C, name="xt_amux")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(*), a(*)
      real(c_double), intent(inout) :: y(*)
      integer(c_int), intent(in) :: ja(*), ia(*)
    end function

    integer(c_int) function xt_ilut(n, a, ja, ia, lfil, droptol, alu, jlu, ju, iwk, w, jw, &
        ierr) bind(C, name="xt_ilut")
      import :: c_int, c_double
      integer(c_int), value :: n, lfil, iwk
      real(c_double), value :: droptol
      real(c_double), intent(in) :: a(*)
      integer(c_int), intent(in) :: ja(*), ia(*)
      real(c_double), intent(inout) :: alu(*), w(*)
      integer(c_int), intent(inout) :: jlu(*), ju(*), jw(*)
      integer(c_int), intent(out) :: ierr
    end function

    integer(c_int) function xt_lusol(n, y, x, alu, jlu, ju) bind(C, name="xt_lusol")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(in) :: y(*), alu(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), intent(in) :: jlu(*), ju(*)
    end function

    integer(c_int) function xt_kronlusol(a, q, alu, jlu, ju, p, x, y) &
        bind(C, name="xt_kronlusol")
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), alu(*)
      integer(c_int), value :: q, p
      integer(c_int), intent(in) :: jlu(*), ju(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
    end function

    integer(c_int) function xt_kronpgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, alu, jlu, ju, &
        rhs, sol, im, eps, maxits, hist, hist_cap, nhist, its, ierr) &
        bind(C, name="xt_kronpgmres")
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), b(*), kv(*), mv(*), alu(*)
      integer(c_int), value :: q, p, im, maxits, hist_cap
      real(c_double), value :: eps
      integer(c_int), intent(in) :: ki(*), kj(*), mi(*), mj(*), jlu(*), ju(*)
      real(c_double), intent(inout) :: rhs(*), sol(*), hist(*)
      integer(c_int), intent(out) :: nhist, its, ierr
    end function
  end interface

contains

  subroutine xt_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc == 0) return
    call c_f_pointer(xt_last_error(), msg, [512])
    k = 0
    do while (k < 512)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write(*, '(A,A,A,I0,A)') ' libxtfem_hip: ', what, ' failed (code ', rc, ')'
    if (k > 0) write(*, '(A,512A1)') '   ', msg(1:k)
    stop
  end subroutine

end module xtfem_hip


subroutine kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, x, y)
  use xtfem_hip
  implicit none
  integer, intent(in) :: q, p, ki(*), kj(*), mi(*), mj(*)
  real(8), intent(in) :: a(*), b(*), kv(*), mv(*), x(*)
  real(8), intent(inout) :: y(*)
  call xt_check(xt_kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, x, y), 'kronspmv')
end subroutine kronspmv

subroutine kronspgemv(a, b, alpha, beta, q, ki, kj, kv, mi, mj, mv, p, x, y)
  use xtfem_hip
  implicit none
  integer, intent(in) :: q, p, ki(*), kj(*), mi(*), mj(*)
  real(8), intent(in) :: a(*), b(*), alpha, beta, kv(*), mv(*), x(*)
  real(8), intent(inout) :: y(*)
  call xt_check(xt_kronspgemv(a, b, alpha, beta, q, ki, kj, kv, mi, mj, mv, p, x, y), &
                'kronspgemv')
end subroutine kronspgemv

subroutine amux(n, x, y, a, ja, ia)
  use xtfem_hip
  implicit none
  integer, intent(in) :: n, ja(*), ia(*)
  real(8), intent(in) :: x(*), a(*)
  real(8), intent(inout) :: y(*)
  call xt_check(xt_amux(n, x, y, a, ja, ia), 'amux')
end subroutine amux

subroutine ilut(n, a, ja, ia, lfil, droptol, alu, jlu, ju, iwk, w, jw, ierr)
  use xtfem_hip
  implicit none
  integer, intent(in) :: n, ja(*), ia(*), lfil, iwk
  real(8), intent(in) :: a(*), droptol
  real(8), intent(inout) :: alu(*), w(*)
  integer, intent(inout) :: jlu(*), ju(*), jw(*)
  integer, intent(out) :: ierr
  call xt_check(xt_ilut(n, a, ja, ia, lfil, droptol, alu, jlu, ju, iwk, w, jw, ierr), 'ilut')
end subroutine ilut

subroutine lusol(n, y, x, alu, jlu, ju)
  use xtfem_hip
  implicit none
  integer, intent(in) :: n, jlu(*), ju(*)
  real(8), intent(in) :: y(*), alu(*)
  real(8), intent(inout) :: x(*)
  call xt_check(xt_lusol(n, y, x, alu, jlu, ju), 'lusol')
end subroutine lusol

subroutine kronlusol(a, q, alu, jlu, ju, p, x, y)
  use xtfem_hip
  implicit none
  integer, intent(in) :: q, p, jlu(*), ju(*)
  real(8), intent(in) :: a(*), alu(*), x(*)
  real(8), intent(inout) :: y(*)
  call xt_check(xt_kronlusol(a, q, alu, jlu, ju, p, x, y), 'kronlusol')
end subroutine kronlusol

subroutine kronpgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, alu, jlu, ju, rhs, sol, im, eps, &
    maxits, iout, ierr)
  use xtfem_hip
  implicit none
  integer, intent(in) :: q, p, ki(*), kj(*), mi(*), mj(*), jlu(*), ju(*), im, maxits, iout
  real(8), intent(in) :: a(*), b(*), kv(*), mv(*), alu(*), eps
  real(8), intent(inout) :: rhs(*), sol(*)
  integer, intent(out) :: ierr
  real(8), allocatable :: hist(:)
  integer :: nhist, its, k
  allocate(hist(maxits + im + 2))
  call xt_check(xt_kronpgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, alu, jlu, ju, rhs, sol, &
                              im, eps, maxits, hist, size(hist), nhist, its, ierr), &
                'kronpgmres')
  ! history: entry 0 is the initial residual (written when iout >= 0), entry k the
  ! estimate after iteration k (written when iout > 0)
  if (iout >= 0 .and. nhist > 0) write(iout, 199) 0, hist(1)
  if (iout > 0) then
    do k = 2, nhist
      write(iout, 199) k - 1, hist(k)
    end do
  end if
  deallocate(hist)
199 format('   its =', i4, ' rel. res. norm =', d20.6)
end subroutine kronpgmres
