! xtfem_hip.f90 -- Fortran binding of libxtfem_hip.so for the XTFEM driver.
!
! Module xtfem_hip declares the C ABI of include/xtfem.h with BIND(C) interfaces
! (scalars by value, arrays by reference).  The external subroutines below keep the
! reference's names and argument lists (kronecker.f90:19-365, ilut.f:3-44,
! 152-495, 2383-2432, delelement.f90:1-59, 109-247), so a driver that calls them
! through implicit interfaces (main3d.f90) links against this file and the library
! in place of kronecker.f90, ilut.f and delelement.f90.  Library failures stop the
! program, as the reference's fatal paths do; the reference's own ierr conventions
! are passed through unchanged.  kronpgmres
! writes the residual history to unit iout in the reference's format
! (kronecker.f90:258, 311, 354).  The live path (mumps.f90, kronecker_mumps.f90)
! carries a MUMPS structure in its argument lists; mumps_solver_hip and mpipgmres_hip
! take an opaque factor handle in its place.
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

    integer(c_int) function xt_mpipgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, pr, rhs, sol, &
        im, eps, maxits, hist, hist_cap, nhist, its, ierr) bind(C, name="xt_mpipgmres")
      import :: c_int, c_double, c_ptr
      real(c_double), intent(in) :: a(*), b(*), kv(*), mv(*)
      integer(c_int), value :: q, p, im, maxits, hist_cap
      real(c_double), value :: eps
      integer(c_int), intent(in) :: ki(*), kj(*), mi(*), mj(*)
      type(c_ptr), value :: pr
      real(c_double), intent(inout) :: rhs(*), sol(*), hist(*)
      integer(c_int), intent(out) :: nhist, its, ierr
    end function

    integer(c_int) function xt_prec_create_band(p, rowptr, colidx, vals, block, out) &
        bind(C, name="xt_prec_create_band")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: p, block
      integer(c_int), intent(in) :: rowptr(*), colidx(*)
      real(c_double), intent(in) :: vals(*)
      type(c_ptr), intent(out) :: out
    end function

    subroutine xt_prec_destroy(pr) bind(C, name="xt_prec_destroy")
      import :: c_ptr
      type(c_ptr), value :: pr
    end subroutine

    integer(c_int) function xt_chkelesta(ne, ngpe, dc, d, elesta, gpsta, nef, nefail) &
        bind(C, name="xt_chkelesta")
      import :: c_int, c_double
      integer(c_int), value :: ne, ngpe
      real(c_double), value :: dc
      real(c_double), intent(in) :: d(*)
      integer(c_int), intent(inout) :: elesta(*), gpsta(*)
      integer(c_int), intent(out) :: nef, nefail(*)
    end function

    integer(c_int) function xt_delelement(nn, ne, ndofn, nnse, ngpe, ncmp, nquads, xconn, &
        xcoord, dns, dnxyz, sd, stiff, wqs, nef, nefail, ki, kj, kv, ndof, knz) &
        bind(C, name="xt_delelement")
      import :: c_int, c_double
      integer(c_int), value :: nn, ne, ndofn, nnse, ngpe, ncmp, nquads, nef, ndof
      integer(c_int), intent(in) :: xconn(*), nefail(*)
      real(c_double), intent(in) :: xcoord(*), dns(*), dnxyz(*), sd(*), stiff(*), wqs(*)
      integer(c_int), intent(inout) :: ki(*), kj(*), knz
      real(c_double), intent(inout) :: kv(*)
    end function

    integer(c_int) function xt_chkcnsta(nn, ne, nnse, xconn, nef, nefail, necnt) &
        bind(C, name="xt_chkcnsta")
      import :: c_int
      integer(c_int), value :: nn, ne, nnse, nef
      integer(c_int), intent(in) :: xconn(*), nefail(*)
      integer(c_int), intent(inout) :: necnt(*)
    end function

    integer(c_int) function xt_chkndsta(nn, ne, nnse, xconn, nef, nefail, necnt, ndf, ndfail) &
        bind(C, name="xt_chkndsta")
      import :: c_int
      integer(c_int), value :: nn, ne, nnse, nef
      integer(c_int), intent(in) :: xconn(*), nefail(*), necnt(*)
      integer(c_int), intent(out) :: ndf, ndfail(*)
    end function

    integer(c_int) function xt_adddiagonal(nnse, ndofn, ndf, ndfail, ki, kj, kv, ndof, knz) &
        bind(C, name="xt_adddiagonal")
      import :: c_int, c_double
      integer(c_int), value :: nnse, ndofn, ndf, ndof
      integer(c_int), intent(in) :: ndfail(*)
      integer(c_int), intent(inout) :: ki(*), kj(*), knz
      real(c_double), intent(inout) :: kv(*)
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

! The live path (mumps.f90, kronecker_mumps.f90) with the exact banded factor of K as an
! opaque handle in place of mumps_par (a MUMPS structure cannot cross the C ABI):
! mumps_solver_hip(pr, p, ki, kj, kv, job) factors K as mumps_solver(..., job) does
! (job 0: the first factor, main3d.f90:550; job 1: K changed, the handle's factor is
! replaced, main3d.f90:687 and 783), and mpipgmres_hip takes the handle where
! mpipgmres takes mumps_par, the other arguments as the reference's.
subroutine mumps_solver_hip(pr, p, ki, kj, kv, job)
  use xtfem_hip
  implicit none
  type(c_ptr), intent(inout) :: pr
  integer, intent(in) :: p, ki(*), kj(*), job
  real(8), intent(in) :: kv(*)
  if (job /= 0 .and. c_associated(pr)) call xt_prec_destroy(pr)
  pr = c_null_ptr
  call xt_check(xt_prec_create_band(p, ki, kj, kv, 0, pr), 'mumps_solver_hip')
end subroutine mumps_solver_hip

subroutine mpipgmres_hip(a, b, q, ki, kj, kv, mi, mj, mv, p, pr, rhs, sol, im, eps, &
    maxits, iout, ierr)
  use xtfem_hip
  implicit none
  integer, intent(in) :: q, p, ki(*), kj(*), mi(*), mj(*), im, maxits, iout
  real(8), intent(in) :: a(*), b(*), kv(*), mv(*), eps
  type(c_ptr), intent(in) :: pr
  real(8), intent(inout) :: rhs(*), sol(*)
  integer, intent(out) :: ierr
  real(8), allocatable :: hist(:)
  integer :: nhist, its, k
  allocate(hist(maxits + im + 2))
  call xt_check(xt_mpipgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, pr, rhs, sol, im, eps, &
                             maxits, hist, size(hist), nhist, its, ierr), 'mpipgmres')
  ! the residual log of kronecker_mumps.f90:270-271, 344, as kronpgmres writes it
  if (iout >= 0 .and. nhist > 0) write(iout, 199) 0, hist(1)
  if (iout > 0) then
    do k = 2, nhist
      write(iout, 199) k - 1, hist(k)
    end do
  end if
  deallocate(hist)
199 format('   its =', i4, ' rel. res. norm =', d20.6)
end subroutine mpipgmres_hip

! Element deletion (delelement.f90): reference names and argument lists, so
! main3d.f90:632-691 links against these in place of delelement.o.  As in the
! driver, kj/kv must have room for the edited matrix (main3d.f90:757-760
! allocates ne*(nnse*ndofn)**2 entries).
subroutine chkelesta(nn, ne, nnse, ngpe, Dc, D, elesta, gpsta, nef, nefail)
  use xtfem_hip
  implicit none
  integer, intent(in) :: nn, ne, nnse, ngpe
  real(8), intent(in) :: Dc, D(*)
  integer, intent(inout) :: elesta(*), gpsta(*)
  integer, intent(out) :: nef, nefail(*)
  integer :: i
  do i = 1, ne
    nefail(i) = 0
  end do
  call xt_check(xt_chkelesta(ne, ngpe, Dc, D, elesta, gpsta, nef, nefail), 'chkelesta')
end subroutine chkelesta

subroutine delelement(nn, ne, ndofn, nnse, ngpe, ncmp, nquads, xconn, xcoord, dns, dnxyz, &
    sd, stiff, wqs, nef, nefail, ki, kj, kv, ndof, knz)
  use xtfem_hip
  implicit none
  integer, intent(in) :: nn, ne, ndofn, nnse, ngpe, ncmp, nquads, nef, ndof
  integer, intent(in) :: xconn(*), nefail(*)
  real(8), intent(in) :: xcoord(*), dns(*), dnxyz(*), sd(*), stiff(*), wqs(*)
  integer, intent(inout) :: ki(*), kj(*), knz
  real(8), intent(inout) :: kv(*)
  call xt_check(xt_delelement(nn, ne, ndofn, nnse, ngpe, ncmp, nquads, xconn, xcoord, dns, &
                              dnxyz, sd, stiff, wqs, nef, nefail, ki, kj, kv, ndof, knz), &
                'delelement')
end subroutine delelement

subroutine chkcnsta(nn, ne, nnse, xconn, nef, nefail, necnt)
  use xtfem_hip
  implicit none
  integer, intent(in) :: nn, ne, nnse, nef, xconn(*), nefail(*)
  integer, intent(inout) :: necnt(*)
  call xt_check(xt_chkcnsta(nn, ne, nnse, xconn, nef, nefail, necnt), 'chkcnsta')
end subroutine chkcnsta

subroutine chkndsta(nn, ne, nnse, xconn, nef, nefail, necnt, ndf, ndfail)
  use xtfem_hip
  implicit none
  integer, intent(in) :: nn, ne, nnse, nef, xconn(*), nefail(*), necnt(*)
  integer, intent(out) :: ndf, ndfail(*)
  call xt_check(xt_chkndsta(nn, ne, nnse, xconn, nef, nefail, necnt, ndf, ndfail), 'chkndsta')
end subroutine chkndsta

subroutine adddiagonal(nnse, ndofn, ndf, ndfail, ki, kj, kv, ndof, knz)
  use xtfem_hip
  implicit none
  integer, intent(in) :: nnse, ndofn, ndf, ndof, ndfail(*)
  integer, intent(inout) :: ki(*), kj(*), knz
  real(8), intent(inout) :: kv(*)
  call xt_check(xt_adddiagonal(nnse, ndofn, ndf, ndfail, ki, kj, kv, ndof, knz), 'adddiagonal')
end subroutine adddiagonal
