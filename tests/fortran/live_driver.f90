! The live path through the Fortran binding: mumps_solver_hip factors K (job 0) and
! mpipgmres_hip solves (a K + b M) sol = rhs with that factor (q = 1: K the 1-D
! Laplacian of order 4, M the identity, a = 1, b = 0.5, rhs all ones).  Prints ierr and
! the solution.
program live_driver
  use iso_c_binding
  implicit none
  integer, parameter :: n = 4, q = 1
  integer :: ki(n + 1), kj(3 * n - 2), mi(n + 1), mj(n), ierr, k, nz
  real(8) :: kv(3 * n - 2), mv(n), a(1), b(1), rhs(n), sol(n)
  type(c_ptr) :: pr
  nz = 0
  do k = 1, n
    ki(k) = nz + 1
    if (k > 1) then
      nz = nz + 1
      kj(nz) = k - 1
      kv(nz) = -1d0
    end if
    nz = nz + 1
    kj(nz) = k
    kv(nz) = 2d0
    if (k < n) then
      nz = nz + 1
      kj(nz) = k + 1
      kv(nz) = -1d0
    end if
    mi(k) = k
    mj(k) = k
    mv(k) = 1d0
  end do
  ki(n + 1) = nz + 1
  mi(n + 1) = n + 1
  a(1) = 1d0
  b(1) = 0.5d0
  rhs = 1d0
  sol = 0d0
  pr = c_null_ptr
  call mumps_solver_hip(pr, n, ki, kj, kv, 0)
  call mpipgmres_hip(a, b, q, ki, kj, kv, mi, mj, mv, n, pr, rhs, sol, 10, 1d-12, 100, -1, &
                     ierr)
  print '(A,I0)', 'ierr ', ierr
  print '(A,4ES25.16)', 'sol ', sol
end program live_driver
