! Calls the reference-named ilut through the Fortran binding (xtfem_hip.f90) on a
! 2-D five-point matrix (1-based CSR) and prints the MSR factors for the test to
! compare with the oracle.  ilut runs on the host: no device is needed.
program ilut_driver
  implicit none
  integer, parameter :: m = 8, n = m * m, lfil = 3, iwk = n * (2 * lfil + 2) + n + 2
  integer :: ia(n + 1), ja(5 * n), jlu(iwk), ju(n), jw(2 * n), ierr, i, j, r, k
  real(8) :: a(5 * n), alu(iwk), w(n + 1)
  k = 0
  ia(1) = 1
  do i = 1, m
    do j = 1, m
      r = (i - 1) * m + j
      if (i > 1) then; k = k + 1; ja(k) = r - m; a(k) = -1.0d0 - 0.01d0 * r; end if
      if (j > 1) then; k = k + 1; ja(k) = r - 1; a(k) = -1.0d0; end if
      k = k + 1; ja(k) = r; a(k) = 4.5d0
      if (j < m) then; k = k + 1; ja(k) = r + 1; a(k) = -1.0d0 + 0.02d0 * j; end if
      if (i < m) then; k = k + 1; ja(k) = r + m; a(k) = -1.0d0; end if
      ia(r + 1) = k + 1
    end do
  end do
  call ilut(n, a, ja, ia, lfil, 1.0d-3, alu, jlu, ju, iwk, w, jw, ierr)
  write(*, '(A,I0)') 'ierr ', ierr
  write(*, '(A,I0)') 'n ', n
  write(*, '(A)') 'csr'
  write(*, '(*(I0,1X))') ia
  write(*, '(*(I0,1X))') ja(1:ia(n + 1) - 1)
  write(*, '(*(ES25.17E3,1X))') a(1:ia(n + 1) - 1)
  write(*, '(A)') 'msr'
  write(*, '(*(I0,1X))') ju
  write(*, '(*(I0,1X))') jlu(1:jlu(n + 1) - 1)
  write(*, '(*(ES25.17E3,1X))') alu(1:jlu(n + 1) - 1)
end program ilut_driver
