! Runs the reference's deletion sequence (main3d.f90:650-667: delelement, chkcnsta,
! chkndsta, adddiagonal) through the Fortran binding (xtfem_hip.f90) on inputs the
! test writes to a stream file (argument 1), and writes the edited K, the survivor
! counts and the orphan nodes to argument 2.  Host work only: no device is needed.
program delete_driver
  implicit none
  integer :: nn, ne, ndof, knz, nef, cap, ndf, u
  integer, allocatable :: xconn(:), nefail(:), ki(:), kj(:), necnt(:), ndfail(:)
  real(8), allocatable :: xcoord(:), dns(:), dnxyz(:), kv(:)
  real(8) :: sd(54), stiff(36), wqs(2)
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nn, ne, ndof, knz, nef
  cap = knz + nef * 576 + 3 * nn
  allocate(xconn(ne * 8), xcoord(nn * 3), dns(24 * 8), dnxyz(72 * 24), nefail(nef))
  allocate(ki(ndof + 1), kj(cap), kv(cap), necnt(nn), ndfail(nn))
  read(u) xconn, xcoord, dns, dnxyz, sd, stiff, wqs, nefail, ki, kj(1:knz), kv(1:knz), necnt
  close(u)
  call delelement(nn, ne, 3, 8, 8, 6, 2, xconn, xcoord, dns, dnxyz, sd, stiff, wqs, nef, &
                  nefail, ki, kj, kv, ndof, knz)
  call chkcnsta(nn, ne, 8, xconn, nef, nefail, necnt)
  call chkndsta(nn, ne, 8, xconn, nef, nefail, necnt, ndf, ndfail)
  call adddiagonal(8, 3, ndf, ndfail, ki, kj, kv, ndof, knz)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) knz, ndf, ki, kj(1:knz), kv(1:knz), necnt, ndfail(1:ndf)
  close(u)
end program delete_driver
